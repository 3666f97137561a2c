"""CPU checks of the drop-in boundary: the library builds, loads and exports every declared symbol."""
import ctypes
import os
import re

from tests.conftest import ROOT, load_product

HEADER = os.path.join(ROOT, "include", "fastedm_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:edm|fedm)_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_seam():
    names = declared_functions()
    for required in ["edm_predict", "edm_predict_batch", "edm_ctx_create", "edm_ctx_destroy", "fedm_run_command"]:
        assert required in names


def test_library_exports_every_declared_symbol():
    fe = load_product()
    lib = ctypes.CDLL(fe.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_package_exports_match_header():
    fe = load_product()
    assert sorted(fe.EXPORTS) == sorted(n for n in declared_functions() if n in fe.EXPORTS)
    assert set(fe.EXPORTS) <= set(declared_functions())
