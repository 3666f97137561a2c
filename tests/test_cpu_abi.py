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


def _c_layout(tmp_path, struct, names):
    import subprocess
    src = tmp_path / f"layout_{struct}.c"
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "fastedm_hip.h"', "int main(void) {",
             f'  printf("%zu\\n", sizeof({struct}));']
    lines += [f'  printf("%zu\\n", offsetof({struct}, {n}));' for n in names]
    lines += ["  return 0;", "}"]
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / f"layout_{struct}"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    return [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]


def test_ctypes_mirrors_match_header(tmp_path):
    """Field offsets and sizes of the package's ctypes structures equal the C compiler's."""
    fe = load_product()
    for cls, struct in [(fe.Command, "fedm_command"), (fe.Options, "edm_options"),
                        (fe.ManifoldView, "edm_manifold_view"), (fe.Knobs, "fedm_knobs")]:
        names = [f[0] for f in cls._fields_]
        out = _c_layout(tmp_path, struct, names)
        assert out[0] == ctypes.sizeof(cls), struct
        assert out[1:] == [getattr(cls, n).offset for n in names], struct


def test_default_knobs():
    """fedm_default_knobs (no device needed): the documented defaults."""
    fe = load_product()
    L = fe.load_library()
    L.fedm_default_knobs.argtypes = [ctypes.POINTER(fe.Knobs)]
    k = fe.Knobs()
    L.fedm_default_knobs(ctypes.byref(k))
    assert k.as_dict() == {"rank_scale": 1.5, "direct_blocks_per_cu": 0, "direct_alt_waves": 0, "list_blocks": 1024,
                           "mfma_blocks": 2048, "use_mfma": 1, "smap_lds": 0, "direct_prune": 1,
                           "concurrent_routes": 0, "smap_block": 4, "mfma_seed": 3,
                           "mfma_prune": 1, "mfma_split": 0}
