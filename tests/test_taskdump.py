"""Task-group dumps (the reference's saveInputs format, src/r.cpp:275-309) and their replay.

CPU: the reader, the enum decoding of src/common.cpp:13-60, the ctypes mirror of fedm_task_group against
the C header's layout (compiled with gcc), and that the committed fixture is what the dump writer makes
from its calls.  GPU: each dumped group replayed through fedm_run_task_group gives the oracle's result
for the call that the group records (rc, k range, rho, MAE, predictions, copredictions, coefficients)."""
import ctypes
import importlib
import json
import os
import subprocess

import numpy as np
import pytest

from tests.conftest import GOLDEN, ROOT, load_product

DUMPS = os.path.join(GOLDEN, "task_dumps.json")


def _taskdump():
    load_product()
    return importlib.import_module("fastedm_r_amd.taskdump")


def _writer():
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_task_dumps", os.path.join(GOLDEN, "make_task_dumps.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_enum_decoding_falls_through_to_the_last_spelling():
    td = _taskdump()
    assert td.decode("Simplex", td.ALGORITHMS) == 0
    assert td.decode("SMap", td.ALGORITHMS) == 1
    assert td.decode("simplex", td.ALGORITHMS) == 1  # from_json compares the exact spelling
    assert [td.decode(d, td.DISTANCES) for d in ("MeanAbsoluteError", "Euclidean", "Wasserstein", "x")] == [0, 1, 2, 2]
    assert [td.decode(m, td.METRICS) for m in ("Diff", "CheckSame", "")] == [0, 1, 1]


def test_reader_requires_every_field(tmp_path):
    td = _taskdump()
    groups = td.read_task_groups(DUMPS)
    assert len(groups) == len(_writer().calls())
    bad = dict(groups[0])
    del bad["usable"]
    p = tmp_path / "bad.json"
    p.write_text(json.dumps([bad]))
    with pytest.raises(ValueError, match="usable"):
        td.read_task_groups(str(p))
    bad = json.loads(json.dumps(groups[0]))
    del bad["generator"]["_observation_number"]
    p.write_text(json.dumps(bad))
    with pytest.raises(ValueError, match="_observation_number"):
        td.read_task_groups(str(p))


def test_fixture_matches_the_writer():
    """The committed dumps are those the writer makes now (generator state, usable rows, options)."""
    made = json.loads(json.dumps(_writer().dumps()))
    with open(DUMPS) as f:
        assert json.load(f) == made


def test_task_group_layout_matches_header(tmp_path):
    """Every field offset and the size of the ctypes mirror equal the C compiler's for fedm_task_group."""
    td = _taskdump()
    names = [f[0] for f in td.TaskGroup._fields_]
    src = tmp_path / "layout.c"
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "fastedm_hip.h"', "int main(void) {",
             '  printf("%zu\\n", sizeof(fedm_task_group));']
    lines += [f'  printf("%zu\\n", offsetof(fedm_task_group, {n}));' for n in names]
    lines += ["  return 0;", "}"]
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    assert out[0] == ctypes.sizeof(td.TaskGroup)
    assert out[1:] == [getattr(td.TaskGroup, n).offset for n in names]


def test_group_struct_fields():
    td = _taskdump()
    groups = td.read_task_groups(DUMPS)
    names = list(_writer().calls())
    g = groups[names.index("panel_inf")]
    keep = []
    s = td.group_struct(g, precision=1, seed=5, keep=keep)
    n = len(g["generator"]["_t"])
    assert s.n == n and s.panel_mode == 1 and s.idw == -1.0 and s.precision == 1 and s.seed == 5
    assert bool(s.panel) and not bool(s.xmap) and not bool(s.co_x)
    assert [s.E[i] for i in range(s.num_E)] == [2, 3]
    assert s.num_libraries == 0 and s.explore == 1
    assert [s.usable[i] for i in range(n)] == g["usable"]
    g = groups[names.index("extras_missing")]
    s = td.group_struct(g, keep=keep)
    assert s.num_extras == 1 and s.extra_metrics[0] == 0 and s.missing_distance > 0
    ex = np.ctypeslib.as_array(s.extras, shape=(n_ := len(g["generator"]["_t"]),))
    assert np.array_equal(ex, np.asarray(g["generator"]["_extras"][0]))
    assert n_ == s.n


def _same(a, b, key):
    x, y = np.asarray(a[key], dtype=np.float64), np.asarray(b[key], dtype=np.float64)
    assert x.shape == y.shape, key
    assert np.array_equal(np.isnan(x), np.isnan(y)), key
    ok = ~np.isnan(y)
    assert np.all(np.abs(x[ok] - y[ok]) <= 1e-12 + 1e-10 * np.abs(y[ok])), key


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1], ids=["plain", "mixed"])
def test_replay_matches_the_recorded_call(precision):
    from oracle import oracle
    td = _taskdump()
    groups = td.read_task_groups(DUMPS)
    for (name, kw), g in zip(_writer().calls().items(), groups):
        print(f"task dump {name} precision {precision}", flush=True)
        got = td.replay(g, precision=precision, seed=kw.get("seed", 0))
        want = oracle.edm(**kw)
        assert got["rc"] == want["rc"], name
        assert (got["kMin"], got["kMax"]) == (want["kMin"], want["kMax"]), name
        for key in ("E", "library", "theta"):
            np.testing.assert_array_equal(got["stats"][key], want["stats"][key])
        _same(got["stats"], want["stats"], "rho")
        _same(got["stats"], want["stats"], "mae")
        if "copredStats" in want:
            _same(got["copredStats"], want["copredStats"], "rho")
        for key in ("predictions", "copredictions", "coeffs", "missingdistance"):
            assert (key in got) == (key in want), (name, key)
            if key in want:
                _same(got, want, key)
