"""The benchmark configurations' inputs are pinned by SHA-256 of their float64 bytes.

The generators (tests/configs.py, bench.py) use numpy element-wise arithmetic, one rounding per
operation, so the series are the same on every IEEE-754 host; a different checksum means a
generator or the host's arithmetic changed, and every measured figure would then describe other data."""
import hashlib

import numpy as np

from tests import configs

SHA256 = {
    "cfg1": "cccbc84505b25fc37cd2a7d1c2ad16d5af260d6525acf8bb6494221f1fc98c59",
    "cfg2": "f86918ecff2f25d6a5e993ed0f56730a6543b09b3ea8a42befa59c1477b5bf47",
    "cfg3": "a7c54e29ab21cea49eb230dec48e3a6391177798cc960023dca51935d5199612",
    "cfg4": "f052a4b3747d7b7b2ecb734a807c85286149f5c75d09bce4c44a44468c903b23",
    "cfg5": "38d548d43f24138699a80342458c976ea366b802b42799e9294d220af7e3e5bd",
}


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def test_cfg1_logistic_with_noise():
    t, x = configs.cfg1_series()
    np.testing.assert_array_equal(t, np.arange(1, 501))
    assert _sha(x) == SHA256["cfg1"]


def test_cfg2_cfg3_coupled_maps():
    assert _sha(np.stack(configs.coupled_maps(5000))) == SHA256["cfg2"]
    assert _sha(np.stack(configs.coupled_maps(20000))) == SHA256["cfg3"]


def test_bench_uses_the_cfg3_series():
    import bench
    assert _sha(np.stack(bench.coupled_logistic(20000))) == SHA256["cfg3"]


def test_cfg4_panels():
    t, x, panel = configs.cfg4_panels()
    assert _sha(np.stack([t, x, panel.astype(np.float64)])) == SHA256["cfg4"]
    assert panel.max() == 511 and 0.04 < np.isnan(x).mean() < 0.06


def test_cfg5_lorenz96():
    s = configs.lorenz96()
    assert s.shape == (100000, 40)
    assert _sha(s) == SHA256["cfg5"]
