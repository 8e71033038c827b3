import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "lpa-ici-cuda_amd")
GOLDEN = os.path.join(REPO, "tests", "golden")
for p in (REPO, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

import lpa_ici  # noqa: E402  (imports torch first, then liblpa_ici.so)
import oracle  # noqa: E402

SIGMA = 20.0 / 255.0
GAMMA = 2.0
SEED = 20250910
SYMMETRIC = "lpa_kernels_m_1_0_d_16_h12_1_2_4_8_16_24_32_symmetric.txt"
GAUSSIAN_M20 = "lpa_kernels_m_2_0_d_16_h12_1_2_4_8_16_24_32_gaussian.txt"
REF_KERNEL_DIR = "/root/reference/kernels"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def kernel_path(name):
    return os.path.join(lpa_ici.KERNEL_DIR, name + ".gz")


_kcache = {}


def kernels(name=SYMMETRIC):
    if name not in _kcache:
        _kcache[name] = lpa_ici.load_kernels(kernel_path(name))
    return _kcache[name]


@pytest.fixture(scope="session")
def ks():
    return kernels(SYMMETRIC)


def noisy_synth(w, h, seed=1, sigma=SIGMA, noise_seed=SEED):
    return lpa_ici.make_noisy(lpa_ici.synth_image(w, h, seed), sigma, noise_seed)


def golden(name):
    return np.load(os.path.join(GOLDEN, name))


def dense_index(ks, r, i):
    """(kernel, dense weight index) of run weight i; r = ks.runs()."""
    run = int(np.searchsorted(r["run_woff"], i, side="right") - 1)
    k = int(np.searchsorted(r["kernel_first_run"], run, side="right") - 1)
    R, C = ks.rows[k], ks.cols[k]
    row = r["run_row"][run] + R // 2
    col = r["run_col"][run] + C // 2 + (i - r["run_woff"][run])
    return k, int(ks.offsets[k] + row * C + col)
