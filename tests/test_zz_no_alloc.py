"""lpa_ici_forward_noalloc (no device allocation; its workspace is a static
array per device) against lpa_ici_forward, through the C ABI: the same
outputs and scale maps bit for bit, on two calls in a row, the second with a
smaller batch (the array keeps the first call's contents), for the batched
and plain forms, the exact walk and a weighted kernel set.  Also: nothing is
taken from the stream-ordered pool, and the refusals (a set not resident
yet, a batch beyond the workspace, LPA_FWD_GRAPH) are status codes."""
import ctypes

import numpy as np
import pytest
import torch

import lpa_ici
from conftest import GAUSSIAN_M20, SIGMA, kernels, noisy_synth

pytestmark = pytest.mark.gpu


def stream_ptr():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def run(entry, ks, xs, sig, gam, flags):
    """Call a forward entry point on device images xs; returns (status, ys, maps)."""
    ys = [lpa_ici.same_layout(x) for x in xs]
    ms = [torch.empty((ks.D,) + tuple(x.shape), dtype=torch.uint8, device="cuda") for x in xs]
    for y, m in zip(ys, ms):
        y.fill_(float("nan"))
        m.fill_(255)
    arr = (lpa_ici.LpaImage * len(xs))(*[lpa_ici.image_struct(x, y, m, s, g)
                                         for x, y, m, s, g in zip(xs, ys, ms, sig, gam)])
    st = getattr(lpa_ici.lib(), entry)(ks.handle, arr, len(xs), int(flags), stream_ptr())
    torch.cuda.synchronize()
    return st, ys, ms


def batch(shapes, seed):
    rng = np.random.default_rng(seed)
    xs = [torch.from_numpy(noisy_synth(w, h, seed=seed + i)).cuda() for i, (w, h) in enumerate(shapes)]
    sig = [float(rng.choice([10, 20, 30, 40])) / 255.0 for _ in shapes]
    gam = [float(rng.choice([1.5, 2.0, 2.5])) for _ in shapes]
    return xs, sig, gam


@pytest.mark.parametrize("name,flags", [("symmetric", 0), ("symmetric", lpa_ici.FWD_PLAIN),
                                        ("symmetric", lpa_ici.FWD_EXACT), ("gaussian", 0)])
def test_noalloc_equals_forward(name, flags):
    ks = kernels() if name == "symmetric" else kernels(GAUSSIAN_M20)
    big = batch([(300, 200), (130, 70), (1, 1), (64, 513), (257, 33)], seed=3)
    small = batch([(40, 31), (96, 64)], seed=9)
    for xs, sig, gam in (big, small):
        st, ya, ma = run("lpa_ici_forward", ks, xs, sig, gam, flags)  # also makes the set resident
        assert st == 0
        lpa_ici.workspace_peak(reset=True)
        st, yb, mb = run("lpa_ici_forward_noalloc", ks, xs, sig, gam, flags)
        assert st == 0, lpa_ici.lib().lpa_last_error()
        assert lpa_ici.workspace_peak() == 0, "the no-alloc forward took memory from the stream-ordered pool"
        for a, b, c, d in zip(ya, yb, ma, mb):
            assert torch.equal(a.view(torch.int32), b.view(torch.int32)) and torch.equal(c, d)


def test_noalloc_refusals():
    xs, sig, gam = batch([(32, 24)], seed=1)
    fresh = lpa_ici.load_kernels()  # never used on a device: not resident
    st, _, _ = run("lpa_ici_forward_noalloc", fresh, xs, sig, gam, 0)
    assert st == 1 and b"not resident" in lpa_ici.lib().lpa_last_error()
    fresh.close()
    ks = kernels()
    assert run("lpa_ici_forward", ks, xs, sig, gam, 0)[0] == 0
    st, _, _ = run("lpa_ici_forward_noalloc", ks, xs, sig, gam, lpa_ici.FWD_GRAPH)
    assert st == 1
    # 4000 descriptors and tables (about 5.6 MB) do not fit the 4 MiB
    # workspace: refused on the host before anything is launched
    x, y = xs[0], lpa_ici.same_layout(xs[0])
    n = 4000
    arr = (lpa_ici.LpaImage * n)(*[lpa_ici.image_struct(x, y, None, sig[0], gam[0]) for _ in range(n)])
    st = lpa_ici.lib().lpa_ici_forward_noalloc(ks.handle, arr, n, 0, stream_ptr())
    assert st == 2 and b"static workspace" in lpa_ici.lib().lpa_last_error()
