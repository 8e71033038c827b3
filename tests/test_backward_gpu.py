"""The HIP adjoint (selection frozen at the forward's scale maps) against
the FP64 CPU adjoint and FP64 central differences.

Bar: gradients within 1e-3 relative of FP64 central differences (eps 1e-3)
at perturbations that flip no selection of the FP32 walk; against the exact
FP64 adjoint the FP32 kernels are held to 1e-4 of the gradient's scale.

The output is linear in x, so two-point central differences are exact for
dL/dx.  It is rational in the taps (through sum w^2), where two-point
differences with eps 1e-3 carry O(eps^2) truncation of up to ~1e-3 relative
on some taps; the taps test uses the fourth-order central difference on the
same step (values at +-eps and +-2 eps; within 3e-7 relative of the exact
FP64 adjoint).  The relative test has a floor proportional to the gradient's
scale, |g - fd| <= 1e-3 * max(|fd|, 1e-2 * max|g|): FP32 sums are accurate
to ~1e-6 of the scale, not of components far below it.
"""
import numpy as np
import pytest
import torch

import lpa_ici
import oracle
from conftest import GAMMA, GAUSSIAN_M20, SIGMA, dense_index, kernels, noisy_synth
from lpa_ici.autograd import lpa_ici as lpa_ici_fn, template_taps

pytestmark = pytest.mark.gpu
EPS = 1e-3
TOL_FD = 1e-3


def forward_gpu(ks, x, sigma=SIGMA):
    xd = torch.from_numpy(x).cuda()
    y, m = lpa_ici.denoise(ks, xd, sigma, GAMMA, scale_maps=True)
    return xd, y, m


class Modified:
    """Kernel arrays with one dense weight changed (for the oracle)."""

    def __init__(self, ks, idx, delta):
        self.D, self.S, self.offsets, self.rows, self.cols = ks.D, ks.S, ks.offsets, ks.rows, ks.cols
        self.weights = ks.weights.copy()
        self.weights[idx] = np.float32(self.weights[idx] + delta)
        self.K = ks.K

    def variances(self, sigma):
        return lpa_ici.kernel_variances(self, sigma)


def loss64(x, g, ks, maps, sigma=SIGMA):
    y, _ = oracle.walk_f64(x, ks, sigma, GAMMA, frozen=maps)
    return float((g.astype(np.float64) * y).sum())


def test_adjoint_matches_fp64_adjoint(ks):
    x = noisy_synth(72, 56, seed=31)
    xd, y, m = forward_gpu(ks, x)
    g = torch.randn_like(xd)
    dxs, dk = lpa_ici.backward(ks, [xd], [y], [m], [g], SIGMA, GAMMA)
    torch.cuda.synchronize()
    dx64, dkd64 = oracle.adjoint_f64(x, g.cpu().numpy(), ks, SIGMA, m.cpu().numpy())
    dx = dxs[0].cpu().numpy()
    e_dx = np.abs(dx - dx64).max() / np.abs(dx64).max()
    r = ks.runs()
    idx = np.array([dense_index(ks, r, i)[1] for i in range(len(r["weights"]))])
    dk64 = dkd64[idx]
    e_dk = np.abs(dk.cpu().numpy() - dk64).max() / np.abs(dk64).max()
    print(f"adjoint vs FP64: dx rel {e_dx:.2e}, dk rel {e_dk:.2e}")
    assert e_dx < 1e-4 and e_dk < 1e-4


def adjoint_error(ks, x):
    """Largest deviations of the HIP adjoint from the FP64 adjoint, relative
    to each gradient's scale."""
    xd, y, m = forward_gpu(ks, x)
    g = torch.randn_like(xd)
    dxs, dk = lpa_ici.backward(ks, [xd], [y], [m], [g], SIGMA, GAMMA)
    torch.cuda.synchronize()
    dx64, dkd64 = oracle.adjoint_f64(x, g.cpu().numpy(), ks, SIGMA, m.cpu().numpy())
    r = ks.runs()
    dk64 = dkd64[np.array([dense_index(ks, r, i)[1] for i in range(len(r["weights"]))])]
    e_dx = np.abs(dxs[0].cpu().numpy() - dx64).max() / np.abs(dx64).max()
    e_dk = np.abs(dk.cpu().numpy() - dk64).max() / np.abs(dk64).max()
    return e_dx, e_dk


def test_adjoint_weighted_kernel_set():
    """A kernel set with no ring form (Gaussian weights): yhat and E come from
    the same weighted walk that forms A."""
    ks_g = kernels(GAUSSIAN_M20)
    assert not ks_g.info()["ring_form"]
    e_dx, e_dk = adjoint_error(ks_g, noisy_synth(72, 56, seed=33))
    print(f"weighted set vs FP64: dx rel {e_dx:.2e}, dk rel {e_dk:.2e}")
    assert e_dx < 1e-4 and e_dk < 1e-4


@pytest.mark.parametrize("D,h", [(8, [1, 2, 4, 8]), (32, [1, 3, 6, 12, 24]), (4, [1, 2, 3, 5, 8, 13, 21, 32]),
                                 (8, [1, 2, 3, 4, 5, 6, 8, 10, 12, 16])])
def test_adjoint_generated_wedge_sets(D, h):
    """Generated ring-form sets with other direction counts, radii (7 to 31)
    and scale counts (up to 10, so E takes two groups of 8 scales): the dx
    pass's window and both passes' walks at these shapes, on an image whose
    sides are not multiples of the tiles."""
    g = lpa_ici.wedge_kernels(D, h)
    assert g.info()["ring_form"] == 1
    e_dx, e_dk = adjoint_error(g, noisy_synth(150, 140, seed=D + len(h)))
    print(f"D={D} h={h} vs FP64: dx rel {e_dx:.2e}, dk rel {e_dk:.2e}")
    assert e_dx < 1e-4 and e_dk < 1e-4


@pytest.mark.parametrize("name", ["lpa_kernels_m_1_0_d_8_h12_1_2_4_8.txt", "lpa_kernels_m_1_0_d_16_h12_1_2_4_8_16.txt",
                                  "lpa_kernels_m_1_0_d_8_h12_1_15_15_30.txt", "lpa_kernels_m_0_0_d_4_square.txt"])
def test_adjoint_shipped_sets(name):
    """Shipped sets with radii 7, 15 and 29, weighted (the dx pass's 4-pixel
    lanes on 64 x 64 tiles) and in ring form (a square set), against the FP64
    adjoint."""
    k = kernels(name)
    e_dx, e_dk = adjoint_error(k, noisy_synth(150, 140, seed=7))
    print(f"{name} vs FP64: dx rel {e_dx:.2e}, dk rel {e_dk:.2e}")
    assert e_dx < 1e-4 and e_dk < 1e-4


@pytest.mark.parametrize("D,h,m,window", [(16, [1, 2, 4, 8, 16, 24, 32], 2, "gaussian"), (4, [1, 3, 6, 11], 1, "square")])
def test_adjoint_generated_lpa_sets(D, h, m, window):
    """Generated LPA sets of order m (negative taps, no ring form) against the
    FP64 adjoint."""
    g = lpa_ici.make_kernels(D, h, m, window)
    e_dx, e_dk = adjoint_error(g, noisy_synth(120, 100, seed=D + m))
    print(f"generated D={D} m={m} {window} vs FP64: dx rel {e_dx:.2e}, dk rel {e_dk:.2e}")
    assert e_dx < 1e-4 and e_dk < 1e-4


@pytest.mark.parametrize("w,h", [(70, 45), (131, 9)])
def test_adjoint_ragged_width(ks, w, h):
    """Widths that are not a multiple of 8: the first pass reads the scale
    maps byte by byte, and the last tile of every row is partial."""
    e_dx, e_dk = adjoint_error(ks, noisy_synth(w, h, seed=w + h))
    print(f"{w}x{h} vs FP64: dx rel {e_dx:.2e}, dk rel {e_dk:.2e}")
    assert e_dx < 1e-4 and e_dk < 1e-4


@pytest.mark.parametrize("w,h", [(65535, 20), (20, 65535)])
def test_adjoint_widest_and_tallest_images(ks, w, h):
    """The largest width and height the boundary accepts: dL/dx at a pixel
    depends on the upstream gradient and the scale maps within 31 px, so on
    windows at the start, middle and end of the long side it equals the FP64
    adjoint of a crop with that margin; one call over this image and a small
    one gives this image's dL/dx bit for bit and the sum of the two dL/dk."""
    R = 31
    x = noisy_synth(w, h, seed=7)
    xd, y, m = forward_gpu(ks, x)
    g = torch.randn_like(xd)
    dxs, dk = lpa_ici.backward(ks, [xd], [y], [m], [g], SIGMA, GAMMA)
    torch.cuda.synchronize()
    dx, gn, mn = dxs[0].cpu().numpy(), g.cpu().numpy(), m.cpu().numpy()
    scale = np.abs(dx).max()
    long_side = max(w, h)
    worst = 0.0
    for a in (0, long_side // 2 - 24, long_side - 48):
        lo, hi = max(0, a - R), min(long_side, a + 48 + R)
        if w > h:
            crop, win, inner = (slice(None), slice(lo, hi)), (slice(None), slice(a, a + 48)), \
                (slice(None), slice(a - lo, a - lo + 48))
        else:
            crop, win, inner = (slice(lo, hi), slice(None)), (slice(a, a + 48), slice(None)), \
                (slice(a - lo, a - lo + 48), slice(None))
        dx64, _ = oracle.adjoint_f64(np.ascontiguousarray(x[crop]), np.ascontiguousarray(gn[crop]), ks, SIGMA,
                                     np.ascontiguousarray(mn[(slice(None),) + crop]))
        worst = max(worst, float(np.abs(dx[win] - dx64[inner]).max() / scale))
    print(f"{w}x{h}: dL/dx vs FP64 adjoint on windows, max rel {worst:.2e}")
    assert worst < 1e-4
    xs2, y2, m2 = forward_gpu(ks, noisy_synth(37, 29, seed=8))
    g2 = torch.randn_like(xs2)
    dxb, dkb = lpa_ici.backward(ks, [xd, xs2], [y, y2], [m, m2], [g, g2], SIGMA, GAMMA)
    _, dk2 = lpa_ici.backward(ks, [xs2], [y2], [m2], [g2], SIGMA, GAMMA)
    torch.cuda.synchronize()
    assert torch.equal(dxb[0], dxs[0])
    assert torch.allclose(dkb, dk + dk2, rtol=1e-4, atol=1e-4 * dkb.abs().max().item())


@pytest.mark.parametrize("w,h", [(72, 56), (200, 120)])
def test_adjoint_with_kept_estimates(ks, w, h):
    """The forward's kept estimates in place of the recomputation: dL/dx
    bit-identical, dL/dkernel equal up to float reassociation of the
    per-block partial rows."""
    x = noisy_synth(w, h, seed=w)
    if w == 200:
        x[4:9, 6:12] = 4.0  # windows outside the fixed-point range too
    xd = torch.from_numpy(x).cuda()
    y, m, e = lpa_ici.denoise(ks, xd, SIGMA, GAMMA, scale_maps=True, keep_estimates=True)
    g = torch.randn_like(xd)
    dx0, dk0 = lpa_ici.backward(ks, [xd], [y], [m], [g], SIGMA, GAMMA)
    dx1, dk1 = lpa_ici.backward(ks, [xd], [y], [m], [g], SIGMA, GAMMA, estimates=[e])
    torch.cuda.synchronize()
    assert torch.equal(dx0[0], dx1[0])
    assert torch.allclose(dk0, dk1, rtol=1e-5, atol=1e-6 * dk0.abs().max().item())


def test_adjoint_fixed_point_fallback(ks):
    """Values outside the fixed-point range in one corner: the windows within
    reach of them recompute yhat with float ring sums, as the forward does,
    the others from the prefix tables."""
    x = noisy_synth(200, 120, seed=34)
    x[4:9, 6:12] = 4.0
    assert np.abs(x).max() > ks.info()["ring_range"]
    e_dx, e_dk = adjoint_error(ks, x)
    print(f"fixed-point fallback vs FP64: dx rel {e_dx:.2e}, dk rel {e_dk:.2e}")
    assert e_dx < 1e-4 and e_dk < 1e-4


def test_adjoint_central_differences_image(ks):
    x = noisy_synth(48, 40, seed=5)
    xd, y, m = forward_gpu(ks, x)
    maps = m.cpu().numpy()
    g = torch.randn_like(xd)
    gh = g.cpu().numpy()
    dx = lpa_ici.backward(ks, [xd], [y], [m], [g], SIGMA, GAMMA)[0][0].cpu().numpy()
    var = ks.variances(SIGMA)
    rng = np.random.default_rng(0)
    checked = 0
    for _ in range(64):
        r, c = int(rng.integers(0, 40)), int(rng.integers(0, 48))
        xp, xm = x.copy(), x.copy()
        xp[r, c] += EPS
        xm[r, c] -= EPS
        # skip perturbations that flip a selection of the FP32 walk
        if not (np.array_equal(oracle.walk_f32(xp, ks, var, GAMMA)[1], maps) and
                np.array_equal(oracle.walk_f32(xm, ks, var, GAMMA)[1], maps)):
            continue
        # the float32 inputs are not exactly x +- EPS: divide by the true step
        step = float(np.float64(xp[r, c]) - np.float64(xm[r, c]))
        fd = (loss64(xp, gh, ks, maps) - loss64(xm, gh, ks, maps)) / step
        assert abs(dx[r, c] - fd) <= TOL_FD * max(abs(fd), 1e-2 * np.abs(dx).max()), (r, c, dx[r, c], fd)
        checked += 1
    assert checked >= 32


def test_adjoint_central_differences_taps(ks):
    x = noisy_synth(40, 40, seed=6)
    xd, y, m = forward_gpu(ks, x)
    maps = m.cpu().numpy()
    g = torch.randn(xd.shape, generator=torch.Generator(device="cuda").manual_seed(7), device="cuda")
    gh = g.cpu().numpy()
    dk = lpa_ici.backward(ks, [xd], [y], [m], [g], SIGMA, GAMMA)[1].cpu().numpy()
    scale = np.abs(dk).max()
    r = ks.runs()
    used = set()
    for d in range(ks.D):
        for s in np.unique(maps[d]):
            used.add(d * ks.S + int(s))
    rng = np.random.default_rng(1)
    checked = 0
    for i in rng.permutation(len(r["weights"])):
        k, di = dense_index(ks, r, int(i))
        if k not in used or ks.weights[di] == 0:
            continue
        mods = {h: Modified(ks, di, h * EPS) for h in (-2, -1, 1, 2)}
        if not all(np.array_equal(oracle.walk_f32(x, mk, mk.variances(SIGMA), GAMMA)[1], maps)
                   for mk in mods.values()):
            continue
        # fourth-order central difference through the four (float-rounded) abscissae
        xs = np.array([np.float64(mods[h].weights[di]) - np.float64(ks.weights[di]) for h in (-2, -1, 1, 2)])
        ys = np.array([loss64(x, gh, mods[h], maps) for h in (-2, -1, 1, 2)])
        fd = float(np.polyfit(xs, ys, 3)[2])
        assert abs(dk[i] - fd) <= TOL_FD * max(abs(fd), 1e-2 * scale), (int(i), k, dk[i], fd)
        checked += 1
        if checked == 64:
            break
    assert checked >= 16


def test_adjoint_batched_uneven(ks):
    """One call over uneven images equals per-image calls (dx bit-identical,
    dkernel the sum up to float reassociation)."""
    shapes = [(70, 33), (128, 96), (31, 150)]
    sig = [10 / 255, 30 / 255, 20 / 255]
    xs, ys, ms, gs = [], [], [], []
    for (w, h), s in zip(shapes, sig):
        xd, y, m = forward_gpu(ks, noisy_synth(w, h, seed=w, sigma=s), s)
        xs.append(xd), ys.append(y), ms.append(m), gs.append(torch.randn_like(xd))
    dxs, dk = lpa_ici.backward(ks, xs, ys, ms, gs, sig, GAMMA)
    dk_sum = torch.zeros_like(dk)
    for i in range(3):
        dx1, _ = lpa_ici.backward(ks, [xs[i]], [ys[i]], [ms[i]], [gs[i]], sig[i], GAMMA, dkernel=dk_sum)
        assert torch.equal(dx1[0], dxs[i])
    torch.cuda.synchronize()
    assert torch.allclose(dk, dk_sum, rtol=1e-4, atol=1e-4 * dk.abs().max().item())


@pytest.mark.parametrize("keep", [False, True])
def test_autograd_harness(ks, keep):
    """The autograd function equals the C ABI calls: y bit for bit, dL/dx
    bit for bit, dL/dtaps up to float reassociation; with the uint8 maps
    alone (the default) and with the kept estimates."""
    x = noisy_synth(40, 36, seed=9)
    xd = torch.from_numpy(x).cuda().requires_grad_(True)
    taps = template_taps(ks).requires_grad_(True)
    y = lpa_ici_fn(xd, taps, ks, SIGMA, GAMMA, keep_estimates=keep)
    y_ref = lpa_ici.denoise(ks, torch.from_numpy(x).cuda(), SIGMA, GAMMA)
    assert torch.equal(y.detach(), y_ref)
    g = torch.randn_like(y)
    (y * g).sum().backward()
    _, m = lpa_ici.denoise(ks, xd.detach(), SIGMA, GAMMA, scale_maps=True)
    dxs, dk = lpa_ici.backward(ks, [xd.detach()], [y_ref], [m], [g], SIGMA, GAMMA)
    assert torch.equal(xd.grad, dxs[0])
    assert torch.allclose(taps.grad, dk, rtol=1e-5, atol=1e-6 * dk.abs().max().item())


def test_autograd_pitched_input_and_set_reuse(ks):
    """A column slice of a wider buffer goes through the autograd function
    (rows pitched, unit stride along a row); the kernel set is built once
    while the taps keep their values, and rebuilt when they change."""
    from lpa_ici.autograd import kernel_set_for
    x = noisy_synth(70, 50, seed=19)
    buf = torch.full((50, 96), float("nan"), device="cuda")
    buf[:, :70] = torch.from_numpy(x).cuda()
    xv = buf[:, :70].requires_grad_(True)
    taps = template_taps(ks).requires_grad_(True)
    y = lpa_ici_fn(xv, taps, ks, SIGMA, GAMMA)
    g = torch.randn_like(y)
    (y * g).sum().backward()
    xd = torch.from_numpy(x).cuda()
    y0, m0 = lpa_ici.denoise(ks, xd, SIGMA, GAMMA, scale_maps=True)
    dx0, _ = lpa_ici.backward(ks, [xd], [y0], [m0], [g.contiguous()], SIGMA, GAMMA)
    assert torch.equal(y.detach().contiguous(), y0)
    assert torch.equal(xv.grad.contiguous(), dx0[0])
    first = kernel_set_for(ks, taps)
    assert kernel_set_for(ks, taps.detach().clone()) is first
    changed = taps.detach().clone()
    changed[0] += 1e-3
    assert kernel_set_for(ks, changed) is not first


def test_backward_needs_scale_map(ks):
    xd = torch.from_numpy(noisy_synth(16, 16)).cuda()
    y = lpa_ici.denoise(ks, xd)
    with pytest.raises(lpa_ici.LpaError) as e:
        lpa_ici.backward(ks, [xd], [y], [None], [torch.ones_like(y)])
    assert e.value.status == 1 and "scale map" in str(e.value)


def test_adjoint_pitched_rows(ks):
    """Images, upstream gradients and dL/dx with a row pitch wider than the
    image (NaN in the inputs' padding, a sentinel in dL/dx's): dL/dx equals
    the dense run bit for bit and its padding is never written; dL/dkernel
    matches up to float reassociation."""
    import ctypes
    w, h, pitch = 133, 70, 160
    x = noisy_synth(w, h, seed=45)

    def view(a, fill):
        buf = torch.full((h, pitch), fill, dtype=torch.float32, device="cuda")
        if a is not None:
            buf[:, :w] = a
        return buf, buf[:, :w]

    _, xv = view(torch.from_numpy(x).cuda(), float("nan"))
    yv, m = lpa_ici.denoise(ks, xv, SIGMA, GAMMA, scale_maps=True)
    g = torch.randn((h, w), generator=torch.Generator(device="cuda").manual_seed(9), device="cuda")
    _, gv = view(g, float("nan"))
    dx_buf, dxv = view(None, 7.0)
    dk = torch.zeros(ks.info()["num_weights"], dtype=torch.float32, device="cuda")
    arr = (lpa_ici.LpaImage * 1)(lpa_ici.image_struct(xv, yv, m, SIGMA, GAMMA))
    lpa_ici.check(lpa_ici.lib().lpa_ici_backward(ks.handle, arr, (ctypes.c_void_p * 1)(gv.data_ptr()),
                                                 (ctypes.c_void_p * 1)(dxv.data_ptr()),
                                                 ctypes.c_void_p(dk.data_ptr()), 1,
                                                 ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    xd = torch.from_numpy(x).cuda()
    y0, m0 = lpa_ici.denoise(ks, xd, SIGMA, GAMMA, scale_maps=True)
    dx0, dk0 = lpa_ici.backward(ks, [xd], [y0], [m0], [g.contiguous()], SIGMA, GAMMA)
    torch.cuda.synchronize()
    assert torch.equal(m, m0)
    assert torch.equal(dxv.contiguous().view(torch.int32), dx0[0].view(torch.int32))
    assert bool((dx_buf[:, w:] == 7.0).all())
    assert torch.allclose(dk, dk0, rtol=1e-5, atol=1e-6 * dk0.abs().max().item())
    # the Python wrapper allocates dL/dx at the image's pitch
    dxs, _ = lpa_ici.backward(ks, [xv], [yv], [m], [gv], SIGMA, GAMMA)
    torch.cuda.synchronize()
    assert dxs[0].stride() == xv.stride()
    assert torch.equal(dxs[0].contiguous().view(torch.int32), dx0[0].view(torch.int32))
