"""Parity of the HIP forward with the oracle (CPU restatement of the
reference's v0_0 path) and with fixtures made by the reference itself.

Bars (the reference's algorithm in FP32 on the same seeded noisy image):
  * exact mode (separately rounded mul/add, the reference's tap order):
    bit-identical output and scale maps;
  * fast modes (fixed-point ring sums, the default for ring-form sets; float
    ring sums; weighted FMA): per-direction scale agreement >= 99.9 %, every
    disagreement at an interval margin below 1e-6, |dy| <= 1e-5 wherever all
    directions agree, |dPSNR| <= 0.005 dB; the error against the FP64 oracle
    is printed for every case;
  * the fixed-point walk is also bit-identical to the oracle's restatement of
    it (integer sums do not depend on the order of the taps).
"""
import os

import numpy as np
import pytest
import torch

import lpa_ici
import oracle
from conftest import (GAMMA, GAUSSIAN_M20, SEED, SIGMA, SYMMETRIC, golden, kernels, noisy_synth)

pytestmark = pytest.mark.gpu

TOL_Y = 1e-5
TOL_MARGIN = 1e-6
MIN_AGREE = 0.999
TOL_PSNR = 0.005


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def run(ks, x, flags=0, maps=True, sigma=SIGMA, gamma=GAMMA, variances=None):
    out = lpa_ici.denoise(ks, dev(x), sigmas=sigma, gammas=gamma, scale_maps=maps, flags=flags, variances=variances)
    torch.cuda.synchronize()
    if maps:
        return out[0].cpu().numpy(), out[1].cpu().numpy()
    return out.cpu().numpy()


def check_fast_parity(ks, x, clean, y, maps, label, sigma=SIGMA, gamma=GAMMA):
    var = ks.variances(sigma)
    yo, mo, marg = oracle.walk_f32(x, ks, var, gamma)
    agree = maps == mo
    per_dir = agree.reshape(ks.D, -1).mean(axis=1)
    assert per_dir.min() >= MIN_AGREE, f"{label}: scale agreement {per_dir.min():.6f}"
    bad = ~agree
    assert np.all(marg[bad] < TOL_MARGIN), f"{label}: disagreement at margin {marg[bad].max()}"
    all_agree = agree.all(axis=0)
    dy = np.abs(y.astype(np.float64) - yo)
    assert dy[all_agree].max(initial=0.0) <= TOL_Y, f"{label}: max |dy| {dy[all_agree].max()}"
    y64, _ = oracle.walk_f64(x, ks, sigma, gamma)
    e64 = np.abs(y - y64).max()
    line = f"{label}: agree={per_dir.min():.6f} flips={int(bad.sum())} max|dy|={dy[all_agree].max():.2e} " \
           f"max|y-y64|={e64:.2e}"
    if clean is not None:
        dp = abs(lpa_ici.psnr(clean, y) - lpa_ici.psnr(clean, yo))
        assert dp <= TOL_PSNR, f"{label}: dPSNR {dp}"
        line += f" dPSNR={dp:.2e}"
    print(line)


@pytest.mark.parametrize("tag,name", [("symmetric", SYMMETRIC), ("m20_gaussian", GAUSSIAN_M20),
                                      ("d4_square", "lpa_kernels_m_1_0_d_4_square.txt")])
def test_exact_matches_reference_fixture(tag, name):
    """Exact mode reproduces the reference's v0_0 output bit for bit."""
    ks = kernels(name)
    x = golden("synth64_noisy.npy")
    y = run(ks, x, flags=lpa_ici.FWD_EXACT, maps=False)
    ref = golden(f"synth64_v0_0_{tag}.npy")
    assert np.array_equal(y.view(np.uint32), ref.view(np.uint32))


def test_exact_cameraman_bitexact(ks):
    x = golden("cameraman_noisy.npy")
    y, maps = run(ks, x, flags=lpa_ici.FWD_EXACT)
    assert np.array_equal(y.view(np.uint32), golden("cameraman_v0_0.npy").view(np.uint32))
    _, mo, _ = oracle.walk_f32(x, ks, ks.variances(SIGMA), GAMMA)
    assert np.array_equal(maps, mo)


WALKS = [(0, "fixed-point ring sums"), (lpa_ici.FWD_RING_FLOAT, "float ring sums"),
         (lpa_ici.FWD_WEIGHTED, "weighted FMA")]


def fixed_point(ks):
    info = ks.info()
    return info["ring_frac_bits"], info["ring_range"]


@pytest.mark.parametrize("flags,label", WALKS)
def test_fast_parity_config1_512(ks, flags, label):
    """Config 1: one 512x512 synthetic image, sigma 20/255, gamma 2; the
    default set has the ring form, so both fast walks are held to the bar."""
    assert ks.info()["ring_form"] == 1
    clean = lpa_ici.synth_image(512, 512, 1)
    x = lpa_ici.make_noisy(clean, SIGMA, SEED)
    y, maps = run(ks, x, flags=flags)
    check_fast_parity(ks, x, clean, y, maps, f"config1 512^2 symmetric, {label}")


def test_fast_parity_gaussian_m20(ks):
    ksg = kernels(GAUSSIAN_M20)
    clean = lpa_ici.synth_image(256, 192, 3)
    x = lpa_ici.make_noisy(clean, SIGMA, SEED)
    y, maps = run(ksg, x)
    check_fast_parity(ksg, x, clean, y, maps, "256x192 m20 gaussian")


@pytest.mark.parametrize("flags,label", WALKS)
def test_fast_parity_barbara(ks, flags, label):
    """The reference's published case: PSNR 25.0736 on barbara."""
    clean = lpa_ici.load_image(lpa_ici.KERNEL_DIR + "/../../tests/golden/images/barbara.png")
    x = lpa_ici.make_noisy(clean, SIGMA, SEED)
    y, maps = run(ks, x, flags=flags)
    check_fast_parity(ks, x, clean, y, maps, f"barbara, {label}")
    assert round(lpa_ici.psnr(clean, y), 4) == 25.0736


@pytest.mark.parametrize("w,h", [(512, 512), (1, 1), (9, 1), (33, 65), (129, 47), (300, 200)])
def test_fixed_point_bitexact(ks, w, h):
    """The default walk of a ring-form set equals the oracle's fixed-point
    restatement bit for bit: output and scale maps."""
    F, rng = fixed_point(ks)
    clean = lpa_ici.synth_image(w, h, 1)
    x = lpa_ici.make_noisy(clean, SIGMA, SEED)
    y, maps = run(ks, x)
    yo, mo = oracle.walk_q(x, ks, ks.variances(SIGMA), GAMMA, F, rng)
    assert np.array_equal(maps, mo)
    assert np.array_equal(y.view(np.uint32), yo.view(np.uint32))


@pytest.mark.parametrize("D,h", [(8, [1, 2, 4, 8, 16]), (32, [1, 3, 6, 12, 24, 32])])
def test_generated_wedge_sets(D, h):
    """Generated uniform-wedge sets with other direction counts and scales:
    exact mode equals the oracle's v0_0 bits, the default walk (fixed-point
    ring sums) the oracle's fixed-point restatement, output and maps."""
    g = lpa_ici.wedge_kernels(D, h)
    assert g.info()["ring_form"] == 1
    x = noisy_synth(160, 96, seed=D)
    var = g.variances(SIGMA)
    y, maps = run(g, x, flags=lpa_ici.FWD_EXACT)
    yo, mo, _ = oracle.walk_f32(x, g, var, GAMMA)
    assert np.array_equal(maps, mo)
    assert np.array_equal(y.view(np.uint32), yo.view(np.uint32))
    F, rng = fixed_point(g)
    y, maps = run(g, x)
    yq, mq = oracle.walk_q(x, g, var, GAMMA, F, rng)
    assert np.array_equal(maps, mq)
    assert np.array_equal(y.view(np.uint32), yq.view(np.uint32))


@pytest.mark.parametrize("D,h,m,window", [(16, [1, 2, 4, 8, 16, 24, 32], 2, "gaussian"), (8, [1, 3, 7, 15], 1, "uniform"),
                                         (4, [1, 2, 5, 12], 2, "square")])
def test_generated_lpa_sets(D, h, m, window):
    """Generated LPA sets of order m (weighted, no ring form): exact mode
    equals the oracle's v0_0 bits, the default weighted walk meets the bar."""
    g = lpa_ici.make_kernels(D, h, m, window)
    assert g.info()["ring_form"] == 0
    clean = lpa_ici.synth_image(160, 96, 5)
    x = lpa_ici.make_noisy(clean, SIGMA, SEED)
    y, maps = run(g, x, flags=lpa_ici.FWD_EXACT)
    yo, mo, _ = oracle.walk_f32(x, g, g.variances(SIGMA), GAMMA)
    assert np.array_equal(maps, mo)
    assert np.array_equal(y.view(np.uint32), yo.view(np.uint32))
    y, maps = run(g, x)
    check_fast_parity(g, x, clean, y, maps, f"generated D={D} m={m} {window}")


@pytest.mark.parametrize("flags,label", [(0, "fixed point"), (lpa_ici.FWD_EXACT, "exact"),
                                         (lpa_ici.FWD_RING_FLOAT, "float ring"), (lpa_ici.FWD_WEIGHTED, "weighted")])
def test_forward_keeps_estimates(ks, flags, label):
    """The estimates a forward keeps are the values it aggregated: equal, bit
    for bit, to lpa_ici_estimates at its scale maps; y and the maps do not
    change when they are kept."""
    xd = dev(noisy_synth(150, 90, seed=12))
    y, m, e = lpa_ici.denoise(ks, xd, SIGMA, GAMMA, scale_maps=True, flags=flags, keep_estimates=True)
    y0, m0 = lpa_ici.denoise(ks, xd, SIGMA, GAMMA, scale_maps=True, flags=flags)
    e_ref = lpa_ici.estimates(ks, xd, m, flags=flags)
    torch.cuda.synchronize()
    assert torch.equal(y, y0) and torch.equal(m, m0), label
    assert torch.equal(e, e_ref), label


def test_fixed_point_fallback_windows(ks):
    """A block of values outside the fixed-point range: windows that reach it
    take float ring sums, the rest stay in fixed point.  Far from the block
    the output equals the fixed-point restatement bit for bit; everywhere it
    meets the fast-mode bar against the reference-order FP32 walk."""
    F, rng = fixed_point(ks)
    x = noisy_synth(512, 384, seed=17)
    x[40:50, 40:50] = 3.0 * rng
    y, maps = run(ks, x)
    check_fast_parity(ks, x, None, y, maps, "512x384 with an out-of-range block")
    yo, mo = oracle.walk_q(x, ks, ks.variances(SIGMA), GAMMA, F, np.inf)
    # a window (tile plus halo) that holds the block is at most the largest
    # tile side plus the halo away from it
    reach = max(max(c["tile_w"], c["tile_h"]) for c in lpa_ici.tile_configs()) + 2 * 31
    far = np.ones(x.shape, bool)
    far[:50 + reach, :50 + reach] = False
    assert far.sum() > 100000
    assert np.array_equal(y.view(np.uint32)[far], yo.view(np.uint32)[far])
    assert np.array_equal(maps[:, far], mo[:, far])
    # every window out of range: float ring sums everywhere
    xs = x[:128, :128] + np.float32(3.0)
    y2, m2 = run(ks, xs)
    check_fast_parity(ks, xs, None, y2, m2, "offset +3 (float ring sums)")
    y3, m3 = run(ks, xs, flags=lpa_ici.FWD_RING_FLOAT)
    assert np.array_equal(y2, y3) and np.array_equal(m2, m3)


def test_non_finite_pixels(ks):
    """Missing pixels stored as NaN, and infinities: exact mode keeps the bits
    of the oracle's walk over the nonzero taps (the reference's sparse GPU
    semantics, src/v4_1_lpa_ici_2D_sparse_shared.cu:826-876 and 1044-1091: a
    non-finite pixel reaches the estimates whose supports hold it; the dense
    CPU loop of v0_0 would also multiply it by the zero taps of each kernel's
    box), the NaN pattern of y included, and its scale maps; the default walk
    takes float ring sums in every window that holds a non-finite value,
    gives the same NaN pattern, and away from those windows equals the
    fixed-point restatement bit for bit."""
    x = noisy_synth(512, 384, seed=23)
    bad = [(30, 40), (31, 41), (32, 41), (60, 70), (50, 20)]
    for (r, c), v in zip(bad, [np.nan, np.nan, np.nan, np.inf, -np.inf]):
        x[r, c] = v
    ye, me = run(ks, x, flags=lpa_ici.FWD_EXACT)
    yf, mf, _ = oracle.walk_f32(x, ks, ks.variances(SIGMA), GAMMA)
    assert np.array_equal(me, mf)
    assert np.array_equal(np.isnan(ye), np.isnan(yf))
    fin = ~np.isnan(yf)
    assert np.array_equal(ye.view(np.uint32)[fin], yf.view(np.uint32)[fin])
    assert np.isnan(ye).any()
    y, maps = run(ks, x)
    assert np.array_equal(np.isnan(y), np.isnan(yf))
    F, rng = fixed_point(ks)
    xq = np.nan_to_num(x, nan=0.0, posinf=0.0, neginf=0.0)
    yq, mq = oracle.walk_q(xq, ks, ks.variances(SIGMA), GAMMA, F, rng)
    # a window (tile plus halo) that holds a non-finite pixel lies within the
    # largest tile side plus the halo of it
    reach = max(max(c["tile_w"], c["tile_h"]) for c in lpa_ici.tile_configs()) + 2 * 31
    far = np.ones(x.shape, bool)
    far[:70 + reach, :70 + reach] = False
    assert far.sum() > 100000
    assert np.array_equal(y.view(np.uint32)[far], yq.view(np.uint32)[far])
    assert np.array_equal(maps[:, far], mq[:, far])


@pytest.mark.parametrize("shape", range(len(lpa_ici.tile_configs())))
def test_tile_shapes_identical(ks, shape):
    """Every tile shape adds each pixel's taps in the same order: same bits."""
    x = noisy_synth(200, 136, seed=5)
    walks = [lpa_ici.FWD_EXACT, 0, lpa_ici.FWD_WEIGHTED, lpa_ici.FWD_RING_FLOAT]
    base = [run(ks, x, flags=f) for f in walks]
    ks.set_tile_config(shape)
    try:
        got = [run(ks, x, flags=f) for f in walks]
    finally:
        ks.set_tile_config(-1)
    for f, (yb, mb), (y, m) in zip(walks, base, got):
        assert np.array_equal(y.view(np.uint32), yb.view(np.uint32)), f
        assert np.array_equal(m, mb), f


@pytest.mark.parametrize("shape", range(len(lpa_ici.tile_configs())))
def test_tile_shapes_identical_weighted_set(shape):
    """The same for a set with no ring form (the m=(2,0) gaussian set): the
    weighted walk, exact and fast, gives the same bits on every tile shape."""
    ksg = lpa_ici.load_kernels(lpa_ici.GAUSSIAN_M20_KERNELS)
    x = noisy_synth(136, 200, seed=6)
    walks = [lpa_ici.FWD_EXACT, 0]
    base = [run(ksg, x, flags=f) for f in walks]
    ksg.set_tile_config(shape)
    got = [run(ksg, x, flags=f) for f in walks]
    for f, (yb, mb), (y, m) in zip(walks, base, got):
        assert np.array_equal(y.view(np.uint32), yb.view(np.uint32)), f
        assert np.array_equal(m, mb), f
    ksg.close()


@pytest.mark.parametrize("w,h", [(1, 1), (1, 9), (9, 1), (3, 3), (33, 65), (100, 3), (31, 31), (129, 47)])
def test_small_and_ragged_exact(ks, w, h):
    x = noisy_synth(w, h, seed=w * 100 + h)
    y, maps = run(ks, x, flags=lpa_ici.FWD_EXACT)
    yo, mo, _ = oracle.walk_f32(x, ks, ks.variances(SIGMA), GAMMA)
    assert np.array_equal(y.view(np.uint32), yo.view(np.uint32))
    assert np.array_equal(maps, mo)


@pytest.mark.parametrize("w,h", [(65535, 20), (20, 65535)])
def test_widest_and_tallest_images(ks, w, h):
    """The largest width and height the boundary accepts (65535; one more is
    LPA_ETOOBIG, tests/test_host.py): the tile indices of the plain and the
    batched forms reach their largest values.  The default walk equals the
    oracle's fixed-point restatement bit for bit, exact mode equals the
    reference-order walk, and one persistent launch over this image and a
    small one equals one launch per image."""
    x = noisy_synth(w, h, seed=5)
    F, rng = fixed_point(ks)
    y, maps = run(ks, x)
    yo, mo = oracle.walk_q(x, ks, ks.variances(SIGMA), GAMMA, F, rng, threads=8)
    assert np.array_equal(maps, mo)
    assert np.array_equal(y.view(np.uint32), yo.view(np.uint32))
    ye, me = run(ks, x, flags=lpa_ici.FWD_EXACT)
    yf, mf, _ = oracle.walk_f32(x, ks, ks.variances(SIGMA), GAMMA, threads=8)
    assert np.array_equal(me, mf)
    assert np.array_equal(ye.view(np.uint32), yf.view(np.uint32))
    xs = [dev(x), dev(noisy_synth(37, 29, seed=6))]
    ya, ma = lpa_ici.denoise(ks, xs, SIGMA, GAMMA, scale_maps=True)
    yb, mb = lpa_ici.denoise(ks, xs, SIGMA, GAMMA, scale_maps=True, flags=lpa_ici.FWD_PLAIN)
    torch.cuda.synchronize()
    for a, b, c, d in zip(ya, yb, ma, mb):
        assert torch.equal(a, b) and torch.equal(c, d)
    assert np.array_equal(ya[0].cpu().numpy().view(np.uint32), yo.view(np.uint32))


def test_pitched_image(ks):
    """x and y with a row stride larger than the width."""
    x = noisy_synth(70, 50, seed=11)
    big = torch.zeros((50, 96), dtype=torch.float32, device="cuda")
    big[:, :70] = dev(x)
    xv = big[:, :70]
    yb = torch.full((50, 96), -7.0, dtype=torch.float32, device="cuda")
    yv = yb[:, :70]
    img = lpa_ici.image_struct(xv, yv, None, SIGMA, GAMMA)
    import ctypes
    arr = (lpa_ici.LpaImage * 1)(img)
    lpa_ici.check(lpa_ici.lib().lpa_ici_forward(kernels().handle, arr, 1, lpa_ici.FWD_EXACT,
                                                ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    torch.cuda.synchronize()
    yo, _, _ = oracle.walk_f32(x, kernels(), kernels().variances(SIGMA), GAMMA)
    out = yb.cpu().numpy()
    assert np.array_equal(out[:, :70], yo)
    assert np.all(out[:, 70:] == -7.0), "pixels past the width were written"


def test_explicit_variances(ks):
    """The reference's interface passes variances in: same result as sigma."""
    x = noisy_synth(64, 40, seed=2)
    a = run(ks, x, flags=lpa_ici.FWD_EXACT, maps=False)
    b = run(ks, x, flags=lpa_ici.FWD_EXACT, maps=False, variances=ks.variances(SIGMA), sigma=0.0)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("name", ["lpa_kernels_python_16_tapered.txt", "lpa_kernels_m_1_0_d_8_h12_1_15_15_30.txt",
                                  "lpa_kernels_m_1_0_d_16_h12_1_2_4_8_16.txt"])
def test_other_kernel_sets_exact(name):
    """Sets beyond the reference's constant-memory budget are accepted."""
    ksx = kernels(name)
    x = noisy_synth(96, 80, seed=4)
    y, maps = run(ksx, x, flags=lpa_ici.FWD_EXACT)
    yo, mo, _ = oracle.walk_f32(x, ksx, ksx.variances(SIGMA), GAMMA)
    assert np.array_equal(y.view(np.uint32), yo.view(np.uint32))
    assert np.array_equal(maps, mo)


def test_kernel_without_taps():
    """A kernel whose weights are all zero has no runs; its estimate is 0, as
    the reference's sum over no taps.  Its variance would be 0, so explicit
    variances stand in.  Exact mode equals the oracle bit for bit; the default
    weighted walk agrees with it at the bar."""
    base = kernels(GAUSSIAN_M20)
    w = base.weights.copy()
    empty = (3 * base.S + 2, 9 * base.S + 5)
    for k in empty:
        w[base.offsets[k]:base.offsets[k] + base.rows[k] * base.cols[k]] = 0.0
    ksz = lpa_ici.KernelSet(base.D, base.S, w, base.offsets, base.rows, base.cols)
    assert all(ksz.runs()["kernel_num_runs"][k] == 0 for k in empty)
    var = ksz.variances(SIGMA)
    var[list(empty)] = np.float32(1e-3)
    x = noisy_synth(128, 96, seed=9)
    y, maps = run(ksz, x, flags=lpa_ici.FWD_EXACT, variances=var, sigma=0.0)
    yo, mo, marg = oracle.walk_f32(x, ksz, var, GAMMA)
    assert np.array_equal(y.view(np.uint32), yo.view(np.uint32))
    assert np.array_equal(maps, mo)
    yf, mf = run(ksz, x, variances=var, sigma=0.0)
    agree = mf == mo
    assert agree.reshape(ksz.D, -1).mean(axis=1).min() >= MIN_AGREE
    assert np.all(marg[~agree] < TOL_MARGIN)
    all_agree = agree.all(axis=0)
    assert np.abs(yf.astype(np.float64) - yo)[all_agree].max() <= TOL_Y


SHIPPED = sorted(f[:-len(".gz")] for f in os.listdir(lpa_ici.KERNEL_DIR) if f.endswith(".txt.gz"))


@pytest.mark.parametrize("name", SHIPPED)
def test_every_shipped_set_both_modes(name):
    """Every kernel set the reference ships (its kernels/ folder): the exact
    mode bit-identical to the oracle's v0_0 walk (output and scale maps), and
    the default walk (fixed-point ring sums for ring-form sets, FMA for the
    others) at the bar."""
    ksx = kernels(name)
    x = noisy_synth(128, 96, seed=11)
    clean = lpa_ici.synth_image(128, 96, 11)
    y, maps = run(ksx, x, flags=lpa_ici.FWD_EXACT)
    yo, mo, _ = oracle.walk_f32(x, ksx, ksx.variances(SIGMA), GAMMA)
    assert np.array_equal(y.view(np.uint32), yo.view(np.uint32))
    assert np.array_equal(maps, mo)
    walk = "fixed point" if ksx.info()["ring_form"] else "weighted FMA"
    yf, mf = run(ksx, x)
    check_fast_parity(ksx, x, clean, yf, mf, f"{name} ({walk})")


def test_batched_equals_plain(ks):
    """One persistent launch over uneven images equals one launch per image."""
    rng = np.random.default_rng(7)
    shapes = [(int(rng.integers(8, 300)), int(rng.integers(8, 300))) for _ in range(12)]
    sig = [float(rng.choice([10, 20, 30, 40])) / 255.0 for _ in shapes]
    gam = [float(rng.choice([1.5, 2.0, 2.5])) for _ in shapes]
    xs = [dev(noisy_synth(w, h, seed=i, sigma=s)) for i, ((w, h), s) in enumerate(zip(shapes, sig))]
    ya, ma = lpa_ici.denoise(ks, xs, sig, gam, scale_maps=True)
    yb, mb = lpa_ici.denoise(ks, xs, sig, gam, scale_maps=True, flags=lpa_ici.FWD_PLAIN)
    torch.cuda.synchronize()
    for a, b, c, d in zip(ya, yb, ma, mb):
        assert torch.equal(a, b) and torch.equal(c, d)
    # each image also matches the oracle at its own sigma and gamma
    for x, y, m, s, g in zip(xs[:4], ya[:4], ma[:4], sig[:4], gam[:4]):
        check_fast_parity(ks, x.cpu().numpy(), None, y.cpu().numpy(), m.cpu().numpy(), "batched", sigma=s, gamma=g)


def test_graph_replay_equals_eager(ks):
    xs = [dev(noisy_synth(w, h, seed=w)) for w, h in [(130, 70), (64, 200), (257, 33)]]
    ys_e = lpa_ici.denoise(ks, xs)
    ys = [torch.empty_like(x) for x in xs]
    stream = torch.cuda.Stream()
    plan = lpa_ici.Plan(ks, xs, ys, flags=lpa_ici.FWD_GRAPH)
    for i in range(4):
        for y in ys:
            y.fill_(float("nan"))
        torch.cuda.synchronize()
        with torch.cuda.stream(stream):
            plan.run(stream)
        stream.synchronize()
        for a, b in zip(ys, ys_e):
            assert torch.equal(a, b), f"replay {i}"
        _ = (xs[0] * 2).sum().item()  # unrelated eager work between replays
    plan.close()


def test_autotune_installs_fastest(ks, tmp_path):
    from lpa_ici.autotune import autotune
    xs = [dev(noisy_synth(256, 192, seed=3))]
    best, times = autotune(ks, xs, cache=str(tmp_path / "cache.json"), reps=1)
    try:
        assert ks.tile_config() == best and times[best] == min(times.values())
        assert len(times) == len(lpa_ici.tile_configs())
        again, t2 = autotune(ks, xs, cache=str(tmp_path / "cache.json"))
        assert again == best and t2 == {}  # served from the cache
    finally:
        ks.set_tile_config(-1)


def reaggregate(ks, est, maps, sigma=SIGMA):
    """y from per-direction estimates in the forward's order and rounding:
    num += fl(est * inv), den += inv over d, then num / den (float32)."""
    var = ks.variances(sigma)
    inv = (np.float32(1.0) / var).astype(np.float32)
    num = np.zeros(est.shape[1:], np.float32)
    den = np.zeros(est.shape[1:], np.float32)
    for d in range(ks.D):
        w = inv[d * ks.S + maps[d].astype(np.int64)]
        num = (num + (est[d] * w).astype(np.float32)).astype(np.float32)
        den = (den + w).astype(np.float32)
    return (num / den).astype(np.float32)


@pytest.mark.parametrize("name,flags", [(SYMMETRIC, 0), (SYMMETRIC, lpa_ici.FWD_RING_FLOAT),
                                        (SYMMETRIC, lpa_ici.FWD_WEIGHTED), (SYMMETRIC, lpa_ici.FWD_EXACT),
                                        (GAUSSIAN_M20, 0)])
def test_estimates_are_the_aggregated_values(name, flags):
    """The per-direction estimates at the selected scales, aggregated on the
    host in the forward's order, give the forward's output bit for bit."""
    ksx = kernels(name)
    x = noisy_synth(150, 110, seed=9)
    xd = dev(x)
    y, m = lpa_ici.denoise(ksx, xd, SIGMA, GAMMA, scale_maps=True, flags=flags)
    est = lpa_ici.estimates(ksx, xd, m, flags=flags)
    torch.cuda.synchronize()
    maps = m.cpu().numpy()
    assert np.array_equal(reaggregate(ksx, est.cpu().numpy(), maps).view(np.uint32), y.cpu().numpy().view(np.uint32))


def test_estimates_fixed_point_bitexact(ks):
    """Default walk: equal to the oracle's fixed-point restatement, estimate by
    estimate (the reference's per-direction debugging outputs)."""
    F, rng = fixed_point(ks)
    clean = lpa_ici.synth_image(300, 200, 2)
    x = lpa_ici.make_noisy(clean, SIGMA, SEED)
    xd = dev(x)
    _, m = lpa_ici.denoise(ks, xd, SIGMA, GAMMA, scale_maps=True)
    est = lpa_ici.estimates(ks, [xd], [m])[0].cpu().numpy()
    _, mo, eo = oracle.walk_q(x, ks, ks.variances(SIGMA), GAMMA, F, rng, want_est=True)
    assert np.array_equal(m.cpu().numpy(), mo)
    assert np.array_equal(est.view(np.uint32), eo.view(np.uint32))


@pytest.mark.parametrize("shape", [3, 5, 8, 9])
def test_estimates_follow_the_forward_tile_config(ks, shape):
    """With a block of values outside the fixed-point range, windows that
    reach it take float ring sums and the others fixed point: the estimates
    are taken on the forward's tile configuration, so they equal the values
    the forward kept, bit for bit, whatever the configuration."""
    _, rng = fixed_point(ks)
    x = noisy_synth(300, 200, seed=21)
    x[60:70, 100:110] = 3.0 * rng
    xd = dev(x)
    ks.set_tile_config(shape)
    try:
        _, m, e = lpa_ici.denoise(ks, xd, SIGMA, GAMMA, scale_maps=True, keep_estimates=True)
        e2 = lpa_ici.estimates(ks, xd, m)
        torch.cuda.synchronize()
    finally:
        ks.set_tile_config(-1)
    assert torch.equal(e.view(torch.int32), e2.view(torch.int32))


def test_estimates_need_a_scale_map(ks):
    xd = dev(noisy_synth(16, 16, seed=1))
    import ctypes
    arr = (lpa_ici.LpaImage * 1)(lpa_ici.image_struct(xd, xd, None, SIGMA, GAMMA))
    out = torch.empty((ks.D, 16, 16), dtype=torch.float32, device="cuda")
    est_p = (ctypes.c_void_p * 1)(out.data_ptr())
    st = lpa_ici.lib().lpa_ici_estimates(ks.handle, arr, est_p, 1, 0, None)
    assert st == 1 and b"scale map" in lpa_ici.lib().lpa_last_error()


def test_full_size_windows_bitexact(ks):
    """Config 2's size (one 4096x4096 image): the walk is local (a pixel reads
    at most 31 px away), so the oracle's fixed-point restatement on a crop
    with a 31-px margin must give the GPU's output in the crop's interior bit
    for bit.  Three windows: a corner, an edge and the interior."""
    F, rng = fixed_point(ks)
    clean = lpa_ici.synth_image(4096, 4096, 7)
    x = lpa_ici.make_noisy(clean, SIGMA, SEED)
    y, maps = run(ks, x)
    var = ks.variances(SIGMA)
    R = 31
    for (r0, c0, n) in [(0, 0, 96), (0, 2000, 96), (3000, 1234, 160)]:
        ra, ca = max(0, r0 - R), max(0, c0 - R)
        rb, cb = min(4096, r0 + n + R), min(4096, c0 + n + R)
        yo, mo = oracle.walk_q(np.ascontiguousarray(x[ra:rb, ca:cb]), ks, var, GAMMA, F, rng)
        inner = (slice(r0 - ra, r0 - ra + n), slice(c0 - ca, c0 - ca + n))
        # a crop edge that is not the image edge sees zeros beyond it: only
        # pixels at least R inside the crop's own edges are compared
        assert np.array_equal(y[r0:r0 + n, c0:c0 + n].view(np.uint32), yo[inner].view(np.uint32)), (r0, c0)
        assert np.array_equal(maps[:, r0:r0 + n, c0:c0 + n], mo[(slice(None),) + inner]), (r0, c0)


def pitched(a, pitch, fill=float("nan")):
    """A device view of a (H x W) with row pitch `pitch` floats; the columns
    past W hold `fill`."""
    h, w = a.shape
    buf = torch.full((h, pitch), fill, dtype=torch.float32, device="cuda")
    buf[:, :w] = dev(a)
    return buf[:, :w]


@pytest.mark.parametrize("flags", [0, lpa_ici.FWD_EXACT, lpa_ici.FWD_PLAIN])
def test_forward_pitched_rows(ks, flags):
    """lpa_ici.denoise on rows with a pitch wider than the image (NaN in the
    padding): the output is allocated at the input's pitch, and the result
    is the dense image's, bit for bit, for three walks, so the padding is
    never read."""
    x = noisy_synth(150, 70, seed=44)
    xv = pitched(x, 193)
    y, m = lpa_ici.denoise(ks, xv, SIGMA, GAMMA, scale_maps=True, flags=flags)
    y0, m0 = lpa_ici.denoise(ks, dev(x), SIGMA, GAMMA, scale_maps=True, flags=flags)
    torch.cuda.synchronize()
    assert y.stride() == xv.stride()
    assert torch.equal(m, m0)
    assert torch.equal(y.contiguous().view(torch.int32), y0.view(torch.int32))
