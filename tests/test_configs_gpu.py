"""The benchmark workloads at their stated sizes, on the GPU.

  4096^2 (one image, the symmetric set and the m=(2,0) gaussian set): the
      default walk against the reference-order FP32 walk at the north-star
      bar on windows, and the symmetric set's fixed-point walk bit for bit
      against the oracle's restatement.  The walk is local (a pixel reads at
      most 31 px away), so the oracle run on a crop with a 31-px margin gives
      the full image's values in the crop's interior.
  256 uneven images (sides 128..2048, per-image sigma and gamma): one
      persistent launch, one launch per image and hipGraph replays give the
      same bits; windows of six images against the oracle.
  64 x 512^2 forward + adjoint: dL/dx against the FP64 adjoint on windows, and
      both gradients against FP64 central differences (eps 1e-3) of the
      frozen-selection forward at 64 pixels and 64 taps.
  32 uneven images (sides 1024..8192), forward + adjoint in one call: each
      image's dL/dx equals a single-image call bit for bit, dL/dkernel their
      sum; peak device memory against the 288 GB of HBM.

Bars: scale agreement >= 99.9 % per direction with every disagreement at an
FP32 interval margin below 1e-6, |dy| <= 1e-5 where all directions agree,
|dPSNR| <= 0.005 dB; gradients within 1e-3 relative of the central
differences (components below 1e-3 of the largest held to 1e-6 of the
largest, the accuracy of FP32 sums); the FP64 error is printed.
"""
import numpy as np
import pytest
import torch

import lpa_ici
import oracle
from conftest import GAUSSIAN_M20, SIGMA, GAMMA, dense_index, kernels
from lpa_ici import workloads

pytestmark = pytest.mark.gpu

R = 31
TOL_Y, TOL_MARGIN, MIN_AGREE, TOL_PSNR = 1e-5, 1e-6, 0.999, 0.005
EPS, TOL_FD, FD_FLOOR = 1e-3, 1e-3, 1e-3


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def crop_of(H, W, r0, c0, h, w, margin=R):
    """Crop bounds around the window (r0, c0, h, w) with `margin` px on each
    side (clipped to the image), and the window's slices inside the crop."""
    ra, ca = max(0, r0 - margin), max(0, c0 - margin)
    rb, cb = min(H, r0 + h + margin), min(W, c0 + w + margin)
    return (slice(ra, rb), slice(ca, cb)), (slice(r0 - ra, r0 - ra + h), slice(c0 - ca, c0 - ca + w))


def windows(H, W, n, rng):
    """Four corners-and-edges windows and random ones, each n x n (clipped)."""
    h, w = min(n, H), min(n, W)
    out = [(0, 0), (H - h, W - w), (0, W - w), ((H - h) // 2, 0)]
    out += [(int(rng.integers(0, H - h + 1)), int(rng.integers(0, W - w + 1))) for _ in range(2)]
    return [(r, c, h, w) for r, c in out]


class WindowBar:
    """Accumulates the north-star bar over windows of one or more images."""

    def __init__(self, ks, label):
        self.ks, self.label = ks, label
        self.agree = np.zeros(ks.D)
        self.total = 0
        self.flips = 0
        self.max_dy = 0.0
        self.max_e64 = 0.0
        self.ys, self.yos, self.cleans = [], [], []

    def add(self, x, y, maps, clean, sigma, gamma, win, walk_q=False):
        ks = self.ks
        H, W = x.shape
        r0, c0, h, w = win
        crop, inner = crop_of(H, W, r0, c0, h, w)
        xc = np.ascontiguousarray(x[crop])
        var = ks.variances(sigma)
        yo, mo, marg = oracle.walk_f32(xc, ks, var, gamma)
        yo, mo, marg = yo[inner], mo[(slice(None),) + inner], marg[(slice(None),) + inner]
        yg = y[r0:r0 + h, c0:c0 + w]
        mg = maps[:, r0:r0 + h, c0:c0 + w]
        agree = mg == mo
        self.agree += agree.reshape(ks.D, -1).sum(axis=1)
        self.total += h * w
        bad = ~agree
        self.flips += int(bad.sum())
        assert np.all(marg[bad] < TOL_MARGIN), f"{self.label}: disagreement at margin {marg[bad].max()}"
        all_agree = agree.all(axis=0)
        dy = np.abs(yg.astype(np.float64) - yo)[all_agree]
        self.max_dy = max(self.max_dy, float(dy.max(initial=0.0)))
        y64, _ = oracle.walk_f64(xc, ks, sigma, gamma)
        self.max_e64 = max(self.max_e64, float(np.abs(yg - y64[inner]).max()))
        self.ys.append(yg.ravel())
        self.yos.append(yo.ravel())
        if clean is not None:
            self.cleans.append(clean[r0:r0 + h, c0:c0 + w].ravel())
        if walk_q:
            info = ks.info()
            yq, mq = oracle.walk_q(xc, ks, var, gamma, info["ring_frac_bits"], info["ring_range"])
            assert np.array_equal(mg, mq[(slice(None),) + inner]), (self.label, win)
            assert np.array_equal(yg.view(np.uint32), yq[inner].view(np.uint32)), (self.label, win)

    def check(self):
        per_dir = self.agree / self.total
        assert per_dir.min() >= MIN_AGREE, f"{self.label}: scale agreement {per_dir.min():.6f}"
        assert self.max_dy <= TOL_Y, f"{self.label}: max |dy| {self.max_dy}"
        line = (f"{self.label}: {self.total} px in windows, agree={per_dir.min():.6f} flips={self.flips} "
                f"max|dy|={self.max_dy:.2e} max|y-y64|={self.max_e64:.2e}")
        if self.cleans:
            clean = np.concatenate(self.cleans)
            dp = abs(lpa_ici.psnr(clean, np.concatenate(self.ys)) - lpa_ici.psnr(clean, np.concatenate(self.yos)))
            assert dp <= TOL_PSNR, f"{self.label}: dPSNR {dp}"
            line += f" dPSNR={dp:.2e}"
        print(line)
        return per_dir.min()


# ---------------------------------------------------------------------------
# one 4096 x 4096 image
# ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def image4096():
    ws = workloads.single(4096)
    clean = ws.clean(0)
    return clean, lpa_ici.make_noisy(clean, ws.sigmas[0], ws.noise_seeds[0])


@pytest.mark.parametrize("name", ["symmetric", "m20_gaussian"])
def test_4096_windows_meet_the_bar(image4096, name):
    """The default walk of each set (fixed-point ring sums for the symmetric
    set, the weighted FMA walk for the gaussian set) on one 4096^2 image:
    eight 192^2 windows (corners, edges, random) against the reference-order
    FP32 walk at the bar, and, for the symmetric set, against the fixed-point
    restatement bit for bit."""
    ks = kernels() if name == "symmetric" else kernels(GAUSSIAN_M20)
    clean, x = image4096
    y, maps = lpa_ici.denoise(ks, dev(x), SIGMA, GAMMA, scale_maps=True)
    torch.cuda.synchronize()
    y, maps = y.cpu().numpy(), maps.cpu().numpy()
    bar = WindowBar(ks, f"4096^2 {name}")
    rng = np.random.default_rng(11)
    wins = windows(4096, 4096, 192, rng) + windows(4096, 4096, 192, rng)[4:]
    for win in wins:
        bar.add(x, y, maps, clean, SIGMA, GAMMA, win, walk_q=(name == "symmetric"))
    bar.check()


# ---------------------------------------------------------------------------
# 256 uneven images in one persistent launch
# ---------------------------------------------------------------------------

def test_uneven_256_batched_plain_graph_identical(ks):
    """256 images, sides 128..2048 (about 300 Mpx), per-image sigma and gamma:
    one persistent launch, one launch per image and three replays of the
    captured hipGraph give the same outputs and scale maps, bit for bit.
    Windows of six images meet the bar against the oracle and equal the
    fixed-point restatement."""
    ws = workloads.uneven(256)
    xs_host = ws.noisy_all()
    xs = [dev(x) for x in xs_host]
    sig, gam = ws.sigmas, ws.gammas
    yb, mb = lpa_ici.denoise(ks, xs, sig, gam, scale_maps=True)
    yp, mp = lpa_ici.denoise(ks, xs, sig, gam, scale_maps=True, flags=lpa_ici.FWD_PLAIN)
    torch.cuda.synchronize()
    for i in range(len(xs)):
        assert torch.equal(yb[i], yp[i]) and torch.equal(mb[i], mp[i]), f"image {i}"
    del yp, mp
    ys = [torch.empty_like(x) for x in xs]
    ms = [torch.empty_like(m) for m in mb]
    stream = torch.cuda.Stream()
    plan = lpa_ici.Plan(ks, xs, ys, ms, sig, gam, flags=lpa_ici.FWD_GRAPH)
    for rep in range(3):
        for y in ys:
            y.fill_(float("nan"))
        for m in ms:
            m.fill_(255)
        torch.cuda.synchronize()
        with torch.cuda.stream(stream):
            plan.run(stream)
        stream.synchronize()
        for i in range(len(xs)):
            assert torch.equal(ys[i], yb[i]) and torch.equal(ms[i], mb[i]), f"replay {rep}, image {i}"
        _ = float(xs[0].sum().item())  # unrelated eager work between replays
    plan.close()
    print(f"256 uneven images, {ws.pixels / 1e6:.1f} Mpx: batched == plain == 3 graph replays")
    # windows of six images: the two smallest, the two largest, two others
    area = np.array([w * h for w, h in ws.sides])
    pick = list(np.argsort(area)[:2]) + list(np.argsort(area)[-2:]) + [17, 101]
    rng = np.random.default_rng(12)
    bar = WindowBar(ks, "256 uneven images")
    for i in pick:
        i = int(i)
        y, maps = yb[i].cpu().numpy(), mb[i].cpu().numpy()
        H, W = y.shape
        for win in windows(H, W, 96, rng)[:3]:
            bar.add(xs_host[i], y, maps, ws.clean(i), ws.sigmas[i], ws.gammas[i], win, walk_q=True)
    bar.check()


# ---------------------------------------------------------------------------
# 64 x 512^2: forward + adjoint, finite differences
# ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def batch64(ks):
    ws = workloads.batch_512(64)
    xs_host = ws.noisy_all()
    xs = [dev(x) for x in xs_host]
    ys, maps = lpa_ici.denoise(ks, xs, SIGMA, GAMMA, scale_maps=True)
    gen = torch.Generator(device="cuda").manual_seed(5)
    gs = [torch.randn(x.shape, generator=gen, device="cuda") for x in xs]
    dxs, dk = lpa_ici.backward(ks, xs, ys, maps, gs, SIGMA, GAMMA)
    torch.cuda.synchronize()
    return {"x": xs_host, "maps": [m.cpu().numpy() for m in maps], "g": [g.cpu().numpy() for g in gs],
            "dx": [d.cpu().numpy() for d in dxs], "dk": dk.cpu().numpy()}


def fd_bar(got, fd, scale):
    """|got - fd| <= 1e-3 max(|fd|, 1e-3 scale); returns whether the floor bound."""
    bound = TOL_FD * max(abs(fd), FD_FLOOR * scale)
    return abs(got - fd) <= bound, abs(fd) < FD_FLOOR * scale


def test_batch64_dx_matches_fp64_adjoint_on_windows(ks, batch64):
    """dL/dx at a pixel depends on the upstream gradient and the scale maps
    within 31 px: the FP64 adjoint on a crop with that margin gives the
    window's values exactly.  Two images, three 48^2 windows each."""
    rng = np.random.default_rng(13)
    worst = 0.0
    for i in (0, 37):
        x, maps, g, dx = batch64["x"][i], batch64["maps"][i], batch64["g"][i], batch64["dx"][i]
        scale = np.abs(dx).max()
        for (r0, c0, h, w) in windows(512, 512, 48, rng)[::2]:
            crop, inner = crop_of(512, 512, r0, c0, h, w)
            dx64, _ = oracle.adjoint_f64(np.ascontiguousarray(x[crop]), np.ascontiguousarray(g[crop]), ks, SIGMA,
                                         np.ascontiguousarray(maps[(slice(None),) + crop]))
            worst = max(worst, float(np.abs(dx[r0:r0 + h, c0:c0 + w] - dx64[inner]).max() / scale))
    print(f"64 x 512^2: dL/dx vs FP64 adjoint on windows, max rel {worst:.2e}")
    assert worst < 1e-4


def fd_pixels(ks, batch, n_points, seed):
    """FP64 central differences (eps 1e-3) of L = sum(g y) over one image of
    the batch, scales frozen at the batch's maps, at random pixels: evaluated
    on a crop of 62 px around the pixel (every output its value can reach,
    and every input those outputs read).  Points where +-eps flips a selection
    of the FP32 walk within reach are skipped.  Returns [(i, r, c, fd)] and
    the number skipped."""
    rng = np.random.default_rng(seed)
    var = ks.variances(SIGMA)
    out, skipped = [], 0
    n = len(batch["x"])
    for _ in range(2 * n_points):
        if len(out) == n_points:
            break
        i = int(rng.integers(0, n))
        x, maps, g = batch["x"][i], batch["maps"][i], batch["g"][i]
        H, W = x.shape
        r, c = int(rng.integers(0, H)), int(rng.integers(0, W))
        crop, _ = crop_of(H, W, r, c, 1, 1, margin=2 * R)
        xc = np.ascontiguousarray(x[crop])
        mc = np.ascontiguousarray(maps[(slice(None),) + crop])
        gc = g[crop].astype(np.float64)
        rr, cc = r - crop[0].start, c - crop[1].start
        xp, xm = xc.copy(), xc.copy()
        xp[rr, cc] += EPS
        xm[rr, cc] -= EPS
        reach = (slice(None), slice(max(0, rr - R), rr + R + 1), slice(max(0, cc - R), cc + R + 1))
        m0 = oracle.walk_f32(xc, ks, var, GAMMA, threads=1)[1][reach]
        if not all(np.array_equal(oracle.walk_f32(v, ks, var, GAMMA, threads=1)[1][reach], m0) for v in (xp, xm)):
            skipped += 1
            continue
        step = float(np.float64(xp[rr, cc]) - np.float64(xm[rr, cc]))
        lp = (gc * oracle.walk_f64(xp, ks, SIGMA, GAMMA, frozen=mc, threads=1)[0]).sum()
        lm = (gc * oracle.walk_f64(xm, ks, SIGMA, GAMMA, frozen=mc, threads=1)[0]).sum()
        out.append((i, r, c, (lp - lm) / step))
    return out, skipped


def fd_taps(ks, batch, taps):
    """FP64 fourth-order central differences (eps 1e-3) of L = sum over the
    batch of sum(g y), scales frozen at the batch's maps, at run weights
    `taps`.  The forward at a perturbed tap t of kernel k = (d, s) is
    evaluated exactly in FP64 from the oracle's per-direction estimates: only
    pixels with s_d = s change, their direction-d estimate by delta x[p + t]
    and its weight 1/var_k through sum(w^2)."""
    D, S = ks.D, ks.S
    r = ks.runs()
    ss = np.array([np.sum(ks.kernel(d, s).astype(np.float64) ** 2) for d in range(D) for s in range(S)])
    inv_tab = (1.0 / (SIGMA ** 2 * ss)).reshape(D, S)
    info = []
    for t in taps:
        k, di = dense_index(ks, r, int(t))
        d, s = divmod(k, S)
        kk = ks.kernel(d, s)
        a, b = np.unravel_index(di - ks.offsets[k], kk.shape)
        w0 = np.float32(ks.weights[di])
        ws = [np.float32(np.float64(w0) + h * EPS) for h in (-2, -1, 1, 2)]
        deltas = np.array([np.float64(w) - np.float64(w0) for w in ws])
        # 1/var of the kernel with the tap replaced: sum(w^2) swaps one term
        inv_p = 1.0 / (SIGMA ** 2 * (ss[k] - np.float64(w0) ** 2 + np.array([np.float64(w) ** 2 for w in ws])))
        info.append((d, s, int(a) - kk.shape[0] // 2, int(b) - kk.shape[1] // 2, deltas, inv_p))
    dl = np.zeros((len(taps), 4))
    for i in range(len(batch["x"])):
        x, maps, g = batch["x"][i], batch["maps"][i], batch["g"][i].astype(np.float64)
        H, W = x.shape
        _, _, est = oracle.walk_f64(x, ks, SIGMA, GAMMA, frozen=maps, want_est=True)
        inv = inv_tab[np.arange(D)[:, None, None], maps]
        A, B = (inv * est).sum(axis=0), inv.sum(axis=0)
        xpad = np.zeros((H + 2 * R, W + 2 * R))
        xpad[R:-R, R:-R] = x
        for j, (d, s, dr, dc, deltas, inv_p) in enumerate(info):
            P = maps[d] == s
            if not P.any():
                continue
            e, iv = est[d][P], inv[d][P]
            xs = xpad[R + dr:R + dr + H, R + dc:R + dc + W][P]
            A0, B0, gp = A[P] - iv * e, B[P] - iv, g[P]
            y0 = A[P] / B[P]
            for h in range(4):
                dl[j, h] += (gp * ((A0 + inv_p[h] * (e + deltas[h] * xs)) / (B0 + inv_p[h]) - y0)).sum()
    return np.array([float(np.polyfit(info[j][4], dl[j], 3)[2]) for j in range(len(taps))])


def test_batch64_dx_central_differences(ks, batch64):
    """dL/dx at 64 random pixels of the batch against FP64 central differences
    of the frozen-selection forward (fd_pixels)."""
    points, skipped = fd_pixels(ks, batch64, 64, seed=14)
    floor = 0
    worst = 0.0
    for i, r, c, fd in points:
        dx = batch64["dx"][i]
        ok, low = fd_bar(dx[r, c], fd, np.abs(dx).max())
        assert ok, (i, r, c, dx[r, c], fd)
        worst = max(worst, abs(dx[r, c] - fd) / max(abs(fd), 1e-30))
        floor += low
    print(f"64 x 512^2 dL/dx: {len(points)} pixels within 1e-3 of central differences ({floor} under the floor, "
          f"{skipped} skipped for a flip), worst relative {worst:.2e}")
    assert len(points) == 64


def test_batch64_dkernel_central_differences(ks, batch64):
    """dL/dkernel at 64 random taps against FP64 fourth-order central
    differences of the frozen-selection forward summed over the 64 images
    (fd_taps).  At this size a +-2e-3 change of one tap moves the estimate of
    every pixel that evaluates the kernel (millions), so some selection of the
    FP32 walk flips for every tap (counted below for the first); the
    differences are of the frozen-selection forward, the function whose
    gradient the adjoint is."""
    r = ks.runs()
    taps = np.random.default_rng(15).choice(len(r["weights"]), 64, replace=False)
    fd = fd_taps(ks, batch64, taps)
    dk = batch64["dk"]
    scale = np.abs(dk).max()
    floor = 0
    worst = 0.0
    for t, f in zip(taps, fd):
        ok, low = fd_bar(dk[t], f, scale)
        assert ok, (int(t), dk[t], f)
        worst = max(worst, abs(dk[t] - f) / max(abs(f), 1e-30))
        floor += low
    # flips of the selection for the first tap at +2 eps, on the GPU's default walk
    di = dense_index(ks, r, int(taps[0]))[1]
    mod = ks.weights.copy()
    mod[di] = np.float32(np.float64(mod[di]) + 2 * EPS)
    kmod = lpa_ici.KernelSet(ks.D, ks.S, mod, ks.offsets, ks.rows, ks.cols)
    flips = 0
    for x, m0 in zip(batch64["x"], batch64["maps"]):
        _, m = lpa_ici.denoise(kmod, dev(x), SIGMA, GAMMA, scale_maps=True)
        flips += int((m.cpu().numpy() != m0).sum())
    kmod.close()
    print(f"64 x 512^2 dL/dkernel: 64 taps within 1e-3 of fourth-order central differences ({floor} under the "
          f"floor), worst relative {worst:.2e}; +2 eps on the first tap flips {flips} selections of the batch")


# ---------------------------------------------------------------------------
# 32 images of 1024..8192 px per side: forward + adjoint in one call
# ---------------------------------------------------------------------------

def test_large_uneven_32_fwd_bwd(ks):
    """32 images with sides 1024..8192 (about 630 Mpx), per-image sigma and
    gamma.  Forward: the persistent launch equals one launch per image.
    Adjoint (estimates recomputed from the uint8 maps): each image's dL/dx
    equals a single-image call bit for bit, dL/dkernel the sum of the
    single-image calls up to float reassociation.  Peak device memory in use
    stays far below 288 GB.  One window of the largest image against the
    oracle: output bit for bit (fixed-point restatement), dL/dx within 1e-4
    of the FP64 adjoint."""
    ws = workloads.large_uneven(32)
    torch.cuda.synchronize()
    total = torch.cuda.mem_get_info()[1]
    torch.cuda.reset_peak_memory_stats()
    lpa_ici.workspace_peak(reset=True)
    xs_host = ws.noisy_all()
    xs = [dev(x) for x in xs_host]
    sig, gam = ws.sigmas, ws.gammas
    ys, maps = lpa_ici.denoise(ks, xs, sig, gam, scale_maps=True)
    for i in range(len(xs)):
        yp, mp = lpa_ici.denoise(ks, xs[i], sig[i], gam[i], scale_maps=True, flags=lpa_ici.FWD_PLAIN)
        assert torch.equal(yp, ys[i]) and torch.equal(mp, maps[i]), f"forward, image {i}"
    del yp, mp
    gen = torch.Generator(device="cuda").manual_seed(7)
    gs = [torch.randn(x.shape, generator=gen, device="cuda") for x in xs]
    dxs, dk = lpa_ici.backward(ks, xs, ys, maps, gs, sig, gam)
    dk_sum = torch.zeros_like(dk)
    for i in range(len(xs)):
        dx1, _ = lpa_ici.backward(ks, [xs[i]], [ys[i]], [maps[i]], [gs[i]], sig[i], gam[i], dkernel=dk_sum)
        assert torch.equal(dx1[0], dxs[i]), f"adjoint, image {i}"
        del dx1
    torch.cuda.synchronize()
    assert torch.allclose(dk, dk_sum, rtol=1e-4, atol=1e-4 * dk.abs().max().item())
    # upper bound of the peak: the tensors' high-water mark plus the library
    # workspaces' (each at its own moment)
    tensors, work = torch.cuda.max_memory_allocated(), lpa_ici.workspace_peak()
    peak = tensors + work
    print(f"32 uneven images, {ws.pixels / 1e6:.1f} Mpx: fwd + adjoint, batch == single calls; "
          f"peak device memory <= {peak / 2**30:.1f} GiB (tensors {tensors / 2**30:.1f}, library workspaces "
          f"{work / 2**30:.2f}) of {total / 2**30:.0f} GiB")
    assert work > 0
    assert peak < 0.5 * total
    i = int(np.argmax([w * h for w, h in ws.sides]))
    H, W = xs_host[i].shape
    y, m = ys[i].cpu().numpy(), maps[i].cpu().numpy()
    bar = WindowBar(ks, "32 uneven images, largest")
    win = (H // 2, W // 3, 64, 64)
    bar.add(xs_host[i], y, m, ws.clean(i), sig[i], gam[i], win, walk_q=True)
    bar.check()
    crop, inner = crop_of(H, W, *win)
    g = gs[i][crop].cpu().numpy()
    dx64, _ = oracle.adjoint_f64(np.ascontiguousarray(xs_host[i][crop]), np.ascontiguousarray(g), ks, sig[i],
                                 np.ascontiguousarray(m[(slice(None),) + crop]))
    dxw = dxs[i][win[0]:win[0] + 64, win[1]:win[1] + 64].cpu().numpy()
    err = np.abs(dxw - dx64[inner]).max() / np.abs(dx64[inner]).max()
    print(f"largest image {W}x{H}: dL/dx window vs FP64 adjoint rel {err:.2e}")
    assert err < 1e-4
