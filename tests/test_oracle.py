"""Pins the oracle (oracle/lpa_oracle.c) before anything is checked against it.

* against the reference's own CPU code, compiled from its sources into
  oracle/_ref (bit-exact), where that build exists;
* against the golden fixtures the reference wrote (bit-exact; these travel);
* against what the mathematics fixes: brute-force FP64 correlation, impulse
  and constant images, the ICI invariants, aggregation convexity, and a
  planted mistake that each check must catch.
"""
import json
import os

import numpy as np
import pytest

import lpa_ici
import oracle
from conftest import (GAMMA, GAUSSIAN_M20, GOLDEN, REF_KERNEL_DIR, SEED, SIGMA, SYMMETRIC, golden, kernels,
                      noisy_synth)

needs_ref = pytest.mark.skipif(not oracle.ref_available("v0_0") or not os.path.isdir(REF_KERNEL_DIR),
                               reason="reference build (oracle/_ref) or tree not present")


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


class TinySet:
    """A hand-made 2-direction x 3-scale kernel set for closed-form checks."""

    def __init__(self):
        k = []
        # direction 0: horizontal wedges to the right, uniform weights
        k.append(np.array([[1.0]], np.float32))
        a = np.zeros((3, 3), np.float32); a[1, 1:] = 0.5; k.append(a)
        b = np.zeros((5, 5), np.float32); b[2, 2:] = 1 / 3; k.append(b)
        # direction 1: vertical wedges downwards, non-uniform weights, one negative
        k.append(np.array([[1.0]], np.float32))
        c = np.zeros((3, 3), np.float32); c[1:, 1] = [0.75, 0.25]; k.append(c)
        e = np.zeros((5, 5), np.float32); e[2:, 2] = [0.6, 0.5, -0.1]; k.append(e)
        self.D, self.S = 2, 3
        self.rows = np.array([m.shape[0] for m in k], np.int32)
        self.cols = np.array([m.shape[1] for m in k], np.int32)
        self.offsets = np.cumsum([0] + [m.size for m in k[:-1]]).astype(np.int32)
        self.weights = np.concatenate([m.ravel() for m in k]).astype(np.float32)
        self.mats = k

    def variances(self, sigma):
        s2 = np.float32(sigma) * np.float32(sigma)
        return np.array([s2 * np.float32((m.astype(np.float32) ** 2).sum(dtype=np.float32)) for m in self.mats],
                        np.float32)


def brute_walk_f64(x, ks, sigma, gamma):
    """Pure-Python walk in double for tiny images: dense correlation with zero
    padding, intervals, aggregation."""
    H, W = x.shape
    out = np.zeros((H, W))
    scales = np.zeros((ks.D, H, W), np.uint8)
    for r in range(H):
        for c in range(W):
            num = den = 0.0
            for d in range(ks.D):
                lo, hi, best, var_b = -np.inf, np.inf, 0.0, 1.0
                for s in range(ks.S):
                    k = d * ks.S + s
                    m = ks.weights[ks.offsets[k]:ks.offsets[k] + ks.rows[k] * ks.cols[k]].reshape(ks.rows[k],
                                                                                                 ks.cols[k])
                    v = 0.0
                    for i in range(m.shape[0]):
                        for j in range(m.shape[1]):
                            rr, cc = r - m.shape[0] // 2 + i, c - m.shape[1] // 2 + j
                            if 0 <= rr < H and 0 <= cc < W:
                                v += float(m[i, j]) * float(x[rr, cc])
                    var = sigma * sigma * float((m.astype(np.float64) ** 2).sum())
                    t = gamma * np.sqrt(var)
                    lo, hi = max(lo, v - t), min(hi, v + t)
                    if hi < lo:
                        break
                    best, var_b = v, var
                    scales[d, r, c] = s
                num += best / var_b
                den += 1.0 / var_b
            out[r, c] = num / den
    return out, scales


# ---------------------------------------------------------------------------
# against the reference build and its fixtures
# ---------------------------------------------------------------------------

@needs_ref
@pytest.mark.parametrize("name", [SYMMETRIC, GAUSSIAN_M20, "lpa_kernels_m_1_0_d_4_square.txt",
                                  "lpa_kernels_python_16_tapered.txt"])
def test_v0_0_and_v1_0_match_reference_build(name):
    ks = kernels(name)
    x = noisy_synth(40, 33, seed=9)
    var = ks.variances(SIGMA)
    ref0 = oracle.ref_denoise(x, ks, var, GAMMA, "v0_0")
    ref1 = oracle.ref_denoise(x, ks, var, GAMMA, "v1_0")
    assert np.array_equal(bits(ref0), bits(ref1))
    assert np.array_equal(bits(oracle.v0_0(x, ks, var, GAMMA)), bits(ref0))
    assert np.array_equal(bits(oracle.v1_0(x, ks, var, GAMMA)), bits(ref0))
    assert np.array_equal(bits(oracle.walk_f32(x, ks, var, GAMMA)[0]), bits(ref0))


@pytest.mark.parametrize("tag,name", [("symmetric", SYMMETRIC), ("m20_gaussian", GAUSSIAN_M20),
                                      ("d4_square", "lpa_kernels_m_1_0_d_4_square.txt")])
def test_oracle_matches_reference_fixtures(tag, name):
    ks = kernels(name)
    x = golden("synth64_noisy.npy")
    var = ks.variances(SIGMA)
    ref = golden(f"synth64_v0_0_{tag}.npy")
    y0, maps, marg = oracle.v0_0(x, ks, var, GAMMA, want_maps=True)
    assert np.array_equal(bits(y0), bits(ref))
    yw, mw, mgw = oracle.walk_f32(x, ks, var, GAMMA)
    assert np.array_equal(bits(yw), bits(ref))
    assert np.array_equal(maps, mw)
    assert np.array_equal(marg, mgw)


def test_walk_matches_reference_fixture_cameraman(ks):
    x = golden("cameraman_noisy.npy")
    y, _, _ = oracle.walk_f32(x, ks, ks.variances(SIGMA), GAMMA)
    assert np.array_equal(bits(y), bits(golden("cameraman_v0_0.npy")))


def test_published_psnr_barbara(ks):
    """README:7 and :208: noisy 22.1056 dB, denoised 25.0736 dB."""
    info = json.load(open(os.path.join(GOLDEN, "images.json")))["barbara"]
    clean = lpa_ici.load_image(os.path.join(GOLDEN, "images", "barbara.png"))
    x = lpa_ici.make_noisy(clean, SIGMA, SEED)
    y, _, _ = oracle.walk_f32(x, ks, ks.variances(SIGMA), GAMMA)
    assert round(lpa_ici.psnr(clean, x), 4) == 22.1056
    assert round(lpa_ici.psnr(clean, y), 4) == 25.0736
    assert lpa_ici.psnr(clean, y) == info["psnr_denoised"]


# ---------------------------------------------------------------------------
# against the mathematics
# ---------------------------------------------------------------------------

def test_walk_f64_matches_bruteforce_tiny():
    ks = TinySet()
    rng = np.random.default_rng(0)
    x = rng.random((7, 9)).astype(np.float32)
    y, maps = oracle.walk_f64(x, ks, 0.1, 1.5)
    yb, mb = brute_walk_f64(x, ks, 0.1, 1.5)
    assert np.array_equal(maps, mb)
    assert np.allclose(y, yb, rtol=0, atol=1e-13)
    # the f32 restatement agrees to float rounding on the same decisions
    y32, m32, _ = oracle.walk_f32(x, ks, ks.variances(0.1), 1.5)
    same = (m32 == mb).all(axis=0)
    assert same.mean() > 0.9
    assert np.abs(y32 - yb)[same].max() < 1e-5


def test_correlation_orientation_impulse():
    """An impulse at (r0, c0) gives yhat(p) = k[r0 - p + centre]: the kernel
    read back flipped, which fixes correlation (not convolution)."""
    ks = TinySet()
    x = np.zeros((9, 9), np.float32)
    x[4, 4] = 1.0
    frozen = np.full((ks.D, 9, 9), 2, np.uint8)
    _, _, est = oracle.walk_f64(x, ks, 0.1, 1.5, frozen=frozen, want_est=True)
    # direction 0, scale 2: weights at row offset 0, column offsets 0..2
    for c in range(9):
        expect = 1 / 3 if 0 <= 4 - c <= 2 else 0.0
        assert np.isclose(est[0, 4, c], np.float32(1 / 3) if expect else 0.0, atol=1e-7)
    # direction 1, scale 2: rows 0..2 below the centre with weights 0.6, 0.5, -0.1
    assert np.allclose([est[1, 4, 4], est[1, 3, 4], est[1, 2, 4]], [0.6, 0.5, -0.1], atol=1e-7)


def test_constant_image_interior(ks):
    """A constant image c gives yhat = c * sum(w) away from the borders."""
    x = np.full((80, 80), 0.5, np.float32)
    y64, _, est = oracle.walk_f64(x, ks, SIGMA, GAMMA, want_est=True)
    inner = slice(32, 48)
    assert np.allclose(y64[inner, inner], 0.5, atol=5e-5)  # weights sum to 1 +- 8e-5
    sums = np.array([ks.weights[ks.offsets[k]:ks.offsets[k] + ks.rows[k] * ks.cols[k]].astype(np.float64).sum()
                     for k in range(ks.K)])
    # a noiseless constant never breaks an interval: every direction reaches the top scale
    _, maps = oracle.walk_f64(x, ks, SIGMA, GAMMA)
    assert np.all(maps[:, inner, inner] == ks.S - 1)
    for d in range(ks.D):
        assert np.allclose(est[d, inner, inner], 0.5 * sums[d * ks.S + ks.S - 1], atol=1e-12)


def test_ici_invariants(ks):
    """Scale 0 is the 1x1 identity (always accepted, estimate = the pixel),
    the selected scale is the largest s whose intervals 0..s all intersect,
    and v0_0 (no early exit) equals the early-exit walk."""
    x = noisy_synth(48, 40, seed=21)
    var = ks.variances(SIGMA)
    yw, mw, _, est = oracle.walk_f32(x, ks, var, GAMMA, want_est=True)
    y0, m0, _ = oracle.v0_0(x, ks, var, GAMMA, want_maps=True)
    assert np.array_equal(mw, m0) and np.array_equal(bits(yw), bits(y0))
    assert np.all(ks.weights[ks.offsets[0::ks.S]] == 1.0) and np.all(ks.rows[0::ks.S] == 1)
    # recompute the decision from per-scale estimates in numpy (FP64 intervals)
    _, _, e64 = oracle.walk_f64(x, ks, SIGMA, GAMMA, frozen=np.zeros_like(mw), want_est=True)
    assert np.allclose(e64, x[None], atol=0)  # scale 0 estimate is the pixel itself


def test_aggregation_convex_and_single_direction(ks):
    x = noisy_synth(40, 40, seed=3)
    y, maps, est = oracle.walk_f64(x, ks, SIGMA, GAMMA, want_est=True)
    assert np.all(y >= est.min(axis=0) - 1e-12) and np.all(y <= est.max(axis=0) + 1e-12)

    class One:  # direction 0 alone returns its own estimate
        D, S = 1, ks.S
        weights, offsets = ks.weights, ks.offsets[:ks.S]
        rows, cols = ks.rows[:ks.S], ks.cols[:ks.S]
    y1, _, e1 = oracle.walk_f64(x, One, SIGMA, GAMMA, want_est=True)
    assert np.allclose(y1, e1[0], atol=1e-15)


def test_planted_mistakes_are_caught():
    """Each check above fails for a plausible mistake planted in the oracle's
    inputs: a transposed kernel, a dropped tap, a flipped kernel."""
    ks = TinySet()
    rng = np.random.default_rng(1)
    x = rng.random((7, 9)).astype(np.float32)
    yb, mb = brute_walk_f64(x, ks, 0.1, 1.5)

    def variant(fn):
        v = TinySet()
        mats = [fn(m.copy()) for m in v.mats]
        v.weights = np.concatenate([m.ravel() for m in mats]).astype(np.float32)
        return v

    for name, fn in [("transposed", lambda m: np.ascontiguousarray(m.T)),
                     ("flipped", lambda m: np.ascontiguousarray(m[::-1, ::-1])),
                     ("dropped tap", lambda m: (m.__setitem__(np.unravel_index(np.argmax(np.abs(m)), m.shape), 0.0)
                                                or m))]:
        y, maps = oracle.walk_f64(x, variant(fn), 0.1, 1.5)
        assert not (np.allclose(y, yb, atol=1e-9) and np.array_equal(maps, mb)), name


def test_margins_mark_near_ties(ks):
    """The margin is the smallest |U-L| over the decisions taken; a
    disagreement between two FP32 evaluations can only sit where it is tiny."""
    x = noisy_synth(64, 64, seed=8)
    var = ks.variances(SIGMA)
    _, maps, marg = oracle.walk_f32(x, ks, var, GAMMA)
    _, m64 = oracle.walk_f64(x, ks, SIGMA, GAMMA)
    flips = maps != m64
    assert flips.mean() < 1e-3
    assert np.all(marg[flips] < 1e-5)
    assert np.all(marg >= 0)


def test_useful_taps_count(ks):
    x = noisy_synth(50, 30, seed=4)
    _, maps, _ = oracle.walk_f32(x, ks, ks.variances(SIGMA), GAMMA)
    nnz = np.array([(ks.weights[ks.offsets[k]:ks.offsets[k] + ks.rows[k] * ks.cols[k]] != 0).sum()
                    for k in range(ks.K)])
    brute = 0
    for d in range(ks.D):
        for s_sel in maps[d].ravel():
            brute += nnz[d * ks.S:d * ks.S + min(int(s_sel) + 1, ks.S - 1) + 1].sum()
    assert oracle.useful_taps(maps, nnz) == brute
    from lpa_ici.metrics import useful_taps
    assert useful_taps(maps, nnz, ks.S) == brute


# ---------------------------------------------------------------------------
# the fixed-point ring-sum walk (the HIP path's default fast walk)
# ---------------------------------------------------------------------------

def fixed_point(ks):
    info = ks.info()
    assert info["ring_form"] == 1
    return info["ring_frac_bits"], info["ring_range"]


def test_fixed_point_parameters(ks):
    """F is the largest power with n_max * 2 * 2^F < 2^31 (n_max = 204 taps for
    the default set), and the range keeps every support sum inside 31 bits."""
    F, rng = fixed_point(ks)
    nmax = max(int((ks.weights[ks.offsets[k]:ks.offsets[k] + ks.rows[k] * ks.cols[k]] != 0).sum())
               for k in range(ks.K))
    assert nmax == 204 and F == 22
    assert nmax * 2 * 2.0 ** F < 2 ** 31 <= nmax * 2 * 2.0 ** (F + 1)
    assert nmax * (rng * 2.0 ** F + 0.5) < 2 ** 31 and 2.5 < rng < 2.6


def test_walk_q_constant_image_closed_form(ks):
    """A constant c away from the borders: every direction reaches the top
    scale and yhat = fl(fl(n * q(c)) * (w * 2^-F)) with n the support size.
    The same check fails when one tap of a support is dropped."""
    F, rng = fixed_point(ks)
    c = np.float32(0.3)
    x = np.full((80, 80), c, np.float32)
    qc = int(np.rint(np.float64(c) * 2.0 ** F))
    inner = slice(32, 48)

    def expected(kset, k):
        m = kset.weights[kset.offsets[k]:kset.offsets[k] + kset.rows[k] * kset.cols[k]]
        n = int((m != 0).sum())
        w = m[m != 0][0]
        return np.float32(np.int64(n * qc)) * (np.float32(w) * np.float32(2.0 ** -F))

    _, maps, est = oracle.walk_q(x, ks, ks.variances(SIGMA), GAMMA, F, rng, want_est=True)
    assert np.all(maps[:, inner, inner] == ks.S - 1)
    for d in range(ks.D):
        e = expected(ks, d * ks.S + ks.S - 1)
        assert np.all(bits(est[d, inner, inner]) == bits(e)), d

    class Dropped:  # kernel (0, S-1) loses one tap
        D, S = ks.D, ks.S
        offsets, rows, cols = ks.offsets, ks.rows, ks.cols
        weights = ks.weights.copy()
    k = ks.S - 1
    nz = np.flatnonzero(Dropped.weights[ks.offsets[k]:ks.offsets[k] + ks.rows[k] * ks.cols[k]])
    Dropped.weights[ks.offsets[k] + nz[-1]] = 0.0
    _, _, est2 = oracle.walk_q(x, Dropped, ks.variances(SIGMA), GAMMA, F, rng, want_est=True)
    assert not np.all(bits(est2[0, inner, inner]) == bits(expected(ks, k)))


def test_walk_q_impulse_orientation():
    """Uniform 2-tap wedges (right of the centre; below the centre): an impulse
    at (4, 4) shows up at the pixels the reflected supports reach."""
    class Uniform:
        D, S = 2, 2
        mats = [np.array([[1.0]], np.float32), np.zeros((3, 3), np.float32),
                np.array([[1.0]], np.float32), np.zeros((3, 3), np.float32)]
        mats[1][1, 1:] = 0.5
        mats[3][1:, 1] = 0.5
        rows = np.array([m.shape[0] for m in mats], np.int32)
        cols = np.array([m.shape[1] for m in mats], np.int32)
        offsets = np.cumsum([0] + [m.size for m in mats[:-1]]).astype(np.int32)
        weights = np.concatenate([m.ravel() for m in mats]).astype(np.float32)
    F, rng = 22, 2.5
    x = np.zeros((9, 9), np.float32)
    x[4, 4] = 0.7
    var = np.full(4, 100.0, np.float32)  # wide intervals: scale 1 is always taken
    _, maps, est = oracle.walk_q(x, Uniform, var, 1.0, F, rng, want_est=True)
    assert np.all(maps == 1)
    hit = np.float32(np.rint(0.7 * 2.0 ** F)) * np.float32(0.5 * 2.0 ** -F)
    e0 = np.zeros((9, 9), np.float32); e0[4, 3] = e0[4, 4] = hit   # pixel p sees x[p], x[p + (0, 1)]
    e1 = np.zeros((9, 9), np.float32); e1[3, 4] = e1[4, 4] = hit   # pixel p sees x[p], x[p + (1, 0)]
    assert np.array_equal(bits(est[0]), bits(e0)) and np.array_equal(bits(est[1]), bits(e1))


def test_walk_q_meets_the_bar_against_f32(ks):
    """Against the reference-order FP32 walk: the fast-mode bar (per-direction
    scale agreement >= 99.9 %, disagreements only at margins < 1e-6,
    |dy| <= 1e-5 where every direction agrees), and close to FP64."""
    F, rng = fixed_point(ks)
    x = noisy_synth(160, 120, seed=12)
    var = ks.variances(SIGMA)
    yq, mq = oracle.walk_q(x, ks, var, GAMMA, F, rng)
    yf, mf, marg = oracle.walk_f32(x, ks, var, GAMMA)
    agree = mq == mf
    assert agree.reshape(ks.D, -1).mean(axis=1).min() >= 0.999
    assert np.all(marg[~agree] < 1e-6)
    all_agree = agree.all(axis=0)
    assert np.abs(yq.astype(np.float64) - yf)[all_agree].max() <= 1e-5
    y64, m64 = oracle.walk_f64(x, ks, SIGMA, GAMMA)
    same = (mq == m64).all(axis=0)
    assert np.abs(yq - y64)[same].max() < 2e-6


def test_walk_q_refuses_out_of_range(ks):
    F, rng = fixed_point(ks)
    x = noisy_synth(20, 20, seed=1)
    x[3, 3] = rng
    with pytest.raises(ValueError):
        oracle.walk_q(x, ks, ks.variances(SIGMA), GAMMA, F, rng)
    x[3, 3] = np.nan
    with pytest.raises(ValueError):
        oracle.walk_q(x, ks, ks.variances(SIGMA), GAMMA, F, rng)


# ---------------------------------------------------------------------------
# oracle_adjoint_f64, pinned on its own by central differences of the FP64
# walk with frozen scales
# ---------------------------------------------------------------------------

FD_SETS = [SYMMETRIC, GAUSSIAN_M20, "lpa_kernels_m_1_0_d_4_square.txt", "lpa_kernels_m_0_0_d_4_square.txt"]
FD_EPS = 1e-3


class _Tapped:
    """A kernel set with one dense weight replaced (the FD abscissa)."""

    def __init__(self, ks, idx, value):
        self.D, self.S, self.offsets, self.rows, self.cols = ks.D, ks.S, ks.offsets, ks.rows, ks.cols
        self.weights = ks.weights.copy()
        self.weights[idx] = np.float32(value)


def _fd_reference(ks, seed):
    """FP64 central differences of L = sum(g * y) at 48 pixels (two-point:
    y is linear in x once the scales are frozen) and 48 dense taps (fourth
    order: y is rational in the taps through sum w^2), on a 26 x 22 image with
    random frozen scales, so every kernel, large ones at the borders
    included, is differentiated."""
    rng = np.random.default_rng(seed)
    H, W = 22, 26
    x = noisy_synth(W, H, seed=seed)
    maps = rng.integers(0, ks.S, size=(ks.D, H, W)).astype(np.uint8)
    g = rng.standard_normal((H, W)).astype(np.float32)
    g64 = g.astype(np.float64)

    def loss(xx, kk):
        y, _ = oracle.walk_f64(xx, kk, SIGMA, GAMMA, frozen=maps, threads=1)
        return float((g64 * y).sum())

    pix = [(0, 0), (H - 1, W - 1), (0, W - 1)] + [(int(rng.integers(0, H)), int(rng.integers(0, W)))
                                                 for _ in range(45)]
    fd_x = []
    for r, c in pix:
        xp, xm = x.copy(), x.copy()
        xp[r, c] += FD_EPS
        xm[r, c] -= FD_EPS
        step = float(np.float64(xp[r, c]) - np.float64(xm[r, c]))
        fd_x.append((loss(xp, ks) - loss(xm, ks)) / step)
    # taps: 32 nonzero and 16 zero ones inside the kernels' bounding boxes
    nz = np.flatnonzero(ks.weights != 0)
    z = np.flatnonzero(ks.weights == 0)
    taps = list(rng.choice(nz, 32, replace=False)) + (list(rng.choice(z, 16, replace=False)) if z.size >= 16 else [])
    fd_k = []
    for t in taps:
        w0 = ks.weights[t]
        hs = (-2, -1, 1, 2)
        sets = [_Tapped(ks, t, np.float64(w0) + h * FD_EPS) for h in hs]
        xs = np.array([np.float64(s.weights[t]) - np.float64(w0) for s in sets])
        ys = np.array([loss(x, s) for s in sets])
        fd_k.append(float(np.polyfit(xs, ys, 3)[2]))
    return x, g, maps, pix, np.array(fd_x), np.array(taps), np.array(fd_k)


def _fd_errors(ks, ref, mistake=0):
    x, g, maps, pix, fd_x, taps, fd_k = ref
    dx, dk = oracle.adjoint_f64(x, g, ks, SIGMA, maps, planted_mistake=mistake)
    ex = max(abs(dx[r, c] - f) for (r, c), f in zip(pix, fd_x)) / np.abs(fd_x).max()
    ek = np.abs(dk[taps] - fd_k).max() / np.abs(fd_k).max()
    return ex, ek


_fd_cache = {}


def _fd_for(name):
    if name not in _fd_cache:
        _fd_cache[name] = _fd_reference(kernels(name), seed=len(name))
    return _fd_cache[name]


@pytest.mark.parametrize("name", FD_SETS)
def test_adjoint_f64_central_differences(name):
    """dL/dx within 1e-9 and dL/dk within 1e-6 of FP64 central differences,
    relative to each gradient's largest component checked."""
    ks = kernels(name)
    ex, ek = _fd_errors(ks, _fd_for(name))
    print(f"{name}: adjoint_f64 vs central differences: dx {ex:.1e}, dk {ek:.1e}")
    assert ex < 1e-9 and ek < 1e-6


def test_adjoint_f64_planted_mistakes_are_caught():
    """Each planted mistake fails the finite-difference pin on at least one
    set, by far more than its tolerance: a wrong tap's weight in the
    variance term (invisible on uniform sets, caught on the gaussian set), a
    rotated kernel in dL/dx, the variance term's factor 2 dropped, y dropped
    from yhat - y, taps outside the image read as edge copies."""
    caught = {}
    for name in FD_SETS:
        ks = kernels(name)
        ref = _fd_for(name)
        for m in range(1, 6):
            ex, ek = _fd_errors(ks, ref, m)
            if ex > 1e-3 or ek > 1e-3:
                caught.setdefault(m, []).append(name)
    print({m: len(v) for m, v in caught.items()})
    assert sorted(caught) == [1, 2, 3, 4, 5], caught
    assert GAUSSIAN_M20 in caught[1]
