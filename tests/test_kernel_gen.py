"""Kernel-set generation (uniform wedges and directional LPA kernels of order
m with uniform, Gaussian and square windows) and the kernel-file writer.

Pins: the shipped default set (its loaded weights, offsets and sizes, which
the generator must give bit for bit), the file round trip, and what the
geometry fixes for any direction count: quarter-turn and mirror symmetry of
the supports, nesting across scales, the closed-form tap count of direction
0, and uniform weights that sum to one within the files' six decimals.
LPA kernels: the shipped LASIP-made sets (axis directions weight for weight,
every support; the square and symmetric sets whole), and what the
construction fixes without them: the fit's moment conditions along the axis,
the Gaussian factor across it, polynomial reproduction on the square
window, quarter-turn and mirror symmetry of the weights.
"""
import math

import numpy as np
import pytest

import lpa_ici
from conftest import SYMMETRIC, kernels

DEFAULT_H = [1, 2, 4, 8, 16, 24, 32]


def support(ks, d, s):
    """Lattice points (x, y) of kernel (d, s), y pointing up (y = -row)."""
    k = ks.kernel(d, s)
    R, C = k.shape
    rows, cols = np.nonzero(k)
    return {(int(c - C // 2), int(R // 2 - r)) for r, c in zip(rows, cols)}


def test_wedges_reproduce_the_shipped_default_set():
    g = lpa_ici.wedge_kernels(16, DEFAULT_H)
    ref = kernels(SYMMETRIC)
    assert (g.D, g.S) == (ref.D, ref.S)
    assert np.array_equal(g.offsets, ref.offsets)
    assert np.array_equal(g.rows, ref.rows) and np.array_equal(g.cols, ref.cols)
    assert np.array_equal(g.weights.view(np.uint32), ref.weights.view(np.uint32))


@pytest.mark.parametrize("D,h", [(16, DEFAULT_H), (8, [1, 2, 4, 8]), (32, [1, 3, 6, 12, 24])])
def test_write_and_load_round_trip(tmp_path, D, h):
    g = lpa_ici.wedge_kernels(D, h)
    path = tmp_path / f"wedges_d{D}.txt"
    lpa_ici.write_kernels(path, g)
    back = lpa_ici.load_kernels(str(path))
    assert (back.D, back.S) == (D, len(h))
    for name in ("weights", "offsets", "rows", "cols"):
        assert np.array_equal(getattr(back, name), getattr(g, name)), name


def test_written_text_format(tmp_path):
    g = lpa_ici.wedge_kernels(4, [1, 2])
    path = tmp_path / "small.txt"
    lpa_ici.write_kernels(path, g)
    lines = path.read_text().splitlines()
    assert lines[0] == "4 2"
    assert lines[1:3] == ["1 1", "1.000000"]
    # direction 0, h = 2: the centre and its right neighbour (45 degree
    # half-width: the diagonal neighbours (1, +-1) project to 0.707 < 1.5 too)
    assert lines[3] == "3 3"
    assert lines[4:7] == ["0.000000 0.000000 0.250000", "0.000000 0.250000 0.250000", "0.000000 0.000000 0.250000"]


@pytest.mark.parametrize("D", [4, 8, 16, 32])
def test_wedge_geometry(D):
    h = [1, 2, 3, 5, 8, 13, 21, 32]
    g = lpa_ici.wedge_kernels(D, h)
    info = g.info()
    assert info["ring_form"] == 1  # uniform weights, supports nested across scales
    for s, hs in enumerate(h):
        for d in range(D):
            sup = support(g, d, s)
            assert (0, 0) in sup
            if s:
                assert support(g, d, s - 1) <= sup
            # a quarter turn maps direction d to d + D/4, a mirror in the x axis to -d
            assert {(-y, x) for x, y in sup} == support(g, (d + D // 4) % D, s)
            assert {(x, -y) for x, y in sup} == support(g, (-d) % D, s)
            k = g.kernel(d, s)
            nz = k[k != 0]
            assert np.all(nz == nz[0])
            assert abs(float(nz.sum(dtype=np.float64)) - 1.0) <= nz.size * 5e-7 + 1e-12
        # direction 0: |y| <= x tan(pi/D), 0 <= x <= h - 1
        t = math.tan(math.pi / D)
        expect = sum(2 * math.floor(x * t + 1e-9) + 1 for x in range(hs))
        assert len(support(g, 0, s)) == expect


def test_wedge_errors(tmp_path):
    with pytest.raises(lpa_ici.LpaError):
        lpa_ici.wedge_kernels(3, [1, 2])
    with pytest.raises(lpa_ici.LpaError):
        lpa_ici.wedge_kernels(16, [1, 33])
    with pytest.raises(lpa_ici.LpaError):
        lpa_ici.wedge_kernels(16, [0])
    g = lpa_ici.wedge_kernels(8, [1, 2])
    with pytest.raises(lpa_ici.LpaError):
        lpa_ici.write_kernels(tmp_path / "missing" / "k.txt", g)


# ---------------------------------------------------------------------------
# directional LPA kernels of order m
# ---------------------------------------------------------------------------

H7 = [1, 2, 4, 8, 16, 24, 32]
# (file, D, window lengths, order, window) of the shipped sets the LASIP demo made
LASIP_SETS = [
    ("lpa_kernels_m_2_0_d_16_h12_1_2_4_8_16_24_32_gaussian.txt", 16, H7, 2, "gaussian"),
    ("lpa_kernels_m_1_0_d_16_h12_1_2_4_8_16_24_32_gaussian.txt", 16, H7, 1, "gaussian"),
    ("lpa_kernels_m_1_0_d_16_h12_1_2_4_8_16_24_32.txt", 16, H7, 1, "gaussian"),
    ("lpa_kernels_m_2_0_d_16_h12_1_2_4_8_16_24_32.txt", 16, H7, 2, "uniform"),
    ("lpa_kernels_m_1_0_d_16_h12_1_2_4_8_16.txt", 16, [1, 2, 4, 8, 16], 1, "uniform"),
    ("lpa_kernels_m_1_0_d_8_h12_1_2_4_8.txt", 8, [1, 2, 4, 8], 1, "uniform"),
    ("lpa_kernels_m_1_0_d_8_h12_1_2_4_8_16.txt", 8, [1, 2, 4, 8, 16], 1, "uniform"),
    ("lpa_kernels_m_1_0_d_8_h12_1_2_4_8_16_24_32.txt", 8, H7, 1, "uniform"),
    ("lpa_kernels_m_1_0_d_8_h12_1_15_15_30.txt", 8, [1, 5, 15, 30], 1, "uniform"),
    ("lpa_kernels_m_1_0_d_16_h12_1_2_4_8_16_24_32_symmetric.txt", 16, H7, 0, "uniform"),
    ("lpa_kernels_m_1_0_d_4_square.txt", 4, [1, 2, 4, 8, 16], 1, "square"),
    ("lpa_kernels_m_0_0_d_4_square.txt", 4, [1, 2, 4, 8, 16], 0, "square"),
]


@pytest.mark.parametrize("name,D,h,m,window", LASIP_SETS)
def test_lpa_kernels_reproduce_the_shipped_sets(name, D, h, m, window):
    """Every kernel of the axis directions (phi a multiple of 90 degrees)
    equals the file weight for weight, and every kernel of every direction has
    the file's support (the wedge); for order 0 and for the square sets that is
    every kernel.  Oblique directions of order >= 1 are the one place the LASIP
    discretization is not restated: their largest deviation is printed, and
    for the Gaussian sets held below 5 % of the kernel's largest weight."""
    f = kernels(name)
    g = lpa_ici.make_kernels(D, h, m, window)
    wedges = lpa_ici.wedge_kernels(D, h) if window != "square" else lpa_ici.make_kernels(D, h, 0, "square")
    assert (g.D, g.S) == (f.D, f.S)
    assert np.array_equal(g.rows, f.rows) and np.array_equal(g.cols, f.cols)
    oblique = 0.0
    for d in range(D):
        for s in range(len(h)):
            a, b = f.kernel(d, s), g.kernel(d, s)
            # the files write some support points as -0.000000: every nonzero
            # of the file is one here, and every nonzero here is on the support
            assert not np.any((a != 0) & (b == 0)), (d, s)
            assert not np.any((b != 0) & (wedges.kernel(d, s) == 0)), (d, s)
            if (4 * d) % D == 0 or m == 0:
                assert np.array_equal(a, b), (d, s)
            else:
                oblique = max(oblique, float(np.abs(a - b).max() / np.abs(a).max()))
    print(f"{name}: axis kernels equal, oblique max deviation {oblique:.3f} of the kernel's largest weight")
    if window == "gaussian":
        assert oblique < 0.05


@pytest.mark.parametrize("m,window", [(1, "gaussian"), (2, "gaussian"), (2, "uniform"), (3, "uniform")])
def test_lpa_kernels_moment_conditions(m, window):
    """Direction 0: the weights on the axis are the 1-D LPA weights up to one
    factor, so sum_u k(u, 0) u^j = 0 for j = 1..min(m, h-1) (the fit
    reproduces polynomials of that degree along the direction); off the axis
    k(u, v) = k(u, 0) w(v) (w the window; 1 for uniform)."""
    h = [2, 3, 5, 8, 13, 21]
    g = lpa_ici.make_kernels(16, h, m, window)
    for s, hs in enumerate(h):
        k = g.kernel(0, s).astype(np.float64)
        r = hs - 1
        axis = k[r, r:]
        u = np.arange(hs, dtype=np.float64)
        scale = np.abs(axis).sum()
        for j in range(1, min(m, hs - 1) + 1):
            assert abs((axis * u ** j).sum()) <= 1e-6 * hs ** (j + 1) * 4, (hs, j)
        sig = 0.3 * hs
        for row in range(k.shape[0]):
            v = float(r - row)
            for col in range(r, k.shape[1]):
                if k[row, col] == 0 or v == 0:
                    continue
                wv = math.exp(-v * v / (2 * sig * sig)) if window == "gaussian" else 1.0
                assert abs(k[row, col] - axis[col - r] * wv) <= 2e-6 * max(1.0, scale), (hs, row, col)


@pytest.mark.parametrize("m", [0, 1, 2, 3])
def test_square_window_reproduces_polynomials(m):
    """Square window: sum_x k(x) u(x)^j = 1 for j = 0 and 0 for j = 1..min(m,
    h-1) over the quadrant, in every direction (u along the direction)."""
    h = [1, 2, 3, 6, 9, 16]
    g = lpa_ici.make_kernels(4, h, m, "square")
    for d in range(4):
        c, sn = round(math.cos(math.pi * d / 2)), round(math.sin(math.pi * d / 2))
        for s, hs in enumerate(h):
            k = g.kernel(d, s).astype(np.float64)
            r = hs - 1
            rows, cols = np.nonzero(k)  # within the quadrant (weights of 0 contribute nothing)
            assert rows.size <= hs * hs
            u = (cols - r) * c + (r - rows) * sn
            for j in range(0, min(m, hs - 1) + 1):
                mom = float((k[rows, cols] * u ** j).sum())
                assert abs(mom - (1.0 if j == 0 else 0.0)) <= 1e-6 * hs ** (j + 2), (d, hs, j, mom)


@pytest.mark.parametrize("D,m,window", [(16, 2, "gaussian"), (8, 1, "uniform"), (32, 1, "gaussian")])
def test_lpa_kernels_symmetry(D, m, window):
    """A quarter turn maps direction d to d + D/4 and a mirror in the x axis
    maps it to -d, weight for weight (within the files' last decimal)."""
    h = [1, 2, 5, 11, 23]
    g = lpa_ici.make_kernels(D, h, m, window)
    for s in range(len(h)):
        for d in range(D):
            k = g.kernel(d, s)
            # quarter turn counter-clockwise with y up: (x, y) -> (-y, x)
            assert np.abs(np.rot90(k) - g.kernel((d + D // 4) % D, s)).max() <= 1.5e-6
            assert np.abs(np.flipud(k) - g.kernel((-d) % D, s)).max() <= 1.5e-6


def test_order_zero_uniform_is_the_wedge_set():
    a = lpa_ici.make_kernels(16, H7, 0, "uniform")
    b = lpa_ici.wedge_kernels(16, H7)
    assert np.array_equal(a.weights.view(np.uint32), b.weights.view(np.uint32))


def test_lpa_kernel_errors():
    with pytest.raises(lpa_ici.LpaError):
        lpa_ici.make_kernels(8, [1, 2], 1, "square")  # square windows need D = 4
    with pytest.raises(lpa_ici.LpaError):
        lpa_ici.make_kernels(16, [1, 2], 5, "uniform")
    with pytest.raises(lpa_ici.LpaError):
        lpa_ici.make_kernels(16, [33], 1, "gaussian")
    with pytest.raises(lpa_ici.LpaError):
        lpa_ici.make_kernels(2, [4], 1, "gaussian")
