"""Kernel-set generation (uniform wedges) and the kernel-file writer.

Pins: the shipped default set (its loaded weights, offsets and sizes, which
the generator must give bit for bit), the file round trip, and what the
geometry fixes for any direction count: quarter-turn and mirror symmetry of
the supports, nesting across scales, the closed-form tap count of direction
0, and uniform weights that sum to one within the files' six decimals.
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
