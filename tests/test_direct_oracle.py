"""Pins of the oracle's direct correlation (CorrelPythran / CorrelPyCuda,
src/fluidimage/calcul/correl.py:284-451, correl_pycuda.py:26-240): the sums
against scipy.signal.correlate2d (a library routine), the overlap areas
against their closed form, and the reference tests' known answers
(src/fluidimage/calcul/test_correl.py:45-238)."""

import numpy as np
import pytest
from scipy.signal import correlate2d

from oracle.correl import OracleCorrelDirect, correl_direct, default_displacement_max_direct
from tests.helpers import shifted_pair


@pytest.mark.parametrize("s0, s1, dmax", [((16, 16), (16, 16), 7), ((24, 32), (24, 32), 5),
                                          ((96, 48), (64, 32), 7)])
def test_direct_sums_match_scipy(s0, s1, dmax):
    rng = np.random.default_rng(sum(s0) + dmax)
    a = rng.random(s0)
    b = rng.random(s1)
    for kind in ("pythran", "pycuda"):
        m, norm = correl_direct(a, b, dmax, kind)
        z = correlate2d(a, b, "full")  # z[k] = sum_n a[n + k - (n_b - 1)] b[n]
        cy, cx = s0[0] // 2 - s1[0] // 2, s0[1] // 2 - s1[1] // 2
        for i in range(2 * dmax + 1):
            for j in range(2 * dmax + 1):
                dy, dx = i - dmax, j - dmax
                s = z[cy + dy + s1[0] - 1, cx + dx + s1[1] - 1]
                if kind == "pythran":
                    if s0 == s1:
                        area = (s1[0] - abs(dy)) * (s1[1] - abs(dx))
                    else:  # im1 inside im0 for |d| <= dmax: full overlap
                        area = s1[0] * s1[1]
                    s = s / area * b.size
                assert np.isclose(m[i, j], np.float32(s), rtol=2e-7, atol=0), (kind, i, j)
        expect = np.sqrt(np.sum(a**2) * np.sum(b**2)) * (a.size if kind == "pycuda" else 1)
        assert np.isclose(norm, expect, rtol=1e-14)


def test_default_displacement_max():
    assert default_displacement_max_direct((32, 32), (32, 32), "pythran") == 15
    assert default_displacement_max_direct((32, 32), (32, 32), "pycuda") == 16
    assert default_displacement_max_direct((96, 48), (64, 32), "pythran") == 7
    assert default_displacement_max_direct((96, 48), (64, 32), "pycuda") == 7


@pytest.mark.parametrize("kind", ["pythran", "pycuda"])
@pytest.mark.parametrize("shape", [(32, 32), (64, 32)])
def test_direct_known_answers(kind, shape):
    rng = np.random.default_rng(7)
    im0, im1 = shifted_pair(rng, shape, (3.3, 5.8))
    op = OracleCorrelDirect(shape, shape, kind=kind)
    c, norm = op(im0, im0)
    dx, dy, cmax, _ = op.compute_displacements_from_correl(c, norm=norm)
    assert (dx, dy) == (0, 0)
    if kind == "pythran":  # the zero-shift value is sum im0^2 = norm
        assert np.isclose(cmax, 1.0, rtol=1e-6)
    c, norm = op(im0, im1)
    dx, dy, cmax, _ = op.compute_displacements_from_correl(c, norm=norm)
    dx, dy = op.apply_subpix(dx, dy, c)
    assert np.allclose([dx, dy], [3.3, 5.8], atol=0.8)


def test_direct_different_sizes():
    """test_correl.py:192-238: im0 96 x 48, im1 64 x 32 (valid mode)."""
    rng = np.random.default_rng(8)
    big0, big1 = shifted_pair(rng, (96, 48), (3.3, 5.8))
    im1 = big1[16:80, 8:40].astype(np.float32)
    op = OracleCorrelDirect(big0.shape, im1.shape, kind="pythran")
    assert op.displacement_max == 7
    c, norm = op(big0, im1)
    dx, dy, _, _ = op.compute_displacements_from_correl(c, norm=norm)
    dx, dy = op.apply_subpix(dx, dy, c)
    assert np.allclose([dx, dy], [3.3, 5.8], atol=0.8)
