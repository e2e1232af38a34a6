"""Host arithmetic of the batched TPS: the subdomain grids of many fields
computed at once equal, bit for bit, the oracle's per-field restatement of
thin_plate_spline_subdom.py:53-90 (np.linspace limits, rescaled smoothing,
20 % buffers) and the scalar host function."""

import numpy as np

from fluidimage_amd.calcul.tps import SingularSystemError, subdomain_grid, subdomain_grids


def _oracle_grid(rng, n_centers, x0, x1, y0, y1):
    from oracle.tps import ThinPlateSplineSubdom

    xs = np.concatenate([[x0, x1], rng.uniform(x0, x1, n_centers - 2)])
    ys = np.concatenate([[y0, y1], rng.uniform(y0, y1, n_centers - 2)])
    return ThinPlateSplineSubdom(np.vstack([ys, xs]), 200, 2.0)


def test_vectorised_grids_match_the_oracle_bitwise():
    rng = np.random.default_rng(3)
    cases = [(int(rng.integers(300, 3000)), *rng.uniform(0, 30, 1), *rng.uniform(200, 900, 1),
              *rng.uniform(0, 30, 1), *rng.uniform(200, 900, 1)) for _ in range(12)]
    counts = [c[0] for c in cases]
    x0 = [c[1] for c in cases]
    x1 = [c[2] for c in cases]
    y0 = [c[3] for c in cases]
    y1 = [c[4] for c in cases]
    grids = subdomain_grids(counts, x0, x1, y0, y1, 200, 2.0)
    for k, (n, a, b, c, d) in enumerate(cases):
        ref = _oracle_grid(rng, n, a, b, c, d)
        nsx, nsy, smoothing, lim = grids[k]
        assert (nsx, nsy) == (ref.nb_subdomx, ref.nb_subdomy)
        assert smoothing == ref.smoothing_coef
        expect = np.concatenate([ref.limits_min_x, ref.limits_max_x, ref.limits_min_y,
                                 ref.limits_max_y])
        assert np.array_equal(lim, expect)


def test_vectorised_grids_match_the_scalar_function():
    rng = np.random.default_rng(0)
    n = 200
    counts = rng.integers(100, 70000, n).astype(float)
    counts[:50] = 65025 - rng.integers(0, 800, 50)
    x0 = rng.uniform(0, 20, n)
    x1 = x0 + rng.uniform(100, 2100, n)
    y0 = rng.uniform(0, 20, n)
    y1 = y0 + rng.uniform(100, 2100, n)
    x0[:50], x1[:50], y0[:50], y1[:50] = 8, 2040, 8, 2040
    x1[7] = x0[7]  # degenerate field
    grids = subdomain_grids(counts, x0, x1, y0, y1, 200, 2.0)
    assert isinstance(grids[7], SingularSystemError)
    for k in range(n):
        if k == 7:
            continue
        a = subdomain_grid(counts[k], x0[k], x1[k], y0[k], y1[k], 200, 2.0)
        b = grids[k]
        assert a[:3] == b[:3] and np.array_equal(a[3], b[3]), k
