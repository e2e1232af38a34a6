"""Pins of the CPU oracle against closed forms, brute force, library routines
and the reference's own known-answer tests (CPU only)."""

import numpy as np
import pytest
from scipy import ndimage

from oracle import correl as oc
from oracle import fix as ofix
from oracle import subpix as osub
from oracle import tps as otps
from oracle.piv import OracleWorkPIVPass, default_params
from tests.helpers import shifted_pair


@pytest.mark.parametrize("shape", [(24, 32), (21, 32), (12, 13), (7, 9)])
def test_like_fftshift_closed_form(shape):
    # src/fluidimage/calcul/test_correl.py:233-245
    arr = np.arange(shape[0] * shape[1], dtype=np.float32).reshape(shape)
    expected = np.ascontiguousarray(np.fft.fftshift(arr[::-1, ::-1]))
    assert np.array_equal(oc.like_fftshift(arr), expected)


@pytest.mark.parametrize("shape", [(8, 8), (12, 10), (9, 7), (16, 16)])
def test_correl_fft_brute_force(shape):
    rng = np.random.default_rng(1)
    im0 = rng.random(shape) * 100
    im1 = rng.random(shape) * 100
    correl, norm = oc.correl_fft(im0, im1, "f64")
    brute = oc.brute_force_correl(im0, im1)
    assert np.allclose(correl, brute, rtol=1e-10, atol=1e-9 * np.abs(brute).max())
    if shape[1] % 2 == 0:
        # norm = n0 n1 sqrt(sum x0'^2 sum x1'^2) for even n1
        x0 = im0 - im0.mean()
        x1 = im1 - im1.mean()
        closed = shape[0] * shape[1] * np.sqrt((x0**2).sum() * (x1**2).sum())
        assert norm == pytest.approx(closed, rel=1e-12)


def test_correl_fft_f32_close_to_f64():
    rng = np.random.default_rng(2)
    im0 = rng.random((5, 32, 32)).astype(np.float32) * 4095
    im1 = rng.random((5, 32, 32)).astype(np.float32) * 4095
    c32, n32 = oc.correl_fft(im0, im1, "f32")
    c64, n64 = oc.correl_fft(im0, im1, "f64")
    assert c32.dtype == np.float32
    for k in range(5):
        assert np.abs(c32[k] - c64[k]).max() <= 2e-6 * np.abs(c64[k]).max()
    assert np.allclose(n32, n64, rtol=1e-6)


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_energy_from_fourier_closed_form(dtype):
    # src/fluidimage/calcul/test_fft.py:66-89
    rng = np.random.default_rng(3)
    n0, n1, coef_norm = 12, 8, 10
    field = (rng.random(n0 * n1) + 1j * rng.random(n0 * n1)).reshape(n0, n1).astype(dtype)
    expected = 0.5 / coef_norm * (
        np.sum(abs(field[:, 0]) ** 2 + abs(field[:, -1]) ** 2)
        + 2 * np.sum(abs(field[:, 1:-1]) ** 2)
    )
    assert oc.energy_from_fourier(field, coef_norm) == pytest.approx(expected, rel=1e-6)


def test_parseval_and_autocorrelation():
    # src/fluidimage/calcul/test_fft.py:8-34 (nx, ny = 20, 24)
    rng = np.random.default_rng(4)
    arr = rng.random((24, 20))
    coef = 24 * 20
    fft = np.fft.rfft2(arr)
    energy_k = oc.energy_from_fourier(fft, coef**2)
    energy_x = np.mean(np.abs(arr) ** 2) / 2
    assert energy_x == pytest.approx(energy_k, rel=1e-10)
    correl = np.fft.irfft2(fft.conj() * fft, s=arr.shape) * coef
    assert correl.max() / (2 * energy_k * coef**2) == pytest.approx(1.0, rel=1e-10)


def test_search_mask_counts():
    wld, ss0, ss1 = oc.search_mask((32, 32), 8)
    assert ss0 == (7, 24) and ss1 == (7, 24)
    # the disk of radius 8 around (iy0, ix0) = (15, 15)
    assert (~wld).sum() == sum(
        1 for a in range(-8, 9) for b in range(-8, 9) if a * a + b * b <= 64
    )
    assert oc.parse_displacement_max("50%", (32, 32)) == 16.0
    w2, s0, s1 = oc.search_mask((64, 32), 5.5)
    assert s0 == (31 - 5, 31 + 6) and s1 == (15 - 5, 15 + 6)


def test_nan_indices_max_matches_sequential_scan():
    rng = np.random.default_rng(5)
    for trial in range(300):
        n0, n1 = rng.integers(5, 14, 2)
        arr = rng.integers(0, 6, (n0, n1)).astype(np.float32)  # many ties
        arr[rng.random((n0, n1)) < 0.15] = np.nan
        a0 = int(rng.integers(0, n0 - 1))
        a1 = int(rng.integers(0, n1 - 1))
        b0 = int(rng.integers(a0 + 1, n0 + 1))
        b1 = int(rng.integers(a1 + 1, n1 + 1))
        assert oc.nan_indices_max(arr, a0, b0, a1, b1) == oc.nan_indices_max_loop(
            arr, a0, b0, a1, b1)


def test_tie_rule_last_maximum():
    arr = np.zeros((8, 8), np.float32)
    arr[3, 3] = arr[3, 5] = arr[4, 2] = 2.0
    assert oc.nan_indices_max(arr, 0, 8, 0, 8) == (4, 2, oc.NO_ERROR)
    arr[6, 6] = 2.0
    assert oc.nan_indices_max(arr, 0, 8, 0, 6)[:2] == (4, 2)


@pytest.mark.parametrize("shape", [(32, 32), (64, 32)])
def test_known_answers_reference_tests(shape):
    # src/fluidimage/calcul/test_correl.py:45-175
    rng = np.random.default_rng(6)
    im0, im1 = shifted_pair(rng, shape, (3.3, 5.8))
    for dmax in ("50%", None):
        c = oc.OracleCorrelFFT(shape, shape, displacement_max=dmax, method_subpix="2d_gaussian2")
        m, norm = c(im0, im0)
        dx, dy, cmax, _ = c.compute_displacements_from_correl(m, norm=norm)
        assert (dx, dy) == (0, 0)
        assert cmax == pytest.approx(1.0, rel=1e-5)
        m, norm = c(im0, im1)
        dx, dy, cmax, _ = c.compute_displacements_from_correl(m, norm=norm)
        assert np.allclose([dx, dy], [3.3, 5.8], atol=0.8)
        sx, sy = c.apply_subpix(dx, dy, m)
        assert np.allclose([sx, sy], [3.3, 5.8], atol=0.2)


def test_second_chance_and_wrapping_square():
    c = oc.OracleCorrelFFT((16, 16), displacement_max=None, particle_radius=3)
    m = np.zeros((16, 16), np.float32)
    m[0, 5] = 10.0  # on the border: first attempt fails
    m[8, 8] = 5.0
    # the square rows [0-3 : 4] wrap to an empty slice, so the second attempt
    # finds the same peak and the error carries the first peak's results
    with pytest.raises(oc.PIVError) as info:
        c.compute_displacements_from_correl(m, norm=1.0)
    assert info.value.explanation == "Correlation peak touching boundary."
    assert info.value.results == (c.ix0 - 5, c.iy0 - 0, np.float32(10.0))
    assert not np.isnan(m).any()
    m3 = np.zeros((16, 16), np.float32)
    m3[4, 15] = 10.0
    m3[8, 8] = 5.0
    dx, dy, cmax, _ = c.compute_displacements_from_correl(m3, norm=1.0)
    assert (dx, dy) == (c.ix0 - 8, c.iy0 - 8) and cmax == np.float32(5.0)
    # displacement_max None: the square is written into the caller's map
    assert np.isnan(m3[1:8, 12:16]).all() and np.isnan(m3).sum() == 7 * 4


def test_subpix_gaussian2_is_lstsq_stationary_point():
    rng = np.random.default_rng(7)
    x = np.array([-1.0, 0.0, 1.0])
    X, Y = np.meshgrid(x, x)
    A = np.stack([np.ones(9), X.ravel(), Y.ravel(), X.ravel() ** 2, Y.ravel() ** 2,
                  (X * Y).ravel()], axis=1)
    for _ in range(50):
        m = rng.random((5, 5)).astype(np.float32)
        m[2, 2] = 1.5
        dpx, dpy = osub.subpix_2d_gaussian2(m, 2, 2)
        crop = m[1:4, 1:4].astype(np.float64)
        crop = crop - crop.min()
        crop = crop / crop.max()
        crop[crop == 0] = 1e-8
        coef, *_ = np.linalg.lstsq(A, np.log(crop).ravel(), rcond=None)
        c0, cx, cy, cxx, cyy, cxy = coef
        sol = np.linalg.solve([[2 * cxx, cxy], [cxy, 2 * cyy]], [-cx, -cy])
        assert dpx == pytest.approx(sol[0], abs=1e-9)
        assert dpy == pytest.approx(sol[1], abs=1e-9)


def test_subpix_fallback_and_errors():
    sp = osub.SubPix("2d_gaussian2")
    m = np.zeros((7, 7), np.float32)
    with pytest.raises(osub.SubPixError):
        sp.compute_subpix(m, 0, 3)
    m = np.ones((7, 7), np.float32)
    m[3, 3] = 2.0
    ix, iy = sp.compute_subpix(m, 3, 3)
    assert (ix, iy) == (3.0, 3.0)
    # centroid direct sums
    m = np.zeros((5, 5), np.float32)
    m[1:4, 1:4] = [[0, 1, 0], [0, 2, 1], [0, 0, 0]]
    assert osub.subpix_centroid(m, 2, 2) == (pytest.approx(0.25), pytest.approx(-0.25))


def test_mean_neighbors_vs_ndimage():
    rng = np.random.default_rng(8)
    for shape in [(7, 9), (2, 5), (12, 12)]:
        dx = rng.normal(size=shape).astype(np.float32)
        dy = rng.normal(size=shape).astype(np.float32)
        dx[rng.random(shape) < 0.2] = np.nan
        mx, my = ofix.mean_neighbors_xy(dx.ravel(), dy.ravel(), *shape)
        ok = ~np.isnan(dx)
        k = np.ones((3, 3))
        k[1, 1] = 0
        s = ndimage.correlate(np.where(ok, dx, 0).astype(np.float64), k, mode="constant")
        n = ndimage.correlate(ok.astype(np.float64), k, mode="constant")
        with np.errstate(invalid="ignore", divide="ignore"):
            expected = np.where(ok & (n > 0), s / n, np.nan)
        assert np.allclose(mx.reshape(shape), expected, rtol=1e-5, equal_nan=True)


def test_tps_interpolates_and_reproduces_affine():
    rng = np.random.default_rng(9)
    centers = rng.uniform(0, 100, (2, 40))
    values = rng.normal(size=40)
    smoothed, w = otps.compute_tps_weights(centers, values, 0.0)
    assert np.allclose(smoothed, values, atol=1e-8)
    affine = 0.3 + 0.02 * centers[0] - 0.05 * centers[1]
    for s in (0.0, 2.0, 30.0):
        smoothed, w = otps.compute_tps_weights(centers, affine, s)
        assert np.allclose(smoothed, affine, atol=1e-8)
        new = rng.uniform(0, 100, (2, 15))
        interp = w @ otps.compute_tps_matrix(new, centers)
        assert np.allclose(interp, 0.3 + 0.02 * new[0] - 0.05 * new[1], atol=1e-7)


def test_tps_subdom_blends_constants():
    ys, xs = np.mgrid[0:300:8, 0:400:8]
    centers = np.vstack([ys.ravel(), xs.ravel()]).astype(float)
    tps = otps.ThinPlateSplineSubdom(centers, 200, 2.0, threshold=1.5)
    assert tps.nb_subdom > 1
    values = np.full(centers.shape[1], 1.25)
    smoothed, weights, summary = tps.compute_tps_weights_subdom(values)
    assert np.allclose(smoothed, 1.25, atol=1e-9)
    tps.init_with_new_positions(np.vstack([ys.ravel()[::3] + 3.0, xs.ravel()[::3] + 1.0]))
    assert np.allclose(tps.interpolate(weights), 1.25, atol=1e-8)


@pytest.mark.parametrize(
    "imshape, window, expected",
    [((512, 512), 48, 400), ((2048, 2048), 32, 16129), ((2048, 2048), 64, 3969),
     ((2048, 2048), 16, 65025), ((1200, 1500), 40, 59 * 74)],
)
def test_grid_counts(imshape, window, expected):
    p = default_params()
    p["piv0"]["shape_crop_im0"] = window
    w = OracleWorkPIVPass(p, 0)
    w.prepare_with_image(imshape)
    assert len(w.ixvecs_grid) == expected


def test_positions_from_displacement_fallbacks():
    p = default_params()
    p["piv0"]["shape_crop_im0"] = 16
    w = OracleWorkPIVPass(p, 1)
    w.prepare_with_image((64, 64))
    w.npad = 8
    n = len(w.ixvecs_grid)
    dx = np.zeros(n, np.int32)
    dy = np.zeros(n, np.int32)
    k_low = 0  # grid point (8, 8)
    dx[k_low] = -20  # ixs1 = 8 - (-10) - 20 = -2 < 0 -> case A
    dy[k_low] = 3
    k_hi = n - 1  # grid point (56, 56)
    dx[k_hi] = 11  # ixs0 = 56 - 5 = 51, ixs1 = 62: inside
    xs, ys, ixs0, iys0, ixs1, iys1 = w.positions(dx, dy)
    gx, gy = w.ixvecs_grid[k_low], w.iyvecs_grid[k_low]
    assert ixs0[k_low] - 8 == gx + 20 and ixs1[k_low] - 8 == gx
    assert iys0[k_low] - 8 == gy - 3 and iys1[k_low] - 8 == gy
    assert xs[k_low] == (gx + 20 + gx) / 2
    assert ixs0[k_hi] - 8 == 56 - 5 and ixs1[k_hi] - 8 == 62
    # floor division of negative predictors
    dx2 = np.full(n, -3, np.int32)
    _, _, ixs0, _, ixs1, _ = w.positions(dx2, np.zeros(n, np.int32))
    mid = n // 2
    assert ixs0[mid] - 8 == w.ixvecs_grid[mid] + 2 and ixs1[mid] - 8 == w.ixvecs_grid[mid] - 1


# ---- image deformation (oracle/deform.py): closed-form pins -------------------------------


def test_keys_weights_partition_and_quadratic_reproduction():
    from oracle.deform import keys_weights

    t = np.random.default_rng(0).random(200)
    w = keys_weights(t)
    assert np.allclose(w.sum(-1), 1.0, atol=1e-14)
    taps = np.array([-1.0, 0.0, 1.0, 2.0])
    for coefs in [(1.0, 0.0, 0.0), (0.3, -1.7, 0.0), (2.0, 0.5, -0.25)]:
        poly = np.polynomial.Polynomial(coefs)
        assert np.allclose((w * poly(taps)[None, :]).sum(-1), poly(t), atol=1e-12)
    # a cubic is not reproduced (Keys a = -0.5 is third-order accurate, not exact)
    cubic = np.polynomial.Polynomial((0, 0, 0, 1.0))
    assert np.abs((w * cubic(taps)[None, :]).sum(-1) - cubic(t)).max() > 1e-3


def test_bicubic_sample_integer_positions_and_quadratic_image():
    from oracle.deform import bicubic_sample

    rng = np.random.default_rng(1)
    img = rng.random((20, 24)) * 100
    yy, xx = np.mgrid[0:20, 0:24]
    assert np.array_equal(bicubic_sample(img, yy.astype(float), xx.astype(float)), img)
    quad = 3.0 + 0.5 * yy - 0.2 * xx + 0.01 * yy * xx + 0.03 * yy**2 - 0.02 * xx**2
    py = rng.uniform(2, 16, 300)
    px = rng.uniform(2, 20, 300)
    expect = 3.0 + 0.5 * py - 0.2 * px + 0.01 * py * px + 0.03 * py**2 - 0.02 * px**2
    assert np.allclose(bicubic_sample(quad, py, px), expect, atol=1e-10)
    # zero outside the image
    assert bicubic_sample(img, np.array([-5.0]), np.array([3.0]))[0] == 0.0


def test_bilinear_field_reproduces_affine():
    from oracle.deform import bilinear_field

    iy = np.array([8, 16, 24, 32, 37])
    ix = np.array([8, 16, 24, 31])
    gy, gx = np.meshgrid(iy, ix, indexing="ij")
    u = 0.5 + 0.1 * gx - 0.05 * gy
    rng = np.random.default_rng(2)
    y = rng.uniform(8, 37, 100)
    x = rng.uniform(8, 31, 100)
    assert np.allclose(bilinear_field(iy, ix, u, y, x), 0.5 + 0.1 * x - 0.05 * y, atol=1e-12)
    # clamped (constant) outside the grid box
    assert np.isclose(bilinear_field(iy, ix, u, np.array([0.0]), np.array([0.0]))[0], u[0, 0])


def test_deformed_crops_zero_and_even_integer_fields():
    from oracle.deform import deformed_crops

    rng = np.random.default_rng(3)
    im0 = rng.integers(0, 4000, (64, 72)).astype(np.uint16)
    im1 = rng.integers(0, 4000, (64, 72)).astype(np.uint16)
    iy = np.array([16, 24, 32, 40])
    ix = np.array([16, 24, 32, 40, 48])
    n = (16, 16)
    zero = np.zeros((4, 5))
    c0, c1 = deformed_crops(im0, im1, iy, ix, zero, zero, n)
    a0 = im0.astype(float) - im0.min()
    a1 = im1.astype(float) - im1.min()
    k = 0
    for y in iy:
        for x in ix:
            assert np.array_equal(c0[k], a0[y - 8:y + 8, x - 8:x + 8])
            assert np.array_equal(c1[k], a1[y - 8:y + 8, x - 8:x + 8])
            k += 1
    # u = 2, v = -2: im0 sampled at (y + 1, x - 1), im1 at (y - 1, x + 1)
    c0, c1 = deformed_crops(im0, im1, iy, ix, zero + 2.0, zero - 2.0, n)
    k = 0
    for y in iy:
        for x in ix:
            assert np.allclose(c0[k], a0[y - 7:y + 9, x - 9:x + 7], atol=1e-9)
            assert np.allclose(c1[k], a1[y - 9:y + 7, x - 7:x + 9], atol=1e-9)
            k += 1


# ---- refinement loss (oracle/refine.py): pins ----------------------------------------------


def test_refine_loss_at_zero_is_one_minus_ncc_of_crops():
    import torch

    from oracle.refine import refine_loss

    rng = np.random.default_rng(11)
    im0 = rng.integers(0, 4000, (48, 56)).astype(np.uint16)
    im1 = rng.integers(0, 4000, (48, 56)).astype(np.uint16)
    iy = np.array([8, 16, 24, 32, 40])
    ix = np.array([8, 16, 24, 32, 40, 48])
    z = torch.zeros(iy.size * ix.size, dtype=torch.float64)
    got = float(refine_loss(im0, im1, iy, ix, (16, 16), z, z))
    a0 = im0.astype(float) - im0.min()
    a1 = im1.astype(float) - im1.min()
    expect = 0.0
    for y in iy:
        for x in ix:
            a = a0[y - 8:y + 8, x - 8:x + 8].ravel()
            b = a1[y - 8:y + 8, x - 8:x + 8].ravel()
            expect += 1.0 - np.corrcoef(a, b)[0, 1]
    assert np.isclose(got, expect, rtol=1e-12)


def test_refine_loss_autograd_matches_central_differences():
    import torch

    from fluidimage_amd.synthetic import config_pair
    from oracle.refine import refine_loss

    im0, im1 = config_pair(3, shape=(72, 80))
    iy = np.array([8, 16, 24, 32, 40, 48, 56, 63])
    ix = np.array([8, 16, 24, 32, 40, 48, 56, 64, 71])
    rng = np.random.default_rng(5)
    d0 = rng.normal(0, 1.0, (2, iy.size * ix.size))
    x = torch.tensor(d0[0], requires_grad=True)
    y = torch.tensor(d0[1], requires_grad=True)
    loss = refine_loss(im0, im1, iy, ix, (16, 16), x, y)
    gx, gy = torch.autograd.grad(loss, (x, y))
    h = 1e-3
    for k in rng.choice(iy.size * ix.size, 6, replace=False):
        for comp, g in ((0, gx), (1, gy)):
            e = np.zeros_like(d0)
            e[comp, k] = h
            lp = float(refine_loss(im0, im1, iy, ix, (16, 16), torch.tensor(d0[0] + e[0]),
                                   torch.tensor(d0[1] + e[1])))
            lm = float(refine_loss(im0, im1, iy, ix, (16, 16), torch.tensor(d0[0] - e[0]),
                                   torch.tensor(d0[1] - e[1])))
            fd = (lp - lm) / (2 * h)
            assert abs(fd - float(g[k])) <= 1e-4 * max(1.0, abs(fd)), (k, comp, fd, float(g[k]))


def test_subpix_gaussian_recovers_an_exact_gaussian():
    """A sampled Gaussian has an exactly quadratic log: the '2d_gaussian' fit
    (subpix.py:157-181) returns its centre, for every nsubpix."""
    rng = np.random.default_rng(21)
    for n in (1, 2, 3):
        for _ in range(20):
            x0, y0 = rng.uniform(-0.45, 0.45, 2)
            sx, sy = rng.uniform(0.8, 2.5, 2)
            size = 2 * n + 5
            c = size // 2
            yy, xx = np.mgrid[0:size, 0:size].astype(np.float64)
            m = np.exp(-((xx - c - x0) ** 2) / (2 * sx**2) - ((yy - c - y0) ** 2) / (2 * sy**2))
            d = osub.subpix_2d_gaussian(m.astype(np.float32), c, c, n)
            assert d is not None
            assert d[0] == pytest.approx(x0, abs=2e-5)
            assert d[1] == pytest.approx(y0, abs=2e-5)


def test_subpix_gaussian_pinv_is_the_decoupled_least_squares():
    """The reference's pseudo-inverse of [X^2, Y^2, X, Y, 1] equals the
    decoupled normal-equation solution of the symmetric grid (the form the
    device code evaluates)."""
    for n in (1, 2, 4):
        P = osub.gaussian_fit_matrix(n)
        N = 2 * n + 1
        xs = np.arange(-n, n + 1, dtype=float)
        S2 = (xs**2).sum()
        S4 = (xs**4).sum()
        X, Y = np.meshgrid(xs, xs)
        X, Y = X.ravel(), Y.ravel()
        den = N * S4 - S2**2
        rows = np.stack([(X**2 - S2 / N) / den, (Y**2 - S2 / N) / den,
                         X / (N * S2), Y / (N * S2)])
        assert np.allclose(P[:4], rows, atol=1e-12)


def test_subpix_gaussian_fallbacks_and_errors():
    sp = osub.SubPix("2d_gaussian", nsubpix=1)
    # a valley (positive curvature): the centroid is taken
    m = np.full((5, 5), 2.0, np.float32)
    m[2, 2] = 1.0
    m[2, 3] = 1.5
    assert osub.subpix_2d_gaussian(m, 2, 2, 1) is None
    assert sp.compute_subpix(m, 2, 2) == pytest.approx(
        tuple(np.array(osub.subpix_centroid(m, 2, 2, 1)) + 2))
    # values <= 0 enter the fit as 1e-6
    m = np.zeros((5, 5), np.float32)
    m[2, 2] = 1.0
    m[2, 3] = 0.5
    d = osub.subpix_2d_gaussian(m, 2, 2, 1)
    l = np.log(np.float32(1e-6))
    logs = np.full((3, 3), l, np.float64)
    logs[1, 1] = 0.0
    logs[1, 2] = np.log(np.float32(0.5))
    xs = np.array([-1.0, 0.0, 1.0])
    X, Y = np.meshgrid(xs, xs)
    a = ((X**2 - 2 / 3) * logs).sum() / 2
    c = (X * logs).sum() / 6
    assert d[0] == pytest.approx(c / (-2 * a), rel=1e-12)
    # close boundary, as every method
    with pytest.raises(osub.SubPixError, match="close boundary"):
        sp.compute_subpix(m, 0, 2)
    # a fit whose displacement exceeds 2 (0.5 + n)^2 is an error (no fallback)
    yy, xx = np.mgrid[-3:4, -3:4].astype(np.float64)
    m = np.exp(-0.01 * xx**2 - yy**2 + 0.1 * xx).astype(np.float32)  # vertex at x = 5
    d = osub.subpix_2d_gaussian(m, 3, 3, 1)
    assert d is not None and d[0] ** 2 + d[1] ** 2 > 4.5
    with pytest.raises(osub.SubPixError, match="wrong subpix"):
        sp.compute_subpix(m, 3, 3)
