"""GPU sliding filters (piv_sliding_filter) against the oracle's
scipy.ndimage restatement of the reference's toolbox, bit for bit."""

import numpy as np
import pytest

from oracle import preproc as opp

pytestmark = pytest.mark.gpu


def _image(rng, dtype, shape):
    if dtype == np.uint8:
        return rng.integers(0, 256, shape).astype(np.uint8)
    if dtype == np.uint16:
        return rng.integers(0, 65536, shape).astype(np.uint16)
    return (rng.random(shape) * 4000 - 100).astype(np.float32)


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.float32])
@pytest.mark.parametrize("mode", ["reflect", "mirror", "nearest", "wrap", "constant"])
def test_sliding_filters_equal_the_oracle(dtype, mode):
    from fluidimage_amd.works.preproc import toolbox as tb

    print("case: sliding filters", dtype.__name__, mode, flush=True)
    rng = np.random.default_rng(31)
    img = _image(rng, dtype, (67, 150))
    for ws, weight in ((3, 1.0), (8, 0.5), ((5, 12), 0.8), (31, 1.0)):
        for name, kw in (("sliding_median", {}), ("sliding_minima", {}),
                         ("sliding_percentile", {"percentile": 10.0})):
            got = getattr(tb, name)(img, weight=weight, window_size=ws,
                                    boundary_condition=mode, **kw)
            exp = getattr(opp, name)(img, weight=weight, window_size=ws,
                                     boundary_condition=mode, **kw)
            assert got.dtype == exp.dtype, (name, ws)
            assert np.array_equal(got, exp), (name, ws, np.abs(got - exp).max())


def test_sliding_median_default_window_on_a_particle_image():
    """The reference's default (30 x 30, reflect) on a 12-bit particle image
    of the benchmark's kind, and several images in one launch (float64, as
    iterate_multiple_imgs returns them)."""
    from fluidimage_amd.synthetic import config_pair
    from fluidimage_amd.works.preproc import toolbox as tb

    im0, im1 = config_pair(2, shape=(384, 512))
    got = tb.sliding_median(im0)
    assert np.array_equal(got, opp.sliding_median(im0))
    both = tb.sliding_minima(np.stack([im0, im1]), weight=0.9)
    assert both.dtype == np.float64 and both.shape == (2, 384, 512)
    assert np.array_equal(both[1], opp.sliding_minima(im1, weight=0.9))
    lst = tb.sliding_percentile([im0, im1], percentile=25.0, window_size=(9, 17))
    assert np.array_equal(lst[0], opp.sliding_percentile(im0, 25.0, window_size=(9, 17)))
