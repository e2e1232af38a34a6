"""Pins of the preprocessing oracle (oracle/preproc.py): scipy.ndimage's
filters, as the reference calls them, against a brute-force rank filter
(window offsets, boundary modes, the rank of a median and of a percentile),
and the output types of the reference's numpy expression."""

import numpy as np
import pytest
import scipy.ndimage as ndi

from oracle import preproc as opp


@pytest.mark.parametrize("mode", ["reflect", "mirror", "nearest", "wrap", "constant"])
@pytest.mark.parametrize("wy,wx", [(3, 3), (4, 4), (2, 5), (6, 3)])
def test_scipy_filters_equal_the_brute_force(mode, wy, wx):
    rng = np.random.default_rng(wy * 10 + wx)
    img = rng.integers(0, 4096, (11, 13)).astype(np.uint16)
    size = wy * wx
    got = ndi.median_filter(img, size=(wy, wx), mode=mode)
    assert np.array_equal(got, opp.brute_rank_filter(img, wy, wx, size // 2, mode))
    got = ndi.minimum_filter(img, size=(wy, wx), mode=mode)
    assert np.array_equal(got, opp.brute_rank_filter(img, wy, wx, 0, mode))
    for p in (10.0, 37.5, 100.0):
        rank = size - 1 if p == 100.0 else int(size * p / 100.0)
        got = ndi.percentile_filter(img, p, size=(wy, wx), mode=mode)
        assert np.array_equal(got, opp.brute_rank_filter(img, wy, wx, rank, mode))


def test_filter_rank_follows_scipy():
    from fluidimage_amd.works.preproc.toolbox import filter_rank

    assert filter_rank("median", 30, 30) == 450
    assert filter_rank("median", 3, 3) == 4
    assert filter_rank("minimum", 5, 7) == 0
    assert filter_rank("percentile", 30, 30, 10.0) == 90
    assert filter_rank("percentile", 30, 30, -90.0) == 90
    assert filter_rank("percentile", 4, 4, 100.0) == 15
    with pytest.raises(RuntimeError):
        filter_rank("percentile", 3, 3, 120.0)


def test_output_types_of_the_reference_expression():
    img16 = np.arange(64, dtype=np.uint16).reshape(8, 8)
    out = opp.sliding_minima(img16, weight=0.5, window_size=3)
    assert out.dtype == np.float64 and out.min() >= 0
    img32 = img16.astype(np.float32)
    out = opp.sliding_median(img32, weight=0.7, window_size=3)
    assert out.dtype == np.float32
    # negatives are clipped to 0: a pixel below its window median
    img = np.full((5, 5), 10, np.uint8)
    img[2, 2] = 3
    out = opp.sliding_median(img, window_size=3)
    assert out[2, 2] == 0 and out[0, 0] == 0
