"""Oracle: sliding-window preprocessing (test infrastructure).

Restates the python toolbox's spatial filters,
src/fluidimage/works/preproc/_toolbox_py.py:66-165 (sliding_median,
sliding_percentile, sliding_minima), on scipy.ndimage, the library the
reference calls, and gives a brute-force rank filter (explicit padding, a
sort per pixel) that pins scipy's conventions: the window offsets
-(w // 2) .. w - 1 - w // 2, the boundary modes and the rank of a median or
a percentile.
"""

import numpy as np
import scipy.ndimage as ndi


def sliding_median(img, weight=1.0, window_size=30, boundary_condition="reflect"):
    """_toolbox_py.py:66-94"""
    img_out = img - weight * ndi.median_filter(img, size=window_size, mode=boundary_condition)
    img_out[img_out < 0] = 0
    return img_out


def sliding_percentile(img, percentile=10.0, weight=1.0, window_size=30,
                       boundary_condition="reflect"):
    """_toolbox_py.py:97-131"""
    img_out = img - weight * ndi.percentile_filter(
        img, percentile, size=window_size, mode=boundary_condition)
    img_out[img_out < 0] = 0
    return img_out


def sliding_minima(img, weight=1.0, window_size=30, boundary_condition="reflect"):
    """_toolbox_py.py:134-160"""
    img_out = img - weight * ndi.minimum_filter(img, size=window_size, mode=boundary_condition)
    img_out[img_out < 0] = 0
    return img_out


# scipy.ndimage boundary modes as numpy.pad modes
PAD_MODES = {"reflect": "symmetric", "mirror": "reflect", "nearest": "edge", "wrap": "wrap",
             "constant": "constant"}


def brute_rank_filter(img, wy, wx, rank, mode="reflect", cval=0.0):
    """The rank-th smallest value of the wy x wx box at offsets
    -(w // 2) .. w - 1 - w // 2 around each pixel, the image extended by
    numpy.pad's equivalent of the scipy mode; small images only."""
    oy, ox = wy // 2, wx // 2
    pad = ((oy, wy - 1 - oy), (ox, wx - 1 - ox))
    kw = {"constant_values": cval} if mode == "constant" else {}
    p = np.pad(img, pad, mode=PAD_MODES[mode], **kw)
    h, w = img.shape
    out = np.empty_like(img)
    for y in range(h):
        for x in range(w):
            out[y, x] = np.sort(p[y:y + wy, x:x + wx], axis=None)[rank]
    return out
