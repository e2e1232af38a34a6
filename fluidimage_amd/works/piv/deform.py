"""Image deformation between passes (an extension: the reference's later
passes shift the windows by the rounded predictor,
src/fluidimage/works/piv/singlepass.py:752-753,822-825).

With ``params.multipass.image_deformation = "bicubic"``, pass k > 0
resamples both images with the predictor field interpolated to every pixel
(``piv_deform_images``: bilinear predictor on the pass grid, Keys cubic
convolution, central differences im0(x - u/2), im1(x + u/2)), correlates the
windows centred on the grid nodes of the deformed images with the fused
kernel, and adds the predictor at the window centre to the residual. The
vector positions are the grid nodes themselves.
"""

import ctypes

import numpy as np

from ... import _native
from ...device import ImageSet

METHODS = (None, "bicubic")


def check_method(method):
    if method not in METHODS:
        raise ValueError(f"params.multipass.image_deformation must be one of {METHODS}")
    return method


def cell_tables(nodes, length):
    """For every pixel coordinate 0 <= p < length: the index j of the grid
    cell [nodes[j], nodes[j+1]] holding p (clamped to the nodes' range) and
    the fraction of p in it."""
    nodes = np.asarray(nodes, dtype=np.float64)
    p = np.arange(length, dtype=np.float64)
    if nodes.size == 1:
        return np.zeros(length, np.int32), np.zeros(length)
    q = np.clip(p, nodes[0], nodes[-1])
    j = np.clip(np.searchsorted(nodes, q, side="right") - 1, 0, nodes.size - 2)
    t = (q - nodes[j]) / (nodes[j + 1] - nodes[j])
    return j.astype(np.int32), t


class DeformedPass:
    """Device tables of one pass (one image shape) for the deformation."""

    def __init__(self, work, device):
        import torch

        self.work = work
        H, W = work.imshape0
        cy, ty = cell_tables(work.iyvecs, H)
        cx, tx = cell_tables(work.ixvecs, W)
        self.iyv, self.ixv = work._device_grid(device)
        self.cell_y = torch.from_numpy(cy).to(device)
        self.t_y = torch.from_numpy(ty).to(device)
        self.cell_x = torch.from_numpy(cx).to(device)
        self.t_x = torch.from_numpy(tx).to(device)
        self.grid = _native.PivDeformGrid(
            self.iyv.data_ptr(), self.ixv.data_ptr(), len(work.iyvecs), len(work.ixvecs),
            self.cell_y.data_ptr(), self.t_y.data_ptr(), self.cell_x.data_ptr(),
            self.t_x.data_ptr())

    def run(self, imageset, npairs, u, v, stream=None):
        """Deformed images of every pair (an ImageSet of 2 npairs float32
        images, pair k at 2k, 2k + 1); u, v: float64 device tensors
        (npairs x nvec, pair-major) of the predictor on this pass's grid."""
        import torch

        work = self.work
        H, W = work.imshape0
        dev = imageset.device
        # the deformed images live until the pass's correlation has read
        # them: the buffer and its image table are kept for the next call
        key = (2 * npairs, H, W, dev)
        if getattr(self, "_out_key", None) != key:
            self._out = torch.empty((2 * npairs, H, W), dtype=torch.float32, device=dev)
            self._out_set = ImageSet([self._out[i] for i in range(2 * npairs)],
                                     subtract_min=False, stream=stream)
            self._out_key = key
        out = self._out
        lib = _native.load()
        _native.check(
            lib.piv_deform_images(int(imageset.dtype), ctypes.c_void_p(imageset.table.data_ptr()),
                                  int(npairs), int(H), int(W), ctypes.byref(self.grid),
                                  ctypes.c_void_p(u.data_ptr()), ctypes.c_void_p(v.data_ptr()),
                                  ctypes.c_void_p(out.data_ptr()), _native.stream_ptr(stream)),
            "piv_deform_images")
        return self._out_set


def add_predictor(dx, dy, u, v, stream=None):
    lib = _native.load()
    _native.check(
        lib.piv_add_predictor(int(dx.numel()), ctypes.c_void_p(dx.data_ptr()),
                              ctypes.c_void_p(dy.data_ptr()), ctypes.c_void_p(u.data_ptr()),
                              ctypes.c_void_p(v.data_ptr()), _native.stream_ptr(stream)),
        "piv_add_predictor")
