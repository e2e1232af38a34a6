"""Image deformation between passes (an extension: the reference's later
passes shift the windows by the rounded predictor,
src/fluidimage/works/piv/singlepass.py:752-753,822-825).

With ``params.multipass.image_deformation = "bicubic"``, pass k > 0 builds
its interrogation windows by resampling both images with the predictor field
interpolated to every pixel (``piv_deform_windows``: bilinear predictor on
the pass grid, Keys cubic convolution, central differences im0(x - u/2),
im1(x + u/2)), correlates the deformed windows with the fused kernel, and
adds the predictor at the window centre to the residual. Windows are centred
on the grid nodes, so the vector positions are the nodes themselves.
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
        self._windows = {}

    def windows(self, npairs, device):
        """Window records reading window w of the two stacks (rows
        [w n0, (w + 1) n0) of stack images 0 and 1)."""
        import torch

        if npairs not in self._windows:
            n0 = self.work.shape_crop_im0[0]
            nwin = npairs * self.work.nb_vectors
            rec = np.zeros((nwin, 10), np.int32)
            rec[:, 1] = 1
            rec[:, 2] = rec[:, 4] = np.arange(nwin, dtype=np.int64) * n0
            self._windows[npairs] = torch.from_numpy(rec).to(device)
        return self._windows[npairs]

    def run(self, imageset, npairs, u, v, stream=None):
        """Deformed windows of every pair; returns (stack ImageSet, window
        records). u, v: float64 device tensors (npairs x nvec, pair-major)."""
        import torch

        work = self.work
        n0, n1 = work.shape_crop_im0
        if tuple(work.shape_crop_im1) != (n0, n1):
            raise ValueError("image deformation needs shape_crop_im1 == shape_crop_im0")
        nwin = npairs * work.nb_vectors
        dev = imageset.device
        stack0 = torch.empty((nwin * n0, n1), dtype=torch.float32, device=dev)
        stack1 = torch.empty_like(stack0)
        lib = _native.load()
        _native.check(
            lib.piv_deform_windows(int(imageset.dtype), ctypes.c_void_p(imageset.table.data_ptr()),
                                   int(npairs), ctypes.byref(self.grid),
                                   ctypes.c_void_p(u.data_ptr()), ctypes.c_void_p(v.data_ptr()),
                                   int(n0), int(n1), ctypes.c_void_p(stack0.data_ptr()),
                                   ctypes.c_void_p(stack1.data_ptr()), _native.stream_ptr(stream)),
            "piv_deform_windows")
        stacks = ImageSet([stack0, stack1], subtract_min=False)
        return stacks, self.windows(npairs, dev)


def add_predictor(dx, dy, u, v, stream=None):
    lib = _native.load()
    _native.check(
        lib.piv_add_predictor(int(dx.numel()), ctypes.c_void_p(dx.data_ptr()),
                              ctypes.c_void_p(dy.data_ptr()), ctypes.c_void_p(u.data_ptr()),
                              ctypes.c_void_p(v.data_ptr()), _native.stream_ptr(stream)),
        "piv_add_predictor")
