"""Device-side tables for the C ABI: image descriptors and window records.

Images stay where the caller put them (torch tensors in HBM); only their
descriptors (``piv_image``) and the window records (``piv_window``) are
built here, as small device arrays.
"""

import contextlib
import ctypes

import numpy as np

from . import _native

IMAGE_DTYPE = np.dtype(
    [
        ("data", np.uint64),
        ("dtype", np.int32),
        ("height", np.int32),
        ("width", np.int32),
        ("reserved", np.int32),
        ("pitch", np.int64),
        ("minval", np.uint64),
    ],
    align=True,
)
assert IMAGE_DTYPE.itemsize == ctypes.sizeof(_native.PivImage)

WINDOW_FIELDS = ("image0", "image1", "y0", "x0", "y1", "x1", "dx_in", "dy_in", "flags",
                 "reserved")
assert len(WINDOW_FIELDS) * 4 == ctypes.sizeof(_native.PivWindow)

_TORCH_TO_PIV = {}


@contextlib.contextmanager
def on_stream(stream):
    """Run the body with ``stream`` as torch's current stream, after the
    work already queued on the caller's current stream: the library's
    launches, torch's allocations and copies of the body and its device-to-
    host reads are then all ordered on ``stream``. Results are ready on
    ``stream`` (the caller orders its own stream after it). No-op when
    ``stream`` is None or already current."""
    if stream is None:
        yield
        return
    import torch

    cur = torch.cuda.current_stream()
    if stream == cur:
        yield
        return
    stream.wait_stream(cur)
    with torch.cuda.stream(stream):
        yield


def piv_dtype_of(tensor):
    import torch

    table = {torch.uint8: _native.PIV_U8, torch.float32: _native.PIV_F32}
    if hasattr(torch, "uint16"):
        table[torch.uint16] = _native.PIV_U16
    if tensor.dtype in table:
        return table[tensor.dtype]
    raise TypeError(f"unsupported image dtype {tensor.dtype} (uint8, uint16, float32)")


def to_device_image(arr, device):
    """numpy/torch 2-D image -> contiguous device tensor (uint8/uint16/float32)."""
    import torch

    if isinstance(arr, torch.Tensor):
        t = arr
    else:
        a = np.ascontiguousarray(arr)
        if a.dtype == np.uint16:
            t = torch.from_numpy(a.view(np.int16)).view(torch.uint16)
        elif a.dtype in (np.uint8, np.float32):
            t = torch.from_numpy(a)
        elif np.issubdtype(a.dtype, np.integer) and a.min() >= 0 and a.max() < 65536:
            t = torch.from_numpy(a.astype(np.uint16).view(np.int16)).view(torch.uint16)
        else:
            t = torch.from_numpy(a.astype(np.float32))
    t = t.to(device)
    if t.dim() != 2:
        raise ValueError("images must be 2-D")
    if t.stride(1) != 1:
        t = t.contiguous()
    return t


class ImageSet:
    """A device table of image descriptors (+ their minimum, computed on the
    GPU, when ``subtract_min``)."""

    def __init__(self, images, subtract_min=True, stream=None):
        import torch

        self.images = list(images)
        if not self.images:
            raise ValueError("empty image set")
        device = self.images[0].device
        self.device = device
        dts = {piv_dtype_of(t) for t in self.images}
        if len(dts) != 1:
            raise TypeError("all images of a set must share one dtype")
        self.dtype = dts.pop()
        n = len(self.images)
        # the flat (vectorised) minimum when every image is contiguous and
        # 16-byte aligned
        flat = all(t.stride(1) == 1 and t.stride(0) == t.shape[1] and t.data_ptr() % 16 == 0
                   for t in self.images)
        self._flat_dtype = self.dtype if flat else -1
        self.mins = torch.empty(n, dtype=torch.float32, device=device)
        host = np.zeros(n, dtype=IMAGE_DTYPE)
        for i, t in enumerate(self.images):
            host[i]["data"] = t.data_ptr()
            host[i]["dtype"] = self.dtype
            host[i]["height"] = t.shape[0]
            host[i]["width"] = t.shape[1]
            host[i]["pitch"] = t.stride(0)
            host[i]["minval"] = self.mins.data_ptr() + 4 * i if subtract_min else 0
        # pinned staging and an asynchronous copy on the launch stream: a
        # pageable copy would make the host wait for all queued GPU work
        staged = torch.empty(host.nbytes, dtype=torch.uint8, pin_memory=True)
        staged.numpy()[:] = host.view(np.uint8)
        with torch.cuda.stream(stream if stream is not None else torch.cuda.current_stream()):
            self.table = staged.to(device, non_blocking=True)
        if subtract_min:
            lib = _native.load()
            _native.check(
                lib.piv_images_min(ctypes.c_void_p(self.table.data_ptr()), n,
                                   ctypes.c_void_p(self.mins.data_ptr()), self._flat_dtype,
                                   _native.stream_ptr(stream)),
                "piv_images_min",
            )

    def recompute_min(self, stream=None):
        lib = _native.load()
        _native.check(
            lib.piv_images_min(ctypes.c_void_p(self.table.data_ptr()), len(self.images),
                               ctypes.c_void_p(self.mins.data_ptr()), self._flat_dtype,
                               _native.stream_ptr(stream)),
            "piv_images_min",
        )


def window_array(image0, image1, y0, x0, y1, x1, dx_in=0, dy_in=0, flags=0):
    """Host (nwin, 10) int32 array of piv_window records (vectorised)."""
    cols = [image0, image1, y0, x0, y1, x1, dx_in, dy_in, flags, 0]
    n = max(np.size(c) for c in cols)
    out = np.zeros((n, 10), dtype=np.int32)
    for k, c in enumerate(cols):
        out[:, k] = c
    return out


def window_table(rows, device):
    """Device tensor of piv_window records from rows or an (n, 10) array."""
    import torch

    arr = np.asarray(rows, dtype=np.int32)
    if arr.ndim == 1:
        arr = arr[None]
    if arr.shape[1] == 9:
        arr = np.concatenate([arr, np.zeros((arr.shape[0], 1), np.int32)], axis=1)
    return torch.from_numpy(np.ascontiguousarray(arr)).to(device)
