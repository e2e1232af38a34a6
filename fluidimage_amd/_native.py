"""ctypes binding of libpiv_hip.so (the C ABI declared in include/piv_hip.h).

The library is built in-tree (``fluidimage_amd/libpiv_hip.so``) by
``fluidimage_amd.build``. There is no fallback: if the library cannot be
loaded every compute entry point raises ``NativeLibraryError``.
"""

import ctypes
import os
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libpiv_hip.so"
ARCH = "gfx950"
# launch tuning table written by ``python -m fluidimage_amd.tuning --write``
# on the target GPU; applied at load (PIV_HIP_TUNING: another file, or "none")
TUNING_PATH = HERE / f"tuning_{ARCH}.json"

PIV_OK = 0
PIV_U8, PIV_U16, PIV_F32 = 0, 1, 2
PIV_SUBPIX = {"2d_gaussian2": 0, "centroid": 1, "no_subpix": 2, "2d_gaussian": 3}
PIV_WINFLAG_INVALID = 1
PIV_OPT_SUBPIX = 1
PIV_OPT_MAP_SQUARE = 2
PIV_DIRECT_PYTHRAN, PIV_DIRECT_PYCUDA = 0, 1
# scipy.ndimage boundary modes of the sliding filters
PIV_BOUNDARY = {"reflect": 0, "constant": 1, "nearest": 2, "mirror": 3, "wrap": 4}
PIV_TPS_PART_PREP, PIV_TPS_PART_SHARED, PIV_TPS_PART_DIRECT, PIV_TPS_PART_ALL = 1, 2, 4, 7

WIN_OK = 0
WIN_PEAK_BOUNDARY = 1
WIN_PEAK_NAN = 2
WIN_SUBPIX_BOUNDARY = 3
WIN_SUBPIX_WRONG = 4
WIN_BAD_CROP = 5
WIN_NO_PEAK = 6

EXPLANATIONS = {
    WIN_PEAK_BOUNDARY: "Correlation peak touching boundary.",
    WIN_PEAK_NAN: "Correlation peak touching nan.",
    WIN_SUBPIX_BOUNDARY: "close boundary",
    WIN_SUBPIX_WRONG: "wrong subpix",
    WIN_BAD_CROP: "Bad im_crop shape.",
    WIN_NO_PEAK: "No valid correlation value in the search box.",
}


class NativeLibraryError(RuntimeError):
    pass


class PivImage(ctypes.Structure):
    _fields_ = [
        ("data", ctypes.c_void_p),
        ("dtype", ctypes.c_int32),
        ("height", ctypes.c_int32),
        ("width", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("pitch", ctypes.c_int64),
        ("minval", ctypes.c_void_p),
    ]


class PivWindow(ctypes.Structure):
    _fields_ = [
        ("image0", ctypes.c_int32),
        ("image1", ctypes.c_int32),
        ("y0", ctypes.c_int32),
        ("x0", ctypes.c_int32),
        ("y1", ctypes.c_int32),
        ("x1", ctypes.c_int32),
        ("dx_in", ctypes.c_int32),
        ("dy_in", ctypes.c_int32),
        ("flags", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class PivResults(ctypes.Structure):
    _fields_ = [
        ("dx", ctypes.c_void_p),
        ("dy", ctypes.c_void_p),
        ("correl_max", ctypes.c_void_p),
        ("err", ctypes.c_void_p),
        ("ix", ctypes.c_void_p),
        ("iy", ctypes.c_void_p),
        ("norm", ctypes.c_void_p),
        ("sec", ctypes.c_void_p),
        ("nsec", ctypes.c_void_p),
        ("sec_sub", ctypes.c_void_p),
    ]


class PivDeformGrid(ctypes.Structure):
    _fields_ = [
        ("iyvecs", ctypes.c_void_p),
        ("ixvecs", ctypes.c_void_p),
        ("ny", ctypes.c_int32),
        ("nx", ctypes.c_int32),
        ("cell_y", ctypes.c_void_p),
        ("t_y", ctypes.c_void_p),
        ("cell_x", ctypes.c_void_p),
        ("t_x", ctypes.c_void_p),
    ]


class PivTpsLattice(ctypes.Structure):
    _fields_ = [
        ("table", ctypes.c_void_p),
        ("step", ctypes.c_double),
        ("a_max", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


# exported symbol -> (restype, argtypes); mirrors include/piv_hip.h
_P = ctypes.c_void_p
_I = ctypes.c_int32
SIGNATURES = {
    "piv_version": (ctypes.c_char_p, []),
    "piv_last_error": (ctypes.c_char_p, []),
    "piv_shape_supported": (ctypes.c_int, [_I, _I]),
    "piv_plan_create": (
        ctypes.c_int,
        [ctypes.POINTER(_P), _I, _I, ctypes.c_double, _P, _P, _I, _I, _I, _I],
    ),
    "piv_plan_destroy": (ctypes.c_int, [_P]),
    "piv_graph_begin": (ctypes.c_int, [_P]),
    "piv_graph_end": (ctypes.c_int, [_P, ctypes.POINTER(_P)]),
    "piv_graph_launch": (ctypes.c_int, [_P, _P]),
    "piv_graph_destroy": (ctypes.c_int, [_P]),
    "piv_images_min": (ctypes.c_int, [_P, _I, _P, _I, _P]),
    "piv_sliding_filter": (
        ctypes.c_int,
        [_P, _I, _I, _I, _I, _I, _I, _I, _I, ctypes.c_double, ctypes.c_double, _P, _P]),
    "piv_correlate_batch": (
        ctypes.c_int,
        [_P, _I, _P, _P, _I, _I, _P, ctypes.POINTER(PivResults), _P],
    ),
    "piv_peaks_from_maps": (ctypes.c_int, [_P, _P, _P, _I, _I, ctypes.POINTER(PivResults), _P]),
    "piv_subpix_at": (ctypes.c_int, [_P, _P, _P, _P, _I, _P, _P, _P, _P]),
    "piv_correlate_direct": (ctypes.c_int, [_I, _P, _I, _I, _P, _I, _I, _I, _I, _P, _P, _P]),
    "piv_peaks_any": (
        ctypes.c_int, [_I, _P, _P, _I, _I, _I, _I, _I, _I, _I, _I, _I,
                       ctypes.POINTER(PivResults), _P]),
    "piv_subpix_any": (ctypes.c_int, [_I, _P, _I, _I, _P, _P, _I, _I, _P, _P, _P, _P]),
    "piv_make_windows": (
        ctypes.c_int,
        [_I, _I, _I, _P, _P, _P, _P, _I, _I, _I, _P, _I, _I, _P, _P, _P, _P, _P, _P],
    ),
    "piv_tps_smooth": (
        ctypes.c_int,
        [_I, _I, _P, _P, _P, _P, _P, _P, _P, _P, ctypes.c_double, _P, _P, _P, _P, _P,
         ctypes.c_int64, _P, ctypes.c_int64, _P, _P, _P, _P, _I, _P],
    ),
    "piv_tps_phi_table": (ctypes.c_int, [ctypes.c_double, _I, _P, _P]),
    "piv_tune_candidates": (ctypes.c_int, [_I, _I, _I, _P, _I]),
    "piv_tune_get": (ctypes.c_int, [_I, _I, _I]),
    "piv_tune_set": (ctypes.c_int, [_I, _I, _I, _I]),
    "piv_tps_select": (
        ctypes.c_int, [_I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, ctypes.c_double, _P]),
    "piv_tps_count": (ctypes.c_int, [_I, _P, _P, _P, _P, _P, _P, _P]),
    "piv_tps_fill": (ctypes.c_int, [_I, _P, _P, _P, _P, _P, _P, _P, _P]),
    "piv_tps_point_slots": (
        ctypes.c_int, [_I, _P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "piv_tps_blend": (ctypes.c_int, [_I, _P, _P, _P, _P, ctypes.c_int64, _P, _P]),
    "piv_tps_influence": (ctypes.c_int, [_I, _I, _P, _P, _P, _P, _P, _P, _P, ctypes.c_int64, _P]),
    "piv_tps_node_field": (
        ctypes.c_int,
        [_I, _I, _I, _I, _I, _P, _P, ctypes.c_int64, _P, _P, _P, ctypes.c_int64, _P, _P, _P,
         ctypes.c_int64, _P, _P, _P, _P]),
    "piv_tps_grid": (
        ctypes.c_int,
        [_I, _I, _P, ctypes.c_double, ctypes.c_double, ctypes.c_double, _I, ctypes.c_double, _I,
         _P, _P, _P, _P, _P, _P]),
    "piv_tps_offsets": (
        ctypes.c_int,
        [_I, _I, _P, _P, _P, ctypes.c_int64, ctypes.c_int64, _I, _P, _P, _P, _P, _P, _P]),
    "piv_tps_at_centers": (ctypes.c_int, [_I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "piv_tps_shared_workspace": (ctypes.c_int64, [_I, _I, _I, _I, _I]),
    "piv_tps_smooth_shared": (
        ctypes.c_int,
        [_I, _I, _P, _P, _P, _P, _P, _P, _I, _I, _P, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P,
         _P, _P, _P, ctypes.c_double, _P, _P, _P, _P, _P, ctypes.c_int64, _P, ctypes.c_int64,
         _P, _P, _P, _P, _I, _I, _P, _P]),
    "piv_tps_interpolate": (
        ctypes.c_int,
        [_I, _P, _P, _P, _P, _P, _P, ctypes.c_int64, _I, _I, _P, _P, ctypes.c_int64, _P, _P,
         _P, _P, _P],
    ),
    "piv_tps_interpolate_noalloc": (
        ctypes.c_int,
        [_I, _P, _P, _P, _P, _P, _P, ctypes.c_int64, _I, _I, _P, _P, ctypes.c_int64, _P, _P,
         _P, _P],
    ),
    "piv_tps_matrix_size": (ctypes.c_int64, [_I]),
    "piv_tps_interpolate_workspace": (ctypes.c_int64, [_I, _I]
    ),
    "piv_deform_images": (
        ctypes.c_int, [_I, _P, _I, _I, _I, ctypes.POINTER(PivDeformGrid), _P, _P, _P, _P]),
    "piv_add_predictor": (ctypes.c_int, [_I, _P, _P, _P, _P, _P]),
    "piv_refine_forward": (
        ctypes.c_int,
        [_I, _P, _I, ctypes.POINTER(PivDeformGrid), _I, _I, _P, _P, _P, _P, _P]),
    "piv_refine_backward": (
        ctypes.c_int,
        [_I, _P, _I, ctypes.POINTER(PivDeformGrid), _I, _I, _P, _P, _P, _P, _P, _P, _P]),
    "piv_fix": (
        ctypes.c_int,
        [_I, _P, _P, _P, _P, _I, _I, ctypes.c_float, ctypes.c_float, ctypes.c_float, _P, _P, _P,
         _P, _P],
    ),
}

_lib = None


def load(path=None):
    """Load (once) and return the ctypes library handle."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path is not None else Path(os.environ.get("PIV_HIP_LIB", LIB_PATH))
    if not p.exists():
        raise NativeLibraryError(
            f"{p} not found: build it with `python -m fluidimage_amd.build` "
            "(the HIP library is required; there is no CPU path)"
        )
    # torch first: its HIP runtime (torch/lib/libamdhip64.so, SONAME
    # libamdhip64.so.7) must be the process's. Loaded after this library,
    # which needs libamdhip64.so.7, torch would bring a second runtime
    # (it asks for libamdhip64.so by another name) and one of the two would
    # find no device.
    import torch  # noqa: F401

    try:
        lib = ctypes.CDLL(str(p))
    except OSError as exc:
        raise NativeLibraryError(f"cannot load {p}: {exc}") from exc
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _apply_tuning(lib)
    if path is None:
        _lib = lib
    return lib


def _apply_tuning(lib):
    """Launch parameters of the tuning table (JSON data), if there is one."""
    import json

    choice = os.environ.get("PIV_HIP_TUNING", str(TUNING_PATH))
    if choice == "none" or not Path(choice).exists():
        return
    table = json.loads(Path(choice).read_text())
    for key, wpb in table.get("correlate_wpb", {}).items():
        n0, n1 = (int(v) for v in key.split("x"))
        check(lib.piv_tune_set(1, n0, n1, int(wpb)), f"tuning table {choice}: {key}", lib)
    if "tps_chol_threads" in table:
        check(lib.piv_tune_set(2, 0, 0, int(table["tps_chol_threads"])),
              f"tuning table {choice}: tps_chol_threads", lib)
    if "tps_shared_solve" in table:
        check(lib.piv_tune_set(3, 0, 0, int(table["tps_shared_solve"])),
              f"tuning table {choice}: tps_shared_solve", lib)


def check(status, what="", lib=None):
    if status != PIV_OK:
        msg = (lib if lib is not None else load()).piv_last_error().decode()
        raise RuntimeError(f"{what} failed (status {status}): {msg}")


def stream_ptr(stream=None):
    """Raw hipStream_t of a torch stream (current stream when None)."""
    import torch

    if stream is None:
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)
