"""Autotuner of the launch parameters of libpiv_hip on the GPU at hand.

The tunables (``piv_tune_*`` in include/piv_hip.h) change how the work is
mapped onto the GPU, never the results:

* windows per workgroup of the fused correlation kernel, per window shape
  (the variants built for the multipass window sizes 16, 24, 32, 48, 64);
* threads per subdomain of the FP64 TPS Cholesky in launches of more than
  two subdomains per CU (256: two workgroups per CU, 512: one; 0 means 256);
  smaller launches pick the width that runs them in one round.

Every candidate is timed on a synthetic batch of the real shape: the
correlation by HIP events around back-to-back launches on the launch
stream, the TPS by the wall time of whole batched calls (they end with a
device-to-host read). The fastest candidates form a table, written as JSON
beside the library (``tuning_<arch>.json``) and applied by
``_native.load()`` whenever the library is loaded.

    python -m fluidimage_amd.tuning [--write] [--shapes 16 24 32 48 64]

The reference has no counterpart: its FFTW plans pick their own algorithm
at plan time (FFTW_MEASURE, src/fluidimage/calcul/fft.py:221-234) and its
TPS solve is LAPACK's.
"""

import argparse
import ctypes
import json
import sys
import time

import numpy as np

from . import _native

CORRELATE_WPB = 1
TPS_CHOL_THREADS = 2
TPS_SHARED_SOLVE = 3
DEFAULT_SHAPES = (16, 24, 32, 48, 64)


def candidates(param, n0=0, n1=0):
    """Values the library accepts for a tunable (list of int)."""
    lib = _native.load()
    buf = (ctypes.c_int32 * 16)()
    n = lib.piv_tune_candidates(int(param), int(n0), int(n1), buf, 16)
    if n < 0:
        raise ValueError(lib.piv_last_error().decode())
    return list(buf[:n])


def get(param, n0=0, n1=0):
    lib = _native.load()
    v = lib.piv_tune_get(int(param), int(n0), int(n1))
    if v < 0:
        raise ValueError(lib.piv_last_error().decode())
    return v


def set_value(param, n0, n1, value):
    lib = _native.load()
    _native.check(lib.piv_tune_set(int(param), int(n0), int(n1), int(value)), "piv_tune_set")


def apply(table, lib=None):
    """Set the values of a tuning table (dict as written by ``autotune``)."""
    lib = lib if lib is not None else _native.load()
    for key, wpb in table.get("correlate_wpb", {}).items():
        n0, n1 = (int(v) for v in key.split("x"))
        _native.check(lib.piv_tune_set(CORRELATE_WPB, n0, n1, int(wpb)), "piv_tune_set")
    if "tps_chol_threads" in table:
        _native.check(lib.piv_tune_set(TPS_CHOL_THREADS, 0, 0, int(table["tps_chol_threads"])),
                      "piv_tune_set")
    if "tps_shared_solve" in table:
        _native.check(lib.piv_tune_set(TPS_SHARED_SOLVE, 0, 0, int(table["tps_shared_solve"])),
                      "piv_tune_set")


def current():
    """The values in use, in the table format."""
    table = {"correlate_wpb": {}, "tps_chol_threads": get(TPS_CHOL_THREADS),
             "tps_shared_solve": get(TPS_SHARED_SOLVE)}
    for n in DEFAULT_SHAPES:
        table["correlate_wpb"][f"{n}x{n}"] = get(CORRELATE_WPB, n, n)
    return table


def _noise_images(npairs, shape, dev, seed=5):
    """Uniform 12-bit noise pairs (the kernels' work does not depend on the
    particle content): an ImageSet on the device."""
    import torch

    from .device import ImageSet, to_device_image

    g = torch.Generator(device=dev).manual_seed(seed)
    ims = []
    for _ in range(2 * npairs):
        t = torch.randint(0, 4096, shape, generator=g, device=dev, dtype=torch.int32)
        ims.append(to_device_image(t.to(torch.int16).view(torch.uint16), dev))
    return ImageSet(ims)


def _events_ms(fn, reps):
    import torch

    fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def tune_correlation(n, npairs=16, image_shape=(2048, 2048), reps=10, keep_maps=True):
    """Time every windows-per-workgroup candidate of the n x n fused kernel
    on npairs image pairs (50 % overlap grid, maps written as in the
    multipass), set the fastest. Returns {"best": wpb, "ms": {wpb: ms}}."""
    import torch

    from .works.piv.multipass import WorkPIV

    dev = torch.device("cuda")
    params = WorkPIV.create_default_params()
    params.piv0.shape_crop_im0 = n
    work = WorkPIV(params).works_piv[0]
    work._prepare_with_image(imshape=image_shape)
    images = _noise_images(npairs, image_shape, dev)
    nwin = work.nb_vectors * npairs
    windows = work.make_windows(0, 1, None, None, device=dev, npairs=npairs)["windows"]
    maps = (torch.empty((nwin, n, n), dtype=torch.float32, device=dev) if keep_maps else None)
    out = work.correl._alloc_results(nwin, dev)
    ms = {}
    before = get(CORRELATE_WPB, n, n)
    try:
        for wpb in candidates(CORRELATE_WPB, n, n):
            set_value(CORRELATE_WPB, n, n, wpb)
            ms[wpb] = _events_ms(
                lambda: work.correl.correlate_windows(images.table, windows, nwin, images.dtype,
                                                      True, maps=maps, out=out), reps)
    except Exception:
        set_value(CORRELATE_WPB, n, n, before)
        raise
    best = min(ms, key=ms.get)
    set_value(CORRELATE_WPB, n, n, best)
    return {"best": best, "ms": ms}


def _grid_fields(nfields, shape, step, dev, seed=7):
    """Vector fields on a PIV grid with 1 % NaN holes and 2 % outliers."""
    import torch

    rng = np.random.default_rng(seed)
    ny, nx = shape
    ys, xs = np.mgrid[step // 2:ny - step // 2 + 1:step, step // 2:nx - step // 2 + 1:step]
    xs = xs.ravel().astype(np.float64)
    ys = ys.ravel().astype(np.float64)
    fields = []
    for _ in range(nfields):
        u = 2.0 * np.sin(xs / 90.0) + 0.5 * np.cos(ys / 70.0)
        v = 1.5 * np.cos(xs / 60.0 + ys / 110.0)
        bad = rng.random(xs.size) < 0.02
        u[bad] += rng.normal(0, 6, bad.sum())
        nan = rng.random(xs.size) < 0.01
        u[nan] = np.nan
        v[nan] = np.nan
        t = lambda a, d: torch.from_numpy(np.ascontiguousarray(a, dtype=d)).to(dev)  # noqa: E731
        fields.append((t(u, np.float32), t(v, np.float32), t(xs, np.float64),
                       t(ys, np.float64)))
    return fields, xs, ys


def tune_tps(nfields=32, shape=(2048, 2048), step=8, reps=3, subdom_size=200):
    """Time the batched TPS (last-pass size: 65,025 vectors per field) with
    every Cholesky width, set the fastest. Returns {"best": threads, "ms": {...}}."""
    import torch

    from .calcul.tps import lattice_step, tps_device_batch

    dev = torch.device("cuda")
    fields, xs, ys = _grid_fields(nfields, shape, step, dev)
    gx = fields[0][2]
    gy = fields[0][3]
    lat = lattice_step(xs, ys)

    def run():
        tps_device_batch(fields, gx, gy, subdom_size, 2.0, 1.5, lattice=lat)

    ms = {}
    before = get(TPS_CHOL_THREADS)
    try:
        for threads in candidates(TPS_CHOL_THREADS):
            set_value(TPS_CHOL_THREADS, 0, 0, threads)
            run()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                run()
            torch.cuda.synchronize()
            ms[threads] = (time.perf_counter() - t0) / reps * 1e3
    except Exception:
        set_value(TPS_CHOL_THREADS, 0, 0, before)
        raise
    best = min(ms, key=ms.get)
    set_value(TPS_CHOL_THREADS, 0, 0, best)
    return {"best": best, "ms": ms}


def autotune(shapes=DEFAULT_SHAPES, npairs=16, tps_fields=32, log=print):
    """Tune every parameter; returns the table (with the measured times)."""
    import torch

    table = {"arch": _native.ARCH, "device": torch.cuda.get_device_name(0),
             "correlate_wpb": {}, "measured_ms": {}}
    for n in shapes:
        r = tune_correlation(n, npairs=npairs)
        table["correlate_wpb"][f"{n}x{n}"] = r["best"]
        table["measured_ms"][f"correlate_{n}x{n}"] = {str(k): round(v, 4)
                                                      for k, v in r["ms"].items()}
        log(f"correlation {n}x{n}: {r['ms']} -> {r['best']}")
    r = tune_tps(nfields=tps_fields)
    table["tps_chol_threads"] = r["best"]
    table["measured_ms"]["tps_batch"] = {str(k): round(v, 3) for k, v in r["ms"].items()}
    log(f"TPS Cholesky threads: {r['ms']} -> {r['best']}")
    return table


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--write", action="store_true",
                    help=f"write the table to {_native.TUNING_PATH.name}")
    ap.add_argument("--shapes", type=int, nargs="*", default=list(DEFAULT_SHAPES))
    ap.add_argument("--pairs", type=int, default=16)
    ap.add_argument("--out", help="also write the table to this file")
    args = ap.parse_args()
    table = autotune(tuple(args.shapes), npairs=args.pairs,
                     log=lambda m: print(m, file=sys.stderr, flush=True))
    text = json.dumps(table, indent=1, sort_keys=True)
    print(text, flush=True)
    if args.write:
        _native.TUNING_PATH.write_text(text + "\n")
    if args.out:
        with open(args.out, "w") as f:
            f.write(text + "\n")


if __name__ == "__main__":
    main()
