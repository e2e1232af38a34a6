"""Benchmark of the PIV path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--pairs P]
                    [--workload multipass|single]

Workload "multipass" (default): P (32 by default) seeded 2048x2048 uint16 pairs (12-bit,
0.03 particles/px, 8 Fourier modes up to 6 px, 2 % of the 64 px windows of
im1 replaced by noise), WorkPIV with 3 passes (64, 32, 16 px windows, 50 %
overlap), ThinPlateSplineSubdom between passes and at the end (use_tps True,
subdom_size 200, smoothing_coef 2.0, threshold_tps 1.5), bicubic image
deformation by the predictor in passes 2 and 3, WorkFIX after every pass,
2d_gaussian2 on the last pass, correlation maps kept (the reference's
results carry them). One step = the whole multipass of the P pairs as one
batch (image minima, window records, one correlation launch per pass over
the windows of every pair, deformed windows in passes 2 and 3, fix, one
batched FP64 TPS per pass, final interpolation); the images are resident in HBM before the timed region.
value = final-pass vectors (65,025 per pair) x P per second. The same
pipeline on one pair (the plain per-pair path) is timed in the same run.
The TPS checks of a step (singular systems, capacities) are read while the
next step runs, and those of the last step after the timed region; a failed
check fails the bench. The per-stage times come from two more steps with
HIP events around the stages, run after the timed ones.

Workload "preproc": the sliding median / minima of 2048x2048 uint16 images.
Workload "single": the single-pass correlation of the same kind of pairs
(32x32 windows, displacement_max 8, 16,129 windows per pair). Workload
"direct": the direct spatial correlation (CorrelPythran) of the 32x32
windows of such pairs. Workload "refine": the differentiable refinement of
1024x1024 float32 pairs.

The roofline object is the first pass's correlation launch (the one that
reads the caller's images; its bytes are the algorithmic HBM figure: unique
image bytes + maps + 16 B peak records), timed as back-to-back relaunches
between two HIP events on the launch stream. The deformed passes' per-pass
figures count the float32 images of the deformation as their source. traffic comes from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary
of that kernel when one exists for the workload.

Prints one JSON line. --copy-calibration runs a plain 256 MiB device copy,
the reference for scaling the FETCH_SIZE / WRITE_SIZE counters.
"""

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "vectors/s, 2048² multipass PIV: batched vs plain HIP at <=1e-3 px; HBM GB/s"
HBM_PEAK_GBS = 8000.0
# rocprofv3 counter scaling measured with --copy-calibration (256 MiB copy):
# FETCH_SIZE reads 131,085 KiB per dispatch (half the bytes moved),
# WRITE_SIZE 262,144 KiB
FETCH_SCALE = 2.0
WRITE_SCALE = 1.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--pairs", type=int, default=32)
    ap.add_argument("--workload", default="multipass",
                    choices=["multipass", "single", "refine", "direct", "preproc",
                             "topology", "uneven"])
    ap.add_argument("--no-maps", action="store_true")
    ap.add_argument("--window", type=int, default=32,
                    help="window size of the single-pass workload (sizes without a "
                         "register-FFT kernel run the DFT kernel of any shape)")
    ap.add_argument("--graph", action="store_true",
                    help="replay the step captured in one HIP graph (its TPS checks are read "
                         "after each replay)")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the one-pair (plain per-pair path) measurement")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 256 MiB device-to-device copy (5 times), the "
                         "reference for scaling FETCH_SIZE / WRITE_SIZE counters")
    return ap.parse_args()


def copy_calibration():
    import torch

    nbytes = 256 * 1024 * 1024
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_per_launch": nbytes, "launches": 5}))


def multipass_params():
    from fluidimage_amd.works.piv.multipass import WorkPIV

    p = WorkPIV.create_default_params()
    p.piv0.shape_crop_im0 = 64
    p.multipass.number = 3
    p.multipass.coeff_zoom = 2
    p.multipass.use_tps = True
    p.multipass.subdom_size = 200
    p.multipass.smoothing_coef = 2.0
    p.multipass.threshold_tps = 1.5
    p.multipass.image_deformation = "bicubic"
    return p


def _oracle_multipass_pair(pair):
    """The oracle (numpy/pocketfft float32 + LAPACK restatement of the
    reference) on one pair of the multipass workload, one thread; returns
    (final vectors, seconds)."""
    from threadpoolctl import threadpool_limits

    from oracle.piv import OracleWorkPIV, default_params

    p = default_params()
    p["piv0"]["shape_crop_im0"] = 64
    p["multipass"].update(number=3, use_tps=True, subdom_size=200, smoothing_coef=2.0,
                          threshold_tps=1.5, image_deformation="bicubic")
    with threadpool_limits(1):
        t0 = time.perf_counter()
        res = OracleWorkPIV(p, precision="f32").calcul(*pair)
        dt = time.perf_counter() - t0
    return len(res[-1]["deltaxs"]), dt


def _oracle_single_pair(pair):
    from threadpoolctl import threadpool_limits

    from oracle.piv import OracleWorkPIVPass, default_params

    p = default_params()
    p["piv0"]["shape_crop_im0"] = 32
    p["piv0"]["displacement_max"] = 8
    w = OracleWorkPIVPass(p, 0, precision="f32")
    w.prepare_with_image(pair[0].shape)
    with threadpool_limits(1):
        t0 = time.perf_counter()
        res = w.loop_vectors(*pair, keep_maps=True)
        dt = time.perf_counter() - t0
    return len(res["deltaxs"]), dt


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def cpu_baseline(pairs, workload, max_cores=16):
    """The reference's CPU path restated (the oracle), timed on the host's
    cores as the reference runs it: one process per allowed core
    (os.sched_getaffinity, at most max_cores, the host's share of a one-GPU
    box), each on its own pair with one thread (OMP_NUM_THREADS=1), like
    multi_exec_async's split of a series; plus one core alone. Run before
    the GPU is touched: the workers are forked."""
    import multiprocessing as mp

    fn = _oracle_multipass_pair if workload == "multipass" else _oracle_single_pair
    cores = max(1, min(len(os.sched_getaffinity(0)), max_cores, len(pairs)))
    n1, dt1 = fn(pairs[0])
    one = n1 / dt1
    os.environ["OMP_NUM_THREADS"] = "1"
    ctx = mp.get_context("fork")
    t0 = time.perf_counter()
    with ctx.Pool(cores) as pool:
        out = pool.map(fn, pairs[:cores], chunksize=1)
    wall = time.perf_counter() - t0
    nall = sum(n for n, _ in out)
    kind = ("one 2048x2048 pair per process, 3 passes + TPS" if workload == "multipass"
            else "one 2048x2048 pair per process, 16,129 windows")
    return {
        "value": round(nall / wall, 1),
        "unit": "vectors/s",
        "cores": cores,
        "kind": "port",
        "one_core_value": round(one, 1),
        "cpu": cpu_model(),
        "sample": f"{cores} pairs of the workload on {cores} processes ({kind}, {nall} final "
                  f"vectors, {wall:.1f} s wall) and one pair on one core ({n1} vectors, "
                  f"{dt1:.1f} s); oracle float32 FFT + LAPACK TPS, one thread per process",
    }


def correl_bytes(work, imshape, itemsize, nwin, npairs, maps):
    """Algorithmic bytes of one correlation launch: unique image bytes
    (2 H W b per pair) + maps (nwin n0 n1 4, when written) + 16 B peak
    record per window."""
    H, W = imshape
    n0, n1 = work.shape_crop_im0
    return 2 * H * W * itemsize * npairs + (nwin * n0 * n1 * 4 if maps else 0) + 16 * nwin


def correl_flops(work, nwin):
    import numpy as np

    n0, n1 = work.shape_crop_im0
    n = n0 * n1
    return nwin * (7.5 * n * np.log2(n) + 14 * n0 * (n1 // 2 + 1))


def traffic_from_profile(kernel_tag):
    """HBM bytes per launch from the committed PMC summary of the kernel
    (lines 'FETCH_SIZE_KiB_per_dispatch: X' and 'WRITE_SIZE_KiB_per_dispatch: Y'),
    scaled by the copy calibration; None without a summary."""
    path = os.path.join(ROOT, "profiles", f"{kernel_tag}_pmc.md")
    if not os.path.exists(path):
        return None
    text = open(path).read()
    f = re.search(r"FETCH_SIZE_KiB_per_dispatch:\s*([0-9.]+)", text)
    w = re.search(r"WRITE_SIZE_KiB_per_dispatch:\s*([0-9.]+)", text)
    if not (f and w):
        return None
    return int((float(f.group(1)) * FETCH_SCALE + float(w.group(1)) * WRITE_SCALE) * 1024)


def tuning_in_use():
    """Launch tunables in effect (from the committed tuning table, if any)."""
    from fluidimage_amd import tuning

    t = tuning.current()
    t["table"] = os.path.basename(str(tuning._native.TUNING_PATH)) if os.path.exists(
        tuning._native.TUNING_PATH) else None
    return t


def timed(step, steps, warmup, breakdown=2):
    """Wall time of `steps` steps after `warmup` ones, and the per-stage
    averages (ms) of `breakdown` more steps run after the timed ones with
    HIP events around their stages (the timed steps record no events)."""
    import torch

    for _ in range(warmup):
        step(None)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step(None)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    timers = {}
    for _ in range(breakdown):
        step(timers)
    torch.cuda.synchronize()
    avg = {k: sum(a.elapsed_time(b) for a, b in v) / len(v) for k, v in timers.items()}
    return wall, avg


def relaunch_ms(fn, reps=20):
    """Average duration of fn's launch over reps back-to-back launches
    between two events on the current stream (the launch stream)."""
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


def make_pipeline(args, shape):
    from fluidimage_amd.works.piv.multipass import WorkPIV
    from fluidimage_amd.works.piv.pipeline import MultipassPipeline

    if args.workload == "multipass":
        params = multipass_params()
    else:
        params = WorkPIV.create_default_params()
        params.piv0.shape_crop_im0 = args.window
        params.piv0.displacement_max = 8 if args.window == 32 else None
    work = WorkPIV(params)
    work._prepare_with_image(imshape=shape)
    return work, MultipassPipeline(work)


def run_bench(args, pairs, keep_maps, steps, warmup):
    """Time `steps` pipeline steps over the pairs as one batch."""
    import torch

    from fluidimage_amd.device import ImageSet, to_device_image

    dev = torch.device("cuda:0")
    work, pipe = make_pipeline(args, pairs[0][0].shape)
    ims = []
    for a, b in pairs:
        ims += [to_device_image(a, dev), to_device_image(b, dev)]
    images = ImageSet(ims)
    state = {}

    if getattr(args, "graph", False):
        from fluidimage_amd.works.piv.pipeline import GraphedPasses

        graphed = GraphedPasses(work, images, len(pairs), keep_maps=keep_maps,
                                final_interp=args.workload == "multipass")

        def step(timers):
            state["passes"] = graphed.replay()

        wall, avg = timed(step, steps, warmup)
        return work, images, state["passes"], wall / steps * 1e3, avg

    def check_tps(pending):
        failed = pipe.check_pending(pending)
        if failed:
            raise RuntimeError(f"bench: TPS checks failed for pairs {sorted(failed)}")

    def step(timers):
        # the TPS checks of a step are read while the next step runs (as a
        # caller processing a series would: pairs failing one would be run
        # again), so the host stays ahead of the GPU
        images.recompute_min()
        state["passes"] = pipe.run(images, len(pairs), keep_maps=keep_maps, timers=timers,
                                   final_interp=args.workload == "multipass", check=False)
        prev = state.get("pending")
        state["pending"] = list(pipe._pending)
        if prev is not None:
            check_tps(prev)

    wall, avg = timed(step, steps, warmup)
    check_tps(state["pending"])
    return work, images, state["passes"], wall / steps * 1e3, avg


def pcie_inclusive(args, pairs, images, work, passes, keep_maps, nvec, npairs):
    """The same step with the host transfers a caller holding host arrays
    pays: the images uploaded from pinned host memory into the device images
    before the step, and the last pass's vectors (dx, dy, correl_max, errors)
    and the final interpolated fields copied back after it; separately, the
    copy of every kept map of the step to pinned host memory (the reference's
    ``correls``). Timed with HIP events on the launch stream, a few steps."""
    import torch

    from fluidimage_amd.works.piv.pipeline import MultipassPipeline

    pipe = MultipassPipeline(work)
    import numpy as np

    host_ims = []
    for a, b in pairs:
        for im in (a, b):
            a16 = im.dtype == np.uint16
            host_ims.append(torch.from_numpy(im.view(np.int16) if a16 else im).pin_memory())
    # uint16 device images are filled through an int16 view (same bytes)
    dev_ims = [d.view(torch.int16) if d.dtype != torch.uint8 and d.element_size() == 2 else d
               for d in images.images]
    steps = max(2, args.steps // 2)

    def results_of(ps):
        last = ps[-1]
        ts = [last.out[k] for k in ("dx", "dy", "correl_max", "err")]
        if args.workload == "multipass":
            ts += [t for f in last.final if f is not None for t in f]
        return ts

    hosts = None
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    up = comp = down = 0.0
    for it in range(steps + 1):
        ev[0].record()
        for h, d in zip(host_ims, dev_ims):
            d.copy_(h, non_blocking=True)
        ev[1].record()
        images.recompute_min()
        ps = pipe.run(images, npairs, keep_maps=keep_maps,
                      final_interp=args.workload == "multipass")
        ev[2].record()
        res = results_of(ps)
        if hosts is None:
            hosts = [torch.empty(t.shape, dtype=t.dtype, pin_memory=True) for t in res]
        for h, t in zip(hosts, res):
            h.copy_(t, non_blocking=True)
        ev[3].record()
        torch.cuda.synchronize()
        if it > 0:  # the first one warms the pinned buffers
            up += ev[0].elapsed_time(ev[1])
            comp += ev[1].elapsed_time(ev[2])
            down += ev[2].elapsed_time(ev[3])
    up, comp, down = up / steps, comp / steps, down / steps
    rec = {"upload_ms": round(up, 3), "step_ms": round(comp, 3),
           "results_download_ms": round(down, 3),
           "vectors/s_with_upload_and_results": round(nvec * npairs / ((up + comp + down) / 1e3), 1)}
    if keep_maps:
        maps = [p.out["maps"] for p in ps if p.out.get("maps") is not None]
        nbytes = sum(m.numel() * m.element_size() for m in maps)
        hmaps = [torch.empty(m.shape, dtype=m.dtype, pin_memory=True) for m in maps]
        torch.cuda.synchronize()
        ev[0].record()
        for h, m in zip(hmaps, maps):
            h.copy_(m, non_blocking=True)
        ev[1].record()
        torch.cuda.synchronize()
        mms = ev[0].elapsed_time(ev[1])
        rec["maps_download_ms"] = round(mms, 3)
        rec["maps_GB"] = round(nbytes / 1e9, 3)
        rec["vectors/s_with_upload_results_and_maps"] = round(
            nvec * npairs / ((up + comp + down + mms) / 1e3), 1)
        del hmaps
    return rec


def bench_refine(args):
    """Differentiable refinement: P 1024x1024 float32 pairs, multipass of
    32x32 windows (2 passes 64/32, TPS), then 20 gradient steps over the batch
    on TPS node values (the node displacements are their ThinPlateSplineSubdom
    smoothing; a step is kept per pair only when it lowers the pair's loss
    below that of its multipass field). value = refined vectors x pairs / time
    of the 20 steps (the TPS influence matrices are built once, before)."""
    import torch

    from fluidimage_amd.device import ImageSet, to_device_image
    from fluidimage_amd.synthetic import config_pair
    from fluidimage_amd.works.piv.multipass import WorkPIV
    from fluidimage_amd.works.piv.pipeline import MultipassPipeline
    from fluidimage_amd.works.piv.refine import RefineProblem, node_field, refine

    npairs = max(1, args.pairs)
    print(f"bench: generating {npairs} 1024x1024 float32 pairs", file=sys.stderr, flush=True)
    pairs = [config_pair(5, pair_idx=k) for k in range(npairs)]
    p = WorkPIV.create_default_params()
    p.piv0.shape_crop_im0 = 64
    p.multipass.number = 2
    p.multipass.use_tps = True
    p.multipass.smoothing_coef = 2.0
    work = WorkPIV(p)
    work._prepare_with_image(imshape=pairs[0][0].shape)
    dev = torch.device("cuda:0")
    ims = []
    for a, b in pairs:
        ims += [to_device_image(a, dev), to_device_image(b, dev)]
    images = ImageSet(ims)
    passes = MultipassPipeline(work).run(images, npairs)
    last = work.works_piv[-1]
    prob = RefineProblem(last, images, npairs)
    field = node_field(prob, p.multipass.subdom_size, p.multipass.smoothing_coef)
    dx0 = torch.stack([passes[-1].final[k][0] for k in range(npairs)])
    dy0 = torch.stack([passes[-1].final[k][1] for k in range(npairs)])
    for _ in range(args.warmup):
        refine(prob, dx0, dy0, steps=1, field=field)
    torch.cuda.synchronize()
    reps = max(1, args.steps // 10)
    t0 = time.perf_counter()
    for _ in range(reps):
        _, _, l0, l1, acc = refine(prob, dx0, dy0, steps=20, field=field)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    nvec = last.nb_vectors
    line = {
        "metric": "refined vectors/s (20 gradient steps, batched pairs)",
        "value": round(nvec * npairs / dt, 1),
        "unit": "vectors/s",
        "n_gpus": 1,
        "steps": reps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"differentiable refinement of {npairs} 1024x1024 float32 pairs: "
                        "windowed NCC over Keys-warped 32x32 windows, bilinear warp field of "
                        "the TPS-smoothed node values (subdom 200, smoothing 2.0), "
                        "20 accept-if-better gradient steps",
            "pairs": npairs,
            "vectors_per_pair": int(nvec),
            "mean_loss_before": round(float(l0.mean()), 4),
            "mean_loss_after": round(float(l1.mean()), 4),
            "mean_accepted_steps": round(float(acc.double().mean()), 2),
        },
    }
    print(json.dumps(line), flush=True)


def bench_preproc(args):
    """Sliding-window preprocessing (the reference's python toolbox,
    _toolbox_py.py:66-165): sliding_median and sliding_minima with the
    reference's default 30 x 30 window and reflect boundaries on P seeded
    2048x2048 uint16 images (12-bit particle images), resident in HBM; one
    launch filters the P images. value = images/s of the median. The CPU
    baseline is scipy.ndimage (what the reference calls) on a 512x512
    crop of one image, one thread, scaled to a whole image by the pixel
    count."""
    import ctypes

    import numpy as np
    import torch

    from fluidimage_amd import _native
    from fluidimage_amd.device import ImageSet, to_device_image
    from fluidimage_amd.synthetic import config_pair
    from fluidimage_amd.works.preproc.toolbox import filter_rank

    npairs = max(1, args.pairs // 2)
    print(f"bench: generating {2 * npairs} 2048x2048 uint16 images", file=sys.stderr, flush=True)
    ims = [im for k in range(npairs) for im in config_pair(2, pair_idx=k)]
    cpu = None
    if not args.no_cpu_baseline:
        from threadpoolctl import threadpool_limits

        from oracle import preproc as opp

        crop = ims[0][:512, :512]
        with threadpool_limits(1):
            t0 = time.perf_counter()
            opp.sliding_median(crop)
            dt = time.perf_counter() - t0
        scale = crop.size / ims[0].size
        cpu = {"value": round(scale / dt, 4), "unit": "images/s", "cores": 1, "kind": "port",
               "sample": f"sliding_median 30x30 of a 512x512 crop of one image with "
                         f"scipy.ndimage, {dt:.2f} s on one thread, scaled by the pixel count"}
    torch.cuda.set_device(0)
    dev = torch.device("cuda:0")
    dims = [to_device_image(a, dev) for a in ims]
    images = ImageSet(dims, subtract_min=False)
    n = len(dims)
    h, w = ims[0].shape
    out = torch.empty((n, h, w), dtype=torch.float64, device=dev)
    lib = _native.load()

    def run(rank):
        _native.check(lib.piv_sliding_filter(
            ctypes.c_void_p(images.table.data_ptr()), n, h, w, images.dtype, 30, 30, rank,
            _native.PIV_BOUNDARY["reflect"], 0.0, 1.0, ctypes.c_void_p(out.data_ptr()),
            _native.stream_ptr(None)), "piv_sliding_filter")

    for _ in range(max(1, args.warmup)):
        run(filter_rank("median", 30, 30))
    med_ms = relaunch_ms(lambda: run(filter_rank("median", 30, 30)), reps=max(2, args.steps))
    min_ms = relaunch_ms(lambda: run(0), reps=max(2, args.steps))
    value = n / (med_ms / 1e3)
    line = {
        "metric": "images/s, sliding_median 30x30 (reflect) of 2048x2048 uint16 images",
        "value": round(value, 1),
        "unit": "images/s",
        "n_gpus": 1,
        "steps": max(2, args.steps),
        "warmup": args.warmup,
        "ms_per_step": round(med_ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u16",
        "data": "synthetic",
        "config": {"workload": f"sliding_median and sliding_minima (30x30, reflect, weight 1) "
                               f"of {n} 2048x2048 uint16 images in one launch each",
                   "images": n,
                   "median_ms_per_launch": round(med_ms, 3),
                   "minima_ms_per_launch": round(min_ms, 3),
                   "minima_images/s": round(n / (min_ms / 1e3), 1)},
    }
    if cpu is not None:
        line["cpu_baseline"] = cpu
    print(json.dumps(line), flush=True)


def bench_topology(args):
    """The GPU-aware TopologyPIV over a series of 2 P seeded 2048x2048
    uint16 TIFF files on the local disk (P couples 2k, 2k + 1: str_subset
    "i:i+2", ind_step 2), the multipass
    workload's parameters, batches of 16 couples: images read by a thread
    pool while the GPU computes, results saved by another thread. value =
    final vectors of the series / wall time of compute() (file reading,
    upload, GPU, download and saving included); one warm-up run on a
    separate series first."""
    import shutil
    import tempfile

    from PIL import Image

    from fluidimage_amd.synthetic import config_pair
    from fluidimage_amd.topologies.piv import TopologyPIV

    npairs = max(2, args.pairs)
    root = Path(tempfile.mkdtemp(prefix="piv_topo_"))
    try:
        print(f"bench: writing {2 * npairs} 2048x2048 uint16 TIFF files", file=sys.stderr,
              flush=True)
        for name, n in (("warm", 4), ("Images", 2 * npairs)):
            d = root / name
            d.mkdir()
            for i in range(n):
                im = config_pair(3, pair_idx=i // 2)[i % 2]
                Image.fromarray(im).save(d / f"im_{i:04d}.tif")

        def topo(name):
            p = TopologyPIV.create_default_params()
            mp = multipass_params()
            for sect in ("piv0", "multipass"):
                for k, v in getattr(mp, sect)._as_dict().items():
                    if not isinstance(v, dict):
                        setattr(getattr(p, sect), k, v)
            p.series.path = str(root / name)
            p.series.ind_step = 2
            p.saving.how = "recompute"
            p.gpu.batch_pairs = 16
            return TopologyPIV(p, logging_level="warning")

        topo("warm").compute()
        t = topo("Images")
        t0 = time.perf_counter()
        t.compute()
        wall = time.perf_counter() - t0
        nvec = 65025
        line = {
            "metric": "vectors/s, TopologyPIV over a series of 2048x2048 uint16 TIFF files "
                      "(multipass, read and save included)",
            "value": round(nvec * len(t.results) / wall, 1),
            "unit": "vectors/s",
            "n_gpus": 1,
            "steps": 1,
            "warmup": 1,
            "ms_per_step": round(wall * 1e3, 1),
            "higher_is_better": True,
            "scaling": "weak",
            "vs_baseline": None,
            "dtype": "f32",
            "data": "synthetic",
            "config": {"workload": f"{len(t.results)} couples of a series of {2 * npairs} "
                                   "files, batches of 16, 4 reading threads",
                       "couples": len(t.results),
                       "ms_per_couple": round(wall * 1e3 / max(1, len(t.results)), 2),
                       "read_s_summed_over_threads": round(sum(t.durations["read"]), 3),
                       "compute_s": round(sum(t.durations["compute_piv"]), 3),
                       "save_s": round(sum(t.durations["save_piv"]), 3)},
        }
        print(json.dumps(line), flush=True)
    finally:
        shutil.rmtree(root, ignore_errors=True)


def bench_uneven(args):
    """The uneven batch: P seeded pairs of sizes drawn from 1024x1024,
    1600x1200, 2048x2048, 2560x2160 (uint16) with window schedules drawn
    from 128/64/32, 64/32/16, 48/24 and Oseen + shear fields of 3-12 px,
    the multipass settings of the default workload otherwise (TPS between
    passes and at the end, bicubic deformation). The pairs are bucketed by
    (shape, schedule), each bucket one batched multipass, all buckets
    captured in one HIP graph; a step = one replay (with the host read of the
    TPS checks). Beside it: the same buckets run eagerly one after the other,
    and the pairs one at a time (the plain path). value = final vectors of
    all pairs per second."""
    import torch

    from fluidimage_amd.synthetic import uneven_batch
    from fluidimage_amd.works.piv.pipeline import UnevenGraph

    npairs = max(1, args.pairs)
    print(f"bench: generating {npairs} uneven pairs", file=sys.stderr, flush=True)
    items = uneven_batch(npairs)
    params = multipass_params()
    torch.cuda.set_device(0)
    print("bench: capturing the uneven batch", file=sys.stderr, flush=True)
    g = UnevenGraph(items, params, keep_maps=not args.no_maps)
    nvec = sum(work.works_piv[-1].nb_vectors * len(idx) for work, _, idx in g.groups)

    def graph_step(_timers):
        g.replay()

    wall, _ = timed(graph_step, args.steps, args.warmup)
    ms_graph = wall / args.steps * 1e3

    def eager_step(_timers):
        for (work, images, idx), pipe in zip(g.groups, g.pipes):
            images.recompute_min()
            pipe.run(images, len(idx), keep_maps=not args.no_maps)

    wall, _ = timed(eager_step, max(2, args.steps // 2), 1)
    ms_eager = wall / max(2, args.steps // 2) * 1e3
    plain_ms = None
    if not args.no_plain:
        from fluidimage_amd.works.piv.pipeline import MultipassPipeline

        singles = []
        for work, images, idx in g.groups:
            for j in range(len(idx)):
                from fluidimage_amd.device import ImageSet

                one = ImageSet(images.images[2 * j:2 * j + 2])
                singles.append((MultipassPipeline(work), one))

        def plain_step(_timers):
            for pipe, one in singles:
                one.recompute_min()
                pipe.run(one, 1, keep_maps=not args.no_maps)

        wall, _ = timed(plain_step, 2, 1)
        plain_ms = wall / 2 * 1e3
    buckets = {f"{tuple(work.works_piv[0].imshape0)} {tuple(w.shape_crop_im0[0] for w in work.works_piv)}":
               len(idx) for work, _, idx in g.groups}
    line = {
        "metric": "vectors/s, uneven batch of 2048-class pairs (mixed sizes and window "
                  "schedules) in one captured HIP graph",
        "value": round(nvec / (ms_graph / 1e3), 1),
        "unit": "vectors/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_graph, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{npairs} pairs of 1024x1024 to 2560x2160 uint16, schedules "
                               "128/64/32, 64/32/16, 48/24, TPS between passes and at the end, "
                               "bicubic deformation, maps " +
                               ("off" if args.no_maps else "materialised"),
                   "pairs": npairs, "final_vectors": int(nvec), "buckets": buckets,
                   "eager_ms_per_step": round(ms_eager, 3),
                   "graph_over_eager": round(ms_eager / ms_graph, 3)},
    }
    if plain_ms is not None:
        line["config"]["plain_per_pair_ms_per_step"] = round(plain_ms, 3)
        line["config"]["batched_over_plain"] = round(plain_ms / ms_graph, 2)
    print(json.dumps(line), flush=True)


def bench_direct(args):
    """Direct (spatial) correlation, the reference's CorrelPythran: the 32x32
    windows (50 % overlap, 16,129 per pair) of P seeded 2048x2048 uint16
    pairs, displacement_max None (15: 31x31 maps), stacked on the device
    before the timed region; a step = maps + norms of every window, then
    the peak search and 2d_gaussian2 fit. value = windows/s. FP64 sums:
    2 x (sum over the map of the overlap areas) flops per window, against
    the FP64 vector peak."""
    import torch

    from fluidimage_amd.calcul.correl import CorrelPythranHIP
    from fluidimage_amd.synthetic import config_pair
    from fluidimage_amd.works.piv.multipass import WorkPIV

    npairs = max(1, args.pairs)
    torch.cuda.set_device(0)
    dev = torch.device("cuda:0")
    print(f"bench: generating {npairs} 2048x2048 uint16 pairs (config 2)", file=sys.stderr,
          flush=True)
    pairs = [config_pair(2, pair_idx=k) for k in range(npairs)]
    p = WorkPIV.create_default_params()
    p.piv0.shape_crop_im0 = 32
    work = WorkPIV(p).works_piv[0]
    work._prepare_with_image(imshape=pairs[0][0].shape)
    n = 32
    iy = torch.from_numpy(work.iyvecs_grid.astype("int64")).to(dev) - n // 2
    ix = torch.from_numpy(work.ixvecs_grid.astype("int64")).to(dev) - n // 2
    rows = (iy[:, None] + torch.arange(n, device=dev))[:, :, None]
    cols = (ix[:, None] + torch.arange(n, device=dev))[:, None, :]
    w0, w1 = [], []
    for a, b in pairs:
        for im, out in ((a, w0), (b, w1)):
            t = torch.from_numpy(im.astype("float32")).to(dev)
            t = t - t.min()
            out.append(t[rows.clamp(0, t.shape[0] - 1), cols.clamp(0, t.shape[1] - 1)])
    w0 = torch.cat(w0).contiguous()
    w1 = torch.cat(w1).contiguous()
    nwin = int(w0.shape[0])
    op = CorrelPythranHIP((n, n), (n, n), method_subpix="2d_gaussian2", nsubpix=None)
    d = op.displacement_max
    area = sum(n - abs(k) for k in range(-d, d + 1)) ** 2
    flops_per_window = 2 * area + 4 * n * n

    def step():
        maps, norms = op.correlate_stack(w0, w1)
        return op.peaks_stack(maps, norms, apply_subpix=True)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / args.steps * 1e3
    e0.record()
    op.correlate_stack(w0, w1)
    e1.record()
    torch.cuda.synchronize()
    k_ms = e0.elapsed_time(e1)
    nerr = int((out["err"] != 0).sum().item())
    achieved = flops_per_window * nwin / (k_ms / 1e3) / 1e12
    line = {
        "metric": "windows/s, direct correlation (CorrelPythran) of 2048² pairs, 32x32 windows",
        "value": round(nwin / (ms / 1e3), 1),
        "unit": "windows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"direct spatial correlation of the 32x32 windows of {npairs} "
                               "2048x2048 uint16 pairs (50% overlap, displacement_max 15, "
                               "31x31 maps), peak search and 2d_gaussian2",
                   "pairs": npairs, "windows": nwin, "window_errors": nerr},
        "roofline": {"bound": "valu", "achieved": round(achieved, 2), "peak": 78.6,
                     "unit": "TFLOP/s", "frac": round(achieved / 78.6, 4), "traffic": None,
                     "kernel": "k_correlate_direct", "kernel_ms": round(k_ms, 4),
                     "flops_per_window": flops_per_window},
    }
    if not args.no_cpu_baseline:
        from threadpoolctl import threadpool_limits

        from oracle.correl import OracleCorrelDirect

        ref = OracleCorrelDirect((n, n), (n, n), method_subpix="2d_gaussian2", nsubpix=None)
        h0 = w0[:3000].cpu().numpy()  # about 10 s of oracle work
        h1 = w1[:3000].cpu().numpy()
        with threadpool_limits(1):
            t0 = time.perf_counter()
            for k in range(len(h0)):
                c, norm = ref(h0[k], h1[k])
                try:
                    ref.compute_displacements_from_correl(c, norm)
                except Exception:
                    pass
            dt = time.perf_counter() - t0
        line["cpu_baseline"] = {"value": round(len(h0) / dt, 1), "unit": "windows/s",
                                "cores": 1, "kind": "port",
                                "sample": f"{len(h0)} windows of the workload, oracle "
                                          f"(numpy per displacement), {dt:.1f} s on 1 thread"}
    print(json.dumps(line), flush=True)


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.workload in ("refine", "direct", "preproc", "topology", "uneven"):
        if args.gpus != 1:
            print("bench.py runs on one GPU only", file=sys.stderr)
            sys.exit(2)
        {"refine": bench_refine, "direct": bench_direct, "preproc": bench_preproc,
         "topology": bench_topology, "uneven": bench_uneven}[args.workload](args)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is described in DESIGN.md", file=sys.stderr)
        sys.exit(2)
    import torch

    from fluidimage_amd.synthetic import config_pair

    keep_maps = not args.no_maps
    cfg = 3 if args.workload == "multipass" else 2
    npairs = max(1, args.pairs)
    print(f"bench: generating {npairs} 2048x2048 uint16 pairs (config {cfg})", file=sys.stderr,
          flush=True)
    pairs = [config_pair(cfg, pair_idx=k) for k in range(npairs)]
    im0 = pairs[0][0]
    cpu_line = None
    if not args.no_cpu_baseline:
        # before the GPU is initialised: the worker processes are forked
        print("bench: CPU baseline", file=sys.stderr, flush=True)
        cpu_line = cpu_baseline(pairs, args.workload)
    torch.cuda.set_device(0)

    print("bench: batched steps", file=sys.stderr, flush=True)
    work, images, passes, ms_step, avg = run_bench(args, pairs, keep_maps, args.steps,
                                                   args.warmup)
    last = work.works_piv[-1]
    nvec = last.nb_vectors
    nerr = int((passes[-1].out["err"] != 0).sum().item())

    # correlation launches: event-bracketed per pass in the steps, and the
    # first pass's (the one that reads the caller's images; the deformed
    # passes read the float32 images of the deformation) relaunched back to
    # back for its clean average duration
    deformed = getattr(work.params.multipass, "image_deformation", None) is not None

    def src_itemsize(w):
        return 4 if deformed and w.index_pass > 0 else im0.dtype.itemsize

    per_pass = {}
    for ip, w in enumerate(work.works_piv):
        key = f"correlate_pass{ip}"
        if key in avg:
            nwin = w.nb_vectors * npairs
            per_pass[key] = {"window": list(w.shape_crop_im0), "windows": int(nwin),
                             "ms_in_step": round(avg[key], 4)}
    if not per_pass:  # graph replays: no per-pass events
        w = work.works_piv[0]
        per_pass["correlate_pass0"] = {"window": list(w.shape_crop_im0),
                                       "windows": int(w.nb_vectors * npairs), "ms_in_step": None}
    ck = "correlate_pass0" if "correlate_pass0" in per_pass else min(per_pass)
    cip = int(ck[len("correlate_pass"):])
    cw = work.works_piv[cip]
    po = passes[cip]
    nwin = cw.nb_vectors * npairs
    maps = po.out["maps"]

    def relaunch():
        cw.correl.correlate_windows(images.table, po.out["windows"], nwin, images.dtype,
                                    cw.has_to_apply_subpix, maps=maps)

    k_ms = relaunch_ms(relaunch)
    cbytes = correl_bytes(cw, im0.shape, im0.dtype.itemsize, nwin, npairs, keep_maps)
    achieved = cbytes / (k_ms / 1e3) / 1e9
    n0, n1 = cw.shape_crop_im0
    dt_tag = {1: "u8", 2: "u16", 4: "f32"}[im0.dtype.itemsize]
    ktag = f"k_correlate_{n0}x{n1}_{dt_tag}_2048_{args.workload}_{npairs}pairs"
    traffic = traffic_from_profile(ktag)
    for key, v in per_pass.items():
        if v["ms_in_step"] is None:
            continue
        w = work.works_piv[int(key[len("correlate_pass"):])]
        b = correl_bytes(w, im0.shape, src_itemsize(w), v["windows"], npairs, keep_maps)
        v["GB/s_in_step"] = round(b / (v["ms_in_step"] / 1e3) / 1e9, 1)
        v["valu_TFLOP/s_in_step"] = round(
            correl_flops(w, v["windows"]) / (v["ms_in_step"] / 1e3) / 1e12, 2)

    if args.workload == "multipass":
        desc = ("multipass PIV of {p} 2048x2048 uint16 pairs as one batch: 64/32/16 px "
                "windows, 50% overlap, WorkFIX, GPU FP64 ThinPlateSplineSubdom (subdom 200, "
                "smoothing 2.0, threshold 1.5) between passes and at the end, bicubic image "
                "deformation in passes 2-3, 2d_gaussian2, "
                "maps ").format(p=npairs)
    else:
        w = args.window
        desc = ("single-pass batched correlation of {p} 2048x2048 uint16 pairs: {w}x{w} "
                "windows 50% overlap, displacement_max {d}, 2d_gaussian2, maps ").format(
                    p=npairs, w=w, d=8 if w == 32 else None)
    value = nvec * npairs / (ms_step / 1e3)
    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "vectors/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": desc + ("materialised" if keep_maps else "off"),
            "pairs": npairs,
            "final_vectors_per_pair": int(nvec),
            "final_window_errors": nerr,
            "stage_ms_in_step": {k: round(v, 4) for k, v in sorted(avg.items())},
            "correlation_passes": per_pass,
            "tuning": tuning_in_use(),
            "hip_graph": bool(args.graph),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": f"k_correlate<{n0},{n1},{dt_tag}> ({ck}, {nwin} windows)",
            "kernel_ms": round(k_ms, 4),
            "bytes_per_launch": int(cbytes),
            "profile_tag": ktag,
        },
    }

    if not args.graph:
        print("bench: host-transfer-inclusive steps", file=sys.stderr, flush=True)
        line["config"]["pcie"] = pcie_inclusive(args, pairs, images, work, passes, keep_maps,
                                                nvec, npairs)

    if not args.no_plain and npairs > 1:
        print("bench: plain per-pair steps", file=sys.stderr, flush=True)
        # the plain per-pair path: the same kernels on one pair, eager
        plain_args = argparse.Namespace(**{**vars(args), "graph": False})
        _, _, _, ms1, _ = run_bench(plain_args, pairs[:1], keep_maps,
                                    max(3, args.steps // 2), args.warmup)
        plain = nvec / (ms1 / 1e3)
        line["config"]["plain_per_pair"] = {"ms_per_step": round(ms1, 3),
                                            "vectors/s": round(plain, 1)}
        line["config"]["batched_over_plain"] = round(value / plain, 2)

    if cpu_line is not None:
        line["cpu_baseline"] = cpu_line
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
