"""Benchmark of the PIV correlation path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload single]

Workload "single" (default): one 2048x2048 uint16 pair (12-bit, 0.03
particles/px, 8-mode random field up to 6 px), 32x32 windows at 50 %
overlap (16,129 windows), displacement_max 8, nb_peaks_to_search 1,
2d_gaussian2 sub-pixel fit, correlation maps materialised (the reference's
``_loop_vectors`` keeps every map). One step = image minima + window records
+ the fused correlation kernel over all windows; the images are resident in
HBM before the timed region. Prints one JSON line.
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "vectors/s, 2048² multipass PIV: batched vs plain HIP at <=1e-3 px; HBM GB/s"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="single", choices=["single"])
    ap.add_argument("--no-maps", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-rows", type=int, default=127,
                    help="vector rows of the CPU-baseline sample")
    return ap.parse_args()


def cpu_baseline(im0, im1, rows):
    """The oracle (numpy/pocketfft float32 restatement of the reference's
    per-window loop) on the first ``rows`` rows of vectors of the pair,
    one core."""
    import numpy as np

    from oracle.piv import OracleWorkPIVPass, default_params

    p = default_params()
    p["piv0"]["shape_crop_im0"] = 32
    p["piv0"]["displacement_max"] = 8
    height = 16 * (rows - 1) + 32
    a0 = np.ascontiguousarray(im0[:height])
    a1 = np.ascontiguousarray(im1[:height])
    w = OracleWorkPIVPass(p, 0, precision="f32")
    w.prepare_with_image(a0.shape)
    t0 = time.perf_counter()
    res = w.loop_vectors(a0, a1, keep_maps=True)
    dt = time.perf_counter() - t0
    n = len(res["deltaxs"])
    return {
        "value": n / dt,
        "unit": "vectors/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{n} windows (first {rows} vector rows of the 2048x2048 pair, "
        f"{a0.shape[0]}x{a0.shape[1]} px), oracle float32, {dt:.1f} s on 1 core",
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is described in DESIGN.md", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from fluidimage_amd.device import ImageSet, to_device_image
    from fluidimage_amd.synthetic import config_pair
    from fluidimage_amd.works.piv.multipass import WorkPIV
    from fluidimage_amd.works.piv.singlepass import FirstWorkPIV

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    print("bench: generating the 2048x2048 uint16 pair", file=sys.stderr, flush=True)
    im0, im1 = config_pair(2)
    params = WorkPIV.create_default_params()
    params.piv0.shape_crop_im0 = 32
    params.piv0.displacement_max = 8
    params.piv0.nb_peaks_to_search = 1
    work = FirstWorkPIV(params)
    work._prepare_with_image(im0)
    nvec = work.nb_vectors
    t0 = to_device_image(im0, dev)
    t1 = to_device_image(im1, dev)
    images = ImageSet([t0, t1])
    maps = None if args.no_maps else torch.empty((nvec, 32, 32), dtype=torch.float32, device=dev)
    out = work.correl._alloc_results(nvec, dev)
    win = work.make_windows(0, 1, device=dev)
    stream = torch.cuda.current_stream()

    def step(ev=None):
        images.recompute_min()
        work.make_windows(0, 1, windows=win["windows"], device=dev)
        if ev is not None:
            ev[0].record(stream)
        work.correl.correlate_windows(images.table, win["windows"], nvec, images.dtype, True,
                                      maps=maps, out=out)
        if ev is not None:
            ev[1].record(stream)

    print("bench: warmup", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(args.steps)]
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_wall = time.perf_counter()
    start.record(stream)
    for k in range(args.steps):
        step(evs[k])
    stop.record(stream)
    torch.cuda.synchronize()
    t_wall = time.perf_counter() - t_wall
    ms_total = start.elapsed_time(stop)
    ms_step = ms_total / args.steps
    kern_ms = float(np.mean([a.elapsed_time(b) for a, b in evs]))
    err = out["err"].cpu().numpy()
    vectors_per_s = nvec / (ms_step / 1e3)

    # algorithmic bytes of one correlation launch: unique image bytes
    # (2 H W b) + maps (nwin n0 n1 4, when written) + 16 B peak record
    H, W = im0.shape
    img_bytes = 2 * H * W * im0.dtype.itemsize
    map_bytes = 0 if maps is None else nvec * 32 * 32 * 4
    bytes_launch = img_bytes + map_bytes + 16 * nvec
    achieved = bytes_launch / (kern_ms / 1e3) / 1e9
    flops_window = 7.5 * 1024 * np.log2(1024) + 14 * 32 * (32 // 2 + 1)
    line = {
        "metric": METRIC,
        "value": round(vectors_per_s, 1),
        "unit": "vectors/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "single-pass batched correlation, one 2048x2048 uint16 pair, "
            "32x32 windows 50% overlap, displacement_max 8, 2d_gaussian2, maps "
            + ("off" if maps is None else "materialised"),
            "windows": int(nvec),
            "pairs": 1,
            "window_errors": int((err != 0).sum()),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "kernel": "k_correlate<32,32,u16>",
            "kernel_ms": round(kern_ms, 4),
            "bytes_per_launch": int(bytes_launch),
            "valu_tflops": round(nvec * flops_window / (kern_ms / 1e3) / 1e12, 2),
        },
        "wall_s": round(t_wall, 3),
    }
    if not args.no_cpu_baseline:
        print("bench: CPU baseline", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(im0, im1, args.cpu_rows)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
