"""Benchmark of the PIV path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload multipass|single]

Workload "multipass" (default): one 2048x2048 uint16 pair (12-bit, 0.03
particles/px, 8 Fourier modes up to 6 px, 2 % of the 64 px windows replaced
by noise), WorkPIV with 3 passes (64, 32, 16 px windows, 50 % overlap),
ThinPlateSplineSubdom between passes and at the end (use_tps True,
subdom_size 200, smoothing_coef 2.0, threshold_tps 1.5), WorkFIX after
every pass, 2d_gaussian2 on the last pass, correlation maps kept (the
reference's results carry them). One step = the whole multipass of the pair
(image minima, window records, correlation, fix, FP64 TPS, final
interpolation); the images are resident in HBM before the timed region.
value = final-pass vectors (65,025) per second.

Workload "single": the single-pass correlation of the same kind of pair
(32x32 windows, displacement_max 8, 16,129 windows).

Prints one JSON line. --copy-calibration runs a plain 256 MiB device copy,
the reference for scaling the FETCH_SIZE / WRITE_SIZE counters.
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
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="multipass", choices=["multipass", "single"])
    ap.add_argument("--no-maps", action="store_true")
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
    return p


def cpu_baseline_multipass(im0, im1):
    """The oracle (numpy/pocketfft float32 + LAPACK restatement of the
    reference) on the whole pair, one thread."""
    from threadpoolctl import threadpool_limits

    from oracle.piv import OracleWorkPIV, default_params

    p = default_params()
    p["piv0"]["shape_crop_im0"] = 64
    p["multipass"].update(number=3, use_tps=True, subdom_size=200, smoothing_coef=2.0,
                          threshold_tps=1.5)
    with threadpool_limits(1):
        t0 = time.perf_counter()
        res = OracleWorkPIV(p, precision="f32").calcul(im0, im1)
        dt = time.perf_counter() - t0
    n = len(res[-1]["deltaxs"])
    return {
        "value": round(n / dt, 1),
        "unit": "vectors/s",
        "cores": 1,
        "kind": "port",
        "sample": f"the whole workload (one 2048x2048 pair, 3 passes + TPS, {n} final "
        f"vectors), oracle float32 FFT + LAPACK TPS, {dt:.1f} s on 1 thread",
    }


def cpu_baseline_single(im0, im1):
    from threadpoolctl import threadpool_limits

    from oracle.piv import OracleWorkPIVPass, default_params

    p = default_params()
    p["piv0"]["shape_crop_im0"] = 32
    p["piv0"]["displacement_max"] = 8
    w = OracleWorkPIVPass(p, 0, precision="f32")
    w.prepare_with_image(im0.shape)
    with threadpool_limits(1):
        t0 = time.perf_counter()
        res = w.loop_vectors(im0, im1, keep_maps=True)
        dt = time.perf_counter() - t0
    n = len(res["deltaxs"])
    return {"value": round(n / dt, 1), "unit": "vectors/s", "cores": 1, "kind": "port",
            "sample": f"the whole workload ({n} windows), oracle float32, {dt:.1f} s on 1 thread"}


def correl_bytes(work, imshape, itemsize, nwin, maps):
    """Algorithmic bytes of one correlation launch: unique image bytes
    (2 H W b) + maps (nwin n0 n1 4, when written) + 16 B peak record."""
    H, W = imshape
    n0, n1 = work.shape_crop_im0
    return 2 * H * W * itemsize + (nwin * n0 * n1 * 4 if maps else 0) + 16 * nwin


def correl_flops(work, nwin):
    import numpy as np

    n0, n1 = work.shape_crop_im0
    n = n0 * n1
    return nwin * (7.5 * n * np.log2(n) + 14 * n0 * (n1 // 2 + 1))


def timed(step, steps, warmup):
    import torch

    for _ in range(warmup):
        step(None)
    torch.cuda.synchronize()
    timers = {}
    t0 = time.perf_counter()
    for _ in range(steps):
        step(timers)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    avg = {k: sum(a.elapsed_time(b) for a, b in v) / len(v) for k, v in timers.items()}
    return wall, avg


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is described in DESIGN.md", file=sys.stderr)
        sys.exit(2)
    import torch

    from fluidimage_amd.device import ImageSet, to_device_image
    from fluidimage_amd.synthetic import config_pair
    from fluidimage_amd.works.piv.multipass import WorkPIV
    from fluidimage_amd.works.piv.pipeline import MultipassPipeline

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    keep_maps = not args.no_maps
    cfg = 3 if args.workload == "multipass" else 2
    print(f"bench: generating the 2048x2048 uint16 pair (config {cfg})", file=sys.stderr,
          flush=True)
    im0, im1 = config_pair(cfg)
    if args.workload == "multipass":
        params = multipass_params()
    else:
        params = WorkPIV.create_default_params()
        params.piv0.shape_crop_im0 = 32
        params.piv0.displacement_max = 8
    work = WorkPIV(params)
    work._prepare_with_image(imshape=im0.shape)
    images = ImageSet([to_device_image(im0, dev), to_device_image(im1, dev)])
    pipe = MultipassPipeline(work)
    state = {}

    def step(timers):
        images.recompute_min()
        state["passes"] = pipe.run(images, 1, keep_maps=keep_maps, timers=timers,
                                   final_interp=args.workload == "multipass")

    print("bench: warmup + timed steps", file=sys.stderr, flush=True)
    wall, avg = timed(step, args.steps, args.warmup)
    ms_step = wall / args.steps * 1e3
    last = work.works_piv[-1]
    nvec = last.nb_vectors
    passes = state["passes"]
    nerr = int((passes[-1].out["err"] != 0).sum().item())
    dom_key = max(avg, key=avg.get)
    per_pass = {}
    for ip, w in enumerate(work.works_piv):
        key = f"correlate_pass{ip}"
        if key in avg:
            b = correl_bytes(w, im0.shape, im0.dtype.itemsize, w.nb_vectors, keep_maps)
            per_pass[key] = {
                "window": list(w.shape_crop_im0),
                "windows": int(w.nb_vectors),
                "ms": round(avg[key], 4),
                "GB/s": round(b / (avg[key] / 1e3) / 1e9, 1),
                "valu_TFLOP/s": round(correl_flops(w, w.nb_vectors) / (avg[key] / 1e3) / 1e12,
                                      2),
            }
    # roofline of the longest correlation launch (HBM-bound with maps)
    ck = max(per_pass, key=lambda k: per_pass[k]["ms"])
    cw = work.works_piv[int(ck[-1])]
    cbytes = correl_bytes(cw, im0.shape, im0.dtype.itemsize, cw.nb_vectors, keep_maps)
    achieved = cbytes / (avg[ck] / 1e3) / 1e9
    if args.workload == "multipass":
        desc = ("multipass PIV of one 2048x2048 uint16 pair: 64/32/16 px windows, 50% "
                "overlap, WorkFIX, GPU FP64 ThinPlateSplineSubdom (subdom 200, smoothing 2.0) "
                "between passes and at the end, 2d_gaussian2, maps ")
    else:
        desc = ("single-pass batched correlation of one 2048x2048 uint16 pair: 32x32 "
                "windows 50% overlap, displacement_max 8, 2d_gaussian2, maps ")
    line = {
        "metric": METRIC,
        "value": round(nvec / (ms_step / 1e3), 1),
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
            "final_vectors": int(nvec),
            "pairs": 1,
            "final_window_errors": nerr,
            "stage_ms": {k: round(v, 4) for k, v in sorted(avg.items())},
            "correlation_passes": per_pass,
            "dominant": dom_key,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "kernel": f"k_correlate<{cw.shape_crop_im0[0]},{cw.shape_crop_im0[1]},u16> ({ck})",
            "kernel_ms": round(avg[ck], 4),
            "bytes_per_launch": int(cbytes),
        },
    }
    if not args.no_cpu_baseline:
        print("bench: CPU baseline", file=sys.stderr, flush=True)
        if args.workload == "multipass":
            line["cpu_baseline"] = cpu_baseline_multipass(im0, im1)
        else:
            line["cpu_baseline"] = cpu_baseline_single(im0, im1)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
