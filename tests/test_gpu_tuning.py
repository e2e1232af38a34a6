"""Launch tuning changes the execution, not the arithmetic: every
windows-per-workgroup variant of the fused correlation kernel and both TPS
Cholesky widths give bit-identical results; the autotuner picks one of the
candidates."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pairs(rng, shape, npairs):
    from fluidimage_amd.device import ImageSet, to_device_image
    import torch

    dev = torch.device("cuda")
    ims = []
    for _ in range(npairs):
        im0 = rng.integers(0, 4096, shape).astype(np.uint16)
        im1 = np.roll(im0, (2, -3), axis=(0, 1))
        im1 = np.clip(im1.astype(np.int32) + rng.integers(-40, 40, shape), 0, 4095)
        ims += [to_device_image(im0, dev), to_device_image(im1.astype(np.uint16), dev)]
    return ImageSet(ims)


@pytest.mark.parametrize("n", [16, 24, 32, 48, 64])
def test_correlation_variants_bit_identical(n):
    import torch

    from fluidimage_amd import tuning
    from fluidimage_amd.works.piv.multipass import WorkPIV

    print("case: correlation variants", n, flush=True)
    rng = np.random.default_rng(100 + n)
    shape = (256, 320)
    npairs = 3
    images = _pairs(rng, shape, npairs)
    params = WorkPIV.create_default_params()
    params.piv0.shape_crop_im0 = n
    work = WorkPIV(params).works_piv[0]
    work._prepare_with_image(imshape=shape)
    dev = torch.device("cuda")
    nwin = work.nb_vectors * npairs
    windows = work.make_windows(0, 1, None, None, device=dev, npairs=npairs)["windows"]
    before = tuning.get(tuning.CORRELATE_WPB, n, n)
    results = {}
    try:
        for wpb in tuning.candidates(tuning.CORRELATE_WPB, n, n):
            tuning.set_value(tuning.CORRELATE_WPB, n, n, wpb)
            maps = torch.full((nwin, n, n), -7.0, dtype=torch.float32, device=dev)
            out = work.correl.correlate_windows(images.table, windows, nwin, images.dtype, True,
                                                maps=maps)
            torch.cuda.synchronize()
            results[wpb] = {k: v.clone() for k, v in out.items() if torch.is_tensor(v)}
            results[wpb]["maps"] = maps
    finally:
        tuning.set_value(tuning.CORRELATE_WPB, n, n, before)
    ref = results[before]
    assert int((ref["err"] == 0).sum()) > nwin // 2
    for wpb, r in results.items():
        for key, t in ref.items():
            assert torch.equal(torch.nan_to_num(r[key], nan=12345.0),
                               torch.nan_to_num(t, nan=12345.0)), (wpb, key)


def test_tps_cholesky_widths_bit_identical():
    import torch

    from fluidimage_amd import tuning
    from fluidimage_amd.calcul.tps import lattice_step, tps_device_batch

    print("case: tps cholesky widths", flush=True)
    dev = torch.device("cuda")
    # 16 fields of 16,129 vectors: about 1,300 subdomains, more than two per
    # CU, where the tuned width applies
    fields, xs, ys = tuning._grid_fields(16, (1024, 1024), 8, dev)
    lat = lattice_step(xs, ys)
    before = tuning.get(tuning.TPS_CHOL_THREADS)
    outs = {}
    try:
        for threads in (256, 512):
            tuning.set_value(tuning.TPS_CHOL_THREADS, 0, 0, threads)
            outs[threads] = tps_device_batch(fields, fields[0][2], fields[0][3], 200, 2.0, 1.5,
                                             lattice=lat)
    finally:
        tuning.set_value(tuning.TPS_CHOL_THREADS, 0, 0, before)
    for a, b in zip(outs[256], outs[512]):
        assert torch.equal(a["smoothed"], b["smoothed"])
        assert torch.equal(a["approx"], b["approx"])
        assert torch.equal(a["stats"], b["stats"])


def test_autotune_picks_a_candidate():
    from fluidimage_amd import tuning

    print("case: autotune 32x32", flush=True)
    before = tuning.get(tuning.CORRELATE_WPB, 32, 32)
    try:
        r = tuning.tune_correlation(32, npairs=2, image_shape=(512, 512), reps=2)
        assert r["best"] in tuning.candidates(tuning.CORRELATE_WPB, 32, 32)
        assert set(r["ms"]) == set(tuning.candidates(tuning.CORRELATE_WPB, 32, 32))
        assert all(v > 0 for v in r["ms"].values())
        assert tuning.get(tuning.CORRELATE_WPB, 32, 32) == r["best"]
    finally:
        tuning.set_value(tuning.CORRELATE_WPB, 32, 32, before)
