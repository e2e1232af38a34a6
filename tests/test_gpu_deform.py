"""GPU image deformation between passes against the oracle (oracle/deform.py).

Bars: deformed windows within float32 rounding of the FP64 oracle (4e-6 of
the largest value: 0.016 for 12-bit images); multipass displacements within 1e-3 px with
identical error strings for every vector except the listed near-ties of the
FP64 maps and the vectors downstream of them; the deformation lowers the
error against the synthetic truth field."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pass_work(shape, window):
    from fluidimage_amd.works.piv.multipass import WorkPIV

    p = WorkPIV.create_default_params()
    p.piv0.shape_crop_im0 = window
    w = WorkPIV(p).works_piv[0]
    w._prepare_with_image(imshape=shape)
    return w


@pytest.mark.parametrize("window, dtype", [(32, np.uint16), (16, np.uint8), (24, np.float32)])
def test_deformed_windows_match_oracle(window, dtype):
    import torch

    from fluidimage_amd.device import ImageSet, to_device_image
    from fluidimage_amd.synthetic import config_pair
    from fluidimage_amd.works.piv.deform import DeformedPass
    from oracle.deform import deformed_crops

    print("case: deformed windows", window, np.dtype(dtype).name, flush=True)
    shape = (200, 232)
    im0, im1 = config_pair(3, shape=shape)
    if dtype == np.uint8:
        im0, im1 = (im0 >> 4).astype(np.uint8), (im1 >> 4).astype(np.uint8)
    elif dtype == np.float32:
        im0, im1 = im0.astype(np.float32) * 0.25, im1.astype(np.float32) * 0.25
    w = _pass_work(shape, window)
    ny, nx = len(w.iyvecs), len(w.ixvecs)
    rng = np.random.default_rng(7)
    gy, gx = np.meshgrid(w.iyvecs, w.ixvecs, indexing="ij")
    u = 3.0 * np.sin(gx / 50.0 + 0.3) + rng.normal(0, 0.3, gx.shape)
    v = -2.5 * np.cos(gy / 40.0) + rng.normal(0, 0.3, gx.shape)
    dev = torch.device("cuda")
    images = ImageSet([to_device_image(im0, dev), to_device_image(im1, dev)])
    dp = DeformedPass(w, dev)
    ut = torch.from_numpy(u.ravel().copy()).to(dev)
    vt = torch.from_numpy(v.ravel().copy()).to(dev)
    deformed = dp.run(images, 1, ut, vt)
    torch.cuda.synchronize()
    n0, n1 = w.shape_crop_im0
    d0 = deformed.images[0].cpu().numpy()
    d1 = deformed.images[1].cpu().numpy()
    # the deformed window of a node is the crop of the deformed images there
    g0 = np.stack([d0[y - n0 // 2:y - n0 // 2 + n0, x - n1 // 2:x - n1 // 2 + n1]
                   for y in w.iyvecs for x in w.ixvecs])
    g1 = np.stack([d1[y - n0 // 2:y - n0 // 2 + n0, x - n1 // 2:x - n1 // 2 + n1]
                   for y in w.iyvecs for x in w.ixvecs])
    e0, e1 = deformed_crops(im0, im1, w.iyvecs, w.ixvecs, u, v, (n0, n1))
    # float32 weights and 16-tap sums against the FP64 oracle: a few
    # float32 roundings of terms up to the image's largest value
    scale = max(1.0, float(np.abs(e0).max()))
    err0, err1 = np.abs(g0 - e0).max(), np.abs(g1 - e1).max()
    print(f"deformed windows: max |diff| {err0:.3g}, {err1:.3g} (scale {scale:.0f})", flush=True)
    assert err0 <= 4e-6 * scale + 1e-4
    assert err1 <= 4e-6 * scale + 1e-4


def _setup(shape, number=3, window=64):
    from fluidimage_amd.synthetic import config_pair
    from fluidimage_amd.works.piv.multipass import WorkPIV
    from tests.test_gpu_piv import _oracle_params

    im0, im1 = config_pair(3, shape=shape)
    p = WorkPIV.create_default_params()
    p.piv0.shape_crop_im0 = window
    p.multipass.number = number
    p.multipass.use_tps = True
    p.multipass.smoothing_coef = 2.0
    p.multipass.subdom_size = 200
    p.multipass.image_deformation = "bicubic"
    op = _oracle_params(p)
    op["multipass"]["image_deformation"] = "bicubic"
    return im0, im1, p, op


@pytest.mark.parametrize("share", [False, True])
def test_multipass_with_deformation_vs_oracle(share, monkeypatch):
    """share: the TPS calls on the deformed passes' grids share their
    factorisations between subdomains of one node pattern (forced at this
    size)."""
    from fluidimage_amd.calcul import tps_planned
    from fluidimage_amd.works.piv.multipass import WorkPIV
    from oracle.piv import OracleWorkPIV
    from tests.test_gpu_multipass import check_passes_against_oracle

    monkeypatch.setattr(tps_planned, "SHARED_MIN_SUBDOMAINS", 0 if share else 1 << 30)
    monkeypatch.setattr(tps_planned, "SHARED_MIN_BATCH", 0)
    print("case: multipass with deformation, shared factorisations", share, flush=True)
    im0, im1, p, op = _setup((512, 512))
    gpu = WorkPIV(p).calcul_from_arrays(im0, im1)
    ow = OracleWorkPIV(op, precision="f64")
    ref = ow.calcul(im0, im1)
    for ip, (g, e) in enumerate(zip(gpu.passes, ref)):
        if ip > 0:
            assert np.allclose(g.xs, e["xs"]) and np.allclose(g.ys, e["ys"])
    check_passes_against_oracle(gpu.passes, ref, ow.works_piv, deform=True,
                                label=f"deform, shared {share}")


def test_deformation_lowers_error_against_truth():
    """Known field: the deformed multipass is closer to the synthetic truth
    than the reference's shifted windows."""
    from fluidimage_amd.synthetic import fourier_modes, make_pair
    from fluidimage_amd.works.piv.multipass import WorkPIV

    shape = (512, 512)
    rng = np.random.default_rng(123)
    field = fourier_modes(shape, rng, lmin=96.0, lmax=160.0, umax=6.0)
    im0, im1 = make_pair(shape, field, rng, density=0.03, dtype=np.uint16)
    errs = {}
    for deform in (None, "bicubic"):
        p = WorkPIV.create_default_params()
        p.piv0.shape_crop_im0 = 64
        p.multipass.number = 3
        p.multipass.use_tps = True
        p.multipass.smoothing_coef = 0.5
        p.multipass.image_deformation = deform
        res = WorkPIV(p).calcul_from_arrays(im0, im1)
        last = res.passes[-1]
        # displacement of the particles seen at the window centre: field at x - d/2
        ux, uy = field(last.xs - 0.5 * last.deltaxs, last.ys - 0.5 * last.deltays)
        ok = ~np.isnan(last.deltaxs)
        errs[deform] = float(np.sqrt(np.mean((last.deltaxs[ok] - ux[ok]) ** 2
                                             + (last.deltays[ok] - uy[ok]) ** 2)))
    print("rms error vs truth:", errs, flush=True)
    assert errs["bicubic"] < errs[None]
