"""piv_tps_interpolate_noalloc (the evaluation at new positions with the
library's fixed scratch, no workspace argument) gives bit for bit what
piv_tps_interpolate with a workspace gives, on two calls in a row, the second
smaller, and refuses a problem beyond its scratch."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _field(rng, shape, step, outliers=0.03, jitter=0.4):
    ny, nx = shape
    ys, xs = np.mgrid[step / 2:ny:step, step / 2:nx:step]
    ys = ys.ravel() + rng.uniform(-jitter, jitter, ys.size)
    xs = xs.ravel() + rng.uniform(-jitter, jitter, xs.size)
    u = 2.0 * np.sin(xs / 90.0) + 0.5 * np.cos(ys / 70.0)
    v = 1.5 * np.cos(xs / 60.0 + ys / 110.0)
    bad = rng.random(xs.size) < outliers
    u[bad] += rng.normal(0, 6, bad.sum())
    v[bad] += rng.normal(0, 6, bad.sum())
    return ys, xs, u.astype(np.float32), v.astype(np.float32)


def _run(entry, field, new):
    import torch

    from fluidimage_amd.calcul import tps

    ys, xs, u, v = field
    dev = torch.device("cuda")
    t = lambda a, d: torch.from_numpy(np.ascontiguousarray(a, dtype=d)).to(dev)  # noqa: E731
    saved = tps.INTERPOLATE_ENTRY
    tps.INTERPOLATE_ENTRY = entry
    try:
        out = tps.tps_device(t(u, np.float32), t(v, np.float32), t(xs, np.float64),
                             t(ys, np.float64), t(new[1], np.float64), t(new[0], np.float64),
                             200, 2.0, 1.5)
        torch.cuda.synchronize()
    finally:
        tps.INTERPOLATE_ENTRY = saved
    return out["approx"].cpu().numpy(), out["nb_subdom"]


@pytest.mark.parametrize("case", ["large_then_small"])
def test_noalloc_matches_workspace_entry(case):
    rng = np.random.default_rng(7)
    # first call: 1536 x 1536 of centres every 6 px (65,536 of them), 18,496
    # new positions every 11 px; second call smaller: 512 x 640 every 16 px
    calls = [((1536, 1536), 6.0, 11.0), ((512, 640), 16.0, 13.0)]
    for shape, step, new_step in calls:
        print(f"case: noalloc {case} shape={shape} step={step}", flush=True)
        field = _field(rng, shape, step)
        gy, gx = np.mgrid[new_step / 2:shape[0]:new_step, new_step / 2:shape[1]:new_step]
        new = np.vstack([gy.ravel(), gx.ravel()])
        ref, nd_ref = _run("workspace", field, new)
        got, nd_got = _run("noalloc", field, new)
        assert nd_ref == nd_got and nd_ref > 1
        assert np.isfinite(ref).all()
        assert np.array_equal(got, ref), np.abs(got - ref).max()


def test_noalloc_refuses_more_than_its_scratch():
    """A problem whose workspace exceeds the fixed scratch is refused with
    PIV_EUNSUPPORTED before anything is launched (no pointer is read)."""
    import ctypes

    import torch  # noqa: F401  (initialises the device)

    from fluidimage_amd import _native

    lib = _native.load()
    npts, nd = 1_000_000, 500  # 80 bytes of workspace per point: beyond 64 MiB
    assert lib.piv_tps_interpolate_workspace(npts, nd) > (64 << 20)
    dummy = ctypes.c_void_p(16)
    st = lib.piv_tps_interpolate_noalloc(npts, dummy, dummy, dummy, dummy, dummy, dummy, 1, nd,
                                         16, dummy, dummy, 1, None, dummy, None, None)
    assert st == 2
    assert b"scratch" in lib.piv_last_error()
