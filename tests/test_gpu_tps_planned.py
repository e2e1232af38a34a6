"""Device-sized TPS batches (piv_tps_grid / piv_tps_offsets): the subdomain
grid computed on the GPU equals the host's numpy arithmetic bit for bit, a
device-sized batch gives the bits of the host-sized path field by field, and
a field beyond the capacities is flagged (and solved again by the host-sized
path inside the multipass pipeline, with the same results)."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def test_device_grid_equals_host_linspace_arithmetic():
    import torch

    from fluidimage_amd import _native
    from fluidimage_amd.calcul.tps import subdomain_grid

    rng = np.random.default_rng(3)
    nf, n, dom_cap = 400, 100000, 4096
    counts = rng.integers(50, n, nf)
    x0 = rng.uniform(-50, 500, nf)
    y0 = rng.uniform(-50, 500, nf)
    # ranges of every kind, including grid-like integers
    rx = np.where(rng.random(nf) < 0.5, rng.integers(64, 4000, nf).astype(float),
                  rng.uniform(10, 4000, nf))
    ry = np.where(rng.random(nf) < 0.5, rng.integers(64, 4000, nf).astype(float),
                  rng.uniform(10, 4000, nf))
    info = np.zeros((nf, 7))
    info[:, 0] = counts
    info[:, 1], info[:, 2] = x0, x0 + rx
    info[:, 3], info[:, 4] = y0, y0 + ry
    dev = torch.device("cuda")
    lib = _native.load()
    tinfo = torch.from_numpy(info).to(dev)
    ftab = torch.empty((nf, 7), dtype=torch.int32, device=dev)
    lim = torch.empty(nf * 4 * dom_cap, dtype=torch.float64, device=dev)
    dom_field = torch.empty(nf * dom_cap, dtype=torch.int32, device=dev)
    smoothing = torch.empty(nf * dom_cap, dtype=torch.float64, device=dev)
    fstat = torch.empty(nf, dtype=torch.int32, device=dev)
    for sub, coef in ((200, 2.0), (137.5, 5.0), (1000, 0.5)):
        _native.check(lib.piv_tps_grid(nf, n, _p(tinfo), float(sub), float(coef), 0.2, dom_cap,
                                       0.0, 0, _p(ftab), _p(lim), _p(dom_field), _p(smoothing),
                                       _p(fstat), None), "piv_tps_grid")
        torch.cuda.synchronize()
        ft, li, sm, st = (t.cpu().numpy() for t in (ftab, lim, smoothing, fstat))
        for f in range(nf):
            nsx, nsy, s_h, lim_h = subdomain_grid(int(counts[f]), info[f, 1], info[f, 2],
                                                  info[f, 3], info[f, 4], sub, coef, 20)
            assert st[f] == 0
            assert (ft[f, 0], ft[f, 1]) == (nsx, nsy), f
            got = li[f * 4 * dom_cap:f * 4 * dom_cap + lim_h.size]
            assert np.array_equal(got, lim_h), f
            assert sm[f * dom_cap] == s_h, f


def _fields(rng, nf, shape=(640, 720), step=8, holes=0.02):
    from tests.test_gpu_tps import _grid_field

    out = []
    for _ in range(nf):
        out.append(_grid_field(rng, step, shape, holes))
    return out


def test_planned_batch_equals_host_sized_batch():
    import torch

    from fluidimage_amd.calcul.tps import lattice_step, tps_device_batch
    from fluidimage_amd.calcul.tps_planned import (make_lattice_table, plan_capacity,
                                                    tps_planned)

    print("case: planned vs host-sized TPS batch", flush=True)
    rng = np.random.default_rng(51)
    dev = torch.device("cuda")
    hf = _fields(rng, 5)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    fields = [(t(u), t(v), t(xs), t(ys)) for xs, ys, u, v in hf]
    gx, gy = fields[0][2], fields[0][3]
    xs0, ys0 = hf[0][0], hf[0][1]
    for gen in (True, False):
        step = lattice_step(xs0, ys0)
        cap = plan_capacity(xs0, ys0, len(fields), 200, 2.0, 20, step=step, generated=gen)
        table = make_lattice_table(cap, dev)
        cat = lambda i: torch.cat([f[i] for f in fields])  # noqa: E731
        pl = tps_planned(cap, cat(0), cat(1), cat(2), cat(3), gx, gy, 200, 2.0, 1.5,
                         lattice_table=table)
        host = tps_device_batch(fields, gx, gy, 200, 2.0, 1.5, lattice=step)
        errs = pl.field_errors()
        tot = pl.totals.cpu().numpy()
        print(f"generated={gen}: totals {tot.tolist()}, capacity slots {cap.slots}, "
              f"mats {cap.mats}, n_max {cap.n_max}", flush=True)
        for k in range(len(fields)):
            assert errs[k] is None
            ax, ay = pl.approx(k)
            assert torch.equal(ax, host[k]["approx"][0]) and torch.equal(ay, host[k]["approx"][1])
            idx, sm = pl.host_field(k)
            assert np.array_equal(idx, host[k]["index"].cpu().numpy())
            assert np.array_equal(sm, host[k]["smoothed"].cpu().numpy())


def test_planned_batch_flags_failures_field_by_field():
    """An all-NaN field, a field with an empty subdomain and a field pushed
    beyond a small slot capacity are flagged; the others keep their bits."""
    import dataclasses

    import torch

    from fluidimage_amd.calcul.tps import lattice_step
    from fluidimage_amd.calcul.tps_planned import (FIELD_CAPACITY, FIELD_EMPTY_SUBDOMAIN,
                                                    FIELD_NO_VECTOR, make_lattice_table,
                                                    plan_capacity, tps_planned)

    print("case: planned TPS failures", flush=True)
    rng = np.random.default_rng(52)
    dev = torch.device("cuda")
    hf = _fields(rng, 5)
    xs, ys = hf[0][0], hf[0][1]
    hf[1][2][:] = np.nan
    hf[1][3][:] = np.nan
    gap = (xs > 100) & (xs < 620)
    hf[3][2][gap] = np.nan
    hf[3][3][gap] = np.nan
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    cat = lambda i: torch.cat([t(f[i]) for f in hf])  # noqa: E731
    gx, gy = t(xs), t(ys)
    step = lattice_step(xs, ys)
    # matrices read from the build (with generated ones, the gap field's
    # wider subdomains are flagged first, as off the table's reach)
    cap = plan_capacity(xs, ys, 5, 200, 2.0, 20, step=step, generated=False)
    table = make_lattice_table(cap, dev)
    args = (cat(2), cat(3), cat(0), cat(1), gx, gy, 200, 2.0, 1.5)
    full = tps_planned(cap, *args, lattice_table=table)
    fstat = full.fstat.cpu().numpy()
    assert fstat.tolist() == [0, FIELD_NO_VECTOR, 0, FIELD_EMPTY_SUBDOMAIN, 0]
    # room for two valid fields only: the third (field 4) is refused
    per_field = int(full.totals[0].item()) // 3
    small = dataclasses.replace(cap, slots=2 * per_field + per_field // 2)
    part = tps_planned(small, *args, lattice_table=table)
    assert part.fstat.cpu().numpy().tolist() == [0, FIELD_NO_VECTOR, 0, FIELD_EMPTY_SUBDOMAIN,
                                                 FIELD_CAPACITY]
    for k in (0, 2):
        assert torch.equal(part.approx(k)[0], full.approx(k)[0])
        assert torch.equal(part.approx(k)[1], full.approx(k)[1])
    # a failed field interpolates to 0 (no predictor downstream)
    assert torch.count_nonzero(part.approx(4)[0]) == 0


def test_pipeline_reruns_fields_beyond_capacity(monkeypatch):
    """Capacities too small for every field: the multipass pipeline flags
    them on the device, runs the pairs again with host sizing, and gives
    the bits of the host-sized pipeline."""
    from fluidimage_amd.calcul import tps_planned
    from fluidimage_amd.synthetic import config_pair
    from fluidimage_amd.works.piv.multipass import WorkPIV
    from fluidimage_amd.works.piv.pipeline import MultipassPipeline
    from tests.test_gpu_multipass import _setup

    print("case: pipeline rerun beyond capacity", flush=True)
    im0, im1, p, _ = _setup(True, number=3, window=64, shape=(384, 384))
    a0, a1 = config_pair(2, shape=(384, 384))
    pairs = [(im0, im1), (a0, a1)]
    monkeypatch.setattr(MultipassPipeline, "__init__", _host_sized_init(MultipassPipeline))
    host = WorkPIV(p).calcul_batch(pairs)
    monkeypatch.undo()
    monkeypatch.setattr(tps_planned, "MARGIN_SLOTS", 0.3)
    small = WorkPIV(p).calcul_batch(pairs)
    for a, b in zip(host, small):
        for pa, pb in zip(a.passes, b.passes):
            assert np.array_equal(pa.deltaxs, pb.deltaxs, equal_nan=True)
            assert pa.errors == pb.errors
        assert np.array_equal(a.passes[-1].deltaxs_final, b.passes[-1].deltaxs_final)


def _host_sized_init(cls):
    orig = cls.__init__

    def init(self, *args, **kwargs):
        orig(self, *args, **kwargs)
        self.sizing = "host"

    return init


@pytest.mark.parametrize("solve", [0, 1], ids=["explicit", "substitution"])
@pytest.mark.parametrize("holes, big_hole", [(0.01, False), (0.03, True)])
def test_shared_factorisations_match_oracle(holes, big_hole, solve):
    """Centers on a grid with missing nodes: subdomains of one node pattern
    share a factorisation and correct for their missing nodes (a large hole
    sends its subdomains to their own factorisation); smoothed and
    interpolated values within 1e-6 px of the oracle's LAPACK solves, outlier
    iteration counts identical, and within 1e-9 px of every subdomain
    factorised alone; with the shared solves by the explicit inverse of the
    pattern's saddle-point matrix (the default) and by blocked
    substitution."""
    from fluidimage_amd import tuning

    before = tuning.get(tuning.TPS_SHARED_SOLVE)
    tuning.set_value(tuning.TPS_SHARED_SOLVE, 0, 0, solve)
    try:
        _check_shared_factorisations(holes, big_hole, solve)
    finally:
        tuning.set_value(tuning.TPS_SHARED_SOLVE, 0, 0, before)


def _check_shared_factorisations(holes, big_hole, solve):
    import torch

    from fluidimage_amd.calcul.tps import lattice_step
    from fluidimage_amd.calcul.tps_planned import (PassGrid, make_lattice_table, plan_capacity,
                                                    tps_planned)
    from oracle.tps import ThinPlateSplineSubdom
    from tests.test_gpu_tps import _grid_field

    print("case: shared factorisations", holes, big_hole, solve, flush=True)
    rng = np.random.default_rng(61)
    dev = torch.device("cuda")
    shape, step = (640, 720), 8
    hf = [_grid_field(rng, step, shape, holes) for _ in range(3)]
    xs, ys = hf[0][0], hf[0][1]
    if big_hole:
        hole = (np.hypot(xs - 300, ys - 300) < 40)
        for f in hf:
            f[2][hole] = np.nan
            f[3][hole] = np.nan
    gxs = np.unique(xs)
    gys = np.unique(ys)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    cat = lambda i: torch.cat([t(f[i]) for f in hf])  # noqa: E731
    # new positions: a finer grid (the next pass)
    ny, nx = np.mgrid[2:shape[0] - 1:4, 2:shape[1] - 1:4]
    nxs, nys = nx.ravel().astype(float), ny.ravel().astype(float)
    lstep = lattice_step(xs, ys, nxs, nys)
    cap = plan_capacity(xs, ys, 3, 200, 2.0, 20, step=lstep, generated=True)
    table = make_lattice_table(cap, dev)
    args = (cap, cat(2), cat(3), cat(0), cat(1), t(nxs), t(nys), 200, 2.0, 1.5)
    grid = PassGrid(gxs, gys, dev)
    shared = tps_planned(*args, lattice_table=table, grid=grid)
    alone = tps_planned(*args, lattice_table=table)
    assert all(e is None for e in shared.field_errors())
    assert all(e is None for e in alone.field_errors())
    for k, (fx, fy, u, v) in enumerate(hf):
        ok = ~np.isnan(u)
        centers = np.vstack([fy[ok], fx[ok]])
        ref = ThinPlateSplineSubdom(centers, 200, 2.0, threshold=1.5)
        esx, wx, sumx = ref.compute_tps_weights_subdom(u[ok])
        esy, wy, sumy = ref.compute_tps_weights_subdom(v[ok])
        ref.init_with_new_positions(np.vstack([nys, nxs]))
        eix, eiy = ref.interpolate(wx), ref.interpolate(wy)
        idx, sm = shared.host_field(k)
        ax, ay = (a.cpu().numpy() for a in shared.approx(k))
        _, sm_alone = alone.host_field(k)
        bx, by = (a.cpu().numpy() for a in alone.approx(k))
        errs = (np.abs(sm[0] - esx).max(), np.abs(sm[1] - esy).max(), np.abs(ax - eix).max(),
                np.abs(ay - eiy).max())
        print(f"field {k}: |shared - LAPACK| {max(errs):.2e}, |shared - alone| "
              f"{max(np.abs(sm - sm_alone).max(), np.abs(ax - bx).max()):.2e}", flush=True)
        assert max(errs) < 1e-6
        assert np.abs(sm - sm_alone).max() < 1e-9 and np.abs(ax - bx).max() < 1e-9
        st = shared.stats[k * cap.dom_cap:(k + 1) * cap.dom_cap].cpu().numpy()
        active = st[:, 2] > 0
        assert st[active, 5].astype(int).tolist() == sumy["nb_iterations"]
        assert st[active, 2].astype(int).tolist() == sumx["nb_iterations"]
