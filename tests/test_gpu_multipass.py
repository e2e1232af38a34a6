"""GPU multipass pipeline against the oracle's multipass chain.

Bars: every pass's displacements within 1e-3 px and identical error
strings, and the final interpolated fields within 1e-3 px, except the
listed windows whose FP64 map has a near-tie and the vectors downstream of
them (their WorkFIX neighbourhood, the interpolation's reach of them in the
next pass): every miss must be one of those."""

import numpy as np
import pytest

from tests.helpers import allowed_mismatches, near_tie_windows

pytestmark = pytest.mark.gpu


def _setup(use_tps, number=3, window=64, shape=(512, 512), nb_peaks=1):
    from fluidimage_amd.synthetic import config_pair
    from fluidimage_amd.works.piv.multipass import WorkPIV
    from tests.test_gpu_piv import _oracle_params

    im0, im1 = config_pair(3, shape=shape)
    p = WorkPIV.create_default_params()
    p.piv0.shape_crop_im0 = window
    p.piv0.nb_peaks_to_search = nb_peaks
    p.multipass.number = number
    p.multipass.use_tps = use_tps
    p.multipass.smoothing_coef = 2.0
    p.multipass.subdom_size = 200
    return im0, im1, p, _oracle_params(p)


def oracle_ties(passes, rel=2e-5):
    """Near-tie windows of every pass of the oracle's run (FP64 maps; a
    window without a map, a bad crop, has no tie)."""
    out = []
    for e in passes:
        maps = e["correls"]
        shape = next(m.shape for m in maps if m is not None)
        full = np.stack([m if m is not None else np.zeros(shape) for m in maps])
        out.append(near_tie_windows(full, rel=rel))
    return out


def check_passes_against_oracle(gpu_passes, ref, works, deform=False, label=""):
    """Every vector of every pass within 1e-3 px of the oracle with the same
    error string, and the final field within 1e-3 px, except the listed
    near-tie windows and what lies downstream of them (allowed_mismatches);
    prints the ties and the misses."""
    ties = oracle_ties(ref)
    allowed, allowed_final = allowed_mismatches(ref, works, ties, deform=deform)
    for ip, (g, e) in enumerate(zip(gpu_passes, ref)):
        both_nan = np.isnan(g.deltaxs) & np.isnan(e["deltaxs"])
        close = (np.abs(g.deltaxs - e["deltaxs"]) <= 1e-3) & (
            np.abs(g.deltays - e["deltays"]) <= 1e-3)
        bad = set(np.flatnonzero(~(both_nan | close)).tolist())
        err_bad = {i for i in range(len(g.deltaxs)) if g.errors.get(i) != e["errors"].get(i)}
        print(f"{label} pass {ip}: {len(g.deltaxs)} vectors, near-ties {sorted(ties[ip])}, "
              f"allowed {len(allowed[ip])}, misses {sorted(bad)[:20]}, "
              f"error-string misses {sorted(err_bad)[:20]}", flush=True)
        assert bad <= allowed[ip], sorted(bad - allowed[ip])[:20]
        assert err_bad <= allowed[ip], sorted(err_bad - allowed[ip])[:20]
    gf, ef = gpu_passes[-1], ref[-1]
    if getattr(gf, "deltaxs_final", None) is not None and "deltaxs_final" in ef:
        df = np.maximum(np.abs(gf.deltaxs_final - ef["deltaxs_final"]),
                        np.abs(gf.deltays_final - ef["deltays_final"]))
        fbad = set(np.flatnonzero(~(df <= 1e-3)).tolist())
        print(f"{label} final field: max |diff| {np.nanmax(df):.3g}, misses {sorted(fbad)[:20]}, "
              f"allowed {len(allowed_final)}", flush=True)
        assert fbad <= allowed_final, sorted(fbad - allowed_final)[:20]


@pytest.mark.parametrize("use_tps", [True, "last", False])
def test_multipass_vs_oracle(use_tps):
    from fluidimage_amd.works.piv.multipass import WorkPIV
    from oracle.piv import OracleWorkPIV

    print("case: multipass", use_tps, flush=True)
    im0, im1, p, op = _setup(use_tps)
    gpu = WorkPIV(p).calcul_from_arrays(im0, im1)
    ow = OracleWorkPIV(op, precision="f64")
    ref = ow.calcul(im0, im1)
    assert len(gpu.passes) == len(ref) == 3
    check_passes_against_oracle(gpu.passes, ref, ow.works_piv, label=f"use_tps={use_tps}")


def test_multipass_window_shapes_without_register_kernel():
    """A 60 -> 30 px multipass (neither size has a register-FFT kernel: the
    DFT kernel of any shape serves both passes) against the oracle."""
    from fluidimage_amd.works.piv.multipass import WorkPIV
    from oracle.piv import OracleWorkPIV

    print("case: multipass 60/30", flush=True)
    im0, im1, p, op = _setup("last", number=2, window=60, shape=(384, 416))
    gpu = WorkPIV(p).calcul_from_arrays(im0, im1)
    ow = OracleWorkPIV(op, precision="f64")
    ref = ow.calcul(im0, im1)
    assert [w.shape_crop_im0 for w in ow.works_piv] == [(60, 60), (30, 30)]
    check_passes_against_oracle(gpu.passes, ref, ow.works_piv, label="60/30")


def test_batch_equals_single_pairs():
    """Batching pairs changes the execution, not the arithmetic."""
    from fluidimage_amd.synthetic import config_pair
    from fluidimage_amd.works.piv.multipass import WorkPIV

    im0, im1, p, _ = _setup(True, number=2, window=32, shape=(384, 384))
    a0, a1 = config_pair(2, shape=(384, 384))
    w = WorkPIV(p)
    batch = w.calcul_batch([(im0, im1), (a0, a1), (im1, im0)])
    for (x0, x1), b in zip([(im0, im1), (a0, a1), (im1, im0)], batch):
        single = WorkPIV(p).calcul_from_arrays(x0, x1)
        for ps, pb in zip(single.passes, b.passes):
            assert np.array_equal(ps.deltaxs, pb.deltaxs, equal_nan=True)
            assert np.array_equal(ps.deltays, pb.deltays, equal_nan=True)
            assert ps.errors == pb.errors
        assert np.array_equal(single.passes[-1].deltaxs_final, b.passes[-1].deltaxs_final)


def test_deferred_failure_reruns_the_pair(monkeypatch):
    """A pair whose deferred TPS check fails is run again alone and its
    results replace the batch's; forced here for pair 1 (whose system is
    fine), so the spliced results equal the plain batch's bit for bit."""
    from fluidimage_amd.synthetic import config_pair
    from fluidimage_amd.works.piv.multipass import WorkPIV
    from fluidimage_amd.works.piv.pipeline import MultipassPipeline

    print("case: deferred rerun", flush=True)
    im0, im1, p, _ = _setup(True, number=2, window=32, shape=(384, 384))
    a0, a1 = config_pair(2, shape=(384, 384))
    pairs = [(im0, im1), (a0, a1), (im1, im0)]
    plain = WorkPIV(p).calcul_batch(pairs)
    calls = []
    orig = MultipassPipeline._check_deferred

    def forced(self):
        failed = orig(self)
        if not calls:
            calls.append(1)
            failed = failed | {1}
        return failed

    monkeypatch.setattr(MultipassPipeline, "_check_deferred", forced)
    spliced = WorkPIV(p).calcul_batch(pairs)
    assert calls == [1]
    for a, b in zip(plain, spliced):
        for pa, pb in zip(a.passes, b.passes):
            assert np.array_equal(pa.deltaxs, pb.deltaxs, equal_nan=True)
            assert np.array_equal(pa.deltays, pb.deltays, equal_nan=True)
            assert pa.errors == pb.errors
        assert np.array_equal(a.passes[-1].deltaxs_final, b.passes[-1].deltaxs_final)
        assert np.array_equal(a.passes[-1].deltays_smooth, b.passes[-1].deltays_smooth)


def test_strcrop_mask_vs_oracle():
    """params.mask.strcrop (src/fluidimage/works/with_mask.py:15-47): the
    images are cropped before the passes and the positions are reported in
    the coordinates of the original images."""
    from fluidimage_amd.works.piv.multipass import WorkPIV
    from oracle.piv import OracleWorkPIV

    print("case: strcrop", flush=True)
    im0, im1, p, _ = _setup(True, number=2, window=32, shape=(448, 512))
    p.mask.strcrop = "20:430, 37:"
    from tests.test_gpu_piv import _oracle_params

    op = _oracle_params(p)
    gpu = WorkPIV(p).calcul_from_arrays(im0, im1)
    ref = OracleWorkPIV(op, precision="f64").calcul(im0, im1)
    for g, e in zip(gpu.passes, ref):
        assert np.array_equal(np.asarray(g.xs, float), np.asarray(e["xs"], float))
        assert np.array_equal(np.asarray(g.ys, float), np.asarray(e["ys"], float))
        assert float(np.min(g.xs)) >= 37 and float(np.min(g.ys)) >= 20
        both_nan = np.isnan(g.deltaxs) & np.isnan(e["deltaxs"])
        ok = both_nan | ((np.abs(g.deltaxs - e["deltaxs"]) <= 1e-3)
                         & (np.abs(g.deltays - e["deltays"]) <= 1e-3))
        print(f"strcrop pass: {ok.mean() * 100:.3f} % within 1e-3 px", flush=True)
        assert ok.mean() >= 0.995
    df = np.abs(gpu.passes[-1].deltaxs_final - ref[-1]["deltaxs_final"])
    assert (df <= 1e-3).mean() >= 0.99


def test_side_stream_matches_default_stream():
    """run_pairs on a fresh side stream (allocations, uploads, launches and
    host reads all ordered on it) gives bit for bit the default stream's
    results."""
    import torch

    from fluidimage_amd.synthetic import config_pair
    from fluidimage_amd.works.piv.multipass import WorkPIV
    from fluidimage_amd.works.piv.pipeline import run_pairs

    print("case: side stream", flush=True)
    im0, im1, p, _ = _setup(True, number=3, window=64, shape=(384, 384))
    a0, a1 = config_pair(2, shape=(384, 384))
    pairs = [(im0, im1), (a0, a1)]
    w = WorkPIV(p)
    pipe, ref = run_pairs(w, pairs)
    expected = [pipe.to_host(ref, k) for k in range(2)]
    side = torch.cuda.Stream()
    pipe2, got = run_pairs(WorkPIV(p), pairs, stream=side)
    side.synchronize()
    for k in range(2):
        r = pipe2.to_host(got, k)
        for pa, pb in zip(expected[k].passes, r.passes):
            assert np.array_equal(pa.deltaxs, pb.deltaxs, equal_nan=True)
            assert np.array_equal(pa.deltays, pb.deltays, equal_nan=True)
            assert pa.errors == pb.errors
        assert np.array_equal(expected[k].passes[-1].deltaxs_final, r.passes[-1].deltaxs_final)
