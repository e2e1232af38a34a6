"""The uneven batch at its stated sizes: 32 uint16 pairs of 1024 x 1024,
1200 x 1600, 2048 x 2048 and 2160 x 2560 px with window schedules drawn from
128/64/32, 64/32/16 and 48/24, Oseen vortices plus shear flows of 3-12 px.

* every pair of the batch gives the bits it gives when run alone;
* every pair's last pass follows the truth field it was rendered with;
* two pairs of different (shape, schedule) match the oracle's multipass
  (every miss a listed near-tie or downstream of one)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_CACHE = {}


def _batch():
    if "items" not in _CACHE:
        from fluidimage_amd.synthetic import uneven_batch

        _CACHE["items"], _CACHE["fields"] = uneven_batch(32, return_fields=True)
    return _CACHE["items"], _CACHE["fields"]


def _params():
    from fluidimage_amd.works.piv.multipass import WorkPIV

    p = WorkPIV.create_default_params()
    p.multipass.use_tps = True
    p.multipass.smoothing_coef = 2.0
    p.multipass.subdom_size = 200
    return p


def _results():
    if "results" not in _CACHE:
        from fluidimage_amd.works.piv.pipeline import run_uneven

        items, _ = _batch()
        print("case: uneven batch of 32 at full size", flush=True)
        _CACHE["results"] = run_uneven(items, _params())
    return _CACHE["results"]


def test_uneven_full_size_pairs_equal_alone_and_follow_truth():
    from fluidimage_amd.works.piv.pipeline import run_uneven

    items, fields = _batch()
    results = _results()
    assert len(results) == 32
    for k, ((im0, im1, sched), res, field) in enumerate(zip(items, results, fields)):
        alone = run_uneven([(im0, im1, sched)], _params())[0]
        for pa, pb in zip(alone.passes, res.passes):
            assert np.array_equal(pa.deltaxs, pb.deltaxs, equal_nan=True), k
            assert np.array_equal(pa.deltays, pb.deltays, equal_nan=True), k
            assert pa.errors == pb.errors, k
        assert np.array_equal(alone.passes[-1].deltaxs_final, res.passes[-1].deltaxs_final)
        last = res.passes[-1]
        ux, uy = field(last.xs - 0.5 * last.deltaxs, last.ys - 0.5 * last.deltays)
        ok = ~np.isnan(last.deltaxs)
        err = np.hypot(last.deltaxs[ok] - ux[ok], last.deltays[ok] - uy[ok])
        med, p95 = float(np.median(err)), float(np.percentile(err, 95))
        print(f"pair {k} {im0.shape} {sched}: {ok.mean() * 100:.2f} % valid, error vs truth "
              f"median {med:.3f} px, 95th percentile {p95:.3f} px", flush=True)
        assert ok.mean() > 0.9
        # Oseen + shear fields of 3-12 px: about 0.15 px median and 0.35 px
        # 95th percentile at the last windows (16-32 px)
        assert med < 0.25 and p95 < 0.5


def _pick_two(items):
    """The smallest pair of each of two different (shape, schedule) groups."""
    seen = {}
    for k, (im0, _, sched) in enumerate(items):
        seen.setdefault((im0.shape, tuple(sched)), k)
    keys = sorted(seen, key=lambda key: key[0][0] * key[0][1] * len(key[1]))
    first = keys[0]
    second = next(key for key in keys if key[0] != first[0] and key[1] != first[1])
    return [seen[first], seen[second]]


def test_uneven_full_size_two_pairs_vs_oracle():
    from oracle.piv import OracleWorkPIV
    from tests.test_gpu_multipass import check_passes_against_oracle
    from tests.test_gpu_piv import _oracle_params

    items, _ = _batch()
    results = _results()
    for k in _pick_two(items):
        im0, im1, sched = items[k]
        print(f"case: oracle of pair {k} {im0.shape} {sched}", flush=True)
        op = _oracle_params(_params())
        op["piv0"]["shape_crop_im0"] = int(sched[0])
        op["multipass"]["number"] = len(sched)
        op["multipass"]["coeff_zoom"] = [sched[j] // sched[j + 1] for j in range(len(sched) - 1)]
        ow = OracleWorkPIV(op, precision="f64")
        ref = ow.calcul(im0, im1)
        check_passes_against_oracle(results[k].passes, ref, ow.works_piv,
                                    label=f"pair {k} {im0.shape} {tuple(sched)}")
