"""Uneven batch (config-4 style): pairs of different sizes and window
schedules (128/64/32, 64/32/16, 48/24) with Oseen + shear fields, bucketed by
(shape, schedule) into batched pipelines, against the oracle's multipass of
each pair. Smaller image sizes than the benchmark configuration keep the
oracle to seconds. Bar: every vector of every pass within 1e-3 px with the
same error string, except the listed near-ties of the FP64 maps and the
vectors downstream of them."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_uneven_batch_vs_oracle():
    from fluidimage_amd.synthetic import uneven_batch
    from fluidimage_amd.works.piv.multipass import WorkPIV
    from fluidimage_amd.works.piv.pipeline import run_uneven
    from oracle.piv import OracleWorkPIV
    from tests.test_gpu_multipass import check_passes_against_oracle
    from tests.test_gpu_piv import _oracle_params

    items = uneven_batch(6, seed_offset=1, sizes=[(384, 384), (320, 448), (512, 512)])
    print("case: uneven batch", [(a.shape, s) for a, _, s in items], flush=True)
    p = WorkPIV.create_default_params()
    p.multipass.use_tps = True
    p.multipass.smoothing_coef = 2.0
    results = run_uneven(items, p)
    for (im0, im1, sched), res in zip(items, results):
        op = _oracle_params(p)
        op["piv0"]["shape_crop_im0"] = int(sched[0])
        op["multipass"]["number"] = len(sched)
        op["multipass"]["coeff_zoom"] = [sched[k] // sched[k + 1] for k in range(len(sched) - 1)]
        ow = OracleWorkPIV(op, precision="f64")
        ref = ow.calcul(im0, im1)
        assert len(res.passes) == len(ref) == len(sched)
        check_passes_against_oracle(res.passes, ref, ow.works_piv,
                                    label=f"{im0.shape} {tuple(sched)}")
