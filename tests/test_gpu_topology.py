"""The GPU-aware TopologyPIV end to end on a series of 16-bit TIFF files:
every couple's saved results equal WorkPIV on the same two images (batching
changes the execution, not the arithmetic), and "complete" computes only the
missing couples."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_topology_series_equals_workpiv(tmp_path):
    from PIL import Image

    from fluidimage_amd.synthetic import config_pair
    from fluidimage_amd.topologies.piv import TopologyPIV
    from fluidimage_amd.works.piv.multipass import WorkPIV

    print("case: topology on 5 TIFF files", flush=True)
    d = tmp_path / "Images"
    d.mkdir()
    a, b = config_pair(2, shape=(256, 320))
    c, e = config_pair(2, shape=(256, 320), pair_idx=1)
    ims = [a, b, c, e, a]
    for i, im in enumerate(ims):
        Image.fromarray(im).save(d / f"Oseen_center_{i + 1:02d}.tif")
    params = TopologyPIV.create_default_params()
    params.series.path = str(d)
    params.piv0.shape_crop_im0 = 32
    params.multipass.number = 2
    params.multipass.use_tps = True
    params.saving.how = "recompute"
    params.saving.postfix = "test_piv"
    params.gpu.batch_pairs = 3
    topo = TopologyPIV(params, logging_level="warning")
    topo.compute()
    assert len(topo.results) == 4
    work = WorkPIV(params)
    for k in range(4):
        ref = work.calcul_from_arrays(ims[k], ims[k + 1])
        path = topo.path_dir_result / f"piv_{k + 1:02d}-{k + 2:02d}.npz"
        with np.load(path) as z:
            for ip, p in enumerate(ref.passes):
                assert np.array_equal(z[f"piv{ip}/deltaxs"], p.deltaxs, equal_nan=True)
                assert np.array_equal(z[f"piv{ip}/deltays"], p.deltays, equal_nan=True)
                assert list(z[f"piv{ip}/ierrors"]) == sorted(p.errors)
            last = ref.passes[-1]
            assert np.array_equal(z[f"piv{len(ref.passes) - 1}/deltaxs_final"],
                                  last.deltaxs_final, equal_nan=True)
    for k in (0, 2):
        (topo.path_dir_result / f"piv_{k + 1:02d}-{k + 2:02d}.npz").unlink()
    params.saving.how = "complete"
    topo = TopologyPIV(params, logging_level="warning")
    topo.compute()
    assert sorted(p.name for p in topo.results) == ["piv_01-02.npz", "piv_03-04.npz"]
