"""Host logic of the GPU-aware TopologyPIV: series of couples from file
names (topologies/piv.py:82-88, fluiddyn's SeriesOfArrays), result names
(data_objects/piv.py:37-57), saving modes (topologies/base.py:49-104) and
the "complete" resume (topologies/piv.py:160-167), with a stand-in for the
GPU work (no GPU here)."""

import numpy as np
import pytest


def _write(dirpath, names, shape=(48, 64), dtype=np.uint16):
    from PIL import Image

    rng = np.random.default_rng(3)
    for n in names:
        arr = rng.integers(0, 4000, shape).astype(dtype)
        Image.fromarray(arr).save(dirpath / n)


def test_series_numbered_and_lettered(tmp_path):
    from fluidimage_amd.topologies.piv import SeriesOfCouples

    d = tmp_path / "num"
    d.mkdir()
    _write(d, [f"Oseen_center_{i:02d}.tif" for i in range(1, 5)])
    s = SeriesOfCouples(d)
    assert s.str_subset == "i:i+2"
    assert s.couples == [("Oseen_center_01.tif", "Oseen_center_02.tif"),
                         ("Oseen_center_02.tif", "Oseen_center_03.tif"),
                         ("Oseen_center_03.tif", "Oseen_center_04.tif")]
    assert s.names == ["piv_01-02", "piv_02-03", "piv_03-04"]
    s = SeriesOfCouples(d, "i:i+3:2", ind_start=1, ind_stop=3)
    assert s.couples == [("Oseen_center_01.tif", "Oseen_center_03.tif"),
                         ("Oseen_center_02.tif", "Oseen_center_04.tif")]
    s = SeriesOfCouples(d, "i:i+2", ind_start=1, ind_step=2)
    assert s.names == ["piv_01-02", "piv_03-04"]
    with pytest.raises(ValueError):
        SeriesOfCouples(d, "i:i+3")
    d = tmp_path / "let"
    d.mkdir()
    _write(d, ["c060a.png", "c060b.png", "c061a.png", "c061b.png"], dtype=np.uint8)
    s = SeriesOfCouples(d)
    assert s.str_subset == "i, 0:2"
    assert s.couples == [("c060a.png", "c060b.png"), ("c061a.png", "c061b.png")]
    assert s.names == ["piv_060ab", "piv_061ab"]


def test_imread_keeps_the_image_type(tmp_path):
    from PIL import Image

    from fluidimage_amd.topologies.piv import imread

    a = (np.arange(12 * 10).reshape(12, 10) * 500).astype(np.uint16)
    Image.fromarray(a).save(tmp_path / "a.tif")
    b = imread(tmp_path / "a.tif")
    assert b.dtype == np.uint16 and np.array_equal(a, b)
    c = (np.arange(120).reshape(12, 10) * 2).astype(np.uint8)
    Image.fromarray(c).save(tmp_path / "c.png")
    assert np.array_equal(imread(tmp_path / "c.png"), c)


class _Result:
    def __init__(self, a, b):
        self.deltaxs = np.array([float(a.mean())], np.float32)
        self.deltays = np.array([float(b.mean())], np.float32)
        self.xs = self.ys = np.zeros(1)
        self.correls_max = np.ones(1, np.float32)
        self.errors = {0: "none"}


class _Multi:
    def __init__(self, r):
        self.passes = [r]


class _StubWork:
    """Stand-in for the GPU WorkPIV: records its batches."""

    def __init__(self):
        self.batches = []

    def calcul_batch(self, pairs, keep_maps=False):
        self.batches.append(len(pairs))
        return [_Multi(_Result(a, b)) for a, b in pairs]


def test_topology_recompute_then_complete(tmp_path):
    from fluidimage_amd.topologies.piv import TopologyPIV

    d = tmp_path / "Images"
    d.mkdir()
    _write(d, [f"im_{i:03d}.tif" for i in range(9)])
    params = TopologyPIV.create_default_params()
    params.series.path = str(d)
    params.saving.how = "recompute"
    params.saving.postfix = "test_piv"
    params.gpu.batch_pairs = 3
    topo = TopologyPIV(params, logging_level="warning")
    assert topo.path_dir_result == tmp_path / "Images.test_piv"
    topo.work_piv = _StubWork()
    topo.compute()
    assert len(topo.results) == 8 and topo.work_piv.batches == [3, 3, 2]
    files = sorted(topo.path_dir_result.glob("piv_*.npz"))
    assert [f.name for f in files][:2] == ["piv_000-001.npz", "piv_001-002.npz"]
    with np.load(files[0]) as z:
        assert list(z["names"]) == ["im_000.tif", "im_001.tif"]
        assert z["piv0/errors"][0] == "none" and z["piv0/ierrors"][0] == 0
    # "complete": only the couples whose result is missing
    files[1].unlink()
    files[5].unlink()
    params.saving.how = "complete"
    topo = TopologyPIV(params, logging_level="warning")
    topo.work_piv = _StubWork()
    topo.compute()
    assert len(topo.results) == 2 and topo.work_piv.batches == [2]
    topo = TopologyPIV(params, logging_level="warning")
    topo.work_piv = _StubWork()
    topo.compute()  # work already done
    assert topo.results == [] and topo.work_piv.batches == []
    # "new_dir" makes Images.test_piv0
    params.saving.how = "new_dir"
    topo = TopologyPIV(params, logging_level="warning")
    assert topo.path_dir_result.name == "Images.test_piv0"


class _FailingWork(_StubWork):
    """A stand-in whose work raises on one couple (its first image's mean
    is the marker)."""

    def calcul_batch(self, pairs, keep_maps=False):
        self.batches.append(len(pairs))
        if any(a[0, 0] == 7 for a, _ in pairs):
            raise ValueError("Only nan.")
        return [_Multi(_Result(a, b)) for a, b in pairs]


def test_topology_goes_on_after_a_failed_couple(tmp_path):
    from PIL import Image

    from fluidimage_amd.topologies.piv import TopologyPIV

    d = tmp_path / "Images"
    d.mkdir()
    _write(d, [f"im_{i:03d}.tif" for i in range(5)])
    a = np.full((48, 64), 7, np.uint16)
    Image.fromarray(a).save(d / "im_002.tif")
    params = TopologyPIV.create_default_params()
    params.series.path = str(d)
    params.saving.how = "recompute"
    topo = TopologyPIV(params, logging_level="critical")
    topo.work_piv = _FailingWork()
    topo.compute()
    # couple 002-003 fails: the batch of 4 is run again couple by couple
    assert sorted(p.name for p in topo.results) == ["piv_000-001.npz", "piv_001-002.npz",
                                                    "piv_003-004.npz"]
    assert list(topo.errors) == ["piv_002-003"]
    assert topo.work_piv.batches == [4, 1, 1, 1, 1]
