"""HIP-graph capture of single-pass and multipass batches (and of uneven
batches, bucket after bucket): every replay, with eager work
and synchronisations between replays, gives bit for bit the eager results
(no memset nodes or host reads are recorded: the fills are kernels)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_graph_replays_equal_eager():
    import torch

    from fluidimage_amd.device import ImageSet, to_device_image
    from fluidimage_amd.synthetic import config_pair
    from fluidimage_amd.works.piv.multipass import WorkPIV
    from fluidimage_amd.works.piv.pipeline import GraphedPasses, MultipassPipeline

    print("case: graph capture single pass", flush=True)
    dev = torch.device("cuda")
    shape = (320, 384)
    pairs = [config_pair(2, shape=shape, pair_idx=k) for k in range(3)]
    others = [config_pair(2, shape=shape, pair_idx=10 + k) for k in range(3)]
    p = WorkPIV.create_default_params()
    p.piv0.shape_crop_im0 = 32
    p.piv0.displacement_max = 8
    work = WorkPIV(p)
    work.works_piv[0]._prepare_with_image(imshape=shape)
    ims = []
    for a, b in pairs:
        ims += [to_device_image(a, dev), to_device_image(b, dev)]
    imageset = ImageSet(ims)
    g = GraphedPasses(work, imageset, len(pairs), keep_maps=True)
    keys = ("dx", "dy", "correl_max", "err", "flags", "maps")

    def eager(image_pairs):
        e_ims = []
        for a, b in image_pairs:
            e_ims += [to_device_image(a, dev), to_device_image(b, dev)]
        out = MultipassPipeline(work).run(ImageSet(e_ims), len(image_pairs), keep_maps=True,
                                          final_interp=False)[0].out
        return {k: out[k].cpu().numpy() for k in keys}

    expect = {0: eager(pairs), 1: eager(others)}
    for rep, which in enumerate([0, 1, 0, 1]):
        src = pairs if which == 0 else others
        for k, (a, b) in enumerate(src):  # new image contents, same tensors
            imageset.images[2 * k].copy_(to_device_image(a, dev))
            imageset.images[2 * k + 1].copy_(to_device_image(b, dev))
        out = g.replay()[0].out
        torch.cuda.synchronize()
        got = {k: out[k].cpu().numpy() for k in keys}
        _ = eager(pairs[:1])  # eager work between replays
        for k in keys:
            assert np.array_equal(got[k], expect[which][k], equal_nan=True), (rep, k)


def test_graph_refuses_host_interpolation():
    from fluidimage_amd.works.piv.multipass import WorkPIV
    from fluidimage_amd.works.piv.pipeline import GraphedPasses

    p = WorkPIV.create_default_params()
    p.multipass.number = 2
    p.multipass.use_tps = "last"  # griddata between the passes: host scipy
    with pytest.raises(NotImplementedError, match="griddata"):
        GraphedPasses(WorkPIV(p), None, 1)


def _multipass_params(deform):
    from fluidimage_amd.works.piv.multipass import WorkPIV

    p = WorkPIV.create_default_params()
    p.piv0.shape_crop_im0 = 64
    p.multipass.number = 3
    p.multipass.use_tps = True
    p.multipass.subdom_size = 200
    p.multipass.smoothing_coef = 2.0
    p.multipass.image_deformation = "bicubic" if deform else None
    return p


def _host(passes, pipe, npairs):
    return [pipe.to_host(passes, k) for k in range(npairs)]


def _same(a, b, label):
    for ra, rb in zip(a, b):
        for pa, pb in zip(ra.passes, rb.passes):
            assert np.array_equal(pa.deltaxs, pb.deltaxs, equal_nan=True), label
            assert np.array_equal(pa.deltays, pb.deltays, equal_nan=True), label
            assert pa.errors == pb.errors, label
        assert np.array_equal(ra.passes[-1].deltaxs_final, rb.passes[-1].deltaxs_final), label
        assert np.array_equal(ra.passes[-1].deltays_smooth, rb.passes[-1].deltays_smooth), label


@pytest.mark.parametrize("deform, share", [(True, False), (False, False), (True, True)])
def test_graph_multipass_replays_equal_eager(deform, share, monkeypatch):
    """The whole 3-pass multipass with TPS between the passes and at the
    end (and image deformation) in one graph: four replays on alternating
    image contents, with eager work between them, give the eager bits
    (share: TPS factorisations shared between subdomains, forced at this
    size)."""
    import torch

    from fluidimage_amd.calcul import tps_planned
    from fluidimage_amd.device import ImageSet, to_device_image
    from fluidimage_amd.synthetic import config_pair
    from fluidimage_amd.works.piv.multipass import WorkPIV
    from fluidimage_amd.works.piv.pipeline import GraphedPasses, MultipassPipeline

    monkeypatch.setattr(tps_planned, "SHARED_MIN_SUBDOMAINS", 0 if share else 1 << 30)
    monkeypatch.setattr(tps_planned, "SHARED_MIN_BATCH", 0)
    print("case: graph capture multipass, deformation", deform, "shared", share, flush=True)
    dev = torch.device("cuda")
    shape = (384, 448)
    pairs = [config_pair(3, shape=shape, pair_idx=k) for k in range(3)]
    others = [config_pair(3, shape=shape, pair_idx=20 + k) for k in range(3)]
    p = _multipass_params(deform)
    work = WorkPIV(p)
    work._prepare_with_image(imshape=shape)

    def eager(image_pairs):
        e_ims = []
        for a, b in image_pairs:
            e_ims += [to_device_image(a, dev), to_device_image(b, dev)]
        pipe = MultipassPipeline(work)
        return _host(pipe.run(ImageSet(e_ims), len(image_pairs)), pipe, len(image_pairs))

    expect = {0: eager(pairs), 1: eager(others)}
    ims = []
    for a, b in pairs:
        ims += [to_device_image(a, dev), to_device_image(b, dev)]
    imageset = ImageSet(ims)
    g = GraphedPasses(work, imageset, len(pairs))
    for rep, which in enumerate([0, 1, 0, 1]):
        src = pairs if which == 0 else others
        for k, (a, b) in enumerate(src):
            imageset.images[2 * k].copy_(to_device_image(a, dev))
            imageset.images[2 * k + 1].copy_(to_device_image(b, dev))
        passes = g.replay()
        got = _host(passes, g.pipe, len(pairs))
        _ = eager(pairs[:1])
        _same(got, expect[which], f"replay {rep}")
    assert g.reruns == 0


def test_graph_reruns_pairs_beyond_capacity(monkeypatch):
    """Capacities too small for the fields: every replay flags them and runs
    the pairs again with host sizing, and the results are the eager bits."""
    import torch

    from fluidimage_amd.calcul import tps_planned
    from fluidimage_amd.device import ImageSet, to_device_image
    from fluidimage_amd.synthetic import config_pair
    from fluidimage_amd.works.piv.multipass import WorkPIV
    from fluidimage_amd.works.piv.pipeline import GraphedPasses, MultipassPipeline

    print("case: graph rerun beyond capacity", flush=True)
    dev = torch.device("cuda")
    shape = (320, 384)
    pairs = [config_pair(3, shape=shape, pair_idx=k) for k in range(2)]
    p = _multipass_params(True)
    work = WorkPIV(p)
    work._prepare_with_image(imshape=shape)
    ims = []
    for a, b in pairs:
        ims += [to_device_image(a, dev), to_device_image(b, dev)]
    # the pairs are run again with host sizing: the eager bits to expect are
    # those of a host-sized run (every subdomain factorised alone)
    pipe = MultipassPipeline(work)
    pipe.sizing = "host"
    expect = _host(pipe.run(ImageSet(ims), 2), pipe, 2)
    monkeypatch.setattr(tps_planned, "MARGIN_SLOTS", 0.3)
    imageset = ImageSet(ims)
    g = GraphedPasses(work, imageset, 2)
    for rep in range(2):
        got = _host(g.replay(), g.pipe, 2)
        _same(got, expect, f"replay {rep}")
    assert g.reruns == 4


def test_uneven_graph_equals_run_uneven():
    """Pairs of different shapes and schedules captured bucket after bucket
    in one graph: a replay gives the bits of the eager run_uneven."""
    from fluidimage_amd.synthetic import uneven_batch
    from fluidimage_amd.works.piv.multipass import WorkPIV
    from fluidimage_amd.works.piv.pipeline import UnevenGraph, run_uneven

    print("case: uneven batch graph", flush=True)
    items = uneven_batch(6, seed_offset=3, sizes=[(384, 384), (320, 448)])
    p = WorkPIV.create_default_params()
    p.multipass.use_tps = True
    p.multipass.smoothing_coef = 2.0
    expect = run_uneven(items, p)
    for concurrent in (False, True):
        g = UnevenGraph(items, p, concurrent=concurrent)
        for rep in range(3):
            g.replay()
            got = g.results()
            _same(got, expect, f"uneven replay {rep} (concurrent {concurrent})")
