"""HIP-graph capture of a single-pass batch: every replay, with eager work
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


def test_graph_refuses_the_multipass():
    from fluidimage_amd.works.piv.multipass import WorkPIV
    from fluidimage_amd.works.piv.pipeline import GraphedPasses

    p = WorkPIV.create_default_params()
    p.multipass.number = 2
    with pytest.raises(NotImplementedError, match="interpolation"):
        GraphedPasses(WorkPIV(p), None, 1)
