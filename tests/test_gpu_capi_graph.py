"""Graph capture through the C ABI (piv_graph_begin / end / launch /
destroy): the image minima and the fused correlation of a batch, recorded
on a caller stream and replayed on new image contents, give bit for bit the
results of eager calls on the same images."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _images(rng, shape, npairs, shift):
    im0 = [rng.integers(0, 4096, shape).astype(np.uint16) for _ in range(npairs)]
    im1 = [np.clip(np.roll(a, shift, axis=(0, 1)).astype(np.int32)
                   + rng.integers(-40, 40, shape), 0, 4095).astype(np.uint16) for a in im0]
    return [x for pair in zip(im0, im1) for x in pair]


def test_capi_graph_replays_equal_eager():
    import torch

    from fluidimage_amd import _native
    from fluidimage_amd.device import ImageSet, to_device_image
    from fluidimage_amd.works.piv.multipass import WorkPIV

    print("case: C-ABI graph, minima + correlation 32x32, 3 pairs", flush=True)
    dev = torch.device("cuda")
    lib = _native.load()
    rng = np.random.default_rng(31)
    shape, npairs = (256, 320), 3
    contents = [_images(rng, shape, npairs, (2, -3)), _images(rng, shape, npairs, (-1, 4))]
    params = WorkPIV.create_default_params()
    params.piv0.shape_crop_im0 = 32
    work = WorkPIV(params).works_piv[0]
    work._prepare_with_image(imshape=shape)
    correl = work.correl
    nwin = work.nb_vectors * npairs
    windows = work.make_windows(0, 1, None, None, device=dev, npairs=npairs)["windows"]
    imageset = ImageSet([to_device_image(a, dev) for a in contents[0]])
    s = torch.cuda.Stream()
    sp = ctypes.c_void_p(s.cuda_stream)
    maps = torch.empty((nwin, 32, 32), dtype=torch.float32, device=dev)
    out = correl._alloc_results(nwin, dev)

    def step(stream):
        imageset.recompute_min(stream)
        correl.correlate_windows(imageset.table, windows, nwin, imageset.dtype, True, maps=maps,
                                 out=out, stream=stream)

    def load(k):
        for i, a in enumerate(contents[k]):
            imageset.images[i].copy_(to_device_image(a, dev))
        torch.cuda.synchronize()

    def snapshot():
        got = {k: v.cpu().numpy() for k, v in out.items() if torch.is_tensor(v)}
        got["maps"] = maps.cpu().numpy()
        return got

    expect = {}
    for k in (0, 1):  # eager results, also the warm-up before the capture
        load(k)
        step(s)
        s.synchronize()
        expect[k] = snapshot()
    assert int((expect[0]["err"] == 0).sum()) > nwin // 2

    _native.check(lib.piv_graph_begin(sp), "piv_graph_begin")
    graph = ctypes.c_void_p()
    try:
        step(s)  # recorded, not run
    finally:  # a failed capture is still ended
        rc = lib.piv_graph_end(sp, ctypes.byref(graph))
    _native.check(rc, "piv_graph_end")
    assert graph.value
    try:
        for rep, k in enumerate([0, 1, 1, 0]):
            load(k)
            maps.fill_(-7.0)  # stale results would show
            for v in out.values():
                if torch.is_tensor(v):
                    v.zero_()
            torch.cuda.synchronize()
            _native.check(lib.piv_graph_launch(graph, sp), "piv_graph_launch")
            s.synchronize()
            got = snapshot()
            for key, ref in expect[k].items():
                assert np.array_equal(got[key], ref, equal_nan=True), (rep, key)
    finally:
        _native.check(lib.piv_graph_destroy(graph), "piv_graph_destroy")
