"""Differentiable refinement on the GPU against the FP64 torch-CPU oracle
(oracle/refine.py) and central finite differences.

Bars: loss within 1e-9 relative of the oracle; gradient on 256 random nodes
within 1e-7 relative (to the largest component) of the oracle's autograd
gradient; central differences (step 1e-3 px) of the GPU loss within 1e-4;
refinement never raises the loss of a pair (steps are kept only when the
loss drops) and lowers it on the synthetic pairs."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem(shape, window, npairs, seed=0):
    import torch

    from fluidimage_amd.device import ImageSet, to_device_image
    from fluidimage_amd.synthetic import config_pair
    from fluidimage_amd.works.piv.multipass import WorkPIV
    from fluidimage_amd.works.piv.refine import RefineProblem

    p = WorkPIV.create_default_params()
    p.piv0.shape_crop_im0 = window
    work = WorkPIV(p).works_piv[0]
    work._prepare_with_image(imshape=shape)
    pairs = [config_pair(3, pair_idx=seed + k, shape=shape) for k in range(npairs)]
    dev = torch.device("cuda")
    ims = []
    for a, b in pairs:
        ims += [to_device_image(a, dev), to_device_image(b, dev)]
    images = ImageSet(ims)
    return work, pairs, RefineProblem(work, images, npairs)


def test_loss_and_gradient_match_oracle():
    import torch

    from fluidimage_amd.works.piv.refine import refine_loss
    from oracle.refine import refine_loss as oracle_loss

    print("case: refine loss/gradient vs oracle", flush=True)
    work, pairs, prob = _problem((160, 176), 32, 2)
    nvec = work.nb_vectors
    rng = np.random.default_rng(3)
    d = rng.normal(0, 1.5, (2, 2, nvec))
    dev = torch.device("cuda")
    x = torch.tensor(d[:, 0], device=dev, requires_grad=True)
    y = torch.tensor(d[:, 1], device=dev, requires_grad=True)
    loss = refine_loss(x, y, prob)
    gx, gy = torch.autograd.grad(loss.sum(), (x, y))
    for k, (a, b) in enumerate(pairs):
        ox = torch.tensor(d[k, 0], requires_grad=True)
        oy = torch.tensor(d[k, 1], requires_grad=True)
        ol = oracle_loss(a, b, work.iyvecs, work.ixvecs, work.shape_crop_im0, ox, oy)
        ogx, ogy = torch.autograd.grad(ol, (ox, oy))
        assert abs(float(loss[k].detach()) - float(ol.detach())) <= 1e-9 * abs(float(ol.detach()))
        g = np.concatenate([gx[k].cpu().numpy(), gy[k].cpu().numpy()])
        og = np.concatenate([ogx.numpy(), ogy.numpy()])
        idx = rng.choice(g.size, min(256, g.size), replace=False)
        scale = np.abs(og).max()
        assert np.abs(g[idx] - og[idx]).max() <= 1e-7 * scale, np.abs(g - og).max() / scale


def test_gradient_matches_central_differences():
    import torch

    from fluidimage_amd.works.piv.refine import refine_loss

    print("case: refine gradient vs central differences", flush=True)
    work, pairs, prob = _problem((160, 176), 32, 1, seed=4)
    nvec = work.nb_vectors
    rng = np.random.default_rng(8)
    dev = torch.device("cuda")
    d = torch.tensor(rng.normal(0, 1.0, (2, 1, nvec)), device=dev)
    x = d[0].clone().requires_grad_(True)
    y = d[1].clone().requires_grad_(True)
    gx, gy = torch.autograd.grad(refine_loss(x, y, prob).sum(), (x, y))
    h = 1e-3
    for k in rng.choice(nvec, 16, replace=False):
        for comp, g in ((0, gx), (1, gy)):
            e = torch.zeros_like(d)
            e[comp, 0, k] = h
            lp, _ = prob.forward((d[0] + e[0]).contiguous(), (d[1] + e[1]).contiguous())
            lm, _ = prob.forward((d[0] - e[0]).contiguous(), (d[1] - e[1]).contiguous())
            fd = float(lp[0] - lm[0]) / (2 * h)
            assert abs(fd - float(g[0, k])) <= 1e-4 * max(1.0, abs(fd)), (k, comp, fd)


def test_refinement_keeps_only_improving_steps():
    import torch

    from fluidimage_amd.synthetic import config_pair
    from fluidimage_amd.works.piv.multipass import WorkPIV
    from fluidimage_amd.works.piv.refine import RefineProblem, refine
    from fluidimage_amd.device import ImageSet, to_device_image

    print("case: refinement on multipass results", flush=True)
    shape = (256, 256)
    p = WorkPIV.create_default_params()
    p.piv0.shape_crop_im0 = 32
    p.multipass.number = 2
    p.multipass.use_tps = True
    p.multipass.smoothing_coef = 2.0
    pairs = [config_pair(2, pair_idx=k, shape=shape) for k in range(3)]
    work = WorkPIV(p)
    results = work.calcul_batch(pairs)
    last = work.works_piv[-1]
    dev = torch.device("cuda")
    ims = []
    for a, b in pairs:
        ims += [to_device_image(a, dev), to_device_image(b, dev)]
    prob = RefineProblem(last, ImageSet(ims), len(pairs))
    dx0 = torch.tensor(np.stack([r.passes[-1].deltaxs_final for r in results]), device=dev)
    dy0 = torch.tensor(np.stack([r.passes[-1].deltays_final for r in results]), device=dev)
    dx, dy, l0, l1, acc = refine(prob, dx0, dy0, steps=20)
    print("loss before", l0.cpu().numpy(), "after", l1.cpu().numpy(), "accepted",
          acc.cpu().numpy(), flush=True)
    assert bool((l1 <= l0).all())
    assert bool((l1 < l0).any())
    # the returned field has the returned loss
    chk, _ = prob.forward(dx.contiguous(), dy.contiguous())
    assert torch.allclose(chk, l1, rtol=1e-12, atol=0)


def test_workpiv_refine_steps_parameter():
    """params.multipass.refine_steps runs the refinement after the last pass;
    the refined field is stored beside the multipass one, never worse."""
    from fluidimage_amd.synthetic import config_pair
    from fluidimage_amd.works.piv.multipass import WorkPIV

    im0, im1 = config_pair(2, shape=(256, 256))
    p = WorkPIV.create_default_params()
    p.piv0.shape_crop_im0 = 32
    p.multipass.number = 2
    p.multipass.use_tps = True
    p.multipass.refine_steps = 10
    res = WorkPIV(p).calcul_from_arrays(im0, im1)
    last = res.passes[-1]
    l0, l1 = last.refine_loss
    assert l1 <= l0
    assert last.deltaxs_refined.shape == last.deltaxs_final.shape
    if last.refine_accepted_steps == 0:
        assert np.array_equal(last.deltaxs_refined, last.deltaxs_final)
