"""Differentiable refinement with TPS node values (the 1024 x 1024 float32
workload of the refinement): the node displacements are the
ThinPlateSplineSubdom smoothing T of the node values, then bilinear warp and
windowed NCC as in test_gpu_refine.py.

Oracle: oracle.tps.smoothing_operator (T as a dense FP64 matrix from the
restated reference's saddle-point systems, pinned in test_oracle_pins.py)
composed with oracle.refine.refine_loss, differentiated by torch's autograd
on the CPU in FP64.

Bars: T and T^T on the device within 1e-9 (relative to the largest value)
of the dense operator; loss within 1e-9 relative and the gradient on 256
random nodes within 1e-7 (relative to its largest component) of the
oracle's; central differences (step 1e-3 px) of the device loss within
1e-4; 20 refinement steps on 64 pairs never raise a pair's loss above that
of its multipass field and lower most of them."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPE = (1024, 1024)
SUBDOM, SMOOTH = 200, 2.0
_CACHE = {}


def _work():
    from fluidimage_amd.works.piv.multipass import WorkPIV

    p = WorkPIV.create_default_params()
    p.piv0.shape_crop_im0 = 64
    p.multipass.number = 2
    p.multipass.use_tps = True
    p.multipass.subdom_size = SUBDOM
    p.multipass.smoothing_coef = SMOOTH
    work = WorkPIV(p)
    work._prepare_with_image(imshape=SHAPE)
    return p, work


def _operator(work):
    if "T" not in _CACHE:
        from oracle.tps import smoothing_operator

        last = work.works_piv[-1]
        centers = np.vstack([last.iyvecs_grid.astype(float), last.ixvecs_grid.astype(float)])
        _CACHE["T"] = smoothing_operator(centers, SUBDOM, SMOOTH)
    return _CACHE["T"]


def _problem(pairs, work):
    import torch

    from fluidimage_amd.device import ImageSet, to_device_image
    from fluidimage_amd.works.piv.refine import RefineProblem, node_field

    dev = torch.device("cuda")
    ims = []
    for a, b in pairs:
        ims += [to_device_image(a, dev), to_device_image(b, dev)]
    prob = RefineProblem(work.works_piv[-1], ImageSet(ims), len(pairs))
    return prob, node_field(prob, SUBDOM, SMOOTH)


def test_node_field_and_transpose_match_the_dense_operator():
    import torch

    from fluidimage_amd.works.piv.refine import RefineProblem, node_field  # noqa: F401

    print("case: TPS node field vs dense operator (63 x 63 nodes)", flush=True)
    p, work = _work()
    T = _operator(work)
    from fluidimage_amd.synthetic import config_pair

    prob, field = _problem([config_pair(5, pair_idx=0)], work)
    nc = T.shape[0]
    rng = np.random.default_rng(5)
    v = rng.normal(0, 3, (2, 3 * nc))
    dev = torch.device("cuda")
    out = field.forward(torch.tensor(v, device=dev)).cpu().numpy()
    adj = field.adjoint(torch.tensor(v, device=dev)).cpu().numpy()
    for f in range(3):
        for c in range(2):
            x = v[c, f * nc:(f + 1) * nc]
            ref = T @ x
            assert np.abs(out[c, f * nc:(f + 1) * nc] - ref).max() <= 1e-9 * np.abs(ref).max()
            ref_t = T.T @ x
            assert np.abs(adj[c, f * nc:(f + 1) * nc] - ref_t).max() <= 1e-9 * np.abs(ref_t).max()


def test_tps_node_gradient_matches_oracle_on_256_nodes():
    import torch

    from fluidimage_amd.synthetic import config_pair
    from fluidimage_amd.works.piv.refine import refine_loss
    from oracle.refine import refine_loss as oracle_loss

    print("case: TPS node gradient vs FP64 CPU oracle, 1024^2 float32", flush=True)
    p, work = _work()
    T = _operator(work)
    last = work.works_piv[-1]
    pair = config_pair(5, pair_idx=1)
    prob, field = _problem([pair], work)
    nvec = last.nb_vectors
    rng = np.random.default_rng(9)
    # a smooth field of a few px plus node noise
    xs, ys = last.ixvecs_grid.astype(float), last.iyvecs_grid.astype(float)
    v = np.stack([2.0 * np.sin(xs / 170.0) + rng.normal(0, 0.3, nvec),
                  1.5 * np.cos(ys / 230.0) + rng.normal(0, 0.3, nvec)])
    dev = torch.device("cuda")
    x = torch.tensor(v[0][None], device=dev, requires_grad=True)
    y = torch.tensor(v[1][None], device=dev, requires_grad=True)
    loss = refine_loss(x, y, prob, field)
    gx, gy = torch.autograd.grad(loss.sum(), (x, y))
    Tt = torch.from_numpy(T)
    ox = torch.tensor(v[0], requires_grad=True)
    oy = torch.tensor(v[1], requires_grad=True)
    ol = oracle_loss(pair[0], pair[1], last.iyvecs, last.ixvecs, last.shape_crop_im0,
                     Tt @ ox, Tt @ oy)
    ogx, ogy = torch.autograd.grad(ol, (ox, oy))
    print(f"loss device {float(loss[0]):.12g} oracle {float(ol):.12g}", flush=True)
    assert abs(float(loss[0].detach()) - float(ol.detach())) <= 1e-9 * abs(float(ol.detach()))
    g = np.concatenate([gx[0].cpu().numpy(), gy[0].cpu().numpy()])
    og = np.concatenate([ogx.numpy(), ogy.numpy()])
    idx = rng.choice(g.size, 256, replace=False)
    scale = np.abs(og).max()
    err = np.abs(g[idx] - og[idx]).max() / scale
    print(f"gradient: max error on 256 nodes {err:.3g} of the largest component", flush=True)
    assert err <= 1e-7


def test_tps_node_gradient_matches_central_differences():
    import torch

    from fluidimage_amd.synthetic import config_pair
    from fluidimage_amd.works.piv.refine import field_of, refine_loss

    print("case: TPS node gradient vs central differences", flush=True)
    p, work = _work()
    last = work.works_piv[-1]
    prob, field = _problem([config_pair(5, pair_idx=2)], work)
    nvec = last.nb_vectors
    rng = np.random.default_rng(13)
    dev = torch.device("cuda")
    d = torch.tensor(rng.normal(0, 1.0, (2, 1, nvec)), device=dev)
    x = d[0].clone().requires_grad_(True)
    y = d[1].clone().requires_grad_(True)
    gx, gy = torch.autograd.grad(refine_loss(x, y, prob, field).sum(), (x, y))
    h = 1e-3
    for k in rng.choice(nvec, 12, replace=False):
        for comp, g in ((0, gx), (1, gy)):
            e = torch.zeros_like(d)
            e[comp, 0, k] = h
            lp, _ = prob.forward(*field_of(field, d[0] + e[0], d[1] + e[1]))
            lm, _ = prob.forward(*field_of(field, d[0] - e[0], d[1] - e[1]))
            fd = float(lp[0] - lm[0]) / (2 * h)
            assert abs(fd - float(g[0, k])) <= 1e-4 * max(1.0, abs(fd)), (k, comp, fd)


def test_refinement_of_64_pairs_keeps_only_improvements():
    import torch

    from fluidimage_amd.synthetic import config_pair
    from fluidimage_amd.works.piv.refine import refine

    print("case: 20 refinement steps on 64 pairs of 1024^2 float32", flush=True)
    p, work = _work()
    pairs = [config_pair(5, pair_idx=100 + k) for k in range(64)]
    results = work.calcul_batch(pairs)
    prob, field = _problem(pairs, work)
    dev = torch.device("cuda")
    dx0 = torch.tensor(np.stack([r.passes[-1].deltaxs_final for r in results]), device=dev)
    dy0 = torch.tensor(np.stack([r.passes[-1].deltays_final for r in results]), device=dev)
    dx, dy, l0, l1, acc = refine(prob, dx0, dy0, steps=20, field=field)
    torch.cuda.synchronize()
    l0, l1, acc = l0.cpu().numpy(), l1.cpu().numpy(), acc.cpu().numpy()
    print(f"loss of the multipass fields: mean {l0.mean():.4f}; kept: mean {l1.mean():.4f}; "
          f"pairs improved {(l1 < l0).sum()} of 64; kept steps {acc.sum()}", flush=True)
    assert (l1 <= l0).all()
    assert (l1 < l0).sum() >= 32
    # a pair without a kept step returns its multipass field
    same = acc == 0
    assert torch.equal(dx[torch.from_numpy(same).to(dev)], dx0[torch.from_numpy(same).to(dev)])
    chk, _ = prob.forward(dx.contiguous(), dy.contiguous())
    assert np.allclose(chk.cpu().numpy(), l1, rtol=1e-12, atol=0)
