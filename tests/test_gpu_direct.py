"""Direct (spatial) correlation on the GPU (CorrelPythranHIP /
CorrelPyCudaHIP) against the oracle's restatement of correl_numpy /
correl_pycuda and CorrelBase's search (src/fluidimage/calcul/correl.py:
226-451, correl_pycuda.py:26-240).

Bars: maps within 1e-6 relative of the FP64 oracle (FP64 sums on both
sides, the GPU map rounded once to float32), norms within 1e-12 relative,
integer peaks identical except near-ties (listed), sub-pixel displacements
within 1e-4 px, the reference tests' known answers."""

import numpy as np
import pytest

from tests.helpers import near_tie_windows, shifted_pair

pytestmark = pytest.mark.gpu


def _classes(kind):
    from fluidimage_amd.calcul.correl import CorrelPyCudaHIP, CorrelPythranHIP

    return CorrelPythranHIP if kind == "pythran" else CorrelPyCudaHIP


@pytest.mark.parametrize("kind", ["pythran", "pycuda"])
@pytest.mark.parametrize("s0, s1, dmax", [((32, 32), (32, 32), None), ((24, 24), (24, 24), 6),
                                          ((64, 32), (64, 32), None),
                                          ((96, 48), (64, 32), None)])
def test_direct_stack_vs_oracle(kind, s0, s1, dmax):
    import torch

    from oracle.correl import OracleCorrelDirect, PIVError

    print("case: direct", kind, s0, s1, dmax, flush=True)
    rng = np.random.default_rng(11 + sum(s0))
    nwin = 24
    w0 = []
    w1 = []
    for k in range(nwin):
        a, b = shifted_pair(rng, s0, (rng.uniform(-4, 4), rng.uniform(-4, 4)))
        y, x = (s0[0] - s1[0]) // 2, (s0[1] - s1[1]) // 2
        w0.append(a)
        w1.append(np.ascontiguousarray(b[y:y + s1[0], x:x + s1[1]]))
    cls = _classes(kind)
    op = cls(s0, s1, method_subpix="2d_gaussian2", nsubpix=None, displacement_max=dmax)
    ref = OracleCorrelDirect(s0, s1, method_subpix="2d_gaussian2", nsubpix=None,
                             displacement_max=dmax, kind=kind)
    assert op.displacement_max == ref.displacement_max
    dev = torch.device("cuda")
    maps, norms = op.correlate_stack(torch.from_numpy(np.stack(w0)).to(dev),
                                     torch.from_numpy(np.stack(w1)).to(dev))
    gm = maps.cpu().numpy().copy()
    gn = norms.cpu().numpy()
    out = op.peaks_stack(maps, norms, apply_subpix=True)
    torch.cuda.synchronize()
    host = {k: v.cpu().numpy() for k, v in out.items()}
    exp_maps = []
    for k in range(nwin):
        em, en = ref(w0[k], w1[k])
        exp_maps.append(em)
        assert np.allclose(gm[k], em, rtol=1e-6, atol=1e-6 * np.abs(em).max())
        assert abs(gn[k] - en) <= 1e-12 * en
    ties = set(near_tie_windows(np.stack([m.astype(np.float64) for m in exp_maps])))
    for k in range(nwin):
        if k in ties:
            continue
        em = exp_maps[k].copy()
        try:
            dx, dy, cmax, _ = ref.compute_displacements_from_correl(em, norm=float(gn[k]))
            dx, dy = ref.apply_subpix(dx, dy, em)
            err = 0
        except PIVError:
            err = 1
        assert (host["err"][k] != 0) == bool(err), k
        if not err:
            assert abs(host["dx"][k] - dx) <= 1e-4 and abs(host["dy"][k] - dy) <= 1e-4, k
            assert np.isclose(host["correl_max"][k], cmax, rtol=1e-5), k


@pytest.mark.parametrize("kind", ["pythran", "pycuda"])
def test_direct_per_window_known_answers(kind):
    """test_correl.py:45-124: identical windows give (0, 0) (and correl_max
    1 with the pythran normalisation); the shift (3.3, 5.8) is found within
    0.8 px; the second-chance square lands in the caller's map."""
    print("case: direct per-window", kind, flush=True)
    rng = np.random.default_rng(21)
    im0, im1 = shifted_pair(rng, (32, 32), (3.3, 5.8))
    op = _classes(kind)((32, 32), (32, 32))
    c, norm = op(im0, im0)
    assert c.shape == op.map_shape and op.get_indices_no_displacement() == (op.iy0, op.ix0)
    dx, dy, cmax, _ = op.compute_displacements_from_correl(c, norm=norm)
    assert (dx, dy) == (0, 0)
    if kind == "pythran":
        assert np.isclose(cmax, 1.0, rtol=1e-5)
    c, norm = op(im0, im1)
    dx, dy, cmax, _ = op.compute_displacements_from_correl(c, norm=norm)
    dx, dy = op.apply_subpix(dx, dy, c)
    assert np.allclose([dx, dy], [3.3, 5.8], atol=0.8)


def test_direct_rejects_bad_sizes():
    from fluidimage_amd.calcul.correl import CorrelPythranHIP

    with pytest.raises(ValueError):
        CorrelPythranHIP((32, 32), (32, 32), displacement_max=0)
    op = CorrelPythranHIP((16, 16), (16, 16), displacement_max=20)
    with pytest.raises(RuntimeError, match="too large"):
        op(np.ones((16, 16), np.float32), np.ones((16, 16), np.float32))
