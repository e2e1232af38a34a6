"""The benchmark's workload at full size: 2048 x 2048 uint16 pairs, passes of
64/32/16 px windows with FP64 TPS between them and at the end, bicubic image
deformation in the later passes. Properties that do not depend on the size:

* a batch of two pairs gives, pair by pair, the bits of each pair run alone;
  the launches cross the small/large width thresholds of the TPS kernels
  (one pair's first TPS call has about 20 subdomains, the batch's last one
  about 600);
* the last pass follows the synthetic truth field the images were rendered
  with (the displacement of the particles seen at the window centre)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _params():
    from fluidimage_amd.works.piv.multipass import WorkPIV

    p = WorkPIV.create_default_params()
    p.piv0.shape_crop_im0 = 64
    p.multipass.number = 3
    p.multipass.coeff_zoom = 2
    p.multipass.use_tps = True
    p.multipass.subdom_size = 200
    p.multipass.smoothing_coef = 2.0
    p.multipass.threshold_tps = 1.5
    p.multipass.image_deformation = "bicubic"
    return p



def test_fullsize_batch_equals_single_and_follows_truth():
    from fluidimage_amd.synthetic import config_pair, config_truth
    from fluidimage_amd.works.piv.multipass import WorkPIV

    print("case: full-size multipass, 2 pairs of 2048x2048 uint16", flush=True)
    pairs = [config_pair(3, pair_idx=k) for k in range(2)]
    p = _params()
    batch = WorkPIV(p).calcul_batch(pairs)
    for k, ((im0, im1), b) in enumerate(zip(pairs, batch)):
        single = WorkPIV(p).calcul_from_arrays(im0, im1)
        assert len(single.passes) == len(b.passes) == 3
        for ps, pb in zip(single.passes, b.passes):
            assert np.array_equal(ps.deltaxs, pb.deltaxs, equal_nan=True)
            assert np.array_equal(ps.deltays, pb.deltays, equal_nan=True)
            assert ps.errors == pb.errors
        assert np.array_equal(single.passes[-1].deltaxs_final, b.passes[-1].deltaxs_final)
        assert np.array_equal(single.passes[-1].deltays_final, b.passes[-1].deltays_final)

        last = b.passes[-1]
        assert last.deltaxs.size == 255 * 255
        field, boxes = config_truth(3, k)
        ux, uy = field(last.xs - 0.5 * last.deltaxs, last.ys - 0.5 * last.deltays)
        # vectors whose 16 px windows (plus the largest displacement) stay
        # clear of the noise boxes of im1
        clear = np.ones(last.deltaxs.shape, bool)
        for (y0, x0, h, w) in boxes:
            clear &= ~((last.ys > y0 - 16) & (last.ys < y0 + h + 16)
                       & (last.xs > x0 - 16) & (last.xs < x0 + w + 16))
        ok = ~np.isnan(last.deltaxs)
        sel = ok & clear
        err = np.hypot(last.deltaxs[sel] - ux[sel], last.deltays[sel] - uy[sel])
        print(f"pair {k}: {ok.mean() * 100:.2f} % valid, {clear.mean() * 100:.1f} % clear of "
              f"noise boxes; error vs truth there: median {np.median(err):.4f} px, 95th "
              f"percentile {np.percentile(err, 95):.4f} px, max {err.max():.3f} px", flush=True)
        assert ok.mean() > 0.95
        # 16 px windows on Fourier modes of 128-512 px wavelength and 6 px
        # amplitude: about 0.2 px median, 0.4 px 95th percentile (the
        # method's own error: the oracle's field agrees with this one)
        assert np.median(err) < 0.3
        assert np.percentile(err, 95) < 0.6
