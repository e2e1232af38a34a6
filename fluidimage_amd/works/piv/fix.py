"""WorkFIX on the GPU (src/fluidimage/works/piv/fix.py:21-243).

The NaN marking (errors, ``correl_min``, ``displacement_max``) and the
8-neighbour mean test run in libpiv_hip (``piv_fix``); the secondary-peak
replacement (fix.py:135-243, only with ``nb_peaks_to_search > 1``) is
per-vector host logic over the secondary peaks and their sub-pixel
positions, both computed by the correlation kernel.
"""

import ctypes
from math import sqrt

import numpy as np

from ... import _native
from ...data_objects import error_strings


def run_fix_device(out, ny, nx, params_fix, stream=None, npairs=1):
    """Apply the NaN marking and neighbour test in place on device results
    ``out`` (dx, dy, correl_max, err; npairs pair-major grids of ny x nx).
    Adds flags, diff2, mean_x, mean_y."""
    import torch

    dev = out["dx"].device
    n = ny * nx * npairs
    flags = torch.empty(n, dtype=torch.int32, device=dev)
    diff2 = torch.empty(n, dtype=torch.float32, device=dev)
    mean_x = torch.empty(n, dtype=torch.float32, device=dev)
    mean_y = torch.empty(n, dtype=torch.float32, device=dev)
    dmax = params_fix.displacement_max
    thr = params_fix.threshold_diff_neighbour
    lib = _native.load()
    _native.check(
        lib.piv_fix(
            int(npairs), ctypes.c_void_p(out["dx"].data_ptr()),
            ctypes.c_void_p(out["dy"].data_ptr()),
            ctypes.c_void_p(out["correl_max"].data_ptr()), ctypes.c_void_p(out["err"].data_ptr()),
            ny, nx, float(params_fix.correl_min), float(dmax) if dmax else 0.0,
            float(thr) if thr is not None else -1.0, ctypes.c_void_p(flags.data_ptr()),
            ctypes.c_void_p(diff2.data_ptr()), ctypes.c_void_p(mean_x.data_ptr()),
            ctypes.c_void_p(mean_y.data_ptr()), _native.stream_ptr(stream)),
        "piv_fix",
    )
    out.update(flags=flags, diff2=diff2, mean_x=mean_x, mean_y=mean_y)
    return out


class WorkFIX:
    """Fix a displacement vector field (fix.py:21-243)."""

    @classmethod
    def _complete_params_with_default(cls, params, tag="fix"):
        params._set_child(
            tag,
            attribs={"correl_min": 0.2, "threshold_diff_neighbour": 10, "displacement_max": None},
        )

    def __init__(self, params_fix, piv_work):
        self.params_fix = params_fix
        self.piv_work = piv_work

    def calcul(self, piv_results):
        """Host-level WorkFIX on a HeavyPIVResults (numpy attributes); the
        arithmetic runs on the GPU."""
        import torch

        dev = torch.device("cuda")
        ny, nx = len(self.piv_work.iyvecs), len(self.piv_work.ixvecs)
        codes = np.zeros(len(piv_results.deltaxs), np.int32)
        for i, expl in piv_results.errors.items():
            codes[i] = 1  # any error: only "non-zero" matters to piv_fix
        out = {
            "dx": torch.from_numpy(np.asarray(piv_results.deltaxs, np.float32).copy()).to(dev),
            "dy": torch.from_numpy(np.asarray(piv_results.deltays, np.float32).copy()).to(dev),
            "correl_max": torch.from_numpy(
                np.asarray(piv_results.correls_max, np.float32).copy()).to(dev),
            "err": torch.from_numpy(codes).to(dev),
        }
        run_fix_device(out, ny, nx, self.params_fix)
        torch.cuda.current_stream().synchronize()
        h = {k: v.cpu().numpy() for k, v in out.items()}
        self.finish_host(piv_results, h)
        return piv_results

    def finish_host(self, piv_results, h):
        """Host bookkeeping after piv_fix: NaN'd fields, errors dict,
        deltaxs_wrong, then the secondary-peak replacement."""
        deltaxs_wrong = {}
        deltays_wrong = {}
        flags = h["flags"]
        old_x = np.asarray(piv_results.deltaxs)
        old_y = np.asarray(piv_results.deltays)
        for i in np.flatnonzero(flags):
            deltaxs_wrong[int(i)] = old_x[i]
            deltays_wrong[int(i)] = old_y[i]
        # errors keep the pass's explanations; WorkFIX appends its own
        errors = piv_results.errors
        labels = ((2, "correl < correl_min"), (4, "delta2 < displacement_max2"),
                  (8, "diff neighbour too large"))
        for bit, text in labels:
            for i in np.flatnonzero(flags & bit):
                i = int(i)
                if i in errors:
                    errors[i] += " + " + text
                else:
                    errors[i] = text
        for i in np.flatnonzero(flags & 8):
            errors[int(i)] += f" (diff = {sqrt(float(h['diff2'][i])):.2f})"
        piv_results.deltaxs = h["dx"]
        piv_results.deltays = h["dy"]
        if self.params_fix.threshold_diff_neighbour is not None:
            piv_results.deltaxs_wrong = deltaxs_wrong
            piv_results.deltays_wrong = deltays_wrong
        if self.piv_work.params.piv0.nb_peaks_to_search >= 2:
            self._secondary(piv_results, h, deltaxs_wrong, deltays_wrong)

    def _secondary(self, piv_results, h, deltaxs_wrong, deltays_wrong):
        """Secondary-peak replacement (fix.py:135-243) with the sub-pixel
        positions the kernel computed for every secondary peak."""
        ratio_correl_peaks = 0.6
        correl_min = self.params_fix.correl_min
        thr = self.params_fix.threshold_diff_neighbour
        dmax = self.params_fix.displacement_max
        mean_x, mean_y, diff2 = h["mean_x"], h["mean_y"], h["diff2"]
        sec_sub = getattr(piv_results, "_sec_sub", None)
        dxin = getattr(piv_results, "deltaxs_input", None)
        dyin = getattr(piv_results, "deltays_input", None)
        nb_replaced = 0
        for ivec, other_peaks in enumerate(piv_results.secondary_peaks):
            if other_peaks is None or len(other_peaks) == 0:
                continue
            correl0 = piv_results.correls_max[ivec]
            dx_input = dxin[ivec] if dxin is not None else 0
            dy_input = dyin[ivec] if dyin is not None else 0
            good = []
            for k, (dx, dy, corr) in enumerate(other_peaks):
                if corr / correl0 > ratio_correl_peaks and corr > correl_min:
                    good.append((dx + dx_input, dy + dy_input, corr, k))
            if not good:
                continue
            diff_neighbours = np.empty(len(good) + 1)
            diff_neighbours[0] = sqrt(float(diff2[ivec]))
            for i, (dx, dy, corr, k) in enumerate(good):
                diff_neighbours[i + 1] = np.sqrt((mean_x[ivec] - dx) ** 2
                                                 + (mean_y[ivec] - dy) ** 2)
            argmin = diff_neighbours.argmin()
            if argmin == 0:
                continue
            dx, dy, corr, k = good[argmin - 1]
            if thr is not None and diff_neighbours[argmin] > thr:
                continue
            if dmax and dx**2 + dy**2 > dmax:
                continue
            if sec_sub is None:
                continue
            sx, sy = sec_sub[ivec, k]
            if np.isnan(sx):
                continue  # PIVError in apply_subpix
            dx = sx + dx_input
            dy = sy + dy_input
            nb_replaced += 1
            replaced = piv_results.__dict__.setdefault("replaced_vectors", {})
            old_dx = deltaxs_wrong.get(ivec, piv_results.deltaxs[ivec])
            old_dy = deltays_wrong.get(ivec, piv_results.deltays[ivec])
            replaced[ivec] = (old_dx, old_dy, piv_results.correls_max[ivec])
            piv_results.deltaxs[ivec] = dx
            piv_results.deltays[ivec] = dy
            piv_results.correls_max[ivec] = corr
            if ivec in piv_results.errors:
                deltaxs_wrong.pop(ivec, None)
                deltays_wrong.pop(ivec, None)
                del piv_results.errors[ivec]
        if nb_replaced > 0:
            print(f"Secondary peaks: {nb_replaced} bad peak(s) replaced")
