"""The TPS smoothing of a set of centers as a linear map of the values,
with its transpose: the differentiable node field of the refinement.

ThinPlateSplineSubdom without outlier iterations
(src/fluidimage/calcul/interpolate/thin_plate_spline_subdom.py:128-159 with
threshold None) maps the values v at the centers to the smoothed values

    T v = (sum over subdomains d of coef_d (I - s Sw_d) v_d) / norm,

where Sw_d = (M_d^-1)[:n, :n] is the value block of the inverse of the
subdomain's saddle-point matrix (thin_plate_spline.py:168-216: [w; a] =
M^-1 [v; 0], smoothed = v - s w) and coef_d / norm the blend weights. The
structure (subdomains, members, factorisations) depends on the centers
only, so one structure serves any number of fields; Sw_d is built once on
the device (``piv_tps_influence``) and T, T^T cost one matrix-vector
product per subdomain and field (``piv_tps_node_field``).
"""

import ctypes

from .. import _native
from .tps import SingularSystemError, lattice_step
from .tps_planned import make_lattice_table, plan_capacity, tps_planned


def _p(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


class TpsNodeField:
    """T and T^T on the centers (xs, ys) (float64 numpy arrays, the same for
    every field) for subdomain size ``subdom_size`` and ``smoothing_coef``,
    as ThinPlateSplineSubdom builds them (percent_buffer_area 20)."""

    def __init__(self, xs, ys, subdom_size, smoothing_coef, device, percent_buffer_area=20):
        import numpy as np
        import torch

        xs = np.ascontiguousarray(xs, dtype=np.float64)
        ys = np.ascontiguousarray(ys, dtype=np.float64)
        self.nc = int(xs.size)
        self.device = device
        step = lattice_step(xs, ys)
        cap = plan_capacity(xs, ys, 1, subdom_size, smoothing_coef, percent_buffer_area,
                            step=step, generated=True)
        table = make_lattice_table(cap, device)
        t = lambda a: torch.from_numpy(a).to(device)  # noqa: E731
        gx, gy = t(xs), t(ys)
        zeros = torch.zeros(self.nc, dtype=torch.float32, device=device)
        tps = tps_planned(cap, zeros, zeros, gx, gy, gx, gy, subdom_size, smoothing_coef, None,
                          percent_buffer_area, lattice_table=table)
        err = tps.field_errors()[0]
        if err is not None:
            raise SingularSystemError(f"TPS node field: {err}")
        self.cap = cap
        self.tps = tps
        self.nd = int(cap.nd)
        self.n_max = int(cap.n_max)
        self.nslots = int(cap.slots)
        self.nused = int(tps.dom_off[self.nd].item())
        self.stride = self.n_max * self.n_max
        self.infl = torch.empty(self.nd * self.stride, dtype=torch.float64, device=device)
        lib = _native.load()
        _native.check(lib.piv_tps_influence(self.nd, self.n_max, _p(tps.dom_off),
                                            _p(tps.mat_off), _p(tps.mats), _p(tps.refl),
                                            _p(tps.vwork), _p(tps.status), _p(self.infl),
                                            self.stride, None), "piv_tps_influence")

    def _apply(self, values, transpose, stream=None):
        import torch

        v = values.to(torch.float64).contiguous()
        if v.dim() != 2 or v.shape[0] != 2 or v.shape[1] % self.nc:
            raise ValueError("values: (2, nfields * nc) float64")
        nfields = v.shape[1] // self.nc
        out = torch.empty_like(v)
        work = torch.empty(4 * nfields * self.nslots, dtype=torch.float64, device=v.device)
        t = self.tps
        lib = _native.load()
        _native.check(lib.piv_tps_node_field(
            int(transpose), nfields, self.nc, self.nd, self.n_max, _p(t.dom_off), _p(t.slot_pt),
            self.nused, _p(t.status), _p(t.smoothing), _p(self.infl), self.stride, _p(t.c_slot),
            _p(t.c_coef), _p(t.c_norm), self.nslots, _p(v), _p(out), _p(work),
            _native.stream_ptr(stream)), "piv_tps_node_field")
        return out

    def forward(self, values, stream=None):
        """T values: the smoothed values at the centers, (2, nfields nc)."""
        return self._apply(values, 0, stream)

    def adjoint(self, grads, stream=None):
        """T^T grads: the gradient of the values from that of the smoothed
        values, (2, nfields nc)."""
        return self._apply(grads, 1, stream)
