"""Device-sized TPS batches: the subdomain smoothing and evaluation of
``tps.tps_device_batch`` without any device-to-host read.

The host sizes a batch once, from the geometry it knows in advance (the
pass grid where the centers lie, the new positions, the number of fields):
``plan_capacity`` computes the reference's subdomain grid of a field whose
every vector is valid (thin_plate_spline_subdom.py:53-126) and takes
margins over its subdomain count, largest subdomain, member slots and
matrix doubles. Everything data-dependent is then computed on the GPU
(``piv_tps_grid``: the subdomain grids; ``piv_tps_offsets``: member counts,
offsets, largest-first order), so a whole multipass can be enqueued at once
or captured in one graph. A field the capacities cannot hold, or that the
reference would not solve (no vector, an empty subdomain, a singular
system), is flagged on the device; ``PlannedTps.field_errors`` reads the
flags once, when the caller is ready to wait, and such a field is solved
again by the host-sized path (with the reference's fallbacks). Fields that
pass get the same bits as from the host-sized path: the kernels are the
same, only the layout of the workspaces differs.
"""

import ctypes
import math
from dataclasses import dataclass

import numpy as np

from .. import _native
from .tps import (MAX_CENTERS_PER_SUBDOM, MAX_LATTICE_A, MAX_SLOTS, SingularSystemError,
                  matrix_sizes, subdomain_grid)

FIELD_OK = 0
FIELD_NO_VECTOR = 1
FIELD_DEGENERATE = 2
FIELD_EMPTY_SUBDOMAIN = 3
FIELD_CAPACITY = 4
FIELD_TOO_MANY = 5
FIELD_OFF_LATTICE = 6
_FIELD_MESSAGES = {
    FIELD_NO_VECTOR: "Only nan.",
    FIELD_DEGENERATE: "degenerate set of centers",
    FIELD_EMPTY_SUBDOMAIN: "empty subdomain",
    FIELD_CAPACITY: "beyond the capacities of the device-sized TPS batch",
    FIELD_TOO_MANY: f"more than {MAX_CENTERS_PER_SUBDOM} centers in a TPS subdomain",
    FIELD_OFF_LATTICE: "center off the lattice of the generated TPS matrices",
}

# margins over the all-valid geometry: fewer valid vectors or narrower
# bounds change the subdomain counts by one or so, hence the subdomain
# sizes by about 1 / 17 per direction at the benchmark sizes
MARGIN_SUBDOMAIN = 1.3
MARGIN_SLOTS = 1.3
MARGIN_MATS = 1.75
MARGIN_EXTENT = 1.35
# subdomains sharing a factorisation (centers on the pass grid): patterns
# per field, and missing nodes of a box solved by the capacitance correction
# (beyond: the subdomain is factorised alone)
SHARED_PATTERNS = 32
SHARED_MAX_MISSING = 32
# sharing pays where a field has many subdomains per pattern and the batch
# many more subdomains than the GPU runs at once (the patterns' explicit
# inverses come first, a short chain of launches): measured in the 32-pair
# multipass of 2048 x 2048 px images (subdom_size 200), 32 fields of 110
# subdomains take 4.45 ms shared against 5.70 ms alone, 32 fields of 42
# take 2.58 ms against 1.54 ms
SHARED_MIN_SUBDOMAINS = 100  # per field
SHARED_MIN_BATCH = 2048  # subdomains of the batch


@dataclass(frozen=True)
class TpsCapacity:
    """Capacities of a device-sized TPS batch of ``nfields`` fields of
    ``n`` vectors."""

    nfields: int
    n: int
    dom_cap: int      # subdomain slots per field
    n_max: int        # centers per subdomain (even, at most 2048)
    slots: int        # member slots of the batch
    mats: int         # matrix doubles of the batch
    step: float       # lattice step of the kernel-function table (0: none)
    a_max: int        # extent of the table (lattice steps)
    generated: bool   # matrices generated from the table (every center on it)

    @property
    def nd(self):
        return self.nfields * self.dom_cap


def plan_capacity(cx, cy, nfields, subdom_size, smoothing_coef, percent_buffer_area=20,
                  step=None, generated=False):
    """Capacities for fields whose centers lie at (about) the host positions
    (cx, cy) (all of them valid): subdomain grid and member counts of that
    field, with margins."""
    cx = np.asarray(cx, dtype=np.float64)
    cy = np.asarray(cy, dtype=np.float64)
    n = cx.size
    nsx, nsy, _, lim = subdomain_grid(n, cx.min(), cx.max(), cy.min(), cy.max(), subdom_size,
                                      smoothing_coef, percent_buffer_area)
    lx0, lx1 = lim[:nsx], lim[nsx:2 * nsx]
    ly0, ly1 = lim[2 * nsx:2 * nsx + nsy], lim[2 * nsx + nsy:]
    in_x = ((cx[None, :] >= lx0[:, None]) & (cx[None, :] < lx1[:, None])).astype(np.int32)
    in_y = ((cy[None, :] >= ly0[:, None]) & (cy[None, :] < ly1[:, None])).astype(np.int32)
    counts = (in_y @ in_x.T).ravel()
    n_max = int(counts.max())
    n_max_cap = min(MAX_CENTERS_PER_SUBDOM, int(math.ceil(n_max * MARGIN_SUBDOMAIN)))
    n_max_cap = max(4, n_max_cap + (n_max_cap & 1))
    slots = int(math.ceil(int(counts.sum()) * MARGIN_SLOTS)) * nfields
    mats = int(math.ceil(int(matrix_sizes(counts).sum()) * MARGIN_MATS)) * nfields
    a_max = 0
    if step:
        extent = max(float((lx1 - lx0).max()), float((ly1 - ly0).max()))
        a_max = min(MAX_LATTICE_A, int(math.ceil(extent * MARGIN_EXTENT / step)) + 1)
    return TpsCapacity(nfields=nfields, n=n, dom_cap=(nsx + 2) * (nsy + 2), n_max=n_max_cap,
                       slots=slots, mats=mats, step=float(step or 0.0), a_max=a_max,
                       generated=bool(generated and step))


def _p(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


_SIDE_STREAMS = {}


def _side_stream(dev):
    """One side stream per device (default priority) for the branch of a
    TPS call that runs beside the main stream's work."""
    import torch

    key = torch.device(dev).index
    s = _SIDE_STREAMS.get(key)
    if s is None:
        s = _SIDE_STREAMS[key] = torch.cuda.Stream(device=dev)
    return s


class PlannedTps:
    """Device results of one device-sized TPS batch (see the module
    docstring); nothing is read back until ``field_errors``."""

    def __init__(self, cap, m, tensors):
        import torch

        self.cap = cap
        self.m = m
        self.__dict__.update(tensors)
        # the checks copied to pinned host memory behind the call's kernels,
        # with an event: field_errors then waits for this call alone, not
        # for work queued on the stream after it (not under graph capture,
        # where field_errors reads the device tensors)
        self._checks = None
        if not torch.cuda.is_current_stream_capturing():
            flat = torch.cat([self.fstat.to(torch.int64), self.status.to(torch.int64)])
            host = torch.empty(flat.shape, dtype=torch.int64, pin_memory=True)
            host.copy_(flat, non_blocking=True)
            done = torch.cuda.Event()
            done.record()
            self._checks = (host, done)

    def approx(self, k):
        """(x, y) float64 device tensors of field k at the new positions."""
        m = self.m
        return self.approx_all[0, k * m:(k + 1) * m], self.approx_all[1, k * m:(k + 1) * m]

    def field_errors(self):
        """Per field: None, or the exception the host-sized path would give
        (SingularSystemError), or a capacity failure (the host-sized path
        then solves it); NotImplementedError when a point lay in more than
        4 subdomains. One device-to-host read."""
        import torch

        if self._checks is not None:
            pinned, done = self._checks
            done.synchronize()
            host = pinned.numpy()
        else:
            host = torch.cat([self.fstat.to(torch.int64),
                              self.status.to(torch.int64)]).cpu().numpy()
        nf, nd = self.cap.nfields, self.cap.nd
        fstat, status = host[:nf], host[nf:]
        if status[nd]:
            raise NotImplementedError("a point lies in more than 4 TPS subdomains")
        out = []
        for f in range(nf):
            st = int(fstat[f])
            if st != FIELD_OK:
                out.append(SingularSystemError(_FIELD_MESSAGES.get(st, f"field status {st}")))
                continue
            bad = status[f * self.cap.dom_cap:(f + 1) * self.cap.dom_cap]
            if bad.any():
                out.append(SingularSystemError(
                    f"{int((bad != 0).sum())} subdomain system(s) not positive definite"))
            else:
                out.append(None)
        return out

    def host_field(self, k):
        """Host copies of field k's selected indices and smoothed values
        (reads the device)."""
        n = self.cap.n
        count = int(self.info[k, 0].item())
        idx = self.idx[k * n:k * n + count].cpu().numpy()
        sm = self.smoothed[:, k * n:k * n + count].cpu().numpy()
        return idx, sm


class PassGrid:
    """The grid of a pass on the device (float64 columns and rows) and
    whether its last column / row is appended at a smaller step; the
    centers of a batch solved with shared factorisations lie on it."""

    def __init__(self, ixvecs, iyvecs, device):
        import torch

        self.gx = torch.from_numpy(np.asarray(ixvecs, dtype=np.float64)).to(device)
        self.gy = torch.from_numpy(np.asarray(iyvecs, dtype=np.float64)).to(device)
        self.irr_x = int(_irregular_last(ixvecs))
        self.irr_y = int(_irregular_last(iyvecs))


def _irregular_last(v):
    v = np.asarray(v, dtype=np.float64)
    if v.size < 3:
        return v.size == 2
    return bool(v[-1] - v[-2] != v[1] - v[0])


def tps_planned(cap, dx, dy, xs, ys, new_x, new_y, subdom_size, smoothing_coef, threshold,
                percent_buffer_area=20, stream=None, lattice_table=None, grid=None):
    """TPS smoothing of ``cap.nfields`` fields of ``cap.n`` vectors (dx, dy
    float32; xs, ys float64; device tensors holding the fields one after
    the other) and evaluation at the ``m`` new positions (new_x, new_y)
    shared by the fields, all enqueued without a host read. Returns a
    ``PlannedTps``. ``lattice_table``: the device table of the kernel
    function on the lattice of ``cap.step`` (``make_lattice_table``), or
    None. ``grid``: a ``PassGrid`` on whose nodes every center lies (the
    first pass, deformed passes): subdomains with the same node pattern
    then share a factorisation (piv_tps_smooth_shared)."""
    import torch

    lib = _native.load()
    dev = dx.device
    sp = _native.stream_ptr(stream)
    f64 = dict(dtype=torch.float64, device=dev)
    i32 = dict(dtype=torch.int32, device=dev)
    nf, n = cap.nfields, cap.n
    ntot = nf * n
    if dx.numel() != ntot:
        raise ValueError("the fields of a TPS batch have one number of vectors")
    nd = cap.nd
    m = int(new_x.numel())
    thr = -1.0 if threshold is None else float(threshold)
    cx = torch.empty(ntot, **f64)
    cy = torch.empty(ntot, **f64)
    vx = torch.empty(ntot, dtype=torch.float32, device=dev)
    vy = torch.empty(ntot, dtype=torch.float32, device=dev)
    idx = torch.empty(ntot, **i32)
    info = torch.empty((nf, 7), **f64)
    step = cap.step if lattice_table is not None else 0.0
    _native.check(lib.piv_tps_select(nf, n, _p(dx), _p(dy), _p(xs), _p(ys), _p(cx), _p(cy),
                                     _p(vx), _p(vy), _p(idx), _p(info), float(step), sp),
                  "piv_tps_select")
    ftab = torch.empty((nf, 7), **i32)
    lim = torch.empty(nf * 4 * cap.dom_cap, **f64)
    dom_field = torch.empty(nd, **i32)
    smoothing = torch.empty(nd, **f64)
    fstat = torch.empty(nf, **i32)
    gen = cap.generated and lattice_table is not None
    gen_width = float(cap.a_max * cap.step) if gen else 0.0
    _native.check(lib.piv_tps_grid(nf, n, _p(info), float(subdom_size), float(smoothing_coef),
                                   percent_buffer_area / 100, cap.dom_cap, gen_width, int(gen),
                                   _p(ftab), _p(lim), _p(dom_field), _p(smoothing), _p(fstat),
                                   sp), "piv_tps_grid")
    counts = torch.empty(nd, **i32)
    _native.check(lib.piv_tps_count(nd, _p(ftab), _p(dom_field), _p(lim), _p(cx), _p(cy),
                                    _p(counts), sp), "piv_tps_count")
    dom_off = torch.empty(nd + 1, dtype=torch.int64, device=dev)
    mat_off = torch.empty(nd + 1, dtype=torch.int64, device=dev)
    order = torch.empty(nd, **i32)
    totals = torch.empty(4, dtype=torch.int64, device=dev)
    _native.check(lib.piv_tps_offsets(nf, cap.dom_cap, _p(ftab), _p(counts), _p(dom_field),
                                      int(cap.slots), int(cap.mats), int(cap.n_max), _p(fstat),
                                      _p(dom_off), _p(mat_off), _p(order), _p(totals), sp),
                  "piv_tps_offsets")
    nslots = int(cap.slots)
    slot_pt = torch.empty(nslots, **i32)
    _native.check(lib.piv_tps_fill(nd, _p(ftab), _p(dom_field), _p(lim), _p(cx), _p(cy),
                                   _p(dom_off), _p(slot_pt), sp), "piv_tps_fill")
    status = torch.zeros(nd + 1, **i32)
    c_slot = torch.empty(ntot * MAX_SLOTS, **i32)
    c_coef = torch.empty(ntot * MAX_SLOTS, **f64)
    c_norm = torch.empty(ntot, **f64)
    _native.check(lib.piv_tps_point_slots(
        nf, _p(ftab), n, _p(lim), _p(cx), _p(cy), _p(dom_off), _p(slot_pt), _p(c_slot),
        _p(c_coef), _p(c_norm), _p(status[nd:]), sp), "piv_tps_point_slots")
    n_dom = torch.empty(nf * m * MAX_SLOTS, **i32)
    n_coef = torch.empty(nf * m * MAX_SLOTS, **f64)
    n_norm = torch.empty(nf * m, **f64)
    _native.check(lib.piv_tps_point_slots(
        nf, _p(ftab), m, _p(lim), _p(new_x), _p(new_y), None, None, _p(n_dom), _p(n_coef),
        _p(n_norm), _p(status[nd:]), sp), "piv_tps_point_slots")
    mats = torch.empty(int(cap.mats), **f64)
    refl = torch.empty((nd, 16), **f64)
    vwork = torch.empty(nslots * 6, **f64)
    smooth_slots = torch.empty((2, nslots), **f64)
    wstride = nslots + 3 * nd
    weights = torch.empty((2, wstride), **f64)
    stats = torch.zeros((nd, 6), **f64)
    slot_xy = torch.empty((2, nslots), **f64)
    lat = None
    if lattice_table is not None:
        lat = _native.PivTpsLattice(lattice_table.data_ptr(), float(cap.step), int(cap.a_max), 0)
    lat_p = None if lat is None else ctypes.byref(lat)
    if grid is None:
        _native.check(
            lib.piv_tps_smooth(nd, int(cap.n_max), _p(cx), _p(cy), _p(vx), _p(vy), _p(dom_off),
                               _p(slot_pt), _p(mat_off), _p(smoothing), thr, _p(mats), _p(refl),
                               _p(vwork), _p(slot_xy), _p(smooth_slots), nslots, _p(weights),
                               wstride, _p(stats), _p(status), _p(order), lat_p, int(gen), sp),
            "piv_tps_smooth")
    else:
        wbytes = int(lib.piv_tps_shared_workspace(nf, cap.dom_cap, SHARED_PATTERNS,
                                                  int(cap.n_max), SHARED_MAX_MISSING))
        swork = torch.empty(max(wbytes, 8), dtype=torch.uint8, device=dev)

        def shared(parts, s):
            _native.check(
                lib.piv_tps_smooth_shared(
                    nf, cap.dom_cap, _p(ftab), _p(lim), _p(fstat), _p(dom_field), _p(counts),
                    _p(grid.gx), int(grid.gx.numel()), grid.irr_x, _p(grid.gy),
                    int(grid.gy.numel()), grid.irr_y, SHARED_PATTERNS, SHARED_MAX_MISSING,
                    int(cap.n_max), _p(cx), _p(cy), _p(idx), _p(vx), _p(vy), _p(dom_off),
                    _p(slot_pt), _p(mat_off), _p(smoothing), thr, _p(mats), _p(refl),
                    _p(vwork), _p(slot_xy), _p(smooth_slots), nslots, _p(weights), wstride,
                    _p(stats), _p(status), _p(order), lat_p, int(gen), parts, _p(swork), s),
                "piv_tps_smooth_shared")

        # the subdomains factorised alone on a side stream, beside the
        # patterns' factorisations (one workgroup per pattern, a small launch)
        # and the shared solves (bound by the latency of their triangular
        # solves): the two touch disjoint subdomains
        shared(_native.PIV_TPS_PART_PREP, sp)
        main = stream if stream is not None else torch.cuda.current_stream()
        side = _side_stream(dev)
        side.wait_stream(main)
        shared(_native.PIV_TPS_PART_DIRECT, _native.stream_ptr(side))
        shared(_native.PIV_TPS_PART_SHARED, sp)
        main.wait_stream(side)
    smoothed = torch.empty((2, ntot), **f64)
    _native.check(lib.piv_tps_blend(ntot, _p(c_slot), _p(c_coef), _p(c_norm), _p(smooth_slots),
                                    nslots, _p(smoothed), sp), "piv_tps_blend")
    approx = torch.zeros((2, nf * m), **f64)
    skip = torch.zeros(nf * m, dtype=torch.uint8, device=dev)
    if m == n:
        _native.check(lib.piv_tps_at_centers(nf, n, _p(info), _p(idx), _p(xs), _p(ys),
                                             _p(new_x), _p(new_y), _p(smoothed), _p(approx),
                                             _p(skip), sp), "piv_tps_at_centers")
    px = new_x.repeat(nf)
    py = new_y.repeat(nf)
    wbytes = int(lib.piv_tps_interpolate_workspace(nf * m, nd))
    iwork = torch.empty(max(wbytes, 8), dtype=torch.uint8, device=dev)
    _native.check(lib.piv_tps_interpolate(nf * m, _p(px), _p(py), _p(n_dom), _p(n_coef),
                                          _p(n_norm), _p(slot_xy), nslots, nd, int(cap.n_max),
                                          _p(dom_off), _p(weights), wstride, _p(skip),
                                          _p(approx), _p(iwork), lat_p, sp),
                  "piv_tps_interpolate")
    return PlannedTps(cap, m, {
        "info": info, "idx": idx, "fstat": fstat, "status": status, "smoothed": smoothed,
        "approx_all": approx, "stats": stats, "totals": totals, "counts": counts,
        "ftab": ftab, "lim": lim, "smoothing": smoothing, "dom_off": dom_off,
        "mat_off": mat_off, "mats": mats, "refl": refl, "vwork": vwork, "slot_pt": slot_pt,
        "c_slot": c_slot, "c_coef": c_coef, "c_norm": c_norm})


def make_lattice_table(cap, device, stream=None):
    """The device table of the kernel function for ``cap`` (None without a
    lattice step)."""
    import torch

    if not cap.step:
        return None
    table = torch.empty((cap.a_max + 1) ** 2, dtype=torch.float64, device=device)
    lib = _native.load()
    _native.check(lib.piv_tps_phi_table(float(cap.step), int(cap.a_max), _p(table),
                                        _native.stream_ptr(stream)), "piv_tps_phi_table")
    return table
