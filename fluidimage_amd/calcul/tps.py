"""Thin-plate-spline smoothing on subdomains, FP64 on the GPU.

Same method and parameters as the reference's ``ThinPlateSplineSubdom``
(src/fluidimage/calcul/interpolate/thin_plate_spline_subdom.py:25-280) and
``compute_tps_weights`` (src/fluidimage/calcul/interpolate/thin_plate_spline.py:168-216):

* selection of the finite vectors, subdomain membership (half-open limits
  with 20 % buffers) and partition-of-unity weights
  (1 - x^2)(1 - y^2) + 0.001 run on the device; the host only turns five
  scalars (count and bounds) into the subdomain grid with the reference's
  formulas (subdomain counts, ``np.linspace`` limits, rescaled smoothing);
* per subdomain, the saddle-point system [[K + sI, P], [P^T, 0]] with
  K = r^2 ln(r^2) / 2 and P = [1, y, x] is solved in FP64 by libpiv_hip:
  Householder QR of P, congruence of K + sI onto the null space of P^T
  (symmetric positive definite there) and a blocked Cholesky factorisation,
  reused for the x and y components and for every outlier iteration
  (threshold_tps, at most 11 solves, thin_plate_spline_subdom.py:236-280);
* blending of the smoothed values and evaluation at the new positions.

The reference solves the same systems with LAPACK's LU (``np.linalg.solve``);
both give the unique solution up to FP64 rounding.
"""

import ctypes
from math import sqrt

import numpy as np

from .. import _native

MAX_SLOTS = 4
MAX_CENTERS_PER_SUBDOM = 2048
# order in which the per-subdomain kernels take the subdomains:
# "largest_first" (cost grows like n^3: the last round of a launch is made of
# the small ones) or "natural"; results do not depend on it
LAUNCH_ORDER = "largest_first"
MAX_LATTICE_A = 511  # largest lattice table: 512 x 512 doubles (2 MiB)
# entry point of the evaluation at the new positions: "workspace"
# (piv_tps_interpolate with a workspace allocated here) or "noalloc"
# (piv_tps_interpolate_noalloc, the library's fixed scratch; at most
# NOALLOC_SCRATCH_BYTES of workspace); the results are the same
INTERPOLATE_ENTRY = "workspace"
NOALLOC_SCRATCH_BYTES = 64 << 20


class SingularSystemError(ArithmeticError):
    """A subdomain system is singular (the reference's LinAlgError)."""


def subdomain_grid(count, x_min, x_max, y_min, y_max, subdom_size, smoothing_coef,
                   percent_buffer_area=20):
    """Host arithmetic of thin_plate_spline_subdom.py:53-90 on the bounds."""
    range_x = x_max - x_min
    range_y = y_max - y_min
    if range_x == 0 or range_y == 0:
        raise SingularSystemError("degenerate set of centers")
    aspect_ratio = range_y / range_x
    nb_subdom = count / subdom_size
    nsx = int(np.floor(sqrt(nb_subdom / aspect_ratio))) or 1
    nsy = int(np.ceil(sqrt(nb_subdom * aspect_ratio))) or 1
    x_dom = np.linspace(x_min, x_max, nsx + 1)
    y_dom = np.linspace(y_min, y_max, nsy + 1)
    num_vectors_per_subdom = count / (nsx * nsy)
    smoothing = smoothing_coef * sqrt(
        (x_dom[1] - x_dom[0]) * (y_dom[1] - y_dom[0]) / num_vectors_per_subdom)
    coef_buffer = percent_buffer_area / 100
    bx = coef_buffer * range_x / nsx
    by = coef_buffer * range_y / nsy
    lim = np.concatenate([x_dom[:-1] - bx, x_dom[1:] + bx, y_dom[:-1] - by, y_dom[1:] + by])
    return nsx, nsy, smoothing, lim


def subdomain_grids(counts, x_min, x_max, y_min, y_max, subdom_size, smoothing_coef,
                    percent_buffer_area=20):
    """``subdomain_grid`` for many fields at once (arrays of per-field
    scalars): fields with the same subdomain counts share one vectorised
    ``np.linspace`` (the same arithmetic, element by element, as one call per
    field). Returns one (nsx, nsy, smoothing, lim) tuple per field, or the
    SingularSystemError of a degenerate field."""
    counts = np.asarray(counts, dtype=np.float64)
    x_min, x_max, y_min, y_max = (np.asarray(a, dtype=np.float64)
                                  for a in (x_min, x_max, y_min, y_max))
    range_x = x_max - x_min
    range_y = y_max - y_min
    bad = (range_x == 0) | (range_y == 0)
    with np.errstate(divide="ignore", invalid="ignore"):
        aspect_ratio = range_y / range_x
        nb_subdom = counts / subdom_size
        nsx = np.floor(np.sqrt(nb_subdom / aspect_ratio))
        nsy = np.ceil(np.sqrt(nb_subdom * aspect_ratio))
    out = [SingularSystemError("degenerate set of centers") if b else None
           for b in bad.tolist()]
    ok = ~bad
    nsx = np.where(ok & (nsx > 0), nsx, 1).astype(np.int64)
    nsy = np.where(ok & (nsy > 0), nsy, 1).astype(np.int64)
    coef_buffer = percent_buffer_area / 100
    idx_ok = np.flatnonzero(ok)
    if idx_ok.size == 0:
        return out
    # fields grouped by their (nsx, nsy): one vectorised pass per group
    keys, inverse = np.unique(nsx[idx_ok] * (1 << 32) + nsy[idx_ok], return_inverse=True)
    for g, key in enumerate(keys.tolist()):
        sx, sy = key >> 32, key & 0xFFFFFFFF
        ks = idx_ok[inverse.ravel() == g]
        x_dom = np.linspace(x_min[ks], x_max[ks], sx + 1, axis=1)
        y_dom = np.linspace(y_min[ks], y_max[ks], sy + 1, axis=1)
        num_vectors_per_subdom = counts[ks] / (sx * sy)
        smoothing = smoothing_coef * np.sqrt(
            (x_dom[:, 1] - x_dom[:, 0]) * (y_dom[:, 1] - y_dom[:, 0]) / num_vectors_per_subdom)
        bx = (coef_buffer * range_x[ks] / sx)[:, None]
        by = (coef_buffer * range_y[ks] / sy)[:, None]
        lim = np.concatenate([x_dom[:, :-1] - bx, x_dom[:, 1:] + bx, y_dom[:, :-1] - by,
                              y_dom[:, 1:] + by], axis=1)
        for k, sm, row in zip(ks.tolist(), smoothing.tolist(), lim):
            out[k] = (sx, sy, sm, row)
    return out


def matrix_sizes(sizes):
    """Doubles of the matrix region of subdomains of `sizes` centers (as
    piv_tps_matrix_size): (n-3) x 3 then (n-3) rows of the factorised matrix,
    both padded to multiples of 16 doubles (128-byte rows)."""
    m = np.maximum(np.asarray(sizes, dtype=np.int64) - 3, 0)
    pad = lambda v: (v + 15) // 16 * 16  # noqa: E731
    return pad(3 * m) + m * pad(m)


def _p(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def lattice_step(*arrays):
    """Largest step h = g / 2**k (g a positive integer, k <= 2) of which every
    value of the host arrays is an integer multiple, or None. The PIV grids
    (window centres w // 2 + i step) give the grid step, or its gcd with the
    next pass's step."""
    vals = np.concatenate([np.ravel(np.asarray(a, dtype=np.float64)) for a in arrays])
    if vals.size == 0 or not np.all(np.isfinite(vals)):
        return None
    for k in range(3):
        scaled = vals * (1 << k)
        ints = np.rint(scaled)
        if np.all(ints == scaled) and np.all(np.abs(ints) < 2**31):
            g = int(np.gcd.reduce(np.abs(ints).astype(np.int64)))
            return None if g == 0 else g / (1 << k)
    return None


def _lattice(lib, step, fields_active, dev, sp):
    """Device table of the kernel function on the lattice of step ``step``,
    large enough for any two points of one subdomain (centers and new
    positions lie inside its limits); None when there is no step or the
    table would exceed MAX_LATTICE_A."""
    import torch

    if not step:
        return None, None
    extent = 0.0
    for f in fields_active:
        nsx = f.nsx
        lx = f.lim_h[nsx:2 * nsx] - f.lim_h[:nsx]
        ly = f.lim_h[2 * nsx + f.nsy:] - f.lim_h[2 * nsx:2 * nsx + f.nsy]
        extent = max(extent, float(lx.max()), float(ly.max()))
    a_max = int(np.ceil(extent / step)) + 1
    if a_max > MAX_LATTICE_A:
        return None, None
    table = torch.empty((a_max + 1) ** 2, dtype=torch.float64, device=dev)
    _native.check(lib.piv_tps_phi_table(float(step), a_max, _p(table), sp), "piv_tps_phi_table")
    lat = _native.PivTpsLattice(table.data_ptr(), float(step), a_max, 0)
    return lat, table


def _upload(arrays, device):
    """Host numpy arrays to device tensors in one pinned, asynchronous copy
    (8-byte aligned slices of one byte buffer)."""
    import torch

    offs, total = [], 0
    for a in arrays:
        offs.append(total)
        total += (a.nbytes + 7) // 8 * 8
    host = torch.empty(max(total, 8), dtype=torch.uint8, pin_memory=True)
    hv = host.numpy()
    for a, o in zip(arrays, offs):
        hv[o:o + a.nbytes] = np.ascontiguousarray(a).view(np.uint8).ravel()
    dev = host.to(device, non_blocking=True)
    out = []
    for a, o in zip(arrays, offs):
        t = dev[o:o + a.nbytes].view(getattr(torch, str(a.dtype)))
        out.append(t.view(a.shape))
    return out


def tps_device(dx, dy, xs, ys, new_x, new_y, subdom_size, smoothing_coef, threshold,
               percent_buffer_area=20, stream=None, keep=False, lattice=None):
    """TPS smoothing of the finite vectors of (dx, dy) at (xs, ys) and
    evaluation at (new_x, new_y); all tensors on the device (dx, dy float32,
    positions float64). Small device-to-host reads size the problem.

    Returns a dict of device tensors: ``smoothed`` (2, count) at the
    selected centers, ``index`` (count) of the selected vectors, ``approx``
    (2, M) at the new positions, plus host ``nb_subdom``, ``stats`` (nd, 6).
    ``lattice``: a step of which the positions are multiples (see
    ``lattice_step``), or None; it changes the speed, not the result.
    """
    out = tps_device_batch([(dx, dy, xs, ys)], new_x, new_y, subdom_size, smoothing_coef,
                           threshold, percent_buffer_area, stream=stream, keep=keep,
                           lattice=lattice)[0]
    if isinstance(out, Exception):
        raise out
    return out


class _Field:
    """Host bookkeeping of one vector field of a batch."""

    def __init__(self, k, n, base):
        self.k = k
        self.n = n          # vectors of the field
        self.base = base    # position of its first center in the batch arrays
        self.sorted_y = 0   # centers sorted by y (membership scans a y range)
        self.error = None


def tps_device_batch(fields, new_x, new_y, subdom_size, smoothing_coef, threshold,
                     percent_buffer_area=20, stream=None, keep=False, lattice=None,
                     defer=False):
    """``tps_device`` for several vector fields at once (one image pair
    each): each field gets its own subdomain grid (thin_plate_spline_subdom.py:53-90),
    and the subdomains of all fields go through one launch of every
    per-subdomain kernel. ``fields`` is a list of (dx, dy, xs, ys) device
    tensors; (new_x, new_y) are the new positions, the same for every field.

    Returns one entry per field: the dict of ``tps_device`` or the
    ``SingularSystemError`` of that field alone ("Only nan." when it has no
    finite vector). ``lattice``: as in ``tps_device``. With ``defer`` the
    call does not wait for the solves: the check for singular systems is
    left to ``deferred_error`` on each returned dict (its "pending" entry).
    With a ``stream``, every allocation, copy, launch and host read of the
    call is ordered on it (results are ready on that stream).
    """
    from ..device import on_stream

    with on_stream(stream):
        return _tps_device_batch(fields, new_x, new_y, subdom_size, smoothing_coef, threshold,
                                 percent_buffer_area, stream, keep, lattice, defer)


def _tps_device_batch(fields, new_x, new_y, subdom_size, smoothing_coef, threshold,
                      percent_buffer_area, stream, keep, lattice, defer):
    import torch

    lib = _native.load()
    dev = fields[0][0].device
    sp = _native.stream_ptr(stream)
    f64 = dict(dtype=torch.float64, device=dev)
    i32 = dict(dtype=torch.int32, device=dev)
    thr = -1.0 if threshold is None else float(threshold)
    n = int(fields[0][0].numel())
    if any(int(f[0].numel()) != n for f in fields):
        raise ValueError("the fields of a TPS batch have one number of vectors")
    if n == 0:
        return [SingularSystemError("Only nan.") for _ in fields]
    flist = [_Field(k, n, k * n) for k in range(len(fields))]
    ntot = n * len(fields)

    def stacked(i):
        t = [f[i] for f in fields]
        if len(t) == 1:
            return t[0].contiguous()
        # consecutive slices of one tensor (the pairs of a batch): a view
        t0 = t[0]
        if all(x.is_contiguous() and x.dtype == t0.dtype and x.device == t0.device
               and x.untyped_storage().data_ptr() == t0.untyped_storage().data_ptr()
               and x.storage_offset() == t0.storage_offset() + k * n
               for k, x in enumerate(t)):
            return torch.as_strided(t0, (len(t) * n,), (1,), t0.storage_offset())
        return torch.cat(t)

    dx, dy, xs, ys = (stacked(i) for i in range(4))
    cx = torch.empty(ntot, **f64)
    cy = torch.empty(ntot, **f64)
    vx = torch.empty(ntot, dtype=torch.float32, device=dev)
    vy = torch.empty(ntot, dtype=torch.float32, device=dev)
    idx = torch.empty(ntot, **i32)
    info = torch.empty((len(fields), 7), **f64)
    _native.check(lib.piv_tps_select(len(fields), n, _p(dx), _p(dy), _p(xs), _p(ys), _p(cx),
                                     _p(cy), _p(vx), _p(vy), _p(idx), _p(info),
                                     float(lattice or 0.0), sp),
                  "piv_tps_select")
    info_h = info.cpu().numpy()
    # subdomain grids (host scalars, all fields at once)
    grids = subdomain_grids(info_h[:, 0], info_h[:, 1], info_h[:, 2], info_h[:, 3],
                            info_h[:, 4], subdom_size, smoothing_coef, percent_buffer_area)
    active = []
    for f, g in zip(flist, grids):
        f.count = int(info_h[f.k, 0])
        f.sorted_y = int(info_h[f.k, 5])
        f.on_lattice = bool(info_h[f.k, 6])
        if f.count == 0:
            f.error = SingularSystemError("Only nan.")
        elif isinstance(g, SingularSystemError):
            f.error = g
        else:
            f.nsx, f.nsy, f.smoothing, f.lim_h = g
            f.nd = f.nsx * f.nsy
            active.append(f)
    own = (xs, ys, idx, n)
    results = _solve_active(lib, active, cx, cy, vx, vy, new_x, new_y, thr, dev, sp, keep, own,
                            lattice, defer, stream)
    out = []
    for f in flist:
        if f.error is not None:
            out.append(f.error)
        else:
            r = results[f.k]
            r["index"] = idx[f.base:f.base + f.count]
            out.append(r)
    return out


def _at_centers(active, own, smoothed, new_x, new_y, dev):
    """Positions of the evaluation that are centers: where the new position
    q of field j is the position of its vector q and that vector is valid,
    the evaluation equals the blended smoothed value of that vector
    ([K P][w; a] = v - s w at the centers). Returns (approx prefilled there,
    uint8 mask of those positions), or (None, None) when the new positions
    are not the fields' own grid."""
    import torch

    xs, ys, idx, n = own
    m = int(new_x.numel())
    if m != n:
        return None, None
    na = len(active)
    ks, counts = _upload([np.array([f.k for f in active], np.int64),
                          np.array([f.count for f in active], np.int64)], dev)
    r = torch.arange(na * n, device=dev)
    j = torch.div(r, n, rounding_mode="floor")
    o = r - j * n
    src = ks[j] * n + o  # stacked valid-list position (field base k n + rank)
    valid = o < counts[j]
    tgt = torch.where(valid, j * m + idx[src].long(), na * m)
    pos = torch.full((na * m + 1,), -1, dtype=torch.int64, device=dev)
    pos.index_put_((tgt,), torch.where(valid, src, -1))
    pos = pos[:na * m]
    sel = ks[j] * n + o  # the field's own position of vector o
    done = (pos >= 0) & (new_x.repeat(na) == xs[sel]) & (new_y.repeat(na) == ys[sel])
    approx = torch.where(done, smoothed[:, pos.clamp(min=0)], 0.0)
    return approx.contiguous(), done.to(torch.uint8)


class PendingStatus:
    """Device status array of one batched TPS call whose check is deferred:
    one value per subdomain (nonzero: singular system), then the flag of a
    point in more than 4 subdomains."""

    def __init__(self, status, nd, keep_alive=()):
        self.status = status
        self.nd = nd
        self._keep_alive = keep_alive
        self._host = None

    def host(self):
        if self._host is None:
            self._host = self.status.cpu().numpy()
            self._keep_alive = ()
        return self._host


def deferred_error(entry):
    """Check of a field solved with ``defer=True``: None, or the
    SingularSystemError the eager call would have returned for it; raises
    NotImplementedError when a point lay in more than 4 subdomains."""
    pending, base, nd = entry["pending"]
    st = pending.host()
    if st[pending.nd]:
        raise NotImplementedError("a point lies in more than 4 TPS subdomains")
    bad = st[base:base + nd]
    if bad.any():
        return SingularSystemError(
            f"{int((bad != 0).sum())} subdomain system(s) not positive definite")
    return None


def _solve_active(lib, active, cx, cy, vx, vy, new_x, new_y, thr, dev, sp, keep, own,
                  lattice=None, defer=False, stream=None):
    """Membership, FP64 solves, blend and evaluation for the fields whose
    grid is valid; fields with an empty subdomain get their error and the
    others are solved again without them."""
    import torch

    f64 = dict(dtype=torch.float64, device=dev)
    i32 = dict(dtype=torch.int32, device=dev)
    if not active:
        return {}
    nds = np.array([f.nd for f in active], dtype=np.int64)
    nd_bases = np.concatenate([[0], np.cumsum(nds)])
    for f, b in zip(active, nd_bases.tolist()):
        f.nd_base = b
    nd = int(nd_bases[-1])
    lim_h = np.concatenate([f.lim_h for f in active])
    lim_off = np.concatenate([[0], np.cumsum([f.lim_h.size for f in active])])
    cols = [[f.nsx for f in active], [f.nsy for f in active], [f.count for f in active],
            [f.base for f in active], nd_bases[:-1], lim_off[:-1], [f.sorted_y for f in active]]
    table = np.array(cols, dtype=np.int32).T.copy()
    dom_field_h = np.repeat(np.arange(len(active), dtype=np.int32), nds)
    smoothing_h = np.repeat(np.array([f.smoothing for f in active], dtype=np.float64), nds)
    lim, ftab, dom_field, smoothing = _upload([lim_h, table, dom_field_h, smoothing_h], dev)
    counts = torch.empty(nd, **i32)
    _native.check(lib.piv_tps_count(nd, _p(ftab), _p(dom_field), _p(lim), _p(cx), _p(cy),
                                    _p(counts), sp), "piv_tps_count")
    sizes = counts.cpu().numpy().astype(np.int64)
    smin = np.minimum.reduceat(sizes, nd_bases[:-1]).tolist()
    smax = np.maximum.reduceat(sizes, nd_bases[:-1]).tolist()
    retry = False
    for f, lo, hi in zip(active, smin, smax):
        if lo < 1:
            f.error = SingularSystemError("empty subdomain")
            retry = True
        elif hi > MAX_CENTERS_PER_SUBDOM:
            raise NotImplementedError(
                f"{hi} centers in a TPS subdomain (limit {MAX_CENTERS_PER_SUBDOM}): "
                "use a smaller params.multipass.subdom_size")
    if retry:
        return _solve_active(lib, [f for f in active if f.error is None], cx, cy, vx, vy,
                             new_x, new_y, thr, dev, sp, keep, own, lattice, defer, stream)
    n_max = int(sizes.max())
    dom_off_h = np.concatenate([[0], np.cumsum(sizes)])
    mat_off_h = np.concatenate([[0], np.cumsum(matrix_sizes(sizes))])
    nslots = int(dom_off_h[-1])
    # per-subdomain work launched largest first (its cost grows like n^3):
    # the last round of a launch is then made of the small subdomains
    if LAUNCH_ORDER == "largest_first":
        order_h = np.argsort(-sizes, kind="stable").astype(np.int32)
    elif LAUNCH_ORDER == "natural":
        order_h = None
    else:
        raise ValueError(f"unknown LAUNCH_ORDER {LAUNCH_ORDER!r}")
    uploads = [dom_off_h.astype(np.int64), mat_off_h.astype(np.int64)]
    if order_h is not None:
        uploads.append(order_h)
    dom_off, mat_off, *rest = _upload(uploads, dev)
    order = rest[0] if rest else None
    slot_pt = torch.empty(nslots, **i32)
    status = torch.zeros(nd + 1, **i32)
    ntot = int(cx.numel())
    # slots of every center of the batch (-1: none, so that the centers
    # beyond a field's count blend to nothing in the one blend launch below):
    # the centers mode of piv_tps_point_slots writes every entry of an active
    # field (empty ones beyond its count); the fields left out are prefilled
    n_field = own[3]
    if ntot == n_field * len(active):
        c_slot = torch.empty(ntot * MAX_SLOTS, **i32)
        c_coef = torch.empty(ntot * MAX_SLOTS, **f64)
        c_norm = torch.empty(ntot, **f64)
    else:
        c_slot = torch.full((ntot * MAX_SLOTS,), -1, **i32)
        c_coef = torch.zeros(ntot * MAX_SLOTS, **f64)
        c_norm = torch.ones(ntot, **f64)
    m = int(new_x.numel())
    na = len(active)
    n_dom = torch.empty(na * m * MAX_SLOTS, **i32)
    n_coef = torch.empty(na * m * MAX_SLOTS, **f64)
    n_norm = torch.empty(na * m, **f64)
    _native.check(lib.piv_tps_fill(nd, _p(ftab), _p(dom_field), _p(lim), _p(cx), _p(cy),
                                   _p(dom_off), _p(slot_pt), sp), "piv_tps_fill")
    _native.check(lib.piv_tps_point_slots(
        na, _p(ftab), n_field, _p(lim), _p(cx), _p(cy), _p(dom_off),
        _p(slot_pt), _p(c_slot), _p(c_coef), _p(c_norm), _p(status[nd:]), sp),
        "piv_tps_point_slots")
    _native.check(lib.piv_tps_point_slots(
        na, _p(ftab), m, _p(lim), _p(new_x), _p(new_y), None, None, _p(n_dom), _p(n_coef),
        _p(n_norm), _p(status[nd:]), sp), "piv_tps_point_slots")
    mats = torch.empty(int(mat_off_h[-1]), **f64)
    refl = torch.empty((nd, 16), **f64)
    vwork = torch.empty(nslots * 6, **f64)
    smooth_slots = torch.empty((2, nslots), **f64)
    wstride = nslots + 3 * nd
    weights = torch.empty((2, wstride), **f64)
    stats = torch.zeros((nd, 6), **f64)
    slot_xy = torch.empty((2, nslots), **f64)
    lat, lat_table = _lattice(lib, lattice, active, dev, sp)
    lat_p = None if lat is None else ctypes.byref(lat)
    # every center of every field on the lattice: the factorisation generates
    # the matrix from the table instead of reading it from the build
    generated = int(lat is not None and all(f.on_lattice for f in active))
    _native.check(
        lib.piv_tps_smooth(nd, n_max, _p(cx), _p(cy), _p(vx), _p(vy), _p(dom_off), _p(slot_pt),
                           _p(mat_off), _p(smoothing), thr, _p(mats), _p(refl), _p(vwork),
                           _p(slot_xy), _p(smooth_slots), nslots, _p(weights), wstride,
                           _p(stats), _p(status), _p(order), lat_p, generated, sp),
        "piv_tps_smooth")
    smoothed = torch.empty((2, ntot), **f64)
    _native.check(lib.piv_tps_blend(ntot, _p(c_slot), _p(c_coef), _p(c_norm), _p(smooth_slots),
                                    nslots, _p(smoothed), sp), "piv_tps_blend")
    px = new_x.repeat(na)
    py = new_y.repeat(na)
    approx, skip = _at_centers(active, own, smoothed, new_x, new_y, dev)
    if approx is None:
        approx = torch.empty((2, na * m), **f64)
    if INTERPOLATE_ENTRY == "noalloc":
        iwork = None
        _native.check(lib.piv_tps_interpolate_noalloc(
            na * m, _p(px), _p(py), _p(n_dom), _p(n_coef), _p(n_norm), _p(slot_xy), nslots, nd,
            n_max, _p(dom_off), _p(weights), wstride, _p(skip), _p(approx), lat_p, sp),
            "piv_tps_interpolate_noalloc")
    elif INTERPOLATE_ENTRY == "workspace":
        wbytes = int(lib.piv_tps_interpolate_workspace(na * m, nd))
        iwork = torch.empty(max(wbytes, 8), dtype=torch.uint8, device=dev)
        _native.check(lib.piv_tps_interpolate(na * m, _p(px), _p(py), _p(n_dom), _p(n_coef),
                                              _p(n_norm), _p(slot_xy), nslots, nd, n_max,
                                              _p(dom_off), _p(weights), wstride, _p(skip),
                                              _p(approx), _p(iwork), lat_p, sp),
                      "piv_tps_interpolate")
    else:
        raise ValueError(f"unknown INTERPOLATE_ENTRY {INTERPOLATE_ENTRY!r}")
    if defer:
        # no synchronisation here: the caller checks the status later
        # (deferred_error); the workspaces stay referenced until then when
        # the launches are not on the current stream (on the current stream
        # the allocator orders any reuse of their memory after the launches)
        keep_alive = ()
        if stream is not None and stream != torch.cuda.current_stream():
            keep_alive = (mats, refl, vwork, smooth_slots, weights, stats, slot_xy, dom_off,
                          mat_off, slot_pt, c_slot, c_coef, c_norm, n_dom, n_coef, n_norm,
                          lat_table, iwork, lim, ftab, dom_field, smoothing, counts, px, py,
                          order)
        pending = PendingStatus(status, nd, keep_alive)
        st = None
    else:
        pending = None
        st = status.cpu().numpy()  # synchronises: the workspaces are free to go after this
        del lat_table, iwork
        if st[nd]:
            raise NotImplementedError("a point lies in more than 4 TPS subdomains")
    results = {}
    for j, f in enumerate(active):
        if st is not None:
            bad = st[f.nd_base:f.nd_base + f.nd]
            if bad.any():
                f.error = SingularSystemError(
                    f"{int((bad != 0).sum())} subdomain system(s) not positive definite")
                continue
        r = {"smoothed": smoothed[:, f.base:f.base + f.count],
             "approx": approx[:, m * j:m * (j + 1)],
             "nb_subdom": f.nd, "nsx": f.nsx, "nsy": f.nsy, "smoothing_coef": f.smoothing,
             "count": f.count, "stats": stats[f.nd_base:f.nd_base + f.nd]}
        if pending is not None:
            r["pending"] = (pending, f.nd_base, f.nd)
        if keep:
            # single-field use: subdomain offsets and center indices are the field's own
            r.update(weights=weights, dom_off=dom_off_h, slot_pt=slot_pt)
        results[f.k] = r
    return results


def summaries(stats):
    """Reference-style summaries (x then y) from the per-subdomain stats."""
    s = stats.cpu().numpy() if hasattr(stats, "cpu") else np.asarray(stats)
    out = []
    for c in range(2):
        out.append({
            "nb_fixed_vectors": [int(v) for v in s[:, 3 * c]],
            "max(Udiff)": [float(v) for v in s[:, 3 * c + 1]],
            "nb_iterations": [int(v) for v in s[:, 3 * c + 2]],
            "nb_fixed_vectors_tot": int(s[:, 3 * c].sum()),
        })
    return out


class ThinPlateSplineSubdomHIP:
    """Host-array front end of ``tps_device`` (``centers = vstack([ys, xs])``)."""

    def __init__(self, centers, subdom_size, smoothing_coef, threshold=None,
                 percent_buffer_area=20):
        self.centers = np.asarray(centers, dtype=np.float64)
        self.subdom_size = subdom_size
        self.smoothing_coef_param = smoothing_coef
        self.threshold = threshold
        self.percent_buffer_area = percent_buffer_area
        self.num_centers = self.centers.shape[1]

    def smooth_and_interpolate(self, deltaxs, deltays, new_positions, stream=None):
        """compute_tps_weights_subdom for x and y, then init_with_new_positions
        + interpolate on ``new_positions = vstack([ys, xs])``.

        Returns (dxs_smooth, dys_smooth, dxs_interp, dys_interp, summary)
        where summary is the y-component summary (the one the reference
        keeps after its second call).
        """
        import torch

        dev = torch.device("cuda")

        def t(a, dtype):
            return torch.from_numpy(np.ascontiguousarray(a, dtype=dtype)).to(dev)

        new_positions = np.asarray(new_positions, dtype=np.float64)
        out = tps_device(t(deltaxs, np.float32), t(deltays, np.float32),
                         t(self.centers[1], np.float64), t(self.centers[0], np.float64),
                         t(new_positions[1], np.float64), t(new_positions[0], np.float64),
                         self.subdom_size, self.smoothing_coef_param, self.threshold,
                         self.percent_buffer_area, stream=stream, keep=True)
        self.nb_subdom = out["nb_subdom"]
        self.nb_subdomx = out["nsx"]
        self.nb_subdomy = out["nsy"]
        self.smoothing_coef = out["smoothing_coef"]
        dom_off = out["dom_off"]
        slot_pt = out["slot_pt"].cpu().numpy()
        index = out["index"].cpu().numpy()
        self.indices_domains = [index[slot_pt[dom_off[d]:dom_off[d + 1]]]
                                for d in range(self.nb_subdom)]
        w = out["weights"].cpu().numpy()
        self.weights_x, self.weights_y = [], []
        for d in range(self.nb_subdom):
            a, b = dom_off[d], dom_off[d + 1]
            base = a + 3 * d
            self.weights_x.append(w[0, base:base + (b - a) + 3].copy())
            self.weights_y.append(w[1, base:base + (b - a) + 3].copy())
        sx, sy = summaries(out["stats"])
        self.summary_x = sx
        sm = out["smoothed"].cpu().numpy()
        ap = out["approx"].cpu().numpy()
        return sm[0], sm[1], ap[0], ap[1], sy
