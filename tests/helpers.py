"""Shared helpers of the parity tests (oracle-side checks run on the CPU)."""

import numpy as np


def near_tie_windows(maps64, mask=None, box=None, rel=2e-5):
    """Indices of windows whose two best valid map values (FP64 oracle maps)
    are closer than ``rel`` times the map's largest magnitude: a float32 map
    may order them either way, so their integer peak is not pinned."""
    maps64 = np.asarray(maps64)
    nwin, n0, n1 = maps64.shape
    valid = np.ones((n0, n1), bool)
    if mask is not None:
        valid &= ~mask
    if box is not None:
        (a, b), (c, d) = box
        bm = np.zeros((n0, n1), bool)
        bm[a:b, c:d] = True
        valid &= bm
    flat = np.where(valid[None], maps64, -np.inf).reshape(nwin, -1)
    top2 = -np.sort(-flat, axis=1)[:, :2]
    scale = np.abs(maps64).reshape(nwin, -1).max(axis=1)
    scale = np.where(scale > 0, scale, 1.0)
    return np.flatnonzero((top2[:, 0] - top2[:, 1]) <= rel * scale)


def random_windows(rng, nwin, shape, dtype=np.float32, scale=255.0):
    im = rng.random((nwin,) + tuple(shape)) * scale
    return im.astype(dtype)


def shifted_pair(rng, shape, shift, nb_particles=None, pad=8):
    """Gaussian-particle windows, im1 = im0 shifted by (dx, dy) px (the
    pattern of src/fluidimage/synthetic.py:6-61, seeded)."""
    ny, nx = shape
    if nb_particles is None:
        nb_particles = (max(nx, ny) // 4) ** 2
    xs = rng.uniform(-pad, nx + pad, nb_particles)
    ys = rng.uniform(-pad, ny + pad, nb_particles)
    X, Y = np.meshgrid(np.arange(nx, dtype=np.float64), np.arange(ny, dtype=np.float64))
    im0 = np.zeros(shape)
    im1 = np.zeros(shape)
    for x, y in zip(xs, ys):
        im0 += np.exp(-0.5 * ((X - x) ** 2 + (Y - y) ** 2))
        im1 += np.exp(-0.5 * ((X - x - shift[0]) ** 2 + (Y - y - shift[1]) ** 2))
    return im0.astype(np.float32), im1.astype(np.float32)


def grid_neighbourhood(indices, ny, nx):
    """The 3 x 3 grid neighbourhoods (ny, nx row-major vector grid) of the
    given vector indices, themselves included: the vectors whose WorkFIX
    neighbour mean reads them (mean_neighbors.py:5-125)."""
    out = set()
    for i in indices:
        r, c = divmod(int(i), nx)
        for dr in (-1, 0, 1):
            for dc in (-1, 0, 1):
                if 0 <= r + dr < ny and 0 <= c + dc < nx:
                    out.add((r + dr) * nx + c + dc)
    return out


def _influence_rects(res, changed, griddata_radius):
    """Rectangles (x0, x1, y0, y1) of the next interpolation that the changed
    vectors of a pass result can move: the limits of every TPS subdomain that
    holds one of them (thin_plate_spline_subdom.py:98-116, the oracle's
    object kept in res["tps"]), or a square of half-side griddata_radius
    around each of them for the piecewise-linear interpolation (its Delaunay
    star lies within two grid steps of a regular grid's node)."""
    xs = np.asarray(res["xs"], float)
    ys = np.asarray(res["ys"], float)
    rects = []
    tps = res.get("tps")
    for i in changed:
        x, y = xs[i], ys[i]
        if tps is not None:
            for iy in range(tps.nb_subdomy):
                y0, y1 = tps.limits_min_y[iy], tps.limits_max_y[iy]
                if not (y0 <= y < y1):
                    continue
                for ix in range(tps.nb_subdomx):
                    x0, x1 = tps.limits_min_x[ix], tps.limits_max_x[ix]
                    if x0 <= x < x1:
                        rects.append((x0, x1, y0, y1))
        else:
            r = griddata_radius
            rects.append((x - r, x + r, y - r, y + r))
    return rects


def _in_rects(xs, ys, rects, margin):
    hit = np.zeros(len(xs), bool)
    for x0, x1, y0, y1 in rects:
        hit |= (xs >= x0 - margin) & (xs <= x1 + margin) & (ys >= y0 - margin) & (ys <= y1 + margin)
    return set(np.flatnonzero(hit).tolist())


def allowed_mismatches(oracle_passes, works, ties, deform=False):
    """Vectors of each pass (and of the final field) that may differ from the
    oracle because of listed near-ties: per pass, the near-tie windows and
    the windows downstream of an earlier pass's allowed vectors (inside the
    interpolation's reach of them; with image deformation, windows whose crop
    reaches a moved node's bilinear cells), each with the 3 x 3
    neighbourhood that WorkFIX's neighbour mean couples to it.

    ``oracle_passes``: the oracle's per-pass result dicts (with their "tps"
    objects); ``works``: the oracle pass works (grids, window shapes);
    ``ties``: near-tie window indices per pass. Returns (list of sets per
    pass, set of allowed indices of the final field)."""
    allowed = []
    rects = []
    for p, (res, work) in enumerate(zip(oracle_passes, works)):
        ny, nx = len(work.iyvecs), len(work.ixvecs)
        gx = np.asarray(work.ixvecs_grid, float)
        gy = np.asarray(work.iyvecs_grid, float)
        margin = 0.0
        if deform and p > 0:
            step = max(np.diff(work.ixvecs).max(), np.diff(work.iyvecs).max())
            margin = max(work.shape_crop_im0) / 2 + step
        seeds = set(int(t) for t in ties[p]) | _in_rects(gx, gy, rects, margin)
        a = grid_neighbourhood(seeds, ny, nx)
        allowed.append(a)
        prev_step = max(np.diff(work.ixvecs).max(), np.diff(work.iyvecs).max())
        rects = _influence_rects(res, sorted(a), 2.0 * prev_step)
    last = works[-1]
    final = _in_rects(np.asarray(last.ixvecs_grid, float), np.asarray(last.iyvecs_grid, float),
                      rects, 0.0)
    return allowed, final
