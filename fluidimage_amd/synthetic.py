"""Seeded synthetic PIV image pairs.

Plays the role of src/fluidimage/synthetic.py:6-61 (Gaussian particle
images shifted by a known displacement), with a seeded generator and
displacement fields:

* particles: I(r) = I0 exp(-8 r^2 / d^2), d (the e^-2 diameter) uniform in
  [2, 3] px, I0 uniform in [150, 255] (x16 for 12-bit images);
* im1 renders every particle at its advected position x + u(x) (no image
  resampling), then both images get additive Gaussian noise (sigma 3, x16
  for 12-bit), are rounded and clipped to the dtype range.
"""

import numpy as np

SEED_BASE = 20261203


def lamb_oseen(shape, core_radius=60.0, umax=4.0, shift=(1.3, -0.7), center=None):
    """Lamb-Oseen vortex scaled to ``umax`` plus a uniform shift (ux, uy)."""
    ny, nx = shape
    if center is None:
        center = ((nx - 1) / 2.0, (ny - 1) / 2.0)
    xc, yc = center
    # max of (1 - exp(-s^2)) / s at s = r / rc ~ 1.1209
    s_star = 1.1209
    g_peak = (1 - np.exp(-(s_star**2))) / s_star

    def field(x, y):
        dx = x - xc
        dy = y - yc
        r = np.hypot(dx, dy)
        s = r / core_radius
        with np.errstate(invalid="ignore", divide="ignore"):
            g = np.where(r > 0, (1 - np.exp(-(s**2))) / np.where(s > 0, s, 1), 0.0)
        utheta = umax * g / g_peak
        with np.errstate(invalid="ignore", divide="ignore"):
            ux = np.where(r > 0, -utheta * dy / np.where(r > 0, r, 1), 0.0)
            uy = np.where(r > 0, utheta * dx / np.where(r > 0, r, 1), 0.0)
        return ux + shift[0], uy + shift[1]

    return field


def fourier_modes(shape, rng, nmodes=8, lmin=128.0, lmax=512.0, umax=6.0):
    """Sum of ``nmodes`` plane-wave modes, wavelengths uniform in
    [lmin, lmax], random direction, phase and amplitude vector, normalised
    so that max |u| over the image grid equals ``umax``."""
    ny, nx = shape
    lam = rng.uniform(lmin, lmax, nmodes)
    theta = rng.uniform(0, 2 * np.pi, nmodes)
    phase = rng.uniform(0, 2 * np.pi, nmodes)
    amp = rng.normal(size=(nmodes, 2))
    kx = 2 * np.pi / lam * np.cos(theta)
    ky = 2 * np.pi / lam * np.sin(theta)

    def raw(x, y):
        ux = np.zeros(np.broadcast(x, y).shape)
        uy = np.zeros_like(ux)
        for m in range(nmodes):
            c = np.cos(kx[m] * x + ky[m] * y + phase[m])
            ux += amp[m, 0] * c
            uy += amp[m, 1] * c
        return ux, uy

    ys, xs = np.mgrid[0:ny:4, 0:nx:4]
    ux, uy = raw(xs.astype(float), ys.astype(float))
    scale = umax / np.sqrt(ux**2 + uy**2).max()

    def field(x, y):
        ux, uy = raw(x, y)
        return ux * scale, uy * scale

    return field


def shear_flow(shape, rng, umax):
    ny, nx = shape
    angle = rng.uniform(0, 2 * np.pi)
    ca, sa = np.cos(angle), np.sin(angle)

    def field(x, y):
        t = (-(x - nx / 2) * sa + (y - ny / 2) * ca) / (0.5 * max(nx, ny))
        u = umax * np.tanh(2 * t)
        return u * ca, u * sa

    return field


def render_particles(shape, xp, yp, diam, inten):
    """Accumulate Gaussian particles on a float64 (ny, nx) image."""
    ny, nx = shape
    img = np.zeros(ny * nx)
    ix = np.floor(xp).astype(np.int64)
    iy = np.floor(yp).astype(np.int64)
    coef = -8.0 / diam**2
    for oy in range(-2, 4):
        for ox in range(-2, 4):
            px = ix + ox
            py = iy + oy
            ok = (px >= 0) & (px < nx) & (py >= 0) & (py < ny)
            val = inten * np.exp(coef * ((px - xp) ** 2 + (py - yp) ** 2))
            img += np.bincount(py[ok] * nx + px[ok], weights=val[ok], minlength=ny * nx)
    return img.reshape(shape)


def make_pair(shape, field, rng, density=0.02, dtype=np.uint8, noise=3.0,
              outlier_windows=None):
    """Return (im0, im1) of ``dtype``; ``field(x, y) -> (ux, uy)`` in px.

    ``outlier_windows``: optional list of (y0, x0, h, w) boxes of im1 that
    are replaced by an independent particle image (uncorrelated noise).
    """
    ny, nx = shape
    bits12 = np.dtype(dtype) == np.uint16
    gain = 16.0 if bits12 else 1.0
    margin = 16
    area = (ny + 2 * margin) * (nx + 2 * margin)
    npart = rng.poisson(density * area)
    xp = rng.uniform(-margin, nx + margin, npart)
    yp = rng.uniform(-margin, ny + margin, npart)
    diam = rng.uniform(2.0, 3.0, npart)
    inten = rng.uniform(150, 255, npart) * gain
    im0 = render_particles(shape, xp, yp, diam, inten)
    ux, uy = field(xp, yp)
    im1 = render_particles(shape, xp + ux, yp + uy, diam, inten)
    if outlier_windows:
        for (y0, x0, h, w) in outlier_windows:
            k = rng.poisson(density * h * w)
            sub = render_particles((h, w), rng.uniform(0, w, k), rng.uniform(0, h, k),
                                   rng.uniform(2.0, 3.0, k), rng.uniform(150, 255, k) * gain)
            im1[y0:y0 + h, x0:x0 + w] = sub
    im0 = im0 + rng.normal(0.0, noise * gain, shape)
    im1 = im1 + rng.normal(0.0, noise * gain, shape)
    top = 4095 if bits12 else np.iinfo(dtype).max if np.issubdtype(dtype, np.integer) else None
    if np.issubdtype(np.dtype(dtype), np.integer):
        im0 = np.clip(np.round(im0), 0, top).astype(dtype)
        im1 = np.clip(np.round(im1), 0, top).astype(dtype)
    else:
        im0 = im0.astype(dtype)
        im1 = im1.astype(dtype)
    return im0, im1


def _field_and_boxes(config_id, rng, shape):
    """Displacement field and noise boxes of configs 2 and 3 (first draws of
    the pair's generator)."""
    field = fourier_modes(shape, rng)
    boxes = None
    if config_id == 3:
        ny, nx = shape
        step = 32
        cand = [(y, x) for y in range(0, ny - 64 + 1, step)
                for x in range(0, nx - 64 + 1, step)]
        nsel = max(1, int(0.02 * len(cand)))
        sel = rng.choice(len(cand), nsel, replace=False)
        boxes = [(cand[i][0], cand[i][1], 64, 64) for i in sel]
    return field, boxes


def config_truth(config_id, pair_idx=0, shape=None):
    """(field, boxes): the displacement field ``field(x, y) -> (ux, uy)`` the
    pair ``config_pair(config_id, pair_idx, shape)`` was rendered with, and the
    (y0, x0, h, w) boxes of its second image replaced by noise (None).
    Configs 2 and 3."""
    if config_id not in (2, 3):
        raise ValueError("truth fields are kept for configs 2 and 3")
    rng = np.random.default_rng(SEED_BASE + 1000 * config_id + pair_idx)
    return _field_and_boxes(config_id, rng, shape or (2048, 2048))


def config_pair(config_id, pair_idx=0, shape=None):
    """Seeded pairs of the benchmark workloads.

    1: 512x512 uint8, Lamb-Oseen (rc 60, max 4 px) + shift (1.3, -0.7),
       density 0.02.
    2: 2048x2048 uint16 (12-bit), 8 Fourier modes (wavelength 128-512 px,
       max 6 px), density 0.03.
    3: as 2, plus 2 % of the 64 px windows of im1 replaced by noise.
    5: 1024x1024 float32, same field family as 2.
    ``shape`` overrides the image size (smaller test cases).
    """
    rng = np.random.default_rng(SEED_BASE + 1000 * config_id + pair_idx)
    if config_id == 1:
        shape = shape or (512, 512)
        return make_pair(shape, lamb_oseen(shape), rng, density=0.02, dtype=np.uint8)
    if config_id in (2, 3):
        shape = shape or (2048, 2048)
        field, boxes = _field_and_boxes(config_id, rng, shape)
        return make_pair(shape, field, rng, density=0.03, dtype=np.uint16,
                         outlier_windows=boxes)
    if config_id == 5:
        shape = shape or (1024, 1024)
        return make_pair(shape, fourier_modes(shape, rng), rng, density=0.03,
                         dtype=np.float32)
    raise ValueError(f"unknown config {config_id}")


UNEVEN_SIZES = [(1024, 1024), (1200, 1600), (2048, 2048), (2160, 2560)]


def uneven_batch(npairs=32, seed_offset=0, sizes=UNEVEN_SIZES, return_fields=False):
    """Config-4 style batch: random sizes, schedules and Oseen + shear fields.

    Returns a list of (im0, im1, windows_schedule) with uint16 images
    (``sizes`` overrides the image sizes, for smaller test cases), and with
    ``return_fields`` also the list of the displacement fields (callables
    (x, y) -> (u, v)) the pairs were rendered with.
    """
    rng = np.random.default_rng(SEED_BASE + 4000 + seed_offset)
    schedules = [(128, 64, 32), (64, 32, 16), (48, 24)]
    out = []
    fields = []
    for _ in range(npairs):
        shape = sizes[rng.integers(len(sizes))]
        sched = schedules[rng.integers(len(schedules))]
        umax = rng.uniform(3, 12)
        ny, nx = shape
        vortex = lamb_oseen(shape, core_radius=rng.uniform(40, 120), umax=umax / 2,
                            shift=(0.0, 0.0),
                            center=(rng.uniform(0, nx), rng.uniform(0, ny)))
        shear = shear_flow(shape, rng, umax / 2)

        def field(x, y, vortex=vortex, shear=shear):
            a = vortex(x, y)
            b = shear(x, y)
            return a[0] + b[0], a[1] + b[1]

        im0, im1 = make_pair(shape, field, rng, density=0.03, dtype=np.uint16)
        out.append((im0, im1, sched))
        fields.append(field)
    if return_fields:
        return out, fields
    return out
