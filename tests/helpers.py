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
