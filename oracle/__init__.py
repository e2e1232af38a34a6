"""CPU oracle: a numpy restatement of fluidimage's PIV correlation path.

TEST INFRASTRUCTURE. Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker (or as the timed CPU baseline). The product package
``fluidimage_amd`` never imports it: its compute path is the HIP library.

Every function cites the reference file:line it restates; paths are relative
to the fluidimage 0.5.3 source tree (``src/fluidimage/...``).

Parity pinning (what fixes each function independently of this code):

* ``correl.like_fftshift``: closed form ``fftshift(arr[::-1, ::-1])`` on the
  reference test shapes (src/fluidimage/calcul/test_correl.py:233-245).
* ``correl.correl_fft``: brute-force FP64 circular cross-correlation of the
  mean-removed windows; Parseval / autocorrelation invariants of
  src/fluidimage/calcul/test_fft.py:8-89; identical windows give
  ``correl_max == 1`` at (iy0, ix0) (test_correl.py:91-92).
* ``correl.energy_from_fourier``: closed form of test_fft.py:66-89.
* ``correl.search_mask``: hand-counted boxes (32 px windows, dmax 8 -> 7..23).
* ``correl.nan_indices_max``: the literal sequential scan
  ``nan_indices_max_loop`` on planted ties / NaNs.
* ``subpix``: 2d_gaussian2 equals the stationary point of a least-squares
  quadratic fitted by ``numpy.linalg.lstsq``; centroid by direct sums.
* ``piv`` grid/positions: hand-computed vector counts and offsets.
* ``fix.mean_neighbors_xy``: ``scipy.ndimage.correlate`` formulation.
* ``fix.work_fix``: hand-built 3 x 3 cases with literal outcomes for every
  branch (strict thresholds on squares, the " + " error composition, the
  secondary-peak ratio, correl_min, tie and NaN rules and the unsquared
  displacement_max quirk of fix.py:192) (tests/test_oracle_pins.py).
* ``correl`` secondary peaks: planted maps with literal peak lists, including
  the empty NaN square of a negative start.
* ``tps``: interpolation at s = 0, affine reproduction for any s,
  partition-of-unity blending of constants; ``compute_tps_weights`` and the
  per-subdomain evaluation against scipy's ``RBFInterpolator``
  (thin_plate_spline kernel, smoothing, degree 1); subdomain counts, limits,
  buffers, rescaled smoothing and half-open membership against
  hand-computed literals (2048 x 2048 at 16 px: 18 x 19; 64,180 vectors:
  17 x 18); the outlier loop's cap (11 solves) and spike replacement.
* ``piv.OracleWorkPIVPass.apply_interp``: the use_tps selector against
  scipy's griddata and ``RBFInterpolator`` on one-subdomain fields; the
  multipass window schedule against literals.

Parity unpinned (no independent check available in this environment): the
exact FFTW float32 bits of the reference maps and the Pythran accumulation
types; both sit below the parity tolerances.
"""
