/*
 * piv_hip.h -- C ABI of the MI355X (gfx950) PIV correlation library
 * (libpiv_hip.so). Plain pointers and sizes only; every buffer argument is a
 * device pointer unless stated otherwise, every call is asynchronous on the
 * caller's HIP stream and returns a status code (PIV_OK == 0). The message
 * of the last failure on the calling thread is piv_last_error().
 *
 * Reference interfaces replaced (paths in the fluidimage 0.5.3 tree):
 *   CorrelFFTW.__init__ / _finalize_init    src/fluidimage/calcul/correl.py:172-200,518-549,643-646
 *   CorrelFFTW.__call__                     src/fluidimage/calcul/correl.py:648-662
 *   compute_displacements_from_correl       src/fluidimage/calcul/correl.py:226-275,561-568
 *   apply_subpix / SubPix.compute_subpix    src/fluidimage/calcul/correl.py:277-281, subpix.py:107-218
 *   BaseWorkPIV._pad_images + _loop_vectors src/fluidimage/works/piv/singlepass.py:196-208,262-361
 *   WorkFIX.calcul                          src/fluidimage/works/piv/fix.py:66-133
 *   mean_neighbors_xy                       src/fluidimage/calcul/mean_neighbors.py:128-155
 *   _calcul_positions_vectors_subimages     src/fluidimage/works/piv/singlepass.py:210-260,787-870
 *
 * Not thread-safe per plan: a plan may be used by one host thread at a time
 * (the reference's FFTW plans have the same restriction, fft.py:218-234).
 */
#ifndef FLUIDIMAGE_AMD_PIV_HIP_H
#define FLUIDIMAGE_AMD_PIV_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define PIV_OK 0
#define PIV_EINVAL 1       /* bad argument */
#define PIV_EUNSUPPORTED 2 /* window shape or option not built into the library */
#define PIV_EHIP 3         /* HIP runtime error */

/* ---- image element types ---- */
#define PIV_U8 0
#define PIV_U16 1
#define PIV_F32 2

/* ---- sub-pixel methods (SubPix.methods, subpix.py:73) ---- */
#define PIV_SUBPIX_2D_GAUSSIAN2 0
#define PIV_SUBPIX_CENTROID 1
#define PIV_SUBPIX_NONE 2
#define PIV_SUBPIX_2D_GAUSSIAN 3 /* least-squares fit on the (2 nsubpix + 1)^2 crop, subpix.py:157-181 */

/* ---- per-window error codes (strings of correl.py:128,130,
 *      subpix.py:154,215, singlepass.py:310) ---- */
#define PIV_WIN_OK 0
#define PIV_WIN_PEAK_BOUNDARY 1 /* "Correlation peak touching boundary." */
#define PIV_WIN_PEAK_NAN 2      /* "Correlation peak touching nan." */
#define PIV_WIN_SUBPIX_BOUNDARY 3 /* "close boundary" */
#define PIV_WIN_SUBPIX_WRONG 4  /* "wrong subpix" */
#define PIV_WIN_BAD_CROP 5      /* "Bad im_crop shape." */
#define PIV_WIN_NO_PEAK 6       /* no finite value in the search box */

/* options of piv_correlate_batch */
#define PIV_OPT_SUBPIX 1     /* sub-pixel fit of the peak (last pass) */
#define PIV_OPT_MAP_SQUARE 2 /* with displacement_max None, the maps keep the
                                second-chance NaN square, as the reference's
                                _loop_vectors maps do (correl.py:234-237) */

/* window flag bits */
#define PIV_WINFLAG_INVALID 1 /* crop outside the padded image: no compute */

/* one image (device memory); rows are `pitch` elements apart */
typedef struct piv_image {
  const void* data;
  int32_t dtype;   /* PIV_U8 / PIV_U16 / PIV_F32 */
  int32_t height;
  int32_t width;
  int32_t reserved;
  int64_t pitch;
  const float* minval; /* device scalar subtracted from in-image pixels
                          (im - im.min(), singlepass.py:206), or NULL */
} piv_image;

/* one interrogation window: top-left corners of the two crops in image
 * coordinates (negative or beyond the image reads as the zero padding) */
typedef struct piv_window {
  int32_t image0, image1; /* indices into the piv_image array */
  int32_t y0, x0;
  int32_t y1, x1;
  int32_t dx_in, dy_in;   /* integer predictor added to the result */
  int32_t flags;          /* PIV_WINFLAG_* */
  int32_t reserved;
} piv_window;

/* per-window outputs (device arrays of length nwin; any may be NULL except
 * dx, dy, correl_max, err). Secondary peaks (nb_peaks_to_search > 1) are
 * stored as [nwin][nb_peaks - 1] triples (dx, dy, correl) in `sec`, their
 * count in `nsec`, and their sub-pixel refinement (dx, dy, NaN on error)
 * in `sec_sub`. */
typedef struct piv_results {
  float* dx;
  float* dy;
  float* correl_max;
  int32_t* err;
  int32_t* ix;
  int32_t* iy;
  double* norm;
  float* sec;
  int32_t* nsec;
  float* sec_sub;
} piv_results;

typedef struct piv_plan piv_plan;

/* Version string of the library. */
const char* piv_version(void);
/* Message of the last failed call on this thread ("" if none). */
const char* piv_last_error(void);
/* 1 if the library correlates windows of n0 x n1 (rows x cols): the shapes
 * of its register-FFT and LDS-staged kernels, and any other shape of 2 to
 * 256 rows and columns and at most 16384 elements (the DFT kernel of any
 * shape; the reference takes any shape through FFTW, fft.py:205-244). */
int piv_shape_supported(int32_t n0, int32_t n1);

/* ---- launch tuning (no counterpart in the reference, whose FFTW plans
 *      are chosen by FFTW_MEASURE, fft.py:221-234) ----
 * Tunable launch parameters: they change how work is mapped onto the GPU,
 * not the results (but for PIV_TUNE_TPS_SHARED_SOLVE, in rounding).
 * PIV_TUNE_CORRELATE_WPB: windows per workgroup of the fused correlation
 *   kernel of window n0 x n1 (the candidates are the variants built);
 * PIV_TUNE_TPS_CHOL_THREADS: threads per subdomain of the TPS Cholesky
 *   for launches of more than two subdomains per compute unit, 256 or 512
 *   (0: 256); smaller launches take 512 up to one subdomain per CU and 256
 *   up to two, so that they run in one round (n0, n1 ignored);
 * PIV_TUNE_TPS_SHARED_SOLVE: how the subdomains that share a pattern's
 *   factorisation are solved: 0 with the explicit inverse of the pattern's
 *   saddle-point matrix (built once per pattern, then one product per
 *   outlier iteration for up to four subdomains at once), 1 by blocked
 *   triangular substitution with the pattern's Cholesky factor (n0, n1
 *   ignored). The two differ in rounding only, the one tunable whose
 *   results are not bit-identical.
 * piv_tune_candidates writes at most cap values and returns their number
 * (negative status on error); piv_tune_get returns the value in use
 * (negative status on error); piv_tune_set returns a status. The setting is
 * process-wide and read at each launch. */
#define PIV_TUNE_CORRELATE_WPB 1
#define PIV_TUNE_TPS_CHOL_THREADS 2
#define PIV_TUNE_TPS_SHARED_SOLVE 3
int piv_tune_candidates(int32_t param, int32_t n0, int32_t n1, int32_t* values, int32_t cap);
int piv_tune_get(int32_t param, int32_t n0, int32_t n1);
int piv_tune_set(int32_t param, int32_t n0, int32_t n1, int32_t value);

/* Create a correlation plan for n0 x n1 windows (CorrelFFTW.__init__).
 * displacement_max < 0 means None. `mask` (host, n0*n1 bytes, may be NULL
 * when displacement_max < 0) is where_large_displacement, `box` (host, 4
 * ints) the search box (i0_start, i0_stop, i1_start, i1_stop). */
int piv_plan_create(piv_plan** plan, int32_t n0, int32_t n1,
                    double displacement_max, const uint8_t* mask,
                    const int32_t* box, int32_t subpix_method,
                    int32_t nsubpix, int32_t particle_radius,
                    int32_t nb_peaks_to_search);
int piv_plan_destroy(piv_plan* plan);

/* out_min[i] = min over image i (device float per image). `images` is a
 * device array of nimages descriptors. flat_dtype: the images' dtype when
 * every image is contiguous (pitch == width) with 16-byte aligned data (a
 * vectorised pass over the flat images), else -1 (row by row). */
int piv_images_min(const piv_image* images, int32_t nimages, float* out_min,
                   int32_t flat_dtype, void* stream);

/* Fused crop -> mean removal -> FFT cross-correlation -> normalisation ->
 * masked peak search (second chance, secondary peaks) -> optional sub-pixel
 * fit, for nwin windows. `images` and `windows` are device arrays; every
 * image of the call has element type `dtype`; `options` is a set of
 * PIV_OPT_* bits. `maps`
 * (device, nwin*n0*n1 float32, may be NULL) receives the correlation maps in
 * the reference layout (after _like_fftshift). */
int piv_correlate_batch(piv_plan* plan, int32_t dtype, const piv_image* images,
                        const piv_window* windows, int32_t nwin,
                        int32_t options, float* maps,
                        const piv_results* res, void* stream);

/* Peak search (+ optional sub-pixel fit) on existing maps (nwin*n0*n1,
 * reference layout) with their norms. With displacement_max None the
 * second-chance NaN square is written into `maps`, as the reference does. */
int piv_peaks_from_maps(piv_plan* plan, float* maps, const double* norms,
                        int32_t nwin, int32_t apply_subpix,
                        const piv_results* res, void* stream);

/* Sub-pixel position at integer indices (ix, iy) of each map:
 * out_x/out_y = ix + deplx / iy + deply, err = PIV_WIN_* . */
int piv_subpix_at(piv_plan* plan, const float* maps, const int32_t* ix,
                  const int32_t* iy, int32_t nwin, float* out_x, float* out_y,
                  int32_t* err, void* stream);

/* ---- direct (spatial) correlation, replacing CorrelPythran / CorrelPyCuda
 *      (correl.py:284-451, correl_numpy; correl_pycuda.py:26-240) ----
 * nwin pairs of windows im0 (nwin x n0y x n0x float32) and im1
 * (nwin x n1y x n1x, no larger than im0); maps (nwin x (2 dmax + 1)^2,
 * element [dmax + dy, dmax + dx]) hold the sum of im1[y, x] im0[y + oy, x + ox]
 * over the reference's overlap for the displacement (dy, dx), and norms
 * the norm the reference returns with it. normalisation
 * PIV_DIRECT_PYTHRAN: sum / overlap area * size(im1), norm
 * sqrt(sum im1^2 sum im0^2) (correl_numpy); PIV_DIRECT_PYCUDA: the plain
 * sum, norm sqrt(sum im1^2 sum im0^2) size(im0) (correl_pycuda). Sums in
 * FP64. At most 16384 window pixels in all and 2 dmax + 1 <= 127; a dmax
 * whose overlap would be empty or leave a window is rejected. */
#define PIV_DIRECT_PYTHRAN 0
#define PIV_DIRECT_PYCUDA 1
int piv_correlate_direct(int32_t nwin, const float* im0, int32_t n0y, int32_t n0x,
                         const float* im1, int32_t n1y, int32_t n1x,
                         int32_t displacement_max, int32_t normalisation, float* maps,
                         double* norms, void* stream);

/* Peak search with the second chance, secondary peaks and (optional)
 * sub-pixel fit on nwin maps of any shape n0 x n1 (3..127) whose zero
 * displacement is at (iy0, ix0): CorrelBase.compute_displacements_from_correl
 * and apply_subpix (correl.py:226-281) for the maps of the direct
 * correlation ((dmax, dmax), correl.py:409-410,444-445). There is no search
 * mask: the second-chance NaN square is written into `maps`, as the
 * reference writes it into the caller's array. */
int piv_peaks_any(int32_t nwin, float* maps, const double* norms, int32_t n0,
                  int32_t n1, int32_t iy0, int32_t ix0, int32_t subpix_method,
                  int32_t nsubpix, int32_t particle_radius, int32_t nb_peaks,
                  int32_t apply_subpix, const piv_results* res, void* stream);

/* Sub-pixel position at (ix, iy) of each n0 x n1 map, as piv_subpix_at. */
int piv_subpix_any(int32_t nwin, const float* maps, int32_t n0, int32_t n1,
                   const int32_t* ix, const int32_t* iy, int32_t subpix_method,
                   int32_t nsubpix, float* out_x, float* out_y, int32_t* err,
                   void* stream);

/* Window records of one pass on a ny x nx vector grid at (iyvecs[i],
 * ixvecs[j]) (BaseWorkPIV._prepare_with_image grid, singlepass.py:108-141).
 * First pass: dx_approx == dy_approx == NULL (singlepass.py:210-260). Later
 * passes: the float64 predictor on the grid, rounded half-to-even
 * (singlepass.py:752-753) and split between the two crops with the
 * reference's out-of-image fallbacks (singlepass.py:787-870). `crop` (host,
 * 8 ints): _start_for_crop0 (y, x), _stop_for_crop0 (y, x),
 * _start_for_crop1 (y, x), _stop_for_crop1 (y, x). Windows whose crop leaves
 * the padded image get PIV_WINFLAG_INVALID. xs/ys (float64 vector
 * positions) and dx_in/dy_in (int32 rounded predictor) may be NULL.
 * npairs pairs of a batch in one call: pair k reads images image0 + 2k,
 * image1 + 2k and its ny nx vectors at offset k ny nx of the predictor and
 * output arrays. */
int piv_make_windows(int32_t npairs, int32_t ny, int32_t nx, const int32_t* iyvecs,
                     const int32_t* ixvecs, const double* dx_approx,
                     const double* dy_approx, int32_t height, int32_t width,
                     int32_t npad, const int32_t* crop, int32_t image0,
                     int32_t image1, piv_window* windows, double* xs,
                     double* ys, int32_t* dx_in, int32_t* dy_in, void* stream);

/* WorkFIX on a (ny, nx) vector grid (fix.py:66-133): NaN the vectors with
 * err != 0, correl_max < correl_min, |d|^2 > displacement_max^2 (if > 0),
 * then the 8-neighbour mean test against threshold (if >= 0). flags receives
 * bit 0: piv error, 1: correl < correl_min, 2: delta2 < displacement_max2,
 * 3: diff neighbour too large; diff2 the squared difference to the
 * neighbour mean; mean_x / mean_y the neighbour means (may be NULL).
 * npairs grids of a batch in one call, pair k at offset k ny nx of every
 * array. */
int piv_fix(int32_t npairs, float* dx, float* dy, const float* correl_max, const int32_t* err,
            int32_t ny, int32_t nx, float correl_min, float displacement_max,
            float threshold, int32_t* flags, float* diff2, float* mean_x,
            float* mean_y, void* stream);

/* Optional lattice of the TPS positions (NULL, or table NULL: none): a
 * table written by piv_tps_phi_table with this step and a_max. Point
 * differences that are exact multiples (a step, b step) with a, b <= a_max
 * read the kernel function from the table, every other difference is
 * evaluated directly; both give the same value. */
typedef struct piv_tps_lattice {
  const double* table;
  double step;
  int32_t a_max;
  int32_t reserved;
} piv_tps_lattice;

/* ---- thin-plate-spline subdomains (FP64), replacing
 *      ThinPlateSplineSubdom.compute_tps_weights_subdom / interpolate
 *      (thin_plate_spline_subdom.py:128-159,224-280) and
 *      compute_tps_weights (thin_plate_spline.py:168-216) ----
 * Subdomain d holds the slots [dom_off[d], dom_off[d+1]) (at most 2048);
 * slot s is center slot_pt[s] of xs/ys (float64) with values vx/vy
 * (float32). mats: workspace of sum_d piv_tps_matrix_size(n_d) doubles,
 * subdomain d at mat_off[d] (offsets multiples of 16 doubles, mats aligned
 * to 128 bytes); refl:
 * 16 doubles per subdomain; vwork: 6 doubles per slot; slot_xy receives the
 * center coordinates by slot (x[nslots] then y[nslots]). Outputs: the
 * smoothed value of every slot (smooth_slots[2][nslots]), the weights of
 * every subdomain (weights[2][wstride], n_d + 3 values at
 * dom_off[d] + 3 d: the n_d center weights then the [1, y, x] terms),
 * per-subdomain stats[6] (x then y: fixed vectors, max |diff| of the first
 * solve, solves) and status (non-zero: singular system). `smoothing` is
 * the device array of the (rescaled) smoothing coefficient of every
 * subdomain, so that the subdomains of several vector fields, each with its
 * own grid, go through one call. threshold < 0 disables the outlier
 * iterations. order (optional, nd entries): a permutation of the
 * subdomains giving the order in which the per-subdomain work is launched
 * (largest first shortens the tail of a launch; results do not depend on
 * it); NULL: 0..nd-1. lattice: see piv_tps_lattice (NULL: direct
 * evaluation).
 * generated (with a lattice): every center is a multiple of the lattice
 * step (info[f][6] of piv_tps_select for every field) and the table covers
 * the differences within a subdomain; the matrix build then stores only
 * B[:, :3] and the factorisation generates the other elements from the
 * table where it first reads them (the same values; mats still needs its
 * full size for the factor). 0: the build stores the lower triangle. */
/* Doubles of the matrix region of a subdomain of n centers (no reference
 * counterpart: the reference's LAPACK call takes a dense n x n array):
 * B'[3:, :3] as (n-3) x 3, padded to a multiple of 16, then the (n-3) x
 * (n-3) matrix that is factorised in place, rows padded to a multiple of 16
 * doubles. -1 for n outside [0, 2048]. */
int64_t piv_tps_matrix_size(int32_t n);

int piv_tps_smooth(int32_t nd, int32_t n_max, const double* xs,
                   const double* ys, const float* vx, const float* vy,
                   const int64_t* dom_off, const int32_t* slot_pt,
                   const int64_t* mat_off, const double* smoothing, double threshold,
                   double* mats, double* refl, double* vwork, double* slot_xy,
                   double* smooth_slots, int64_t nslots, double* weights,
                   int64_t wstride, double* stats, int32_t* status,
                   const int32_t* order, const piv_tps_lattice* lattice,
                   int32_t generated, void* stream);

/* One vector field (one image pair) of a TPS batch: its subdomain grid
 * (nsx x nsy, limits at lim + lim_off: lmin_x[nsx], lmax_x[nsx],
 * lmin_y[nsy], lmax_y[nsy]), its number of selected centers, the position
 * of its first center in the batch's center arrays and the index of its
 * first subdomain. A device array of these records describes a batch. */
typedef struct piv_tps_field {
  int32_t nsx, nsy;
  int32_t count;
  int32_t center_base;
  int32_t dom_base;
  int32_t lim_off;
  int32_t sorted_y; /* nonzero: the centers are sorted by y (info[f][5] of
                       piv_tps_select); membership then scans only the
                       centers within a subdomain's y limits */
} piv_tps_field;

/* piv_tps_select: for each of nfields fields of n vectors (elements
 * [f n, (f + 1) n) of dx, dy, xs, ys), ordered compaction of the vectors
 * whose dx and dy are finite (singlepass.py:450-463) into cx, cy
 * (positions), vx, vy (values) and idx (index in the field), starting at
 * element f n; info[f][7] = count, x_min, x_max, y_min, y_max, 1.0
 * when ys is nondecreasing over the field (0.0 otherwise), and 1.0 when
 * lattice_step > 0 and every selected position is a multiple of it.
 * piv_tps_count / piv_tps_fill: members of every subdomain d of the batch
 * (field dom_field[d], local index d - dom_base = iy * nsx + ix) with the
 * half-open limits of its field, in increasing center order
 * (thin_plate_spline_subdom.py:98-116); fill writes the batch center
 * indices at dom_off[d] (prefix sum of the counts) into slot_pt.
 * piv_tps_point_slots: for each point, its (at most 4) subdomains in
 * increasing order, their partition-of-unity weights and the sum of the
 * weights (thin_plate_spline_subdom.py:118-126,187-196,208-222). With
 * dom_off/slot_pt (centers): the points are the centers of each field
 * (px = cx), outputs at the center's batch index, the slot being its
 * position in the member list of the subdomain; the entries of a field from
 * its count up to max_points (at most the distance between the fields'
 * first centers) are written empty (slots -1, weights 0, norm 1); without:
 * the max_points
 * new positions (px, py) shared by every field, outputs at
 * f * max_points + p, the slot being the batch subdomain index.
 * status[0] is set when a point lies in more than 4 subdomains. */
int piv_tps_select(int32_t nfields, int32_t n, const float* dx, const float* dy,
                   const double* xs, const double* ys, double* cx, double* cy,
                   float* vx, float* vy, int32_t* idx, double* info,
                   double lattice_step, void* stream);
int piv_tps_count(int32_t nd, const piv_tps_field* fields,
                  const int32_t* dom_field, const double* lim, const double* cx,
                  const double* cy, int32_t* counts, void* stream);
int piv_tps_fill(int32_t nd, const piv_tps_field* fields,
                 const int32_t* dom_field, const double* lim, const double* cx,
                 const double* cy, const int64_t* dom_off, int32_t* slot_pt,
                 void* stream);
int piv_tps_point_slots(int32_t nfields, const piv_tps_field* fields,
                        int32_t max_points, const double* lim, const double* px,
                        const double* py, const int64_t* dom_off,
                        const int32_t* slot_pt, int32_t* out_slot,
                        double* out_coef, double* out_norm, int32_t* status,
                        void* stream);

/* Bytes of the workspace of piv_tps_smooth_shared. */
int64_t piv_tps_shared_workspace(int32_t nfields, int32_t dom_cap, int32_t npat_cap,
                                 int32_t n_max, int32_t cap_r);

/* piv_tps_smooth for a device-sized batch (see piv_tps_grid below) whose
 * centers are nodes of the pass grid gx[ngx] x gy[ngy] (the vectors'
 * positions; irr_x / irr_y: the grid's last column / row is appended at a
 * smaller step). No counterpart in the reference, which factorises every
 * subdomain (thin_plate_spline.py:214): subdomains whose boxes hold the
 * same grid-node pattern share one factorisation per field (the matrix of a
 * box is the same up to a translation, with the polynomial part relative
 * to the box's first node), and a subdomain missing at most cap_r (<= 32;
 * <= 12 when PIV_TUNE_TPS_SHARED_SOLVE selects the substitution) nodes of
 * its box (its invalid vectors) is solved with it and a capacitance
 * correction; the others are factorised alone as in piv_tps_smooth. Same outputs as piv_tps_smooth (weights in image
 * coordinates); status[d] = 2 when a center is off the grid. At most
 * npat_cap (<= 64) patterns per field. `order`: the order of the nfields
 * dom_cap subdomain slots (piv_tps_offsets); idx: the vector index of
 * every center (piv_tps_select); work: piv_tps_shared_workspace bytes.
 * `parts`: the PIV_TPS_PART_* steps to launch: PREP (patterns and orders)
 * first, then SHARED (the patterns' factorisations and the shared solves)
 * and DIRECT (the subdomains factorised alone), which touch disjoint
 * subdomains and may run on two streams after PREP (the caller orders
 * them); PIV_TPS_PART_ALL launches the three in order on `stream`. */
#define PIV_TPS_PART_PREP 1
#define PIV_TPS_PART_SHARED 2
#define PIV_TPS_PART_DIRECT 4
#define PIV_TPS_PART_ALL 7
int piv_tps_smooth_shared(int32_t nfields, int32_t dom_cap, const piv_tps_field* fields,
                          const double* lim, const int32_t* fstat, const int32_t* dom_field,
                          const int32_t* counts, const double* gx, int32_t ngx, int32_t irr_x,
                          const double* gy, int32_t ngy, int32_t irr_y, int32_t npat_cap,
                          int32_t cap_r, int32_t n_max, const double* xs, const double* ys,
                          const int32_t* idx, const float* vx, const float* vy,
                          const int64_t* dom_off, const int32_t* slot_pt, const int64_t* mat_off,
                          const double* smoothing, double threshold, double* mats, double* refl,
                          double* vwork, double* slot_xy, double* smooth_slots, int64_t nslots,
                          double* weights, int64_t wstride, double* stats, int32_t* status,
                          const int32_t* order, const piv_tps_lattice* lattice,
                          int32_t generated, int32_t parts, void* work, void* stream);

/* ---- device-sized TPS batches (no counterpart in the reference, whose
 *      ThinPlateSplineSubdom.__init__ sizes each field on the host,
 *      thin_plate_spline_subdom.py:53-126): the subdomain grids, member
 *      offsets and launch order of a batch computed on the GPU from
 *      capacities fixed beforehand, so that a whole multipass can run
 *      without a device-to-host read and be captured in one graph ----
 * Field status (fstat[f]): */
#define PIV_TPS_FIELD_OK 0
#define PIV_TPS_FIELD_NO_VECTOR 1       /* no finite vector ("Only nan.") */
#define PIV_TPS_FIELD_DEGENERATE 2      /* centers on one line (zero range) */
#define PIV_TPS_FIELD_EMPTY_SUBDOMAIN 3 /* a subdomain without center */
#define PIV_TPS_FIELD_CAPACITY 4        /* beyond the batch's capacities */
#define PIV_TPS_FIELD_TOO_MANY 5        /* more than 2048 centers in a subdomain */
#define PIV_TPS_FIELD_OFF_LATTICE 6     /* generated matrices asked for, but a center or a
                                           subdomain is off the kernel-function table */
/* piv_tps_grid: from info[f][7] of piv_tps_select (count, bounds, sorted,
 * on-lattice), for each of nfields fields of n vectors: its subdomain grid
 * as thin_plate_spline_subdom.py:53-90 computes it (counts floor/ceil of
 * sqrt, np.linspace limits rounded as numpy rounds them, buffers
 * coef_buffer * range / count, smoothing rescaled by
 * sqrt(area / vectors per subdomain)), laid out in dom_cap subdomain slots
 * per field: fields[f] (dom_base f dom_cap, lim_off 4 f dom_cap, center_base
 * f n), lim (4 dom_cap doubles per field), dom_field (nfields dom_cap;
 * -1 beyond the field's grid), smoothing (per subdomain slot) and fstat.
 * gen_width > 0: a subdomain wider than that (the reach of the lattice
 * table) fails its field with PIV_TPS_FIELD_OFF_LATTICE, as does a field
 * not on the lattice when require_lattice. At most 2048 fields.
 * piv_tps_offsets: after piv_tps_count over the nfields dom_cap slots,
 * fails the fields with an empty subdomain (first), more than 2048 or more
 * than n_max_cap centers in one, admits the others in order while their
 * slots and matrix doubles fit slots_cap and mats_cap (else
 * PIV_TPS_FIELD_CAPACITY), then writes dom_off and mat_off (nfields dom_cap
 * + 1 each, the slots of failed fields empty), order (largest subdomain
 * first) and totals[4] (slots, matrix doubles, largest subdomain,
 * subdomains with centers). A failed field's record loses its grid, so its
 * points get no slot (piv_tps_point_slots writes them empty with norm 1)
 * and its blended and interpolated values are 0. */
int piv_tps_grid(int32_t nfields, int32_t n, const double* info, double subdom_size,
                 double smoothing_coef, double coef_buffer, int32_t dom_cap, double gen_width,
                 int32_t require_lattice, piv_tps_field* fields, double* lim,
                 int32_t* dom_field, double* smoothing, int32_t* fstat, void* stream);
int piv_tps_offsets(int32_t nfields, int32_t dom_cap, piv_tps_field* fields,
                    const int32_t* counts, const int32_t* dom_field, int64_t slots_cap,
                    int64_t mats_cap, int32_t n_max_cap, int32_t* fstat, int64_t* dom_off,
                    int64_t* mat_off, int32_t* order, int64_t* totals, void* stream);
/* The evaluation at the new positions that are a field's own centers (new
 * positions shared by the nfields fields of n vectors, new_x[q] against
 * xs[f n + q]): out[c][f n + q] receives the blended smoothed value of the
 * vector (smoothed[c][f n + rank], the rank of vector q among the field's
 * selected vectors idx) and skip[f n + q] = 1, so that piv_tps_interpolate
 * leaves it (it equals the evaluation: [K P][w; a] = v - s w at the
 * centers). out and skip start at zero; info as piv_tps_select writes it. */
int piv_tps_at_centers(int32_t nfields, int32_t n, const double* info, const int32_t* idx,
                       const double* xs, const double* ys, const double* new_x,
                       const double* new_y, const double* smoothed, double* out,
                       uint8_t* skip, void* stream);

/* out[c][i] = sum_k coef[4 i + k] * smooth_slots[c][slots[4 i + k]] /
 * norm[i] over the point's slots (-1 terminated): the blend of
 * thin_plate_spline_subdom.py:145-158. out[2][npts]. */
int piv_tps_blend(int32_t npts, const int32_t* slots, const double* coef,
                  const double* norm, const double* smooth_slots,
                  int64_t nslots, double* out, void* stream);

/* Differentiable TPS node field (no counterpart in the reference, which
 * does not differentiate; the map is its smoothing without the outlier
 * iterations, thin_plate_spline.py:168-216 per subdomain and the blend of
 * thin_plate_spline_subdom.py:145-158). The influence matrix of subdomain d,
 * Sw = (M^-1)[:n, :n] of its factorised saddle-point matrix (from
 * piv_tps_smooth: dom_off, mat_off, mats, refl, vwork, status), row-major
 * n x n at infl + d * stride, stride >= n_max^2 doubles. */
int piv_tps_influence(int32_t nd, int32_t n_max, const int64_t* dom_off,
                      const int64_t* mat_off, const double* mats, const double* refl,
                      const double* vwork, const int32_t* status, double* infl,
                      int64_t stride, void* stream);

/* The smoothed field at the nc centers of one TPS structure (nd subdomains,
 * members slot_pt, used slots nused, smoothing per subdomain, influence
 * matrices of piv_tps_influence, the centers' slots / blend coefficients /
 * norms c_slot, c_coef, c_norm of piv_tps_point_slots) as a linear map of
 * the values, for nfields fields at once: transpose 0: out = T in (in: the
 * values, out: the smoothed values, both [2][nfields nc]); transpose 1:
 * out = T^T in (the gradient of the values from that of the smoothed
 * values). work: 4 nfields nslots doubles. Every status must be 0. */
int piv_tps_node_field(int32_t transpose, int32_t nfields, int32_t nc, int32_t nd,
                       int32_t n_max, const int64_t* dom_off, const int32_t* slot_pt,
                       int64_t nused, const int32_t* status, const double* smoothing,
                       const double* infl, int64_t stride, const int32_t* c_slot,
                       const double* c_coef, const double* c_norm, int64_t nslots,
                       const double* in, double* out, double* work, void* stream);

/* TPS evaluation at npts new positions (px, py): for each subdomain
 * doms[4 p + k] (-1 terminated) of the point, the weighted sum of
 * w_j r^2 ln(r^2) / 2 over its centers plus the linear part, blended with
 * coef[4 p + k] and divided by norm[p] (thin_plate_spline_subdom.py:224-234);
 * slot_xy: the center coordinates by slot written by piv_tps_smooth; nd
 * subdomains (dom_off[nd + 1]), none with more than n_max (<= 2048)
 * centers. out[2][npts]. skip (optional, one byte per position): positions
 * with a nonzero byte are not evaluated and keep what out holds (the caller
 * puts there the blended smoothed value of a position that is a center,
 * which the evaluation equals: [K P][w; a] = v - s w at the centers).
 * work: device workspace of piv_tps_interpolate_workspace(npts, nd) bytes
 * (the points are sorted by subdomain there, one workgroup evaluating the
 * points of one subdomain). lattice: see piv_tps_lattice. */
int piv_tps_interpolate(int32_t npts, const double* px, const double* py,
                        const int32_t* doms, const double* coef,
                        const double* norm, const double* slot_xy, int64_t nslots,
                        int32_t nd, int32_t n_max, const int64_t* dom_off,
                        const double* weights, int64_t wstride, const uint8_t* skip,
                        double* out, void* work, const piv_tps_lattice* lattice,
                        void* stream);
int64_t piv_tps_interpolate_workspace(int32_t npts, int32_t nd);

/* piv_tps_interpolate without the workspace argument, for callers that
 * cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc). No counterpart in the reference.
 * It allocates nothing: its scratch (the point lists and partial sums) is a
 * fixed device array of PIV_TPS_NOALLOC_SCRATCH_BYTES per GPU, held until the
 * process ends; a call whose piv_tps_interpolate_workspace(npts, nd) exceeds
 * it fails with PIV_EUNSUPPORTED. Every call on one GPU shares that array:
 * calls that overlap on one GPU (concurrent streams, threads) race. Same
 * arguments and results as piv_tps_interpolate otherwise. */
#define PIV_TPS_NOALLOC_SCRATCH_BYTES (64LL << 20)
int piv_tps_interpolate_noalloc(int32_t npts, const double* px, const double* py,
                                const int32_t* doms, const double* coef, const double* norm,
                                const double* slot_xy, int64_t nslots, int32_t nd,
                                int32_t n_max, const int64_t* dom_off, const double* weights,
                                int64_t wstride, const uint8_t* skip, double* out,
                                const piv_tps_lattice* lattice, void* stream);

/* Table of the kernel function r^2 ln(r^2) / 2 on a lattice of step
 * `step`: table[(a_max + 1) b + a] at r^2 = (a step)^2 + (b step)^2,
 * 0 <= a, b <= a_max (at most 4095), (a_max + 1)^2 doubles. No counterpart
 * in the reference, which evaluates the logarithm for every pair of points
 * (thin_plate_spline.py:81-94); here it serves piv_tps_smooth and
 * piv_tps_interpolate when the positions lie on a lattice (the PIV grids
 * of the passes), with results equal to the direct evaluation. */
int piv_tps_phi_table(double step, int32_t a_max, double* table, void* stream);

/* ---- image deformation between passes (no counterpart in the reference,
 *      whose later passes shift the windows by the rounded predictor,
 *      singlepass.py:752-753,822-825) ----
 * Pass grid: nodes iyvecs[ny] x ixvecs[nx] (row-major vectors); for every
 * image row y and column x, the cell of the predictor's bilinear
 * interpolation and the fraction in it (cell_y[y], t_y[y], cell_x[x],
 * t_x[x]; clamped to the grid's bounding box). Device pointers. */
typedef struct piv_deform_grid {
  const int32_t* iyvecs;
  const int32_t* ixvecs;
  int32_t ny, nx;
  const int32_t* cell_y;
  const double* t_y;
  const int32_t* cell_x;
  const double* t_x;
} piv_deform_grid;

/* Deformed images of npairs pairs (images 2k, 2k + 1 of `images`, all
 * height x width of element type dtype, minima subtracted):
 * im0(y - v/2, x - u/2) and im1(y + v/2, x + u/2) at every pixel (y, x) by
 * Keys cubic convolution (a = -0.5), zero outside the image, with (u, v) the
 * bilinear interpolation at the pixel of the predictor (device,
 * npairs x ny x nx float64) on the grid. `out` (device float32,
 * 2 npairs x height x width) receives them in image order; the deformed
 * window of a grid node is the crop of these images around the node.
 * Weights and tap sums are float32, positions FP64. `grid` is a host
 * pointer to a record of device pointers; grid->nx is at most 4096. */
int piv_deform_images(int32_t dtype, const piv_image* images, int32_t npairs,
                      int32_t height, int32_t width, const piv_deform_grid* grid,
                      const double* u, const double* v, float* out,
                      void* stream);

/* dx[i] += u[i], dy[i] += v[i] (float32 results, float64 predictor): the
 * residual of a deformed pass plus the predictor at the window centre. */
int piv_add_predictor(int32_t n, float* dx, float* dy, const double* u,
                      const double* v, void* stream);

/* ---- differentiable refinement of the final field (no counterpart in the
 *      reference, which stops after the last pass, multipass.py:185-188) ----
 * Parameters (dx, dy): displacement at every node of `grid` (device,
 * npairs x ny x nx float64). Field at a pixel: bilinear interpolation of
 * the nodes. Window w of a pair: n0 x n1 pixels centred on node w, samples
 * a = im0(p - d/2), b = im1(p + d/2) by Keys cubic convolution (images 2k,
 * 2k + 1, minima subtracted, zero outside). loss[k] += sum over the windows
 * of pair k of 1 - NCC(a, b) (caller zeroes loss); stats (nwin x 5 float64)
 * keeps the window sums for the backward pass, which adds
 * grad_out[k] * d loss[k] / d (dx, dy) into (gdx, gdy), recomputing the
 * samples window by window. */
int piv_refine_forward(int32_t dtype, const piv_image* images, int32_t npairs,
                       const piv_deform_grid* grid, int32_t n0, int32_t n1,
                       const double* dx, const double* dy, double* stats,
                       double* loss, void* stream);
int piv_refine_backward(int32_t dtype, const piv_image* images, int32_t npairs,
                        const piv_deform_grid* grid, int32_t n0, int32_t n1,
                        const double* dx, const double* dy, const double* stats,
                        const double* grad_out, double* gdx, double* gdy,
                        void* stream);

/* ---- sliding-window preprocessing, replacing the python toolbox's
 *      sliding_minima / sliding_median / sliding_percentile
 *      (src/fluidimage/works/preproc/_toolbox_py.py:66-165) ---- */
/* scipy.ndimage boundary modes (boundary_condition) */
#define PIV_BOUNDARY_REFLECT 0 /* d c b a | a b c d | d c b a */
#define PIV_BOUNDARY_CONSTANT 1 /* cval outside (cast to the image type) */
#define PIV_BOUNDARY_NEAREST 2 /* a a a | a b c d | d d d */
#define PIV_BOUNDARY_MIRROR 3 /* d c b | a b c d | c b a */
#define PIV_BOUNDARY_WRAP 4   /* b c d | a b c d | a b c */
/* out[i] = max(img_i - weight * F(img_i), 0) for the `nimages` images of
 * the device table `images` (all height x width, element type `dtype`; the
 * minval fields are not used), F the rank-th smallest value of the
 * wy x wx box around each pixel (offsets -(w / 2) .. w - 1 - w / 2, as
 * scipy.ndimage with origin 0): rank 0 is minimum_filter, rank
 * wy * wx / 2 median_filter, int(wy * wx * percentile / 100)
 * percentile_filter. `out` is nimages x height x width, float64 for uint8 and
 * uint16 images and float32 for float32 images, the types and roundings of
 * the reference's numpy expression. Windows of 1 to 128 rows and columns. */
int piv_sliding_filter(const piv_image* images, int32_t nimages, int32_t height,
                       int32_t width, int32_t dtype, int32_t wy, int32_t wx, int32_t rank,
                       int32_t mode, double cval, double weight, void* out, void* stream);

/* ---- capture and replay as one HIP graph (no reference counterpart: the
 *      reference's per-window loop, src/fluidimage/works/piv/singlepass.py:
 *      262-361, pays its host cost per window; a replayed graph pays the
 *      launch cost of a whole batch once) ----
 * The calls issued on `stream` (not NULL) by the capturing thread between
 * piv_graph_begin and piv_graph_end are recorded, not run. Every compute
 * entry point can be captured (piv_images_min, piv_make_windows,
 * piv_correlate_batch, piv_peaks_from_maps, piv_subpix_at, piv_fix,
 * piv_correlate_direct, piv_peaks_any, piv_subpix_any, piv_deform_images,
 * piv_add_predictor, piv_refine_*, piv_tps_*): each only launches kernels.
 * piv_plan_create / piv_plan_destroy allocate and free: call them outside
 * the capture, as every allocation of the buffers a graph uses. A replay
 * (piv_graph_launch, on any stream) runs the recorded launches with the
 * recorded pointers and sizes: change buffer contents between replays, not
 * addresses or shapes; buffers must outlive the graph. A capture whose
 * calls failed still has to be ended; piv_graph_end then fails and returns
 * no graph. */
typedef struct piv_graph piv_graph;
int piv_graph_begin(void* stream);
int piv_graph_end(void* stream, piv_graph** graph);
int piv_graph_launch(piv_graph* graph, void* stream);
int piv_graph_destroy(piv_graph* graph);

#ifdef __cplusplus
}
#endif

#endif /* FLUIDIMAGE_AMD_PIV_HIP_H */
