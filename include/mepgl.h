/*
 * mepgl.h - C ABI of the MI355X gauged-string-method relaxation library
 * (libmepgl.so, HIP kernels for gfx950).
 *
 * The library replaces the solver layer of the reference (mepgl/src/nlcg.cu,
 * interpolation.cu, parallel_sum.cu) and its relaxation loop (loop() in
 * mepgl/src/mepgl.cu:38-391).  The reference passes raw device pointers and
 * compile-time sizes between host functions; here an opaque context owns all
 * device memory for one or many strings, sizes are runtime values, and every
 * entry point returns a status code instead of exiting.
 *
 * Host arrays use the reference's device layout (io.cu:23-30): per-site
 * planes are N*N values with x fastest ([j][i]), fields are F*N*N pairs
 * (real, imaginary / x, y) with x fastest then y then image, i.e. the
 * layout of a C-contiguous numpy array of shape (F, N, N, 2).  Host arrays
 * are borrowed for the duration of a call only.
 *
 * Threading: a context is driven by one host thread; mepgl_set_control may
 * be called from any thread (the stdin thread of the CLI).
 */
#ifndef MEPGL_H
#define MEPGL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  0 is success; mepgl_last_error() describes failures. */
#define MEPGL_OK 0
#define MEPGL_ERR_ARG (-1)       /* invalid argument or size            */
#define MEPGL_ERR_HIP (-2)       /* HIP runtime error                   */
#define MEPGL_ERR_NONFINITE (-3) /* NaN/Inf in a per-image energy or |g|^2 */
#define MEPGL_ERR_STATE (-4)     /* call not valid in the current state */
#define MEPGL_ERR_NOMEM (-5)     /* device allocation failed            */

/* Precision policies.  FP64: every stored field and every operation in
 * double (DOUBLE_PRECISION build of the reference, real.cuh:7-14).
 * MIXED: fields, directions and gradients stored and updated in float; all
 * reductions, NLCG scalars, line search, arclength and reparametrisation in
 * double (the reductions with two-sum compensation). */
#define MEPGL_FP64 0
#define MEPGL_MIXED 1

/* Flags for mepgl_create. */
#define MEPGL_POLY_EXACT 1u     /* exact line-search quartic instead of the
                                   reference's (nlcg.cu:1162-1165, 1251)   */
#define MEPGL_CUBIC_TEXTBOOK 2u /* (1-w)^3-(1-w) spline weight instead of the
                                   reference's (interpolation.cu:394-400)  */
#define MEPGL_NO_GRAPH 4u       /* mepgl_loop launches eagerly instead of
                                   replaying a captured hipGraph           */
#define MEPGL_GAUGE_ALIGN 8u    /* modes L and C align the constant phase of
                                   neighbouring images before the arclength
                                   (mepgl_gauge_align); off = reference    */
#define MEPGL_GAUGE_ALIGN_LOCAL 32u /* modes L and C align neighbouring images
                                   by a local gauge transformation before the
                                   arclength (mepgl_gauge_align_local), and
                                   before the constant-phase alignment when
                                   both are set; off = reference           */
#define MEPGL_ROLLBACK 16u      /* mepgl_loop keeps a device copy of the
                                   solver state from the start of each outer
                                   iteration and restores it when that
                                   iteration produces a non-finite value   */

/* A per-site parameter: a uniform value, or an N*N plane when plane != NULL. */
typedef struct {
  double value;
  const double *plane;
} mepgl_param;

/* One string of images.  Replaces the compile-time configuration of the
 * reference (config.cuh:14-34: N, dx, MULTICOMPONENT, NO_SELF_FIELD) and the
 * parameter files loaded by loadInput (io.cu:10-126). */
typedef struct {
  int N;                 /* grid points per side                          */
  int F;                 /* images, 1 .. 2048                             */
  int ncomp;             /* 1 or 2 order-parameter components             */
  int no_self_field;     /* 1: A fixed, magnetic term and dF/dA omitted   */
  double dx;             /* lattice spacing                               */
  mepgl_param a_1, b_1, m_xx_1, m_yy_1;
  mepgl_param a_2, b_2, m_xx_2, m_yy_2;
  mepgl_param h;         /* applied field                                 */
  double q_1, q_2, eta, gamma, delta; /* q.npy (io.cu:60-67)              */
  const int32_t *comp_domain; /* N*N, computational domain (nonzero = in) */
  const int32_t *sc_domain;   /* N*N, superconducting domain              */
  int segment_rows;      /* rows one wavefront marches; 0 = library picks */
} mepgl_string_desc;

typedef struct mepgl_ctx mepgl_ctx;

/* Creates a context for n_strings strings on HIP device `device`, builds the
 * neighbour maps (buildNearestNeighbourMap, nlcg.cu:48-103, launched from
 * mepgl.cu:598-599) and puts the solver in the state loop() starts from
 * (mepgl.cu:124-140).  Fields start at zero until mepgl_upload.
 * Returns MEPGL_ERR_ARG, before touching the device and with *out NULL,
 * when one image's field set of a string would exceed 2^31 - 1 bytes
 * ((N + 1) * (2 + 2 ncomp) * roundup(N, 64) stored reals: N above about
 * 8,190 for one component in FP64, 6,680 for two). */
int mepgl_create(const mepgl_string_desc *strings, int n_strings, int precision,
                 unsigned flags, int device, mepgl_ctx **out);
void mepgl_destroy(mepgl_ctx *ctx);
const char *mepgl_last_error(const mepgl_ctx *ctx);

/* Host -> device copy of one string's fields (the host-to-device copies of
 * mepgl.cu:567-595).  psi_2 may be NULL for one component. */
int mepgl_upload(mepgl_ctx *ctx, int string, const double *a, const double *psi_1,
                 const double *psi_2);
/* Device -> host copy of the current fields (mepgl.cu:791-796). */
int mepgl_download(mepgl_ctx *ctx, int string, double *a, double *psi_1, double *psi_2);

/* Solver state of loop()'s start (mepgl.cu:124-140): zero gradients and
 * direction, infinite gradient norms. */
int mepgl_reset_solver(mepgl_ctx *ctx);

/* The boundary phase winding of every image (compute_winding,
 * gl_utils.py:10-41, as post.py:140-144 reports it: 2 pi times the winding
 * number, not divided by 2 pi), for both components; read with
 * MEPGL_SCALAR_WINDING_1 and MEPGL_SCALAR_WINDING_2. */
int mepgl_winding(mepgl_ctx *ctx);

/* computeFreeEnergy (nlcg.cu:217-240) for every image of every string,
 * evaluated in double; read with MEPGL_SCALAR_ENERGY. */
int mepgl_free_energy(mepgl_ctx *ctx);

/* The gradient at the current fields (gradientPsi + gradientA,
 * nlcg.cu:504-888) into host arrays of one string; does not disturb the
 * solver state.  psi_2_g may be NULL for one component. */
int mepgl_gradient(mepgl_ctx *ctx, int string, double *a_g, double *psi_1_g, double *psi_2_g);

/* nlcgSteps (nlcg.cu:1462-1586) on every string: iter_n NLCG steps,
 * Polak-Ribiere beta when nlcg != 0, psi frozen when maxwell_solver != 0.
 * Returns MEPGL_ERR_NONFINITE if an energy or gradient norm is not finite. */
int mepgl_nlcg_steps(mepgl_ctx *ctx, int iter_n, int nlcg, int maxwell_solver);

/* computeJB (nlcg.cu:243-380): magnetic field b and supercurrents j_1, j_2
 * of one string into host arrays (F*N*N, F*N*N*2, F*N*N*2). */
int mepgl_compute_jb(mepgl_ctx *ctx, int string, double *b, double *j_1, double *j_2);

/* arcLength + computeJB + arcLengthJB (interpolation.cu:143-169,
 * mepgl.cu:193-201) on every string; read with MEPGL_SCALAR_S and
 * MEPGL_SCALAR_S_JB. */
int mepgl_arc_length(mepgl_ctx *ctx);

/* redistributeLinear / redistributeCubic (interpolation.cu:229-250,
 * 404-481) of every string on its last s_jb, keeping F images. */
int mepgl_redistribute(mepgl_ctx *ctx, int cubic);

/* The end-of-run resampling of mepgl.cu:699-766: arclength, then linear
 * redistribution of one string to F_new images (reallocates its buffers).
 * That string restarts from loop()'s initial solver state (zero gradients
 * and direction, infinite norms); the other strings of the context keep
 * their gradients, directions and per-image scalars. */
int mepgl_resample(mepgl_ctx *ctx, int string, int F_new);

/* Aligns the constant phase of neighbouring images: image n is rotated by
 * Theta_n = sum_{m<=n} arg(sum conj(psi_m) psi_{m-1}), both components by
 * the same angle, together with its direction and gradients.  No counterpart
 * in the reference, which measures image distances with the gauge-invariant
 * metric s_jb (mepgl.cu:31-32); the rotation is an exact symmetry of the
 * discrete energy and never increases a neighbour distance.  Read the
 * angles with MEPGL_SCALAR_THETA. */
int mepgl_gauge_align(mepgl_ctx *ctx);

/* Aligns neighbouring images by a local gauge transformation: for image
 * n >= 1, xi_n minimises |A_n + D xi - A_{n-1}|^2 (D: forward differences,
 * zero in the last column for x and the last row for y), solved exactly in
 * the DCT-II eigenbasis of the grid's Neumann Laplacian; chi_n = chi_{n-1} +
 * xi_n is applied as A_n += D chi_n and psi_a,n *= exp(i q_a chi_n) to the
 * fields, and to the psi parts of the direction and both gradients.  No
 * counterpart in the reference, which measures image distances with the
 * gauge-invariant metric s_jb (mepgl.cu:31-32) and interpolates the raw
 * fields (interpolation.cu:190-227).  The plaquette curls of every image
 * are unchanged and no neighbour A distance grows; the energy changes at
 * first order in dx (the discretisation is gauge covariant to that order).
 * Strings with a fixed A (no self field) are left unchanged. */
int mepgl_gauge_align_local(mepgl_ctx *ctx);

/* Resets the Polak-Ribiere memory of the interior images
 * (mepgl.cu:231-234). */
int mepgl_reset_interior_beta(mepgl_ctx *ctx);

/* loop() (mepgl.cu:148-306): `iterations` outer iterations of mode
 * 'S' (relaxation), 'M' (Maxwell: psi frozen), 'L' (linear
 * reparametrisation then relaxation) or 'C' (cubic).  Each outer iteration
 * is one hipGraph replay unless MEPGL_NO_GRAPH.  As in the reference
 * (mepgl.cu:357-359), the stop control is checked after each iteration, so
 * a call runs at least one; the control stays set until cleared.
 * Every outer iteration is checked for a NaN or +-Inf energy, line-search
 * coefficient or gradient norm; the first such iteration ends the call with
 * MEPGL_ERR_NONFINITE (mepgl_last_error names it and the first bad image).
 * Without MEPGL_ROLLBACK at most one further iteration has run by then;
 * with it, the state from the start of the bad iteration is restored. */
int mepgl_loop(mepgl_ctx *ctx, int iterations, int relaxation_steps, char mode, int nlcg);

/* Per-image scalars of one string (F values, 4F for MEPGL_SCALAR_POLY). */
#define MEPGL_SCALAR_FENERGY 0   /* line-search A: energy before the update */
#define MEPGL_SCALAR_G_NORM2 1   /* caller's g_norm2                        */
#define MEPGL_SCALAR_ALPHA 2
#define MEPGL_SCALAR_S 3
#define MEPGL_SCALAR_S_JB 4
#define MEPGL_SCALAR_BETA 5
#define MEPGL_SCALAR_G_DELTA_G 6
#define MEPGL_SCALAR_POLY 7      /* B, C, D, E per image                    */
#define MEPGL_SCALAR_ENERGY 8    /* last mepgl_free_energy                  */
#define MEPGL_SCALAR_G_NORM2_OLD 9
#define MEPGL_SCALAR_LS_INDEX 10 /* scan index where the line search stopped */
#define MEPGL_SCALAR_THETA 11    /* phase applied by the last alignment     */
#define MEPGL_SCALAR_WINDING_1 12 /* boundary winding of psi_1 (mepgl_winding) */
#define MEPGL_SCALAR_WINDING_2 13 /* boundary winding of psi_2                 */
int mepgl_get_scalars(mepgl_ctx *ctx, int string, int which, double *out);

/* Internal solver buffers of one string, for state checks:
 * 0 = caller's gradient g, 1 = caller's g_old, 2 = direction d. */
int mepgl_download_state(mepgl_ctx *ctx, int string, int which, double *a, double *psi_1,
                         double *psi_2);

/* Restores saved solver state, for an exact resume of the relaxation: the
 * same buffers as mepgl_download_state (0 g, 1 g_old, 2 d), and the
 * per-image g_norm2, g_norm2_old, s and s_jb.  The reference restarts each
 * run from zero gradients and directions (mepgl.cu:124-140); a resume from
 * its reloaded output files is mepgl_upload + mepgl_reset_solver.  A call
 * sequence download, download_state x3, get_scalars x2 ... upload,
 * upload_state x3, set_scalars x2 continues bit for bit.  At sites whose
 * neighbour map is empty in both domains the gradient and direction must be
 * zero, as every state the loop produces has them: the step kernels skip
 * such sites' stencils and keep those values. */
int mepgl_upload_state(mepgl_ctx *ctx, int string, int which, const double *a,
                       const double *psi_1, const double *psi_2);
int mepgl_set_scalars(mepgl_ctx *ctx, int string, int which, const double *values);

/* Controls of the stdin thread (mepgl.cu:635-671), safe from any thread. */
#define MEPGL_CONTROL_STOP 0
#define MEPGL_CONTROL_RELAXATION_STEPS 1
#define MEPGL_CONTROL_NLCG 2
/* No counterpart in the reference: where P's line search runs.  1 (the
 * default): in the G of the next step of the same call or outer iteration
 * (every G wavefront scans the table from P's partials), so only the last
 * step's P has a finishing launch, when every string has at most 64 tiles
 * per image (ceil(N / 64) * ceil(N / segment rows)); 0, or more tiles: in a
 * finishing launch after every P.  Results are identical bit for bit.  Set
 * it between calls, from the calling thread (a change drops the captured
 * graphs). */
#define MEPGL_CONTROL_FOLD_LINE_SEARCH 3
/* No counterpart in the reference: how the NLCG steps of an outer iteration
 * are launched.  Split (the default): the images of each class are cut into
 * parts at image boundaries, and the parts' steps run as launch sequences on
 * their own streams, so that one part's P (mostly arithmetic) runs beside
 * another part's G (mostly memory traffic).  A context starts with 2 parts;
 * mepgl_autotune picks 2 or 4 together with the segment heights.  Splitting
 * applies when an outer iteration has more than one step, and not while
 * per-kernel timing is on.
 * 0: one launch sequence.  1: split, with the current part count.  2 to 4:
 * split into that many parts.  Other values: MEPGL_ERR_ARG.
 * mepgl_get_control returns the part count, or 0 when not split.  Results are
 * identical bit for bit.  Set it between calls, from the calling thread (a
 * change drops the captured graphs). */
#define MEPGL_CONTROL_SPLIT_IMAGES 4
/* No counterpart in the reference: in mixed precision, the four strips of a
 * G or P workgroup trade the columns just outside each strip through LDS
 * when they are one segment of one image (1, the default), instead of
 * loading them (0).  Results are identical bit for bit.  Set it between
 * calls (a change drops the captured graphs). */
#define MEPGL_CONTROL_EDGE_EXCHANGE 5
int mepgl_set_control(mepgl_ctx *ctx, int key, int value);
int mepgl_get_control(mepgl_ctx *ctx, int key, int *value);

/* Per-kernel timing with HIP events on the context's stream.  enable != 0
 * brackets every launch with events and disables graph replay; read returns
 * the average duration in ms and the launch count of kernel `kernel`:
 * 0 = G with the pending update, 1 = P (direction + quartic),
 * 2 = G without an update, 3 = reserved (P reduces beta itself, so there
 * is no finishing launch after G; reads a count of 0),
 * 4 = finishing launch after P (quartic totals + line search; with the
 * line search folded into G, only after the last step of a call or outer
 * iteration). */
int mepgl_profile(mepgl_ctx *ctx, int enable);
int mepgl_profile_read(mepgl_ctx *ctx, int kernel, double *avg_ms, int *count);

/* computeFreeEnergy (nlcg.cu:217-240) of one string held in caller-owned
 * device memory, allocating nothing.  No counterpart in the reference (its
 * computeFreeEnergy allocates a density buffer, nlcg.cu:226, and its sums
 * their partials, parallel_sum.cu:52-66): this entry point is for callers
 * that cannot allocate device memory, such as code under HIP graph capture,
 * serving loops and harnesses that forbid hipMalloc.  It needs no context.
 * For this call every pointer is a device pointer: comp_domain and
 * sc_domain in the descriptor (N*N int32, [j][i]) and any parameter plane
 * (N*N double, [j][i]); a, psi_1, psi_2 (psi_2 only for two components) are
 * (F, N, N, 2) double arrays in the layout above; energy receives F doubles.
 * stream is a hipStream_t (NULL: the default stream); the call is
 * asynchronous on it.  The per-tile partial sums go to one static device
 * array per GPU of 65,536 (sum, compensation) pairs, so
 * F * ceil(N / 64) * ceil(N / rows) may not exceed 65,536, where rows is
 * desc->segment_rows (16 when 0); larger inputs return MEPGL_ERR_ARG before
 * any launch.  The array keeps the last call's partials (each call
 * overwrites every partial it reads), and calls that overlap on one GPU
 * race on it.  With the same segment height the energies equal those of
 * mepgl_free_energy on an FP64 context bit for bit. */
int mepgl_free_energy_noalloc(const mepgl_string_desc *desc, const double *a,
                              const double *psi_1, const double *psi_2, double *energy,
                              void *stream);

/* Waits for all work of the context. */
int mepgl_synchronize(mepgl_ctx *ctx);

/* Schedule.  The reference launches one thread per cell with a fixed block
 * size (config.cuh:25-31, blockSizeX); here one wavefront marches a
 * 64-column strip down `rows` rows (heights of 3k + 1 rows waste no work).
 * mepgl_set_segment_rows sets one string's height (clamped to
 * [min(N, 4), N]).  mepgl_autotune times five rounds per candidate height
 * (4 .. 127) of an outer iteration of mode L with `steps` NLCG steps
 * (arclength, linear redistribution, the steps; only the steps when a string
 * has one image), each round visiting every candidate in turn, for each
 * class of strings with equal N, components and fixed-A setting.  It keeps
 * the height with the fastest round and restores the solver state, so
 * tuning leaves fields, scalars and reparametrisation plans as it found
 * them; it needs device memory for one copy of the fields, gradients and
 * directions.  rows_out
 * (n_strings entries) may be NULL.  Sums over cells are taken per strip, so
 * a height change can move results in the last bits (well inside 1e-12
 * relative). */
int mepgl_set_segment_rows(mepgl_ctx *ctx, int string, int rows);
int mepgl_autotune(mepgl_ctx *ctx, int steps, int *rows_out);

/* Geometry and memory facts: work items (wavefronts) per NLCG launch,
 * bytes of device memory held, and the segment height of a string. */
int mepgl_info(mepgl_ctx *ctx, int string, int *tile_rows, int *work_items, long long *bytes);

#ifdef __cplusplus
}
#endif
#endif /* MEPGL_H */
