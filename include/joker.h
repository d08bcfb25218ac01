/*
 * joker.h -- C ABI of the MI355X (gfx950) implementation of The Joker's
 * marginal-likelihood and rejection-sampling path.
 *
 * Conventions
 *   - Every function returns an int status (JK_OK == 0); jk_status_string()
 *     names it.  Per-sample numeric failures are not errors: they are written
 *     into the output exactly as the reference writes them (+inf when the
 *     L x L system is singular, NaN for invalid rows).
 *   - Array arguments of the compute entry points are DEVICE pointers (HBM,
 *     e.g. torch tensors' data_ptr()).  `stream` is a hipStream_t passed as
 *     void* (NULL = the default stream).  Calls are asynchronous on `stream`.
 *   - Prior rows are structure-of-arrays columns P [day], e, omega [rad],
 *     M0 [rad], s [RV unit] -- the packed order of fast_likelihood.pyx:41 with
 *     the array-of-structs chunk split into columns.
 *   - A jk_star is immutable after creation; any number of streams may use it
 *     concurrently (the reference's CJokerHelper is not reentrant,
 *     fast_likelihood.pyx:100-114).
 *
 * Reference citations are relative to the adrn/thejoker v1.3.0 source tree.
 */
#ifndef THEJOKER_AMD_JOKER_H
#define THEJOKER_AMD_JOKER_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  Replaces the reference's Python exceptions at the C level:
 * ValueError on bad shapes (fast_likelihood.pyx:174-179) -> JK_EINVAL. */
enum {
    JK_OK = 0,
    JK_EINVAL = 1,   /* bad argument (shape, size, mode)                     */
    JK_ENOMEM = 2,   /* device allocation failed                             */
    JK_EHIP = 3,     /* a HIP runtime call failed                            */
    JK_ENODEV = 4,   /* no usable gfx950 device                              */
    JK_ERANGE = 5    /* a size exceeds a documented maximum                  */
};

/* Jitter handling.  JK_JITTER_REFERENCE reproduces fast_likelihood.pyx, where
 * the jitter-inflated inverse variance s_ivar is computed (:458) but never
 * read, so lnL does not depend on s.  JK_JITTER_PHYSICAL uses
 * C = diag(sigma^2 + s^2), i.e. ivar / (1 + s^2 ivar) as in
 * src/tests/py_likelihood.py:17-29. */
enum { JK_JITTER_REFERENCE = 0, JK_JITTER_PHYSICAL = 1 };

/* Prior on K.  JK_KPRIOR_NORMAL: Lambda[0] fixed (fixed_K_prior == 1 in
 * fast_likelihood.pyx:225-228).  JK_KPRIOR_FIXED_COMPANION_MASS:
 * Lambda[0] = min(max_K^2, sigma_K0^2 (P/P0)^(-2/3) / (1 - e^2)) per sample
 * (fast_likelihood.pyx:461-464; distributions.py:195-244). */
enum { JK_KPRIOR_NORMAL = 0, JK_KPRIOR_FIXED_COMPANION_MASS = 1 };

/* Largest supported number of linear parameters (K, v0, offsets, trends). */
#define JK_LMAX 8

typedef struct jk_star jk_star;

/* Status name.  No reference counterpart (Python exceptions). */
const char *jk_status_string(int status);

/* ABI version, for the Python loader's check. */
int jk_abi_version(void);

/* Select and initialise the device; idempotent.  Replaces nothing in the
 * reference (CPU only); required before any other call on `device`.
 *
 * Scratch.  The rejection, maximum, polishing and settling calls keep their
 * temporary device memory (ballot words, block counts, partial maxima, FP64
 * lists) in one buffer per (device, stream), allocated with hipMalloc on
 * first use and reused by later calls on that stream in stream order.  It only
 * grows (x1.5, in 2 MiB steps, after synchronising the stream) and is kept
 * until jk_release_scratch or the end of the process: a server keeps the peak
 * scratch of every stream it has used.  A call holds its stream's buffer from
 * its first to its last launch, so host threads calling on one stream queue
 * their launch sequences one after the other.  Under stream capture the call
 * takes stream-ordered pool memory instead (hipMallocAsync / hipFreeAsync);
 * jk_init sets the release threshold of the device's default pool to 32 GiB
 * so that those blocks are reused. */
int jk_init(int device);

/* Free every scratch buffer kept for `device` (see jk_init), after waiting
 * for the work queued on its stream.  Later calls allocate afresh.  No
 * reference counterpart. */
int jk_release_scratch(int device);

/* Build the per-star constants.  Replaces CJokerHelper.__init__
 * (fast_likelihood.pyx:125-253) as constructed by TheJoker._make_joker_helper
 * (thejoker.py:87-91).  Host arrays:
 *   t[n]       observation times, BMJD days (data._t_bmjd, :163)
 *   rv[n]      RVs (data.rv.value, :164)
 *   ivar[n]    1/sigma^2 (data.ivar, :165-166)
 *   trend[n*(L-1)] row-major trend_M (likelihood_helpers.py:28-37; the
 *              shape check is :174-179)
 *   mu[L], lambda_diag[L]  prior mean / variance of the linear parameters
 *              in M_T row order (:202-252; lambda_diag[0] is ignored for
 *              JK_KPRIOR_FIXED_COMPANION_MASS)
 *   t_ref      data._t_ref_bmjd (:162)
 * The star lives on the device selected by jk_init.  Its arrays are uploaded
 * on a stream of the library's own, so the call does not wait for the
 * launches queued on the caller's streams. */
int jk_star_create(int n, const double *t, const double *rv, const double *ivar,
                   int L, const double *trend, const double *mu, const double *lambda_diag,
                   double t_ref, int k_prior_mode, double sigma_K0, double P0, double max_K,
                   int jitter_mode, jk_star **out);

/* A star whose data have a full covariance matrix: cov[n*n] row-major, in
 * (RV unit)^2, symmetric positive definite (its symmetric part is used);
 * n <= 64.  Other arguments as jk_star_create.  The reference's fast path
 * reads a 1-D ivar (fast_likelihood.pyx:165-166, :317) and rejects 2-D data
 * (data.py:146-150 makes ivar a matrix); its slow test twin
 * (tests/py_likelihood.py:74-78) is diagonal too, so this entry point has no
 * counterpart: the likelihood is ln N(y | M mu, C + s^2 I + M Lambda M^T)
 * (reference jitter mode: C), computed in the eigenbasis of C, where
 * C + s^2 I is diagonal.  The eigen-decomposition (cyclic Jacobi) runs here
 * on the host.  Such a star is accepted by jk_lnl_f64, jk_lnl_reject_f64 and
 * jk_linear_posterior; the mixed-precision, polishing and many-star calls
 * return JK_EINVAL for it.  Not positive definite: JK_EINVAL; n > 64:
 * JK_ERANGE. */
int jk_star_create_cov(int n, const double *t, const double *rv, const double *cov, int L,
                       const double *trend, const double *mu, const double *lambda_diag,
                       double t_ref, int k_prior_mode, double sigma_K0, double P0, double max_K,
                       int jitter_mode, jk_star **out);

int jk_star_destroy(jk_star *star);

/* Plain FP64 marginal log-likelihood.  Replaces
 * CJokerHelper.batch_marginal_ln_likelihood (fast_likelihood.pyx:428-469):
 * ll_out[i] = lnL of prior row rows[i] (or row i when rows == NULL).
 * Device arrays P, e, omega, M0, s index the prior cache; ll_out has n
 * entries.  Matches the reference's lnL to 1e-10 relative. */
int jk_lnl_f64(const jk_star *star, int64_t n, const double *P, const double *e,
               const double *omega, const double *M0, const double *s,
               const int64_t *rows, double *ll_out, void *stream);

/* Mixed precision marginal log-likelihood: FP64 phase 2 pi (t - t_ref)/P - M0
 * and its 2 pi wrap, FP32 Kepler solve and design column, FP64 moment
 * accumulation, jitter weights, Cholesky log-det and lnL.  Same arguments
 * and semantics as jk_lnl_f64 (the FP64 columns: the kernel is bound by
 * arithmetic, and an FP32 e would cost sqrt(1 - e^2) its accuracy as
 * e -> 1).
 * bound_out (device, n floats, optional) receives for every row a bound on
 * |lnL - lnL of jk_lnl_f64| (first order in the FP32 design-column error,
 * doubled), +inf where none is known; jk_settle_batch uses it.
 * No reference counterpart beyond fast_likelihood.pyx:428-469. */
int jk_lnl_mixed(const jk_star *star, int64_t n, const double *P, const double *e,
                 const double *omega, const double *M0, const double *s,
                 const int64_t *rows, double *ll_out, float *bound_out, void *stream);

/* FP64 re-evaluation of the live set after a mixed-precision pass: every
 * entry i >= first of ll_inout (n rows, as in jk_lnl_mixed) that is not below
 * max(ll[0..n)) - window -- all of them when the maximum is NaN or infinite --
 * is recomputed with the FP64 kernel in place; *n_polished_out (device int64)
 * receives how many.  Entries before `first` are taken as already FP64 (an
 * earlier call polished them against a maximum no larger than today's, so
 * none of them can have entered the window since): they count toward the
 * maximum only.  With window >= 40, exp(lnL - max) of every row left in
 * mixed precision is below the smallest nonzero uniform (2^-53), so the
 * rejection step's accepted set equals the plain FP64 path's.  No reference
 * counterpart (the reference has one precision). */
int jk_polish_f64(const jk_star *star, int64_t n, const double *P, const double *e,
                  const double *omega, const double *M0, const double *s,
                  const int64_t *rows, double *ll_inout, double window, int64_t first,
                  int64_t *n_polished_out, void *stream);

/* NaN-propagating maximum of ll[0..n) into *max_out (device), as np.max at
 * likelihood_helpers.py:108.  n must be > 0. */
int jk_max_f64(const double *ll, int64_t n, double *max_out, void *stream);

/* Rejection step, likelihood_helpers.py:107-109 / multiproc_helpers.py:256-258:
 *   u_i = i-th output of numpy's Generator(PCG64).uniform after skipping
 *         `stream_offset` draws from the generator whose bit_generator.state
 *         is {state = state_hi:state_lo, inc = inc_hi:inc_lo}
 *   keep i  iff  exp(ll_i - max(ll)) > u_i
 * idx_out (device, max_keep entries) receives the first max_keep kept
 * indices in ascending order; *n_pass_out (device int64) the number kept
 * before truncation; *llmax_out (device double) max(ll).  The caller then
 * advances its host generator by n draws (bit_generator.advance(n)). */
int jk_reject(const double *ll, int64_t n, uint64_t state_hi, uint64_t state_lo,
              uint64_t inc_hi, uint64_t inc_lo, uint64_t stream_offset, int64_t max_keep,
              int64_t *idx_out, int64_t *n_pass_out, double *llmax_out, void *stream);

/* The same rejection step against a caller-supplied maximum llmax (device
 * scalar), used when the maximum comes from elsewhere (an FP64 re-evaluated
 * live set, or a previous launch). */
int jk_reject_with_max(const double *ll, int64_t n, const double *llmax,
                       uint64_t state_hi, uint64_t state_lo, uint64_t inc_hi, uint64_t inc_lo,
                       uint64_t stream_offset, int64_t max_keep, int64_t *idx_out,
                       int64_t *n_pass_out, void *stream);

/* One rejection-sampling pass of one star in a single call: the plain FP64
 * lnL of n prior rows into ll_out (exactly as jk_lnl_f64) followed by the
 * rejection step over them (exactly as jk_reject, same arguments and
 * outputs), with max(ll) reduced inside the likelihood launch instead of by
 * a second pass over ll_out.  Replaces marginal_ln_likelihood_inmem plus the
 * rejection step of rejection_sample_inmem (likelihood_helpers.py:104-109). */
int jk_lnl_reject_f64(const jk_star *star, int64_t n, const double *P, const double *e,
                      const double *omega, const double *M0, const double *s,
                      const int64_t *rows, double *ll_out, uint64_t state_hi, uint64_t state_lo,
                      uint64_t inc_hi, uint64_t inc_lo, uint64_t stream_offset, int64_t max_keep,
                      int64_t *idx_out, int64_t *n_pass_out, double *llmax_out, void *stream);

/* jk_lnl_reject_f64 without any allocation: the same arguments, kernels and
 * outputs, with the scratch (per-wave maxima, ballot words, block counts) in a
 * file-scope device array of 24 MiB that each GPU holds from jk_init to the
 * end of the process.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory -- code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc.  n <= JK_NOALLOC_MAX_ROWS, else
 * JK_ERANGE.  All calls on one GPU share its array, whatever the stream or
 * thread: calls that overlap in time on one GPU race and corrupt each other's
 * results, so the caller serialises them (one stream, one thread).  Nothing
 * in the array needs clearing between calls. */
#define JK_NOALLOC_MAX_ROWS ((int64_t)1 << 27)
int jk_lnl_reject_f64_noalloc(const jk_star *star, int64_t n, const double *P,
                              const double *e, const double *omega, const double *M0,
                              const double *s, const int64_t *rows, double *ll_out,
                              uint64_t state_hi, uint64_t state_lo, uint64_t inc_hi,
                              uint64_t inc_lo, uint64_t stream_offset, int64_t max_keep,
                              int64_t *idx_out, int64_t *n_pass_out, double *llmax_out,
                              void *stream);

/* One prior row in full detail, for tests.  Replaces
 * CJokerHelper.test_likelihood_worker and the public members it fills
 * (fast_likelihood.pyx:109-114, 547-576).  row: device array of 5 doubles
 * (P, e, omega, M0, s).  Device outputs:
 *   MT_out[L*n]     design matrix M_T, row 0 the unit-K Keplerian column
 *   Ainv_out[L*L]   Lambda^-1 + M_T C^-1 M_T^T (the reference's Ainv)
 *   A_out[L*L]      its inverse (A)
 *   a_out[L]        Ainv^-1 (M_T C^-1 y + Lambda^-1 mu) (a)
 *   b_out[n]        M mu (b)
 *   B_out[n*n]      C + M Lambda M^T (B); optional (NULL)
 *   Binv_out[n*n]   C^-1 - C^-1 M A M^T C^-1 (Binv); optional (NULL)
 *   scalars_out[4]  lnL, chi2 = (b - y)^T Binv (b - y), ln det(2 pi B),
 *                   Lambda[0]
 * C = diag(1/ivar) in the reference jitter mode (as the reference), else
 * diag(1/ivar + s^2).  Lambda[0] is not clipped to max_K^2 on this path, as in
 * test_likelihood_worker (:567-569).  lnL is +inf when Ainv is not positive
 * definite, NaN for an invalid row.  Diagonal-covariance stars only
 * (JK_EINVAL otherwise).  One lane: not for bulk use. */
int jk_debug_single(const jk_star *star, const double *row, double *MT_out, double *Ainv_out,
                    double *A_out, double *a_out, double *b_out, double *B_out, double *Binv_out,
                    double *scalars_out, void *stream);

/* Per-sample numeric failures of an lnL array ll[0..n) (device):
 * counts_out[0] = entries at +inf (the L x L system was singular, where the
 * reference writes INF: fast_likelihood.pyx:283-291, 347-348, 381-382),
 * counts_out[1] = entries at -inf, counts_out[2] = NaN entries (invalid
 * prior rows).  counts_out: device, 3 uint64, overwritten. */
int jk_count_failures(const double *ll, int64_t n, uint64_t *counts_out, void *stream);

/* numpy-identical uniforms u[i] = output offset+i of the PCG64 stream
 * (Generator.uniform, likelihood_helpers.py:107), for tests and callers that
 * want the draws themselves. */
int jk_pcg64_uniform(uint64_t state_hi, uint64_t state_lo, uint64_t inc_hi, uint64_t inc_lo,
                     uint64_t offset, int64_t n, double *u_out, void *stream);

/* Linear-parameter posterior of selected rows.  Replaces the per-row linear
 * algebra of CJokerHelper.batch_get_posterior_samples
 * (fast_likelihood.pyx:471-531): for row rows[i] (or i), a_out[i*L..] = a
 * (the solve of :394-420) and cov_out[i*L*L..] = inv(Ainv) (the covariance
 * passed to rng.multivariate_normal at :529-530), ll_out[i] = lnL.  Lambda[0]
 * is not clipped to max_K^2 on this path (fast_likelihood.pyx:519-521).  The
 * multivariate-normal draws stay on the host so numpy's stream is unchanged. */
int jk_linear_posterior(const jk_star *star, int64_t n, const double *P, const double *e,
                        const double *omega, const double *M0, const double *s,
                        const int64_t *rows, double *a_out, double *cov_out, double *ll_out,
                        void *stream);


/* ------------------------------------------------------------------------
 * Many stars in one launch.  No reference counterpart: the reference handles
 * one star per CJokerHelper (fast_likelihood.pyx:70) and parallelises only
 * over prior rows (multiproc_helpers.py:17-60).
 * ------------------------------------------------------------------------ */

typedef struct jk_batch jk_batch;

/* A unit of work of jk_lnl_batch: rows [row_begin, row_begin + n_rows) of the
 * prior cache (or of `rows`) against star `star`, written to
 * ll_out[out_offset .. out_offset + n_rows). */
typedef struct {
    int32_t star;
    int32_t n_rows;
    int64_t row_begin;
    int64_t out_offset;
} jk_work_item;

/* One rejection segment of jk_reject_batch: ll[offset .. offset + n) with its
 * own numpy PCG64 state (state/inc as in jk_reject, skipping stream_offset
 * draws); kept indices (relative to offset, ascending, at most max_keep)
 * go to idx_out[idx_offset ..]. */
typedef struct {
    int64_t offset;
    int64_t n;
    int64_t max_keep;
    int64_t idx_offset;
    uint64_t state_hi, state_lo, inc_hi, inc_lo;
    uint64_t stream_offset;
} jk_segment;

/* Gather stars (all with the same L and jitter mode) into a device-resident
 * star table; stars[i] is addressed as star index i by work items.  The
 * jk_star objects may be destroyed afterwards. */
int jk_batch_create(const jk_star *const *stars, int n_stars, jk_batch **out);

/* jk_linear_posterior for the accepted rows of many stars of a table in one
 * launch: row i (cache row rows[i]) belongs to table entry star_idx[i]
 * (device int32); outputs laid out as jk_linear_posterior's with the
 * table's L.  star_idx entries must be valid table indices (slots in use).
 * No reference counterpart (the reference draws per star). */
int jk_linear_posterior_batch(const jk_batch *batch, int64_t n, const int32_t *star_idx,
                              const double *P, const double *e, const double *omega,
                              const double *M0, const double *s, const int64_t *rows,
                              double *a_out, double *cov_out, double *ll_out, void *stream);
int jk_batch_destroy(jk_batch *batch);

/* A star table of `capacity` empty slots for stars with L linear parameters
 * and jitter mode `jitter_mode`, filled and refilled by jk_batch_set_star:
 * the form a request loop uses, where stars come and go between launches.
 * No work item may name a slot that was never set. */
int jk_batch_create_slots(int capacity, int L, int jitter_mode, jk_batch **out);

/* Point slot `slot` of a table made by jk_batch_create_slots at `star` (same
 * L, jitter mode and device).  The star's device arrays are referenced, not
 * copied, so the star must outlive every launch that reads the slot.  The
 * update is ordered on `stream` after the work already queued there. */
int jk_batch_set_star(jk_batch *batch, int slot, const jk_star *star, void *stream);

/* Mark slot `slot` of a star table exact (exact != 0) or not: mixed-precision
 * launches over the table (jk_lnl_batch and jk_lnl_prune_batch with
 * precision 1) evaluate an exact slot's rows in FP64 and write error bound 0
 * for them, so the settling calls leave them alone.  For stars whose rows
 * nearly all lie within the settling margin of the maximum (flat
 * likelihoods), where every mixed value would be re-evaluated in FP64
 * anyway.  jk_batch_set_star clears the mark.  Ordered on `stream`.  No
 * counterpart in the reference, which evaluates everything in FP64
 * (fast_likelihood.pyx:255-469); a request loop that serves
 * iterative_rejection_sample (thejoker.py:259-370) uses it per request. */
int jk_batch_set_exact(jk_batch *batch, int slot, int exact, void *stream);

/* Persistent work-queue launch over device-resident work items.
 * precision 0: FP64; 1: mixed (as jk_lnl_mixed).  The columns are the FP64
 * ones in both cases.  bound_out (mixed only, optional): per-row error bounds as in
 * jk_lnl_mixed, at the same offsets as ll_out.  block: threads per workgroup
 * (64, 128 or 256); grid: workgroups (0 = 8 per compute unit).  workspace:
 * NULL or 256 device bytes (the queue counter; pass one to keep the call free
 * of allocation). */
int jk_lnl_batch(const jk_batch *batch, int precision, int64_t n_items,
                 const jk_work_item *items, const double *P, const void *e, const void *omega,
                 const void *M0, const void *s, const int64_t *rows, double *ll_out,
                 float *bound_out, int block, int grid, void *workspace, void *stream);

/* jk_lnl_batch with pruning by an upper bound.  No reference counterpart:
 * it returns the same rejection decisions as evaluating every row
 * (likelihood_helpers.py:107-109) with less arithmetic.
 * `probes` is a table with the same L and jitter mode whose entry k holds
 * star k's data on a subset S of its epochs (jk_star_create on those
 * epochs).  Because the rest's conditional density is at most that of its
 * noise alone,
 *   lnL(row) <= lnL_S(row) - 1/2 sum_{n not in S} ln(2 pi sigma_n^2).
 * prune holds two doubles per table entry k: (thr_k, add_k).  For a row of
 * star k whose bound lnL_S + add_k (plus, for precision 1, lnL_S's error
 * bound) is finite and below thr_k, ll_out receives the bound instead of the row's
 * lnL (bound_out -1: the entry is an upper bound on lnL, not an estimate;
 * the settling calls re-evaluate such a row in FP64 when it comes within
 * their margin).  bound_out is required for precision 1; for precision 0 it
 * is optional and receives only these marks (0 for a row evaluated in FP64,
 * -1 for an upper bound).  Every other row gets exactly what jk_lnl_batch
 * writes.  thr_k = -inf evaluates all of star k's rows.
 * A caller that sets thr_k = (a value known not to exceed the final maximum)
 * minus a margin of 746 or more leaves every decision
 * exp(ll - max) > u unchanged for all u in [0, 1).
 * With a smaller margin (at least 37), only u == 0 can differ (see
 * jk_reject_batch_zero_check).
 * n_full_out (optional, device uint64, one entry per table entry): entry k
 * accumulates the number of star k's rows evaluated in full.  Rows are read from the cache directly: work items
 * address cache rows row_begin .. row_begin + n_rows. */
int jk_lnl_prune_batch(const jk_batch *batch, const jk_batch *probes, const double *prune,
                       int precision, int64_t n_items, const jk_work_item *items, const double *P,
                       const void *e, const void *omega, const void *M0, const void *s,
                       double *ll_out, float *bound_out, uint64_t *n_full_out, int block,
                       int grid, void *workspace, void *stream);

/* Segmented rejection step (jk_reject per segment, one launch sequence for
 * all): segs is a HOST array of n_seg descriptors; n_pass_out[i] and
 * llmax_out[i] (device) receive segment i's pass count and maximum. */
int jk_reject_batch(const double *ll, int n_seg, const jk_segment *segs, int64_t *idx_out,
                    int64_t *n_pass_out, double *llmax_out, void *stream);

/* jk_reject_batch that also sets zero_u_out[i] (device int64, cleared first)
 * to 1 when one of segment i's uniforms is exactly 0 (probability 2^-53 per
 * row): the one case that settling with a margin below 746 leaves open
 * (jk_settle_live_batch).  No reference counterpart. */
int jk_reject_batch_zero_check(const double *ll, int n_seg, const jk_segment *segs,
                               int64_t *idx_out, int64_t *n_pass_out, double *llmax_out,
                               int64_t *zero_u_out, void *stream);

/* One segment of jk_polish_batch: ll[offset .. offset + n) holds the lnL of
 * star `star` (star-table index) on cache rows row_begin .. (or
 * rows[row_begin ..] when a rows array is given); entries from `first` on are
 * mixed precision, earlier ones already FP64 (as in jk_polish_f64). */
typedef struct {
    int64_t offset;
    int64_t n;
    int32_t star;
    int32_t reserved;
    int64_t row_begin;
    int64_t first;
} jk_polish_segment;

/* jk_polish_f64 for many segments in one launch sequence: each segment's
 * entries not below its own maximum - window are recomputed in FP64 in
 * place; n_polished_out[i] (device) counts segment i's.  segs is a HOST
 * array; FP64 columns as in jk_lnl_f64. */
int jk_polish_batch(const jk_batch *batch, int n_seg, const jk_polish_segment *segs,
                    const double *P, const double *e, const double *omega, const double *M0,
                    const double *s, const int64_t *rows, double *ll_inout, double window,
                    int64_t *n_polished_out, void *stream);

/* One segment of jk_settle_batch: ll[offset .. offset + n) and
 * bound[offset .. offset + n) of star `star` (star-table index) on cache rows
 * row_begin .. (or rows[row_begin ..]), and the numpy PCG64 stream its
 * rejection step will draw (state / inc / stream_offset as in jk_segment). */
typedef struct {
    int64_t offset;
    int64_t n;
    int32_t star;
    int32_t reserved;
    int64_t row_begin;
    uint64_t state_hi, state_lo, inc_hi, inc_lo;
    uint64_t stream_offset;
} jk_settle_segment;

/* Make a mixed-precision lnL buffer decide like the FP64 one.  For every
 * segment, with bound (from jk_lnl_batch / jk_lnl_mixed; 0 for rows already
 * FP64) bounding |lnL - lnL(FP64)| per row, the rows that could hold the
 * FP64 maximum, and then the rows whose rejection decision
 * exp(lnL - max) > u (u: the segment's uniform for that row) the bound leaves
 * open, are re-evaluated in FP64 in place and their bound set to 0.  A
 * following jk_reject_batch on the same segments and streams then accepts
 * exactly the rows the FP64 buffer would.  n_settled_out[i] (device) counts
 * segment i's re-evaluated rows.  segs is a HOST array; FP64 columns as in
 * jk_lnl_f64.  No reference counterpart (the reference has one precision). */
int jk_settle_batch(const jk_batch *batch, int n_seg, const jk_settle_segment *segs,
                    const double *P, const double *e, const double *omega, const double *M0,
                    const double *s, const int64_t *rows, double *ll_inout, float *bound_inout,
                    int64_t *n_settled_out, void *stream);

/* jk_settle_batch followed by jk_reject_batch (max_keep for every segment;
 * idx_out holds the segments' kept indices one after the other, min(n,
 * max_keep) places each), in one launch sequence that draws each row's
 * uniform once: after the rows that could hold the maximum are settled, one
 * pass decides every row the bound leaves no doubt about and flags the
 * others, which are re-evaluated in FP64 and decided from their own uniform.
 * Same accepted set, n_pass and maximum as the two calls; n_settled_out[i]
 * counts segment i's re-evaluated rows.  segs is a HOST array.  No reference
 * counterpart. */
int jk_settle_reject_batch(const jk_batch *batch, int n_seg, const jk_settle_segment *segs,
                           int64_t max_keep, const double *P, const double *e,
                           const double *omega, const double *M0, const double *s,
                           const int64_t *rows, double *ll_inout, float *bound_inout,
                           int64_t *idx_out, int64_t *n_pass_out, double *llmax_out,
                           int64_t *n_settled_out, void *stream);

/* One segment of jk_settle_live_batch: entries [0, n) of the lnL buffer at
 * offset (cache rows row_begin + i, or rows[row_begin + i]) of table star;
 * entries [first, n) are new since the segment's previous call, and floor is
 * the FP64 maximum of entries [0, first) (the maximum the rejection step
 * found over them; -inf when first = 0). */
typedef struct {
    int64_t offset;
    int64_t n;
    int64_t first;
    int32_t star;
    int32_t reserved;
    int64_t row_begin;
    double floor;
} jk_live_segment;

/* Settle a growing mixed-precision lnL buffer for EVERY future uniform, as
 * the iterative rejection loop (likelihood_helpers.py:164-201) needs: it
 * draws new uniforms for all entries each round.  Of each segment's new
 * entries, the ones with lnL + bound >= lo - margin are re-evaluated in FP64
 * in place (bound := 0), where lo = max(floor, max over the new entries of
 * lnL - bound) is a lower bound on the FP64 maximum.  An entry whose bound is
 * negative (pruned by jk_lnl_prune_batch: lnL is an upper bound) counts with
 * bound 0 in the first test and not at all in lo.  An entry left in mixed
 * precision then lies more than margin below the maximum in both precisions,
 * so exp(lnL - max) is 0 and the entry is rejected whatever the uniform;
 * after a call on every new entry, jk_reject_batch over [0, n) accepts
 * exactly the FP64 set under any stream.  That needs margin >= 746 (exp
 * underflows to 0 below -745.14; the library uses 750).  A margin from 40 up
 * covers every uniform but exactly 0 (exp(-40) < 2^-53, the smallest
 * positive uniform): jk_reject_batch_zero_check reports that case, and the
 * caller then settles the segment with margin 750 from first = 0 and repeats
 * the rejection step.  Work is proportional to the new entries only.
 * n_settled_out[i] (device, int64) receives segment i's count.  segs is a HOST array.  No reference counterpart. */
int jk_settle_live_batch(const jk_batch *batch, int n_seg, const jk_live_segment *segs,
                         const double *P, const double *e, const double *omega, const double *M0,
                         const double *s, const int64_t *rows, double *ll_inout,
                         float *bound_inout, double margin, int64_t *n_settled_out, void *stream);

/* Settle a mixed-precision lnL buffer to a tolerance: of each segment's
 * entries [first, n) (jk_live_segment; floor unused), every one whose bound
 * exceeds tol, is NaN, or is negative (an upper bound left by
 * jk_lnl_prune_batch) is re-evaluated in FP64 in place (bound := 0).
 * Afterwards |lnL - lnL(FP64)| <= tol on every entry, whatever its distance
 * from the maximum: the form for lnL arrays handed back to a caller (the
 * reference returns FP64 values, thejoker.py:93-138, 252-257).  tol >= 0.
 * n_settled_out[i] (device) receives segment i's count.  segs is a HOST
 * array.  No reference counterpart (the reference has one precision). */
int jk_settle_tol_batch(const jk_batch *batch, int n_seg, const jk_live_segment *segs,
                        const double *P, const double *e, const double *omega, const double *M0,
                        const double *s, const int64_t *rows, double *ll_inout,
                        float *bound_inout, float tol, int64_t *n_settled_out, void *stream);

/* Settle a buffer of values and upper bounds: a jk_lnl_prune_batch launch
 * with threshold +inf leaves, for every probed row, an upper bound on its lnL
 * (bound -1); this call evaluates in FP64 (in place, bound := 0) the rows that
 * can matter, for each segment's entries [first, n) (jk_live_segment, floor
 * unused):
 *   1. with hi = max of ll + max(bound, 0), an upper bound on the FP64
 *      maximum, every upper-bound row with ll >= hi - window;
 *   2. with lo = max over the value rows (bound >= 0) of ll - bound, a lower
 *      bound on the FP64 maximum, every upper-bound row with ll >= lo - margin.
 * Afterwards every upper-bound row lies more than margin below the maximum,
 * so (margin >= 40) a rejection step decides it like its FP64 value for
 * every uniform but 0 -- which jk_settle_reject_batch handles by evaluating
 * such a row, as any row its bound leaves open.  n_settled_out[i] (device)
 * counts segment i's evaluated rows.  segs is a HOST array.  No reference
 * counterpart: the same rejection decisions as fast_likelihood.pyx:428-469
 * plus likelihood_helpers.py:107-109 with most rows never evaluated in full. */
int jk_settle_pruned_batch(const jk_batch *batch, int n_seg, const jk_live_segment *segs,
                           const double *P, const double *e, const double *omega,
                           const double *M0, const double *s, const int64_t *rows,
                           double *ll_inout, float *bound_inout, double window, double margin,
                           int64_t *n_settled_out, void *stream);

/* A segment plan: the descriptors and scratch of jk_settle_batch and
 * jk_reject_batch for a fixed layout of segments (jk_settle_segment: offsets,
 * sizes, stars, first rows, first PCG64 streams), kept on the device, so that
 * the launches of jk_seg_plan_settle / jk_seg_plan_reject allocate nothing,
 * copy nothing from the host and never wait -- the form a HIP graph captures,
 * to be replayed step after step.  max_keep: the rejection step's limit per
 * segment (idx_out of jk_seg_plan_reject holds the segments' kept indices one
 * after the other, min(n, max_keep) places each).  The batch must outlive the
 * plan.  No reference counterpart. */
typedef struct jk_seg_plan jk_seg_plan;
int jk_seg_plan_create(const jk_batch *batch, int n_seg, const jk_settle_segment *segs,
                       int64_t max_keep, jk_seg_plan **out);
int jk_seg_plan_destroy(jk_seg_plan *plan);

/* New numpy PCG64 streams for the plan's segments: streams holds n_seg groups
 * of {state_hi, state_lo, inc_hi, inc_lo, stream_offset}.  Staged in pinned
 * memory and copied on `stream`, ordered before the work queued there later
 * (e.g. the next replay of a graph that holds the plan's launches); waits only
 * for the previous update's copy. */
int jk_seg_plan_set_streams(jk_seg_plan *plan, const uint64_t *streams, void *stream);

/* jk_settle_batch and jk_reject_batch on the plan's segments. */
int jk_seg_plan_settle(const jk_seg_plan *plan, const double *P, const double *e,
                       const double *omega, const double *M0, const double *s,
                       const int64_t *rows, double *ll_inout, float *bound_inout,
                       int64_t *n_settled_out, void *stream);
int jk_seg_plan_reject(const jk_seg_plan *plan, const double *ll, int64_t *idx_out,
                       int64_t *n_pass_out, double *llmax_out, void *stream);
/* jk_settle_reject_batch on the plan's segments (max_keep of the plan). */
int jk_seg_plan_settle_reject(const jk_seg_plan *plan, const double *P, const double *e,
                              const double *omega, const double *M0, const double *s,
                              const int64_t *rows, double *ll_inout, float *bound_inout,
                              int64_t *idx_out, int64_t *n_pass_out, double *llmax_out,
                              int64_t *n_settled_out, void *stream);

/* Fill rows [row0, row0 + n) of a device prior cache (arrays indexed from
 * 0 for row row0) with draws of the default nonlinear prior
 * (prior.py:410-493): P log-uniform on [P_min, P_max] (distributions.py:
 * 24-28), e from the inverse CDF given as n_tab + 1 equally spaced quantiles
 * (device array; e.g. of Kipping13Global's Beta(0.867, 3.03),
 * distributions.py:263-268), omega and M0 uniform on [0, 2 pi),
 * ln s ~ N(ln_s_mu, ln_s_sigma) or s = 0 when ln_s_sigma < 0.  Draws come
 * from a counter-based Philox4x32-10 stream keyed by `seed`, so row r is the
 * same whatever chunk it is generated in.  The FP32 columns are optional
 * (all four or none).  No reference counterpart: the reference draws priors
 * on the host through pymc (prior.py:297-407). */
int jk_prior_fill(int64_t n, int64_t row0, uint64_t seed, double P_min, double P_max,
                  const double *e_quantiles, int n_tab, double ln_s_mu, double ln_s_sigma,
                  double *P, double *e, double *omega, double *M0, double *s, float *e32,
                  float *omega32, float *M032, float *s32, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* THEJOKER_AMD_JOKER_H */
