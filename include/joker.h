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
 * reference (CPU only); required before any other call on `device`. */
int jk_init(int device);

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
 * The star lives on the device selected by jk_init. */
int jk_star_create(int n, const double *t, const double *rv, const double *ivar,
                   int L, const double *trend, const double *mu, const double *lambda_diag,
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
 * accumulation, Cholesky log-det and lnL.  Same semantics as jk_lnl_f64; the
 * angular / eccentricity / jitter columns are the FP32 copy of the cache.
 * No reference counterpart beyond fast_likelihood.pyx:428-469. */
int jk_lnl_mixed(const jk_star *star, int64_t n, const double *P, const float *e,
                 const float *omega, const float *M0, const float *s,
                 const int64_t *rows, double *ll_out, void *stream);

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

#ifdef __cplusplus
}
#endif

#endif /* THEJOKER_AMD_JOKER_H */
