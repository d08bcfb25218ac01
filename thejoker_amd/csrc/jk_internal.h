// jk_internal.h -- host-side launch interface between the C ABI (jk_capi.hip)
// and the kernels (jk_kernels.hip).  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "jk_device.h"

// Largest epoch count of a full-covariance star (jk_star_create_cov): its
// kernels keep a row's design column, one double per epoch and lane, in LDS.
#define JK_COV_NMAX 64

// Host mirror of one star: the StarDev block passed by value to kernels,
// plus the device allocation that holds the epoch arrays it points into.
struct jk_star {
    StarDev dev;        // by-value kernel argument (pointers are device)
    void *d_epochs;     // device allocation behind dev.dt / rp / y / ivar / T
    int device;
};

// Device-resident table of stars for the batched launch.
struct jk_batch {
    StarDev *d_stars;
    int n_stars;
    int L;
    int jitter_mode;
    int device;
    void *d_epochs;     // epoch arrays of all stars (StarDev pointers point here)
};

// Internal segment descriptor of the segmented rejection step.
struct SegDev {
    int64_t off, n, max_keep, idx_off;
    int64_t mask_off, bsum_off, nb, max_off;
    u128 s_base, inc;
    int64_t star, row0;   // segmented polish: star-table index, first cache row
    int64_t first;        // segmented polish: first entry still in mixed precision
    double floor;         // live settling: FP64 maximum of the rows settled before
};

// Jump table of numpy's PCG64 LCG: (A^(2^i), D_(2^i)) for i = 0..63, where
// a jump by k steps maps state -> A^k state + inc * D_k.
struct PcgJump {
    u128 A[64];
    u128 D[64];
};

// Host-side 128-bit helpers (same arithmetic as the device u128 functions).
void jk_pcg_jump_coeffs(uint64_t k, u128 *A, u128 *D);

hipError_t jk_launch_lnl_f64(const jk_star *st, int64_t n, const double *P, const double *e,
                             const double *om, const double *M0, const double *s,
                             const int64_t *rows, double *ll, hipStream_t stream);
hipError_t jk_launch_lnl_mixed(const jk_star *st, int64_t n, const double *P, const double *e,
                               const double *om, const double *M0, const double *s,
                               const int64_t *rows, double *ll, float *bound, hipStream_t stream);
hipError_t jk_launch_max(const double *ll, int64_t n, double *partials, int n_partials,
                         double *max_out, hipStream_t stream);
int jk_max_partials_count(int64_t n);
hipError_t jk_launch_reject(const double *ll, int64_t n, const double *llmax, u128 state, u128 inc,
                            uint64_t offset, int64_t max_keep, int64_t *idx_out,
                            int64_t *n_pass_out, uint64_t *mask_scratch,
                            int64_t *bsum_scratch, hipStream_t stream);
int64_t jk_reject_mask_words(int64_t n);
int64_t jk_reject_bsum_count(int64_t n);
hipError_t jk_launch_uniform(u128 state, u128 inc, uint64_t offset, int64_t n, double *u,
                             hipStream_t stream);
hipError_t jk_launch_linear_posterior(const jk_star *st, int64_t n, const double *P,
                                      const double *e, const double *om, const double *M0,
                                      const double *s, const int64_t *rows, double *a_out,
                                      double *cov_out, double *ll_out, hipStream_t stream);
hipError_t jk_launch_settle_tol_batch(const jk_batch *b, int n_seg, const SegDev *d_segs,
                                      int64_t max_n, int64_t max_nb, uint64_t *mask,
                                      int64_t *bsum, int64_t *idx, int64_t *cum, const double *P,
                                      const double *e, const double *om, const double *M0,
                                      const double *s, const int64_t *rows, double *ll,
                                      float *bound, float tol, int64_t *n_settled,
                                      hipStream_t stream);
hipError_t jk_launch_settle_pruned_batch(const jk_batch *b, int n_seg, const SegDev *d_segs,
                                         int64_t max_n, int64_t max_nb, double *partials, int pb,
                                         uint64_t *mask, int64_t *bsum, int64_t *idx,
                                         double *level, int64_t *cum, int64_t *count1,
                                         const double *P, const double *e, const double *om,
                                         const double *M0, const double *s, const int64_t *rows,
                                         double *ll, float *bound, double window, double margin,
                                         int64_t *n_settled, hipStream_t stream);
hipError_t jk_launch_debug_row(const jk_star *st, const double *row, double *MT, double *Ainv,
                               double *A, double *a, double *b, double *B, double *Binv,
                               double *sc, hipStream_t stream);
hipError_t jk_launch_count_failures(const double *ll, int64_t n, unsigned long long *counts,
                                    hipStream_t stream);
hipError_t jk_launch_linear_posterior_batch(const jk_batch *b, int64_t n, const int32_t *star_idx,
                                            const double *P, const double *e, const double *om,
                                            const double *M0, const double *s,
                                            const int64_t *rows, double *a_out, double *cov_out,
                                            double *ll_out, hipStream_t stream);
hipError_t jk_upload_pcg_table(int device);
hipError_t jk_upload_sincos_table(int device);
int64_t jk_lnl_reject_scratch_bytes(int64_t n);
hipError_t jk_launch_lnl_reject_f64(const jk_star *st, int64_t n, const double *P, const double *e,
                                    const double *om, const double *M0, const double *s,
                                    const int64_t *rows, double *ll, u128 state, u128 inc,
                                    uint64_t offset, int64_t max_keep, int64_t *idx_out,
                                    int64_t *n_pass_out, double *llmax_out, void *scratch,
                                    hipStream_t stream);
int64_t jk_polish_scratch_bytes(int64_t n);
hipError_t jk_launch_lnl_batch(const jk_batch *b, int precision, int64_t n_items,
                               const jk_work_item *items, const double *P, const void *e,
                               const void *om, const void *M0, const void *s,
                               const int64_t *rows, double *ll, float *bound, int block, int grid,
                               unsigned *counter, hipStream_t stream);
hipError_t jk_launch_set_exact(const jk_batch *b, int slot, int exact, hipStream_t stream);
hipError_t jk_launch_lnl_prune_batch(const jk_batch *b, const jk_batch *pb, const double *prune,
                                     int precision, int64_t n_items, const jk_work_item *items,
                                     const double *P, const void *e, const void *om,
                                     const void *M0, const void *s, double *ll, float *bound,
                                     unsigned long long *n_full, int block, int grid,
                                     unsigned *counter, hipStream_t stream);
hipError_t jk_launch_reject_batch(const double *ll, int n_seg, const SegDev *d_segs,
                                  int64_t max_n, int64_t max_nb, double *partials, int pb,
                                  uint64_t *mask, int64_t *bsum, int64_t *idx_out,
                                  int64_t *n_pass_out, double *llmax_out, hipStream_t stream,
                                  int64_t *zero_u = nullptr);
void jk_host_pcg_jump(u128 state, u128 inc, uint64_t k, u128 *out);
hipError_t jk_launch_polish_batch(const jk_batch *b, int n_seg, const SegDev *d_segs,
                                  int64_t max_n, int64_t max_nb, double *partials, int pb,
                                  uint64_t *mask, int64_t *bsum, int64_t *idx, double *llmax,
                                  int64_t *cum, const double *P, const double *e, const double *om,
                                  const double *M0, const double *s, const int64_t *rows,
                                  double *ll, double window, int64_t *n_polished,
                                  hipStream_t stream);
hipError_t jk_launch_prior_fill(int64_t n, int64_t row0, uint64_t seed, double P_min, double P_max,
                                const double *e_tab, int n_tab, double s_mu, double s_sigma,
                                double *P, double *e, double *om, double *M0, double *s,
                                float *e32, float *om32, float *M032, float *s32,
                                hipStream_t stream);
int jk_seg_flag_blocks(int64_t max_n);
hipError_t jk_launch_settle_batch(const jk_batch *b, int n_seg, const SegDev *d_segs,
                                  int64_t max_n, int64_t max_nb, double *partials, int pb,
                                  uint64_t *mask, int64_t *bsum, int64_t *idx, double *lo,
                                  double *llmax, int64_t *cum, int64_t *count1, const double *P,
                                  const double *e, const double *om, const double *M0,
                                  const double *s, const int64_t *rows, double *ll, float *bound,
                                  int64_t *n_settled, hipStream_t stream);
hipError_t jk_launch_settle_live_batch(const jk_batch *b, int n_seg, const SegDev *d_segs,
                                       int64_t max_n, int64_t max_nb, double *partials, int pb,
                                       uint64_t *mask, int64_t *bsum, int64_t *idx, double *lo,
                                       int64_t *cum, const double *P, const double *e,
                                       const double *om, const double *M0, const double *s,
                                       const int64_t *rows, double *ll, float *bound,
                                       double margin, int64_t *n_settled, hipStream_t stream);
hipError_t jk_launch_settle_reject_batch(const jk_batch *b, int n_seg, const SegDev *d_settle,
                                         const SegDev *d_reject, int64_t max_n, int64_t max_nb,
                                         double *partials, int pb, uint64_t *mask,
                                         uint64_t *open_mask, int64_t *bsum, int64_t *idx,
                                         double *lo, int64_t *cum, int64_t *count1,
                                         const double *P, const double *e, const double *om,
                                         const double *M0, const double *s, const int64_t *rows,
                                         double *ll, float *bound, int64_t *idx_out,
                                         int64_t *n_pass_out, double *llmax_out,
                                         int64_t *n_settled, hipStream_t stream);
hipError_t jk_launch_polish(const jk_star *st, int64_t n, const double *P, const double *e,
                            const double *om, const double *M0, const double *s,
                            const int64_t *rows, double *ll, double window, int64_t first,
                            void *scratch, int64_t *n_polished, hipStream_t stream);
