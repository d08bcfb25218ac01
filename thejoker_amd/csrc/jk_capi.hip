// jk_capi.hip -- extern "C" entry points of libjoker (declared in
// include/joker.h).  Host-side validation, star construction and scratch
// management; all arithmetic is in jk_kernels.hip.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <memory>
#include <mutex>
#include <vector>

#include "jk_internal.h"

#define JK_ABI_VERSION 5
#define JK_POOL_KEEP (32ull << 30)

// Scratch of jk_lnl_reject_f64_noalloc: a file-scope device array, so that
// the entry point allocates nothing.  Each GPU the library runs on holds its
// own copy, at its own address, until the process ends.  The array's name is
// a host-side stand-in, not a device address: jk_init looks the address up
// once per device (hipGetSymbolAddress, with that device current) and every
// launch takes that pointer.  No kernel needs any of it cleared: every byte
// a launch reads (per-wave maxima, ballot words, block counts) is written
// earlier in the same launch sequence.
#define JK_NOALLOC_SCRATCH_BYTES ((size_t)24 << 20)
static __device__ __align__(256) unsigned char scratch[JK_NOALLOC_SCRATCH_BYTES];

namespace {

std::mutex g_init_mu;
bool g_inited[64] = {false};
void *g_noalloc_scratch[64] = {nullptr};   // device address of `scratch` per device
// Per device, a non-blocking stream for the upload of a new star's epoch
// arrays: a copy on the default stream waits for every launch queued there.
// On its own stream the upload waits for its copy only (though where the
// runtime copies with a kernel, that kernel still needs a free compute unit,
// which a persistent launch occupying the device does not leave).
hipStream_t g_upload_stream[64] = {nullptr};

int check_hip(hipError_t e) { return e == hipSuccess ? JK_OK : JK_EHIP; }

int current_device_inited()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return -1;
    if (dev < 0 || dev >= 64 || !g_inited[dev]) return -1;
    return dev;
}

u128 make_u128(uint64_t hi, uint64_t lo)
{
    u128 r;
    r.lo = lo;
    r.hi = hi;
    return r;
}

// Scratch kept per (device, stream) between calls.  The settling and
// rejection calls need up to a few GB of temporary device memory each; taken
// from the stream-ordered pool at every call, it costs a new mapping whenever
// the pool holds no free block of the size asked for, and such calls were
// measured at up to seconds with the device idle meanwhile.  The next call on
// the same stream reuses the buffer, after this call's work in stream order.
//
// A call holds its slot (the slot's mutex) from scratch_get until
// scratch_put, i.e. until its last launch is queued: two host threads that
// call on the same stream therefore queue their launch sequences one after the
// other instead of interleaving them over one buffer, and a slot grows (after
// synchronising its stream) only while no other call holds its old pointer.
// jk_release_scratch frees a device's slots; otherwise they live until the
// process ends.  Under stream capture the stream-ordered pool is used instead,
// so that a captured graph never refers to memory a later call may replace.
struct ScratchSlot {
    int device;
    hipStream_t stream;
    void *ptr;
    size_t bytes;
    bool dead;          // released by jk_release_scratch; look the slot up again
    std::mutex mu;      // held by the call using the slot
};
std::mutex g_scratch_mu;
std::vector<std::shared_ptr<ScratchSlot>> g_scratch;

struct ScratchLease {
    std::shared_ptr<ScratchSlot> slot;   // locked while held
    bool pooled = false;                 // stream-ordered pool memory (capture)
};

hipError_t scratch_get(hipStream_t st, size_t bytes, void **out, ScratchLease *lease)
{
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    hipError_t e = hipStreamIsCapturing(st, &cs);
    if (e != hipSuccess) return e;
    if (cs != hipStreamCaptureStatusNone) {
        lease->pooled = true;
        return hipMallocAsync(out, bytes, st);
    }
    lease->pooled = false;
    int dev = 0;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
    std::shared_ptr<ScratchSlot> slot;
    for (;;) {
        {
            std::lock_guard<std::mutex> lock(g_scratch_mu);
            slot.reset();
            for (auto &sl : g_scratch)
                if (sl->device == dev && sl->stream == st) slot = sl;
            if (!slot) {
                slot = std::make_shared<ScratchSlot>();
                slot->device = dev;
                slot->stream = st;
                slot->ptr = nullptr;
                slot->bytes = 0;
                slot->dead = false;
                g_scratch.push_back(slot);
            }
        }
        slot->mu.lock();
        if (!slot->dead) break;
        slot->mu.unlock();
    }
    if (slot->bytes < bytes || !slot->ptr) {
        size_t want = bytes < 256 ? 256 : bytes;
        if (slot->ptr) {
            want = std::max(want, slot->bytes + slot->bytes / 2);
            if ((e = hipStreamSynchronize(st)) == hipSuccess) e = hipFree(slot->ptr);
            if (e != hipSuccess) {
                slot->mu.unlock();
                return e;
            }
            slot->ptr = nullptr;
            slot->bytes = 0;
        }
        const size_t align = (size_t)2 << 20;
        want = (want + align - 1) / align * align;
        void *ptr = nullptr;
        if ((e = hipMalloc(&ptr, want)) != hipSuccess) {
            slot->mu.unlock();
            return e;
        }
        slot->ptr = ptr;
        slot->bytes = want;
    }
    *out = slot->ptr;
    lease->slot = std::move(slot);
    return hipSuccess;
}

hipError_t scratch_put(hipStream_t st, void *ptr, ScratchLease &lease)
{
    if (lease.pooled) return hipFreeAsync(ptr, st);
    if (lease.slot) {
        lease.slot->mu.unlock();
        lease.slot.reset();
    }
    return hipSuccess;
}

}  // namespace

extern "C" {

const char *jk_status_string(int status)
{
    switch (status) {
    case JK_OK: return "ok";
    case JK_EINVAL: return "invalid argument";
    case JK_ENOMEM: return "device allocation failed";
    case JK_EHIP: return "HIP runtime error";
    case JK_ENODEV: return "no usable gfx950 device (call jk_init first)";
    case JK_ERANGE: return "size exceeds the supported maximum";
    default: return "unknown status";
    }
}

int jk_abi_version(void) { return JK_ABI_VERSION; }

int jk_init(int device)
{
    std::lock_guard<std::mutex> lock(g_init_mu);
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return JK_ENODEV;
    if (device < 0 || device >= count || device >= 64) return JK_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return JK_EHIP;
    if (g_inited[device]) return JK_OK;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return JK_EHIP;
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) return JK_ENODEV;
    if (jk_upload_pcg_table(device) != hipSuccess) return JK_EHIP;
    if (jk_upload_sincos_table(device) != hipSuccess) return JK_EHIP;
    if (!g_upload_stream[device] &&
        hipStreamCreateWithFlags(&g_upload_stream[device], hipStreamNonBlocking) != hipSuccess)
        return JK_EHIP;
    void *sp = nullptr;
    if (hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)) != hipSuccess || !sp) return JK_EHIP;
    g_noalloc_scratch[device] = sp;
    // The device's default stream-ordered pool serves the small workspaces of
    // calls made without one and the scratch of calls under graph capture;
    // keep up to JK_POOL_KEEP bytes of it instead of returning them at every
    // synchronisation, so that the next call maps nothing.
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess) {
        uint64_t keep = JK_POOL_KEEP;
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    }
    if (hipDeviceSynchronize() != hipSuccess) return JK_EHIP;
    g_inited[device] = true;
    return JK_OK;
}

int jk_release_scratch(int device)
{
    if (device < 0 || device >= 64) return JK_EINVAL;
    if (!g_inited[device]) return JK_OK;
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) return JK_EHIP;
    std::vector<std::shared_ptr<ScratchSlot>> mine;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        std::vector<std::shared_ptr<ScratchSlot>> keep;
        for (auto &sl : g_scratch) (sl->device == device ? mine : keep).push_back(sl);
        g_scratch.swap(keep);
    }
    if (mine.empty()) return JK_OK;
    if (hipSetDevice(device) != hipSuccess) return JK_EHIP;
    hipError_t e = hipSuccess;
    for (auto &sl : mine) {
        // waits for a call still queuing on the slot, then for its work
        std::lock_guard<std::mutex> hold(sl->mu);
        sl->dead = true;
        if (sl->ptr) {
            hipError_t e1 = hipStreamSynchronize(sl->stream);
            if (e1 == hipSuccess) e1 = hipFree(sl->ptr);
            if (e1 != hipSuccess && e == hipSuccess) e = e1;
            sl->ptr = nullptr;
            sl->bytes = 0;
        }
    }
    (void)hipSetDevice(cur);
    return check_hip(e);
}

// The star from epoch times t (epoch order) and rv, ivar, trend (N x (L-1)
// row-major) either in epoch order (Qt == nullptr) or rotated into the
// eigenbasis of a full covariance whose eigenvectors are the rows of Qt.
static int star_build(int n, const double *t, const double *rv, const double *ivar, int L,
                      const double *trend, const double *mu, const double *lambda_diag,
                      double t_ref, int k_prior_mode, double sigma_K0, double P0, double max_K,
                      int jitter_mode, const double *Qt, jk_star **out)
{
    if (!out) return JK_EINVAL;
    *out = nullptr;
    if (n <= 0 || L < 1 || L > JK_LMAX || !t || !rv || !ivar || !mu || !lambda_diag)
        return JK_EINVAL;
    if (L > 1 && !trend) return JK_EINVAL;
    if (jitter_mode != JK_JITTER_REFERENCE && jitter_mode != JK_JITTER_PHYSICAL) return JK_EINVAL;
    if (k_prior_mode != JK_KPRIOR_NORMAL && k_prior_mode != JK_KPRIOR_FIXED_COMPANION_MASS)
        return JK_EINVAL;
    const int dev = current_device_inited();
    if (dev < 0) return JK_ENODEV;
    for (int i = 0; i < L; ++i)
        if (!(lambda_diag[i] > 0.0) && !(i == 0 && k_prior_mode == JK_KPRIOR_FIXED_COMPANION_MASS))
            return JK_EINVAL;
    if (k_prior_mode == JK_KPRIOR_FIXED_COMPANION_MASS && !(P0 > 0.0)) return JK_EINVAL;

    jk_star *st = new (std::nothrow) jk_star;
    if (!st) return JK_ENOMEM;
    memset(st, 0, sizeof(*st));
    st->device = dev;
    StarDev &d = st->dev;
    d.n = n;
    d.L = L;
    d.jitter_mode = jitter_mode;
    d.k_prior_mode = k_prior_mode;
    d.sigK0sq = sigma_K0 * sigma_K0;
    d.P0 = P0;
    d.maxKsq = max_K * max_K;
    d.mu0 = mu[0];
    for (int i = 0; i < L; ++i) d.mu[i] = mu[i];
    d.lam0 = lambda_diag[0];
    d.lam_inv[0] = k_prior_mode == JK_KPRIOR_NORMAL ? 1.0 / lambda_diag[0] : 0.0;
    d.ln_lam_trend = 0.0;
    for (int i = 1; i < L; ++i) {
        d.lam_inv[i] = 1.0 / lambda_diag[i];
        d.ln_lam_trend += log(lambda_diag[i]);
    }

    // Host epoch arrays: dt, r' = y - T mu_trend, y, ivar, T (L-1 rows of n).
    const int LT = L - 1;
    const size_t n_q = Qt ? (size_t)n * n : 0;
    std::vector<double> h((size_t)n * (5 + LT) + n_q);
    double *h_dt = h.data(), *h_rp = h_dt + n, *h_y = h_rp + n, *h_iv = h_y + n, *h_s2 = h_iv + n,
           *h_T = h_s2 + n;
    d.sum_ln_ivar = 0.0;
    d.S_rr = 0.0;
    for (int k = 0; k < n; ++k) {
        if (!(ivar[k] > 0.0) || !isfinite(ivar[k]) || !isfinite(t[k]) || !isfinite(rv[k])) {
            delete st;
            return JK_EINVAL;
        }
        h_dt[k] = t[k] - t_ref;
        double rp = rv[k];
        for (int j = 0; j < LT; ++j) {
            const double Tv = trend[(size_t)k * LT + j];
            h_T[(size_t)j * n + k] = Tv;
            rp -= Tv * mu[j + 1];
        }
        h_rp[k] = rp;
        h_y[k] = rv[k];
        h_iv[k] = ivar[k];
        h_s2[k] = 1.0 / ivar[k];
        d.sum_ln_ivar += log(ivar[k]);
        d.S_rr += ivar[k] * rp * rp;
        for (int j = 0; j < LT; ++j) {
            const double wT = ivar[k] * h_T[(size_t)j * n + k];
            d.S_tr[j] += wT * rp;
            d.S_ty[j] += wT * rv[k];
            for (int q = 0; q <= j; ++q) {
                d.S_tt[j * JK_LMAX + q] += wT * h_T[(size_t)q * n + k];
            }
        }
    }
    d.t0_one = L >= 2 && !Qt;
    for (int k = 0; k < n && d.t0_one; ++k) d.t0_one = h_T[k] == 1.0;
    d.ln_c_phys = (double)n * JK_LN_TWO_PI + d.ln_lam_trend +
                  (k_prior_mode == JK_KPRIOR_NORMAL ? log(lambda_diag[0]) : 0.0);
    d.ln_c = d.ln_c_phys - d.sum_ln_ivar;
    for (int j = 0; j < LT; ++j)
        for (int q = 0; q < j; ++q) d.S_tt[q * JK_LMAX + j] = d.S_tt[j * JK_LMAX + q];
    if (Qt) std::copy(Qt, Qt + n_q, h.begin() + (size_t)n * (5 + LT));

    void *dmem = nullptr;
    if (hipMalloc(&dmem, h.size() * sizeof(double)) != hipSuccess) {
        delete st;
        return JK_ENOMEM;
    }
    // on the device's upload stream, so that the copy does not wait for the
    // launches queued on the default stream; h must outlive the copy
    hipError_t ce = hipMemcpyAsync(dmem, h.data(), h.size() * sizeof(double),
                                   hipMemcpyHostToDevice, g_upload_stream[dev]);
    if (ce == hipSuccess) ce = hipStreamSynchronize(g_upload_stream[dev]);
    if (ce != hipSuccess) {
        (void)hipFree(dmem);
        delete st;
        return JK_EHIP;
    }
    st->d_epochs = dmem;
    double *base = (double *)dmem;
    d.dt = base;
    d.rp = base + n;
    d.y = base + 2 * n;
    d.ivar = base + 3 * n;
    d.sig2 = base + 4 * n;
    d.T = base + 5 * n;
    d.Qt = Qt ? base + (size_t)n * (5 + LT) : nullptr;
    *out = st;
    return JK_OK;
}

int jk_star_create(int n, const double *t, const double *rv, const double *ivar, int L,
                   const double *trend, const double *mu, const double *lambda_diag,
                   double t_ref, int k_prior_mode, double sigma_K0, double P0, double max_K,
                   int jitter_mode, jk_star **out)
{
    return star_build(n, t, rv, ivar, L, trend, mu, lambda_diag, t_ref, k_prior_mode, sigma_K0,
                      P0, max_K, jitter_mode, nullptr, out);
}

// Eigen-decomposition of a symmetric n x n matrix A (row-major, overwritten)
// by cyclic Jacobi rotations (Golub & Van Loan, Matrix Computations, 4th ed.,
// sec. 8.5): on return the diagonal of A holds the eigenvalues and the rows
// of V the eigenvectors (V A_in V^T = diag).  Sweeps stop when the
// off-diagonal mass falls below 1e-30 of the total, or after 100 sweeps.
static bool jacobi_eigen(int n, std::vector<double> &A, std::vector<double> &V)
{
    V.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
    double total = 0.0;
    for (double a : A) total += a * a;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) off += 2.0 * A[(size_t)p * n + q] * A[(size_t)p * n + q];
        if (off <= 1e-30 * total) return true;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = A[(size_t)p * n + q];
                if (apq == 0.0) continue;
                const double app = A[(size_t)p * n + p], aqq = A[(size_t)q * n + q];
                // the rotation angle that zeroes A[p][q]: t = tan(theta)
                const double th = (aqq - app) / (2.0 * apq);
                const double tn = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                const double c = 1.0 / sqrt(tn * tn + 1.0), sn = tn * c;
                for (int k = 0; k < n; ++k) {       // columns p, q
                    const double akp = A[(size_t)k * n + p], akq = A[(size_t)k * n + q];
                    A[(size_t)k * n + p] = c * akp - sn * akq;
                    A[(size_t)k * n + q] = sn * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {       // rows p, q
                    const double apk = A[(size_t)p * n + k], aqk = A[(size_t)q * n + k];
                    A[(size_t)p * n + k] = c * apk - sn * aqk;
                    A[(size_t)q * n + k] = sn * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {       // eigenvector rows p, q
                    const double vpk = V[(size_t)p * n + k], vqk = V[(size_t)q * n + k];
                    V[(size_t)p * n + k] = c * vpk - sn * vqk;
                    V[(size_t)q * n + k] = sn * vpk + c * vqk;
                }
            }
    }
    return false;
}

int jk_star_create_cov(int n, const double *t, const double *rv, const double *cov, int L,
                       const double *trend, const double *mu, const double *lambda_diag,
                       double t_ref, int k_prior_mode, double sigma_K0, double P0, double max_K,
                       int jitter_mode, jk_star **out)
{
    if (!out) return JK_EINVAL;
    *out = nullptr;
    if (n <= 0 || !cov || !rv || !t) return JK_EINVAL;
    if (n > JK_COV_NMAX) return JK_ERANGE;
    if (L > 1 && !trend) return JK_EINVAL;
    if (L < 1 || L > JK_LMAX) return JK_EINVAL;
    // the symmetric part of cov, checked finite
    std::vector<double> A((size_t)n * n), V;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            const double a = 0.5 * (cov[(size_t)i * n + j] + cov[(size_t)j * n + i]);
            if (!isfinite(a)) return JK_EINVAL;
            A[(size_t)i * n + j] = a;
        }
    if (!jacobi_eigen(n, A, V)) return JK_EINVAL;
    // rotated data: y' = Q^T y, T' = Q^T T, ivar' = 1 / lambda (C positive definite)
    const int LT = L - 1;
    std::vector<double> yr(n), ivr(n), Tr((size_t)n * (LT > 0 ? LT : 1));
    for (int i = 0; i < n; ++i) {
        const double lam = A[(size_t)i * n + i];
        if (!(lam > 0.0)) return JK_EINVAL;
        ivr[i] = 1.0 / lam;
        const double *q = V.data() + (size_t)i * n;
        double y = 0.0;
        for (int k = 0; k < n; ++k) y += q[k] * rv[k];
        yr[i] = y;
        for (int j = 0; j < LT; ++j) {
            double tv = 0.0;
            for (int k = 0; k < n; ++k) tv += q[k] * trend[(size_t)k * LT + j];
            Tr[(size_t)i * LT + j] = tv;
        }
    }
    return star_build(n, t, yr.data(), ivr.data(), L, LT > 0 ? Tr.data() : trend, mu,
                      lambda_diag, t_ref, k_prior_mode, sigma_K0, P0, max_K, jitter_mode,
                      V.data(), out);
}


int jk_star_destroy(jk_star *star)
{
    if (!star) return JK_EINVAL;
    int dev = 0;
    (void)hipGetDevice(&dev);
    (void)hipSetDevice(star->device);
    hipError_t e = hipFree(star->d_epochs);
    (void)hipSetDevice(dev);
    delete star;
    return check_hip(e);
}

int jk_lnl_f64(const jk_star *star, int64_t n, const double *P, const double *e,
               const double *omega, const double *M0, const double *s, const int64_t *rows,
               double *ll_out, void *stream)
{
    if (!star || n < 0 || (n > 0 && (!P || !e || !omega || !M0 || !ll_out))) return JK_EINVAL;
    if (star->dev.jitter_mode == JK_JITTER_PHYSICAL && n > 0 && !s) return JK_EINVAL;
    if (current_device_inited() < 0) return JK_ENODEV;
    return check_hip(jk_launch_lnl_f64(star, n, P, e, omega, M0, s, rows, ll_out,
                                       (hipStream_t)stream));
}

int jk_lnl_mixed(const jk_star *star, int64_t n, const double *P, const double *e,
                 const double *omega, const double *M0, const double *s, const int64_t *rows,
                 double *ll_out, float *bound_out, void *stream)
{
    if (star && star->dev.Qt) return JK_EINVAL;   // full covariance: FP64 only
    if (!star || n < 0 || (n > 0 && (!P || !e || !omega || !M0 || !ll_out))) return JK_EINVAL;
    if (star->dev.jitter_mode == JK_JITTER_PHYSICAL && n > 0 && !s) return JK_EINVAL;
    if (current_device_inited() < 0) return JK_ENODEV;
    return check_hip(jk_launch_lnl_mixed(star, n, P, e, omega, M0, s, rows, ll_out, bound_out,
                                         (hipStream_t)stream));
}

int jk_max_f64(const double *ll, int64_t n, double *max_out, void *stream)
{
    if (!ll || !max_out || n <= 0) return JK_EINVAL;
    if (current_device_inited() < 0) return JK_ENODEV;
    hipStream_t st = (hipStream_t)stream;
    const int np = jk_max_partials_count(n);
    double *partials = nullptr;
    ScratchLease lease;
    if (scratch_get(st, sizeof(double) * np, (void **)&partials, &lease) != hipSuccess)
        return JK_ENOMEM;
    hipError_t err = jk_launch_max(ll, n, partials, np, max_out, st);
    hipError_t err2 = scratch_put(st, partials, lease);
    if (err != hipSuccess) return JK_EHIP;
    return check_hip(err2);
}

static int reject_impl(const double *ll, int64_t n, const double *llmax_in, uint64_t state_hi,
                       uint64_t state_lo, uint64_t inc_hi, uint64_t inc_lo, uint64_t offset,
                       int64_t max_keep, int64_t *idx_out, int64_t *n_pass_out,
                       double *llmax_out, hipStream_t st)
{
    if (!ll || n <= 0 || max_keep < 0 || !n_pass_out || (max_keep > 0 && !idx_out))
        return JK_EINVAL;
    if (!llmax_in && !llmax_out) return JK_EINVAL;
    if (current_device_inited() < 0) return JK_ENODEV;
    const int np = jk_max_partials_count(n);
    const int64_t nwords = jk_reject_mask_words(n);
    const int64_t nb = jk_reject_bsum_count(n);
    const size_t bytes = sizeof(double) * np + sizeof(uint64_t) * nwords + sizeof(int64_t) * nb;
    char *scratch = nullptr;
    ScratchLease lease;
    if (scratch_get(st, bytes, (void **)&scratch, &lease) != hipSuccess) return JK_ENOMEM;
    double *partials = (double *)scratch;
    uint64_t *mask = (uint64_t *)(scratch + sizeof(double) * np);
    int64_t *bsum = (int64_t *)(scratch + sizeof(double) * np + sizeof(uint64_t) * nwords);
    hipError_t err = hipSuccess;
    const double *mx = llmax_in;
    if (!mx) {
        err = jk_launch_max(ll, n, partials, np, llmax_out, st);
        mx = llmax_out;
    }
    if (err == hipSuccess)
        err = jk_launch_reject(ll, n, mx, make_u128(state_hi, state_lo), make_u128(inc_hi, inc_lo),
                               offset, max_keep, idx_out, n_pass_out, mask, bsum, st);
    hipError_t err2 = scratch_put(st, scratch, lease);
    if (err != hipSuccess) return JK_EHIP;
    return check_hip(err2);
}

int jk_reject(const double *ll, int64_t n, uint64_t state_hi, uint64_t state_lo, uint64_t inc_hi,
              uint64_t inc_lo, uint64_t stream_offset, int64_t max_keep, int64_t *idx_out,
              int64_t *n_pass_out, double *llmax_out, void *stream)
{
    if (!llmax_out) return JK_EINVAL;
    return reject_impl(ll, n, nullptr, state_hi, state_lo, inc_hi, inc_lo, stream_offset, max_keep,
                       idx_out, n_pass_out, llmax_out, (hipStream_t)stream);
}

int jk_reject_with_max(const double *ll, int64_t n, const double *llmax, uint64_t state_hi,
                       uint64_t state_lo, uint64_t inc_hi, uint64_t inc_lo, uint64_t stream_offset,
                       int64_t max_keep, int64_t *idx_out, int64_t *n_pass_out, void *stream)
{
    if (!llmax) return JK_EINVAL;
    return reject_impl(ll, n, llmax, state_hi, state_lo, inc_hi, inc_lo, stream_offset, max_keep,
                       idx_out, n_pass_out, nullptr, (hipStream_t)stream);
}

int jk_lnl_reject_f64(const jk_star *star, int64_t n, const double *P, const double *e,
                      const double *omega, const double *M0, const double *s,
                      const int64_t *rows, double *ll_out, uint64_t state_hi, uint64_t state_lo,
                      uint64_t inc_hi, uint64_t inc_lo, uint64_t stream_offset, int64_t max_keep,
                      int64_t *idx_out, int64_t *n_pass_out, double *llmax_out, void *stream)
{
    if (!star || n <= 0 || !P || !e || !omega || !M0 || !ll_out) return JK_EINVAL;
    if (star->dev.jitter_mode == JK_JITTER_PHYSICAL && !s) return JK_EINVAL;
    if (max_keep < 0 || !n_pass_out || !llmax_out || (max_keep > 0 && !idx_out)) return JK_EINVAL;
    if ((n + 255) / 256 > (int64_t)0x7fffffff) return JK_ERANGE;
    if (current_device_inited() < 0) return JK_ENODEV;
    hipStream_t st = (hipStream_t)stream;
    void *scratch = nullptr;
    ScratchLease lease;
    if (scratch_get(st, (size_t)jk_lnl_reject_scratch_bytes(n), (void **)&scratch, &lease) !=
        hipSuccess)
        return JK_ENOMEM;
    hipError_t err = jk_launch_lnl_reject_f64(star, n, P, e, omega, M0, s, rows, ll_out,
                                              make_u128(state_hi, state_lo),
                                              make_u128(inc_hi, inc_lo), stream_offset, max_keep,
                                              idx_out, n_pass_out, llmax_out, scratch, st);
    hipError_t err2 = scratch_put(st, scratch, lease);
    if (err != hipSuccess) return JK_EHIP;
    return check_hip(err2);
}

int jk_lnl_reject_f64_noalloc(const jk_star *star, int64_t n, const double *P,
                              const double *e, const double *omega, const double *M0,
                              const double *s, const int64_t *rows, double *ll_out,
                              uint64_t state_hi, uint64_t state_lo, uint64_t inc_hi,
                              uint64_t inc_lo, uint64_t stream_offset, int64_t max_keep,
                              int64_t *idx_out, int64_t *n_pass_out, double *llmax_out,
                              void *stream)
{
    if (!star || n <= 0 || !P || !e || !omega || !M0 || !ll_out) return JK_EINVAL;
    if (star->dev.jitter_mode == JK_JITTER_PHYSICAL && !s) return JK_EINVAL;
    if (max_keep < 0 || !n_pass_out || !llmax_out || (max_keep > 0 && !idx_out)) return JK_EINVAL;
    if (n > JK_NOALLOC_MAX_ROWS ||
        (size_t)jk_lnl_reject_scratch_bytes(n) > JK_NOALLOC_SCRATCH_BYTES)
        return JK_ERANGE;
    const int dev = current_device_inited();
    if (dev < 0 || !g_noalloc_scratch[dev]) return JK_ENODEV;
    if (star->device != dev) return JK_EINVAL;
    return check_hip(jk_launch_lnl_reject_f64(star, n, P, e, omega, M0, s, rows, ll_out,
                                              make_u128(state_hi, state_lo),
                                              make_u128(inc_hi, inc_lo), stream_offset, max_keep,
                                              idx_out, n_pass_out, llmax_out,
                                              g_noalloc_scratch[dev], (hipStream_t)stream));
}

int jk_debug_single(const jk_star *star, const double *row, double *MT_out, double *Ainv_out,
                    double *A_out, double *a_out, double *b_out, double *B_out, double *Binv_out,
                    double *scalars_out, void *stream)
{
    if (!star || !row || !MT_out || !Ainv_out || !A_out || !a_out || !b_out || !scalars_out)
        return JK_EINVAL;
    if (star->dev.Qt) return JK_EINVAL;   // diagonal data covariance only
    if (current_device_inited() < 0) return JK_ENODEV;
    return check_hip(jk_launch_debug_row(star, row, MT_out, Ainv_out, A_out, a_out, b_out, B_out,
                                         Binv_out, scalars_out, (hipStream_t)stream));
}

int jk_count_failures(const double *ll, int64_t n, uint64_t *counts_out, void *stream)
{
    if (n < 0 || !counts_out || (n > 0 && !ll)) return JK_EINVAL;
    if (current_device_inited() < 0) return JK_ENODEV;
    return check_hip(jk_launch_count_failures(ll, n, (unsigned long long *)counts_out,
                                              (hipStream_t)stream));
}

int jk_pcg64_uniform(uint64_t state_hi, uint64_t state_lo, uint64_t inc_hi, uint64_t inc_lo,
                     uint64_t offset, int64_t n, double *u_out, void *stream)
{
    if (n < 0 || (n > 0 && !u_out)) return JK_EINVAL;
    if (current_device_inited() < 0) return JK_ENODEV;
    return check_hip(jk_launch_uniform(make_u128(state_hi, state_lo), make_u128(inc_hi, inc_lo),
                                       offset, n, u_out, (hipStream_t)stream));
}

int jk_linear_posterior_batch(const jk_batch *batch, int64_t n, const int32_t *star_idx,
                              const double *P, const double *e, const double *omega,
                              const double *M0, const double *s, const int64_t *rows,
                              double *a_out, double *cov_out, double *ll_out, void *stream)
{
    if (!batch || n < 0) return JK_EINVAL;
    if (n > 0 && (!star_idx || !rows || !P || !e || !omega || !M0 || !a_out || !cov_out ||
                  !ll_out))
        return JK_EINVAL;
    if (batch->jitter_mode == JK_JITTER_PHYSICAL && n > 0 && !s) return JK_EINVAL;
    if ((n + 63) / 64 > (int64_t)0x7fffffff) return JK_ERANGE;
    if (current_device_inited() < 0) return JK_ENODEV;
    return check_hip(jk_launch_linear_posterior_batch(batch, n, star_idx, P, e, omega, M0, s, rows,
                                                      a_out, cov_out, ll_out,
                                                      (hipStream_t)stream));
}

int jk_linear_posterior(const jk_star *star, int64_t n, const double *P, const double *e,
                        const double *omega, const double *M0, const double *s,
                        const int64_t *rows, double *a_out, double *cov_out, double *ll_out,
                        void *stream)
{
    if (!star || n < 0) return JK_EINVAL;
    if (n > 0 && (!P || !e || !omega || !M0 || !a_out || !cov_out || !ll_out)) return JK_EINVAL;
    if (star->dev.jitter_mode == JK_JITTER_PHYSICAL && n > 0 && !s) return JK_EINVAL;
    if (current_device_inited() < 0) return JK_ENODEV;
    return check_hip(jk_launch_linear_posterior(star, n, P, e, omega, M0, s, rows, a_out, cov_out,
                                                ll_out, (hipStream_t)stream));
}

int jk_polish_f64(const jk_star *star, int64_t n, const double *P, const double *e,
                  const double *omega, const double *M0, const double *s, const int64_t *rows,
                  double *ll_inout, double window, int64_t first, int64_t *n_polished_out,
                  void *stream)
{
    if (star && star->dev.Qt) return JK_EINVAL;   // full covariance: FP64 only
    if (!star || n < 0 || !(window >= 0.0) || first < 0) return JK_EINVAL;
    if (n == 0) return JK_OK;
    if (!P || !e || !omega || !M0 || !ll_inout || !n_polished_out) return JK_EINVAL;
    if (star->dev.jitter_mode == JK_JITTER_PHYSICAL && !s) return JK_EINVAL;
    if (current_device_inited() < 0) return JK_ENODEV;
    hipStream_t st = (hipStream_t)stream;
    void *scratch = nullptr;
    ScratchLease lease;
    if (scratch_get(st, (size_t)jk_polish_scratch_bytes(n), (void **)&scratch, &lease) !=
        hipSuccess)
        return JK_ENOMEM;
    hipError_t err = jk_launch_polish(star, n, P, e, omega, M0, s, rows, ll_inout, window, first,
                                      scratch, n_polished_out, st);
    hipError_t err2 = scratch_put(st, scratch, lease);
    if (err != hipSuccess) return JK_EHIP;
    return check_hip(err2);
}

int jk_batch_create(const jk_star *const *stars, int n_stars, jk_batch **out)
{
    if (!out || !stars || n_stars <= 0) return JK_EINVAL;
    *out = nullptr;
    const int dev = current_device_inited();
    if (dev < 0) return JK_ENODEV;
    const int L = stars[0] ? stars[0]->dev.L : 0;
    const int jm = stars[0] ? stars[0]->dev.jitter_mode : 0;
    size_t n_epoch_doubles = 0;
    for (int i = 0; i < n_stars; ++i) {
        if (!stars[i] || stars[i]->dev.L != L || stars[i]->dev.jitter_mode != jm) return JK_EINVAL;
        if (stars[i]->dev.Qt) return JK_EINVAL;   // full covariance: single-star calls only
        n_epoch_doubles += (size_t)stars[i]->dev.n * (5 + L - 1);
    }
    jk_batch *b = new (std::nothrow) jk_batch;
    if (!b) return JK_ENOMEM;
    memset(b, 0, sizeof(*b));
    b->n_stars = n_stars;
    b->L = L;
    b->jitter_mode = jm;
    b->device = dev;
    if (hipMalloc((void **)&b->d_stars, sizeof(StarDev) * n_stars) != hipSuccess ||
        hipMalloc(&b->d_epochs, sizeof(double) * n_epoch_doubles) != hipSuccess) {
        (void)hipFree(b->d_stars);
        delete b;
        return JK_ENOMEM;
    }
    std::vector<StarDev> h(n_stars);
    double *base = (double *)b->d_epochs;
    size_t off = 0;
    for (int i = 0; i < n_stars; ++i) {
        h[i] = stars[i]->dev;
        const size_t cnt = (size_t)h[i].n * (5 + L - 1);
        if (hipMemcpy(base + off, stars[i]->d_epochs, sizeof(double) * cnt,
                      hipMemcpyDeviceToDevice) != hipSuccess) {
            (void)hipFree(b->d_stars);
            (void)hipFree(b->d_epochs);
            delete b;
            return JK_EHIP;
        }
        double *p = base + off;
        const int n = h[i].n;
        h[i].dt = p;
        h[i].rp = p + n;
        h[i].y = p + 2 * n;
        h[i].ivar = p + 3 * n;
        h[i].sig2 = p + 4 * n;
        h[i].T = p + 5 * n;
        off += cnt;
    }
    if (hipMemcpy(b->d_stars, h.data(), sizeof(StarDev) * n_stars, hipMemcpyHostToDevice) !=
        hipSuccess) {
        (void)hipFree(b->d_stars);
        (void)hipFree(b->d_epochs);
        delete b;
        return JK_EHIP;
    }
    *out = b;
    return JK_OK;
}

int jk_batch_create_slots(int capacity, int L, int jitter_mode, jk_batch **out)
{
    if (!out || capacity <= 0 || L < 1 || L > JK_LMAX) return JK_EINVAL;
    if (jitter_mode != JK_JITTER_REFERENCE && jitter_mode != JK_JITTER_PHYSICAL) return JK_EINVAL;
    *out = nullptr;
    const int dev = current_device_inited();
    if (dev < 0) return JK_ENODEV;
    jk_batch *b = new (std::nothrow) jk_batch;
    if (!b) return JK_ENOMEM;
    memset(b, 0, sizeof(*b));
    b->n_stars = capacity;
    b->L = L;
    b->jitter_mode = jitter_mode;
    b->device = dev;
    if (hipMalloc((void **)&b->d_stars, sizeof(StarDev) * capacity) != hipSuccess) {
        delete b;
        return JK_ENOMEM;
    }
    // Empty slots hold a zero-epoch star; no work item should name one.
    std::vector<StarDev> h(capacity);
    memset(h.data(), 0, sizeof(StarDev) * capacity);
    for (int i = 0; i < capacity; ++i) {
        h[i].L = L;
        h[i].jitter_mode = jitter_mode;
    }
    if (hipMemcpy(b->d_stars, h.data(), sizeof(StarDev) * capacity, hipMemcpyHostToDevice) !=
        hipSuccess) {
        (void)hipFree(b->d_stars);
        delete b;
        return JK_EHIP;
    }
    *out = b;
    return JK_OK;
}

int jk_batch_set_star(jk_batch *b, int slot, const jk_star *star, void *stream)
{
    if (!b || !star || slot < 0 || slot >= b->n_stars || b->d_epochs) return JK_EINVAL;
    if (star->dev.L != b->L || star->dev.jitter_mode != b->jitter_mode) return JK_EINVAL;
    if (star->dev.Qt) return JK_EINVAL;   // full covariance: single-star calls only
    if (star->device != b->device) return JK_EINVAL;
    if (current_device_inited() < 0) return JK_ENODEV;
    return check_hip(hipMemcpyAsync(b->d_stars + slot, &star->dev, sizeof(StarDev),
                                    hipMemcpyHostToDevice, (hipStream_t)stream));
}

int jk_batch_set_exact(jk_batch *b, int slot, int exact, void *stream)
{
    if (!b || slot < 0 || slot >= b->n_stars) return JK_EINVAL;
    if (current_device_inited() < 0) return JK_ENODEV;
    return check_hip(jk_launch_set_exact(b, slot, exact, (hipStream_t)stream));
}

int jk_batch_destroy(jk_batch *b)
{
    if (!b) return JK_EINVAL;
    int dev = 0;
    (void)hipGetDevice(&dev);
    (void)hipSetDevice(b->device);
    hipError_t e1 = hipFree(b->d_stars);
    hipError_t e2 = hipFree(b->d_epochs);
    (void)hipSetDevice(dev);
    delete b;
    return (e1 == hipSuccess && e2 == hipSuccess) ? JK_OK : JK_EHIP;
}

int jk_lnl_batch(const jk_batch *batch, int precision, int64_t n_items, const jk_work_item *items,
                 const double *P, const void *e, const void *omega, const void *M0, const void *s,
                 const int64_t *rows, double *ll_out, float *bound_out, int block, int grid,
                 void *workspace, void *stream)
{
    if (!batch || n_items < 0 || (precision != 0 && precision != 1)) return JK_EINVAL;
    if (bound_out && precision != 1) return JK_EINVAL;
    if (n_items == 0) return JK_OK;
    if (!items || !P || !e || !omega || !M0 || !ll_out) return JK_EINVAL;
    if (batch->jitter_mode == JK_JITTER_PHYSICAL && !s) return JK_EINVAL;
    if (block != 64 && block != 128 && block != 256) return JK_EINVAL;
    if (n_items > (int64_t)0xFFFFFFF0u) return JK_ERANGE;
    const int dev = current_device_inited();
    if (dev < 0) return JK_ENODEV;
    if (grid <= 0) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            return JK_EHIP;
        grid = 8 * cus;
    }
    hipStream_t st = (hipStream_t)stream;
    void *ws = workspace;
    if (!ws && hipMallocAsync(&ws, 256, st) != hipSuccess) return JK_ENOMEM;
    hipError_t err = jk_launch_lnl_batch(batch, precision, n_items, items, P, e, omega, M0, s, rows,
                                         ll_out, bound_out, block, grid, (unsigned *)ws, st);
    hipError_t err2 = workspace ? hipSuccess : hipFreeAsync(ws, st);
    if (err != hipSuccess) return JK_EHIP;
    return check_hip(err2);
}

int jk_lnl_prune_batch(const jk_batch *batch, const jk_batch *probes, const double *prune,
                       int precision, int64_t n_items, const jk_work_item *items, const double *P,
                       const void *e, const void *omega, const void *M0, const void *s,
                       double *ll_out, float *bound_out, uint64_t *n_full_out, int block,
                       int grid, void *workspace, void *stream)
{
    if (!batch || !probes || n_items < 0 || (precision != 0 && precision != 1)) return JK_EINVAL;
    if (precision == 1 && !bound_out) return JK_EINVAL;
    if (probes->L != batch->L || probes->jitter_mode != batch->jitter_mode ||
        probes->n_stars < batch->n_stars || probes->device != batch->device)
        return JK_EINVAL;
    if (n_items == 0) return JK_OK;
    if (!prune || !items || !P || !e || !omega || !M0 || !ll_out) return JK_EINVAL;
    if (batch->jitter_mode == JK_JITTER_PHYSICAL && !s) return JK_EINVAL;
    if (block != 64 && block != 128 && block != 256) return JK_EINVAL;
    if (n_items > (int64_t)0xFFFFFFF0u) return JK_ERANGE;
    const int dev = current_device_inited();
    if (dev < 0) return JK_ENODEV;
    if (grid <= 0) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            return JK_EHIP;
        grid = 8 * cus;
    }
    hipStream_t st = (hipStream_t)stream;
    void *ws = workspace;
    if (!ws && hipMallocAsync(&ws, 256, st) != hipSuccess) return JK_ENOMEM;
    hipError_t err = jk_launch_lnl_prune_batch(batch, probes, prune, precision, n_items, items, P,
                                               e, omega, M0, s, ll_out, bound_out,
                                               (unsigned long long *)n_full_out, block, grid,
                                               (unsigned *)ws, st);
    hipError_t err2 = workspace ? hipSuccess : hipFreeAsync(ws, st);
    if (err != hipSuccess) return JK_EHIP;
    return check_hip(err2);
}

static int reject_batch_impl(const double *ll, int n_seg, const jk_segment *segs,
                             int64_t *idx_out, int64_t *n_pass_out, double *llmax_out,
                             int64_t *zero_u_out, void *stream)
{
    if (!ll || n_seg <= 0 || !segs || !n_pass_out || !llmax_out) return JK_EINVAL;
    if (current_device_inited() < 0) return JK_ENODEV;
    std::vector<SegDev> h(n_seg);
    int64_t max_n = 1, max_nb = 1, mask_words = 0, bsum_n = 0;
    bool any_keep = false;
    for (int i = 0; i < n_seg; ++i) {
        const jk_segment &g = segs[i];
        if (g.n <= 0 || g.offset < 0 || g.max_keep < 0 || g.idx_offset < 0) return JK_EINVAL;
        any_keep = any_keep || g.max_keep > 0;
        SegDev &d = h[i];
        d.off = g.offset;
        d.n = g.n;
        d.max_keep = g.max_keep;
        d.idx_off = g.idx_offset;
        d.mask_off = mask_words;
        d.nb = jk_reject_bsum_count(g.n);
        d.bsum_off = bsum_n;
        d.max_off = 0;
        mask_words += jk_reject_mask_words(g.n);
        bsum_n += d.nb;
        if (g.n > max_n) max_n = g.n;
        if (d.nb > max_nb) max_nb = d.nb;
        u128 state = make_u128(g.state_hi, g.state_lo);
        d.inc = make_u128(g.inc_hi, g.inc_lo);
        jk_host_pcg_jump(state, d.inc, g.stream_offset + 1, &d.s_base);
    }
    if (any_keep && !idx_out) return JK_EINVAL;
    int pb = (int)((max_n + 4095) / 4096);
    if (pb > 1024) pb = 1024;
    if (pb < 1) pb = 1;
    hipStream_t st = (hipStream_t)stream;
    const size_t seg_bytes = sizeof(SegDev) * n_seg;
    const size_t seg_pad = ((seg_bytes + 255) / 256) * 256;
    const size_t bytes = seg_pad + sizeof(double) * (size_t)pb * n_seg +
                         sizeof(uint64_t) * mask_words + sizeof(int64_t) * bsum_n;
    char *scratch = nullptr;
    ScratchLease lease;
    if (scratch_get(st, bytes, (void **)&scratch, &lease) != hipSuccess) return JK_ENOMEM;
    SegDev *d_segs = (SegDev *)scratch;
    double *partials = (double *)(scratch + seg_pad);
    uint64_t *mask = (uint64_t *)(partials + (size_t)pb * n_seg);
    int64_t *bsum = (int64_t *)(mask + mask_words);
    hipError_t err = hipMemcpyAsync(d_segs, h.data(), seg_bytes, hipMemcpyHostToDevice, st);
    if (err == hipSuccess)
        err = jk_launch_reject_batch(ll, n_seg, d_segs, max_n, max_nb, partials, pb, mask, bsum,
                                     idx_out, n_pass_out, llmax_out, st, zero_u_out);
    // The host descriptor array must outlive the queued copy.
    hipError_t err3 = hipStreamSynchronize(st);
    hipError_t err2 = scratch_put(st, scratch, lease);
    if (err != hipSuccess || err3 != hipSuccess) return JK_EHIP;
    return check_hip(err2);
}

int jk_reject_batch(const double *ll, int n_seg, const jk_segment *segs, int64_t *idx_out,
                    int64_t *n_pass_out, double *llmax_out, void *stream)
{
    return reject_batch_impl(ll, n_seg, segs, idx_out, n_pass_out, llmax_out, nullptr, stream);
}

int jk_reject_batch_zero_check(const double *ll, int n_seg, const jk_segment *segs,
                               int64_t *idx_out, int64_t *n_pass_out, double *llmax_out,
                               int64_t *zero_u_out, void *stream)
{
    if (!zero_u_out) return JK_EINVAL;
    return reject_batch_impl(ll, n_seg, segs, idx_out, n_pass_out, llmax_out, zero_u_out, stream);
}

int jk_polish_batch(const jk_batch *batch, int n_seg, const jk_polish_segment *segs,
                    const double *P, const double *e, const double *omega, const double *M0,
                    const double *s, const int64_t *rows, double *ll_inout, double window,
                    int64_t *n_polished_out, void *stream)
{
    if (!batch || n_seg <= 0 || !segs || !ll_inout || !n_polished_out || !(window >= 0.0))
        return JK_EINVAL;
    if (!P || !e || !omega || !M0) return JK_EINVAL;
    if (batch->jitter_mode == JK_JITTER_PHYSICAL && !s) return JK_EINVAL;
    if (current_device_inited() < 0) return JK_ENODEV;
    std::vector<SegDev> h(n_seg);
    int64_t max_n = 1, max_nb = 1, mask_words = 0, bsum_n = 0, idx_n = 0;
    for (int i = 0; i < n_seg; ++i) {
        const jk_polish_segment &g = segs[i];
        if (g.n <= 0 || g.offset < 0 || g.row_begin < 0 || g.star < 0 ||
            g.star >= batch->n_stars || g.first < 0)
            return JK_EINVAL;
        SegDev &d = h[i];
        memset(&d, 0, sizeof(d));
        d.off = g.offset;
        d.n = g.n;
        d.max_keep = g.n;
        d.idx_off = idx_n;
        d.mask_off = mask_words;
        d.nb = jk_reject_bsum_count(g.n);
        d.bsum_off = bsum_n;
        d.star = g.star;
        d.row0 = g.row_begin;
        d.first = g.first;
        idx_n += g.n;
        mask_words += jk_reject_mask_words(g.n);
        bsum_n += d.nb;
        if (g.n > max_n) max_n = g.n;
        if (d.nb > max_nb) max_nb = d.nb;
    }
    int pb = (int)((max_n + 4095) / 4096);
    if (pb > 1024) pb = 1024;
    if (pb < 1) pb = 1;
    hipStream_t st = (hipStream_t)stream;
    const size_t seg_bytes = sizeof(SegDev) * n_seg;
    const size_t seg_pad = ((seg_bytes + 255) / 256) * 256;
    const size_t bytes = seg_pad + sizeof(double) * ((size_t)pb * n_seg + n_seg) +
                         sizeof(uint64_t) * mask_words +
                         sizeof(int64_t) * (bsum_n + idx_n + n_seg + 1);
    char *scratch = nullptr;
    ScratchLease lease;
    if (scratch_get(st, bytes, (void **)&scratch, &lease) != hipSuccess) return JK_ENOMEM;
    SegDev *d_segs = (SegDev *)scratch;
    double *partials = (double *)(scratch + seg_pad);
    double *llmax = partials + (size_t)pb * n_seg;
    uint64_t *mask = (uint64_t *)(llmax + n_seg);
    int64_t *bsum = (int64_t *)(mask + mask_words);
    int64_t *idx = bsum + bsum_n;
    int64_t *cum = idx + idx_n;
    hipError_t err = hipMemcpyAsync(d_segs, h.data(), seg_bytes, hipMemcpyHostToDevice, st);
    if (err == hipSuccess)
        err = jk_launch_polish_batch(batch, n_seg, d_segs, max_n, max_nb, partials, pb, mask, bsum,
                                     idx, llmax, cum, P, e, omega, M0, s, rows, ll_inout, window,
                                     n_polished_out, st);
    // The host descriptor array must outlive the queued copy.
    hipError_t err3 = hipStreamSynchronize(st);
    hipError_t err2 = scratch_put(st, scratch, lease);
    if (err != hipSuccess || err3 != hipSuccess) return JK_EHIP;
    return check_hip(err2);
}

int jk_settle_batch(const jk_batch *batch, int n_seg, const jk_settle_segment *segs,
                    const double *P, const double *e, const double *omega, const double *M0,
                    const double *s, const int64_t *rows, double *ll_inout, float *bound_inout,
                    int64_t *n_settled_out, void *stream)
{
    if (!batch || n_seg <= 0 || !segs || !ll_inout || !bound_inout || !n_settled_out)
        return JK_EINVAL;
    if (!P || !e || !omega || !M0) return JK_EINVAL;
    if (batch->jitter_mode == JK_JITTER_PHYSICAL && !s) return JK_EINVAL;
    if (current_device_inited() < 0) return JK_ENODEV;
    std::vector<SegDev> h(n_seg);
    int64_t max_n = 1, max_nb = 1, mask_words = 0, bsum_n = 0, idx_n = 0;
    for (int i = 0; i < n_seg; ++i) {
        const jk_settle_segment &g = segs[i];
        if (g.n <= 0 || g.offset < 0 || g.row_begin < 0 || g.star < 0 || g.star >= batch->n_stars)
            return JK_EINVAL;
        SegDev &d = h[i];
        memset(&d, 0, sizeof(d));
        d.off = g.offset;
        d.n = g.n;
        d.max_keep = g.n;
        d.idx_off = idx_n;
        d.mask_off = mask_words;
        d.nb = jk_reject_bsum_count(g.n);
        d.bsum_off = bsum_n;
        d.star = g.star;
        d.row0 = g.row_begin;
        u128 state = make_u128(g.state_hi, g.state_lo);
        d.inc = make_u128(g.inc_hi, g.inc_lo);
        jk_host_pcg_jump(state, d.inc, g.stream_offset + 1, &d.s_base);
        idx_n += g.n;
        mask_words += jk_reject_mask_words(g.n);
        bsum_n += d.nb;
        if (g.n > max_n) max_n = g.n;
        if (d.nb > max_nb) max_nb = d.nb;
    }
    int pb = (int)((max_n + 4095) / 4096);
    if (pb > 1024) pb = 1024;
    if (pb < 1) pb = 1;
    hipStream_t st = (hipStream_t)stream;
    const size_t seg_bytes = sizeof(SegDev) * n_seg;
    const size_t seg_pad = ((seg_bytes + 255) / 256) * 256;
    const size_t bytes = seg_pad + sizeof(double) * ((size_t)pb * n_seg + 2 * (size_t)n_seg) +
                         sizeof(uint64_t) * mask_words +
                         sizeof(int64_t) * (bsum_n + idx_n + 2 * (size_t)n_seg + 1);
    char *scratch = nullptr;
    ScratchLease lease;
    if (scratch_get(st, bytes, (void **)&scratch, &lease) != hipSuccess) return JK_ENOMEM;
    SegDev *d_segs = (SegDev *)scratch;
    double *partials = (double *)(scratch + seg_pad);
    double *lo = partials + (size_t)pb * n_seg;
    double *llmax = lo + n_seg;
    uint64_t *mask = (uint64_t *)(llmax + n_seg);
    int64_t *bsum = (int64_t *)(mask + mask_words);
    int64_t *idx = bsum + bsum_n;
    int64_t *cum = idx + idx_n;
    int64_t *count1 = cum + n_seg + 1;
    hipError_t err = hipMemcpyAsync(d_segs, h.data(), seg_bytes, hipMemcpyHostToDevice, st);
    if (err == hipSuccess)
        err = jk_launch_settle_batch(batch, n_seg, d_segs, max_n, max_nb, partials, pb, mask, bsum,
                                     idx, lo, llmax, cum, count1, P, e, omega, M0, s, rows,
                                     ll_inout, bound_inout, n_settled_out, st);
    // The host descriptor array must outlive the queued copy.
    hipError_t err3 = hipStreamSynchronize(st);
    hipError_t err2 = scratch_put(st, scratch, lease);
    if (err != hipSuccess || err3 != hipSuccess) return JK_EHIP;
    return check_hip(err2);
}

int jk_settle_reject_batch(const jk_batch *batch, int n_seg, const jk_settle_segment *segs,
                           int64_t max_keep, const double *P, const double *e,
                           const double *omega, const double *M0, const double *s,
                           const int64_t *rows, double *ll_inout, float *bound_inout,
                           int64_t *idx_out, int64_t *n_pass_out, double *llmax_out,
                           int64_t *n_settled_out, void *stream)
{
    if (!batch || n_seg <= 0 || !segs || !ll_inout || !bound_inout || !n_pass_out ||
        !llmax_out || !n_settled_out || max_keep < 0 || (max_keep > 0 && !idx_out))
        return JK_EINVAL;
    if (!P || !e || !omega || !M0) return JK_EINVAL;
    if (batch->jitter_mode == JK_JITTER_PHYSICAL && !s) return JK_EINVAL;
    if (current_device_inited() < 0) return JK_ENODEV;
    std::vector<SegDev> hs(n_seg), hr(n_seg);
    int64_t max_n = 1, max_nb = 1, mask_words = 0, bsum_n = 0, idx_n = 0, keep_n = 0;
    for (int i = 0; i < n_seg; ++i) {
        const jk_settle_segment &g = segs[i];
        if (g.n <= 0 || g.offset < 0 || g.row_begin < 0 || g.star < 0 || g.star >= batch->n_stars)
            return JK_EINVAL;
        SegDev &d = hs[i];
        memset(&d, 0, sizeof(d));
        d.off = g.offset;
        d.n = g.n;
        d.max_keep = g.n;
        d.idx_off = idx_n;
        d.mask_off = mask_words;
        d.nb = jk_reject_bsum_count(g.n);
        d.bsum_off = bsum_n;
        d.star = g.star;
        d.row0 = g.row_begin;
        u128 state = make_u128(g.state_hi, g.state_lo);
        d.inc = make_u128(g.inc_hi, g.inc_lo);
        jk_host_pcg_jump(state, d.inc, g.stream_offset + 1, &d.s_base);
        // the final compaction: the caller's max_keep and packed output
        SegDev &r = hr[i];
        r = d;
        r.max_keep = g.n < max_keep ? g.n : max_keep;
        r.idx_off = keep_n;
        keep_n += r.max_keep;
        idx_n += g.n;
        mask_words += jk_reject_mask_words(g.n);
        bsum_n += d.nb;
        if (g.n > max_n) max_n = g.n;
        if (d.nb > max_nb) max_nb = d.nb;
    }
    int pb = (int)((max_n + 4095) / 4096);
    if (pb > 1024) pb = 1024;
    if (pb < 1) pb = 1;
    hipStream_t st = (hipStream_t)stream;
    const size_t seg_bytes = sizeof(SegDev) * n_seg;
    const size_t seg_pad = ((seg_bytes + 255) / 256) * 256;
    const size_t bytes = 2 * seg_pad + sizeof(double) * ((size_t)pb * n_seg + (size_t)n_seg) +
                         2 * sizeof(uint64_t) * mask_words +
                         sizeof(int64_t) * (bsum_n + idx_n + 2 * (size_t)n_seg + 1);
    char *scratch = nullptr;
    ScratchLease lease;
    if (scratch_get(st, bytes, (void **)&scratch, &lease) != hipSuccess) return JK_ENOMEM;
    SegDev *d_settle = (SegDev *)scratch;
    SegDev *d_reject = (SegDev *)(scratch + seg_pad);
    double *partials = (double *)(scratch + 2 * seg_pad);
    double *lo = partials + (size_t)pb * n_seg;
    uint64_t *mask = (uint64_t *)(lo + n_seg);
    uint64_t *open_mask = mask + mask_words;
    int64_t *bsum = (int64_t *)(open_mask + mask_words);
    int64_t *idx = bsum + bsum_n;
    int64_t *cum = idx + idx_n;
    int64_t *count1 = cum + n_seg + 1;
    hipError_t err = hipMemcpyAsync(d_settle, hs.data(), seg_bytes, hipMemcpyHostToDevice, st);
    if (err == hipSuccess)
        err = hipMemcpyAsync(d_reject, hr.data(), seg_bytes, hipMemcpyHostToDevice, st);
    if (err == hipSuccess)
        err = jk_launch_settle_reject_batch(batch, n_seg, d_settle, d_reject, max_n, max_nb,
                                            partials, pb, mask, open_mask, bsum, idx, lo, cum,
                                            count1, P, e, omega, M0, s, rows, ll_inout,
                                            bound_inout, idx_out, n_pass_out, llmax_out,
                                            n_settled_out, st);
    // The host descriptor arrays must outlive the queued copies.
    hipError_t err3 = hipStreamSynchronize(st);
    hipError_t err2 = scratch_put(st, scratch, lease);
    if (err != hipSuccess || err3 != hipSuccess) return JK_EHIP;
    return check_hip(err2);
}

// jk_settle_live_batch, jk_settle_tol_batch and jk_settle_pruned_batch over
// the entries [first, n) of each segment.
enum { SETTLE_LIVE = 0, SETTLE_TOL = 1, SETTLE_PRUNED = 2 };

static int settle_new_rows(int mode, const jk_batch *batch, int n_seg,
                           const jk_live_segment *segs, const double *P, const double *e,
                           const double *omega, const double *M0, const double *s,
                           const int64_t *rows, double *ll_inout, float *bound_inout,
                           double margin, float tol, double window, int64_t *n_settled_out,
                           void *stream)
{
    if (!batch || n_seg <= 0 || !segs || !ll_inout || !bound_inout || !n_settled_out)
        return JK_EINVAL;
    if (!P || !e || !omega || !M0) return JK_EINVAL;
    if (batch->jitter_mode == JK_JITTER_PHYSICAL && !s) return JK_EINVAL;
    if (current_device_inited() < 0) return JK_ENODEV;
    std::vector<SegDev> h(n_seg);
    int64_t max_n = 1, max_nb = 1, mask_words = 0, bsum_n = 0, idx_n = 0;
    for (int i = 0; i < n_seg; ++i) {
        const jk_live_segment &g = segs[i];
        if (g.n < 0 || g.first < 0 || g.first > g.n || g.offset < 0 || g.row_begin < 0 ||
            g.star < 0 || g.star >= batch->n_stars)
            return JK_EINVAL;
        // the segment's new rows [first, n) as a segment of their own
        const int64_t n = g.n - g.first;
        SegDev &d = h[i];
        memset(&d, 0, sizeof(d));
        d.off = g.offset + g.first;
        d.n = n;
        d.max_keep = n;
        d.idx_off = idx_n;
        d.mask_off = mask_words;
        d.nb = jk_reject_bsum_count(n) > 0 ? jk_reject_bsum_count(n) : 1;  // block 0 writes the count
        d.bsum_off = bsum_n;
        d.star = g.star;
        d.row0 = g.row_begin + g.first;
        d.floor = g.floor;
        idx_n += n;
        mask_words += jk_reject_mask_words(n);
        bsum_n += d.nb;
        if (n > max_n) max_n = n;
        if (d.nb > max_nb) max_nb = d.nb;
    }
    int pb = (int)((max_n + 4095) / 4096);
    if (pb > 1024) pb = 1024;
    if (pb < 1) pb = 1;
    hipStream_t st = (hipStream_t)stream;
    const size_t seg_bytes = sizeof(SegDev) * n_seg;
    const size_t seg_pad = ((seg_bytes + 255) / 256) * 256;
    const size_t bytes = seg_pad + sizeof(double) * ((size_t)pb * n_seg + (size_t)n_seg) +
                         sizeof(uint64_t) * mask_words +
                         sizeof(int64_t) * (bsum_n + idx_n + 2 * (size_t)n_seg + 1);
    char *scratch = nullptr;
    ScratchLease lease;
    if (scratch_get(st, bytes, (void **)&scratch, &lease) != hipSuccess) return JK_ENOMEM;
    SegDev *d_segs = (SegDev *)scratch;
    double *partials = (double *)(scratch + seg_pad);
    double *lo = partials + (size_t)pb * n_seg;
    uint64_t *mask = (uint64_t *)(lo + n_seg);
    int64_t *bsum = (int64_t *)(mask + mask_words);
    int64_t *idx = bsum + bsum_n;
    int64_t *cum = idx + idx_n;
    int64_t *count1 = cum + n_seg + 1;
    hipError_t err = hipMemcpyAsync(d_segs, h.data(), seg_bytes, hipMemcpyHostToDevice, st);
    if (err == hipSuccess && mode == SETTLE_LIVE)
        err = jk_launch_settle_live_batch(batch, n_seg, d_segs, max_n, max_nb, partials, pb, mask,
                                          bsum, idx, lo, cum, P, e, omega, M0, s, rows, ll_inout,
                                          bound_inout, margin, n_settled_out, st);
    else if (err == hipSuccess && mode == SETTLE_TOL)
        err = jk_launch_settle_tol_batch(batch, n_seg, d_segs, max_n, max_nb, mask, bsum, idx, cum,
                                         P, e, omega, M0, s, rows, ll_inout, bound_inout, tol,
                                         n_settled_out, st);
    else if (err == hipSuccess)
        err = jk_launch_settle_pruned_batch(batch, n_seg, d_segs, max_n, max_nb, partials, pb,
                                            mask, bsum, idx, lo, cum, count1, P, e, omega, M0, s,
                                            rows, ll_inout, bound_inout, window, margin,
                                            n_settled_out, st);
    // The host descriptor array must outlive the queued copy.
    hipError_t err3 = hipStreamSynchronize(st);
    hipError_t err2 = scratch_put(st, scratch, lease);
    if (err != hipSuccess || err3 != hipSuccess) return JK_EHIP;
    return check_hip(err2);
}

int jk_settle_live_batch(const jk_batch *batch, int n_seg, const jk_live_segment *segs,
                         const double *P, const double *e, const double *omega, const double *M0,
                         const double *s, const int64_t *rows, double *ll_inout,
                         float *bound_inout, double margin, int64_t *n_settled_out, void *stream)
{
    if (!(margin >= 40.0)) return JK_EINVAL;
    return settle_new_rows(SETTLE_LIVE, batch, n_seg, segs, P, e, omega, M0, s, rows, ll_inout,
                           bound_inout, margin, 0.0f, 0.0, n_settled_out, stream);
}

int jk_settle_tol_batch(const jk_batch *batch, int n_seg, const jk_live_segment *segs,
                        const double *P, const double *e, const double *omega, const double *M0,
                        const double *s, const int64_t *rows, double *ll_inout,
                        float *bound_inout, float tol, int64_t *n_settled_out, void *stream)
{
    if (!(tol >= 0.0f)) return JK_EINVAL;
    return settle_new_rows(SETTLE_TOL, batch, n_seg, segs, P, e, omega, M0, s, rows, ll_inout,
                           bound_inout, 750.0, tol, 0.0, n_settled_out, stream);
}

int jk_settle_pruned_batch(const jk_batch *batch, int n_seg, const jk_live_segment *segs,
                           const double *P, const double *e, const double *omega,
                           const double *M0, const double *s, const int64_t *rows,
                           double *ll_inout, float *bound_inout, double window, double margin,
                           int64_t *n_settled_out, void *stream)
{
    if (!(margin >= 40.0) || !(window >= 0.0)) return JK_EINVAL;
    return settle_new_rows(SETTLE_PRUNED, batch, n_seg, segs, P, e, omega, M0, s, rows, ll_inout,
                           bound_inout, margin, 0.0f, window, n_settled_out, stream);
}

// ---------------------------------------------------------------------------
// Segment plans: the descriptors and scratch of jk_settle_batch and
// jk_reject_batch for a fixed segment layout, kept on the device, so that
// their launches allocate nothing, copy nothing from the host and never wait.
// ---------------------------------------------------------------------------
struct jk_seg_plan {
    const jk_batch *batch;
    int n_seg;
    int device;
    int64_t max_n, max_nb;
    int pb;
    std::vector<SegDev> h_settle, h_reject;   // host images (streams rewritten in place)
    SegDev *pinned;                            // staging for the stream updates
    hipEvent_t staged;                         // the last staging copy has been read
    char *mem;                                 // device: descriptors + scratch
    SegDev *d_settle, *d_reject;
    double *partials, *lo, *llmax;
    uint64_t *mask, *open_mask;
    int64_t *bsum, *idx, *cum, *count1;
};

static void plan_fill_streams(jk_seg_plan *p, const uint64_t *streams)
{
    for (int i = 0; i < p->n_seg; ++i) {
        const uint64_t *q = streams + 5 * (size_t)i;
        u128 state = make_u128(q[0], q[1]);
        u128 inc = make_u128(q[2], q[3]);
        u128 base;
        jk_host_pcg_jump(state, inc, q[4] + 1, &base);
        p->h_settle[i].s_base = p->h_reject[i].s_base = base;
        p->h_settle[i].inc = p->h_reject[i].inc = inc;
    }
}

int jk_seg_plan_destroy(jk_seg_plan *p)
{
    if (!p) return JK_EINVAL;
    int dev = 0;
    (void)hipGetDevice(&dev);
    (void)hipSetDevice(p->device);
    hipError_t e1 = p->staged ? hipEventSynchronize(p->staged) : hipSuccess;
    hipError_t e2 = p->mem ? hipFree(p->mem) : hipSuccess;
    hipError_t e3 = p->pinned ? hipHostFree(p->pinned) : hipSuccess;
    if (p->staged) (void)hipEventDestroy(p->staged);
    (void)hipSetDevice(dev);
    delete p;
    return (e1 == hipSuccess && e2 == hipSuccess && e3 == hipSuccess) ? JK_OK : JK_EHIP;
}

int jk_seg_plan_create(const jk_batch *batch, int n_seg, const jk_settle_segment *segs,
                       int64_t max_keep, jk_seg_plan **out)
{
    if (!batch || n_seg <= 0 || !segs || !out || max_keep < 0) return JK_EINVAL;
    *out = nullptr;
    const int dev = current_device_inited();
    if (dev < 0) return JK_ENODEV;
    jk_seg_plan *p = new (std::nothrow) jk_seg_plan();
    if (!p) return JK_ENOMEM;
    p->batch = batch;
    p->n_seg = n_seg;
    p->device = dev;
    p->h_settle.resize(n_seg);
    p->h_reject.resize(n_seg);
    int64_t max_n = 1, max_nb = 1, mask_words = 0, bsum_n = 0, idx_n = 0, keep_n = 0;
    std::vector<uint64_t> streams(5 * (size_t)n_seg);
    for (int i = 0; i < n_seg; ++i) {
        const jk_settle_segment &g = segs[i];
        if (g.n <= 0 || g.offset < 0 || g.row_begin < 0 || g.star < 0 ||
            g.star >= batch->n_stars) {
            delete p;
            return JK_EINVAL;
        }
        SegDev d;
        memset(&d, 0, sizeof(d));
        d.off = g.offset;
        d.n = g.n;
        d.mask_off = mask_words;
        d.nb = jk_reject_bsum_count(g.n);
        d.bsum_off = bsum_n;
        d.star = g.star;
        d.row0 = g.row_begin;
        SegDev r = d;
        d.max_keep = g.n;
        d.idx_off = idx_n;
        r.max_keep = g.n < max_keep ? g.n : max_keep;
        r.idx_off = keep_n;
        p->h_settle[i] = d;
        p->h_reject[i] = r;
        idx_n += g.n;
        keep_n += r.max_keep;
        mask_words += jk_reject_mask_words(g.n);
        bsum_n += d.nb;
        if (g.n > max_n) max_n = g.n;
        if (d.nb > max_nb) max_nb = d.nb;
        uint64_t *q = streams.data() + 5 * (size_t)i;
        q[0] = g.state_hi; q[1] = g.state_lo; q[2] = g.inc_hi; q[3] = g.inc_lo;
        q[4] = g.stream_offset;
    }
    plan_fill_streams(p, streams.data());
    int pb = (int)((max_n + 4095) / 4096);
    if (pb > 1024) pb = 1024;
    if (pb < 1) pb = 1;
    p->max_n = max_n;
    p->max_nb = max_nb;
    p->pb = pb;
    const size_t seg_bytes = sizeof(SegDev) * n_seg;
    const size_t seg_pad = ((seg_bytes + 255) / 256) * 256;
    const size_t bytes = 2 * seg_pad +
                         sizeof(double) * ((size_t)pb * n_seg + 2 * (size_t)n_seg) +
                         2 * sizeof(uint64_t) * mask_words +
                         sizeof(int64_t) * (bsum_n + idx_n + 2 * (size_t)n_seg + 1);
    if (hipMalloc((void **)&p->mem, bytes) != hipSuccess ||
        hipHostMalloc((void **)&p->pinned, 2 * seg_bytes, hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&p->staged, hipEventDisableTiming) != hipSuccess) {
        jk_seg_plan_destroy(p);
        return JK_ENOMEM;
    }
    p->d_settle = (SegDev *)p->mem;
    p->d_reject = (SegDev *)(p->mem + seg_pad);
    p->partials = (double *)(p->mem + 2 * seg_pad);
    p->lo = p->partials + (size_t)pb * n_seg;
    p->llmax = p->lo + n_seg;
    p->mask = (uint64_t *)(p->llmax + n_seg);
    p->open_mask = p->mask + mask_words;
    p->bsum = (int64_t *)(p->open_mask + mask_words);
    p->idx = p->bsum + bsum_n;
    p->cum = p->idx + idx_n;
    p->count1 = p->cum + n_seg + 1;
    if (hipMemcpy(p->d_settle, p->h_settle.data(), seg_bytes, hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipMemcpy(p->d_reject, p->h_reject.data(), seg_bytes, hipMemcpyHostToDevice) !=
            hipSuccess) {
        jk_seg_plan_destroy(p);
        return JK_EHIP;
    }
    *out = p;
    return JK_OK;
}

int jk_seg_plan_set_streams(jk_seg_plan *p, const uint64_t *streams, void *stream)
{
    if (!p || !streams) return JK_EINVAL;
    if (current_device_inited() < 0) return JK_ENODEV;
    hipStream_t st = (hipStream_t)stream;
    // the staging buffer may still be read by the previous update's copy
    if (hipEventSynchronize(p->staged) != hipSuccess) return JK_EHIP;
    plan_fill_streams(p, streams);
    const size_t seg_bytes = sizeof(SegDev) * p->n_seg;
    memcpy(p->pinned, p->h_settle.data(), seg_bytes);
    memcpy(p->pinned + p->n_seg, p->h_reject.data(), seg_bytes);
    if (hipMemcpyAsync(p->d_settle, p->pinned, seg_bytes, hipMemcpyHostToDevice, st) !=
            hipSuccess ||
        hipMemcpyAsync(p->d_reject, p->pinned + p->n_seg, seg_bytes, hipMemcpyHostToDevice,
                       st) != hipSuccess ||
        hipEventRecord(p->staged, st) != hipSuccess)
        return JK_EHIP;
    return JK_OK;
}

int jk_seg_plan_settle(const jk_seg_plan *p, const double *P, const double *e,
                       const double *omega, const double *M0, const double *s,
                       const int64_t *rows, double *ll_inout, float *bound_inout,
                       int64_t *n_settled_out, void *stream)
{
    if (!p || !P || !e || !omega || !M0 || !ll_inout || !bound_inout || !n_settled_out)
        return JK_EINVAL;
    if (p->batch->jitter_mode == JK_JITTER_PHYSICAL && !s) return JK_EINVAL;
    if (current_device_inited() < 0) return JK_ENODEV;
    return check_hip(jk_launch_settle_batch(p->batch, p->n_seg, p->d_settle, p->max_n, p->max_nb,
                                            p->partials, p->pb, p->mask, p->bsum, p->idx, p->lo,
                                            p->llmax, p->cum, p->count1, P, e, omega, M0, s, rows,
                                            ll_inout, bound_inout, n_settled_out,
                                            (hipStream_t)stream));
}

int jk_seg_plan_settle_reject(const jk_seg_plan *p, const double *P, const double *e,
                              const double *omega, const double *M0, const double *s,
                              const int64_t *rows, double *ll_inout, float *bound_inout,
                              int64_t *idx_out, int64_t *n_pass_out, double *llmax_out,
                              int64_t *n_settled_out, void *stream)
{
    if (!p || !P || !e || !omega || !M0 || !ll_inout || !bound_inout || !n_pass_out ||
        !llmax_out || !n_settled_out)
        return JK_EINVAL;
    if (p->batch->jitter_mode == JK_JITTER_PHYSICAL && !s) return JK_EINVAL;
    if (current_device_inited() < 0) return JK_ENODEV;
    return check_hip(jk_launch_settle_reject_batch(
        p->batch, p->n_seg, p->d_settle, p->d_reject, p->max_n, p->max_nb, p->partials, p->pb,
        p->mask, p->open_mask, p->bsum, p->idx, p->lo, p->cum, p->count1, P, e, omega, M0, s, rows,
        ll_inout, bound_inout, idx_out, n_pass_out, llmax_out, n_settled_out,
        (hipStream_t)stream));
}

int jk_seg_plan_reject(const jk_seg_plan *p, const double *ll, int64_t *idx_out,
                       int64_t *n_pass_out, double *llmax_out, void *stream)
{
    if (!p || !ll || !n_pass_out || !llmax_out) return JK_EINVAL;
    if (current_device_inited() < 0) return JK_ENODEV;
    return check_hip(jk_launch_reject_batch(ll, p->n_seg, p->d_reject, p->max_n, p->max_nb,
                                            p->partials, p->pb, p->mask, p->bsum, idx_out,
                                            n_pass_out, llmax_out, (hipStream_t)stream));
}

int jk_prior_fill(int64_t n, int64_t row0, uint64_t seed, double P_min, double P_max,
                  const double *e_quantiles, int n_tab, double ln_s_mu, double ln_s_sigma,
                  double *P, double *e, double *omega, double *M0, double *s, float *e32,
                  float *omega32, float *M032, float *s32, void *stream)
{
    if (n < 0 || row0 < 0 || !(P_min > 0.0) || !(P_max > P_min) || n_tab < 1 || !e_quantiles)
        return JK_EINVAL;
    if (n > 0 && (!P || !e || !omega || !M0 || !s)) return JK_EINVAL;
    const bool any32 = e32 || omega32 || M032 || s32;
    if (any32 && !(e32 && omega32 && M032 && s32)) return JK_EINVAL;
    if (current_device_inited() < 0) return JK_ENODEV;
    return check_hip(jk_launch_prior_fill(n, row0, seed, P_min, P_max, e_quantiles, n_tab, ln_s_mu,
                                          ln_s_sigma, P, e, omega, M0, s, e32, omega32, M032, s32,
                                          (hipStream_t)stream));
}

}  // extern "C"
