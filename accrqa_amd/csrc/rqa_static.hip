// rqa_static.hip -- the no-allocation variant of accrqa_amd_rqa_device_f32
// (include/accrqa_amd.h), on the batched engine's kernels (rqa_batch.hpp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "accrqa_amd.h"
#include "engine.hpp"
#define RQA_HELPERS_ONLY
#include "rqa_batch.hpp"

// ================================================================ no-allocation entry point
// accrqa_amd_rqa_device_static_f32: accrqa_amd_rqa_device_f32 for callers that
// cannot allocate device memory (code under stream capture, serving loops,
// harnesses that forbid hipMalloc).  Its work memory is the file-scope device
// array `scratch` below, one per GPU for the life of the process; the call
// allocates nothing and copies nothing from the host: every per-call table
// travels by value in a kernel argument, a setup kernel writes it into the
// scratch array with the padded copy of the series and zeroes the histograms
// and the work-queue counter, then the batched kernels run (k_batch,
// k_batch_fin) and write the caller's device outputs.  Limits: series of at
// most NA_MAX_N samples, NA_MAX_T thresholds, NA_MAX_L l_min values.  Calls
// that overlap on one GPU share its array and race.
namespace eng {
namespace {

constexpr int NA_MAX_N = 32768;   // samples of the series
constexpr int NA_MAX_T = 32;      // thresholds
constexpr int NA_MAX_L = 16;      // l_min values
constexpr int NA_TG = 16;         // thresholds per problem of k_batch
constexpr int NA_MAX_P = NA_MAX_T / NA_TG;
constexpr int NA_MAX_DEV = 64;
// byte offsets in the scratch array
constexpr size_t NA_OFF_IPRE = 256;
constexpr size_t NA_OFF_PROBS = 512;
constexpr size_t NA_OFF_OUTS = 1024;
constexpr size_t NA_OFF_ROWS = 1280;
constexpr size_t NA_OFF_LMIN = 1792;
constexpr size_t NA_OFF_X = 2048;
constexpr size_t NA_X_BYTES = (size_t)(NA_MAX_N + PAD) * sizeof(float);
constexpr size_t NA_OFF_HIST = NA_OFF_X + NA_X_BYTES;
constexpr size_t NA_HIST_BINS = (size_t)2 * NA_MAX_T * (NA_MAX_N + 1);
constexpr size_t NA_SCRATCH_BYTES = NA_OFF_HIST + NA_HIST_BINS * sizeof(uint32_t);
static_assert(sizeof(rqa::BatchProb<float, NA_TG>) * NA_MAX_P <= NA_OFF_OUTS - NA_OFF_PROBS, "problem slots");
static_assert((NA_MAX_P + 1) * sizeof(int64_t) <= NA_OFF_PROBS - NA_OFF_IPRE, "item prefix slots");
static_assert(NA_MAX_T * sizeof(longlong2) <= NA_OFF_LMIN - NA_OFF_ROWS, "row slots");
static_assert(NA_MAX_L * sizeof(int) <= NA_OFF_X - NA_OFF_LMIN, "l_min slots");
static_assert(NA_OFF_X % 256 == 0 && NA_OFF_HIST % 256 == 0, "aligned sections");

// Everything a call needs besides the series, by value (well under the
// kernel-argument limit).
struct NaArgs {
    const float *src;
    int64_t n;
    int np, T, nl;
    int64_t ipre[NA_MAX_P + 1];
    rqa::BatchProb<float, NA_TG> probs[NA_MAX_P];
    rqa::BatchBlockOut out;
    longlong2 rows[NA_MAX_T];
    int lmins[NA_MAX_L];
    int64_t bins;
};

// Writes the tables and the padded series into the scratch array and zeroes
// the queue counter and the histograms.
__global__ void k_na_setup(NaArgs A, unsigned char *__restrict__ scr) {
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    if (tid == 0) *(unsigned long long *)scr = 0ull;
    if (tid <= (size_t)A.np) ((int64_t *)(scr + NA_OFF_IPRE))[tid] = A.ipre[tid];
    if (tid < (size_t)A.np) ((rqa::BatchProb<float, NA_TG> *)(scr + NA_OFF_PROBS))[tid] = A.probs[tid];
    if (tid == 0) *(rqa::BatchBlockOut *)(scr + NA_OFF_OUTS) = A.out;
    if (tid < (size_t)A.T) ((longlong2 *)(scr + NA_OFF_ROWS))[tid] = A.rows[tid];
    if (tid < (size_t)A.nl) ((int *)(scr + NA_OFF_LMIN))[tid] = A.lmins[tid];
    float *x = (float *)(scr + NA_OFF_X);
    for (size_t i = tid; i < (size_t)A.n + PAD; i += stride) x[i] = i < (size_t)A.n ? A.src[i] : 0.0f;
    uint32_t *h = (uint32_t *)(scr + NA_OFF_HIST);
    for (size_t i = tid; i < (size_t)A.bins; i += stride) h[i] = 0u;
}

// Per device: the scratch array's device address (hipGetSymbolAddress on the
// first call with that device current), the CU count and the resident
// workgroups per CU of the two k_batch instantiations.
struct NaDev {
    bool ready = false;
    unsigned char *ptr = nullptr;
    int ncu = 0;
    int occ[2] = {1, 1};
};
NaDev g_na[NA_MAX_DEV];

}  // namespace
}  // namespace eng

// The no-allocation entry point's work memory (see above); host code uses its
// device address from hipGetSymbolAddress, never this name.
static __device__ __align__(256) unsigned char scratch[eng::NA_SCRATCH_BYTES];

namespace eng {
namespace {

int na_device(NaDev *&out) {
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= NA_MAX_DEV) {
        g_last_error = "device index beyond the no-allocation entry point's table";
        return E_NO_DEVICE;
    }
    NaDev &D = g_na[dev];
    if (!D.ready) {
        void *p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        HIP_TRY(hipDeviceGetAttribute(&D.ncu, hipDeviceAttributeMultiprocessorCount, dev));
        int nb = 0;
        HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &nb, reinterpret_cast<const void *>(rqa::k_batch<float, 1, NA_TG>), rqa::WG, 0));
        D.occ[0] = std::max(nb, 1);
        HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &nb, reinterpret_cast<const void *>(rqa::k_batch<float, 2, NA_TG>), rqa::WG, 0));
        D.occ[1] = std::max(nb, 1);
        D.ptr = (unsigned char *)p;
        D.ready = true;
    }
    out = &D;
    return E_OK;
}

int device_static(const float *d_in, size_t n, int tau, int emb, const float *thr, int T, const int *lm, int nl,
                  int dist, float *d_rr, float *d_det, float *d_lam, void *stream) {
    const bool want_l = d_det || d_lam;
    if (n == 0 || T <= 0 || (want_l && nl <= 0)) return E_INVALID;
    if (!d_in || !thr || (want_l && !lm) || (!d_rr && !d_det && !d_lam)) return E_NULL;
    int e = check_args(n, &tau, 1, &emb, 1, lm, want_l ? nl : 0, dist);
    if (e) return e;
    if (n > (size_t)NA_MAX_N || T > NA_MAX_T || (want_l && nl > NA_MAX_L)) {
        g_last_error = "accrqa_amd_rqa_device_static_f32: beyond its fixed work memory (at most 32768 samples, "
                       "32 thresholds, 16 l_min values)";
        return E_INVALID;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    NaDev *D = nullptr;
    if ((e = na_device(D))) return e;
    hipStream_t st = (hipStream_t)stream;
    const int64_t nc = (int64_t)n - (int64_t)(emb - 1) * tau;
    const int64_t hs = nc + 1;
    NaArgs A;
    std::memset(&A, 0, sizeof(A));
    A.src = d_in;
    A.n = (int64_t)n;
    A.T = T;
    A.nl = want_l ? nl : 0;
    for (int l = 0; l < A.nl; l++) A.lmins[l] = lm[l];
    A.out.nc = (int)nc;
    A.out.rr = A.out.det = A.out.lam = 0;
    // live thresholds sorted by bound, in problems of up to NA_TG
    float bnd[NA_MAX_T];
    int live[NA_MAX_T], nlive = 0;
    for (int t = 0; t < T; t++) {
        A.rows[t] = longlong2{-1, -1};
        const float b = bound_of<float>(thr[t], dist);
        if (never_recurs(b)) continue;
        bnd[t] = b == 0.0f ? 0.0f : b;
        live[nlive++] = t;
    }
    std::stable_sort(live, live + nlive, [&](int a, int b) { return bnd[a] < bnd[b]; });
    const int kinds = ((d_rr || d_lam) ? 1 : 0) | (d_det ? 2 : 0);
    int64_t bins = 0;
    A.ipre[0] = 0;
    for (int g0 = 0; g0 < nlive; g0 += NA_TG) {
        const int nt = std::min(NA_TG, nlive - g0);
        rqa::BatchProb<float, NA_TG> &P = A.probs[A.np];
        P.xoff = 0;
        P.hoff = bins;
        P.nc = (int)nc;
        P.tau = tau;
        P.m = emb;
        P.nt = nt;
        for (int t = 0; t < NA_TG; t++) P.b[t] = bnd[live[g0 + std::min(t, nt - 1)]];
        for (int t = 0; t < nt; t++) A.rows[live[g0 + t]] = longlong2{bins + (int64_t)t * hs, bins + (int64_t)(nt + t) * hs};
        bins += 2 * (int64_t)nt * hs;
        A.ipre[A.np + 1] = A.ipre[A.np] + ((kinds & 1) ? rqa::batch_lam_items(nc) : 0) +
                           ((kinds & 2) ? rqa::batch_det_items(nc) : 0);
        A.np++;
    }
    A.bins = bins;
    unsigned char *scr = D->ptr;
    const int64_t work = std::max<int64_t>(bins, (int64_t)n + PAD);
    const unsigned sb = (unsigned)std::min<int64_t>((work + 255) / 256, 2048);
    hipLaunchKernelGGL(k_na_setup, dim3(sb), dim3(256), 0, st, A, scr);
    HIP_TRY(hipGetLastError());
    const int64_t nitems = A.ipre[A.np];
    if (nitems > 0) {
        const int64_t want = (nitems + rqa::WAVES - 1) / rqa::WAVES;
        const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>(want, (int64_t)D->occ[dist - 1] * D->ncu));
        const float *x = (const float *)(scr + NA_OFF_X);
        const auto *probs = (const rqa::BatchProb<float, NA_TG> *)(scr + NA_OFF_PROBS);
        const int64_t *ipre = (const int64_t *)(scr + NA_OFF_IPRE);
        auto *queue = (unsigned long long *)scr;
        auto *hist = (uint32_t *)(scr + NA_OFF_HIST);
        if (dist == 1)
            hipLaunchKernelGGL((rqa::k_batch<float, 1, NA_TG>), dim3(g), dim3(rqa::WG), 0, st, x, probs, ipre, A.np,
                               kinds, queue, hist);
        else
            hipLaunchKernelGGL((rqa::k_batch<float, 2, NA_TG>), dim3(g), dim3(rqa::WG), 0, st, x, probs, ipre, A.np,
                               kinds, queue, hist);
        HIP_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL((rqa::k_batch_fin<float>), dim3((unsigned)T), dim3(rqa::BFIN), 0, st,
                       (const uint32_t *)(scr + NA_OFF_HIST), (const rqa::BatchBlockOut *)(scr + NA_OFF_OUTS),
                       (const longlong2 *)(scr + NA_OFF_ROWS), T, (const int *)(scr + NA_OFF_LMIN), A.nl, d_rr,
                       d_det, d_lam);
    HIP_TRY(hipGetLastError());
    return E_OK;
}

}  // namespace
}  // namespace eng

extern "C" int accrqa_amd_rqa_device_static_f32(const float *d_input, size_t data_size, int tau, int emb,
                                                const float *threshold_values, int nThresholds,
                                                const int *lmin_values, int nL, int distance_type, float *d_rr,
                                                float *d_det, float *d_lam, void *stream) {
    return eng::device_static(d_input, data_size, tau, emb, threshold_values, nThresholds, lmin_values, nL,
                              distance_type, d_rr, d_det, d_lam, stream);
}
