// rqa_batch.hip -- host side of the batched engine and its C ABI
// (accrqa_amd_rqa_batch_*, include/accrqa_amd.h).
//
// A call covers many series x (tau, emb) blocks.  Blocks with N_c <= BMAX_NC
// become problems of the persistent kernel k_batch (one problem per group of
// up to TG thresholds that recur at all): all of them in one launch per
// chunk, then k_batch_fin turns every histogram into the measures at the
// block's place in the caller's layouts.  A chunk holds as many blocks as
// its u32 histograms fit in CHUNK_BINS; per chunk the launches are a fill
// (histograms and queue counter), k_batch and k_batch_fin.  Longer series
// go through the large-N path (run_block), block by block.  Everything is
// queued on the library's stream and synchronised once, when the outputs
// come back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "accrqa_amd.h"
#include "engine.hpp"
#define RQA_HELPERS_ONLY
#include "rqa_batch.hpp"

static_assert(rqa::WG == rqa::ENG_WG, "one workgroup size for every persistent kernel");

namespace eng {
namespace {

constexpr int64_t CHUNK_BINS = (int64_t)768 << 20;   // u32 histogram bins per chunk (3 GiB)

// A (series, tau, emb) block of the call.
template <typename F>
struct SBlock {
    int64_t xoff;      // first sample in the padded batch buffer
    int64_t n, nc;
    int tau, m;
    int64_t out;       // block index in the caller's layouts
    const F *thr;      // the series' thresholds (caller order)
    std::vector<F> b;  // live bounds, sorted ascending
    std::vector<int> order;   // caller index of each sorted live bound
};

size_t align16(size_t v) { return (v + 15) & ~(size_t)15; }

// Queues one chunk of blocks [s0, s1) of `blocks` (all small, all with the
// same TG) on st: descriptors up, then fill, k_batch, k_batch_fin.
template <typename F, int NORM, int TG>
int run_chunk(const std::vector<SBlock<F>> &blocks, size_t s0, size_t s1, int T, const int *lm, int nl, int kinds,
              F *d_rr, F *d_det, F *d_lam, hipStream_t st) {
    Ctx &c = g_ctx;
    std::vector<rqa::BatchProb<F, TG>> probs;
    std::vector<int64_t> ipre(1, 0);
    std::vector<rqa::BatchBlockOut> outs;
    std::vector<longlong2> rows;
    int64_t bins = 0;
    for (size_t s = s0; s < s1; s++) {
        const SBlock<F> &B = blocks[s];
        const int live = (int)B.b.size();
        const int64_t hs = B.nc + 1;
        rqa::BatchBlockOut o;
        o.nc = (int)B.nc;
        o.rr = B.out * T;
        o.det = B.out * (int64_t)T * nl;
        o.lam = B.out * (int64_t)nl * T;
        outs.push_back(o);
        const size_t r0 = rows.size();
        rows.resize(r0 + T, longlong2{-1, -1});
        for (int g0 = 0; g0 < live; g0 += TG) {
            const int nt = std::min(TG, live - g0);
            rqa::BatchProb<F, TG> P;
            P.xoff = B.xoff;
            P.hoff = bins;
            P.nc = (int)B.nc;
            P.tau = B.tau;
            P.m = B.m;
            P.nt = nt;
            for (int t = 0; t < TG; t++) P.b[t] = B.b[g0 + std::min(t, nt - 1)];
            for (int t = 0; t < nt; t++)
                rows[r0 + B.order[g0 + t]] = longlong2{bins + (int64_t)t * hs, bins + (int64_t)(nt + t) * hs};
            bins += 2 * (int64_t)nt * hs;
            probs.push_back(P);
            const int64_t ni = ((kinds & 1) ? rqa::batch_lam_items(B.nc) : 0) +
                               ((kinds & 2) ? rqa::batch_det_items(B.nc) : 0);
            ipre.push_back(ipre.back() + ni);
        }
    }
    const int np = (int)probs.size();
    const int64_t nitems = ipre.back();
    // one descriptor blob: queue counter, ipre, problems, block outputs, rows, l_min list
    const size_t o_ipre = 16, o_probs = align16(o_ipre + ipre.size() * sizeof(int64_t));
    const size_t o_outs = align16(o_probs + probs.size() * sizeof(probs[0]));
    const size_t o_rows = align16(o_outs + outs.size() * sizeof(outs[0]));
    const size_t o_lm = align16(o_rows + rows.size() * sizeof(longlong2));
    const size_t blob = align16(o_lm + (size_t)std::max(nl, 1) * sizeof(int));
    std::vector<char> h(blob, 0);
    std::memcpy(h.data() + o_ipre, ipre.data(), ipre.size() * sizeof(int64_t));
    if (np) std::memcpy(h.data() + o_probs, probs.data(), probs.size() * sizeof(probs[0]));
    std::memcpy(h.data() + o_outs, outs.data(), outs.size() * sizeof(outs[0]));
    std::memcpy(h.data() + o_rows, rows.data(), rows.size() * sizeof(longlong2));
    if (nl > 0) std::memcpy(h.data() + o_lm, lm, (size_t)nl * sizeof(int));
    int e;
    if ((e = ensure(c.bdesc, blob))) return e;
    if ((e = ensure(c.bhist, (size_t)std::max<int64_t>(bins, 1) * sizeof(uint32_t)))) return e;
    char *d = (char *)c.bdesc.p;
    if ((e = stage_upload(d + 16, h.data() + 16, blob - 16, st))) return e;
    uint32_t *hist = (uint32_t *)c.bhist.p;
    unsigned long long *queue = (unsigned long long *)d;
    const unsigned fb = (unsigned)std::min<int64_t>((bins + 255) / 256 + 1, 8192);
    hipLaunchKernelGGL(rqa::k_fill_zero<uint32_t>, dim3(fb), dim3(256), 0, st, hist, (size_t)bins, queue);
    HIP_TRY(hipGetLastError());
    if (nitems > 0) {
        const unsigned g = grid_for(rqa::k_batch<F, NORM, TG>, nitems, c.ncu);
        hipLaunchKernelGGL((rqa::k_batch<F, NORM, TG>), dim3(g), dim3(rqa::WG), 0, st, (const F *)c.x.p,
                           (const rqa::BatchProb<F, TG> *)(d + o_probs), (const int64_t *)(d + o_ipre), np, kinds,
                           queue, hist);
        HIP_TRY(hipGetLastError());
    }
    const int64_t nfin = (int64_t)outs.size() * T;
    hipLaunchKernelGGL((rqa::k_batch_fin<F>), dim3((unsigned)nfin), dim3(rqa::BFIN), 0, st, (const uint32_t *)hist,
                       (const rqa::BatchBlockOut *)(d + o_outs), (const longlong2 *)(d + o_rows), T,
                       (const int *)(d + o_lm), nl, d_rr, d_det, d_lam);
    HIP_TRY(hipGetLastError());
    return E_OK;
}

template <typename F, int NORM>
int run_small(const std::vector<SBlock<F>> &blocks, int T, const int *lm, int nl, int kinds, F *d_rr, F *d_det,
              F *d_lam, hipStream_t st) {
    int maxlive = 0;
    for (const SBlock<F> &B : blocks) maxlive = std::max(maxlive, (int)B.b.size());
    const int tg = maxlive <= 8 ? 8 : 16;
    size_t s0 = 0;
    while (s0 < blocks.size()) {
        // blocks up to the chunk's histogram budget (at least one)
        int64_t bins = 0;
        size_t s1 = s0;
        while (s1 < blocks.size()) {
            const int64_t bb = 2 * (int64_t)blocks[s1].b.size() * (blocks[s1].nc + 1);
            if (s1 > s0 && bins + bb > CHUNK_BINS) break;
            bins += bb;
            s1++;
        }
        const int e = tg == 8 ? run_chunk<F, NORM, 8>(blocks, s0, s1, T, lm, nl, kinds, d_rr, d_det, d_lam, st)
                              : run_chunk<F, NORM, 16>(blocks, s0, s1, T, lm, nl, kinds, d_rr, d_det, d_lam, st);
        if (e) return e;
        s0 = s1;
    }
    return E_OK;
}

// Many series in one call: every series uploaded once (each followed by its
// own PAD zeros, which the kernels' windows read past its end); blocks with
// N_c <= BMAX_NC through the persistent batched kernels, longer ones block by
// block through the large-N path; every block writes its part of device
// outputs, which come back in one copy each.  Thresholds are per series
// ([nseries][T]); outputs are the per-series reference layouts stacked:
// rr [s][a][b][T], det [s][a][b][T][nl][5], lam [s][a][b][nl][T][5].
template <typename F>
void host_batch(const F *series, const size_t *offsets, int nseries, const int *tau, int ntau, const int *emb,
                int nemb, const int *lm, int nl, const F *thr, int T, int dist, F *rr_out, F *det_out, F *lam_out,
                int *error) {
    if (!error) return;
    *error = E_OK;
    const bool want_l = det_out || lam_out;
    if (nseries <= 0 || ntau <= 0 || nemb <= 0 || T <= 0 || (want_l && nl <= 0)) *error = E_INVALID;
    if (!series || !offsets || !tau || !emb || !thr || (want_l && !lm) || (!rr_out && !det_out && !lam_out))
        *error = E_NULL;
    if (*error) return;
    int e = E_OK;
    size_t total = 0;
    for (int k = 0; k < nseries && !e; k++) {
        if (offsets[k + 1] <= offsets[k]) e = E_INVALID;
        if (!e) e = check_args(offsets[k + 1] - offsets[k], tau, ntau, emb, nemb, lm, want_l ? nl : 0, dist);
        total += offsets[k + 1] - offsets[k] + PAD;
    }
    if (e) {
        *error = e;
        return;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if ((e = init_ctx())) {
        *error = e;
        return;
    }
    Ctx &c = g_ctx;
    hipStream_t st = c.stream;
    CallScope call(st);
    if (call.err) {
        *error = call.err;
        return;
    }
    std::vector<F> host(total, F(0));
    std::vector<size_t> poff(nseries);
    size_t p = 0;
    for (int k = 0; k < nseries; k++) {
        poff[k] = p;
        std::copy(series + offsets[k], series + offsets[k + 1], host.begin() + (ptrdiff_t)p);
        p += offsets[k + 1] - offsets[k] + PAD;
    }
    const size_t nblocks = (size_t)nseries * ntau * nemb;
    const size_t per_rr = (size_t)T, per_dl = (size_t)T * nl * 5;
    if ((e = ensure(c.x, total * sizeof(F))) || (e = ensure(c.out_rr, nblocks * per_rr * sizeof(F))) ||
        (e = ensure(c.out_det, nblocks * (want_l ? per_dl : 1) * sizeof(F))) ||
        (e = ensure(c.out_lam, nblocks * (want_l ? per_dl : 1) * sizeof(F)))) {
        *error = e;
        return;
    }
    hipError_t he = hipMemcpyAsync(c.x.p, host.data(), total * sizeof(F), hipMemcpyHostToDevice, st);
    if (he == hipSuccess) he = hipStreamSynchronize(st);
    if (he != hipSuccess) {
        *error = map_err(he);
        return;
    }
    F *d_rr = rr_out ? (F *)c.out_rr.p : nullptr;
    F *d_det = det_out ? (F *)c.out_det.p : nullptr;
    F *d_lam = lam_out ? (F *)c.out_lam.p : nullptr;

    // the blocks: bounds per series (shared by its (tau, emb) blocks), small
    // ones ordered longest first for the work queue
    std::vector<SBlock<F>> small;
    std::vector<size_t> large;   // block indices for the large-N path
    std::vector<std::vector<F>> sb_bounds(nseries);
    std::vector<std::vector<int>> sb_order(nseries);
    for (int k = 0; k < nseries; k++) {
        const F *tk = thr + (size_t)k * T;
        std::vector<F> bnd(T);
        std::vector<int> live;
        for (int t = 0; t < T; t++) {
            const F bv = bound_of(tk[t], dist);
            if (!never_recurs(bv)) {
                bnd[t] = bv == F(0) ? F(0) : bv;
                live.push_back(t);
            }
        }
        std::stable_sort(live.begin(), live.end(), [&](int a, int b) { return bnd[a] < bnd[b]; });
        for (int t : live) sb_bounds[k].push_back(bnd[t]);
        sb_order[k] = live;
    }
    for (int k = 0; k < nseries; k++)
        for (int a = 0; a < ntau; a++)
            for (int b = 0; b < nemb; b++) {
                const size_t blk = ((size_t)k * ntau + a) * nemb + b;
                const int64_t n = (int64_t)(offsets[k + 1] - offsets[k]);
                const int64_t nc = n - (int64_t)(emb[b] - 1) * tau[a];
                if (nc > rqa::BMAX_NC) {
                    large.push_back(blk);
                    continue;
                }
                SBlock<F> B;
                B.xoff = (int64_t)poff[k];
                B.n = n;
                B.nc = nc;
                B.tau = tau[a];
                B.m = emb[b];
                B.out = (int64_t)blk;
                B.thr = thr + (size_t)k * T;
                B.b = sb_bounds[k];
                B.order = sb_order[k];
                small.push_back(std::move(B));
            }
    std::stable_sort(small.begin(), small.end(), [](const SBlock<F> &x, const SBlock<F> &y) { return x.nc > y.nc; });
    const int kinds = ((rr_out || lam_out) ? 1 : 0) | (det_out ? 2 : 0);
    if (!small.empty()) {
        e = dist == 1 ? run_small<F, 1>(small, T, lm, want_l ? nl : 0, kinds, d_rr, d_det, d_lam, st)
                      : run_small<F, 2>(small, T, lm, want_l ? nl : 0, kinds, d_rr, d_det, d_lam, st);
    }
    for (size_t i = 0; i < large.size() && !e; i++) {
        const size_t blk = large[i];
        const size_t k = blk / ((size_t)ntau * nemb), a = (blk / nemb) % ntau, b = blk % nemb;
        e = run_block<F>((const F *)c.x.p + poff[k], offsets[k + 1] - offsets[k], tau[a], emb[b],
                         thr + (size_t)k * T, T, want_l ? lm : nullptr, want_l ? nl : 0, dist,
                         d_rr ? d_rr + blk * per_rr : nullptr, d_det ? d_det + blk * per_dl : nullptr,
                         d_lam ? d_lam + blk * per_dl : nullptr, st);
    }
    if (e) {
        (void)hipStreamSynchronize(st);
        *error = e;
        return;
    }
    he = hipSuccess;
    if (rr_out) he = hipMemcpyAsync(rr_out, c.out_rr.p, nblocks * per_rr * sizeof(F), hipMemcpyDeviceToHost, st);
    if (he == hipSuccess && det_out)
        he = hipMemcpyAsync(det_out, c.out_det.p, nblocks * per_dl * sizeof(F), hipMemcpyDeviceToHost, st);
    if (he == hipSuccess && lam_out)
        he = hipMemcpyAsync(lam_out, c.out_lam.p, nblocks * per_dl * sizeof(F), hipMemcpyDeviceToHost, st);
    if (he == hipSuccess) he = hipStreamSynchronize(st);
    if (he != hipSuccess) *error = map_err(he);
}

}  // namespace
}  // namespace eng

extern "C" {

void accrqa_amd_rqa_batch_f32(const float *series, const size_t *offsets, int nSeries, const int *tau_values,
                              int nTaus, const int *emb_values, int nEmbs, const int *lmin_values, int nL,
                              const float *threshold_values, int nThresholds, int distance_type, float *rr_out,
                              float *det_out, float *lam_out, int *error) {
    eng::host_batch<float>(series, offsets, nSeries, tau_values, nTaus, emb_values, nEmbs, lmin_values, nL,
                           threshold_values, nThresholds, distance_type, rr_out, det_out, lam_out, error);
}
void accrqa_amd_rqa_batch_f64(const double *series, const size_t *offsets, int nSeries, const int *tau_values,
                              int nTaus, const int *emb_values, int nEmbs, const int *lmin_values, int nL,
                              const double *threshold_values, int nThresholds, int distance_type, double *rr_out,
                              double *det_out, double *lam_out, int *error) {
    eng::host_batch<double>(series, offsets, nSeries, tau_values, nTaus, emb_values, nEmbs, lmin_values, nL,
                            threshold_values, nThresholds, distance_type, rr_out, det_out, lam_out, error);
}

}  // extern "C"

