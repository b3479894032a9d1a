// rqa_batch.hpp -- gfx950 kernels of the batched engine: many small problems
// (series x tau x emb x threshold group, N_c <= BMAX_NC) in one persistent
// launch fed by a device work queue, then one launch that turns every
// histogram into the reference's measures.
//
// A problem's work items are its LAM bands (64 full columns, lane = column,
// rows 0 .. N_c-1) followed by its DET strips (64 whole upper diagonals,
// lane = diagonal).  A sequence is never cut, so there are no segment carries
// and no stitch: every run closes inside its item.  Items are handed out by
// one atomic counter, problems being ordered longest first, so the waves of
// the persistent grid finish together whatever the mix of lengths.
//
// Decisions, ranks and the word-level line scan are those of the large-N
// kernels (dim_step, pack_ranks, rank_planes, GeTree, consume; exact-order
// sums, rqa_device.hpp), so every histogram is the reference's bit for bit.
// Each wave keeps the run-length histogram of its current (problem, kind) in
// its own LDS (lengths < BLCAP) and adds it to the problem's u32 histogram in
// HBM when it moves on; longer runs go straight to HBM.  The bucket table of
// the ranks is built per wave from the problem's sorted bounds.
#pragma once
#include "rqa_device.hpp"

namespace rqa {

constexpr int BMAX_NC = 65535;   // u32 bins cannot overflow: every count is below N_c^2 < 2^32
constexpr int BNB = 128;         // bucket table entries per wave
constexpr int BWIN = 256;        // DET column window per wave (samples)
// Run lengths below BLCAP go to the wave's LDS histogram.  Smaller than the
// large-N kernels' LCAP: a wave's LDS (histogram with a spare bin per lane and
// row, table, window) is 6.0 KB for 8 thresholds, so LDS allows 6 waves per
// SIMD, above what the registers allow; the short series of a batch rarely
// have runs this long.
constexpr int BLCAP = 64;
constexpr int BNSPARE = WAVE;               // spare bins per row, one per lane
constexpr int BLROW = BLCAP + BNSPARE;      // LDS words per threshold row

// One problem: a series at x + xoff (n = N_c + (m-1) tau samples, followed
// by at least PAD zeros in the batch buffer), one (tau, emb), nt <= TG live
// thresholds as sorted exact bounds.  Histograms (u32, N_c + 1 bins per row):
// h_V row t at hoff + t (N_c + 1), h_D row t at hoff + (nt + t)(N_c + 1).
template <typename F, int TG>
struct BatchProb {
    int64_t xoff;
    int64_t hoff;
    int nc, tau, m, nt;
    F b[TG];
};

// Items of a problem: ceil(N_c / 64) LAM bands, then ceil((N_c - 1) / 64)
// DET strips.
__host__ __device__ __forceinline__ int64_t batch_lam_items(int64_t nc) { return (nc + WAVE - 1) / WAVE; }
__host__ __device__ __forceinline__ int64_t batch_det_items(int64_t nc) { return (nc - 1 + WAVE - 1) / WAVE; }

__device__ __forceinline__ uint32_t bkey_hi(float v) { return __float_as_uint(v); }
__device__ __forceinline__ uint32_t bkey_hi(double v) {
    return (uint32_t)((unsigned long long)__double_as_longlong(v) >> 32);
}
__device__ __forceinline__ uint32_t bkey(float v) { return __float_as_uint(v); }
__device__ __forceinline__ unsigned long long bkey(double v) { return (unsigned long long)__double_as_longlong(v); }
// First key of the bucket whose high key bits start at hi.
__device__ __forceinline__ uint32_t bkey_at(float, unsigned long long hi) { return (uint32_t)hi; }
__device__ __forceinline__ unsigned long long bkey_at(double, unsigned long long hi) {
    return hi >= (1ull << 32) ? ~0ull : (hi << 32);
}

// The bucket table of rank_of for sorted non-negative bounds b[0 .. nt), built
// by the 64 lanes of one wave: bucket bk covers high key bits
// [base + bk<<sh, base + (bk+1)<<sh); the last bucket lies past the largest
// bound.  Entry: the key of the first bound at or above the bucket's start
// (all-ones if none), the number of bounds below the start, and a flag if the
// bucket holds two or more bounds (pack_ranks then counts directly).
template <typename F, int TG>
__device__ __forceinline__ LutParams batch_lut(const F *b, int nt, LutEntry<F> *lut) {
    const int lane = threadIdx.x & (WAVE - 1);
    LutParams P;
    P.base = bkey_hi(b[0]);
    const uint32_t range = bkey_hi(b[nt - 1]) - P.base;
    int sh = 0;
    while ((range >> sh) > (uint32_t)(BNB - 2)) sh++;
    P.sh = sh;
    P.nb = (int)(range >> sh) + 2;
    for (int bk = lane; bk < P.nb; bk += WAVE) {
        const unsigned long long s_hi = (unsigned long long)P.base + ((unsigned long long)bk << sh);
        const auto start = bkey_at(F(0), s_hi);
        const auto end = bkey_at(F(0), s_hi + (1ull << sh));
        int r_lo = 0, inside = 0;
#pragma unroll
        for (int t = 0; t < TG; t++) {
            if (t < nt) {
                const auto k = bkey(b[t]);
                if (k < start) r_lo++;
                else if (k < end) inside++;
            }
        }
        LutEntry<F> e;
        e.next = r_lo < nt ? (decltype(e.next))bkey(b[r_lo]) : (decltype(e.next))~0ull;
        e.info = (uint32_t)r_lo | (inside >= 2 ? 0x100u : 0u);
        if constexpr (sizeof(F) == 8) e.pad = 0u;
        lut[bk] = e;
    }
    return P;
}

// Per-wave LDS of the batched kernel.
template <typename F, int TG>
struct BatchWave {
    uint32_t lh[TG * BLROW];   // run lengths < BLCAP of the current (problem, kind), row t at t * BLROW
    LutEntry<F> lut[BNB];
    F bnd[TG];
    F win[BWIN];
};

// The word scan of one 32-step block for every threshold of the group, with
// runs that start at the sequence's first step counted like any other (the
// sequence is never cut) and runs of BLCAP or more added to row t of `hrow`
// (u32) straight away.
template <int TG>
__device__ __forceinline__ void batch_scan(const uint32_t (&R)[8], uint32_t keep, uint32_t force,
                                           uint32_t (&run)[TG], int nt, uint32_t *lh, uint32_t spare,
                                           uint32_t *hrow, int64_t hs) {
    constexpr int B = rank_bits<TG>();
    uint32_t Pl[B];
    rank_planes<B>(R, Pl, keep, force);
    const GeTree<B> G(Pl);
    bool lead_in = true;
#pragma unroll
    for (int t = 0; t < TG; t++) {
        if (t < nt) {
            const uint32_t nw = G.ge(t + 1);
            if (lead_in) {
                if (__builtin_amdgcn_ballot_w64((~nw | run[t]) != 0u) == 0ull) continue;
                lead_in = false;
            }
            uint32_t *lht = lh + t * BLROW;
            consume<BLCAP>(
                nw, run[t], spare, [&](uint32_t b) { lds_inc(&lht[b + 1]); },
                [&](uint32_t len) { lds_inc(&lht[len]); },
                [&](uint32_t first) {
                    const uint32_t tt = opaque(t);
                    if ((int64_t)first < hs) atomicAdd(&hrow[(size_t)tt * hs + first], 1u);
                });
        }
    }
}

// The runs still open at the end of a sequence.
template <int TG>
__device__ __forceinline__ void batch_close(const uint32_t (&run)[TG], int nt, uint32_t *lh, uint32_t *hrow,
                                            int64_t hs) {
#pragma unroll
    for (int t = 0; t < TG; t++) {
        const uint32_t len = run[t];
        if (t < nt && len) {
            if (len < (uint32_t)BLCAP)
                lds_inc(&lh[t * BLROW + len]);
            else if ((int64_t)len < hs)
                atomicAdd(&hrow[(size_t)opaque(t) * hs + len], 1u);
        }
    }
}

// Adds the wave's LDS histogram to rows 0 .. nt-1 of hrow and clears it.
template <int TG>
__device__ __forceinline__ void batch_flush(uint32_t *lh, int nt, uint32_t *hrow, int64_t hs) {
    const int lane = threadIdx.x & (WAVE - 1);
    wave_sync();
    for (int e = lane; e < TG * BLCAP; e += WAVE) {
        const int t = e / BLCAP, len = e - t * BLCAP;
        const uint32_t v = lh[t * BLROW + len];
        if (v) {
            if (len && t < nt && len < hs) atomicAdd(&hrow[(size_t)t * hs + len], v);
            lh[t * BLROW + len] = 0u;
        }
    }
    wave_sync();
}

// The persistent kernel.  queue: the item counter (zero before the launch);
// ipre[p] = the first item of problem p (np + 1 entries); kinds: bit 0 = the
// LAM bands are items, bit 1 = the DET strips are; hist: the u32 histograms
// of every problem.  WPE: the minimum waves per SIMD the register allocation
// must allow (1 = the compiler's choice); the host picks the fastest
// instantiation by measurement.
template <typename F, int NORM, int TG, int WPE = 1>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(WPE))) void k_batch(const F *__restrict__ x, const BatchProb<F, TG> *__restrict__ probs,
                                              const int64_t *__restrict__ ipre, int np, int kinds,
                                              unsigned long long *__restrict__ queue, uint32_t *__restrict__ hist) {
    __shared__ BatchWave<F, TG> wv[WAVES];
    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = wave_id();
    BatchWave<F, TG> &Wv = wv[wave];
    for (int e = lane; e < TG * BLROW; e += WAVE) Wv.lh[e] = 0u;
    const int64_t nitems = ipre[np];
    constexpr int MJ = 8;   // column samples kept in registers (LAM); dimensions past them load per block

    int cur = -1, cur_kind = -1;   // (problem, kind) whose runs the LDS histogram holds
    uint32_t *cur_row = nullptr;
    int64_t cur_hs = 0;
    int cur_nt = 0;
    int p = 0;
    LutParams P{0u, 0, 1};
    while (true) {
        unsigned long long it = 0;
        if (lane == 0) it = atomicAdd(queue, 1ull);
        const int64_t item = (int64_t)__builtin_amdgcn_readfirstlane((int)(it & 0xFFFFFFFFull)) |
                             ((int64_t)__builtin_amdgcn_readfirstlane((int)(it >> 32)) << 32);
        if (item >= nitems) break;
        // the item's problem: the current one, the next, or a binary search
        if (!(ipre[p] <= item && item < ipre[p + 1])) {
            if (p + 1 < np && ipre[p + 1] <= item && item < ipre[p + 2]) {
                p = p + 1;
            } else {
                int lo = 0, hi = np - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (ipre[mid] <= item) lo = mid;
                    else hi = mid - 1;
                }
                p = lo;
            }
        }
        const BatchProb<F, TG> &Q = probs[p];
        const int64_t nc = Q.nc;
        const int tau = Q.tau, m = Q.m, nt = Q.nt;
        const int64_t hs = nc + 1;
        const F *xs = x + Q.xoff;
        const int64_t local = item - ipre[p];
        const int64_t nlam = (kinds & 1) ? batch_lam_items(nc) : 0;
        const int kind = local < nlam ? 0 : 1;
        if (p != cur || kind != cur_kind) {
            if (cur >= 0) batch_flush<TG>(Wv.lh, cur_nt, cur_row, cur_hs);
            if (p != cur) {
                if (lane < TG) Wv.bnd[lane] = Q.b[lane];
                wave_sync();
                P = batch_lut<F, TG>(Wv.bnd, nt, Wv.lut);
            }
            cur = p;
            cur_kind = kind;
            cur_row = hist + Q.hoff + (kind ? (int64_t)nt * hs : 0);
            cur_hs = hs;
            cur_nt = nt;
            wave_sync();
        }
        uint32_t run[TG];
#pragma unroll
        for (int t = 0; t < TG; t++) run[t] = 0u;

        if (kind == 0) {
            // LAM: band `local`, full columns
            const int64_t j = local * WAVE + lane;
            const bool jv = j < nc;
            F xjc[MJ];
#pragma unroll
            for (int k = 0; k < MJ; k++) xjc[k] = xs[j + (int64_t)(k < m ? k : 0) * tau];
            for (int64_t i0 = 0; i0 < nc; i0 += BLK) {
                const int nb = nc - i0 < BLK ? (int)(nc - i0) : BLK;
                F s[BLK];
#pragma unroll
                for (int u = 0; u < BLK; u++) s[u] = F(0);
#pragma unroll
                for (int k = 0; k < MJ; k++) {
                    if (k < m) {
                        const F *xr = xs + i0 + (int64_t)k * tau;
#pragma unroll
                        for (int u = 0; u < BLK; u++) s[u] = dim_step<NORM>(s[u], xr[u], xjc[k]);
                    }
                }
                for (int k = MJ; k < m; k++) {
                    const F xj = xs[j + (int64_t)k * tau];
                    const F *xr = xs + i0 + (int64_t)k * tau;
#pragma unroll
                    for (int u = 0; u < BLK; u++) s[u] = dim_step<NORM>(s[u], xr[u], xj);
                }
                uint32_t R[8];
                pack_ranks<F, TG>(s, R, Wv.lut, P, Wv.bnd, nt);
                // rows past the column's end count as recurrent padding, taken back below
                const uint32_t fill = nb == BLK ? 0u : (0xFFFFFFFFu << nb);
                batch_scan<TG>(R, jv ? ~fill : 0u, jv ? 0u : 0xFFFFFFFFu, run, nt, Wv.lh, BLCAP - 1 + lane,
                               cur_row, hs);
                if (fill && jv) {
#pragma unroll
                    for (int t = 0; t < TG; t++) run[t] -= (uint32_t)(BLK - nb);
                }
            }
            if (jv) batch_close<TG>(run, nt, Wv.lh, cur_row, hs);
        } else {
            // DET: strip local - nlam, whole diagonals r = r0 + lane, pairs (s, s + r)
            const int64_t r0 = 1 + (local - nlam) * WAVE;
            const int64_t r = r0 + lane;
            const bool rv = r < nc;
            const int64_t dlen = rv ? nc - r : 0;
            const int64_t smax = nc - r0;   // lane 0's diagonal, the strip's longest
            const int kc_max = tau > 0 ? (BWIN - (BLK + WAVE - 1)) / tau + 1 : m;
            const int kc = kc_max < m ? kc_max : m;
            for (int64_t s0 = 0; s0 < smax; s0 += BLK) {
                F s[BLK];
#pragma unroll
                for (int u = 0; u < BLK; u++) s[u] = F(0);
                for (int k0 = 0; k0 < m; k0 += kc) {
                    const int k1 = k0 + kc < m ? k0 + kc : m;
                    const F *base = xs + s0 + r0 + (int64_t)k0 * tau;
                    const int span = BLK + WAVE - 1 + (k1 - 1 - k0) * tau;
                    wave_sync();
                    for (int e = lane; e < span; e += WAVE) Wv.win[e] = base[e];
                    wave_sync();
                    for (int k = k0; k < k1; k++) {
                        const F *xr = xs + s0 + (int64_t)k * tau;
                        const F *wl = Wv.win + lane + (k - k0) * tau;
                        // all 32 window reads in flight before the arithmetic
                        // (left to itself the compiler waits on every pair)
                        F wv[BLK];
#pragma unroll
                        for (int u = 0; u < BLK; u++) wv[u] = wl[u];
                        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                        for (int u = 0; u < BLK; u++) s[u] = dim_step<NORM>(s[u], xr[u], wv[u]);
                    }
                }
                uint32_t R[8];
                pack_ranks<F, TG>(s, R, Wv.lut, P, Wv.bnd, nt);
                const int64_t nv64 = dlen - s0;
                const int nvalid = nv64 <= 0 ? 0 : (nv64 >= BLK ? BLK : (int)nv64);
                const uint32_t vmask = nvalid >= BLK ? 0xFFFFFFFFu : ((1u << nvalid) - 1u);
                batch_scan<TG>(R, 0xFFFFFFFFu, ~vmask, run, nt, Wv.lh, BLCAP - 1 + lane, cur_row, hs);
            }
            if (rv) batch_close<TG>(run, nt, Wv.lh, cur_row, hs);
        }
    }
    if (cur >= 0) batch_flush<TG>(Wv.lh, cur_nt, cur_row, cur_hs);
}

// ---------------------------------------------------------------- k_batch_fin
// One workgroup per (block, caller threshold): the reference's five measures
// of DET and of LAM for every l_min, and RR, from the two u32 histograms of
// the block's threshold (AccRQA_functions.cpp:48-133 with the histogram
// processing of AccRQA_CPU_function.cpp:156-204), written at the block's
// place in the caller's layouts.  rows[b*T + c] = (h_V row, h_D row) offsets
// into hist, or x < 0 for a threshold that never recurs (empty histograms).
// DET's histogram is taken as 2 h[l] plus the line of identity at N_c
// (correctDETHistogram).  Lmax and TTmax are the values the reference's
// bisection over the suffix counts returns for a non-increasing scan: the
// largest l whose suffix count exceeds the last bin's (DET: the longest line
// below N_c, 0 if none) or exceeds 0 (LAM: the longest line).  ENTR sums
// p ln p in fp64 (the reference's order is sequential; this is a tree).
constexpr int BFIN = 256;
constexpr int BFIN_L = 4;   // l_min values per pass

struct BatchBlockOut {
    int nc;
    int64_t rr;    // element offset of the block's RR [T]
    int64_t det;   // tuple offset (units of 5 values) of its DET [T][nl][5]
    int64_t lam;   // tuple offset of its LAM [nl][T][5]
};

template <typename F>
__global__ __launch_bounds__(BFIN) void k_batch_fin(const uint32_t *__restrict__ hist,
                                                    const BatchBlockOut *__restrict__ blocks,
                                                    const longlong2 *__restrict__ rows, int T,
                                                    const int *__restrict__ lmins, int nl, F *__restrict__ rr_out,
                                                    F *__restrict__ det_out, F *__restrict__ lam_out) {
    const int64_t bc = blockIdx.x;
    const int64_t b = bc / T;
    const int c = (int)(bc - b * T);
    const BatchBlockOut B = blocks[b];
    const int64_t nc = B.nc, hs = nc + 1;
    const longlong2 rw = rows[bc];
    const int tid = threadIdx.x;
    __shared__ unsigned long long ru[BFIN];
    __shared__ double rd[BFIN];

    auto red_u = [&](unsigned long long v, bool is_max) -> unsigned long long {
        __syncthreads();
        ru[tid] = v;
        __syncthreads();
        for (int off = BFIN / 2; off > 0; off >>= 1) {
            if (tid < off) ru[tid] = is_max ? max(ru[tid], ru[tid + off]) : ru[tid] + ru[tid + off];
            __syncthreads();
        }
        return ru[0];
    };
    auto red_d = [&](double v) -> double {
        __syncthreads();
        rd[tid] = v;
        __syncthreads();
        for (int off = BFIN / 2; off > 0; off >>= 1) {
            if (tid < off) rd[tid] += rd[tid + off];
            __syncthreads();
        }
        return rd[0];
    };

    for (int kind = 0; kind < 2; kind++) {   // 0: LAM (h_V), 1: DET (h_D)
        F *out = kind ? det_out : lam_out;
        if (!out && !(kind == 0 && rr_out)) continue;
        const bool live = rw.x >= 0;
        const uint32_t *h = live ? hist + (kind ? rw.y : rw.x) : nullptr;
        // bin f of the processed histogram
        auto hv = [&](int64_t f) -> unsigned long long {
            const unsigned long long v = live ? (unsigned long long)h[f] : 0ull;
            return kind ? 2ull * v + (f == nc ? 1ull : 0ull) : v;
        };
        // pass 1: M0 = sum f h, the longest line below N_c, and per l_min the
        // suffix sums S_l = sum_{f >= l} h, M_l = sum_{f >= l} f h
        unsigned long long m0 = 0, top = 0;
        for (int64_t f = tid; f < hs; f += BFIN) {
            const unsigned long long v = hv(f);
            m0 += (unsigned long long)f * v;
            if (v && f < nc) top = (unsigned long long)f + 1;
        }
        m0 = red_u(m0, false);
        top = red_u(top, true);
        const unsigned long long hN = hv(nc);
        const double lmax = kind ? (double)(top ? top - 1 : 0) : (hN ? (double)nc : (double)(top ? top - 1 : 0));
        if (kind == 0 && rr_out && tid == 0)
            rr_out[B.rr + c] = ((F)(int64_t)m0) / ((F)(int64_t)(nc * nc));
        if (!out) continue;
        for (int l0 = 0; l0 < nl; l0 += BFIN_L) {
            unsigned long long S[BFIN_L], M[BFIN_L];
            int64_t lm[BFIN_L];
#pragma unroll
            for (int q = 0; q < BFIN_L; q++) {
                S[q] = M[q] = 0;
                lm[q] = l0 + q < nl ? lmins[l0 + q] : hs;
            }
            for (int64_t f = tid; f < hs; f += BFIN) {
                const unsigned long long v = hv(f);
#pragma unroll
                for (int q = 0; q < BFIN_L; q++)
                    if (f >= lm[q]) {
                        S[q] += v;
                        M[q] += (unsigned long long)f * v;
                    }
            }
#pragma unroll
            for (int q = 0; q < BFIN_L; q++) {
                S[q] = red_u(S[q], false);
                M[q] = red_u(M[q], false);
            }
            double E[BFIN_L];
#pragma unroll
            for (int q = 0; q < BFIN_L; q++) E[q] = 0.0;
            for (int64_t f = tid; f < hs; f += BFIN) {
                const unsigned long long v = hv(f);
#pragma unroll
                for (int q = 0; q < BFIN_L; q++)
                    if (f >= lm[q]) {
                        const double pr = (double)v / (double)S[q];
                        if (pr > 0) E[q] += pr * log(pr);
                    }
            }
#pragma unroll
            for (int q = 0; q < BFIN_L; q++) E[q] = red_d(E[q]);
            if (tid == 0) {
                for (int q = 0; q < BFIN_L && l0 + q < nl; q++) {
                    const int l = l0 + q;
                    const double v0 = m0 > 0 ? (double)M[q] / (double)m0 : __builtin_nan("");
                    const double v1 = S[q] > 0 ? (double)M[q] / (double)S[q] : __builtin_nan("");
                    const double v3 = (-1.0) * E[q];
                    const double v4 = (double)m0 / (double)(nc * nc);
                    // DET: [T][nl][5] (l fastest); LAM: [nl][T][5] (threshold fastest)
                    F *o = out + 5 * (kind ? (B.det + (int64_t)c * nl + l) : (B.lam + (int64_t)l * T + c));
                    o[0] = (F)v0;
                    o[1] = (F)v1;
                    o[2] = (F)lmax;
                    o[3] = (F)v3;
                    o[4] = (F)v4;
                }
            }
        }
    }
}

// Zeroes n elements of p and, when given, the work-queue counter q.
template <typename T>
__global__ void k_fill_zero(T *__restrict__ p, size_t n, unsigned long long *__restrict__ q) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    if (q && i == 0) *q = 0ull;
    for (; i < n; i += stride) p[i] = T(0);
}

}  // namespace rqa
