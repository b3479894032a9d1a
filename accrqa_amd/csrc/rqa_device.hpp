// rqa_device.hpp -- gfx950 kernels of the RR / DET / LAM engine.
//
// Decision (one pair, one threshold), restating R_matrix_element
// (src/AccRQA_CPU_function.cpp:40-46) without a square root:
//   Euclidean: s = sum_k (x[i+k*tau] - x[j+k*tau])^2 in the input type, k in
//              order, one rounding per operation (no FMA: this file is
//              compiled with contraction off), then R = [s <= S(eps)] where the
//              host computes S(eps) = the largest s with RN(sqrt(s)) <= eps, so
//              the decision equals the reference's (eps - sqrt(s)) >= 0 bit for bit.
//   Maximal:   s = max(0, |x[i+k*tau] - x[j+k*tau]|...) ignoring NaN components,
//              R = [s <= eps] (eps = +inf mapped to the largest finite value).
//
// Sequences and their kernels (all thresholds of a group in one pass):
//   k_lam  -- lane = column j of the full N_c x N_c matrix, steps = rows i.
//             Column runs equal the reference's row runs because R is
//             bit-symmetric (get_length_histogram_LAM_inplace,
//             src/AccRQA_CPU_function.cpp:105-142).
//   k_det  -- lane = upper diagonal r, steps = s along it, pair (s, s+r)
//             (get_length_histogram_DET_inplace, src/AccRQA_CPU_function.cpp:63-103).
//   k_rr   -- recurrence count over the upper triangle, weight 2 off the
//             diagonal and 1 on it (rqa_CPU_RR_metric_ref_parallel,
//             src/AccRQA_CPU_function.cpp:347-378, counts the full matrix).
//
// Per 32-step block a lane packs one threshold rank per step into eight
// registers, turns them into bit planes and, through a comparator tree, into
// one "not recurrent" word per threshold (bit u = step u, bit 0 oldest); runs
// are read off each word with find-first-bit tricks and the open run is
// carried to the next word (rqa_bits.hpp).  A sequence is cut into segments
// processed independently; each segment reports (prefix run, suffix run) and
// k_stitch joins them in order.  Histogram bins below LCAP are counted in a
// workgroup-private LDS histogram and merged with one global atomic per
// non-zero bin; longer runs go straight to global u64 atomics.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rqa_bits.hpp"

// A translation unit that only needs the device helpers defines
// RQA_HELPERS_ONLY: the non-template kernels below (one definition each, in
// the engine's own unit) are then left out.

#pragma clang fp contract(off)

namespace rqa {

constexpr int WAVE = 64;
constexpr int BLK = 32;            // steps per decision word
constexpr int WG = 256;            // threads per workgroup (4 waves)
constexpr int WAVES = WG / WAVE;
// k_lam, fp32 Euclidean up to 16 thresholds: 4 waves per SIMD (128 VGPRs, a
// few bytes spilled) instead of the compiler's 3 (measured: 61.8 -> 56.0 ms
// on the default workload); the other instantiations keep what their
// registers give (1 = no minimum).
#define RQA_WPE_LAM(F, NORM, TG) \
    __attribute__((amdgpu_waves_per_eu((sizeof(F) == 4 && (NORM) == 1 && (TG) <= 16) ? 4 : 1)))
// k_lam_mx, k_det_mx: three waves per SIMD with the line scan (168 VGPRs for
// 16 thresholds, no spills; left to itself the compiler takes a few registers
// more and halves that), four without it (the LDS allows four workgroups).
#ifndef RQA_WPE_MX
#define RQA_WPE_MX(scan) __attribute__((amdgpu_waves_per_eu((scan) ? 3 : 4)))
#endif
// The tile scans up to 16 thresholds: 7 waves per SIMD (<= 73 VGPRs), one
// or two more than the compiler's own choice (measured on the default
// workload: -0.8 ms for each LAM scan, -1.9 ms for k_det_scan).  At 32
// thresholds the column and row scans take 6 (80 VGPRs, a few spilled; 3 %
// off the million-sample call) and k_det_scan keeps what its registers give.
#define RQA_WPE_SCAN(tg, n, n32) __attribute__((amdgpu_waves_per_eu((tg) == 32 ? (n32) : (n))))
constexpr int DET_WIN = 992;       // LDS samples per wave for the k_det lane window

template <typename F, int TG>
struct Bounds {
    F s[TG];
};

// One embedding dimension of the distance, in the reference's order.
template <int NORM, typename F>
__device__ __forceinline__ F dim_step(F s, F a, F b) {
    F d = a - b;
    if constexpr (NORM == 1) {
        F q = d * d;
        return s + q;
    } else {
        // fmax returns the non-NaN operand, i.e. `if (|d| > M) M = |d|` with
        // NaN components ignored; s is never NaN (it starts at +0).
        return fmax(s, fabs(d));
    }
}

__device__ __forceinline__ void lds_inc(uint32_t *p) {
    __hip_atomic_fetch_add(p, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void glob_add(unsigned long long *p, unsigned long long v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A uniform value the compiler may not hoist out of a rare branch: keeps
// per-threshold addresses from occupying scalar registers for the whole loop.
__device__ __forceinline__ uint32_t opaque(uint32_t v) {
    asm volatile("" : "+s"(v));
    return v;
}

// LDS traffic between the lanes of one wave only.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// OR of x over the wave's 64 lanes, returned as a wave-uniform value:
// prefix ORs within each row of 16 lanes (row_shr 1, 2, 4, 8), then across
// rows (row_bcast 15 into rows 1 and 3, row_bcast 31 into rows 2 and 3), so
// lane 63 ends with the OR of all lanes.  Lanes a shift leaves without a
// source keep 0.
__device__ __forceinline__ uint32_t wave_or(uint32_t x) {
    x |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xF, 0xF, false);
    x |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xF, 0xF, false);
    x |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xF, 0xF, false);
    x |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xF, 0xF, false);
    x |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xA, 0xF, false);
    x |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xC, 0xF, false);
    return (uint32_t)__builtin_amdgcn_readlane((int)x, 63);
}

__device__ __forceinline__ int wave_id() {
    return __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
}

// ---------------------------------------------------------------- line scan
// Run-length histogram of one workgroup, in three tiers: lengths below LCAP
// in LDS; lengths LCAP .. PCAP-1 in the workgroup's own slice of a global
// scratch histogram, counted with workgroup-scope atomics (no other
// workgroup touches the slice, so the adds stay in the XCD's L2 without
// contention); longer runs with device-scope atomics straight into the
// result.  The slice is zeroed on entry and added to the result on exit.
constexpr int PCAP = 2048;
constexpr int PBINS = PCAP - LCAP;   // private bins per threshold

struct Sink {
    uint32_t *lh;                  // LDS: row t at t * lrow<NSP>(), bins 1 .. LCAP-1, then NSP spare bins
    uint32_t spare;                // this lane's spare bin (lane_spare), in consume's numbering
    uint32_t *priv;                // workgroup slice: row t at t * PBINS, bin len - LCAP
    unsigned long long *hist;      // result: row t at t * (nc + 1)
    int64_t nc;
};

__device__ __forceinline__ void wg_add(uint32_t *p) {
    __hip_atomic_fetch_add(p, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <int TG>
__device__ __forceinline__ uint32_t *priv_slice(uint32_t *priv_base) {
    return priv_base + (size_t)blockIdx.x * TG * PBINS;
}

template <int TG>
__device__ __forceinline__ void priv_zero(uint32_t *priv, int nt, int nthreads) {
    for (int e = threadIdx.x; e < nt * PBINS; e += nthreads) priv[e] = 0u;
}

// A kernel's LDS histogram rows hold LCAP bins and NSP spare bins (WAVE / NSP
// lanes share one): NSP = NSPARE where LDS is tight, WAVE where it is not.
template <int NSP>
constexpr int lrow() {
    static_assert(NSP >= 1 && NSP <= WAVE && WAVE % NSP == 0, "spare bins per row");
    return LCAP + NSP;
}
template <int NSP = NSPARE>
__device__ __forceinline__ uint32_t lane_spare() {
    return LCAP - 1 + ((threadIdx.x & (WAVE - 1)) / (WAVE / NSP));
}

// Adds the workgroup's LDS and private bins to the result (after a barrier):
// bins 1 .. LCAP-1 of each LDS row (bin 0 and the spare bins count nothing).
template <int TG, int NSP = NSPARE>
__device__ __forceinline__ void flush_hist(const uint32_t *lh, uint32_t *priv, int nt, unsigned long long *hist,
                                           int64_t nc, int nthreads) {
    for (int e = threadIdx.x; e < TG * LCAP; e += nthreads) {
        const int t = e / LCAP, b = e - t * LCAP;
        const uint32_t v = lh[t * lrow<NSP>() + b];
        if (v && b && t < nt) glob_add(&hist[(size_t)t * (nc + 1) + b], v);
    }
    for (int e = threadIdx.x; e < nt * PBINS; e += nthreads) {
        const uint32_t v = __hip_atomic_load(&priv[e], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const int t = e / PBINS;
        const int64_t len = LCAP + (e - t * PBINS);
        if (v && len <= nc) glob_add(&hist[(size_t)t * (nc + 1) + len], v);
    }
}

// The slow path of consume(): the segment's first run goes to the carry slot
// (for the stitch), runs of LCAP or more to the private or result bins.  A
// run longer than the sequence would be an engine bug and is dropped rather
// than written out of bounds (the parity tests would then fail on the count).
__device__ __forceinline__ void slow_run(uint32_t first, uint2 *slot, const Sink &K, uint32_t t) {
    if (first >= PRE)
        slot->x = first - PRE;
    else if (first < (uint32_t)PCAP)
        wg_add(&K.priv[(size_t)t * PBINS + (first - LCAP)]);
    else if (first <= (uint64_t)K.nc)
        glob_add(&K.hist[(size_t)t * (K.nc + 1) + first], 1ull);
}

__device__ __forceinline__ void close_run(uint2 *slot, uint32_t run) {
    const SegEnds e = seg_ends(run);
    if (e.pre_open)
        *slot = make_uint2(e.pre, e.suf);
    else
        slot->y = e.suf;
}

// All thresholds of one 32-step block of one lane.  nw(t) = (ge(t+1) & keep)
// | force: keep clears steps that count as recurrent padding, force sets
// steps that decide 0 (both applied to the rank planes).  Histogram row t
// starts at hist + t*(nc+1), the segment's carry slot of threshold t at
// cslot + t*cstride.  Called by all lanes of the wave together.
//
// Thresholds are sorted, so a step that does not recur at t does not recur at
// any lower threshold, and a run open at t is open at every higher one: the
// thresholds at which every lane's word is all non-recurrent and no lane has
// a run open form a prefix.  consume() would change nothing there (no run
// closes or opens), so the wave skips them, testing up to the first threshold
// that fails.
template <int TG, int NSP = NSPARE>
__device__ __forceinline__ void scan_words(const uint32_t (&R)[8], uint32_t keep, uint32_t force, uint32_t (&run)[TG],
                                           int nt, const Sink &K, uint2 *cslot, size_t cstride) {
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
            uint32_t *lht = K.lh + t * lrow<NSP>();
            consume(
                nw, run[t], K.spare, [&](uint32_t b) { lds_inc(&lht[b + 1]); },
                [&](uint32_t len) { lds_inc(&lht[len]); },
                [&](uint32_t first) {
                    const uint32_t tt = opaque(t);
                    slow_run(first, cslot + (size_t)tt * cstride, K, tt);
                });
        }
    }
}

template <int TG>
__device__ __forceinline__ void close_segment(const uint32_t (&run)[TG], int nt, uint2 *cslot, size_t cstride) {
#pragma unroll
    for (int t = 0; t < TG; t++)
        if (t < nt) close_run(&cslot[(size_t)opaque(t) * cstride], run[t]);
}


// ---------------------------------------------------------------- threshold ranks
// All thresholds of a group are decided through one integer per pair, its
// rank r = #{t : B_t < s} over the group's bounds sorted ascending: then
// R_t = [s <= B_t] = [t >= r].  The rank comes from a bucket table indexed by
// the high bits of s (exact: a bucket holds at most one bound, whose key is
// stored with the bucket; a bucket with several bounds sends the wave to a
// direct count).  Keys are the IEEE bit patterns, which order non-negative
// values exactly; NaN maps to +inf, i.e. rank T, never recurrent.
constexpr int NB_MAX = 2048;   // bucket table entries (LDS)

struct LutParams {
    uint32_t base;   // high key bits of the first bucket
    int sh;          // bucket width = 1 << sh (in high key bits)
    int nb;          // number of buckets
};

template <typename F>
struct LutEntry;
template <>
struct LutEntry<float> {
    uint32_t next;   // key of the first bound >= bucket start (0xFFFFFFFF if none)
    uint32_t info;   // bits 0-7: bounds below the bucket start; bit 8: several bounds in the bucket
};
template <>
struct LutEntry<double> {
    unsigned long long next;
    uint32_t info;
    uint32_t pad;
};

__device__ __forceinline__ uint32_t skey(float s) { return min(__float_as_uint(s), 0x7F800000u); }
__device__ __forceinline__ unsigned long long skey(double s) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(s);
    return b > 0x7FF0000000000000ull ? 0x7FF0000000000000ull : b;
}
__device__ __forceinline__ uint32_t khi(uint32_t k) { return k; }
__device__ __forceinline__ uint32_t khi(unsigned long long k) { return (uint32_t)(k >> 32); }

template <typename F>
__device__ __forceinline__ uint32_t rank_of(F s, const LutEntry<F> *lut, const LutParams &P, uint32_t &flags) {
    const auto key = skey(s);
    const int b = max(min((int)(khi(key) - P.base) >> P.sh, P.nb - 1), 0);
    const LutEntry<F> e = lut[b];
    flags |= e.info;
    return (e.info & 0xFFu) + (key > e.next ? 1u : 0u);
}

// Direct count against the sorted bounds (buckets holding several bounds);
// bnd is the workgroup's LDS copy of the group's bounds.
template <typename F, int TG>
__device__ __forceinline__ uint32_t rank_direct(F s, const F *bnd, int nt) {
    const auto key = skey(s);
    uint32_t r = 0;
#pragma unroll
    for (int t = 0; t < TG; t++)
        if (t < nt) r += key > skey(bnd[t]) ? 1u : 0u;
    return r;
}

// Ranks of 32 steps, one per byte: register q holds steps q, q+8, q+16, q+24.
template <typename F, int TG>
__device__ __forceinline__ void pack_ranks(const F (&s)[BLK], uint32_t (&R)[8], const LutEntry<F> *lut,
                                           const LutParams &P, const F *bnd, int nt) {
    uint32_t flags = 0;
#pragma unroll
    for (int q = 0; q < 8; q++) R[q] = 0;
#pragma unroll
    for (int u = 0; u < BLK; u++) R[u & 7] |= rank_of(s[u], lut, P, flags) << (8 * (u >> 3));
    if (__any((flags & 0x100u) != 0)) {
#pragma unroll
        for (int q = 0; q < 8; q++) R[q] = 0;
#pragma unroll
        for (int u = 0; u < BLK; u++) R[u & 7] |= rank_direct<F, TG>(s[u], bnd, nt) << (8 * (u >> 3));
    }
}

template <typename F>
__device__ __forceinline__ void load_lut(LutEntry<F> *dst, const LutEntry<F> *src, int nb) {
    for (int e = threadIdx.x; e < nb; e += WG) dst[e] = src[e];
}

// ---------------------------------------------------------------- rank tiles
// When LAM and DET are computed together, the LAM pass keeps its packed ranks
// for k_det_scan.  Tile (row block I, band J) covers rows 32I .. 32I+31 and
// columns 64J .. 64J+63; lane l's 32 bytes (the eight rank registers of
// column 64J + l, register q byte g = row 32I + 8g + q) sit at tile*2048 +
// 32l.  Only tiles with J >= I/2 are kept -- every pair (i, j > i) lies in
// one of them -- packed row block by row block: row block I starts at tile
// I*nbands - floor(I/2)*floor((I-1)/2).
constexpr int TILE_BYTES = 2048;

__device__ __forceinline__ int64_t tile_index(int64_t I, int64_t J, int64_t nbands) {
    return I * nbands - (I / 2) * ((I - 1) / 2) + (J - I / 2);
}

__device__ __forceinline__ void store_tile(uint8_t *tiles, int64_t I, int64_t J, int64_t nbands,
                                           const uint32_t (&R)[8]) {
    if (J < I / 2) return;
    uint4 *dst = (uint4 *)(tiles + tile_index(I, J, nbands) * TILE_BYTES + (threadIdx.x & (WAVE - 1)) * 32);
    dst[0] = make_uint4(R[0], R[1], R[2], R[3]);
    dst[1] = make_uint4(R[4], R[5], R[6], R[7]);
}

// LAM work of the column kernels.  Without a table: every (band, segment)
// over full columns, carry slot q of nseg per column.  With a table (the
// split form, rank tiles kept): (band J, segment q) over rows [0, V_J) only,
// V_J = min(64(J+1), N_c) -- the upper triangle and the band's diagonal
// tiles; k_lam_hscan covers rows >= V_J from the tiles, in carry slots
// nseg .. 2 nseg-1 of the column.
struct LamWork {
    const int2 *items;
    int64_t nitems;
    int nseg;   // segments of a full column
    int cseg;   // carry slots per column
};

// ---------------------------------------------------------------- k_lam
// Work item = (band of 64 columns, row segment q).  x is padded with >= 128
// zeros past N so that the unguarded window reads stay inside the buffer.
// The row window x[i0 + u + k*tau] (u < 32) is uniform across the wave and
// comes through scalar loads; x[j + k*tau] is one vector load per k and block.
template <typename F, int NORM, int TG>
__global__ __launch_bounds__(WG) RQA_WPE_LAM(F, NORM, TG) void k_lam(const F *__restrict__ x, int64_t nc, int tau, int m, Bounds<F, TG> S,
                                            int nt, const LutEntry<F> *__restrict__ lut, LutParams P, int seg_len,
                                            LamWork W, unsigned long long *__restrict__ hist,
                                            uint2 *__restrict__ carry, uint32_t *__restrict__ priv_base,
                                            uint8_t *__restrict__ tiles) {
    __shared__ uint32_t lh[TG * LROW];
    __shared__ LutEntry<F> lut_s[NB_MAX];
    __shared__ F bnd_s[TG];
    for (int e = threadIdx.x; e < TG * LROW; e += WG) lh[e] = 0;
    load_lut(lut_s, lut, P.nb);
    if (threadIdx.x < TG) bnd_s[threadIdx.x] = S.s[threadIdx.x];
    priv_zero<TG>(priv_slice<TG>(priv_base), nt, WG);
    __syncthreads();
    uint32_t *priv = priv_slice<TG>(priv_base);
    const Sink K{lh, lane_spare(), priv, hist, nc};

    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = wave_id();
    const int64_t nbands = (nc + WAVE - 1) / WAVE;
    const int64_t nitems = W.items ? W.nitems : nbands * W.nseg;
    const size_t cstride = (size_t)W.cseg * nc;   // carry stride between thresholds
    constexpr int MJ = sizeof(F) == 4 ? 16 : 8;  // column samples kept in registers

    for (int64_t item = (int64_t)blockIdx.x * WAVES + wave; item < nitems; item += (int64_t)gridDim.x * WAVES) {
        int64_t band;
        int q;
        if (W.items) {
            const int2 it = W.items[item];
            band = it.x;
            q = it.y;
        } else {
            band = item / W.nseg;
            q = (int)(item - band * W.nseg);
        }
        const int64_t vend = W.items ? min((band + 1) * WAVE, nc) : nc;
        const int64_t j = band * WAVE + lane;
        const bool jv = j < nc;
        const int64_t ib = (int64_t)q * seg_len;
        const int64_t ie = ib + seg_len < vend ? ib + seg_len : vend;
        uint2 *cslot = carry + (size_t)q * nc + (jv ? j : 0);

        uint32_t run[TG];
#pragma unroll
        for (int t = 0; t < TG; t++) run[t] = jv ? PRE : 0u;
        // the column's samples x[j + k tau] are the same for every block of the item
        F xjc[MJ];
#pragma unroll
        for (int k = 0; k < MJ; k++) xjc[k] = x[j + (int64_t)(k < m ? k : 0) * tau];

        for (int64_t i0 = ib; i0 < ie; i0 += BLK) {
            const int nb = ie - i0 < BLK ? (int)(ie - i0) : BLK;
            F s[BLK];
#pragma unroll
            for (int u = 0; u < BLK; u++) s[u] = F(0);
#pragma unroll
            for (int k = 0; k < MJ; k++) {
                if (k < m) {
                    const F *xr = x + i0 + (int64_t)k * tau;
#pragma unroll
                    for (int u = 0; u < BLK; u++) s[u] = dim_step<NORM>(s[u], xr[u], xjc[k]);
                }
            }
            for (int k = MJ; k < m; k++) {
                const F xj = x[j + (int64_t)k * tau];
                const F *xr = x + i0 + (int64_t)k * tau;
#pragma unroll
                for (int u = 0; u < BLK; u++) s[u] = dim_step<NORM>(s[u], xr[u], xj);
            }
            uint32_t R[8];
            pack_ranks<F, TG>(s, R, lut_s, P, bnd_s, nt);
            if (tiles) store_tile(tiles, i0 / BLK, band, nbands, R);
            // rows past the segment count as recurrent padding, taken back below
            const uint32_t fill = nb == BLK ? 0u : (0xFFFFFFFFu << nb);
            scan_words<TG>(R, jv ? ~fill : 0u, jv ? 0u : 0xFFFFFFFFu, run, nt, K, cslot, cstride);
            if (fill && jv) {
#pragma unroll
                for (int t = 0; t < TG; t++) run[t] -= (uint32_t)(BLK - nb);
            }
        }
        if (jv) close_segment<TG>(run, nt, cslot, cstride);
    }

    __syncthreads();
    flush_hist<TG>(lh, priv, nt, hist, nc, WG);
}

// ---------------------------------------------------------------- k_det
// Work item = (strip of 64 diagonals r = r0..r0+63, step segment q) taken
// from a host-built table.  Step s of lane l is the pair (s, s + r0 + l);
// steps past the end of a diagonal decide 0.  The row window x[s + k*tau] is
// uniform (scalar loads); the per-lane window x[s + r0 + l + k*tau] is staged
// in LDS once per 32-step block and dimension chunk.
template <typename F, int NORM, int TG>
__global__ __launch_bounds__(WG) void k_det(const F *__restrict__ x, int64_t nc, int tau, int m, Bounds<F, TG> S,
                                            int nt, const LutEntry<F> *__restrict__ lut, LutParams P, int seg_len,
                                            int nseg, const int2 *__restrict__ items, int64_t nitems,
                                            unsigned long long *__restrict__ hist, uint2 *__restrict__ carry,
                                            uint32_t *__restrict__ priv_base) {
    __shared__ uint32_t lh[TG * LROW];
    __shared__ LutEntry<F> lut_s[NB_MAX];
    __shared__ F win[WAVES][DET_WIN];
    __shared__ F bnd_s[TG];
    for (int e = threadIdx.x; e < TG * LROW; e += WG) lh[e] = 0;
    load_lut(lut_s, lut, P.nb);
    if (threadIdx.x < TG) bnd_s[threadIdx.x] = S.s[threadIdx.x];
    priv_zero<TG>(priv_slice<TG>(priv_base), nt, WG);
    __syncthreads();
    uint32_t *priv = priv_slice<TG>(priv_base);
    const Sink K{lh, lane_spare(), priv, hist, nc};

    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = wave_id();
    F *wn = win[wave];
    // dimensions staged per window: span = BLK + WAVE - 1 + (kc-1)*tau <= DET_WIN
    const int kc_max = tau > 0 ? (DET_WIN - (BLK + WAVE - 1)) / tau + 1 : m;
    const int kc = kc_max < m ? kc_max : m;
    const size_t cstride = (size_t)nseg * (nc - 1);

    for (int64_t item = (int64_t)blockIdx.x * WAVES + wave; item < nitems; item += (int64_t)gridDim.x * WAVES) {
        const int2 it = items[item];
        const int64_t r0 = 1 + (int64_t)it.x * WAVE;
        const int q = it.y;
        const int64_t r = r0 + lane;
        const bool rv = r < nc;
        const int64_t dlen = rv ? nc - r : 0;       // steps s < dlen are on the diagonal
        const int64_t sb = (int64_t)q * seg_len;
        const int64_t se_nom = sb + seg_len;
        const int64_t smax = nc - r0;                // lane 0's diagonal length
        const int64_t se = se_nom < smax ? se_nom : smax;
        uint2 *cslot = carry + (size_t)q * (nc - 1) + (rv ? r - 1 : 0);

        uint32_t run[TG];
#pragma unroll
        for (int t = 0; t < TG; t++) run[t] = rv ? PRE : 0u;

        for (int64_t s0 = sb; s0 < se; s0 += BLK) {
            F s[BLK];
#pragma unroll
            for (int u = 0; u < BLK; u++) s[u] = F(0);
            for (int k0 = 0; k0 < m; k0 += kc) {
                const int k1 = k0 + kc < m ? k0 + kc : m;
                const int64_t base = s0 + r0 + (int64_t)k0 * tau;
                const int span = BLK + WAVE - 1 + (k1 - 1 - k0) * tau;
                for (int e = lane; e < span; e += WAVE) wn[e] = x[base + e];
                wave_sync();
                for (int k = k0; k < k1; k++) {
                    const F *xr = x + s0 + (int64_t)k * tau;
                    const F *wl = wn + lane + (k - k0) * tau;
#pragma unroll
                    for (int u = 0; u < BLK; u++) s[u] = dim_step<NORM>(s[u], xr[u], wl[u]);
                }
                wave_sync();
            }
            uint32_t R[8];
            pack_ranks<F, TG>(s, R, lut_s, P, bnd_s, nt);
            // bit u valid iff s0+u < dlen; invalid steps decide 0
            const int64_t nv64 = dlen - s0;
            const int nvalid = nv64 <= 0 ? 0 : (nv64 >= BLK ? BLK : (int)nv64);
            const uint32_t vmask = nvalid >= BLK ? 0xFFFFFFFFu : ((1u << nvalid) - 1u);
            scan_words<TG>(R, 0xFFFFFFFFu, ~vmask, run, nt, K, cslot, cstride);
        }
        // Steps past the strip's end decide 0 for every lane, so a segment cut
        // short by the end of the diagonals reports the same (prefix, suffix)
        // as one padded with zeros.
        if (rv) close_segment<TG>(run, nt, cslot, cstride);
    }

    __syncthreads();
    flush_hist<TG>(lh, priv, nt, hist, nc, WG);
}

// ---------------------------------------------------------------- k_det_scan
// DET pass over the rank tiles the LAM pass kept.  Work item = (group of
// four strips of 64 diagonals, step segment): wave w takes strip w, the
// diagonals r0 = rg + 64w + l (rg = 1 + 256 * group).  The pairs of a 32-step
// block, (s0+u, s0+r0+l+u), lie in row block s0/32 and columns s0+r0 ..
// s0+r0+94: tiles J0+w .. J0+w+2 (J0 = band of s0+rg), so the group's block
// needs the six tiles J0 .. J0+5, each read once and shared through LDS (one
// 36-byte record per tile lane: conflict-free gathers), against three per
// strip when the waves fetch their own.  The next block's records are in
// registers while a block is gathered and scanned.  (One LDS buffer: 22.5 KB
// per workgroup for 16 thresholds, which leaves registers, not LDS, to set
// the number of waves per SIMD.)
constexpr int TREC = 9;              // LDS words per tile lane (8 + 1 pad)
constexpr int DTILES = WAVES + 2;    // tiles of one block of a strip group
constexpr int DRECS = DTILES * WAVE; // their records (tile lanes)

template <int TG>
__global__ __launch_bounds__(WG) RQA_WPE_SCAN(TG, 7, 3) void k_det_scan(const uint8_t *__restrict__ tiles, int64_t nbands, int64_t nc, int nt,
                                                 int seg_len, int nseg, const int2 *__restrict__ items,
                                                 int64_t nitems, unsigned long long *__restrict__ hist,
                                                 uint2 *__restrict__ carry, uint32_t *__restrict__ priv_base) {
    static_assert(DRECS <= 2 * WG, "each thread moves at most two records per block");
    // 32 thresholds: 8 spare bins per row keep the workgroup's LDS under a
    // fifth of the CU's, so five fit (16 spare bins: four)
    constexpr int NSP = TG >= 32 ? 8 : NSPARE;
    __shared__ uint32_t lh[TG * lrow<NSP>()];
    __shared__ uint32_t tl[DRECS * TREC];
    for (int e = threadIdx.x; e < TG * lrow<NSP>(); e += WG) lh[e] = 0;
    priv_zero<TG>(priv_slice<TG>(priv_base), nt, WG);
    __syncthreads();
    uint32_t *priv = priv_slice<TG>(priv_base);
    const Sink K{lh, lane_spare<NSP>(), priv, hist, nc};

    const int tid = threadIdx.x;
    const int lane = tid & (WAVE - 1);
    const int wave = wave_id();
    const size_t cstride = (size_t)nseg * (nc - 1);
    // records this thread moves: tid and tid + WG (tile rec/64, tile lane rec%64)
    const int rec1 = tid + WG;
    const bool has1 = rec1 < DRECS;

    for (int64_t item = blockIdx.x; item < nitems; item += gridDim.x) {
        const int2 it = items[item];
        const int64_t rg = 1 + (int64_t)it.x * WG;
        const int q = it.y;
        const int64_t r0 = rg + (int64_t)wave * WAVE;
        const int64_t r = r0 + lane;
        const bool rv = r < nc;
        const int64_t dlen = rv ? nc - r : 0;
        const int64_t sb = (int64_t)q * seg_len;
        const int64_t se_g = min(sb + (int64_t)seg_len, nc - rg);   // strip 0 is the group's longest
        const int64_t se = min(sb + (int64_t)seg_len, nc - r0);     // this strip's part (none if <= sb)
        const bool wact = sb < se;
        uint2 *cslot = carry + (size_t)q * (nc - 1) + (rv ? r - 1 : 0);

        uint32_t run[TG];
#pragma unroll
        for (int t = 0; t < TG; t++) run[t] = rv ? PRE : 0u;

        uint4 pf[4];
        auto fetch = [&](int64_t s) {
            const int64_t I = s / BLK, J0 = (s + rg) / WAVE;
            if (J0 + tid / WAVE < nbands) {
                const uint4 *src = (const uint4 *)(tiles + tile_index(I, J0 + tid / WAVE, nbands) * TILE_BYTES) +
                                   2 * lane;
                pf[0] = src[0];
                pf[1] = src[1];
            }
            if (has1 && J0 + rec1 / WAVE < nbands) {
                const uint4 *src = (const uint4 *)(tiles + tile_index(I, J0 + rec1 / WAVE, nbands) * TILE_BYTES) +
                                   2 * lane;
                pf[2] = src[0];
                pf[3] = src[1];
            }
        };
        auto put = [&](int64_t s, uint32_t *buf) {
            const int64_t J0 = (s + rg) / WAVE;
            if (J0 + tid / WAVE < nbands) {
                uint32_t *d = buf + tid * TREC;
                d[0] = pf[0].x; d[1] = pf[0].y; d[2] = pf[0].z; d[3] = pf[0].w;
                d[4] = pf[1].x; d[5] = pf[1].y; d[6] = pf[1].z; d[7] = pf[1].w;
            }
            if (has1 && J0 + rec1 / WAVE < nbands) {
                uint32_t *d = buf + rec1 * TREC;
                d[0] = pf[2].x; d[1] = pf[2].y; d[2] = pf[2].z; d[3] = pf[2].w;
                d[4] = pf[3].x; d[5] = pf[3].y; d[6] = pf[3].z; d[7] = pf[3].w;
            }
        };
        if (sb < se_g) fetch(sb);
        for (int64_t s0 = sb; s0 < se_g; s0 += BLK) {
            __syncthreads();   // the previous block's gathers are done
            put(s0, tl);
            if (s0 + BLK < se_g) fetch(s0 + BLK);
            __syncthreads();
            if (s0 < se) {
                // step u of lane l: column s0 + r0 + l + u, i.e. record
                // (s0 + rg - 64 J0) + 64 w + l + u of the group's block
                const int64_t c0g = s0 + rg;
                const int base = (int)(c0g - (c0g / WAVE) * WAVE) + wave * WAVE + lane;
                const uint8_t *tb = (const uint8_t *)tl;
                uint32_t R[8] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
#pragma unroll
                for (int u = 0; u < BLK; u++)
                    R[u & 7] |= (uint32_t)tb[(base + u) * (4 * TREC) + 4 * (u & 7) + (u >> 3)] << (8 * (u >> 3));
                const int64_t nv64 = dlen - s0;
                const int nvalid = nv64 <= 0 ? 0 : (nv64 >= BLK ? BLK : (int)nv64);
                const uint32_t vmask = nvalid >= BLK ? 0xFFFFFFFFu : ((1u << nvalid) - 1u);
                scan_words<TG, NSP>(R, 0xFFFFFFFFu, ~vmask, run, nt, K, cslot, cstride);
            }
        }
        if (rv && wact) close_segment<TG>(run, nt, cslot, cstride);
    }

    __syncthreads();
    flush_hist<TG, NSP>(lh, priv, nt, hist, nc, WG);
}

// ---------------------------------------------------------------- k_lam_hscan
// The rest of each LAM column in the split form: rows i >= V_J of column j
// (band J) are, R being symmetric, columns i >= V_J of row j, and those pairs
// are in the rank tiles.  Work item = (row band J' = rows 64J' .. 64J'+63,
// segment q of the columns [V_J', N_c)); lane l walks row 64J' + l in blocks
// of 32 columns.  A block's bytes are records 32h .. 32h+31 (h = which half
// of column band Jc) of tiles (2J', Jc) and (2J'+1, Jc); the wave copies them
// to LDS and lane l reads byte (l%32) of each record of its row block.
//
// LDS layout of the wave's 64 records (record r = 32a + column, a = the row
// block): word w of record r at w*HW + (r%8)*8 + r/8.  Register q of lane
// (a, row) needs word row%8 of records 32a + q + 8g, g = 0..3, i.e. the four
// words at (row%8)*HW + 8q + 4a: one aligned 16-byte read, then byte row/8 of
// each by byte permutes.  HW = 72 puts the 16 distinct reads of a register
// (8 words x 2 row blocks) on distinct banks, and a store of word w by all 64
// lanes covers 64 consecutive words.
constexpr int HW = 72;
template <int TG>
__global__ __launch_bounds__(WG) RQA_WPE_SCAN(TG, 7, 6) void k_lam_hscan(const uint8_t *__restrict__ tiles, int64_t nbands, int64_t nc, int nt,
                                                  int seg_len, int nseg, const int2 *__restrict__ items,
                                                  int64_t nitems, unsigned long long *__restrict__ hist,
                                                  uint2 *__restrict__ carry, uint32_t *__restrict__ priv_base) {
    // one spare bin per lane up to 16 thresholds; at 32, 8 spare bins per row
    // keep six workgroups' LDS within the CU's
    constexpr int NSP = TG >= 32 ? 8 : WAVE;
    __shared__ uint32_t lh[TG * lrow<NSP>()];
    __shared__ __align__(16) uint32_t tl[WAVES][8 * HW];
    for (int e = threadIdx.x; e < TG * lrow<NSP>(); e += WG) lh[e] = 0;
    priv_zero<TG>(priv_slice<TG>(priv_base), nt, WG);
    __syncthreads();
    uint32_t *priv = priv_slice<TG>(priv_base);
    const Sink K{lh, lane_spare<NSP>(), priv, hist, nc};

    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = wave_id();
    uint32_t *tw = tl[wave];
    const int a = lane >> 5, rr = lane & 31;
    // this lane's record (r = lane) in the layout: word w at wofs + w*HW
    const int wofs = (lane & 7) * 8 + (lane >> 3);
    // the lane's row: word rr%8 of its row block's records, byte rr/8
    const uint4 *rsrc = (const uint4 *)(tw + (rr & 7) * HW + 4 * a);
    const uint32_t bsel = (uint32_t)(rr >> 3) | ((uint32_t)(4 + (rr >> 3)) << 8) | 0x0C0C0000u;
    const size_t cstride = (size_t)2 * nseg * nc;

    for (int64_t item = (int64_t)blockIdx.x * WAVES + wave; item < nitems; item += (int64_t)gridDim.x * WAVES) {
        const int2 it = items[item];
        const int64_t Jr = it.x;
        const int q = it.y;
        const int64_t j = Jr * WAVE + lane;
        const bool jv = j < nc;
        const int64_t hb = min((Jr + 1) * WAVE, nc);   // first column of the part
        const int64_t cb = hb + (int64_t)q * seg_len;
        const int64_t ce = cb + seg_len < nc ? cb + seg_len : nc;
        const int64_t I = 2 * Jr + a;                    // row block of this lane's row
        uint2 *cslot = carry + (size_t)(nseg + q) * nc + (jv ? j : 0);

        uint32_t run[TG];
#pragma unroll
        for (int t = 0; t < TG; t++) run[t] = jv ? PRE : 0u;

        // Row blocks past the matrix hold no tile (their lanes are idle).  The
        // lane's record travels from HBM one block ahead: v0, v1 hold block
        // c0 + BLK's while block c0 is gathered and scanned.
        const bool has_rows = BLK * I < nc;
        uint4 v0 = make_uint4(0u, 0u, 0u, 0u), v1 = v0;
        auto fetch = [&](int64_t c) {
            const uint4 *src = (const uint4 *)(tiles + tile_index(I, c / WAVE, nbands) * TILE_BYTES +
                                               (32 * ((c / BLK) & 1) + rr) * 32);
            v0 = src[0];
            v1 = src[1];
        };
        if (has_rows && cb < ce) fetch(cb);
        for (int64_t c0 = cb; c0 < ce; c0 += BLK) {
            const int nb = ce - c0 < BLK ? (int)(ce - c0) : BLK;
            wave_sync();   // the previous block's gathers are done
            if (has_rows) {
                uint32_t *d = tw + wofs;
                d[0 * HW] = v0.x; d[1 * HW] = v0.y; d[2 * HW] = v0.z; d[3 * HW] = v0.w;
                d[4 * HW] = v1.x; d[5 * HW] = v1.y; d[6 * HW] = v1.z; d[7 * HW] = v1.w;
                if (c0 + BLK < ce) fetch(c0 + BLK);
            }
            wave_sync();
            uint32_t R[8];
#pragma unroll
            for (int q = 0; q < 8; q++) {
                // words g = 0..3 are records 32a + q + 8g: byte rr/8 of each is row rr's step q + 8g
                const uint4 v = rsrc[2 * q];
                const uint32_t lo = __builtin_amdgcn_perm(v.y, v.x, bsel);
                const uint32_t hi = __builtin_amdgcn_perm(v.w, v.z, bsel);
                R[q] = __builtin_amdgcn_perm(hi, lo, 0x05040100u);
            }
            // columns past the part count as recurrent padding, taken back below
            const uint32_t fill = nb == BLK ? 0u : (0xFFFFFFFFu << nb);
            scan_words<TG, NSP>(R, jv ? ~fill : 0u, jv ? 0u : 0xFFFFFFFFu, run, nt, K, cslot, cstride);
            if (fill && jv) {
#pragma unroll
                for (int t = 0; t < TG; t++) run[t] -= (uint32_t)(BLK - nb);
            }
        }
        if (jv) close_segment<TG>(run, nt, cslot, cstride);
    }

    __syncthreads();
    flush_hist<TG, NSP>(lh, priv, nt, hist, nc, WG);
}

// ---------------------------------------------------------------- k_rp
// The recurrence plot (accrqa_RP, src/AccRQA_functions.cpp:996-1060):
// out[i N_c + j] = [s_ij <= B] (1 or 0) for the one threshold's bound B,
// row-major.  Work item = one block of 32 rows x 64 columns, one per wave;
// lane = column.  Row samples are wave-uniform (scalar) loads, as in k_rr;
// each row of the block is one store of 64 consecutive elements.  Reads
// reach x[N + 62] at most (the series buffer is padded with zeros).
template <typename F, int NORM, typename O>
__global__ __launch_bounds__(WG) void k_rp(const F *__restrict__ x, int64_t nc, int tau, int m, F bound,
                                           O *__restrict__ out) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = wave_id();
    const int64_t nbands = (nc + WAVE - 1) / WAVE;
    const int64_t nblk = ((nc + BLK - 1) / BLK) * nbands;
    for (int64_t b = (int64_t)blockIdx.x * WAVES + wave; b < nblk; b += (int64_t)gridDim.x * WAVES) {
        const int64_t I = b / nbands, J = b - I * nbands;
        const int64_t i0 = I * BLK;
        const int64_t j = J * WAVE + lane;
        F s[BLK];
#pragma unroll
        for (int u = 0; u < BLK; u++) s[u] = F(0);
        F xj = x[j];
        for (int k = 0; k < m; k++) {
            const F xjn = x[j + (int64_t)(k + 1 < m ? k + 1 : k) * tau];
            const F *xr = x + i0 + (int64_t)k * tau;
#pragma unroll
            for (int u = 0; u < BLK; u++) s[u] = dim_step<NORM>(s[u], xr[u], xj);
            xj = xjn;
        }
        if (j < nc) {
            const int nr = nc - i0 < BLK ? (int)(nc - i0) : BLK;
            O *o = out + i0 * nc + j;
#pragma unroll
            for (int u = 0; u < BLK; u++)
                if (u < nr) o[(int64_t)u * nc] = (O)(s[u] <= bound ? 1 : 0);
        }
    }
}

// ---------------------------------------------------------------- k_rp4
// k_rp with four columns per lane: work item = RP4_ROWS rows x 256 columns,
// lane l owns columns j0 + 4l .. j0 + 4l + 3, so each row of the block is one
// store of 4 elements per lane (256 consecutive elements per instruction,
// against 64 in k_rp).  The host picks it when N_c % 4 == 0 and the plot is
// aligned for those stores (4 bytes for chars, 16 for ints).  Column samples
// of columns past N_c are read at N_c - 1 (their elements are not stored);
// row reads reach x[N + RP4_ROWS - 2] (the series buffer is padded with zeros).
template <typename F>
constexpr int rp4_rows() { return sizeof(F) == 4 ? 16 : 8; }

template <typename F, int NORM, typename O>
__global__ __launch_bounds__(WG) void k_rp4(const F *__restrict__ x, int64_t nc, int tau, int m, F bound,
                                            O *__restrict__ out) {
    constexpr int ROWS = rp4_rows<F>();
    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = wave_id();
    const int64_t ncb = (nc + 4 * WAVE - 1) / (4 * WAVE);
    const int64_t nblk = ((nc + ROWS - 1) / ROWS) * ncb;
    for (int64_t b = (int64_t)blockIdx.x * WAVES + wave; b < nblk; b += (int64_t)gridDim.x * WAVES) {
        const int64_t I = b / ncb, Jb = b - I * ncb;
        const int64_t i0 = I * ROWS;
        const int64_t j = Jb * 4 * WAVE + 4 * lane;   // first of the lane's four columns
        int64_t jc[4];
#pragma unroll
        for (int c = 0; c < 4; c++) jc[c] = j + c < nc ? j + c : nc - 1;
        F s[ROWS][4];
#pragma unroll
        for (int u = 0; u < ROWS; u++)
#pragma unroll
            for (int c = 0; c < 4; c++) s[u][c] = F(0);
        for (int k = 0; k < m; k++) {
            F xj[4];
#pragma unroll
            for (int c = 0; c < 4; c++) xj[c] = x[jc[c] + (int64_t)k * tau];
            const F *xr = x + i0 + (int64_t)k * tau;
#pragma unroll
            for (int u = 0; u < ROWS; u++)
#pragma unroll
                for (int c = 0; c < 4; c++) s[u][c] = dim_step<NORM>(s[u][c], xr[u], xj[c]);
        }
        if (j < nc) {
            const int nr = nc - i0 < ROWS ? (int)(nc - i0) : ROWS;
            const bool full = j + 3 < nc;
#pragma unroll
            for (int u = 0; u < ROWS; u++) {
                if (u < nr) {
                    O *o = out + (i0 + u) * nc + j;
                    if (full) {
                        if constexpr (sizeof(O) == 1) {
                            uint32_t w = 0u;
#pragma unroll
                            for (int c = 0; c < 4; c++) w |= (s[u][c] <= bound ? 1u : 0u) << (8 * c);
                            *(uint32_t *)o = w;
                        } else {
                            *(uint4 *)o = make_uint4(s[u][0] <= bound ? 1u : 0u, s[u][1] <= bound ? 1u : 0u,
                                                     s[u][2] <= bound ? 1u : 0u, s[u][3] <= bound ? 1u : 0u);
                        }
                    } else {
#pragma unroll
                        for (int c = 0; c < 4; c++)
                            if (j + c < nc) o[c] = (O)(s[u][c] <= bound ? 1 : 0);
                    }
                }
            }
        }
    }
}

// ---------------------------------------------------------------- plot squares
// The recurrence plot is symmetric bit for bit: s_ij and s_ji are the same
// sums, since (a - b)^2 = (b - a)^2 and |a - b| = |b - a| exactly, and every
// form decides the reference's s.  The symmetric kernels (k_rp_sym,
// k_rp_mx_sym) decide the 64 x 64 squares with g <= J only (rows 64g..,
// columns 64J..) and store each one as is
// and mirrored (out[64J.., 64g..]), both as 16-byte row stores staged in
// the wave's LDS:
//   D, one 32-row block of the square, row-major (32 rows x 64 bytes),
//     stored after each block;
//   T, the mirror, one row per column of the square (64 rows x 64 bytes),
//     stored after the square's two blocks, except on the diagonal (g == J),
//     where D covers the square.
// Rows of both are padded to 80 bytes, so the 16-byte accesses of eight
// consecutive rows fall on all 32 banks.  The host takes these kernels only
// where N_c % 16 == 0 and the plot is 16-byte aligned: then every 16-byte
// store lies inside one row of the plot, and one is made only when its row
// and first column are below N_c.  Decisions of rows or columns past N_c
// come from the zero padding of the series and are not stored.
struct RpStage {
    static constexpr int RS = 80;   // bytes per row of D and T
    uint8_t d[BLK * RS];
    uint8_t t[WAVE * RS];
};

// Work item it -> square (g, J).  Squares go in groups of 16 x 16 (rows
// 1024G.., columns 1024H.., G <= H, group number H (H + 1) / 2 + G; square
// (g, J) = (16G + (it % 256) / 16, 16H + it % 16)), so that the waves at work
// at one time store their rows in runs of 1,024 bytes rather than 64
// (measured: groups of 8 and 32 were slower).  False for the squares of a
// group below the diagonal or past the last band.
constexpr int RP_GROUP = 16;
__device__ __forceinline__ bool rp_square(int64_t it, int64_t nbands, int64_t &g, int64_t &J) {
    const int64_t grp = it / (RP_GROUP * RP_GROUP);
    const int loc = (int)(it - grp * (RP_GROUP * RP_GROUP));
    int64_t H = (int64_t)((__builtin_sqrt(8.0 * (double)grp + 1.0) - 1.0) * 0.5);
    while (H > 0 && H * (H + 1) / 2 > grp) H--;
    while ((H + 1) * (H + 2) / 2 <= grp) H++;
    g = RP_GROUP * (grp - H * (H + 1) / 2) + loc / RP_GROUP;
    J = RP_GROUP * H + loc % RP_GROUP;
    return g <= J && J < nbands;
}

// The wave's first work item: its slot among the grid's waves, numbered XCD
// by XCD.  Workgroups go to the eight XCDs in turn (workgroup b to XCD
// b % 8), so with a grid of a multiple of 8 workgroups (the host rounds it)
// the squares of one group are taken by waves of one XCD, whose L2 then
// sees both halves of every 128-byte line the group stores.
__device__ __forceinline__ int64_t rp_first_item(int wave) {
    if (gridDim.x % 8) return (int64_t)blockIdx.x * WAVES + wave;
    return ((int64_t)(blockIdx.x % 8) * (gridDim.x / 8) + blockIdx.x / 8) * WAVES + wave;
}

// Work items of a plot of nbands bands: whole groups of squares.
__host__ __device__ inline int64_t rp_items(int64_t nbands) {
    const int64_t ng = (nbands + RP_GROUP - 1) / RP_GROUP;
    return ng * (ng + 1) / 2 * (RP_GROUP * RP_GROUP);
}

// D to out[i0 + row][j0 + 16 ch ..]: lane -> row (lane >> 2) + 16s, ch = lane & 3.
__device__ __forceinline__ void rp_store_block(const RpStage &S, int lane, int64_t i0, int64_t j0, int64_t nc,
                                               char *__restrict__ out) {
#pragma unroll
    for (int s = 0; s < 2; s++) {
        const int row = (lane >> 2) + 16 * s, ch = lane & 3;
        const uint4 v = *(const uint4 *)(S.d + row * RpStage::RS + 16 * ch);
        const int64_t i = i0 + row, jc = j0 + 16 * ch;
        if (i < nc && jc < nc) *(uint4 *)(out + i * nc + jc) = v;
    }
}

// T to out[j0 + row][r0 + 16 ch ..]: lane -> row (lane >> 2) + 16s, ch = lane & 3.
__device__ __forceinline__ void rp_store_mirror(const RpStage &S, int lane, int64_t j0, int64_t r0, int64_t nc,
                                                char *__restrict__ out) {
#pragma unroll
    for (int s = 0; s < 4; s++) {
        const int row = (lane >> 2) + 16 * s, ch = lane & 3;
        const uint4 v = *(const uint4 *)(S.t + row * RpStage::RS + 16 * ch);
        const int64_t jr = j0 + row, ic = r0 + 16 * ch;
        if (jr < nc && ic < nc) *(uint4 *)(out + jr * nc + ic) = v;
    }
}

// ---------------------------------------------------------------- k_rp_sym
// The symmetric recurrence plot in the exact form, bytes: per 32-row block
// of a square, lane = column, the sums of k_rp; the lane's 32 decisions,
// four rows per word, go to D one byte at a time and to T as two 16-byte
// LDS stores.  Reads are k_rp's.
template <typename F, int NORM>
__global__ __launch_bounds__(WG) void k_rp_sym(const F *__restrict__ x, int64_t nc, int tau, int m, F bound,
                                               char *__restrict__ out) {
    __shared__ __attribute__((aligned(16))) RpStage stg[WAVES];
    constexpr int RS = RpStage::RS;
    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = wave_id();
    RpStage &S = stg[wave];
    const int64_t nbands = (nc + WAVE - 1) / WAVE;
    const int64_t nsq = rp_items(nbands);
    for (int64_t it = rp_first_item(wave); it < nsq; it += (int64_t)gridDim.x * WAVES) {
        int64_t g, J;
        if (!rp_square(it, nbands, g, J)) continue;
        const int64_t j0 = J * WAVE, r0 = g * WAVE;
        const int64_t j = j0 + lane;
        for (int k = 0; k < 2; k++) {
            const int64_t i0 = r0 + BLK * k;
            if (i0 >= nc) break;
            F s[BLK];
#pragma unroll
            for (int u = 0; u < BLK; u++) s[u] = F(0);
            F xj = x[j];
            for (int kk = 0; kk < m; kk++) {
                const F xjn = x[j + (int64_t)(kk + 1 < m ? kk + 1 : kk) * tau];
                const F *xr = x + i0 + (int64_t)kk * tau;
#pragma unroll
                for (int u = 0; u < BLK; u++) s[u] = dim_step<NORM>(s[u], xr[u], xj);
                xj = xjn;
            }
            // byte c of word q: row 4q + c
            uint32_t w[8];
#pragma unroll
            for (int q = 0; q < 8; q++) {
                w[q] = 0u;
#pragma unroll
                for (int c = 0; c < 4; c++) w[q] |= (s[4 * q + c] <= bound ? 1u : 0u) << (8 * c);
            }
            wave_sync();   // the last block's D and the last square's T are stored
            uint8_t *dd = S.d + lane;
#pragma unroll
            for (int u = 0; u < BLK; u++) dd[u * RS] = (uint8_t)(w[u >> 2] >> (8 * (u & 3)));
            uint4 *tt = (uint4 *)(S.t + lane * RS + 32 * k);
            tt[0] = make_uint4(w[0], w[1], w[2], w[3]);
            tt[1] = make_uint4(w[4], w[5], w[6], w[7]);
            wave_sync();
            rp_store_block(S, lane, i0, j0, nc, out);
        }
        if (g < J) rp_store_mirror(S, lane, j0, r0, nc, out);
    }
}

// ---------------------------------------------------------------- k_rr
// Work item = (band of 64 columns, row segment q) over rows i <= max column
// of the band.  Pair weight: 2 if i < j, 1 if i == j, 0 otherwise.
template <typename F, int NORM, int TG>
__global__ __launch_bounds__(WG) void k_rr(const F *__restrict__ x, int64_t nc, int tau, int m, Bounds<F, TG> S,
                                           int nt, const LutEntry<F> *__restrict__ lut, LutParams P, int seg_len,
                                           const int2 *__restrict__ items, int64_t nitems,
                                           unsigned long long *__restrict__ counts) {
    __shared__ unsigned long long acc[TG];
    __shared__ LutEntry<F> lut_s[NB_MAX];
    __shared__ F bnd_s[TG];
    for (int e = threadIdx.x; e < TG; e += WG) acc[e] = 0;
    load_lut(lut_s, lut, P.nb);
    if (threadIdx.x < TG) bnd_s[threadIdx.x] = S.s[threadIdx.x];
    __syncthreads();
    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = wave_id();
    constexpr int B = rank_bits<TG>();

    for (int64_t item = (int64_t)blockIdx.x * WAVES + wave; item < nitems; item += (int64_t)gridDim.x * WAVES) {
        const int2 it = items[item];
        const int64_t j0 = (int64_t)it.x * WAVE;
        const int64_t j = j0 + lane;
        const bool jv = j < nc;
        const int64_t rows = (j0 + WAVE < nc ? j0 + WAVE : nc);   // rows 0..rows-1 can be <= j
        const int64_t ib = (int64_t)it.y * seg_len;
        const int64_t ie = ib + seg_len < rows ? ib + seg_len : rows;

        uint32_t c2[TG], c1[TG];
#pragma unroll
        for (int t = 0; t < TG; t++) c2[t] = c1[t] = 0;

        for (int64_t i0 = ib; i0 < ie; i0 += BLK) {
            F s[BLK];
#pragma unroll
            for (int u = 0; u < BLK; u++) s[u] = F(0);
            F xj = x[j];
            for (int k = 0; k < m; k++) {
                const F xjn = x[j + (int64_t)(k + 1 < m ? k + 1 : k) * tau];
                const F *xr = x + i0 + (int64_t)k * tau;
#pragma unroll
                for (int u = 0; u < BLK; u++) s[u] = dim_step<NORM>(s[u], xr[u], xj);
                xj = xjn;
            }
            uint32_t R[8];
            pack_ranks<F, TG>(s, R, lut_s, P, bnd_s, nt);
            uint32_t Pl[B];
            rank_planes<B>(R, Pl);
            const GeTree<B> G(Pl);
            // rows i0+u with u < nrow are in the segment; weight 2 for i < j, 1 for i == j
            const int64_t nr64 = ie - i0;
            const uint32_t mrow = nr64 >= BLK ? 0xFFFFFFFFu : ((1u << nr64) - 1u);
            const int64_t du = j - i0;   // step index of the diagonal element
            const uint32_t mup = !jv ? 0u : (du >= BLK ? 0xFFFFFFFFu : (du <= 0 ? 0u : ((1u << du) - 1u)));
            const uint32_t mdg = (jv && du >= 0 && du < BLK) ? (1u << du) : 0u;
#pragma unroll
            for (int t = 0; t < TG; t++) {
                if (t < nt) {
                    const uint32_t w = ~G.ge(t + 1) & mrow;
                    c2[t] += __popc(w & mup);
                    c1[t] += __popc(w & mdg);
                }
            }
        }
#pragma unroll
        for (int t = 0; t < TG; t++) {
            if (t < nt) {
                unsigned long long v = 2ull * c2[t] + c1[t];
                for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
                if (lane == 0 && v) atomicAdd(&acc[t], v);
            }
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < nt; e += WG)
        if (acc[e]) glob_add(&counts[e], acc[e]);
}

// ---------------------------------------------------------------- k_fp64_diag
// Diagnostic for fp32 input (no counterpart in the reference): for each
// threshold, the pairs of the full matrix whose decision in the reference's
// fp32 arithmetic differs from the decision in true fp64 arithmetic on the
// same fp32 samples, i.e. [s32 <= B32_t] != [s64 <= B64_t] with s in the
// reference's order in each precision and B the exact bound of each
// precision.  Work items as k_rr (band of 64 columns, row segment, rows up to
// the band's last column), weight 2 off the diagonal.
constexpr int DIAG_MAXT = 32;

template <int NORM>
__global__ __launch_bounds__(WG) void k_fp64_diag(const float *__restrict__ x, int64_t nc, int tau, int m,
                                                  Bounds<float, DIAG_MAXT> B32, Bounds<double, DIAG_MAXT> B64,
                                                  int nt, int seg_len, const int2 *__restrict__ items, int64_t nitems,
                                                  unsigned long long *__restrict__ counts) {
    __shared__ unsigned long long acc[DIAG_MAXT];
    for (int e = threadIdx.x; e < DIAG_MAXT; e += WG) acc[e] = 0;
    __syncthreads();
    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = wave_id();
    for (int64_t item = (int64_t)blockIdx.x * WAVES + wave; item < nitems; item += (int64_t)gridDim.x * WAVES) {
        const int2 it = items[item];
        const int64_t j0 = (int64_t)it.x * WAVE;
        const int64_t j = j0 + lane;
        const bool jv = j < nc;
        const int64_t rows = (j0 + WAVE < nc ? j0 + WAVE : nc);
        const int64_t ib = (int64_t)it.y * seg_len;
        const int64_t ie = ib + seg_len < rows ? ib + seg_len : rows;
        uint32_t c[DIAG_MAXT];
#pragma unroll
        for (int t = 0; t < DIAG_MAXT; t++) c[t] = 0;
        for (int64_t i = ib; i < ie; i++) {
            float s32 = 0.f;
            double s64 = 0.0;
            for (int k = 0; k < m; k++) {
                const float a = x[i + (int64_t)k * tau], b = x[j + (int64_t)k * tau];
                s32 = dim_step<NORM>(s32, a, b);
                s64 = dim_step<NORM>(s64, (double)a, (double)b);
            }
            const uint32_t w = !jv || i > j ? 0u : (i == j ? 1u : 2u);
#pragma unroll
            for (int t = 0; t < DIAG_MAXT; t++)
                if (t < nt) c[t] += ((s32 <= B32.s[t]) != (s64 <= B64.s[t])) ? w : 0u;
        }
#pragma unroll
        for (int t = 0; t < DIAG_MAXT; t++) {
            if (t < nt) {
                unsigned long long v = c[t];
                for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
                if (lane == 0 && v) atomicAdd(&acc[t], v);
            }
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < nt; e += WG)
        if (acc[e]) glob_add(&counts[e], acc[e]);
}

// ================================================================ MFMA form
// Euclidean norm, fp32 input.  The squared distance is approximated as
//   s~ = (n_i + n_j) - 2 g_ij,   g = x_i . x_j on MFMA,
// n_i = |x_i|^2 (accumulated in fp64, rounded to fp32).  The Gram runs in
// FP16 hi/lo form (v_mfma_f32_32x32x16_f16): each sample, scaled by 2^e so
// that |a| < 2^14, is split as a = h + l + r with h = fp16(a), l = fp16(a - h)
// (both round to nearest, subnormals kept), |r| <= 2^-22 |a| + 2^-25; the
// products h.l, l.h and then h.h are accumulated, l.l is dropped.  With
// u = 2^-24, C = ceil(m/16) and A, B the scaled norms, in units of s~:
//  * split: 2 sum_k (|l_a l_b| + |r_a b| + |a r_b| + |r_a r_b|)
//      <= 14.04u (A + B) + m 2^-25.9, using sqrt(mA) <= 2A + m/8;
//  * accumulation, allowing each MFMA up to 16 additions of relative error
//    2u (a faithful adder, any order; h.h of every chunk after all small
//    products): 2 * 32.2C u (A + B) / 2.  Measured on the instruction: its
//    error stays within 4.4u (|c| + sum |p~|), where p~ counts an fp16
//    subnormal operand at 2^-14, the smallest normal (a product with one is
//    aligned by that exponent, so against the true products the error can
//    reach 10^3 u).  With every operand raised to 2^-14 the bound grows by
//    under 0.03u (A + B) and m 2^-32 in scaled units, inside the spare and dA.
// For the reference's exact-order fp32 sum s:
//   |s - s~| <= (3m + 31 + 33C) u (n_i + n_j) + m 2^-25 2^-2e + 1e-30 =: E_ij
// (exact-order sum: gamma_{m+1} * 2(n_i+n_j); rounded norms 2u; the Gram as
// above; the final add and fma 3u; the rest spare).
// A byte table over the bits of s~ gives the rank r = #{t : B_t < s} exactly
// for every bucket that has no bound within E_max = max E_ij of its value
// range; a bucket with one bound within reach names that bound, and its pairs
// are decided against it with their own E_ij (second stage); pairs still
// within E_ij of a bound are re-decided in exact order on VALU.  Inputs whose
// norms are not finite or exceed 1e30 flag every bucket, so the whole matrix
// is then decided exactly.
constexpr int NBX = 16384;       // byte buckets of the MFMA rank table (LDS), DET pass
constexpr int NBX_LAM = 32768;   // the same for the LAM pass (which has the LDS to spare)
constexpr int NBX_RP = 4096;     // the same for the symmetric plot (four workgroups per CU)
constexpr int MXW = 128;     // LDS row window: 32 + (m-1)*tau samples at most
constexpr int NBND = 64;     // bound slots reachable by a 6-bit table index

struct MxParams {
    uint32_t base;   // key bits of bucket 0 (half the smallest bound, rounded down to a bucket edge)
    int sh;          // bucket width in key bits (log2)
    int nb;          // buckets used (<= NBX)
    int qlo;         // base >> sh: bucket 0's number on the key >> sh scale
    float cE;        // (3m + 31 + 33 ceil(m/16)) * 2^-24 (+ 4 * 2^-24 for fp64 input)
    float dA;        // m * 2^-25: the additive part of E_ij in scaled units (x 2^-2e)
    int amax;        // scaled samples lie below 2^amax (14; 8 for the FP8 Gram)
};

// Precision of the Gram's operands.  GRAM_HILO is the form every path uses;
// the others exist for the measured comparison of Gram precisions (split
// form, fp32 input, up to 16 thresholds) and differ only in the window word,
// the matrix instruction and the error bound (cE, dA, amax; engine
// mx_params), so the table, the second stage and the exact re-decision keep
// every decision exact whatever the precision:
//   GRAM_F16   one FP16 half per sample (h.h only): |a - h| <= 2^-11 |a| or
//              2^-25 (subnormal h);
//   GRAM_BF16  one BF16 half per sample: |a - bf16(a)| <= 2^-8 |a| or
//              2^-126 (a flushed fp32 subnormal);
//   GRAM_FP8   one FP8 (E4M3) byte per sample, |a| < 2^8: the bound allows
//              2^-3 |a| (round-to-nearest gives 2^-4) or 2^-6 for the
//              subnormal range (a flushed byte included), and an
//              accumulation error of 2^-9 (|c| + sum |p|) per instruction.
constexpr int GRAM_HILO = 0, GRAM_F16 = 1, GRAM_BF16 = 2, GRAM_FP8 = 3;

// Scale of the FP16 hi/lo Gram, from the largest norm: samples enter the
// windows as a = x 2^e with |a| < 2^amax = 2^14 (every sample is a coordinate of some
// row, so |x| <= sqrt(n_max)); the Gram comes back through c2 = -2^(1-2e) in
// the fma that forms s~ (exact: a power of two); dE = m 2^-25 2^-2e + 1e-30.
// Norms that are zero, not finite or above 1e30 keep e = 0 (the last two
// flag every bucket anyway).
struct MxScale {
    float s, c2, dE;
};

__device__ __forceinline__ MxScale mx_scale(const unsigned int *__restrict__ nmax_bits, const MxParams &P) {
    const uint32_t nb = *nmax_bits;
    const float nmax = __uint_as_float(nb);
    int e = 0;
    if (nmax > 0.0f && nmax <= 1e30f) {
        const int q = (int)(nb >> 23) - 126;   // n_max < 2^q
        e = P.amax - ((q + 1) >> 1);           // sqrt(n_max) 2^e < 2^amax
        e = e < -60 ? -60 : (e > 60 ? 60 : e);
    }
    MxScale Z;
    Z.s = __int_as_float((127 + e) << 23);
    Z.c2 = -__int_as_float((128 - 2 * e) << 23);
    Z.dE = fmaf(P.dA, __int_as_float((127 - 2 * e) << 23), 1e-30f);
    return Z;
}

// n_i = sum_k x[i + k tau]^2 (fp64, rounded to fp32); entries nc .. nc+127 = 0.
// nmax_bits: atomic max of the float bits (non-negative floats order as uints;
// NaN bits are above +inf, so a NaN norm forces the all-exact fallback).
template <typename F>
__global__ void k_norms(const F *__restrict__ x, int64_t nc, int tau, int m, float *__restrict__ nrm,
                        unsigned int *__restrict__ nmax_bits) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc + 2 * WAVE) return;
    float v = 0.f;
    if (i < nc) {
        double a = 0.0;
        for (int k = 0; k < m; k++) {
            const double xv = (double)x[i + (int64_t)k * tau];
            a += xv * xv;
        }
        v = (float)a;
        const float mag = fabsf(v);
        atomicMax(nmax_bits, __float_as_uint(mag > 1e30f || mag != mag ? __int_as_float(0x7F800000) : mag));
    }
    nrm[i] = v;
}

// Byte table.  Bucket without any bound within E_max of its value range:
// bits 0-5 = rank (number of bounds below it), bits 6-7 clear.  Otherwise bit
// 7 set and bits 0-5 = index of the one bound within reach (bit 6 set if there
// are several): such pairs are checked against that bound with their own
// E_ij, and only the truly ambiguous ones are re-decided in exact order.
template <int TG>
__global__ void k_mx_lut(Bounds<float, TG> S, int nt, MxParams P, const unsigned int *__restrict__ nmax_bits,
                         uint8_t *__restrict__ lut) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= P.nb) return;
    const float nmax = __uint_as_float(*nmax_bits);
    const MxScale Z = mx_scale(nmax_bits, P);
    float emax = P.cE * 2.0f * nmax;
    emax = (emax + Z.dE) * 1.0001f;
    if (!(nmax <= 1e30f)) emax = __int_as_float(0x7F800000);
    const uint32_t k0 = P.base + ((uint32_t)b << P.sh);
    const uint64_t k1l = (uint64_t)P.base + ((uint64_t)(b + 1) << P.sh);
    const float v0 = __uint_as_float(k0);
    const float v1 = k1l >= 0x7F800000ull ? __int_as_float(0x7F800000) : __uint_as_float((uint32_t)k1l);
    const bool last = b == P.nb - 1;
    // keys below bucket 0 and above the last bucket are clamped into them
    const float lo = b == 0 ? 0.0f : v0;
    const float hi = last ? __int_as_float(0x7F800000) : v1;
    int r_lo = 0, near = 0, idx = 0;
    for (int t = 0; t < nt; t++) {
        const float B = S.s[t];
        if (B < v0) r_lo++;
        if (B >= lo - emax && B <= hi + emax) {
            if (near == 0) idx = t;
            near++;
        }
    }
    // several bounds within reach: index NBND - 1, the NaN slot, so that the
    // second stage finds every such pair ambiguous without a test of its own
    uint32_t e = (uint32_t)r_lo;
    if (near == 1) e = 0x80u | (uint32_t)idx;
    if (near >= 2) e = 0xC0u | (uint32_t)(NBND - 1);
    if (!(emax < __int_as_float(0x7F800000))) e = 0xC0u | (uint32_t)(NBND - 1);   // non-finite or huge norms: all exact
    lut[b] = (uint8_t)e;
}

// Bucket of s~ (bucket edges are multiples of 2^sh in key bits): key >> sh,
// clamped to the table by one median-of-three.  A slightly negative s~
// (rounding near 0) has a negative bit pattern as an integer and lands in
// bucket 0, like +0; keys below the table saturate to bucket 0 and keys above
// it to the last bucket.
__device__ __forceinline__ uint32_t mx_key_bucket(float st, const MxParams &P) {
    const int q = __float_as_int(st) >> P.sh;
    int qc;
    asm("v_med3_i32 %0, %1, %2, %3" : "=v"(qc) : "v"(q), "v"(P.qlo), "v"(P.qlo + P.nb - 1));
    return (uint32_t)(qc - P.qlo);
}

// Second stage of one flagged element (table byte with bit 7 set): the one
// bound within reach decides unless it lies within E_ij of s~, or several
// bounds were within reach; amb is set for those pairs.  A byte for several
// bounds names bound slot NBND - 1, which holds NaN: d is then NaN and the
// pair ambiguous (its rank byte, NBND - 1 or NBND, is replaced by the exact
// re-decision; padding elements keep it and are masked out).  cE and dE carry
// the margin for the rounding of E (mx_params).  st need not be clamped at 0:
// a negative st decides "below the bound" only when st + E_ij is, and the
// true s lies within E_ij of st.
__device__ __forceinline__ uint32_t mx_second(uint32_t byte, float st, float sum, const float *bnd, float cE,
                                              float dE, bool &amb) {
    const uint32_t idx = byte & 0x3Fu;
    const float d = st - bnd[idx];
    const float E = fmaf(cE, sum, dE);
    amb = !(fabsf(d) > E);
    return idx + (d > 0.0f ? 1u : 0u);
}

// Bound slot t of the MFMA kernels' LDS copy: the group's bounds, +inf past
// them, NaN in the last slot (mx_second).
__device__ __forceinline__ float mx_bound_slot(int t, int nt, const float *s) {
    return t == NBND - 1 ? __int_as_float(0x7FC00000) : (t < nt ? s[t] : __int_as_float(0x7F800000));
}

// Exact-order fp32 squared distance of one pair (the reference's sum).
// Loads go out four dimensions at a time; the sum keeps the order k = 0..m-1.
template <typename F>
__device__ __forceinline__ F exact_sqdist(const F *__restrict__ x, int64_t i, int64_t j, int tau, int m) {
    F s = F(0);
    for (int k0 = 0; k0 < m; k0 += 4) {
        F a[4], b[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int64_t o = (int64_t)(k0 + k < m ? k0 + k : m - 1) * tau;
            a[k] = x[i + o];
            b[k] = x[j + o];
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (k0 + k < m) {
                const F d = a[k] - b[k];
                const F q = d * d;
                s = s + q;
            }
        }
    }
    return s;
}

// Wave-wide exclusive prefix sum of a small per-lane count.
__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v, uint32_t &total) {
    const int lane = threadIdx.x & (WAVE - 1);
    uint32_t inc = v;
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        const uint32_t o = __shfl_up(inc, off);
        if (lane >= off) inc += o;
    }
    total = __shfl(inc, WAVE - 1);
    return inc - v;
}

typedef float floatx16 __attribute__((ext_vector_type(16)));

typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// One window word: the scaled sample a split as h = fp16(a) (bits 0..15) and
// l = fp16(a - h) (bits 16..31); a - h is exact in fp32.
__device__ __forceinline__ float hl_word(float a) {
    const _Float16 hi = (_Float16)a;
    const _Float16 lo = (_Float16)(a - (float)hi);
    return __uint_as_float((uint32_t)__builtin_bit_cast(uint16_t, hi) |
                           ((uint32_t)__builtin_bit_cast(uint16_t, lo) << 16));
}

// The window word of the other Gram precisions (GRAM_*): bf16(a) in bits
// 0..15 (round to nearest even), or the FP8 (E4M3) byte of a in bits 0..7.
template <int G>
__device__ __forceinline__ float win_word(float a) {
    if constexpr (G == GRAM_BF16) {
        return __uint_as_float((uint32_t)__builtin_bit_cast(uint16_t, (__bf16)a));
    } else if constexpr (G == GRAM_FP8) {
        return __uint_as_float((uint32_t)__builtin_amdgcn_cvt_pk_fp8_f32(a, 0.0f, 0, false) & 0xFFu);
    } else {
        return hl_word(a);
    }
}

// Gram tile G[a][b] = x_(i0+a) . x_(c0+b), a, b < 32, on MFMA.  Lane l
// supplies A[row l%32][k] and B[k][col l%32] for the eight dimensions
// k = 8(l/32) + t, t < 8, of each chunk of 16 (slot t; A and B take the same
// slots, so the sum over k does not depend on the instruction's slot order).
// Per-wave operand windows of the MFMA kernels: rw holds the window words of
// the block's rows, cw those of its NT column tiles (tile p at +32).  Both
// are read at one per-lane offset o (cw tile p at o + 32p), so one LDS
// address per operand step serves the row operand and every column tile,
// the tiles being immediate offsets from it.  Slot ZS of rw and slots
// ZS + 32p of cw hold 0: the operands of dimensions k >= m (both factors 0,
// so the product is 0 whatever else the windows hold).
template <int NT>
struct MxWin {
    static constexpr int ZS = MXW + 32 * (NT - 1);   // past the content of both windows
    static constexpr int CWO = ZS + 1;                // cw[0] relative to rw[0], in words
    float rw[ZS + 1];
    float cw[ZS + 32 * (NT - 1) + 1];
};

template <int NT>
__device__ __forceinline__ void mx_win_init(MxWin<NT> &W, int lane) {
    if (lane == 0) W.rw[MxWin<NT>::ZS] = 0.f;
    if (lane < NT) W.cw[MxWin<NT>::ZS + 32 * lane] = 0.f;
}

// Window address of dimension k for lane (col, h): col + k*tau, or the zero
// slot for k >= m.  The first chunk's eight (k = 8h + t) depend only on the
// lane, m and tau.
template <int NT>
__device__ __forceinline__ const float *gram_addr(const MxWin<NT> &W, int col, int k, int m, int tau) {
    return W.rw + (k < m ? col + k * tau : MxWin<NT>::ZS);
}

template <int NT>
__device__ __forceinline__ void gram_addrs(const MxWin<NT> &W, int col, int h, int m, int tau, const float *(&pa)[8]) {
#pragma unroll
    for (int t = 0; t < 8; t++) pa[t] = gram_addr(W, col, 8 * h + t, m, tau);
}

// The h halves (sel 0x05040100) or l halves (0x07060302) of eight window
// words, as the eight fp16 operands of one MFMA.
template <uint32_t SEL>
__device__ __forceinline__ halfx8 hl_halves(const uint32_t (&w)[8]) {
    u32x4 v;
#pragma unroll
    for (int i = 0; i < 4; i++) v[i] = __builtin_amdgcn_perm(w[2 * i + 1], w[2 * i], SEL);
    return __builtin_bit_cast(halfx8, v);
}

__device__ __forceinline__ floatx16 mfma16(halfx8 a, halfx8 b, floatx16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
}

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// The eight low bytes of eight window words as the 64-bit FP8 operand
// (slot t = byte t).
__device__ __forceinline__ long fp8_bytes(const uint32_t (&w)[8]) {
    uint32_t v[2];
#pragma unroll
    for (int i = 0; i < 2; i++) {
        const uint32_t lo = __builtin_amdgcn_perm(w[4 * i + 1], w[4 * i], 0x0C0C0400u);
        const uint32_t hi = __builtin_amdgcn_perm(w[4 * i + 3], w[4 * i + 2], 0x04000C0Cu);
        v[i] = lo | hi;
    }
    return (long)(((unsigned long long)v[1] << 32) | v[0]);
}

// One MFMA of the Gram over one chunk of 16 dimensions in precision G (not
// GRAM_HILO, whose three products gram_tile issues itself).
template <int G>
__device__ __forceinline__ floatx16 mfma_plain(const uint32_t (&a)[8], const uint32_t (&b)[8], floatx16 c) {
    constexpr uint32_t SH = 0x05040100u;
    if constexpr (G == GRAM_BF16) {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, hl_halves<SH>(a)),
                                                       __builtin_bit_cast(bf16x8, hl_halves<SH>(b)), c, 0, 0, 0);
    } else if constexpr (G == GRAM_FP8) {
        return __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(fp8_bytes(a), fp8_bytes(b), c, 0, 0, 0);
    } else {
        return mfma16(hl_halves<SH>(a), hl_halves<SH>(b), c);
    }
}

// One MFMA of the Gram's instruction on given operands (diagnostic, one
// wave): a = A[32][16], b = B[16][32] as fp16 bits, row-major; c = the
// accumulator C[32][32] in, d = C + A B out.  Lanes supply the operands as
// gram_tile does (dimension 8h + t in slot t) and hold C in its layout.
#ifndef RQA_HELPERS_ONLY
__global__ void k_mfma16_probe(const uint16_t *__restrict__ a, const uint16_t *__restrict__ b,
                               const float *__restrict__ c, float *__restrict__ d) {
    const int lane = threadIdx.x & (WAVE - 1), col = lane & 31, h = lane >> 5;
    halfx8 av, bv;
#pragma unroll
    for (int t = 0; t < 8; t++) {
        av[t] = __builtin_bit_cast(_Float16, a[col * 16 + 8 * h + t]);
        bv[t] = __builtin_bit_cast(_Float16, b[(8 * h + t) * 32 + col]);
    }
    floatx16 acc;
#pragma unroll
    for (int e = 0; e < 16; e++) acc[e] = c[(8 * (e >> 2) + 4 * h + (e & 3)) * 32 + col];
    acc = mfma16(av, bv, acc);
#pragma unroll
    for (int e = 0; e < 16; e++) d[(8 * (e >> 2) + 4 * h + (e & 3)) * 32 + col] = acc[e];
}
#endif

// One MFMA of the other Gram precisions on given operands (diagnostic, one
// wave): a = A[32][16], b = B[16][32], c = C[32][32], row-major fp32; each
// operand is rounded as the window words of precision G are (win_word), qa
// and qb get the rounded values as the hardware decodes them, and d = C +
// qa qb from one mfma_plain<G> in the kernels' lane layout.
template <int G>
__global__ void k_gram_probe(const float *__restrict__ a, const float *__restrict__ b, const float *__restrict__ c,
                             float *__restrict__ d, float *__restrict__ qa, float *__restrict__ qb) {
    const int lane = threadIdx.x & (WAVE - 1), col = lane & 31, h = lane >> 5;
    auto dec = [](uint32_t w) -> float {
        if constexpr (G == GRAM_BF16) return __uint_as_float(w << 16);
        else if constexpr (G == GRAM_FP8) return __builtin_amdgcn_cvt_f32_fp8((int)w, 0);
        else return (float)__builtin_bit_cast(_Float16, (uint16_t)(w & 0xFFFFu));
    };
    uint32_t wa[8], wb[8];
#pragma unroll
    for (int t = 0; t < 8; t++) {
        wa[t] = __float_as_uint(win_word<G>(a[col * 16 + 8 * h + t]));
        wb[t] = __float_as_uint(win_word<G>(b[(8 * h + t) * 32 + col]));
        qa[col * 16 + 8 * h + t] = dec(wa[t]);
        qb[(8 * h + t) * 32 + col] = dec(wb[t]);
    }
    floatx16 acc;
#pragma unroll
    for (int e = 0; e < 16; e++) acc[e] = c[(8 * (e >> 2) + 4 * h + (e & 3)) * 32 + col];
    acc = mfma_plain<G>(wa, wb, acc);
#pragma unroll
    for (int e = 0; e < 16; e++) d[(8 * (e >> 2) + 4 * h + (e & 3)) * 32 + col] = acc[e];
}

// One 32x32 Gram tile (column tile p) in scaled units: G[row][32p + col]
// over the m dimensions.  Order of accumulation: l.h and h.l of every chunk,
// then h.h of every chunk (the bound on the adds is taken in that order).
template <int NT, int p, int G = GRAM_HILO>
__device__ __forceinline__ floatx16 gram_tile(const MxWin<NT> &W, const float *const (&pa)[8], int col, int h, int m,
                                              int tau) {
    constexpr int CW = MxWin<NT>::CWO + 32 * p;
    constexpr uint32_t SH = 0x05040100u, SL = 0x07060302u;
    floatx16 acc = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    uint32_t a[8], b[8];
    // words of chunk k0 (the first chunk's addresses are pa)
    auto load = [&](int k0) {
#pragma unroll
        for (int t = 0; t < 8; t++) {
            const float *q = k0 == 0 ? pa[t] : gram_addr(W, col, k0 + 8 * h + t, m, tau);
            a[t] = __float_as_uint(q[0]);
            b[t] = __float_as_uint(q[CW]);
        }
    };
    if constexpr (G != GRAM_HILO) {
        // one product per chunk, in chunk order
        for (int k0 = 0; k0 < m; k0 += 16) {
            load(k0);
            acc = mfma_plain<G>(a, b, acc);
        }
        return acc;
    }
    if (m <= 16) {
        load(0);
        const halfx8 ah = hl_halves<SH>(a), bh = hl_halves<SH>(b);
        acc = mfma16(hl_halves<SL>(a), bh, acc);
        acc = mfma16(ah, hl_halves<SL>(b), acc);
        return mfma16(ah, bh, acc);
    }
    // more chunks: the small products of all, then h.h chunk by chunk
    for (int k0 = 0; k0 < m; k0 += 16) {
        load(k0);
        acc = mfma16(hl_halves<SL>(a), hl_halves<SH>(b), acc);
        acc = mfma16(hl_halves<SH>(a), hl_halves<SL>(b), acc);
    }
    for (int k0 = 0; k0 < m; k0 += 16) {
        load(k0);
        acc = mfma16(hl_halves<SH>(a), hl_halves<SH>(b), acc);
    }
    return acc;
}

// Ranks of one Gram tile's 16 elements of this lane, packed one per byte in
// the layout of the tile (element e -> byte e>>2 of Rq[e&3]).  First stage:
// table bytes, flags included.  Second stage, for flagged bytes only: the
// bound within reach decides with the pair's own E_ij; fm gets the elements
// still ambiguous after that (a few in 10^5 pairs), which the caller
// re-decides in exact order.  ni[e]: norm of element e's row, nj: norm of the
// lane's column.
__device__ __forceinline__ void tile_ranks(const floatx16 &acc, const float (&ni)[16], float nj, const uint8_t *lut_s,
                                           const MxParams &P, const MxScale &Z, const float *bnd_s, uint32_t (&Rq)[4],
                                           uint32_t &fm, uint32_t &flw) {
#pragma unroll
    for (int c = 0; c < 4; c++) Rq[c] = 0u;
#pragma unroll
    for (int e = 0; e < 16; e++) {
        const float st = fmaf(Z.c2, acc[e], ni[e] + nj);
        Rq[e & 3] |= (uint32_t)lut_s[mx_key_bucket(st, P)] << (8 * (e >> 2));
    }
    fm = 0u;
    // this lane's flagged elements: bit 8g + c <=> element 4g + c (table byte bit 7)
    const uint32_t fl = ((Rq[0] >> 7) & 0x01010101u) | ((Rq[1] >> 6) & 0x02020202u) |
                        ((Rq[2] >> 5) & 0x04040404u) | ((Rq[3] >> 4) & 0x08080808u);
    flw = fl;
    if (__builtin_amdgcn_ballot_w64(fl != 0u) != 0ull) {
        // elements flagged in some lane, as one scalar word: the test per
        // element below is then a scalar bit test, not vector compares
        const uint32_t W = wave_or(fl);
#pragma unroll
        for (int e = 0; e < 16; e++) {
            const int c = e & 3, sh = 8 * (e >> 2);
            if (W & (1u << (sh + c))) {
                const uint32_t byte = (Rq[c] >> sh) & 0xFFu;
                const float sum = ni[e] + nj;
                const float st = fmaf(Z.c2, acc[e], sum);
                bool amb;
                const uint32_t rk = mx_second(byte, st, sum, bnd_s, P.cE, Z.dE, amb);
                if (byte & 0x80u) {
                    Rq[c] = (Rq[c] & ~(0xFFu << sh)) | (rk << sh);
                    if (amb) fm |= 1u << e;
                }
            }
        }
    }
}

// Norms of the rows of this lane's tile elements: i0 + 8(e>>2) + 4h + (e&3).
__device__ __forceinline__ void tile_row_norms(const float *__restrict__ nrm, int64_t i0, int h, float (&ni)[16]) {
#pragma unroll
    for (int g = 0; g < 4; g++) {
        const float4 v = *(const float4 *)(nrm + i0 + 8 * g + 4 * h);
        ni[4 * g + 0] = v.x;
        ni[4 * g + 1] = v.y;
        ni[4 * g + 2] = v.z;
        ni[4 * g + 3] = v.w;
    }
}

// ---------------------------------------------------------------- k_lam_mx
// LAM pass of the MFMA form.  Work item = (band of 64 columns, row segment).
// Per 32-row block two 32x32 Gram tiles (columns j0..j0+31, then j0+32..j0+63)
// come from MFMA one after the other; element e of a lane in half h is row
// 8(e>>2) + 4h + (e&3) of column l%32 of its tile, i.e. byte e>>2 of rank
// register (e&3) + 4h of that column.  One cross-half swap per register then
// leaves lane l (half 0) with column j0 + l and lane l (half 1) with column
// j0 + l, all 32 rows, and the scan is k_lam's.  SCAN = false (split form):
// ranks and rank tiles only, the column runs left to k_lam_vscan.  STATS
// (diagnostic instantiation): stats[0..2] += elements ranked, elements the
// table flagged, elements re-decided in exact order.  G: the Gram's operand
// precision (GRAM_HILO everywhere but the precision comparison).
//
// F = double (fp64 input): the windows hold the series rounded to fp32 (then
// scaled and split as for fp32 input, the Gram's bound taken on them), the
// norms are fp64 sums rounded to fp32, and the exact re-decision sums the
// fp64 series in the reference's order against the fp64 bounds SX.  Against
// the reference's fp64 s: each rounded product adds at most 2u|x_i x_j|
// (u = 2^-24), i.e. u (n_i + n_j) to the dot product and 2u (n_i + n_j) to
// s~ after doubling; the reference's own fp64 error is below u (n_i + n_j);
// the table's and the second stage's bounds are the fp64 bounds rounded to
// fp32, off by at most 2u (n_i + n_j) near any s~ that reaches them.  The
// host adds 4u to the fp32 coefficient cE, which covers all of it.
template <typename F, int TG, bool SCAN, bool STATS = false, int G = GRAM_HILO>
__global__ __launch_bounds__(WG) RQA_WPE_MX(SCAN) void k_lam_mx(const F *__restrict__ x, const float *__restrict__ nrm,
                                               const unsigned int *__restrict__ nmax_bits, int64_t nc,
                                               int tau, int m, Bounds<float, TG> S, Bounds<F, TG> SX, int nt,
                                               const uint8_t *__restrict__ lut, MxParams P, int seg_len, LamWork W,
                                               unsigned long long *__restrict__ hist, uint2 *__restrict__ carry,
                                               uint32_t *__restrict__ priv_base, uint8_t *__restrict__ tiles,
                                               unsigned long long *__restrict__ stats) {
    __shared__ uint32_t lh[SCAN ? TG * LROW : 1];
    uint32_t st_rank = 0, st_flag = 0, st_exact = 0;
    __shared__ uint8_t lut_s[NBX_LAM];
    __shared__ float bnd_s[NBND];
    // the exact re-decision's bounds in the input type (fp64 input: the
    // table and the second stage use them rounded to fp32, S, with the
    // rounding inside the error bound cE)
    __shared__ F bndx_s[sizeof(F) == 8 ? NBND : 1];
    // rw: x[i0 .. i0 + 31 + (m-1)tau] of the current block; cw: x[j0 .. j0 + 63 + (m-1)tau] of the item,
    // both rounded to fp32, scaled and split into window words
    __shared__ MxWin<2> win[WAVES];
    if constexpr (SCAN)
        for (int e = threadIdx.x; e < TG * LROW; e += WG) lh[e] = 0;
    for (int e = threadIdx.x; e < P.nb / 4; e += WG) ((uint32_t *)lut_s)[e] = ((const uint32_t *)lut)[e];
    if (threadIdx.x < NBND) bnd_s[threadIdx.x] = mx_bound_slot((int)threadIdx.x, nt, S.s);
    if constexpr (sizeof(F) == 8)
        if (threadIdx.x < NBND) bndx_s[threadIdx.x] = (int)threadIdx.x < nt ? SX.s[threadIdx.x] : (F)__builtin_inf();
    const F *bndx;
    if constexpr (sizeof(F) == 8)
        bndx = bndx_s;
    else
        bndx = bnd_s;
    if constexpr (SCAN) priv_zero<TG>(priv_slice<TG>(priv_base), nt, WG);
    __syncthreads();
    uint32_t *priv = priv_slice<TG>(priv_base);
    const Sink K{lh, lane_spare(), priv, hist, nc};

    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = wave_id();
    const int col = lane & 31;
    const int h = lane >> 5;
    const int rspan = 32 + (m - 1) * tau;   // <= MXW (checked by the host)
    const int cspan = 64 + (m - 1) * tau;   // <= MXW + 32
    const int64_t nbands = (nc + WAVE - 1) / WAVE;
    const int64_t nitems = W.items ? W.nitems : nbands * W.nseg;
    const size_t cstride = (size_t)W.cseg * nc;
    MxWin<2> &Wn = win[wave];
    float *rw = Wn.rw, *cw = Wn.cw;
    mx_win_init(Wn, lane);
    const float *pa[8];
    gram_addrs(Wn, col, h, m, tau, pa);
    const MxScale Z = mx_scale(nmax_bits, P);

    for (int64_t item = (int64_t)blockIdx.x * WAVES + wave; item < nitems; item += (int64_t)gridDim.x * WAVES) {
        int64_t band;
        int q;
        if (W.items) {
            const int2 it = W.items[item];
            band = it.x;
            q = it.y;
        } else {
            band = item / W.nseg;
            q = (int)(item - band * W.nseg);
        }
        const int64_t vend = W.items ? min((band + 1) * WAVE, nc) : nc;
        const int64_t j0 = band * WAVE;
        const int64_t jown = j0 + lane;   // the column this lane scans (after the swap)
        const bool jv = jown < nc;
        const int64_t ib = (int64_t)q * seg_len;
        const int64_t ie = ib + seg_len < vend ? ib + seg_len : vend;
        uint2 *cslot = carry + (size_t)q * nc + (jv ? jown : 0);

        uint32_t run[TG];
#pragma unroll
        for (int t = 0; t < TG; t++) run[t] = jv ? PRE : 0u;

        // column window once per item; row window of the first block
        for (int e = lane; e < cspan; e += WAVE) cw[e] = win_word<G>((float)x[j0 + e] * Z.s);
        float pre0 = lane < rspan ? x[ib + lane] : 0.f;
        float pre1 = lane + WAVE < rspan ? x[ib + WAVE + lane] : 0.f;

        for (int64_t i0 = ib; i0 < ie; i0 += BLK) {
            const int nb = ie - i0 < BLK ? (int)(ie - i0) : BLK;
            wave_sync();
            if (lane < rspan) rw[lane] = win_word<G>(pre0 * Z.s);
            if (lane + WAVE < rspan) rw[WAVE + lane] = win_word<G>(pre1 * Z.s);
            wave_sync();
            // next block's row window, in flight while this block computes
            // (only if there is a next block: its window ends by x[N + 30])
            const bool more = i0 + BLK < ie;
            pre0 = more && lane < rspan ? x[i0 + BLK + lane] : 0.f;
            pre1 = more && lane + WAVE < rspan ? x[i0 + BLK + WAVE + lane] : 0.f;
            float ni[16];
            tile_row_norms(nrm, i0, h, ni);
            // real rows of this lane's elements (rows past the segment are
            // padding); all of them but in a segment's last block
            uint32_t rowm = 0xFFFFu;
            if (nb < BLK) {
                rowm = 0u;
#pragma unroll
                for (int e = 0; e < 16; e++) rowm |= ((8 * (e >> 2) + 4 * h + (e & 3)) < nb ? 1u : 0u) << e;
            }
            uint32_t Rr[8], fm = 0u;
            uint32_t flw0, flw1;
            {
                const floatx16 acc = gram_tile<2, 0, G>(Wn, pa, col, h, m, tau);
                uint32_t Rq[4], f;
                tile_ranks(acc, ni, nrm[j0 + col], lut_s, P, Z, bnd_s, Rq, f, flw0);
#pragma unroll
                for (int c = 0; c < 4; c++) Rr[c] = Rq[c];
                if (j0 + col < nc) fm |= f & rowm;
            }
            {
                const floatx16 acc = gram_tile<2, 1, G>(Wn, pa, col, h, m, tau);
                uint32_t Rq[4], f;
                tile_ranks(acc, ni, nrm[j0 + 32 + col], lut_s, P, Z, bnd_s, Rq, f, flw1);
#pragma unroll
                for (int c = 0; c < 4; c++) Rr[4 + c] = Rq[c];
                if (j0 + 32 + col < nc) fm |= (f & rowm) << 16;
            }
            if constexpr (STATS) {
                // flag words to element bits (element e <-> bit 8(e>>2) + (e&3))
                uint32_t e0 = 0, e1 = 0;
#pragma unroll
                for (int e = 0; e < 16; e++) {
                    e0 |= ((flw0 >> (8 * (e >> 2) + (e & 3))) & 1u) << e;
                    e1 |= ((flw1 >> (8 * (e >> 2) + (e & 3))) & 1u) << e;
                }
                const uint32_t v0 = j0 + col < nc ? rowm : 0u, v1 = j0 + 32 + col < nc ? rowm : 0u;
                st_rank += __builtin_popcount(v0) + __builtin_popcount(v1);
                st_flag += __builtin_popcount(e0 & v0) + __builtin_popcount(e1 & v1);
                st_exact += __builtin_popcount(fm);
            }
            // exact re-decision of the ambiguous elements by their own lane
            while (fm) {
                const uint32_t b = __builtin_ctz(fm);
                fm &= fm - 1u;
                const uint32_t e = b & 15u, p = b >> 4;
                const F sx = exact_sqdist<F>(x, i0 + 4 * h + 8 * (e >> 2) + (e & 3), j0 + 32 * p + col, tau, m);
                const uint32_t rk = rank_direct<F, TG>(sx, bndx, nt);
                const uint32_t c = 4 * p + (e & 3), sh = 8 * (e >> 2), keep = ~(0xFFu << sh), val = rk << sh;
#pragma unroll
                for (int cc = 0; cc < 8; cc++)
                    if (c == (uint32_t)cc) Rr[cc] = (Rr[cc] & keep) | val;
            }
            // one swap per register: half 0 keeps tile 0 rows 8g+c and gets
            // tile 0 rows 8g+4+c from its partner; half 1 the same for tile 1
            uint32_t R[8];
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const auto sw = __builtin_amdgcn_permlane32_swap(Rr[c], Rr[4 + c], false, false);
                R[c] = sw[0];
                R[4 + c] = sw[1];
            }
            if (tiles) store_tile(tiles, i0 / BLK, band, nbands, R);
            if constexpr (SCAN) {
                const uint32_t fill = nb == BLK ? 0u : (0xFFFFFFFFu << nb);
                scan_words<TG>(R, jv ? ~fill : 0u, jv ? 0u : 0xFFFFFFFFu, run, nt, K, cslot, cstride);
                if (fill && jv) {
#pragma unroll
                    for (int t = 0; t < TG; t++) run[t] -= (uint32_t)(BLK - nb);
                }
            }
        }
        if constexpr (SCAN)
            if (jv) close_segment<TG>(run, nt, cslot, cstride);
    }

    if constexpr (SCAN) {
        __syncthreads();
        flush_hist<TG>(lh, priv, nt, hist, nc, WG);
    }
    if constexpr (STATS) {
        if (st_rank) atomicAdd(&stats[0], (unsigned long long)st_rank);
        if (st_flag) atomicAdd(&stats[1], (unsigned long long)st_flag);
        if (st_exact) atomicAdd(&stats[2], (unsigned long long)st_exact);
    }
}

// ---------------------------------------------------------------- k_rp_mx
// The recurrence plot in the MFMA form (fp32, Euclidean, one threshold): the
// decisions of k_lam_mx's tiles -- hi/lo Gram, byte table, second stage and
// exact-order recheck, with the one bound's rank 0 meaning recurrent -- written
// straight to the row-major plot.  Work item = one block of 32 rows x 64
// columns, one per wave (both Gram tiles); element e of lane (col, h) of tile
// p is row 8(e>>2) + 4h + (e&3), column 32p + col, so each store writes two
// runs of 32 consecutive bytes.  Windows read up to x[N + 62] and nrm[N_c + 62]
// (both padded with zeros).
#ifndef RQA_HELPERS_ONLY
__global__ __launch_bounds__(WG) RQA_WPE_MX(false) void k_rp_mx(const float *__restrict__ x,
                                                               const float *__restrict__ nrm,
                                                               const unsigned int *__restrict__ nmax_bits, int64_t nc,
                                                               int tau, int m, float bound,
                                                               const uint8_t *__restrict__ lut, MxParams P,
                                                               char *__restrict__ out) {
    __shared__ uint8_t lut_s[NBX_LAM];
    __shared__ float bnd_s[NBND];
    __shared__ MxWin<2> win[WAVES];
    for (int e = threadIdx.x; e < P.nb / 4; e += WG) ((uint32_t *)lut_s)[e] = ((const uint32_t *)lut)[e];
    if (threadIdx.x < NBND) bnd_s[threadIdx.x] = mx_bound_slot((int)threadIdx.x, 1, &bound);
    __syncthreads();
    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = wave_id();
    const int col = lane & 31;
    const int h = lane >> 5;
    const int rspan = 32 + (m - 1) * tau;   // <= MXW (checked by the host)
    const int cspan = 64 + (m - 1) * tau;
    MxWin<2> &Wn = win[wave];
    float *rw = Wn.rw, *cw = Wn.cw;
    mx_win_init(Wn, lane);
    const float *pa[8];
    gram_addrs(Wn, col, h, m, tau, pa);
    const MxScale Z = mx_scale(nmax_bits, P);
    const int64_t nbands = (nc + WAVE - 1) / WAVE;
    const int64_t nblk = ((nc + BLK - 1) / BLK) * nbands;
    for (int64_t b = (int64_t)blockIdx.x * WAVES + wave; b < nblk; b += (int64_t)gridDim.x * WAVES) {
        const int64_t I = b / nbands, J = b - I * nbands;
        const int64_t i0 = I * BLK, j0 = J * WAVE;
        wave_sync();
        for (int e = lane; e < cspan; e += WAVE) cw[e] = hl_word(x[j0 + e] * Z.s);
        for (int e = lane; e < rspan; e += WAVE) rw[e] = hl_word(x[i0 + e] * Z.s);
        wave_sync();
        float ni[16];
        tile_row_norms(nrm, i0, h, ni);
#pragma unroll
        for (int p = 0; p < 2; p++) {
            const floatx16 acc = p == 0 ? gram_tile<2, 0>(Wn, pa, col, h, m, tau) : gram_tile<2, 1>(Wn, pa, col, h, m, tau);
            uint32_t Rq[4], fm, flw;
            tile_ranks(acc, ni, nrm[j0 + 32 * p + col], lut_s, P, Z, bnd_s, Rq, fm, flw);
            const int64_t j = j0 + 32 * p + col;
            const bool jv = j < nc;
            // exact re-decision of the ambiguous elements by their own lane
            uint32_t f = jv ? fm : 0u;
            while (f) {
                const uint32_t e = __builtin_ctz(f);
                f &= f - 1u;
                const int64_t i = i0 + 8 * (e >> 2) + 4 * h + (e & 3);
                if (i < nc) {
                    const float sx = exact_sqdist<float>(x, i, j, tau, m);
                    const uint32_t rk = rank_direct<float, 1>(sx, bnd_s, 1);
                    const uint32_t sh = 8 * (e >> 2), c = e & 3;
                    Rq[c] = (Rq[c] & ~(0xFFu << sh)) | (rk << sh);
                }
            }
            if (jv) {
#pragma unroll
                for (int e = 0; e < 16; e++) {
                    const int64_t i = i0 + 8 * (e >> 2) + 4 * h + (e & 3);
                    if (i < nc) out[i * nc + j] = ((Rq[e & 3] >> (8 * (e >> 2))) & 0xFFu) == 0u ? 1 : 0;
                }
            }
        }
    }
}
#endif

// ---------------------------------------------------------------- k_rp_mx_sym
// The symmetric recurrence plot (squares as in k_rp_sym) in the MFMA form.
// Each 32-row block of a square is decided as in k_rp_mx; with one bound
// every final rank is 0 or 1, so four elements' decisions (rank 0 -> 1) are
// their rank word XOR 0x01010101.  D gets each lane's column one byte at a
// time.  For T, byte q of rank word c is row 8q + 4h + c of the lane's
// column: a byte transpose of the four words and one cross-half swap per
// two words leave half 0 with rows 0..15 and half 1 with rows 16..31, one
// 16-byte LDS store per tile.  Windows read what k_rp_mx reads.
#ifndef RQA_HELPERS_ONLY
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(4))) void k_rp_mx_sym(
    const float *__restrict__ x, const float *__restrict__ nrm, const unsigned int *__restrict__ nmax_bits,
    int64_t nc, int tau, int m, float bound, const uint8_t *__restrict__ lut, MxParams P, char *__restrict__ out) {
    __shared__ uint8_t lut_s[NBX_RP];
    __shared__ float bnd_s[NBND];
    __shared__ MxWin<2> win[WAVES];
    __shared__ __attribute__((aligned(16))) RpStage stg[WAVES];
    for (int e = threadIdx.x; e < P.nb / 4; e += WG) ((uint32_t *)lut_s)[e] = ((const uint32_t *)lut)[e];
    if (threadIdx.x < NBND) bnd_s[threadIdx.x] = mx_bound_slot((int)threadIdx.x, 1, &bound);
    __syncthreads();
    constexpr int RS = RpStage::RS;
    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = wave_id();
    const int col = lane & 31;
    const int h = lane >> 5;
    const int rspan = 32 + (m - 1) * tau;   // <= MXW (checked by the host)
    const int cspan = 64 + (m - 1) * tau;
    MxWin<2> &Wn = win[wave];
    float *rw = Wn.rw, *cw = Wn.cw;
    mx_win_init(Wn, lane);
    const float *pa[8];
    gram_addrs(Wn, col, h, m, tau, pa);
    const MxScale Z = mx_scale(nmax_bits, P);
    RpStage &S = stg[wave];
    const int64_t nbands = (nc + WAVE - 1) / WAVE;
    const int64_t nsq = rp_items(nbands);
    for (int64_t it = rp_first_item(wave); it < nsq; it += (int64_t)gridDim.x * WAVES) {
        int64_t g, J;
        if (!rp_square(it, nbands, g, J)) continue;
        const int64_t j0 = J * WAVE, r0 = g * WAVE;
        wave_sync();
        for (int e = lane; e < cspan; e += WAVE) cw[e] = hl_word(x[j0 + e] * Z.s);
        for (int k = 0; k < 2; k++) {
            const int64_t i0 = r0 + BLK * k;
            if (i0 >= nc) break;
            for (int e = lane; e < rspan; e += WAVE) rw[e] = hl_word(x[i0 + e] * Z.s);
            wave_sync();
            float ni[16];
            tile_row_norms(nrm, i0, h, ni);
#pragma unroll
            for (int p = 0; p < 2; p++) {
                const floatx16 acc = p == 0 ? gram_tile<2, 0>(Wn, pa, col, h, m, tau) : gram_tile<2, 1>(Wn, pa, col, h, m, tau);
                uint32_t Rq[4], fm, flw;
                tile_ranks(acc, ni, nrm[j0 + 32 * p + col], lut_s, P, Z, bnd_s, Rq, fm, flw);
                const int64_t j = j0 + 32 * p + col;
                // exact re-decision of the ambiguous elements by their own lane
                uint32_t f = j < nc ? fm : 0u;
                while (f) {
                    const uint32_t e = __builtin_ctz(f);
                    f &= f - 1u;
                    const int64_t i = i0 + 8 * (e >> 2) + 4 * h + (e & 3);
                    if (i < nc) {
                        const float sx = exact_sqdist<float>(x, i, j, tau, m);
                        const uint32_t rk = rank_direct<float, 1>(sx, bnd_s, 1);
                        const uint32_t sh = 8 * (e >> 2), c = e & 3;
                        Rq[c] = (Rq[c] & ~(0xFFu << sh)) | (rk << sh);
                    }
                }
                uint32_t Dw[4];
#pragma unroll
                for (int c = 0; c < 4; c++) Dw[c] = Rq[c] ^ 0x01010101u;
                // D: element e is row 8(e>>2) + 4h + (e&3), column 32p + col
                uint8_t *dd = S.d + 4 * h * RS + 32 * p + col;
#pragma unroll
                for (int e = 0; e < 16; e++) dd[(8 * (e >> 2) + (e & 3)) * RS] = (uint8_t)(Dw[e & 3] >> (8 * (e >> 2)));
                // T: byte q of Dw[c] is row 8q + 4h + c; word q of the
                // transpose holds rows 8q + 4h .. 8q + 4h + 3
                const uint32_t lo01 = __builtin_amdgcn_perm(Dw[1], Dw[0], 0x05010400u);
                const uint32_t hi01 = __builtin_amdgcn_perm(Dw[1], Dw[0], 0x07030602u);
                const uint32_t lo23 = __builtin_amdgcn_perm(Dw[3], Dw[2], 0x05010400u);
                const uint32_t hi23 = __builtin_amdgcn_perm(Dw[3], Dw[2], 0x07030602u);
                const uint32_t w0 = __builtin_amdgcn_perm(lo23, lo01, 0x05040100u);
                const uint32_t w1 = __builtin_amdgcn_perm(lo23, lo01, 0x07060302u);
                const uint32_t w2 = __builtin_amdgcn_perm(hi23, hi01, 0x05040100u);
                const uint32_t w3 = __builtin_amdgcn_perm(hi23, hi01, 0x07060302u);
                // half 0 then holds rows 0..15 of its column, half 1 rows 16..31
                const auto s02 = __builtin_amdgcn_permlane32_swap(w0, w2, false, false);
                const auto s13 = __builtin_amdgcn_permlane32_swap(w1, w3, false, false);
                *(uint4 *)(S.t + (32 * p + col) * RS + 32 * k + 16 * h) = make_uint4(s02[0], s02[1], s13[0], s13[1]);
            }
            wave_sync();
            rp_store_block(S, lane, i0, j0, nc, out);
            wave_sync();
        }
        if (g < J) rp_store_mirror(S, lane, j0, r0, nc, out);
    }
}
#endif

// ---------------------------------------------------------------- k_lam_vscan
// The column part of LAM in the split form, read back from the rank tiles
// (the MFMA column kernel then only ranks and stores).  Work item = (band J,
// segment q of the rows [0, V_J)), as that kernel's; lane l walks column
// 64J + l in blocks of 32 rows, each block's ranks being the lane's own
// 32-byte record of tile (I, J), in the packed layout the scan takes: no
// exchange between lanes.  The next block's record is in flight while a
// block is scanned.
template <int TG>
__global__ __launch_bounds__(WG) RQA_WPE_SCAN(TG, 7, 6) void k_lam_vscan(const uint8_t *__restrict__ tiles, int64_t nbands, int64_t nc,
                                                  int nt, int seg_len, LamWork W,
                                                  unsigned long long *__restrict__ hist, uint2 *__restrict__ carry,
                                                  uint32_t *__restrict__ priv_base) {
    __shared__ uint32_t lh[TG * lrow<WAVE>()];
    for (int e = threadIdx.x; e < TG * lrow<WAVE>(); e += WG) lh[e] = 0;
    priv_zero<TG>(priv_slice<TG>(priv_base), nt, WG);
    __syncthreads();
    uint32_t *priv = priv_slice<TG>(priv_base);
    const Sink K{lh, lane_spare<WAVE>(), priv, hist, nc};

    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = wave_id();
    const size_t cstride = (size_t)W.cseg * nc;

    for (int64_t item = (int64_t)blockIdx.x * WAVES + wave; item < W.nitems; item += (int64_t)gridDim.x * WAVES) {
        const int2 it = W.items[item];
        const int64_t band = it.x;
        const int q = it.y;
        const int64_t vend = min((band + 1) * WAVE, nc);
        const int64_t jown = band * WAVE + lane;
        const bool jv = jown < nc;
        const int64_t ib = (int64_t)q * seg_len;
        const int64_t ie = ib + seg_len < vend ? ib + seg_len : vend;
        uint2 *cslot = carry + (size_t)q * nc + (jv ? jown : 0);

        uint32_t run[TG];
#pragma unroll
        for (int t = 0; t < TG; t++) run[t] = jv ? PRE : 0u;

        uint4 v0 = make_uint4(0u, 0u, 0u, 0u), v1 = v0;
        auto fetch = [&](int64_t i) {
            const uint4 *src = (const uint4 *)(tiles + tile_index(i / BLK, band, nbands) * TILE_BYTES) + 2 * lane;
            v0 = src[0];
            v1 = src[1];
        };
        if (ib < ie) fetch(ib);
        for (int64_t i0 = ib; i0 < ie; i0 += BLK) {
            const int nb = ie - i0 < BLK ? (int)(ie - i0) : BLK;
            const uint32_t R[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
            if (i0 + BLK < ie) fetch(i0 + BLK);
            // rows past the part count as recurrent padding, taken back below
            const uint32_t fill = nb == BLK ? 0u : (0xFFFFFFFFu << nb);
            scan_words<TG, WAVE>(R, jv ? ~fill : 0u, jv ? 0u : 0xFFFFFFFFu, run, nt, K, cslot, cstride);
            if (fill && jv) {
#pragma unroll
                for (int t = 0; t < TG; t++) run[t] -= (uint32_t)(BLK - nb);
            }
        }
        if (jv) close_segment<TG>(run, nt, cslot, cstride);
    }

    __syncthreads();
    flush_hist<TG, WAVE>(lh, priv, nt, hist, nc, WG);
}

// ---------------------------------------------------------------- k_det_mx
// DET pass of the MFMA form.  Work item = (strip of 64 diagonals, step
// segment), as k_det.  A 32-step block needs G(s0+u, s0+r0+v) for v = u + l;
// three 32x32 MFMA tiles cover v = 0..95, one after the other.  Each needed
// element's rank byte is stored at (diagonal l, step u) of an LDS byte matrix
// laid out so that lane l then reads its diagonal's eight packed rank
// registers as words: byte u of row l sits at l*SKB + 4(u&7) + (u>>3).
constexpr int SKB = 36;   // bytes per diagonal row (9 words: conflict-free stores and loads)

template <int TG>
__global__ __launch_bounds__(WG) RQA_WPE_MX(true) void k_det_mx(const float *__restrict__ x, const float *__restrict__ nrm,
                                               const unsigned int *__restrict__ nmax_bits, int64_t nc,
                                               int tau, int m, Bounds<float, TG> S, int nt,
                                               const uint8_t *__restrict__ lut, MxParams P, int seg_len, int nseg,
                                               const int2 *__restrict__ items, int64_t nitems,
                                               unsigned long long *__restrict__ hist, uint2 *__restrict__ carry,
                                               uint32_t *__restrict__ priv_base) {
    __shared__ uint32_t lh[TG * LROW];
    __shared__ uint8_t lut_s[NBX];
    __shared__ uint32_t skew[WAVES][WAVE * SKB / 4];
    __shared__ float bnd_s[NBND];
    // rw: x[s0 .. s0 + 31 + (m-1)tau]; cw: x[c0 .. c0 + 95 + (m-1)tau]
    __shared__ MxWin<3> win[WAVES];
    for (int e = threadIdx.x; e < TG * LROW; e += WG) lh[e] = 0;
    for (int e = threadIdx.x; e < P.nb / 4; e += WG) ((uint32_t *)lut_s)[e] = ((const uint32_t *)lut)[e];
    if (threadIdx.x < NBND) bnd_s[threadIdx.x] = mx_bound_slot((int)threadIdx.x, nt, S.s);
    priv_zero<TG>(priv_slice<TG>(priv_base), nt, WG);
    __syncthreads();
    uint32_t *priv = priv_slice<TG>(priv_base);
    const Sink K{lh, lane_spare(), priv, hist, nc};

    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = wave_id();
    const int col = lane & 31;
    const int h = lane >> 5;
    uint8_t *sk8 = (uint8_t *)skew[wave];
    MxWin<3> &Wn = win[wave];
    float *rw = Wn.rw, *cw = Wn.cw;
    mx_win_init(Wn, lane);
    const float *pa[8];
    gram_addrs(Wn, col, h, m, tau, pa);
    const MxScale Z = mx_scale(nmax_bits, P);
    const size_t cstride = (size_t)nseg * (nc - 1);
    const int rspan = 32 + (m - 1) * tau, cspan = 96 + (m - 1) * tau;   // <= MXW, MXW + 64

    for (int64_t item = (int64_t)blockIdx.x * WAVES + wave; item < nitems; item += (int64_t)gridDim.x * WAVES) {
        const int2 it = items[item];
        const int64_t r0 = 1 + (int64_t)it.x * WAVE;
        const int q = it.y;
        const int64_t r = r0 + lane;
        const bool rv = r < nc;
        const int64_t dlen = rv ? nc - r : 0;
        const int64_t sb = (int64_t)q * seg_len;
        const int64_t se_nom = sb + seg_len;
        const int64_t smax = nc - r0;
        const int64_t se = se_nom < smax ? se_nom : smax;
        uint2 *cslot = carry + (size_t)q * (nc - 1) + (rv ? r - 1 : 0);

        uint32_t run[TG];
#pragma unroll
        for (int t = 0; t < TG; t++) run[t] = rv ? PRE : 0u;

        // operand windows of the first block
        float pr[2], pc[3];
        pr[0] = lane < rspan ? x[sb + lane] : 0.f;
        pr[1] = lane + WAVE < rspan ? x[sb + WAVE + lane] : 0.f;
#pragma unroll
        for (int c = 0; c < 3; c++) pc[c] = lane + c * WAVE < cspan ? x[sb + r0 + c * WAVE + lane] : 0.f;

        for (int64_t s0 = sb; s0 < se; s0 += BLK) {
            const int64_t c0 = s0 + r0;   // first column of tile 0
            wave_sync();
#pragma unroll
            for (int c = 0; c < 2; c++)
                if (lane + c * WAVE < rspan) rw[lane + c * WAVE] = hl_word(pr[c] * Z.s);
#pragma unroll
            for (int c = 0; c < 3; c++)
                if (lane + c * WAVE < cspan) cw[lane + c * WAVE] = hl_word(pc[c] * Z.s);
            wave_sync();
            // next block's windows, in flight while this block computes (only
            // if there is a next block: its windows end by x[N + 94])
            const bool more = s0 + BLK < se;
#pragma unroll
            for (int c = 0; c < 2; c++)
                pr[c] = more && lane + c * WAVE < rspan ? x[s0 + BLK + c * WAVE + lane] : 0.f;
#pragma unroll
            for (int c = 0; c < 3; c++)
                pc[c] = more && lane + c * WAVE < cspan ? x[c0 + BLK + c * WAVE + lane] : 0.f;
            float ni[16];
            tile_row_norms(nrm, s0, h, ni);
            // element e of tile p: row u = 8(e>>2) + 4h + (e&3), column v = 32p + col,
            // diagonal l = v - u = 32p + colh - u0 with u0 = u - 4h, colh = col - 4h.
            // Its byte goes to l*SKB + 4(u&7) + (u>>3) = skb + SKB*(32p - u0) + 4(u0&7) + (u0>>3),
            // skb = SKB*colh + 16h (u0&7 <= 3, so adding 4h changes only the low three bits of u).
            // Tile 1 is needed whole; tile 0 where colh >= u0, tile 2 where colh < u0.
            const int colh = col - 4 * h;
            const int skb = SKB * colh + 16 * h;   // may be negative; every stored offset is not
            uint64_t fma = 0;
#pragma unroll
            for (int p = 0; p < 3; p++) {
                const floatx16 acc = p == 0   ? gram_tile<3, 0>(Wn, pa, col, h, m, tau)
                                     : p == 1 ? gram_tile<3, 1>(Wn, pa, col, h, m, tau)
                                              : gram_tile<3, 2>(Wn, pa, col, h, m, tau);
                uint32_t Rq[4], fm, flw;
                tile_ranks(acc, ni, nrm[c0 + 32 * p + col], lut_s, P, Z, bnd_s, Rq, fm, flw);
                uint32_t need = 0;
#pragma unroll
                for (int e = 0; e < 16; e++) {
                    const int u0 = 8 * (e >> 2) + (e & 3);
                    const bool nd = p == 1 ? true : (p == 0 ? colh >= u0 : colh < u0);
                    if (nd) {
                        need |= 1u << e;
                        sk8[skb + SKB * (32 * p - u0) + 4 * (u0 & 7) + (u0 >> 3)] = (uint8_t)(Rq[e & 3] >> (8 * (e >> 2)));
                    }
                }
                // pairs past the end of the matrix are not re-decided (the scan masks them)
                fma |= (uint64_t)(c0 + 32 * p + col < nc ? (fm & need) : 0u) << (16 * p);
            }
            // exact re-decision of the ambiguous elements by their own lane
            while (fma) {
                const uint32_t b = (uint32_t)__builtin_ctzll(fma);
                fma &= fma - 1u;
                const uint32_t e = b & 15u, p = b >> 4;
                const uint32_t u = 8 * (e >> 2) + 4 * h + (e & 3);
                const float sx = exact_sqdist<float>(x, s0 + u, c0 + 32 * p + col, tau, m);
                const uint32_t l = 32 * p + col - u;
                sk8[l * SKB + 4 * (u & 7) + (u >> 3)] = (uint8_t)rank_direct<float, TG>(sx, bnd_s, nt);
            }
            wave_sync();
            uint32_t R[8];
#pragma unroll
            for (int c = 0; c < 8; c++) R[c] = skew[wave][lane * (SKB / 4) + c];
            const int64_t nv64 = dlen - s0;
            const int nvalid = nv64 <= 0 ? 0 : (nv64 >= BLK ? BLK : (int)nv64);
            const uint32_t vmask = nvalid >= BLK ? 0xFFFFFFFFu : ((1u << nvalid) - 1u);
            scan_words<TG>(R, 0xFFFFFFFFu, ~vmask, run, nt, K, cslot, cstride);
        }
        if (rv) close_segment<TG>(run, nt, cslot, cstride);
    }

    __syncthreads();
    flush_hist<TG>(lh, priv, nt, hist, nc, WG);
}

// ---------------------------------------------------------------- k_stitch
// Joins the (prefix, suffix) runs of each sequence's segments in order.
// Block (x, t): threshold t, sequences x*STT + tid, + gridDim.x*STT, ...
// Joined runs shorter than SCAP are counted in an LDS histogram per block.
// A sequence has cseg carry slots, slot q of sequence sq for threshold t at
// carry[(t*cseg + q)*nseq + sq] (the sequences of one slot side by side, so
// a wave's 64 lanes write, and the stitch reads, 512 contiguous bytes).  DET
// sequences are diagonals r = sq + 1 of length nc - r; LAM sequences are full
// columns, either one part (slots 0 ..) or, split (split != 0), rows [0, V)
// in slots 0 .. and rows [V, nc) in slots nseg .., V = min(64 (sq/64 + 1),
// nc).  Part segment q covers [q*seg_len, min((q+1)*seg_len, part length)).
constexpr int STT = 256;
constexpr int SCAP = 4096;

#ifndef RQA_HELPERS_ONLY
__global__ __launch_bounds__(STT) void k_stitch(const uint2 *__restrict__ carry, int nt, int64_t nseq, int64_t nc,
                                                int is_det, int split, int seg_len, int nseg, int cseg,
                                                unsigned long long *__restrict__ hist) {
    __shared__ uint32_t sh[SCAP];
    const int t = blockIdx.y;   // threshold: the grid has nt rows
    if (t >= nt) return;
    for (int e = threadIdx.x; e < SCAP; e += STT) sh[e] = 0u;
    __syncthreads();
    unsigned long long *h = hist + (size_t)t * (nc + 1);
    auto count = [&](uint64_t len) {
        if (len < (uint64_t)SCAP)
            lds_inc(&sh[len]);
        else if (len <= (uint64_t)nc)
            glob_add(&h[len], 1ull);
    };
    for (int64_t sq = (int64_t)blockIdx.x * STT + threadIdx.x; sq < nseq; sq += (int64_t)gridDim.x * STT) {
        const uint2 *c = carry + (size_t)t * cseg * nseq + sq;
        uint64_t cur = 0;
        auto part = [&](int64_t plen, const uint2 *pc0) {
            const int64_t np = (plen + seg_len - 1) / seg_len;
            for (int64_t q = 0; q < np; q++) {
                const uint2 pc = pc0[(size_t)q * nseq];
                const int64_t L = (q + 1) * seg_len <= plen ? seg_len : plen - q * seg_len;
                if ((int64_t)pc.x == L) {   // whole segment recurrent
                    cur += (uint64_t)L;
                    continue;
                }
                const uint64_t tot = cur + pc.x;
                if (tot) count(tot);
                cur = pc.y;
            }
        };
        if (is_det) {
            part(nc - (sq + 1), c);
        } else if (split) {
            const int64_t V = min((sq / WAVE + 1) * WAVE, nc);
            part(V, c);
            part(nc - V, c + (size_t)nseg * nseq);
        } else {
            part(nc, c);
        }
        if (cur) count(cur);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < SCAP; e += STT)
        if (sh[e] && e <= nc) glob_add(&h[e], sh[e]);
}
#endif

// ---------------------------------------------------------------- k_finalize
// One workgroup per threshold.  Turns a line histogram into the reference's
// five measures for every l_min (AccRQA_functions.cpp:48-133 and
// rqa_process_length_histogram, AccRQA_CPU_function.cpp:181-204).
// For DET it first applies correctDETHistogram (AccRQA_CPU_function.cpp:156-162).
// metric/scan are workspaces of nt*(nc+1) u64 each (suffix sums, written up
// to the highest non-zero bin below the last).
// Output per (t, l): 5 values in F, laid out by out_index(t, l).
constexpr int FIN = 1024;

// Before k_finalize, grid (blocks, threshold): for DET, correctDETHistogram
// (AccRQA_CPU_function.cpp:156-162: every bin doubled, then the
// line-of-identity bin nc incremented); for both, the highest non-zero bin
// below nc of each threshold, as tops[t] = bin + 1 (0 if none; tops zeroed
// before the launch).
constexpr int HPT = 256;
#ifndef RQA_HELPERS_ONLY
__global__ __launch_bounds__(HPT) void k_hist_prep(unsigned long long *__restrict__ hist, int64_t nc, int is_det,
                                                   unsigned long long *__restrict__ tops) {
    const int t = blockIdx.y;
    const int64_t hs = nc + 1;
    unsigned long long *h = hist + (size_t)t * hs;
    unsigned long long best = 0;
    for (int64_t f = (int64_t)blockIdx.x * HPT + threadIdx.x; f < hs; f += (int64_t)gridDim.x * HPT) {
        unsigned long long v = h[f];
        if (is_det) {
            v = (unsigned long long)(2.0 * (double)v) + (f == hs - 1 ? 1ull : 0ull);
            h[f] = v;
        }
        if (v && f < hs - 1) best = (unsigned long long)f + 1;   // f grows, so this is the thread's highest
    }
    __shared__ unsigned long long red[HPT];
    red[threadIdx.x] = best;
    __syncthreads();
    for (int off = HPT / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0 && red[0]) atomicMax(&tops[t], red[0]);
}
#endif

template <typename F>
__global__ __launch_bounds__(FIN) void k_finalize(unsigned long long *__restrict__ hist, int64_t nc, int nt,
                                                  const unsigned long long *__restrict__ tops,
                                                  const int *__restrict__ lmins, int nl,
                                                  unsigned long long *__restrict__ metric,
                                                  unsigned long long *__restrict__ scan, F *__restrict__ out,
                                                  int det_layout, const int *__restrict__ perm, int T,
                                                  unsigned long long *__restrict__ m0_out) {
    const int t = blockIdx.x;
    if (t >= nt) return;
    const int64_t hs = nc + 1;
    unsigned long long *h = hist + (size_t)t * hs;
    unsigned long long *M = metric + (size_t)t * hs;
    unsigned long long *Sc = scan + (size_t)t * hs;
    const int tid = threadIdx.x;
    const bool is_det = det_layout != 0;

    // k_hist_prep has applied the DET correction and found top = the highest
    // non-zero bin below hs - 1 (-1 if none): the histogram is zero above its
    // longest line except the line-of-identity bin hs - 1.  Suffix sums are
    // formed over [0, top] only, and read above it through SC / MM, where only
    // bin hs - 1 contributes.
    const int64_t top = (int64_t)tops[t] - 1;
    const unsigned long long hl = h[hs - 1];
    const unsigned long long ml_top = (unsigned long long)(hs - 1) * hl;   // M above top

    // suffix sums M[f] = sum_{g >= f} g h[g], Sc[f] = sum_{g >= f} h[g] for
    // f <= top: tiles of FIN*FPT bins from the top down, FPT consecutive bins
    // per thread, an inclusive suffix scan of the threads' totals inside each
    // tile, the tiles above carried in (cm, cs)
    constexpr int FPT = 8;
    constexpr int64_t TILE = (int64_t)FIN * FPT;
    __shared__ unsigned long long bm[FIN], bs[FIN];
    unsigned long long cm = ml_top, cs = hl;
    for (int64_t b0 = top >= 0 ? (top / TILE) * TILE : -1; b0 >= 0; b0 -= TILE) {
        const int64_t f0 = b0 + (int64_t)tid * FPT;
        unsigned long long v[FPT], lm = 0, ls = 0;
#pragma unroll
        for (int k = 0; k < FPT; k++) {
            v[k] = f0 + k <= top ? h[f0 + k] : 0ull;
            lm += (unsigned long long)(f0 + k) * v[k];
            ls += v[k];
        }
        bm[tid] = lm;
        bs[tid] = ls;
        __syncthreads();
        for (int off = 1; off < FIN; off <<= 1) {
            unsigned long long am = 0, as = 0;
            if (tid + off < FIN) {
                am = bm[tid + off];
                as = bs[tid + off];
            }
            __syncthreads();
            bm[tid] += am;
            bs[tid] += as;
            __syncthreads();
        }
        unsigned long long rm = (tid + 1 < FIN ? bm[tid + 1] : 0ull) + cm;   // bins above my chunk
        unsigned long long rs = (tid + 1 < FIN ? bs[tid + 1] : 0ull) + cs;
#pragma unroll
        for (int k = FPT - 1; k >= 0; k--) {
            rm += (unsigned long long)(f0 + k) * v[k];
            rs += v[k];
            if (f0 + k <= top) {
                M[f0 + k] = rm;
                Sc[f0 + k] = rs;
            }
        }
        const unsigned long long tm = bm[0], ts = bs[0];
        __syncthreads();
        cm += tm;
        cs += ts;
    }
    __syncthreads();
    __threadfence_block();
    auto SC = [&](int64_t f) -> unsigned long long { return f <= top ? Sc[f] : (f <= hs - 1 ? hl : 0ull); };
    auto MM = [&](int64_t f) -> unsigned long long { return f <= top ? M[f] : (f <= hs - 1 ? ml_top : 0ull); };

    // Lmax / TTmax: the reference's bisection (tertiary_metric), one thread.
    __shared__ unsigned long long lmax_s;
    if (tid == 0) {
        unsigned long long last, start, mid, end, length, mv, mv1;
        bool found = false;
        end = hs - 1;
        mid = hs / 2;
        last = is_det ? SC((int64_t)end) : 0ull;
        start = 0;
        length = hs;
        while (!found) {
            mv = SC((int64_t)mid);
            mv1 = mid > 0 ? SC((int64_t)mid - 1) : 0ull;
            if (mv == last && mv1 > mv) {
                mid = mid - 1;
                found = true;
            } else if (length == 0) {
                break;
            } else if (mid == 0) {
                break;
            } else if (mv > last) {
                start = mid;
                length = end - start;
                mid = mid + (length + 1) / 2;
            } else if (mv == last && mv1 == last) {
                end = mid;
                length = end - start;
                mid = start + (length >> 1);
            } else {
                break;   // unreachable for a non-increasing scan
            }
        }
        lmax_s = mid;
        if (m0_out) m0_out[t] = MM(0);   // recurrence count (sum of l * h[l])
    }
    __syncthreads();

    __shared__ double red[FIN];
    const int ot = perm ? perm[t] : t;   // position of this threshold in the caller's order
    for (int l = 0; l < nl; l++) {
        const int64_t lm = lmins[l];
        const unsigned long long sl = SC(lm);
        double e = 0.0;
        for (int64_t f = lm + tid; f <= top; f += FIN) {
            const double p = (double)h[f] / (double)sl;
            if (p > 0) e += p * log(p);
        }
        if (tid == 0 && hs - 1 > top && hs - 1 >= lm) {   // bin hs - 1
            const double p = (double)hl / (double)sl;
            if (p > 0) e += p * log(p);
        }
        red[tid] = e;
        __syncthreads();
        for (int off = FIN / 2; off > 0; off >>= 1) {
            if (tid < off) red[tid] += red[tid + off];
            __syncthreads();
        }
        if (tid == 0) {
            const unsigned long long m0 = MM(0), ml = MM(lm);
            const double v0 = m0 > 0 ? (double)ml / (double)m0 : __builtin_nan("");
            const double v1 = sl > 0 ? (double)ml / (double)sl : __builtin_nan("");
            const double v2 = (double)lmax_s;
            const double v3 = (-1.0) * red[0];
            const double v4 = (double)m0 / (double)(nc * nc);
            // DET: [T][nL][5] (l fastest); LAM: [nL][T][5] (threshold fastest)
            const size_t pos = det_layout ? ((size_t)ot * nl + l) : ((size_t)l * T + ot);
            F *o = out + 5 * pos;
            o[0] = (F)v0;
            o[1] = (F)v1;
            o[2] = (F)v2;
            o[3] = (F)v3;
            o[4] = (F)v4;
        }
        __syncthreads();
    }
}

// RR output: (F)count / (F)(nc*nc), the PLT_CPU formula.
template <typename F>
__global__ void k_rr_out(const unsigned long long *__restrict__ counts, int nt, int64_t nc,
                         const int *__restrict__ perm, F *__restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const int64_t tmp = nc * nc;
    out[perm ? perm[t] : t] = ((F)(int64_t)counts[t]) / ((F)tmp);
}

}  // namespace rqa
