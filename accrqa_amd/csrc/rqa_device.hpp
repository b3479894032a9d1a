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
// Each lane keeps one 32-bit word per threshold: bit (31-u) is the decision
// of step u of the current 32-step block.  Runs inside a word are found with
// count-leading/trailing-zero bit tricks; the open run is carried to the next
// word in a register.  A sequence is cut into segments processed
// independently; each segment reports (prefix run, suffix run) and
// k_stitch joins them in order.  Histogram bins below LCAP are counted in a
// workgroup-private LDS histogram and merged with one global atomic per
// non-zero bin; longer runs go straight to global u64 atomics.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace rqa {

constexpr int WAVE = 64;
constexpr int BLK = 32;            // steps per decision word
constexpr int WG = 256;            // threads per workgroup (4 waves)
constexpr int WAVES = WG / WAVE;
constexpr int LCAP = 128;          // run lengths < LCAP use the LDS histogram
constexpr int DET_WIN = 1024;      // LDS samples per wave for the k_det lane window

template <typename F, int TG>
struct Bounds {
    F s[TG];
};

struct Counts {
    uint32_t c[2];
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

// Histogram row t starts at gh + t*(hmax+1); a run longer than the sequence
// would be an engine bug, and is dropped rather than written out of bounds
// (the parity tests would then fail on the missing count).
__device__ __forceinline__ void emit(uint32_t len, uint32_t *lh, unsigned long long *gh, uint32_t t, uint32_t hmax) {
    if (len < (uint32_t)LCAP)
        lds_inc(&lh[len]);
    else if (len <= hmax)
        glob_add(&gh[(size_t)opaque(t) * (hmax + 1) + len], 1ull);
}

// ---------------------------------------------------------------- threshold ranks
// All thresholds of a group are decided through one integer per pair, its
// rank r = #{t : B_t < s} over the group's bounds sorted ascending: then
// R_t = [s <= B_t] = [t >= r].  The rank comes from a bucket table indexed by
// the high bits of s (exact: a bucket holds at most one bound, whose key is
// stored with the bucket; a bucket with several bounds sends the wave to a
// direct count).  Keys are the IEEE bit patterns, which order non-negative
// values exactly; NaN maps to +inf, i.e. rank T, never recurrent.
//
// Ranks of 32 steps are packed one per byte (register q holds steps q, q+8,
// q+16, q+24), and one subtraction per register yields four decisions:
// byte 0x80 + t - r has its top bit set iff r <= t.  Shifting register q's
// top bits down by 7-q puts step u at bit u of the decision word.
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
    int b = (int)(khi(key) - P.base) >> P.sh;
    b = b < 0 ? 0 : (b > P.nb - 1 ? P.nb - 1 : b);
    const LutEntry<F> e = lut[b];
    flags |= e.info;
    return (e.info & 0xFFu) + (key > e.next ? 1u : 0u);
}

// Direct count against the sorted bounds (buckets holding several bounds).
template <typename F, int TG>
__device__ __forceinline__ uint32_t rank_direct(F s, const Bounds<F, TG> &S, int nt) {
    const auto key = skey(s);
    uint32_t r = 0;
#pragma unroll
    for (int t = 0; t < TG; t++)
        if (t < nt) r += key > skey(S.s[t]) ? 1u : 0u;
    return r;
}

template <typename F, int TG>
__device__ __forceinline__ void pack_ranks(const F (&s)[BLK], uint32_t (&R)[8], const LutEntry<F> *lut,
                                           const LutParams &P, const Bounds<F, TG> &S, int nt) {
    uint32_t flags = 0;
#pragma unroll
    for (int q = 0; q < 8; q++) R[q] = 0;
#pragma unroll
    for (int u = 0; u < BLK; u++) R[u & 7] |= rank_of(s[u], lut, P, flags) << (8 * (u >> 3));
    if (__any((flags & 0x100u) != 0)) {
#pragma unroll
        for (int q = 0; q < 8; q++) R[q] = 0;
#pragma unroll
        for (int u = 0; u < BLK; u++) R[u & 7] |= rank_direct(s[u], S, nt) << (8 * (u >> 3));
    }
}

// Decision word of threshold t: bit u = [rank of step u <= t].
__device__ __forceinline__ uint32_t swar_word(const uint32_t (&R)[8], uint32_t t) {
    const uint32_t K = 0x80808080u + t * 0x01010101u;
    uint32_t w = 0;
#pragma unroll
    for (int q = 0; q < 8; q++) w |= ((K - R[q]) >> (7 - q)) & (0x01010101u << q);
    return w;
}

template <typename F>
__device__ __forceinline__ void load_lut(LutEntry<F> *dst, const LutEntry<F> *src, int nb) {
    for (int e = threadIdx.x; e < nb; e += WG) dst[e] = src[e];
}

// Consume one 32-step word, bit u = step u (bit 0 oldest).  run: open run
// entering the word.  Bit tb of popen is set until the segment's first zero;
// the run it closes is the segment's prefix run, written to *pre (the stitch
// joins it with the previous segment) instead of being counted here.
__device__ __forceinline__ void consume_word(uint32_t w, uint32_t &run, uint32_t &popen, uint32_t t, uint2 *cslot,
                                             size_t cstride, uint32_t *lh, unsigned long long *gh, uint32_t hmax) {
    if (w == 0xFFFFFFFFu) {
        run += BLK;
        return;
    }
    const uint32_t tb = 1u << t;
    const uint32_t lead = __builtin_ctz(~w);     // ones from the oldest step up to the first zero
    const uint32_t first = run + lead;
    if (popen & tb) {
        cslot[(size_t)opaque(t) * cstride].x = first;
        popen &= ~tb;
    } else if (first) {
        emit(first, lh, gh, t, hmax);
    }
    const uint32_t tail = __clz(~w);             // ones from the newest step down to the last zero
    uint32_t mid = w & (0xFFFFFFFFu << lead) & (0xFFFFFFFFu >> tail);
    while (mid) {
        const uint32_t lo = __builtin_ctz(mid);
        const uint32_t len = __builtin_ctz(~(mid >> lo));
        emit(len, lh, gh, t, hmax);
        mid &= ~(((1u << len) - 1u) << lo);
    }
    run = tail;
}

// Segment end: (prefix, suffix) of the segment for the stitch.  A segment
// whose prefix never closed is all ones: prefix = suffix = its length.
__device__ __forceinline__ void close_segment(uint2 *slot, uint32_t run, uint32_t popen, uint32_t tb) {
    if (popen & tb)
        *slot = make_uint2(run, run);
    else
        slot->y = run;
}

__device__ __forceinline__ int wave_id() {
    return __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
}

// ---------------------------------------------------------------- k_lam
// Work item = (band of 64 columns, row segment q).  x is padded with >= 128
// zeros past N so that the unguarded window reads stay inside the buffer.
// The row window x[i0 + u + k*tau] (u < 32) is uniform across the wave and
// comes through scalar loads; x[j + k*tau] is one vector load per k and block.
template <typename F, int NORM, int TG>
__global__ __launch_bounds__(WG) void k_lam(const F *__restrict__ x, int64_t nc, int tau, int m, Bounds<F, TG> S,
                                            int nt, const LutEntry<F> *__restrict__ lut, LutParams P, int seg_len,
                                            int nseg, unsigned long long *__restrict__ hist,
                                            uint2 *__restrict__ carry) {
    __shared__ uint32_t lh[TG * LCAP];
    __shared__ LutEntry<F> lut_s[NB_MAX];
    for (int e = threadIdx.x; e < TG * LCAP; e += WG) lh[e] = 0;
    load_lut(lut_s, lut, P.nb);
    __syncthreads();

    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = wave_id();
    const int64_t nbands = (nc + WAVE - 1) / WAVE;
    const int64_t nitems = nbands * nseg;
    const size_t cstride = (size_t)nc * nseg;   // carry stride between thresholds

    for (int64_t item = (int64_t)blockIdx.x * WAVES + wave; item < nitems; item += (int64_t)gridDim.x * WAVES) {
        const int64_t band = item / nseg;
        const int q = (int)(item - band * nseg);
        const int64_t j = band * WAVE + lane;
        const bool jv = j < nc;
        const int64_t ib = (int64_t)q * seg_len;
        const int64_t ie = ib + seg_len < nc ? ib + seg_len : nc;
        uint2 *cslot = carry + (size_t)(jv ? j : 0) * nseg + q;

        uint32_t run[TG];
#pragma unroll
        for (int t = 0; t < TG; t++) run[t] = 0;
        uint32_t popen = jv ? 0xFFFFFFFFu : 0u;

        for (int64_t i0 = ib; i0 < ie; i0 += BLK) {
            const int nb = ie - i0 < BLK ? (int)(ie - i0) : BLK;
            F s[BLK];
#pragma unroll
            for (int u = 0; u < BLK; u++) s[u] = F(0);
            F xj = x[j];
            for (int k = 0; k < m; k++) {
                // next dimension's column sample, loaded while this one computes
                const F xjn = x[j + (int64_t)(k + 1 < m ? k + 1 : k) * tau];
                const F *xr = x + i0 + (int64_t)k * tau;
#pragma unroll
                for (int u = 0; u < BLK; u++) s[u] = dim_step<NORM>(s[u], xr[u], xj);
                xj = xjn;
            }
            uint32_t R[8];
            pack_ranks(s, R, lut_s, P, S, nt);
            const uint32_t fill = nb == BLK ? 0u : (0xFFFFFFFFu << nb);   // rows past the segment
#pragma unroll
            for (int t = 0; t < TG; t++) {
                if (t < nt) {
                    uint32_t w = swar_word(R, t);
                    w = jv ? (w | fill) : 0u;
                    consume_word(w, run[t], popen, t, cslot, cstride, lh + t * LCAP, hist, (uint32_t)nc);
                    if (fill && jv) run[t] -= (uint32_t)(BLK - nb);
                }
            }
        }
        if (jv) {
#pragma unroll
            for (int t = 0; t < TG; t++)
                if (t < nt) close_segment(&cslot[(size_t)opaque(t) * cstride], run[t], popen, 1u << t);
        }
    }

    __syncthreads();
    for (int e = threadIdx.x; e < TG * LCAP; e += WG) {
        const uint32_t v = lh[e];
        const int t = e / LCAP;
        if (v && t < nt) glob_add(&hist[(size_t)t * (nc + 1) + (e - t * LCAP)], v);
    }
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
                                            unsigned long long *__restrict__ hist, uint2 *__restrict__ carry) {
    __shared__ uint32_t lh[TG * LCAP];
    __shared__ LutEntry<F> lut_s[NB_MAX];
    __shared__ F win[WAVES][DET_WIN];
    for (int e = threadIdx.x; e < TG * LCAP; e += WG) lh[e] = 0;
    load_lut(lut_s, lut, P.nb);
    __syncthreads();

    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = wave_id();
    F *wn = win[wave];
    // dimensions staged per window: span = BLK + WAVE - 1 + (kc-1)*tau <= DET_WIN
    const int kc_max = tau > 0 ? (DET_WIN - (BLK + WAVE - 1)) / tau + 1 : m;
    const int kc = kc_max < m ? kc_max : m;
    const size_t cstride = (size_t)(nc - 1) * nseg;

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
        uint2 *cslot = carry + (size_t)(rv ? r - 1 : 0) * nseg + q;

        uint32_t run[TG];
#pragma unroll
        for (int t = 0; t < TG; t++) run[t] = 0;
        uint32_t popen = rv ? 0xFFFFFFFFu : 0u;

        for (int64_t s0 = sb; s0 < se; s0 += BLK) {
            F s[BLK];
#pragma unroll
            for (int u = 0; u < BLK; u++) s[u] = F(0);
            for (int k0 = 0; k0 < m; k0 += kc) {
                const int k1 = k0 + kc < m ? k0 + kc : m;
                const int64_t base = s0 + r0 + (int64_t)k0 * tau;
                const int span = BLK + WAVE - 1 + (k1 - 1 - k0) * tau;
                for (int e = lane; e < span; e += WAVE) wn[e] = x[base + e];
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                for (int k = k0; k < k1; k++) {
                    const F *xr = x + s0 + (int64_t)k * tau;
                    const F *wl = wn + lane + (k - k0) * tau;
#pragma unroll
                    for (int u = 0; u < BLK; u++) s[u] = dim_step<NORM>(s[u], xr[u], wl[u]);
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
            uint32_t R[8];
            pack_ranks(s, R, lut_s, P, S, nt);
            // bit u valid iff s0+u < dlen; invalid steps decide 0
            const int64_t nv64 = dlen - s0;
            const int nvalid = nv64 <= 0 ? 0 : (nv64 >= BLK ? BLK : (int)nv64);
            const uint32_t vmask = nvalid >= BLK ? 0xFFFFFFFFu : ((1u << nvalid) - 1u);
#pragma unroll
            for (int t = 0; t < TG; t++) {
                if (t < nt) {
                    const uint32_t w = swar_word(R, t) & vmask;
                    consume_word(w, run[t], popen, t, cslot, cstride, lh + t * LCAP, hist, (uint32_t)nc);
                }
            }
        }
        // Steps past the strip's end decide 0 for every lane, so a segment cut
        // short by the end of the diagonals reports the same (prefix, suffix)
        // as one padded with zeros.
        if (rv) {
#pragma unroll
            for (int t = 0; t < TG; t++)
                if (t < nt) close_segment(&cslot[(size_t)opaque(t) * cstride], run[t], popen, 1u << t);
        }
    }

    __syncthreads();
    for (int e = threadIdx.x; e < TG * LCAP; e += WG) {
        const uint32_t v = lh[e];
        const int t = e / LCAP;
        if (v && t < nt) glob_add(&hist[(size_t)t * (nc + 1) + (e - t * LCAP)], v);
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
    for (int e = threadIdx.x; e < TG; e += WG) acc[e] = 0;
    load_lut(lut_s, lut, P.nb);
    __syncthreads();
    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = wave_id();

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
            pack_ranks(s, R, lut_s, P, S, nt);
            // rows i0+u with u < nrow are in the segment; weight 2 for i < j, 1 for i == j
            const int64_t nr64 = ie - i0;
            const uint32_t mrow = nr64 >= BLK ? 0xFFFFFFFFu : ((1u << nr64) - 1u);
            const int64_t du = j - i0;   // step index of the diagonal element
            const uint32_t mup = !jv ? 0u : (du >= BLK ? 0xFFFFFFFFu : (du <= 0 ? 0u : ((1u << du) - 1u)));
            const uint32_t mdg = (jv && du >= 0 && du < BLK) ? (1u << du) : 0u;
#pragma unroll
            for (int t = 0; t < TG; t++) {
                if (t < nt) {
                    const uint32_t w = swar_word(R, t) & mrow;
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

// ================================================================ MFMA form
// Euclidean norm, fp32 input.  The squared distance is approximated as
//   s~ = (n_i + n_j) - 2 g_ij,   g = x_i . x_j on MFMA (v_mfma_f32_32x32x2f32),
// n_i = |x_i|^2 (accumulated in fp64, rounded to fp32).  For the reference's
// exact-order fp32 sum s:
//   |s - s~| <= (4m + 16) * 2^-24 * (n_i + n_j) + 1e-30 =: E_ij
// (exact-order sum: gamma_{m+1} * 2(n_i+n_j); rounded norms 2u; the fp32 dot
// product gamma_m (n_i+n_j)/2 doubled; the final add and fma 3u; 10u spare).
// A byte table over the bits of s~ gives the rank r = #{t : B_t < s} exactly
// for every bucket that has no bound within E_max = max E_ij of its value
// range; pairs in the other buckets are re-decided in exact order on VALU.
// Inputs whose norms are not finite or exceed 1e30 flag every bucket, so the
// whole matrix is then decided exactly.
constexpr int NBX = 32768;   // byte buckets of the MFMA rank table
constexpr int MXW = 128;     // LDS operand window: 32 + (m-1)*tau samples at most

struct MxParams {
    uint32_t base;   // key bits of bucket 0 (half the smallest bound)
    int sh;          // bucket width in key bits
    int nb;          // buckets used (<= NBX)
    float cE;        // (4m + 16) * 2^-24
};

// n_i = sum_k x[i + k tau]^2 (fp64, rounded to fp32); entries nc .. nc+127 = 0.
// nmax_bits: atomic max of the float bits (non-negative floats order as uints;
// NaN bits are above +inf, so a NaN norm forces the all-exact fallback).
__global__ void k_norms(const float *__restrict__ x, int64_t nc, int tau, int m, float *__restrict__ nrm,
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
// bits 0-5 = rank (number of bounds below it), bit 7 clear.  Otherwise bit 7
// set and bits 0-5 = index of the one bound within reach (bit 6 set if there
// are several): such pairs are checked against that bound with their own
// E_ij, and only the truly ambiguous ones are re-decided in exact order.
template <int TG>
__global__ void k_mx_lut(Bounds<float, TG> S, int nt, MxParams P, const unsigned int *__restrict__ nmax_bits,
                         uint8_t *__restrict__ lut) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= P.nb) return;
    const float nmax = __uint_as_float(*nmax_bits);
    float emax = P.cE * 2.0f * nmax;
    emax = emax * 1.0001f + 1e-30f;
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
    uint32_t e = (uint32_t)r_lo;
    if (near == 1) e = 0x80u | (uint32_t)idx;
    if (near >= 2) e = 0xC0u;
    if (!(emax < __int_as_float(0x7F800000))) e = 0xC0u;   // non-finite or huge norms: all exact
    lut[b] = (uint8_t)e;
}

// Second stage for a flagged pair: rank if the one nearby bound is farther
// than E_ij = cE * (n_i + n_j), else flag for exact re-decision.
__device__ __forceinline__ uint32_t mx_resolve(uint32_t ent, float st, float nsum, const float *bnd, float cE,
                                               uint32_t &amb) {
    if (!(ent & 0x80u)) return ent & 0x3Fu;
    if (ent & 0x40u) {
        amb = 1u;
        return 0u;
    }
    const uint32_t idx = ent & 0x3Fu;
    const float d = st - bnd[idx];
    const float E = fmaf(cE, nsum, 1e-30f) * 1.0001f;
    if (fabsf(d) > E) return idx + (d > 0.0f ? 1u : 0u);
    amb = 1u;
    return 0u;
}

__device__ __forceinline__ uint32_t mx_key_bucket(float st, const MxParams &P) {
    const uint32_t key = __float_as_uint(st);   // st >= +0, finite under the norm guard
    int b = (int)(key - P.base) >> P.sh;
    return (uint32_t)(b < 0 ? 0 : (b > P.nb - 1 ? P.nb - 1 : b));
}

// Exact-order fp32 squared distance of one pair (the reference's sum).
// Loads go out eight dimensions at a time; the sum keeps the order k = 0..m-1.
__device__ __forceinline__ float exact_sqdist(const float *__restrict__ x, int64_t i, int64_t j, int tau, int m) {
    float s = 0.f;
    for (int k0 = 0; k0 < m; k0 += 8) {
        float a[8], b[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int64_t o = (int64_t)(k0 + k < m ? k0 + k : m - 1) * tau;
            a[k] = x[i + o];
            b[k] = x[j + o];
        }
#pragma unroll
        for (int k = 0; k < 8; k++) {
            if (k0 + k < m) {
                const float d = a[k] - b[k];
                const float q = d * d;
                s = s + q;
            }
        }
    }
    return s;
}

template <int TG>
__device__ __forceinline__ uint32_t exact_rank(float s, const Bounds<float, TG> &S, int nt) {
    return rank_direct<float, TG>(s, S, nt);
}

typedef float floatx16 __attribute__((ext_vector_type(16)));

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

// ---------------------------------------------------------------- k_lam_mx
// LAM pass of the MFMA form.  Work item = (band of 32 columns, row segment).
// Lane l owns column j = j0 + l%32; the 32x32 MFMA tile gives it 16 rows of
// that column: rows 8g + 4h + r (g, r < 4, h = l/32), i.e. exactly the rank
// registers R[4h .. 4h+3] of the byte packing.  The two halves swap their
// four registers, then each half runs the words of every other threshold
// (t = 2t' + h) over all 32 rows.
template <int TG>
__global__ __launch_bounds__(WG) void k_lam_mx(const float *__restrict__ x, const float *__restrict__ nrm, int64_t nc,
                                               int tau, int m, Bounds<float, TG> S, int nt,
                                               const uint8_t *__restrict__ lut, MxParams P, int seg_len, int nseg,
                                               unsigned long long *__restrict__ hist, uint2 *__restrict__ carry) {
    constexpr int TH = (TG + 1) / 2;
    constexpr int QCAP = 256;                      // recheck queue entries per wave and round
    __shared__ uint32_t lh[TG * LCAP];
    __shared__ uint8_t lut_s[NBX];
    __shared__ uint32_t rbuf[WAVES][WAVE][4];      // packed ranks, re-read after rechecks
    __shared__ uint16_t queue[WAVES][QCAP];
    __shared__ float bnd_s[TG];
    __shared__ float rwin[WAVES][MXW];   // x[i0 .. i0 + 31 + (m-1)tau] of the current block
    __shared__ float cwin[WAVES][MXW];   // x[j0 .. j0 + 31 + (m-1)tau] of the current item
    for (int e = threadIdx.x; e < TG * LCAP; e += WG) lh[e] = 0;
    for (int e = threadIdx.x; e < P.nb / 4; e += WG) ((uint32_t *)lut_s)[e] = ((const uint32_t *)lut)[e];
    if (threadIdx.x < TG) bnd_s[threadIdx.x] = S.s[threadIdx.x];
    __syncthreads();

    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = wave_id();
    const int col = lane & 31;
    const int h = lane >> 5;
    const int span = 32 + (m - 1) * tau;   // <= MXW (checked by the host)
    const int64_t nbands = (nc + 31) / 32;
    const int64_t nitems = nbands * nseg;
    const size_t cstride = (size_t)nc * nseg;
    const int m2 = (m + 1) / 2;

    for (int64_t item = (int64_t)blockIdx.x * WAVES + wave; item < nitems; item += (int64_t)gridDim.x * WAVES) {
        const int64_t band = item / nseg;
        const int q = (int)(item - band * nseg);
        const int64_t j0 = band * 32;
        const int64_t j = j0 + col;
        const bool jv = j < nc;
        const int64_t ib = (int64_t)q * seg_len;
        const int64_t ie = ib + seg_len < nc ? ib + seg_len : nc;
        const float nj = nrm[j];
        uint2 *cslot = carry + (size_t)(jv ? j : 0) * nseg + q;

        uint32_t run[TH];
#pragma unroll
        for (int t = 0; t < TH; t++) run[t] = 0;
        uint32_t popen = jv ? 0xFFFFFFFFu : 0u;

        // column window once per item; row window of the first block
        float *rw = rwin[wave], *cw = cwin[wave];
        for (int e = lane; e < span; e += WAVE) cw[e] = x[j0 + e];
        float pre0 = lane < span ? x[ib + lane] : 0.f;
        float pre1 = lane + WAVE < span ? x[ib + WAVE + lane] : 0.f;

        for (int64_t i0 = ib; i0 < ie; i0 += BLK) {
            const int nb = ie - i0 < BLK ? (int)(ie - i0) : BLK;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (lane < span) rw[lane] = pre0;
            if (lane + WAVE < span) rw[WAVE + lane] = pre1;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // next block's row window, in flight while this block computes
            // (only if there is a next block: its window ends by x[N + 30])
            const bool more = i0 + BLK < ie;
            pre0 = more && lane < span ? x[i0 + BLK + lane] : 0.f;
            pre1 = more && lane + WAVE < span ? x[i0 + BLK + WAVE + lane] : 0.f;
            // Gram tile G[i0 + a][j0 + b] on MFMA: lane supplies A[row=col][k=h], B[k=h][col]
            floatx16 acc = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll 4
            for (int k2 = 0; k2 < m2; k2++) {
                const int dim = 2 * k2 + h;
                const float a = dim < m ? rw[col + dim * tau] : 0.f;
                const float b = dim < m ? cw[col + dim * tau] : 0.f;
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc, 0, 0, 0);
            }
            // rows of element e: i0 + 8(e>>2) + 4h + (e&3)
            float ni[16];
#pragma unroll
            for (int g = 0; g < 4; g++) {
                const float4 v = *(const float4 *)(nrm + i0 + 8 * g + 4 * h);
                ni[4 * g + 0] = v.x;
                ni[4 * g + 1] = v.y;
                ni[4 * g + 2] = v.z;
                ni[4 * g + 3] = v.w;
            }
            uint32_t Rh[4] = {0u, 0u, 0u, 0u};   // R[4h + c]: rows 4h + c + 8g in byte g
            uint32_t fm = 0;                      // elements to re-decide exactly
#pragma unroll
            for (int e = 0; e < 16; e++) {
                const float sum = ni[e] + nj;
                float st = fmaf(-2.0f, acc[e], sum);
                st = fmaxf(st, 0.0f);
                const uint32_t ent = lut_s[mx_key_bucket(st, P)];
                uint32_t amb = 0;
                const uint32_t rk = mx_resolve(ent, st, sum, bnd_s, P.cE, amb);
                Rh[e & 3] |= rk << (8 * (e >> 2));
                fm |= amb << e;
            }
            if (!jv) fm = 0;
            // rows past the segment are filled later; do not re-decide them
            {
                const int nvr = nb;   // valid rows are i0 .. i0+nb-1
                uint32_t vm = 0;
#pragma unroll
                for (int e = 0; e < 16; e++) vm |= ((8 * (e >> 2) + 4 * h + (e & 3)) < nvr ? 1u : 0u) << e;
                fm &= vm;
            }
            if (__any(fm != 0)) {
                uint32_t total;
                const uint32_t cnt = __popc(fm);
                uint32_t off = wave_excl_scan(cnt, total);
#pragma unroll
                for (int c = 0; c < 4; c++) rbuf[wave][lane][c] = Rh[c];
                uint8_t *rb = (uint8_t *)rbuf[wave];
                for (uint32_t base = 0; base < total; base += QCAP) {
                    uint32_t f = fm;
                    uint32_t o = off;
                    while (f) {
                        const uint32_t e = __builtin_ctz(f);
                        f &= f - 1;
                        if (o >= base && o < base + QCAP) queue[wave][o - base] = (uint16_t)((lane << 4) | e);
                        o++;
                    }
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    const uint32_t nq = total - base < (uint32_t)QCAP ? total - base : (uint32_t)QCAP;
                    for (uint32_t qi = lane; qi < nq; qi += WAVE) {
                        const uint32_t it = queue[wave][qi];
                        const uint32_t ow = it >> 4, e = it & 15u;
                        const int64_t i = i0 + 8 * (e >> 2) + 4 * (ow >> 5) + (e & 3);
                        const int64_t jj = j0 + (ow & 31);
                        const float sx = exact_sqdist(x, i, jj, tau, m);
                        const uint32_t r = exact_rank<TG>(sx, S, nt);
                        // byte (e>>2) of register (e&3) of lane ow
                        rb[(ow * 4 + (e & 3)) * 4 + (e >> 2)] = (uint8_t)r;
                    }
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                }
#pragma unroll
                for (int c = 0; c < 4; c++) Rh[c] = rbuf[wave][lane][c];
            }
            // full rank set of this column: R[0..3] from half 0, R[4..7] from half 1
            uint32_t R[8];
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const uint32_t other = __shfl_xor(Rh[c], 32);
                R[c] = h ? other : Rh[c];
                R[4 + c] = h ? Rh[c] : other;
            }
            const uint32_t fill = nb == BLK ? 0u : (0xFFFFFFFFu << nb);
#pragma unroll
            for (int tp = 0; tp < TH; tp++) {
                const uint32_t t = 2 * tp + h;
                if ((int)t < nt) {
                    const uint32_t K = 0x80808080u + t * 0x01010101u;
                    uint32_t w = 0;
#pragma unroll
                    for (int c = 0; c < 8; c++) w |= ((K - R[c]) >> (7 - c)) & (0x01010101u << c);
                    w = jv ? (w | fill) : 0u;
                    consume_word(w, run[tp], popen, (uint32_t)tp, cslot + (size_t)t * cstride - (size_t)tp * cstride,
                                 cstride, lh + t * LCAP, hist + (size_t)t * (nc + 1) - (size_t)tp * (nc + 1),
                                 (uint32_t)nc);
                    if (fill && jv) run[tp] -= (uint32_t)(BLK - nb);
                }
            }
        }
        if (jv) {
#pragma unroll
            for (int tp = 0; tp < TH; tp++) {
                const uint32_t t = 2 * tp + h;
                if ((int)t < nt) close_segment(&cslot[(size_t)t * cstride], run[tp], popen, 1u << tp);
            }
        }
    }

    __syncthreads();
    for (int e = threadIdx.x; e < TG * LCAP; e += WG) {
        const uint32_t v = lh[e];
        const int t = e / LCAP;
        if (v && t < nt) glob_add(&hist[(size_t)t * (nc + 1) + (e - t * LCAP)], v);
    }
}

// ---------------------------------------------------------------- k_det_mx
// DET pass of the MFMA form.  Work item = (strip of 64 diagonals, step
// segment), as k_det.  A 32-step block needs G(s0+u, s0+r0+v) for v = u + l;
// three 32x32 MFMA tiles cover v = 0..95.  Each tile element is written,
// already combined as n_col - 2 g, into a skewed LDS buffer S[l][u] (l = v - u)
// so that lane l then reads its own diagonal's 32 steps, adds n_row, and packs
// the 32 ranks itself (no exchange).  Elements with v - u outside 0..63 are
// not needed and not written.
constexpr int NBX_DET = 16384;
constexpr int SKS = 33;   // skew buffer row stride (floats): conflict-free writes and reads

template <int TG>
__global__ __launch_bounds__(WG) void k_det_mx(const float *__restrict__ x, const float *__restrict__ nrm, int64_t nc,
                                               int tau, int m, Bounds<float, TG> S, int nt,
                                               const uint8_t *__restrict__ lut, MxParams P, int seg_len, int nseg,
                                               const int2 *__restrict__ items, int64_t nitems,
                                               unsigned long long *__restrict__ hist, uint2 *__restrict__ carry) {
    constexpr int QCAP = 256;
    __shared__ uint32_t lh[TG * LCAP];
    __shared__ uint8_t lut_s[NBX_DET];
    __shared__ float skew[WAVES][WAVE * SKS];      // also holds the packed ranks during rechecks
    __shared__ uint16_t queue[WAVES][QCAP];
    __shared__ float nwin[WAVES][BLK + WAVE];      // n_j of the block's columns
    __shared__ float bnd_s[TG];
    __shared__ float rwin[WAVES][MXW];             // x[s0 .. s0 + 31 + (m-1)tau]
    __shared__ float cwin[WAVES][MXW + WAVE];      // x[c0 .. c0 + 95 + (m-1)tau]
    for (int e = threadIdx.x; e < TG * LCAP; e += WG) lh[e] = 0;
    for (int e = threadIdx.x; e < P.nb / 4; e += WG) ((uint32_t *)lut_s)[e] = ((const uint32_t *)lut)[e];
    if (threadIdx.x < TG) bnd_s[threadIdx.x] = S.s[threadIdx.x];
    __syncthreads();

    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = wave_id();
    const int col = lane & 31;
    const int h = lane >> 5;
    float *sk = skew[wave];
    float *rw = rwin[wave], *cw = cwin[wave];
    const size_t cstride = (size_t)(nc - 1) * nseg;
    const int m2 = (m + 1) / 2;
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
        uint2 *cslot = carry + (size_t)(rv ? r - 1 : 0) * nseg + q;

        uint32_t run[TG];
#pragma unroll
        for (int t = 0; t < TG; t++) run[t] = 0;
        uint32_t popen = rv ? 0xFFFFFFFFu : 0u;

        // operand windows of the first block
        float pr[2], pc[3];
        pr[0] = lane < rspan ? x[sb + lane] : 0.f;
        pr[1] = lane + WAVE < rspan ? x[sb + WAVE + lane] : 0.f;
#pragma unroll
        for (int c = 0; c < 3; c++) pc[c] = lane + c * WAVE < cspan ? x[sb + r0 + c * WAVE + lane] : 0.f;

        for (int64_t s0 = sb; s0 < se; s0 += BLK) {
            const int64_t c0 = s0 + r0;   // first column of tile 0
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
            for (int c = 0; c < 2; c++)
                if (lane + c * WAVE < rspan) rw[lane + c * WAVE] = pr[c];
#pragma unroll
            for (int c = 0; c < 3; c++)
                if (lane + c * WAVE < cspan) cw[lane + c * WAVE] = pc[c];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // next block's windows, in flight while this block computes (only
            // if there is a next block: its windows end by x[N + 94])
            const bool more = s0 + BLK < se;
#pragma unroll
            for (int c = 0; c < 2; c++)
                pr[c] = more && lane + c * WAVE < rspan ? x[s0 + BLK + c * WAVE + lane] : 0.f;
#pragma unroll
            for (int c = 0; c < 3; c++)
                pc[c] = more && lane + c * WAVE < cspan ? x[c0 + BLK + c * WAVE + lane] : 0.f;
            floatx16 acc0 = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
            floatx16 acc1 = acc0, acc2 = acc0;
#pragma unroll 2
            for (int k2 = 0; k2 < m2; k2++) {
                const int dim = 2 * k2 + h;
                const bool dv = dim < m;
                const int off = dim * tau;
                const float a = dv ? rw[col + off] : 0.f;
                const float b0 = dv ? cw[col + off] : 0.f;
                const float b1 = dv ? cw[32 + col + off] : 0.f;
                const float b2 = dv ? cw[64 + col + off] : 0.f;
                acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b0, acc0, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b1, acc1, 0, 0, 0);
                acc2 = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b2, acc2, 0, 0, 0);
            }
            // element e of tile p: row u = 8(e>>2) + 4h + (e&3), column v = 32p + col, diagonal l = v - u
            const float n0 = nrm[c0 + col], n1 = nrm[c0 + 32 + col], n2 = nrm[c0 + 64 + col];
            nwin[wave][lane] = h ? n1 : n0;       // columns c0 .. c0+63
            if (lane < 32) nwin[wave][64 + lane] = n2;   // columns c0+64 .. c0+95
#pragma unroll
            for (int e = 0; e < 16; e++) {
                const int u = 8 * (e >> 2) + 4 * h + (e & 3);
                const int l0 = col - u, l1 = 32 + col - u, l2 = 64 + col - u;
                if (l0 >= 0) sk[l0 * SKS + u] = fmaf(-2.0f, acc0[e], n0);
                sk[l1 * SKS + u] = fmaf(-2.0f, acc1[e], n1);          // 1 <= l1 <= 63 always
                if (l2 < WAVE) sk[l2 * SKS + u] = fmaf(-2.0f, acc2[e], n2);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const float *nrow = nrm + s0;   // uniform: scalar loads
            uint32_t R[8] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
            uint32_t fm = 0;
#pragma unroll
            for (int u = 0; u < BLK; u++) {
                float st = sk[lane * SKS + u] + nrow[u];
                st = fmaxf(st, 0.0f);
                const uint32_t ent = lut_s[mx_key_bucket(st, P)];
                uint32_t amb = 0, rk;
                if (ent & 0x80u)
                    rk = mx_resolve(ent, st, nrow[u] + nwin[wave][lane + u], bnd_s, P.cE, amb);
                else
                    rk = ent & 0x3Fu;
                R[u & 7] |= rk << (8 * (u >> 3));
                fm |= amb << u;
            }
            const int64_t nv64 = dlen - s0;
            const int nvalid = nv64 <= 0 ? 0 : (nv64 >= BLK ? BLK : (int)nv64);
            const uint32_t vmask = nvalid >= BLK ? 0xFFFFFFFFu : ((1u << nvalid) - 1u);
            fm &= vmask;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (__any(fm != 0)) {
                uint32_t *rb32 = (uint32_t *)sk;   // 64 lanes x 8 registers, reused skew buffer
#pragma unroll
                for (int c = 0; c < 8; c++) rb32[lane * 8 + c] = R[c];
                uint8_t *rb = (uint8_t *)sk;
                uint32_t total;
                const uint32_t off = wave_excl_scan(__popc(fm), total);
                for (uint32_t base = 0; base < total; base += QCAP) {
                    uint32_t f = fm, o = off;
                    while (f) {
                        const uint32_t u = __builtin_ctz(f);
                        f &= f - 1;
                        if (o >= base && o < base + QCAP) queue[wave][o - base] = (uint16_t)((lane << 5) | u);
                        o++;
                    }
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    const uint32_t nq = total - base < (uint32_t)QCAP ? total - base : (uint32_t)QCAP;
                    for (uint32_t qi = lane; qi < nq; qi += WAVE) {
                        const uint32_t itq = queue[wave][qi];
                        const uint32_t ow = itq >> 5, u = itq & 31u;
                        const int64_t i = s0 + u;
                        const int64_t jj = i + r0 + ow;
                        const float sx = exact_sqdist(x, i, jj, tau, m);
                        // byte (u>>3) of register (u&7) of lane ow
                        rb[(ow * 8 + (u & 7)) * 4 + (u >> 3)] = (uint8_t)exact_rank<TG>(sx, S, nt);
                    }
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                }
#pragma unroll
                for (int c = 0; c < 8; c++) R[c] = rb32[lane * 8 + c];
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
#pragma unroll
            for (int t = 0; t < TG; t++) {
                if (t < nt) {
                    const uint32_t w = swar_word(R, t) & vmask;
                    consume_word(w, run[t], popen, t, cslot, cstride, lh + t * LCAP, hist, (uint32_t)nc);
                }
            }
        }
        if (rv) {
#pragma unroll
            for (int t = 0; t < TG; t++)
                if (t < nt) close_segment(&cslot[(size_t)opaque(t) * cstride], run[t], popen, 1u << t);
        }
    }

    __syncthreads();
    for (int e = threadIdx.x; e < TG * LCAP; e += WG) {
        const uint32_t v = lh[e];
        const int t = e / LCAP;
        if (v && t < nt) glob_add(&hist[(size_t)t * (nc + 1) + (e - t * LCAP)], v);
    }
}

// ---------------------------------------------------------------- k_stitch
// One thread per (threshold, sequence): joins the (prefix, suffix) runs of
// the sequence's segments in order.  seq_len(sq) gives the sequence length;
// segment q spans steps [q*seg_len, min((q+1)*seg_len, seq_len)).
__global__ void k_stitch(const uint2 *__restrict__ carry, int nt, int64_t nseq, int64_t nc, int is_det, int seg_len,
                         int nseg, unsigned long long *__restrict__ hist) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (int64_t)nt * nseq) return;
    const int t = (int)(gid / nseq);
    const int64_t sq = gid - (int64_t)t * nseq;
    // DET sequences are diagonals r = sq + 1 of length nc - r; LAM sequences are full columns.
    const int64_t len = is_det ? nc - (sq + 1) : nc;
    const int64_t nseg_sq = (len + seg_len - 1) / seg_len;
    const uint2 *c = carry + ((size_t)t * nseq + sq) * nseg;
    unsigned long long *h = hist + (size_t)t * (nc + 1);
    uint64_t cur = 0;
    for (int64_t q = 0; q < nseg_sq; q++) {
        const uint2 pc = c[q];
        const int64_t L = (q + 1) * seg_len <= len ? seg_len : len - q * seg_len;
        if ((int64_t)pc.x == L) {   // whole segment recurrent
            cur += (uint64_t)L;
            continue;
        }
        const uint64_t tot = cur + pc.x;
        if (tot && tot <= (uint64_t)nc) glob_add(&h[tot], 1ull);
        cur = pc.y;
    }
    if (cur && cur <= (uint64_t)nc) glob_add(&h[cur], 1ull);
}

// ---------------------------------------------------------------- k_finalize
// One workgroup per threshold.  Turns a line histogram into the reference's
// five measures for every l_min (AccRQA_functions.cpp:48-133 and
// rqa_process_length_histogram, AccRQA_CPU_function.cpp:181-204).
// For DET it first applies correctDETHistogram (AccRQA_CPU_function.cpp:156-162).
// metric/scan are workspaces of nt*(nc+1) u64 each.
// Output per (t, l): 5 values in F, laid out by out_index(t, l).
constexpr int FIN = 1024;

template <typename F>
__global__ __launch_bounds__(FIN) void k_finalize(unsigned long long *__restrict__ hist, int64_t nc, int nt,
                                                  int is_det, const int *__restrict__ lmins, int nl,
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

    if (is_det) {
        for (int64_t f = tid; f < hs; f += FIN) h[f] = (unsigned long long)(2.0 * (double)h[f]);
        __syncthreads();
        if (tid == 0) h[hs - 1] += 1;
        __syncthreads();
    }

    // suffix sums: thread tid owns chunk [tid*C, min((tid+1)*C, hs))
    const int64_t C = (hs + FIN - 1) / FIN;
    const int64_t c0 = (int64_t)tid * C;
    const int64_t c1 = c0 + C < hs ? c0 + C : hs;
    unsigned long long sm = 0, ss = 0;
    for (int64_t f = c0; f < c1; f++) {
        sm += (unsigned long long)f * h[f];
        ss += h[f];
    }
    __shared__ unsigned long long bm[FIN], bs[FIN];
    bm[tid] = sm;
    bs[tid] = ss;
    __syncthreads();
    // inclusive suffix scan over chunk totals (Hillis-Steele)
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
    unsigned long long rm = tid + 1 < FIN ? bm[tid + 1] : 0ull;   // sum of chunks after mine
    unsigned long long rs = tid + 1 < FIN ? bs[tid + 1] : 0ull;
    for (int64_t f = c1 - 1; f >= c0; f--) {
        rm += (unsigned long long)f * h[f];
        rs += h[f];
        M[f] = rm;
        Sc[f] = rs;
    }
    __syncthreads();
    __threadfence_block();

    // Lmax / TTmax: the reference's bisection (tertiary_metric), one thread.
    __shared__ unsigned long long lmax_s;
    if (tid == 0) {
        unsigned long long last, start, mid, end, length, mv, mv1;
        bool found = false;
        end = hs - 1;
        mid = hs / 2;
        last = is_det ? Sc[end] : 0ull;
        start = 0;
        length = hs;
        while (!found) {
            mv = Sc[mid];
            mv1 = mid > 0 ? Sc[mid - 1] : 0ull;
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
        if (m0_out) m0_out[t] = M[0];   // recurrence count (sum of l * h[l])
    }
    __syncthreads();

    __shared__ double red[FIN];
    const int ot = perm ? perm[t] : t;   // position of this threshold in the caller's order
    for (int l = 0; l < nl; l++) {
        const int64_t lm = lmins[l];
        const unsigned long long sl = Sc[lm];
        double e = 0.0;
        for (int64_t f = lm + tid; f < hs; f += FIN) {
            const double p = (double)h[f] / (double)sl;
            if (p > 0) e += p * log(p);
        }
        red[tid] = e;
        __syncthreads();
        for (int off = FIN / 2; off > 0; off >>= 1) {
            if (tid < off) red[tid] += red[tid + off];
            __syncthreads();
        }
        if (tid == 0) {
            const unsigned long long m0 = M[0], ml = M[lm];
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
