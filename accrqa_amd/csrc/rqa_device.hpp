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
constexpr int LCAP = 256;          // run lengths < LCAP use the LDS histogram
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

// gh points at a histogram row of length hmax + 1; a run longer than the
// sequence would be an engine bug, and is dropped rather than written out of
// bounds (the parity tests would then fail on the missing count).
__device__ __forceinline__ void emit(uint32_t len, uint32_t *lh, unsigned long long *gh, uint32_t hmax) {
    if (len < (uint32_t)LCAP)
        lds_inc(&lh[len]);
    else if (len <= hmax)
        glob_add(&gh[len], 1ull);
}

// ---------------------------------------------------------------- decision words
// w = 2w + [s_u <= bound] for u = 0..31, i.e. bit 31-u holds step u.  One
// v_cmp (writes a lane mask) and one v_addc (consumes it as carry-in) per
// decision.  Four compares go to four SGPR pairs before the four adds read
// them, so every mask is read at least four instructions after it is
// written (gfx950 requires two wait states between a VALU mask write and a
// VALU carry read).  NaN s never compares true, as in the reference.
template <typename F>
__device__ __forceinline__ uint32_t decision_word(const F (&s)[BLK], F bound) {
    uint32_t w = 0;
#pragma unroll
    for (int u = 0; u < BLK; u += 4) {
        uint64_t m0, m1, m2, m3, c;
        if constexpr (sizeof(F) == 4) {
            asm("v_cmp_ge_f32_e64 %[m0], %[b], %[s0]\n\t"
                "v_cmp_ge_f32_e64 %[m1], %[b], %[s1]\n\t"
                "v_cmp_ge_f32_e64 %[m2], %[b], %[s2]\n\t"
                "v_cmp_ge_f32_e64 %[m3], %[b], %[s3]\n\t"
                "v_addc_co_u32_e64 %[w], %[c], %[w], %[w], %[m0]\n\t"
                "v_addc_co_u32_e64 %[w], %[c], %[w], %[w], %[m1]\n\t"
                "v_addc_co_u32_e64 %[w], %[c], %[w], %[w], %[m2]\n\t"
                "v_addc_co_u32_e64 %[w], %[c], %[w], %[w], %[m3]"
                : [w] "+v"(w), [m0] "=&s"(m0), [m1] "=&s"(m1), [m2] "=&s"(m2), [m3] "=&s"(m3), [c] "=&s"(c)
                : [b] "v"(bound), [s0] "v"(s[u]), [s1] "v"(s[u + 1]), [s2] "v"(s[u + 2]), [s3] "v"(s[u + 3]));
        } else {
            asm("v_cmp_ge_f64_e64 %[m0], %[b], %[s0]\n\t"
                "v_cmp_ge_f64_e64 %[m1], %[b], %[s1]\n\t"
                "v_cmp_ge_f64_e64 %[m2], %[b], %[s2]\n\t"
                "v_cmp_ge_f64_e64 %[m3], %[b], %[s3]\n\t"
                "v_addc_co_u32_e64 %[w], %[c], %[w], %[w], %[m0]\n\t"
                "v_addc_co_u32_e64 %[w], %[c], %[w], %[w], %[m1]\n\t"
                "v_addc_co_u32_e64 %[w], %[c], %[w], %[w], %[m2]\n\t"
                "v_addc_co_u32_e64 %[w], %[c], %[w], %[w], %[m3]"
                : [w] "+v"(w), [m0] "=&s"(m0), [m1] "=&s"(m1), [m2] "=&s"(m2), [m3] "=&s"(m3), [c] "=&s"(c)
                : [b] "v"(bound), [s0] "v"(s[u]), [s1] "v"(s[u + 1]), [s2] "v"(s[u + 2]), [s3] "v"(s[u + 3]));
        }
    }
    return w;
}

// Weighted count: cnt += [s_u <= bound] for u = 0..31 (same interleaving).
template <typename F>
__device__ __forceinline__ uint32_t count_le(const F (&s)[BLK], F bound, uint32_t cnt) {
#pragma unroll
    for (int u = 0; u < BLK; u += 4) {
        uint64_t m0, m1, m2, m3, c;
        if constexpr (sizeof(F) == 4) {
            asm("v_cmp_ge_f32_e64 %[m0], %[b], %[s0]\n\t"
                "v_cmp_ge_f32_e64 %[m1], %[b], %[s1]\n\t"
                "v_cmp_ge_f32_e64 %[m2], %[b], %[s2]\n\t"
                "v_cmp_ge_f32_e64 %[m3], %[b], %[s3]\n\t"
                "v_addc_co_u32_e64 %[n], %[c], %[n], 0, %[m0]\n\t"
                "v_addc_co_u32_e64 %[n], %[c], %[n], 0, %[m1]\n\t"
                "v_addc_co_u32_e64 %[n], %[c], %[n], 0, %[m2]\n\t"
                "v_addc_co_u32_e64 %[n], %[c], %[n], 0, %[m3]"
                : [n] "+v"(cnt), [m0] "=&s"(m0), [m1] "=&s"(m1), [m2] "=&s"(m2), [m3] "=&s"(m3), [c] "=&s"(c)
                : [b] "v"(bound), [s0] "v"(s[u]), [s1] "v"(s[u + 1]), [s2] "v"(s[u + 2]), [s3] "v"(s[u + 3]));
        } else {
            asm("v_cmp_ge_f64_e64 %[m0], %[b], %[s0]\n\t"
                "v_cmp_ge_f64_e64 %[m1], %[b], %[s1]\n\t"
                "v_cmp_ge_f64_e64 %[m2], %[b], %[s2]\n\t"
                "v_cmp_ge_f64_e64 %[m3], %[b], %[s3]\n\t"
                "v_addc_co_u32_e64 %[n], %[c], %[n], 0, %[m0]\n\t"
                "v_addc_co_u32_e64 %[n], %[c], %[n], 0, %[m1]\n\t"
                "v_addc_co_u32_e64 %[n], %[c], %[n], 0, %[m2]\n\t"
                "v_addc_co_u32_e64 %[n], %[c], %[n], 0, %[m3]"
                : [n] "+v"(cnt), [m0] "=&s"(m0), [m1] "=&s"(m1), [m2] "=&s"(m2), [m3] "=&s"(m3), [c] "=&s"(c)
                : [b] "v"(bound), [s0] "v"(s[u]), [s1] "v"(s[u + 1]), [s2] "v"(s[u + 2]), [s3] "v"(s[u + 3]));
        }
    }
    return cnt;
}

// Consume one 32-step word (bit 31 = oldest step).  run: open run entering
// the word.  Bit tb of popen is set until the segment's first zero; the run
// it closes is the segment's prefix run, written to *pre (the stitch joins
// it with the previous segment) instead of being counted here.
__device__ __forceinline__ void consume_word(uint32_t w, uint32_t &run, uint32_t &popen, uint32_t tb, uint32_t *pre,
                                             uint32_t *lh, unsigned long long *gh, uint32_t hmax) {
    if (w == 0xFFFFFFFFu) {
        run += BLK;
        return;
    }
    const uint32_t lead = __clz(~w);             // ones above the highest zero
    const uint32_t first = run + lead;
    if (popen & tb) {
        *pre = first;
        popen &= ~tb;
    } else if (first) {
        emit(first, lh, gh, hmax);
    }
    const uint32_t tail = __builtin_ctz(~w);     // ones below the lowest zero
    uint32_t mid = w & (0xFFFFFFFFu >> lead) & ~((1u << tail) - 1u);
    while (mid) {
        const uint32_t h = 31u - __clz(mid);
        const uint32_t len = __clz(~(mid << (31u - h)));
        emit(len, lh, gh, hmax);
        mid &= ~(((1u << len) - 1u) << (h + 1u - len));
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

// Copy a uniform value into a VGPR (keeps the threshold bounds out of the
// scalar register file, which the row window and the lane masks need).
template <typename F>
__device__ __forceinline__ F to_vgpr(F v) {
    F r;
    if constexpr (sizeof(F) == 4)
        asm volatile("v_mov_b32 %0, %1" : "=v"(r) : "s"(v));
    else
        asm volatile("v_mov_b64 %0, %1" : "=v"(r) : "s"(v));
    return r;
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
                                            int nt, int seg_len, int nseg, unsigned long long *__restrict__ hist,
                                            uint2 *__restrict__ carry) {
    __shared__ uint32_t lh[TG * LCAP];
    for (int e = threadIdx.x; e < TG * LCAP; e += WG) lh[e] = 0;
    __syncthreads();

    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = wave_id();
    const int64_t nbands = (nc + WAVE - 1) / WAVE;
    const int64_t nitems = nbands * nseg;
    const size_t cstride = (size_t)nc * nseg;   // carry stride between thresholds
    F bnd[TG];
#pragma unroll
    for (int t = 0; t < TG; t++) bnd[t] = to_vgpr(S.s[t]);

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
            const uint32_t fill = nb == BLK ? 0u : ((1u << (BLK - nb)) - 1u);
#pragma unroll
            for (int t = 0; t < TG; t++) {
                if (t < nt) {
                    uint32_t w = decision_word(s, bnd[t]);
                    w = jv ? (w | fill) : 0u;
                    consume_word(w, run[t], popen, 1u << t, &cslot[(size_t)t * cstride].x, lh + t * LCAP,
                                 hist + (size_t)t * (nc + 1), (uint32_t)nc);
                    if (fill && jv) run[t] -= (uint32_t)(BLK - nb);
                }
            }
        }
        if (jv) {
#pragma unroll
            for (int t = 0; t < TG; t++)
                if (t < nt) close_segment(&cslot[(size_t)t * cstride], run[t], popen, 1u << t);
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
                                            int nt, int seg_len, int nseg, const int2 *__restrict__ items,
                                            int64_t nitems, unsigned long long *__restrict__ hist,
                                            uint2 *__restrict__ carry) {
    __shared__ uint32_t lh[TG * LCAP];
    __shared__ F win[WAVES][DET_WIN];
    for (int e = threadIdx.x; e < TG * LCAP; e += WG) lh[e] = 0;
    __syncthreads();

    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = wave_id();
    F *wn = win[wave];
    // dimensions staged per window: span = BLK + WAVE - 1 + (kc-1)*tau <= DET_WIN
    const int kc_max = tau > 0 ? (DET_WIN - (BLK + WAVE - 1)) / tau + 1 : m;
    const int kc = kc_max < m ? kc_max : m;
    const size_t cstride = (size_t)(nc - 1) * nseg;
    F bnd[TG];
#pragma unroll
    for (int t = 0; t < TG; t++) bnd[t] = to_vgpr(S.s[t]);

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
            // bit (31-u) valid iff s0+u < dlen; invalid steps decide 0
            const int64_t nv64 = dlen - s0;
            const int nvalid = nv64 <= 0 ? 0 : (nv64 >= BLK ? BLK : (int)nv64);
            const uint32_t vmask = nvalid == 0 ? 0u : (0xFFFFFFFFu << (BLK - nvalid));
#pragma unroll
            for (int t = 0; t < TG; t++) {
                if (t < nt) {
                    const uint32_t w = decision_word(s, bnd[t]) & vmask;
                    consume_word(w, run[t], popen, 1u << t, &cslot[(size_t)t * cstride].x, lh + t * LCAP,
                                 hist + (size_t)t * (nc + 1), (uint32_t)nc);
                }
            }
        }
        // Steps past the strip's end decide 0 for every lane, so a segment cut
        // short by the end of the diagonals reports the same (prefix, suffix)
        // as one padded with zeros.
        if (rv) {
#pragma unroll
            for (int t = 0; t < TG; t++)
                if (t < nt) close_segment(&cslot[(size_t)t * cstride], run[t], popen, 1u << t);
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
                                           int nt, int seg_len, const int2 *__restrict__ items, int64_t nitems,
                                           unsigned long long *__restrict__ counts) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = wave_id();
    __shared__ unsigned long long acc[TG];
    for (int e = threadIdx.x; e < TG; e += WG) acc[e] = 0;
    __syncthreads();
    F bnd[TG];
#pragma unroll
    for (int t = 0; t < TG; t++) bnd[t] = to_vgpr(S.s[t]);

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
            for (int k = 0; k < m; k++) {
                const F xj = x[j + (int64_t)k * tau];
                const F *xr = x + i0 + (int64_t)k * tau;
#pragma unroll
                for (int u = 0; u < BLK; u++) s[u] = dim_step<NORM>(s[u], xr[u], xj);
            }
            // whole block strictly above every column of a complete band
            const bool full = (i0 + BLK <= j0) && (i0 + BLK <= ie) && (j0 + WAVE <= nc);
            if (full) {
#pragma unroll
                for (int t = 0; t < TG; t++)
                    if (t < nt) c2[t] = count_le(s, bnd[t], c2[t]);
            } else {
#pragma unroll
                for (int t = 0; t < TG; t++) {
                    if (t < nt) {
                        const uint32_t w = decision_word(s, bnd[t]);
#pragma unroll
                        for (int u = 0; u < BLK; u++) {
                            const int64_t i = i0 + u;
                            const bool r = jv && (i < ie) && ((w >> (BLK - 1 - u)) & 1u);
                            c2[t] += (r && i < j) ? 1u : 0u;
                            c1[t] += (r && i == j) ? 1u : 0u;
                        }
                    }
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
