// rqa_bits.hpp -- word-level pieces of the line scan: decision words from
// packed threshold ranks, and run extraction from a 32-step word.  Plain
// integer code, compiled into the kernels and, unchanged, into the host unit
// test of the line scan (tests/test_line_scan_bits.py).
//
// Conventions.  A lane walks one sequence (a column for LAM, a diagonal for
// DET) in blocks of 32 steps.  For each step the engine has a rank
// r = #{t : B_t < s} over the group's sorted squared bounds, so threshold t
// recurs iff r <= t.  Ranks are packed one per byte: register q holds steps
// q, q+8, q+16, q+24 (byte g = step q + 8g).  Words are "not recurrent" masks:
// bit u of nw(t) is set iff step u is not recurrent for threshold t.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define RQA_HD __host__ __device__ __forceinline__
#else
#define RQA_HD inline
#endif

namespace rqa {

constexpr int LCAP = 128;              // run lengths < LCAP go to the workgroup's LDS histogram
// Run counter flag: the segment's first run has not closed yet.  Bit 30, so
// that every run counter (at most PRE plus a segment of at most 2^29 steps and
// a word) is positive as a signed integer and an all-ones word stands apart
// from it.
constexpr uint32_t PRE = 0x40000000u;
// Spare bins after each row of an LDS run histogram: a lane adds one there for
// a word whose carried run counts nothing (none closes, it is empty, or it
// goes to the slow path), so that the wave's adds that count nothing spread
// over NSPARE addresses instead of meeting on one.
constexpr int NSPARE = 16;
constexpr int LROW = LCAP + NSPARE;    // LDS words per threshold row

// Position of the lowest / highest set bit (as the hardware's find-first-bit
// instructions): 0xFFFFFFFF for a zero word.  On the device they are the
// instructions themselves, so the compiler does not split the zero case off
// into a select at every use.
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __forceinline__ uint32_t ffbl(uint32_t v) {
    uint32_t r;
    asm("v_ffbl_b32 %0, %1" : "=v"(r) : "v"(v));
    return r;
}
__device__ __forceinline__ uint32_t ffbh(uint32_t v) {
    uint32_t r;
    asm("v_ffbh_u32 %0, %1" : "=v"(r) : "v"(v));
    return r;
}
#else
inline uint32_t ffbl(uint32_t v) { return v ? (uint32_t)__builtin_ctz(v) : 0xFFFFFFFFu; }
inline uint32_t ffbh(uint32_t v) { return v ? (uint32_t)__builtin_clz(v) : 0xFFFFFFFFu; }
#endif

// All-ones shifted right by (s mod 32), as the hardware's shift takes its
// count (ffbh(0) = 0xFFFFFFFF gives 1).
RQA_HD uint32_t ones_shr(uint32_t s) { return 0xFFFFFFFFu >> (s & 31u); }

// a + b, saturating at 0xFFFFFFFF (one clamped add on the device).
RQA_HD uint32_t add_sat(uint32_t a, uint32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_elementwise_add_sat(a, b);
#else
    const uint32_t s = a + b;
    return s < a ? 0xFFFFFFFFu : s;
#endif
}

RQA_HD uint32_t umin(uint32_t a, uint32_t b) { return a < b ? a : b; }

// Bits of a rank in [0, TG].
template <int TG>
constexpr int rank_bits() {
    return TG < 2 ? 1 : TG < 4 ? 2 : TG < 8 ? 3 : TG < 16 ? 4 : TG < 32 ? 5 : 6;
}

// Exchange of the off-diagonal sub-blocks of an 8x8 bit matrix held one row
// per byte (four matrices side by side): bits (k+s) of row a trade places
// with bits k of row a+s, for the bits k selected by m.
RQA_HD void bit_swap(uint32_t &a, uint32_t &b, int s, uint32_t m) {
    const uint32_t na = (a & ~(m << s)) | ((b << s) & (m << s));
    const uint32_t nb = (b & ~m) | ((a >> s) & m);
    a = na;
    b = nb;
}

// Bit planes of the packed ranks: P[k] bit (q + 8g) = bit k of the rank of
// step q + 8g, i.e. bit u of P[k] is bit k of step u's rank.  Each byte
// column is an 8x8 bit transpose (rows = registers, columns = rank bits),
// done in three exchange stages.  Then steps outside `keep` get rank 0
// (recurrent for every threshold) and steps in `force` the largest rank
// (recurrent for none): B operations for the whole word instead of one per
// threshold.
template <int B>
RQA_HD void rank_planes(const uint32_t (&R)[8], uint32_t (&P)[B], uint32_t keep = 0xFFFFFFFFu,
                        uint32_t force = 0u) {
    uint32_t r[8];
#pragma unroll
    for (int q = 0; q < 8; q++) r[q] = R[q];
#pragma unroll
    for (int q = 0; q < 4; q++) bit_swap(r[q], r[q + 4], 4, 0x0F0F0F0Fu);
#pragma unroll
    for (int q = 0; q < 8; q += 4) {
        bit_swap(r[q], r[q + 2], 2, 0x33333333u);
        bit_swap(r[q + 1], r[q + 3], 2, 0x33333333u);
    }
#pragma unroll
    for (int q = 0; q < 8; q += 2) bit_swap(r[q], r[q + 1], 1, 0x55555555u);
#pragma unroll
    for (int k = 0; k < B; k++) P[k] = (r[k] & keep) | force;
}

// Comparator tree over the planes: ge(c) = mask of steps with rank >= c, so
// nw(t) = ge(t + 1).  lo[c] (c < 2^(B-1)) compares the low B-1 bits; the top
// bit is applied per threshold.  All indices are compile-time constants
// after unrolling, so the tree costs one logic operation per node.
template <int B>
struct GeTree {
    static constexpr int H = 1 << (B - 1);
    uint32_t lo[H];
    uint32_t top;

    RQA_HD GeTree(const uint32_t (&P)[B]) {
        lo[0] = 0xFFFFFFFFu;
#pragma unroll
        for (int j = 0; j < B - 1; j++) {
            const int n = 1 << j;
#pragma unroll
            for (int c = n; c < 2 * n; c++) lo[c] = P[j] & lo[c - n];
#pragma unroll
            for (int c = 0; c < n; c++) lo[c] = P[j] | lo[c];
        }
        top = P[B - 1];
    }
    // rank >= c, for 1 <= c < 2^B
    RQA_HD uint32_t ge(int c) const { return (c & H) ? (top & lo[c & (H - 1)]) : (top | lo[c & (H - 1)]); }
};

// One 32-step word of one threshold.  nw: bit u set = step u not recurrent
// (bit 0 = oldest).  run: recurrent steps carried in from earlier words, with
// PRE added while the segment's first run is open.  Counting goes through
// three callbacks:
//   first(b): the carried run, closed at the word's first non-recurrent step;
//     b <= CAP - 2 counts a run of length b + 1, larger b (at most `spare`,
//     the caller's spare bin CAP - 1 + k) counts nothing: no run closes, the
//     closing run is empty, or it takes the slow path.  Always called, so the
//     common case needs no branch;
//   inner(len): a run strictly inside the word, 1 <= len <= 30;
//   slow(first): the carried run when it is >= CAP long or is the segment's
//     first run (first >= PRE: its length is first - PRE, 0 included).
// CAP (at least 32) is the caller's LDS histogram size: LCAP for the large-N
// kernels.  No select or compare is spent on "does a run close": for nw == 0
// both find-first-bit results are all ones, so the saturating sum is all
// ones (a spare bin, and negative as a signed integer, so not slow) and the
// new carry min(run + 32, tail) is run + 32.
template <uint32_t CAP = LCAP, class First, class Inner, class Slow>
RQA_HD void consume(uint32_t nw, uint32_t &run, uint32_t spare, First &&first_bin, Inner &&inner, Slow &&slow) {
    static_assert(CAP >= 32, "inner runs (up to 30 steps) always take the histogram");
    const uint32_t lead = ffbl(nw);                 // recurrent steps before the first non-recurrent one
    const uint32_t tail = ffbh(nw);                 // recurrent steps after the last non-recurrent one
    const uint32_t first = add_sat(run, lead);      // the carried run, closed by that first step
    first_bin(umin(first - 1u, spare));             // first = 0 wraps to a spare bin too
    if ((int32_t)first >= (int32_t)CAP) slow(first);
    // runs strictly inside the word: recurrent steps above the lowest
    // non-recurrent one (~nw & -nw) and at or below the highest (for nw == 0
    // mid is 0)
    uint32_t mid = ~nw & (0u - nw) & ones_shr(tail);
    while (mid) {
        // both operands are non-zero: mid by the loop test, nw >> lo because
        // an inner run ends before the word's last non-recurrent step
        const uint32_t lo = (uint32_t)__builtin_ctz(mid);
        inner((uint32_t)__builtin_ctz(nw >> lo));   // ends at the next non-recurrent step
        mid &= mid + (1u << lo);                     // clear that run
    }
    run = umin(run + 32u, tail);
}

// Segment end: (prefix run, suffix run) for the stitch.  A segment whose
// first run never closed is recurrent throughout: prefix = suffix = length.
struct SegEnds {
    uint32_t pre, suf;
    bool pre_open;
};
RQA_HD SegEnds seg_ends(uint32_t run) {
    if (run >= PRE) return {run - PRE, run - PRE, true};
    return {0u, run, false};
}

}  // namespace rqa
