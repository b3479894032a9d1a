// Host test of the kernels' word-level line scan (accrqa_amd/csrc/rqa_bits.hpp):
//  1. rank_planes + GeTree reproduce [rank > t] for every step and threshold;
//  2. consume() over a segment of words reports exactly the runs a direct
//     scan finds: inner runs to the histogram, the first run as the segment
//     prefix, the last as the suffix (the layout k_stitch joins).
// Prints "ok" and exits 0 on success.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "rqa_bits.hpp"

using namespace rqa;

static int fails = 0;
#define CHECK(c, ...)                                   \
    do {                                                \
        if (!(c)) {                                     \
            if (fails < 20) std::printf(__VA_ARGS__);   \
            fails++;                                    \
        }                                               \
    } while (0)

template <int TG>
static void test_planes(std::mt19937_64 &g) {
    constexpr int B = rank_bits<TG>();
    for (int it = 0; it < 20000; it++) {
        int rk[32];
        uint32_t R[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        const int mode = it % 4;
        for (int u = 0; u < 32; u++) {
            int v = (int)(g() % (TG + 1));
            if (mode == 1) v = (it / 4) % (TG + 1);   // constant blocks
            if (mode == 2) v = (u * 7 + it) % (TG + 1);
            rk[u] = v;
            R[u & 7] |= (uint32_t)v << (8 * (u >> 3));
        }
        uint32_t P[B];
        rank_planes<B>(R, P);
        for (int k = 0; k < B; k++)
            for (int u = 0; u < 32; u++)
                CHECK(((P[k] >> u) & 1u) == (uint32_t)((rk[u] >> k) & 1), "plane TG=%d k=%d u=%d\n", TG, k, u);
        GeTree<B> G(P);
        for (int c = 1; c < (1 << B); c++) {
            const uint32_t w = G.ge(c);
            for (int u = 0; u < 32; u++)
                CHECK(((w >> u) & 1u) == (uint32_t)(rk[u] >= c), "ge TG=%d c=%d u=%d\n", TG, c, u);
        }
    }
}

struct Result {
    std::vector<long> hist;   // counts by length (inner runs only)
    long pre, suf;
};

// Direct scan of a 0/1 sequence.
static Result direct(const std::vector<int> &b) {
    const long L = (long)b.size();
    Result r{std::vector<long>(L + 2, 0), 0, 0};
    long i = 0;
    while (i < L && b[i]) i++;
    if (i == L) {
        r.pre = r.suf = L;
        return r;
    }
    r.pre = i;
    long j = L;
    while (j > 0 && b[j - 1]) j--;
    r.suf = L - j;
    long run = 0;
    for (long k = i; k < j; k++) {
        if (b[k])
            run++;
        else {
            if (run) r.hist[run]++;
            run = 0;
        }
    }
    return r;
}

template <uint32_t CAP>
static Result scanned(const std::vector<int> &b, uint32_t start_run, uint32_t spare) {
    const long L = (long)b.size();
    Result r{std::vector<long>(L + 2, 0), -1, -1};
    uint32_t run = start_run;
    long pre = -1;
    for (long w0 = 0; w0 < L; w0 += 32) {
        const int nb = (int)(L - w0 < 32 ? L - w0 : 32);
        uint32_t nw = 0;
        for (int u = 0; u < nb; u++)
            if (!b[w0 + u]) nw |= 1u << u;
        // steps past the end are filled as recurrent and taken back afterwards
        consume<CAP>(
            nw, run, spare,
            [&](uint32_t b) {
                // b + 1 is the bin: a counted run below CAP, or a spare bin
                if (b <= CAP - 2) {
                    CHECK((long)b + 1 <= L, "run of %u steps in a sequence of %ld\n", b + 1, L);
                    if ((long)b + 1 <= L) r.hist[b + 1]++;
                }
                CHECK(b <= spare, "first-run bin %u beyond the spare bin %u\n", b, spare);
            },
            [&](uint32_t len) {
                CHECK(len >= 1 && len <= 30, "inner run of %u steps\n", len);
                r.hist[len]++;
            },
            [&](uint32_t first) {
                if (first >= PRE) {
                    CHECK(pre < 0, "prefix written twice\n");
                    pre = first - PRE;
                } else {
                    CHECK((long)first <= L, "slow run of %u steps in a sequence of %ld\n", first, L);
                    if ((long)first <= L) r.hist[first]++;
                }
            });
        if (nb < 32) run -= (uint32_t)(32 - nb);
    }
    const SegEnds e = seg_ends(run);
    if (e.pre_open) {
        CHECK(pre < 0, "open prefix but written\n");
        r.pre = e.pre;
    } else {
        r.pre = pre;
    }
    r.suf = e.suf;
    return r;
}

// CAP: the LDS histogram size of the caller (128 for the large-N kernels, 64
// for the batched kernel).
template <uint32_t CAP>
static void test_consume(std::mt19937_64 &g) {
    for (int it = 0; it < 40000; it++) {
        const long L = 1 + (long)(g() % (it % 10 == 0 ? 2000 : 300));
        std::vector<int> b(L);
        const int mode = it % 6;
        const double p = (double)(g() % 1000) / 1000.0;
        for (long k = 0; k < L; k++) {
            if (mode == 0) b[k] = (g() % 1000) < p * 1000;
            else if (mode == 1) b[k] = 1;
            else if (mode == 2) b[k] = 0;
            else if (mode == 3) b[k] = (k / (1 + it % 37)) % 2;   // regular blocks, some >= LCAP
            else if (mode == 4) b[k] = (g() % 50) != 0;          // long runs
            else b[k] = (g() % 50) == 0;                         // isolated ones
        }
        const Result d = direct(b), s = scanned<CAP>(b, PRE, CAP - 1 + (uint32_t)(g() % NSPARE));
        CHECK(d.pre == s.pre, "L=%ld mode=%d pre %ld vs %ld\n", L, mode, d.pre, s.pre);
        CHECK(d.suf == s.suf, "L=%ld mode=%d suf %ld vs %ld\n", L, mode, d.suf, s.suf);
        for (long k = 0; k <= L; k++)
            CHECK(d.hist[k] == s.hist[k], "L=%ld mode=%d hist[%ld] %ld vs %ld\n", L, mode, k, d.hist[k], s.hist[k]);
    }
}

// Carried runs at the segment-length limit (2^29 steps, the open-prefix flag
// set or not) still take the slow path when they close, and an all-recurrent
// word only adds 32 to them.
static void test_long_carry() {
    for (uint32_t flag : {PRE, 0u}) {
        for (uint32_t nw : {0x1u, 0x10u, 0x80000000u, 0u}) {
            uint32_t run = flag + (1u << 29) - 3u;
            const uint32_t before = run;
            bool slow = false;
            uint32_t got = 0;
            consume<LCAP>(
                nw, run, LCAP - 1u,
                [&](uint32_t b) { CHECK(b >= LCAP - 1u, "long run counted in the LDS bins (b=%u)\n", b); },
                [&](uint32_t) { CHECK(false, "inner run in a word with one non-recurrent step\n"); },
                [&](uint32_t f) {
                    slow = true;
                    got = f;
                });
            if (nw == 0u) {
                CHECK(!slow && run == before + 32u, "all-recurrent word: slow %d run %u\n", (int)slow, run);
            } else {
                const uint32_t lead = (uint32_t)__builtin_ctz(nw);
                CHECK(slow && got == before + lead, "closing long run: slow %d first %u\n", (int)slow, got);
            }
        }
    }
}

int main() {
    std::mt19937_64 g(12345);
    test_long_carry();
    test_planes<4>(g);
    test_planes<8>(g);
    test_planes<16>(g);
    test_planes<32>(g);
    test_consume<LCAP>(g);
    test_consume<64>(g);
    if (fails) {
        std::printf("FAILED %d checks\n", fails);
        return 1;
    }
    std::printf("ok\n");
    return 0;
}
