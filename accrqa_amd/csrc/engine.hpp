// engine.hpp -- host-side pieces shared by the engine's translation units
// (rqa_engine.hip: the large-N path and most of the C ABI; rqa_batch.hip: the
// batched engine).  Internal: not part of the installed interface.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <map>
#include <utility>
#include <mutex>
#include <string>

namespace rqa {
constexpr int ENG_WG = 256;   // threads per workgroup of the persistent kernels
}

namespace eng {

// The reference's Accrqa_Error values used here.
enum : int {
    E_OK = 0,
    E_INVALID = 2,
    E_NULL = 4,
    E_NO_DEVICE = 8,
    E_NO_MEMORY = 9,
    E_DEVICE = 10,
};

extern std::string g_last_error;

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    uint64_t gen = 0;   // allocation number (a new one for every hipMalloc)
};

// Per-process device context: the current device, the library's stream and
// the work buffers, which grow on demand and are reused across calls.
struct Ctx {
    int device = -1;
    int ncu = 0;
    hipStream_t stream = nullptr;
    // second stream of the split form: the DET tile scan runs there beside the
    // LAM scans, after ev_tiles (the tiles are written) and before ev_det
    hipStream_t side = nullptr;
    hipEvent_t ev_tiles = nullptr, ev_det = nullptr;
    DevBuf x, hist, carry, metric, scan, counts, perm, lmins, lut, nrm, nmax, lut8, out_rr, out_det, out_lam, priv,
        tiles, dump, tops, mxstats, probe, out_rp, bhist, bdesc, hist2, carry2, priv2;
    // segment length of the current (tau, emb) block, and whether its split
    // form runs slab by slab (the tile store of all rows does not fit)
    int blk_L = 0;
    bool blk_slab = false;
    // per-slab item ranges of the current item tables (slab form): items
    // [off[s], off[s+1]) of each table belong to slab s
    const int64_t *slab_lamv = nullptr, *slab_lamh = nullptr, *slab_dets = nullptr;
    int nslabs = 0;
    // work-item tables of the current (nc, seg_len): views of an ItemSet
    // entry (owned there, never freed through these)
    DevBuf items_rr, items_det, items_dets, items_lamv, items_lamh;
    // Raw-result sinks of accrqa_amd_histograms_*: when set, launch_group
    // copies each group's counts / h_D rows / h_V rows there, indexed by the
    // caller's threshold order.  Null for every other entry point.
    unsigned long long *dump_rr = nullptr, *dump_det = nullptr, *dump_lam = nullptr;
    int64_t n_items_rr = 0, n_items_det = 0, n_items_dets = 0, n_items_lamv = 0, n_items_lamh = 0;
};

extern std::mutex g_mu;
extern Ctx g_ctx;

int map_err(hipError_t e);

#define HIP_TRY(expr)                             \
    do {                                          \
        hipError_t e_ = (expr);                   \
        if (e_ != hipSuccess) return map_err(e_); \
    } while (0)

// Grows b to at least `bytes` (never shrinks).
int ensure(DevBuf &b, size_t bytes);
// Host -> device copy in stream order through pinned staging.
int stage_upload(void *dst, const void *src, size_t bytes, hipStream_t st);
// Current device's context (a fresh one when the device changed).
int init_ctx();
// Ordering of consecutive calls on different streams (see rqa_engine.hip).
int begin_call(hipStream_t st);
int end_call(hipStream_t st);

// begin_call on construction (its error in .err), end_call on every return.
struct CallScope {
    hipStream_t st;
    int err;
    explicit CallScope(hipStream_t s) : st(s), err(begin_call(s)) {}
    ~CallScope() { (void)end_call(st); }
};

void fill_zero(unsigned long long *p, size_t n, hipStream_t st);

// Argument checks shared by the entry points (N_c >= 1, tau >= 0, emb >= 1,
// l_min in [0, N_c], distance 1 or 2).
int check_args(size_t n, const int *tau, int ntau, const int *emb, int nemb, const int *lm, int nl, int dist);

constexpr int PAD = 128;   // zero samples after N (kernel windows read up to N + 95)

// The padded device copy of a series (c.x).
template <typename F>
int upload_series(const F *src, size_t n, bool src_on_device, hipStream_t st);

// One (tau, emb) block of the large-N path on the padded series dx.
template <typename F>
int run_block(const F *dx, size_t n, int tau, int emb, const F *thr, int T, const int *lmins, int nl, int dist,
              F *d_rr, F *d_det, F *d_lam, hipStream_t st);

// ---------------------------------------------------------------- thresholds
// Euclidean: largest s >= 0 with (eps - RN(sqrt(s))) >= 0 in F, so that the
// GPU decision [s <= S] equals the reference's (eps - sqrt(s)) >= 0
// (AccRQA_CPU_function.cpp:22,45).  Values that never recur map to -1.
template <typename F>
bool euclid_ok(F eps, F s) {
    F d = std::sqrt(s);
    F diff = eps - d;
    return diff >= F(0);
}

template <typename F>
F euclid_bound(F eps) {
    const F big = std::numeric_limits<F>::max();
    const F inf = std::numeric_limits<F>::infinity();
    if (std::isnan(eps) || eps < F(0)) return F(-1);
    if (eps == inf) return big;
    if (eps == F(0)) return F(0);   // +0 or -0: only s == 0 recurs
    double e2 = (double)eps * (double)eps;
    F s = e2 > (double)big ? big : (F)e2;
    while (s < big && euclid_ok(eps, std::nextafter(s, inf))) s = std::nextafter(s, inf);
    while (s > F(0) && !euclid_ok(eps, s)) s = std::nextafter(s, -inf);
    return s;
}

template <typename F>
F bound_of(F eps, int dist) {
    if (dist == 1) return euclid_bound(eps);
    // maximal norm: R = [(eps - d) >= 0] = [d <= eps] except eps = d = +inf
    if (eps == std::numeric_limits<F>::infinity()) return std::numeric_limits<F>::max();
    return eps;   // NaN compares false, negative eps never recurs
}

// A threshold whose bound is negative or NaN never recurs; it skips the pair
// kernels (its histogram is empty).  -0 behaves as +0.
template <typename F>
bool never_recurs(F b) {
    return std::isnan(b) || b < F(0);
}

// Persistent grid: as many workgroups as can be resident at once (from the
// kernel's register and LDS use, with `dyn` bytes of dynamic LDS per
// workgroup), or fewer if there are fewer items.
template <typename Kern>
unsigned grid_for(Kern kern, int64_t nitems, int ncu, size_t dyn = 0) {
    static std::map<std::pair<const void *, size_t>, int> per_cu;
    const void *key = reinterpret_cast<const void *>(kern);
    auto it = per_cu.find({key, dyn});
    int nb = 0;
    if (it == per_cu.end()) {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, key, rqa::ENG_WG, dyn) != hipSuccess || nb < 1) nb = 1;
        per_cu[{key, dyn}] = nb;
    } else {
        nb = it->second;
    }
    constexpr int waves = rqa::ENG_WG / 64;
    const int64_t wgs = (nitems + waves - 1) / waves;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(wgs, (int64_t)nb * ncu));
}

// A persistent grid rounded down to a multiple of the 8 XCDs (kernels that
// number their waves XCD by XCD).
inline unsigned xcd_round(unsigned g) { return g >= 8 ? g / 8 * 8 : g; }

// Persistent grid for a kernel whose work items belong to whole workgroups.
template <typename Kern>
unsigned grid_for_wg(Kern kern, int64_t nitems, int ncu, size_t dyn = 0) {
    return grid_for(kern, nitems * (rqa::ENG_WG / 64), ncu, dyn);
}

// Dynamic LDS that holds a kernel to at most `per_cu` resident workgroups per
// CU (of four waves, so per_cu waves per SIMD): the workgroup's static LDS
// padded past a (per_cu + 1)-th share of the CU's 160 KB.  0 when the kernel's
// own LDS already keeps it there, or per_cu <= 0 (no cap).
constexpr size_t LDS_PER_CU = 160 * 1024;
template <typename Kern>
size_t lds_cap_pad(Kern kern, int per_cu) {
    if (per_cu <= 0) return 0;
    static std::map<const void *, size_t> stat;
    const void *key = reinterpret_cast<const void *>(kern);
    auto it = stat.find(key);
    size_t st = 0;
    if (it == stat.end()) {
        hipFuncAttributes a{};
        if (hipFuncGetAttributes(&a, key) == hipSuccess) st = a.sharedSizeBytes;
        stat[key] = st;
    } else {
        st = it->second;
    }
    const size_t need = LDS_PER_CU / (size_t)(per_cu + 1) + 1;
    return need > st ? need - st : 0;
}

}  // namespace eng
