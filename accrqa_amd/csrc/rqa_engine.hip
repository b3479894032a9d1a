// rqa_engine.hip -- host side of the gfx950 RR / DET / LAM engine and its C ABI
// (include/accrqa_amd.h).
//
// One call handles one (tau, emb) block for all thresholds: thresholds are
// turned into exact comparison bounds, sorted, split into groups of 4/8/16/32
// (one pass over the pairs per group), and each group runs on one HIP stream
//   RR alone:   k_rr -> k_rr_out
//   LAM:        k_lam | k_lam_mx (column pass; keeps the rank tiles when they
//               fit, and then leaves the column runs to k_lam_vscan in the
//               MFMA form) [-> k_lam_hscan] -> k_stitch -> k_finalize, RR from
//               its histogram -> k_rr_out
//   DET:        k_det_scan (from the tiles) | k_det | k_det_mx -> k_stitch
//               -> k_finalize
// (k_lam_mx / k_det_mx: MFMA Gram form, with k_norms and k_mx_lut first).  Device buffers live in a per-process context that grows
// on demand and is reused across calls (the reference allocates and frees per
// call, src/AccRQA-GPU-HST.cu:1855-1945).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "accrqa_amd.h"
#include "engine.hpp"
#include "rqa_device.hpp"

namespace eng {

std::string g_last_error;

uint64_t g_gen = 0;

std::mutex g_mu;
Ctx g_ctx;

// Pinned host staging for the small uploads of a call (threshold
// permutation, l_min lists, bucket tables, work-item tables): copies from
// pinned memory are asynchronous, so the bytes must stay as they are until
// the copies reading them have run.  The arena is filled front to back and
// only reset once every copy issued from it is known complete: the previous
// call's completion event (which, through the ordering below, follows every
// earlier call) and the current call's stream.
struct Staging {
    char *p = nullptr;
    size_t cap = 0, used = 0;
} g_stage;

// Calls are serialised by g_mu.  Consecutive calls on different streams (the
// host entry points use the library's own stream, the device entry points
// the caller's) are ordered on the device as well: each call's stream waits
// for the completion event of the previous call, whose stream waited for the
// one before, so the work buffers shared by all calls are never written
// while an earlier call may still read them.
struct CallOrder {
    hipEvent_t done = nullptr;
    hipStream_t last = nullptr;
    bool any = false;
} g_order;

// Work-item tables kept for the last few (nc, seg_len): a call with several
// (tau, m) blocks, or a loop over many series of a few lengths, alternates
// between them, and rebuilding costs a host pass and five uploads.  Least
// recently used entry replaced, its new tables uploaded in stream order.
struct ItemSet {
    int64_t nc = -1;
    int L = -1;
    bool slab = false;   // lamv, lamh and dets sorted slab by slab (slab form)
    uint64_t used = 0;
    DevBuf rr, det, dets, lamv, lamh;
    int64_t n_rr = 0, n_det = 0, n_dets = 0, n_lamv = 0, n_lamh = 0;
    std::vector<int64_t> off_lamv, off_lamh, off_dets;   // slab form: first item of each slab (+ end)
};
constexpr int NITEMSETS = 8;
ItemSet g_items[NITEMSETS];
uint64_t g_items_clock = 0;

// Euclidean distance form: 0 = automatic (MFMA Gram + exact recheck for fp32
// input with m >= 8, exact VALU sums otherwise), 1 = exact VALU sums only,
// 2 = MFMA + recheck wherever it applies (fp32 input, Euclidean norm).
int g_form = 0;

// LAM/DET split through kept rank tiles: 0 = when the tiles fit, 1 = never.
int g_split = 0;

// Slab form of the split: 0 = when the tile store of all rows does not fit
// (slabs as tall as fit), > 0 = always, with slabs of this many rows (a
// multiple of 64, at least 256; for tests).
int g_slab_rows = 0;

// Gram operand precision of the split form's LAM rank pass (rqa::GRAM_*):
// GRAM_HILO always, but for the measured comparison of precisions.  Another
// precision applies to the passes that run the split MFMA form (the tile
// store), which then need fp32 input and at most 16 thresholds per block
// (ERR_INVALID_ARGUMENT otherwise; groups of up to 16 run at TG = 16); every
// other pass keeps its usual form.
int g_gram = 0;

// Recurrence plot, exact form: four columns per lane (k_rp4, where its
// stores align) or one (k_rp; accrqa_amd_set_rp_wide(0)).
bool g_rp_wide = true;

// Recurrence plot of bytes: the upper triangle, each 64 x 64 square stored
// mirrored as well (k_rp_mx_sym, k_rp_sym; where N_c % 16 == 0 and the plot
// is 16-byte aligned), or every element (k_rp_mx, k_rp4, k_rp;
// accrqa_amd_set_rp_sym(0)).
bool g_rp_sym = true;

// The split form's DET tile scan on the context's second stream, beside the
// LAM column and row scans on the call's stream (1), or after them (0).  Off
// while kernels are profiled one by one.
int g_overlap = 1;

// Resident workgroups (four waves each, so waves per SIMD) of the three tile
// scans, k_lam_vscan, k_lam_hscan and k_det_scan: their LDS is padded with
// dynamic LDS until no more fit per CU; 0 = as many as their registers and
// LDS allow.  Seven measured faster than eight on the scans' VALU, LDS-atomic
// and branch mix (accrqa_amd_set_scan_waves).
int g_scan_waves = 7;

// MFMA-form decision statistics (diagnostic): when on, the split form's rank
// kernel counts elements ranked, flagged by the table and re-decided in exact
// order into the context's mxstats buffer (3 u64).
bool g_mxstats = false;

// MFMA form for F = float and, in the LAM pass, F = double (hi/lo Gram of the
// series rounded to fp32, exact re-decision in fp64); the direct DET kernel is fp32-only.
template <typename F>
bool use_mfma(int dist, int m, int tau) {
    if (dist != 1) return false;
    if ((int64_t)(m - 1) * tau > rqa::MXW - 32) return false;   // operand window must fit in LDS
    if (g_form == 1) return false;
    if (g_form == 2) return true;
    return m >= 8;
}

int map_err(hipError_t e) {
    if (e == hipSuccess) return E_OK;
    g_last_error = hipGetErrorString(e);
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return E_NO_MEMORY;
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return E_NO_DEVICE;
    return E_DEVICE;
}

int ensure(DevBuf &b, size_t bytes) {
    if (bytes == 0) bytes = 256;
    if (b.cap >= bytes) return E_OK;
    if (b.p) {
        // kernels queued by this call or an earlier one, on any stream, may
        // still read it (a rare path: buffers only grow)
        HIP_TRY(hipDeviceSynchronize());
        hipError_t e = hipFree(b.p);
        b.p = nullptr;
        b.cap = 0;
        if (e != hipSuccess) return map_err(e);
    }
    HIP_TRY(hipMalloc(&b.p, bytes));
    b.cap = bytes;
    b.gen = ++g_gen;
    return E_OK;
}

// A small read-only device table uploaded only when its bytes change: the
// (tau, emb) blocks of one call, and the calls of a loop over series, reuse
// the same threshold permutation, l_min list and bucket tables.  Kept per
// allocation, so a reallocated buffer is always filled again.
struct Uploaded {
    uint64_t gen = 0;
    std::vector<char> bytes;
};
Uploaded g_up_perm, g_up_lmins, g_up_lut;

// Copies `bytes` from host memory to device memory `dst` in stream order,
// through the pinned staging arena (the source may be released on return).
int stage_upload(void *dst, const void *src, size_t bytes, hipStream_t st) {
    if (!bytes) return E_OK;
    const size_t need = (bytes + 255) & ~(size_t)255;
    Staging &S = g_stage;
    if (S.used + need > S.cap) {
        // every copy issued from the arena must have run before it is reused
        if (g_order.any) HIP_TRY(hipEventSynchronize(g_order.done));
        HIP_TRY(hipStreamSynchronize(st));
        S.used = 0;
        if (need > S.cap) {
            if (S.p) HIP_TRY(hipHostFree(S.p));
            S.p = nullptr;
            S.cap = 0;
            const size_t cap = std::max<size_t>(need, (size_t)8 << 20);
            HIP_TRY(hipHostMalloc((void **)&S.p, cap, hipHostMallocDefault));
            S.cap = cap;
        }
    }
    char *h = S.p + S.used;
    std::memcpy(h, src, bytes);
    S.used += need;
    HIP_TRY(hipMemcpyAsync(dst, h, bytes, hipMemcpyHostToDevice, st));
    return E_OK;
}

int upload_once(const DevBuf &b, Uploaded &u, const void *src, size_t bytes, hipStream_t st) {
    if (u.gen == b.gen && u.bytes.size() == bytes && std::memcmp(u.bytes.data(), src, bytes) == 0) return E_OK;
    int e = stage_upload(b.p, src, bytes, st);
    if (e) return e;
    u.gen = b.gen;
    u.bytes.assign((const char *)src, (const char *)src + bytes);
    return E_OK;
}

// Start and end of every entry point that queues device work on `st`.
int begin_call(hipStream_t st) {
    if (!g_order.done) HIP_TRY(hipEventCreateWithFlags(&g_order.done, hipEventDisableTiming));
    if (g_order.any && g_order.last != st) HIP_TRY(hipStreamWaitEvent(st, g_order.done, 0));
    return E_OK;
}

int end_call(hipStream_t st) {
    HIP_TRY(hipEventRecord(g_order.done, st));
    g_order.last = st;
    g_order.any = true;
    return E_OK;
}

int init_ctx() {
    int dev = -1;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        g_last_error = "no HIP device";
        return E_NO_DEVICE;
    }
    HIP_TRY(hipGetDevice(&dev));
    if (g_ctx.device != dev) {
        // A different current device: start a fresh context on it.  Buffers of
        // the old device are released on that device.
        if (g_ctx.device >= 0) {
            int cur = dev;
            (void)hipSetDevice(g_ctx.device);
            DevBuf *all[] = {&g_ctx.x,       &g_ctx.hist,       &g_ctx.carry,      &g_ctx.metric,
                             &g_ctx.scan,    &g_ctx.counts,     &g_ctx.perm,       &g_ctx.lmins,
                             &g_ctx.lut,     &g_ctx.nrm,        &g_ctx.nmax,       &g_ctx.lut8,
                             &g_ctx.out_rr,  &g_ctx.out_det,    &g_ctx.out_lam,    &g_ctx.priv,
                             &g_ctx.tiles,   &g_ctx.dump,       &g_ctx.tops,
                             &g_ctx.mxstats, &g_ctx.probe, &g_ctx.out_rp, &g_ctx.bhist, &g_ctx.bdesc,
                             &g_ctx.hist2,   &g_ctx.carry2,     &g_ctx.priv2};
            for (DevBuf *b : all) {
                if (b->p) (void)hipFree(b->p);
                *b = DevBuf();
            }
            for (ItemSet &s : g_items) {
                for (DevBuf *b : {&s.rr, &s.det, &s.dets, &s.lamv, &s.lamh})
                    if (b->p) (void)hipFree(b->p);
                s = ItemSet();
            }
            g_ctx.items_rr = g_ctx.items_det = g_ctx.items_dets = g_ctx.items_lamv = g_ctx.items_lamh = DevBuf();
            if (g_order.done) (void)hipEventSynchronize(g_order.done);
            if (g_order.done) (void)hipEventDestroy(g_order.done);
            g_order = CallOrder();
            if (g_ctx.stream) (void)hipStreamDestroy(g_ctx.stream);
            if (g_ctx.side) (void)hipStreamDestroy(g_ctx.side);
            if (g_ctx.ev_tiles) (void)hipEventDestroy(g_ctx.ev_tiles);
            if (g_ctx.ev_det) (void)hipEventDestroy(g_ctx.ev_det);
            g_ctx.stream = g_ctx.side = nullptr;
            g_ctx.ev_tiles = g_ctx.ev_det = nullptr;
            (void)hipSetDevice(cur);
        }
        hipDeviceProp_t prop;
        HIP_TRY(hipGetDeviceProperties(&prop, dev));
        g_ctx.ncu = prop.multiProcessorCount;
        HIP_TRY(hipStreamCreateWithFlags(&g_ctx.stream, hipStreamNonBlocking));
        HIP_TRY(hipStreamCreateWithFlags(&g_ctx.side, hipStreamNonBlocking));
        HIP_TRY(hipEventCreateWithFlags(&g_ctx.ev_tiles, hipEventDisableTiming));
        HIP_TRY(hipEventCreateWithFlags(&g_ctx.ev_det, hipEventDisableTiming));
        g_ctx.device = dev;
    }
    return E_OK;
}

int tg_for(int nt) { return nt <= 4 ? 4 : (nt <= 8 ? 8 : (nt <= 16 ? 16 : 32)); }

inline uint32_t hkey(float v) {
    uint32_t b;
    std::memcpy(&b, &v, 4);
    return b;
}
inline unsigned long long hkey(double v) {
    unsigned long long b;
    std::memcpy(&b, &v, 8);
    return b;
}
inline uint32_t hhi(uint32_t k) { return k; }
inline uint32_t hhi(unsigned long long k) { return (uint32_t)(k >> 32); }
inline unsigned long long key_at(float, unsigned long long hi) { return hi; }
inline unsigned long long key_at(double, unsigned long long hi) {
    return hi >= (1ull << 32) ? ~0ull : (hi << 32);
}

// Bucket table for the rank of a pair (see rank_of in rqa_device.hpp).  b:
// the group's bounds, ascending, non-negative and finite.  Bucket bk covers
// high key bits [base + bk<<sh, base + (bk+1)<<sh); the last bucket lies past
// the largest bound, so keys clamped into it get rank nt.
template <typename F>
void build_lut(const F *b, int nt, rqa::LutEntry<F> *out, rqa::LutParams &P) {
    std::vector<unsigned long long> key(nt);
    for (int t = 0; t < nt; t++) key[t] = hkey(b[t]);
    const uint32_t base = hhi(decltype(hkey(F(0)))(key[0]));
    const uint32_t range = hhi(decltype(hkey(F(0)))(key[nt - 1])) - base;
    int sh = 0;
    while ((range >> sh) > (uint32_t)(rqa::NB_MAX - 2)) sh++;
    const int nb = (int)(range >> sh) + 2;
    for (int bk = 0; bk < nb; bk++) {
        const unsigned long long s_hi = (unsigned long long)base + ((unsigned long long)bk << sh);
        const unsigned long long start = key_at(F(0), s_hi);
        const unsigned long long end = key_at(F(0), s_hi + (1ull << sh));
        int r_lo = 0, inside = 0;
        for (int t = 0; t < nt; t++) {
            if (key[t] < start) r_lo++;
            else if (key[t] < end) inside++;
        }
        out[bk].next = r_lo < nt ? (decltype(out[bk].next))key[r_lo] : (decltype(out[bk].next))~0ull;
        out[bk].info = (uint32_t)r_lo | (inside >= 2 ? 0x100u : 0u);
    }
    P.base = base;
    P.sh = sh;
    P.nb = nb;
}

// ---------------------------------------------------------------- launches
__global__ void k_fill_u64(unsigned long long *p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = 0ull;
}

template <typename F>
__global__ void k_pad_copy(const F *__restrict__ src, F *__restrict__ dst, size_t n, size_t ntot) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < ntot; i += stride) dst[i] = i < n ? src[i] : F(0);
}

void fill_zero(unsigned long long *p, size_t n, hipStream_t st) {
    if (!n) return;
    unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill_u64, dim3(blocks), dim3(256), 0, st, p, n);
}

// Row t of a group's raw results (hs words) -> out + perm[t] * hs.
__global__ void k_dump_rows(const unsigned long long *__restrict__ src, size_t hs, const int *__restrict__ perm,
                            unsigned long long *__restrict__ out) {
    const int t = blockIdx.y;
    const unsigned long long *s = src + (size_t)t * hs;
    unsigned long long *o = out + (size_t)perm[t] * hs;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < hs; i += (size_t)gridDim.x * blockDim.x)
        o[i] = s[i];
}

// k_hist_prep (DET correction, highest occupied bin) then k_finalize (suffix
// sums, Lmax, ENTR, the five measures) for the nt histograms of a group.
template <typename F>
int finalize(unsigned long long *hist, int64_t nc, int nt, int is_det, const int *d_lmins, int nl, F *out,
             const int *d_perm, int T, unsigned long long *m0_out, hipStream_t st) {
    Ctx &c = g_ctx;
    int e;
    if ((e = ensure(c.tops, std::max<size_t>(64, (size_t)nt) * sizeof(unsigned long long)))) return e;
    unsigned long long *tops = (unsigned long long *)c.tops.p;
    fill_zero(tops, (size_t)nt, st);
    const int64_t hs = nc + 1;
    const unsigned bx = (unsigned)std::min<int64_t>((hs + rqa::HPT - 1) / rqa::HPT, 64);
    hipLaunchKernelGGL(rqa::k_hist_prep, dim3(bx, nt), dim3(rqa::HPT), 0, st, hist, nc, is_det, tops);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL((rqa::k_finalize<F>), dim3(nt), dim3(rqa::FIN), 0, st, hist, nc, nt,
                       (const unsigned long long *)tops, d_lmins, nl, (unsigned long long *)c.metric.p,
                       (unsigned long long *)c.scan.p, out, is_det, d_perm, T, m0_out);
    HIP_TRY(hipGetLastError());
    return E_OK;
}

int dump_rows(const unsigned long long *src, int nt, size_t hs, const int *d_perm, unsigned long long *out,
              hipStream_t st) {
    if (!out || nt <= 0) return E_OK;
    const unsigned bx = (unsigned)std::min<size_t>((hs + 255) / 256, 1024);
    hipLaunchKernelGGL(k_dump_rows, dim3(bx, nt), dim3(256), 0, st, src, hs, d_perm, out);
    HIP_TRY(hipGetLastError());
    return E_OK;
}

// ---------------------------------------------------------------- profiling
// Optional per-kernel timing with HIP events recorded on the launch stream
// (accrqa_amd_profile_*).  Off by default.  The event pairs are folded into
// per-kernel totals when the totals are read (or when many are pending), so a
// timed loop of calls is not synchronised per call.
struct ProfEntry {
    const char *name;
    hipEvent_t a, b;
};
struct Prof {
    bool on = false;
    std::vector<ProfEntry> pending;
    std::vector<hipEvent_t> pool;
    std::vector<std::pair<std::string, std::pair<double, long long>>> acc;
    const char *cur = nullptr;
    hipEvent_t cur_a = nullptr;
} g_prof;

hipEvent_t prof_event() {
    if (!g_prof.pool.empty()) {
        hipEvent_t e = g_prof.pool.back();
        g_prof.pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

void prof_begin(const char *name, hipStream_t st) {
    if (!g_prof.on) return;
    g_prof.cur = name;
    g_prof.cur_a = prof_event();
    (void)hipEventRecord(g_prof.cur_a, st);
}

void prof_end(hipStream_t st) {
    if (!g_prof.on || !g_prof.cur) return;
    hipEvent_t b = prof_event();
    (void)hipEventRecord(b, st);
    g_prof.pending.push_back({g_prof.cur, g_prof.cur_a, b});
    g_prof.cur = nullptr;
}

// Waits for the pending event pairs and adds them to the totals.
void prof_fold() {
    for (ProfEntry &p : g_prof.pending) {
        float ms = 0.f;
        (void)hipEventSynchronize(p.b);
        (void)hipEventElapsedTime(&ms, p.a, p.b);
        bool found = false;
        for (auto &kv : g_prof.acc)
            if (kv.first == p.name) {
                kv.second.first += ms;
                kv.second.second += 1;
                found = true;
            }
        if (!found) g_prof.acc.push_back({p.name, {(double)ms, 1}});
        g_prof.pool.push_back(p.a);
        g_prof.pool.push_back(p.b);
    }
    g_prof.pending.clear();
}

// End of a call: fold only when many pairs are pending (bounded event use).
void prof_collect(hipStream_t) {
    if (g_prof.pending.size() > 16384) prof_fold();
}

// Segment length: enough full-column (band, segment) items for about sixteen
// per resident wave of a full GPU (65536 items; the split LAM and DET passes
// have items of very different sizes, so they get many), in multiples of 32
// steps, at least 256.
// Every segment boundary costs a carry slot per (threshold, sequence) and a
// join in k_stitch, so segments stay as long as the parallelism allows.  At
// most 2^29 steps: a run counter (the open-prefix flag, bit 30, plus at most
// a segment and a word) then stays below 2^31, as the line scan's signed test
// needs.
int seg_len_for(int64_t nc) {
    const int64_t nbands = (nc + rqa::WAVE - 1) / rqa::WAVE;
    const int64_t nseg = std::max<int64_t>(1, (65536 + nbands - 1) / nbands);
    int64_t L = (nc + nseg - 1) / nseg;
    L = std::max<int64_t>(L, 256);
    L = (L + rqa::BLK - 1) / rqa::BLK * rqa::BLK;
    return (int)std::min<int64_t>(L, 1 << 29);
}

// Workgroup-private long-run bins for a launch of `grid` workgroups.
template <int TG>
int ensure_priv(unsigned grid) {
    return ensure(g_ctx.priv, (size_t)grid * TG * rqa::PBINS * sizeof(uint32_t));
}

// Bytes of the rank tiles the LAM pass keeps for k_det_scan: row blocks
// I < ceil(nc/32), bands J >= I/2 (see rqa::tile_index).
size_t tile_store_bytes(int64_t nc) {
    const int64_t nrb = (nc + rqa::BLK - 1) / rqa::BLK;
    const int64_t nbands = (nc + rqa::WAVE - 1) / rqa::WAVE;
    const int64_t ntiles = nrb * nbands - (nrb / 2) * ((nrb - 1) / 2);
    return (size_t)ntiles * rqa::TILE_BYTES;
}

// The tile store is used when it fits in well under half of the free device
// memory; otherwise DET recomputes its pairs (k_det / k_det_mx).
bool tile_store_fits(int64_t nc) {
    const size_t need = tile_store_bytes(nc);
    if (g_ctx.tiles.cap >= need) return true;
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return false;
    return need + g_ctx.tiles.cap < fr / 2;
}

// First tile of row block I in the packed store (rqa::tile_index(I, I/2)).
int64_t tile_start(int64_t I, int64_t nbands) { return I * nbands - (I / 2) * ((I - 1) / 2); }

// Bytes of the largest slab of L rows (the first: row block I holds the
// tiles of bands >= I/2, so earlier slabs hold more).
size_t slab_bytes(int64_t nc, int64_t L) {
    const int64_t nrb = (nc + rqa::BLK - 1) / rqa::BLK;
    const int64_t nbands = (nc + rqa::WAVE - 1) / rqa::WAVE;
    const int64_t ib = std::min<int64_t>(L / rqa::BLK, nrb);
    return (size_t)(tile_start(ib, nbands) - tile_start(0, nbands)) * rqa::TILE_BYTES;
}

// Slab height of the slab form: the tallest multiple of 64 rows (at least
// 4,096, at most the default segment length) whose largest slab fits in under
// half of the free device memory; 0 if none does.
int slab_len_for(int64_t nc) {
    if (g_slab_rows > 0) return g_slab_rows;
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return 0;
    const size_t budget = fr / 2 + g_ctx.tiles.cap;
    for (int64_t L = seg_len_for(nc) / rqa::WAVE * rqa::WAVE; L >= 4096; L = L / 2 / rqa::WAVE * rqa::WAVE)
        if (slab_bytes(nc, L) < budget) return (int)L;
    return 0;
}

unsigned stitch_blocks(int64_t nseq) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((nseq + rqa::STT - 1) / rqa::STT, 64));
}

// Work-item tables (uploaded once per (N_c, segment length)):
//   k_rr : (band b, row segment q) over rows 0 .. min(64(b+1), N_c) - 1
//   k_det: (strip b, step segment q) over steps 0 .. N_c - (1 + 64b) - 1
void select_items(const ItemSet &s) {
    Ctx &c = g_ctx;
    c.items_rr = s.rr;
    c.items_det = s.det;
    c.items_dets = s.dets;
    c.items_lamv = s.lamv;
    c.items_lamh = s.lamh;
    c.n_items_rr = s.n_rr;
    c.n_items_det = s.n_det;
    c.n_items_dets = s.n_dets;
    c.n_items_lamv = s.n_lamv;
    c.n_items_lamh = s.n_lamh;
    c.nslabs = s.slab ? (int)s.off_lamv.size() - 1 : 0;
    c.slab_lamv = s.slab ? s.off_lamv.data() : nullptr;
    c.slab_lamh = s.slab ? s.off_lamh.data() : nullptr;
    c.slab_dets = s.slab ? s.off_dets.data() : nullptr;
}

// Work-item tables of segment length L.  Slab form: the split passes' tables
// (lamv, lamh, dets) are ordered slab by slab, slab s being the rows
// [sL, (s+1)L) -- the row segment q = s of a column band, the row band
// 64b / L of a row part, the step segment q = s of a strip group -- and by
// decreasing size within a slab.
int build_items(int64_t nc, int L, bool slab, hipStream_t st) {
    ItemSet *slot = &g_items[0];
    for (ItemSet &s : g_items) {
        if (s.nc == nc && s.L == L && s.slab == slab) {
            s.used = ++g_items_clock;
            select_items(s);
            return E_OK;
        }
        if (s.used < slot->used) slot = &s;
    }
    ItemSet &S = *slot;
    // (item, steps) lists; each table is sorted by decreasing size so that the
    // persistent waves take the long items first
    std::vector<std::pair<int64_t, int2>> rr, det, dets, lamv, lamh;
    const int64_t nb = (nc + rqa::WAVE - 1) / rqa::WAVE;
    auto seg = [&](std::vector<std::pair<int64_t, int2>> &v, int64_t b, int64_t len) {
        for (int64_t q = 0; q * L < len; q++)
            v.push_back({std::min<int64_t>(L, len - q * L), make_int2((int)b, (int)q)});
    };
    for (int64_t b = 0; b < nb; b++) {
        const int64_t V = std::min<int64_t>((b + 1) * rqa::WAVE, nc);
        seg(rr, b, V);          // rows that can be <= a column of the band
        seg(lamv, b, V);        // split LAM: rows [0, V) of the band's columns
        seg(lamh, b, nc - V);   // split LAM: columns [V, nc) of the band's rows
    }
    const int64_t nstrips = (nc - 1 + rqa::WAVE - 1) / rqa::WAVE;
    for (int64_t b = 0; b < nstrips; b++) seg(det, b, nc - (1 + b * rqa::WAVE));
    // k_det_scan: groups of WAVES strips, segments of the group's longest strip
    const int64_t ngroups = (nc - 1 + rqa::WG - 1) / rqa::WG;
    for (int64_t g = 0; g < ngroups; g++) seg(dets, g, nc - (1 + g * rqa::WG));
    int e;
    const int64_t nsl = (nc + L - 1) / L;
    // slab of an item: -1 = not a slab-ordered table
    auto upload = [&](std::vector<std::pair<int64_t, int2>> &v, DevBuf &buf, int64_t &n, int kind = -1,
                      std::vector<int64_t> *off = nullptr) -> int {
        auto slab_of = [&](const int2 &x) -> int64_t {
            return kind == 1 ? ((int64_t)x.x * rqa::WAVE) / L : (int64_t)x.y;   // lamh: the row band's slab
        };
        if (off) {
            std::stable_sort(v.begin(), v.end(), [&](const auto &x, const auto &y) {
                const int64_t sx = slab_of(x.second), sy = slab_of(y.second);
                return sx != sy ? sx < sy : x.first > y.first;
            });
            off->assign((size_t)nsl + 1, 0);
            for (const auto &x : v) (*off)[(size_t)slab_of(x.second) + 1]++;
            for (int64_t k = 0; k < nsl; k++) (*off)[(size_t)k + 1] += (*off)[(size_t)k];
        } else {
            std::stable_sort(v.begin(), v.end(), [](const auto &x, const auto &y) { return x.first > y.first; });
        }
        std::vector<int2> it(v.size());
        for (size_t i = 0; i < v.size(); i++) it[i] = v[i].second;
        int e2;
        if ((e2 = ensure(buf, std::max<size_t>(it.size(), 1) * sizeof(int2)))) return e2;
        // ordered on the call's stream: after the queued kernels that may
        // still read the entry's previous tables, before those of this block
        if ((e2 = stage_upload(buf.p, it.data(), it.size() * sizeof(int2), st))) return e2;
        n = (int64_t)it.size();
        return E_OK;
    };
    S.nc = -1;
    if ((e = upload(rr, S.rr, S.n_rr))) return e;
    if ((e = upload(det, S.det, S.n_det))) return e;
    if ((e = upload(dets, S.dets, S.n_dets, 0, slab ? &S.off_dets : nullptr))) return e;
    if ((e = upload(lamv, S.lamv, S.n_lamv, 0, slab ? &S.off_lamv : nullptr))) return e;
    if ((e = upload(lamh, S.lamh, S.n_lamh, 1, slab ? &S.off_lamh : nullptr))) return e;
    if (!slab) S.off_dets.clear(), S.off_lamv.clear(), S.off_lamh.clear();
    S.nc = nc;
    S.L = L;
    S.slab = slab;
    S.used = ++g_items_clock;
    select_items(S);
    return E_OK;
}

// Parameters of the MFMA form's byte table for sorted bounds b[0..nt).
// fp64 input: the Gram's inputs and the table's bounds are rounded to fp32;
// extra_u units of 2^-24 (n_i + n_j) cover that (see the fp64 note at
// rqa::k_lam_mx).  gram: the Gram's operand precision (rqa::GRAM_*).  With
// operands rounded as |a - q(a)| <= eps |a| + eta, the dot product moves by at
// most sum (2 eps + eps^2)|a||b| + eta (1 + eps)(|a| + |b|) + eta^2, so s~ by
// R (n_i + n_j) + D with R = 2 eps + eps^2 + 4 eta (1 + eps) and
// D = eta (1 + eps) m / 2 + 2 m eta^2 (scaled units; sum |a| <= 2A + m/8):
//   GRAM_F16  eps = 2^-11, eta = 2^-25: R < 16392 u, D < m 2^-25;
//   GRAM_BF16 eps = 2^-8, eta = 2^-126 (a flushed fp32 subnormal):
//             R < 131336 u, D < m 2^-25;
//   GRAM_FP8  eps = 2^-3, eta = 2^-6 (|a| < 2^8): R <= 5636096 u, D < m 2^-6;
// added to the hi/lo form's coefficient, whose split term they replace (the
// rest of it, over 2m + 30 u, also covers the accumulation's growth with the
// rounded operands, at most (1 + eps)^2 - 1 of its 33C u).  The FP8 matrix
// instruction accumulates more coarsely than the 16-bit ones: measured within
// 5400 u (|c| + sum |p|) per instruction, taken as 2^-9 (|c| + sum |p|), i.e.
// at most 2^-9 C (1 + eps)^2 (A + B) < 90000 C u (A + B) more in s~.
// The float nearest above v (v itself when it is exact).
static float round_up_f(double v) {
    float f = (float)v;
    if ((double)f < v) f = std::nextafter(f, std::numeric_limits<float>::infinity());
    return f;
}

rqa::MxParams mx_params(const float *b, int nt, int m, int nbmax, int extra_u = 0, int gram = rqa::GRAM_HILO) {
    rqa::MxParams P;
    const float lo = b[0] * 0.5f;
    float hi = b[nt - 1] * 2.0f;
    if (!(hi <= std::numeric_limits<float>::max())) hi = std::numeric_limits<float>::max();
    const uint32_t base = hkey(lo), top = hkey(hi);
    int sh = 0;
    while (((top - base) >> sh) > (uint32_t)(nbmax - 8)) sh++;
    // bucket edges at multiples of 2^sh: the device finds a bucket as key >> sh
    const uint32_t base_al = base & ~((1u << sh) - 1u);
    int nb = (int)((top - base_al) >> sh) + 2;
    nb = (nb + 3) & ~3;
    P.base = base_al;
    P.sh = sh;
    P.nb = nb;
    P.qlo = (int)(base_al >> sh);
    const int C = (m + 15) / 16;   // chunks of 16 dimensions in the FP16 hi/lo Gram
    const double gram_u = gram == rqa::GRAM_F16 ? 16392.0
                          : gram == rqa::GRAM_BF16 ? 131336.0
                          : gram == rqa::GRAM_FP8  ? 5636100.0 + 90000.0 * C
                                                   : 0.0;
    // Both coefficients carry a factor 1 + 2^-12 and are rounded up, which covers
    // the rounding of E = fma(cE, n_i + n_j, dE) itself on the device.
    P.cE = round_up_f((3.0 * m + 31.0 + 33.0 * C + extra_u + gram_u) * std::ldexp(1.0, -24) * (1.0 + std::ldexp(1.0, -12)));
    P.dA = round_up_f(m * std::ldexp(1.0, gram == rqa::GRAM_FP8 ? -6 : -25) * (1.0 + std::ldexp(1.0, -12)));
    P.amax = gram == rqa::GRAM_FP8 ? 8 : 14;
    return P;
}

// The split form's LAM rank pass with another Gram precision (the measured
// comparison of precisions; fp32 input, 16 thresholds), with the decision
// counters when they are on.
template <int TG, int G, bool STATS>
void launch_gram_g(const float *dx, int64_t nc, int tau, int m, const rqa::Bounds<float, TG> &Bf, int nt,
                   const rqa::MxParams &MP, int L, const rqa::LamWork &LW, unsigned long long *hist, uint2 *carry,
                   uint8_t *tiles, hipStream_t st) {
    Ctx &c = g_ctx;
    const unsigned g = grid_for(rqa::k_lam_mx<float, TG, false, STATS, G>, LW.nitems, c.ncu);
    hipLaunchKernelGGL((rqa::k_lam_mx<float, TG, false, STATS, G>), dim3(g), dim3(rqa::WG), 0, st, dx,
                       (const float *)c.nrm.p, (const unsigned int *)c.nmax.p, nc, tau, m, Bf, Bf, nt,
                       (const uint8_t *)c.lut8.p, MP, L, LW, hist, carry, (uint32_t *)c.priv.p, tiles,
                       STATS ? (unsigned long long *)c.mxstats.p : (unsigned long long *)nullptr);
}

template <int TG>
int launch_gram_variant(int gram, const float *dx, int64_t nc, int tau, int m, const rqa::Bounds<float, TG> &Bf,
                        int nt, const rqa::MxParams &MP, int L, const rqa::LamWork &LW, unsigned long long *hist,
                        uint2 *carry, uint8_t *tiles, hipStream_t st) {
    int e;
    if (g_mxstats && (e = ensure(g_ctx.mxstats, 3 * sizeof(unsigned long long)))) return e;
    auto go = [&](auto gk, auto sk) {
        launch_gram_g<TG, decltype(gk)::value, decltype(sk)::value>(dx, nc, tau, m, Bf, nt, MP, L, LW, hist, carry,
                                                                   tiles, st);
    };
    using F16 = std::integral_constant<int, rqa::GRAM_F16>;
    using BF16 = std::integral_constant<int, rqa::GRAM_BF16>;
    using FP8 = std::integral_constant<int, rqa::GRAM_FP8>;
    using On = std::true_type;
    using Off = std::false_type;
    switch (gram) {
        case rqa::GRAM_F16: g_mxstats ? go(F16{}, On{}) : go(F16{}, Off{}); break;
        case rqa::GRAM_BF16: g_mxstats ? go(BF16{}, On{}) : go(BF16{}, Off{}); break;
        default: g_mxstats ? go(FP8{}, On{}) : go(FP8{}, Off{}); break;
    }
    HIP_TRY(hipGetLastError());
    return E_OK;
}

// The split form slab by slab, when the tile store of all rows does not fit:
// slab s is the rows [sL, (s+1)L), i.e. the row segment q = s of every
// column band (k_lam_mx + k_lam_vscan, or k_lam), the row bands 64b / L = s
// (k_lam_hscan) and the step segment q = s of every strip group
// (k_det_scan).  Each of them reads and writes only the tiles of the slab's
// row blocks, so one slab's tiles at a time suffice: the tile pointer handed
// to the kernels is moved back by the slab's first tile index.  Segment
// boundaries already fall on slab boundaries, so the carries and k_stitch
// are those of the whole-store form; DET runs slab by slab beside LAM, in its
// own histogram and carry buffers, since the tiles are gone once the slabs
// are done.
bool overlap_on() { return g_overlap != 0 && !g_prof.on; }

// k_det_scan over DET items [d0, d1) into hist2 / carry2 on the context's
// second stream, after everything `st` holds so far (the tiles it scans),
// with its own private bins; ev_det marks its end.  The LAM scans that follow
// on `st` write only hist and carry, so the two run side by side.
template <int TG>
int det_scan_beside(const uint8_t *tiles, int64_t nbands, int64_t nc, int nt, int L, int nseg, int64_t d0,
                    int64_t d1, unsigned long long *hist2, uint2 *carry2, hipStream_t st) {
    Ctx &c = g_ctx;
    int e;
    const size_t pd = lds_cap_pad(rqa::k_det_scan<TG>, g_scan_waves);
    const unsigned g = grid_for_wg(rqa::k_det_scan<TG>, d1 - d0, c.ncu, pd);
    if ((e = ensure(c.priv2, (size_t)g * TG * rqa::PBINS * sizeof(uint32_t)))) return e;
    HIP_TRY(hipEventRecord(c.ev_tiles, st));
    HIP_TRY(hipStreamWaitEvent(c.side, c.ev_tiles, 0));
    hipLaunchKernelGGL((rqa::k_det_scan<TG>), dim3(g), dim3(rqa::WG), pd, c.side, tiles, nbands, nc, nt, L, nseg,
                       (const int2 *)c.items_dets.p + d0, d1 - d0, hist2, carry2, (uint32_t *)c.priv2.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(c.ev_det, c.side));
    return E_OK;
}

template <typename F, int NORM, int TG>
int launch_group_slabs(const F *dx, int64_t nc, int tau, int m, const rqa::Bounds<F, TG> &B, int nt,
                       const rqa::LutEntry<F> *d_lut, const rqa::LutParams &P, bool want_rr, bool want_det,
                       const int *d_lmins, int nl, const int *d_perm, int T, F *d_rr, F *d_det, F *d_lam,
                       hipStream_t st) {
    Ctx &c = g_ctx;
    if (g_gram != rqa::GRAM_HILO) {
        g_last_error = "the Gram precision comparison needs the tile store of all rows";
        return E_INVALID;
    }
    const int L = c.blk_L;
    const int nseg = (int)((nc + L - 1) / L);
    const size_t hs = (size_t)nc + 1;
    const int64_t nbands = (nc + rqa::WAVE - 1) / rqa::WAVE;
    int e;
    unsigned long long *counts = (unsigned long long *)c.counts.p;
    if ((e = ensure(c.tiles, slab_bytes(nc, L)))) return e;
    if ((e = ensure(c.hist, (size_t)nt * hs * 8))) return e;
    if ((e = ensure(c.metric, (size_t)nt * hs * 8))) return e;
    if ((e = ensure(c.scan, (size_t)nt * hs * 8))) return e;
    if ((e = ensure(c.carry, (size_t)nt * (size_t)nc * 2 * nseg * sizeof(uint2)))) return e;
    if (want_det) {
        if ((e = ensure(c.hist2, (size_t)nt * hs * 8))) return e;
        if ((e = ensure(c.carry2, (size_t)nt * (size_t)(nc - 1) * nseg * sizeof(uint2)))) return e;
    }
    unsigned long long *hist = (unsigned long long *)c.hist.p, *hist2 = (unsigned long long *)c.hist2.p;
    uint2 *carry = (uint2 *)c.carry.p, *carry2 = (uint2 *)c.carry2.p;
    fill_zero(hist, (size_t)nt * hs, st);
    if (want_det) fill_zero(hist2, (size_t)nt * hs, st);
    const bool mx = NORM == 1 && use_mfma<F>(1, m, tau);
    rqa::Bounds<float, TG> Bf;
    for (int t = 0; t < TG; t++) Bf.s[t] = (float)B.s[t];
    rqa::MxParams MP{};
    if (mx) {
        MP = mx_params(Bf.s, nt, m, rqa::NBX_LAM, sizeof(F) == 8 ? 4 : 0);
        if ((e = ensure(c.lut8, rqa::NBX_LAM))) return e;
        hipLaunchKernelGGL((rqa::k_mx_lut<TG>), dim3((MP.nb + 255) / 256), dim3(256), 0, st, Bf, nt, MP,
                           (const unsigned int *)c.nmax.p, (uint8_t *)c.lut8.p);
        HIP_TRY(hipGetLastError());
    }
    // the DET scan of slab s runs beside its LAM scans; slab s + 1's rank pass
    // rewrites the tiles, so it waits for that scan
    const bool beside = want_det && overlap_on();
    bool det_pending = false;
    for (int s = 0; s < c.nslabs; s++) {
        const int64_t ia = (int64_t)s * L / rqa::BLK;
        if (det_pending) {
            HIP_TRY(hipStreamWaitEvent(st, c.ev_det, 0));
            det_pending = false;
        }
        // the kernels index tiles from the store's start; this slab's block ia is the buffer's first
        uint8_t *tiles = (uint8_t *)((uintptr_t)c.tiles.p - (uintptr_t)(tile_start(ia, nbands) * rqa::TILE_BYTES));
        const int64_t v0 = c.slab_lamv[s], v1 = c.slab_lamv[s + 1];
        const rqa::LamWork LW{(const int2 *)c.items_lamv.p + v0, v1 - v0, nseg, 2 * nseg};
        const int64_t d0 = c.slab_dets[s], d1 = c.slab_dets[s + 1];
        auto fork_det = [&]() -> int {
            if (!beside || det_pending || d1 <= d0) return E_OK;
            det_pending = true;
            return det_scan_beside<TG>(tiles, nbands, nc, nt, L, nseg, d0, d1, hist2, carry2, st);
        };
        if (LW.nitems > 0) {
            bool done = false;
            if constexpr (NORM == 1) {
                if (mx) {
                    const unsigned g = grid_for(rqa::k_lam_mx<F, TG, false>, LW.nitems, c.ncu);
                    prof_begin("k_lam_mx", st);
                    hipLaunchKernelGGL((rqa::k_lam_mx<F, TG, false>), dim3(g), dim3(rqa::WG), 0, st, dx,
                                       (const float *)c.nrm.p, (const unsigned int *)c.nmax.p, nc, tau, m, Bf, B, nt,
                                       (const uint8_t *)c.lut8.p, MP, L, LW, hist, carry, (uint32_t *)c.priv.p, tiles,
                                       (unsigned long long *)nullptr);
                    prof_end(st);
                    HIP_TRY(hipGetLastError());
                    if ((e = fork_det())) return e;
                    const size_t pv = lds_cap_pad(rqa::k_lam_vscan<TG>, g_scan_waves);
                    const unsigned gv = grid_for(rqa::k_lam_vscan<TG>, LW.nitems, c.ncu, pv);
                    if ((e = ensure_priv<TG>(gv))) return e;
                    prof_begin("k_lam_vscan", st);
                    hipLaunchKernelGGL((rqa::k_lam_vscan<TG>), dim3(gv), dim3(rqa::WG), pv, st, (const uint8_t *)tiles,
                                       nbands, nc, nt, L, LW, hist, carry, (uint32_t *)c.priv.p);
                    prof_end(st);
                    HIP_TRY(hipGetLastError());
                    done = true;
                }
            }
            if (!done) {
                const unsigned g = grid_for(rqa::k_lam<F, NORM, TG>, LW.nitems, c.ncu);
                if ((e = ensure_priv<TG>(g))) return e;
                prof_begin("k_lam", st);
                hipLaunchKernelGGL((rqa::k_lam<F, NORM, TG>), dim3(g), dim3(rqa::WG), 0, st, dx, nc, tau, m, B, nt,
                                   d_lut, P, L, LW, hist, carry, (uint32_t *)c.priv.p, tiles);
                prof_end(st);
                HIP_TRY(hipGetLastError());
            }
        }
        if ((e = fork_det())) return e;
        const int64_t h0 = c.slab_lamh[s], h1 = c.slab_lamh[s + 1];
        if (h1 > h0) {
            const size_t ph = lds_cap_pad(rqa::k_lam_hscan<TG>, g_scan_waves);
            const unsigned g = grid_for(rqa::k_lam_hscan<TG>, h1 - h0, c.ncu, ph);
            if ((e = ensure_priv<TG>(g))) return e;
            prof_begin("k_lam_hscan", st);
            hipLaunchKernelGGL((rqa::k_lam_hscan<TG>), dim3(g), dim3(rqa::WG), ph, st, (const uint8_t *)tiles, nbands,
                               nc, nt, L, nseg, (const int2 *)c.items_lamh.p + h0, h1 - h0, hist, carry,
                               (uint32_t *)c.priv.p);
            prof_end(st);
            HIP_TRY(hipGetLastError());
        }
        if (want_det && d1 > d0 && !det_pending) {
            const size_t pd = lds_cap_pad(rqa::k_det_scan<TG>, g_scan_waves);
            const unsigned g = grid_for_wg(rqa::k_det_scan<TG>, d1 - d0, c.ncu, pd);
            if ((e = ensure_priv<TG>(g))) return e;
            prof_begin("k_det_scan", st);
            hipLaunchKernelGGL((rqa::k_det_scan<TG>), dim3(g), dim3(rqa::WG), pd, st, (const uint8_t *)tiles, nbands,
                               nc, nt, L, nseg, (const int2 *)c.items_dets.p + d0, d1 - d0, hist2, carry2,
                               (uint32_t *)c.priv.p);
            prof_end(st);
            HIP_TRY(hipGetLastError());
        }
    }
    if (det_pending) HIP_TRY(hipStreamWaitEvent(st, c.ev_det, 0));
    prof_begin("k_stitch", st);
    hipLaunchKernelGGL(rqa::k_stitch, dim3(stitch_blocks(nc), nt), dim3(rqa::STT), 0, st, (const uint2 *)carry, nt,
                       nc, nc, 0, 1, L, nseg, 2 * nseg, hist);
    prof_end(st);
    HIP_TRY(hipGetLastError());
    if ((e = dump_rows(hist, nt, hs, d_perm, c.dump_lam, st))) return e;
    prof_begin("k_finalize", st);
    if ((e = finalize<F>(hist, nc, nt, 0, d_lmins, nl, d_lam, d_perm, T, counts, st))) return e;
    prof_end(st);
    if (want_rr) {
        if ((e = dump_rows(counts, nt, 1, d_perm, c.dump_rr, st))) return e;
        hipLaunchKernelGGL((rqa::k_rr_out<F>), dim3(1), dim3(64), 0, st, (const unsigned long long *)counts, nt, nc,
                           d_perm, d_rr);
        HIP_TRY(hipGetLastError());
    }
    if (want_det) {
        prof_begin("k_stitch", st);
        hipLaunchKernelGGL(rqa::k_stitch, dim3(stitch_blocks(nc - 1), nt), dim3(rqa::STT), 0, st,
                           (const uint2 *)carry2, nt, nc - 1, nc, 1, 0, L, nseg, nseg, hist2);
        prof_end(st);
        HIP_TRY(hipGetLastError());
        prof_begin("k_finalize", st);
        if ((e = finalize<F>(hist2, nc, nt, 1, d_lmins, nl, d_det, d_perm, T, nullptr, st))) return e;
        prof_end(st);
        if ((e = dump_rows(hist2, nt, hs, d_perm, c.dump_det, st))) return e;
    }
    return E_OK;
}

// One group of up to 32 possibly-recurrent thresholds (sorted bounds).
template <typename F, int NORM, int TG>
int launch_group(const F *dx, int64_t nc, int tau, int m, const F *bounds, int nt, const rqa::LutEntry<F> *d_lut,
                 const rqa::LutParams &P, bool want_rr, bool want_det, bool want_lam, const int *d_lmins, int nl,
                 const int *d_perm, int T, F *d_rr, F *d_det, F *d_lam, hipStream_t st) {
    Ctx &c = g_ctx;
    rqa::Bounds<F, TG> B;
    for (int t = 0; t < TG; t++) B.s[t] = t < nt ? bounds[t] : bounds[nt - 1];
    const int L = c.blk_L;
    const int nseg = (int)((nc + L - 1) / L);
    const size_t hs = (size_t)nc + 1;
    int e;
    unsigned long long *counts = (unsigned long long *)c.counts.p;
    if (want_lam && nc > 1 && c.blk_slab)
        return launch_group_slabs<F, NORM, TG>(dx, nc, tau, m, B, nt, d_lut, P, want_rr, want_det, d_lmins, nl,
                                               d_perm, T, d_rr, d_det, d_lam, st);
    // Split form (rank tiles fit in memory): the LAM column kernel ranks the
    // pairs of the upper triangle and the diagonal tiles once, keeps them, and
    // k_lam_hscan / k_det_scan read the rest of LAM and all of DET from them.
    uint8_t *tiles = nullptr;
    if (want_lam && nc > 1 && g_split == 0 && tile_store_fits(nc)) {
        if ((e = ensure(c.tiles, tile_store_bytes(nc)))) return e;
        tiles = (uint8_t *)c.tiles.p;
    }
    const int64_t nbands = (nc + rqa::WAVE - 1) / rqa::WAVE;
    const rqa::LamWork LW = tiles ? rqa::LamWork{(const int2 *)c.items_lamv.p, c.n_items_lamv, nseg, 2 * nseg}
                                  : rqa::LamWork{nullptr, nbands * nseg, nseg, nseg};
    const int cseg_lam = LW.cseg;

    if (want_rr && !want_lam) {
        fill_zero(counts, 32, st);
        const int64_t ni = c.n_items_rr;
        prof_begin("k_rr", st);
        hipLaunchKernelGGL((rqa::k_rr<F, NORM, TG>), dim3(grid_for(rqa::k_rr<F, NORM, TG>, ni, c.ncu)), dim3(rqa::WG), 0,
                           st, dx, nc, tau, m,
                           B, nt, d_lut, P, L, (const int2 *)c.items_rr.p, ni, counts);
        prof_end(st);
        HIP_TRY(hipGetLastError());
    }

    if (want_lam || want_det) {
        if ((e = ensure(c.hist, (size_t)nt * hs * 8))) return e;
        if ((e = ensure(c.metric, (size_t)nt * hs * 8))) return e;
        if ((e = ensure(c.scan, (size_t)nt * hs * 8))) return e;
        if ((e = ensure(c.carry, (size_t)nt * (size_t)nc * cseg_lam * sizeof(uint2)))) return e;
    }
    unsigned long long *hist = (unsigned long long *)c.hist.p;
    uint2 *carry = (uint2 *)c.carry.p;
    // with the tile store, the DET scan runs beside the LAM scans, into hist2
    // and carry2, once the rank pass has written the tiles
    bool det_forked = false;
    auto fork_det = [&]() -> int {
        if (!tiles || !want_det || nc <= 1 || det_forked || !overlap_on()) return E_OK;
        int e2;
        const size_t hs2 = (size_t)nc + 1;
        if ((e2 = ensure(c.hist2, (size_t)nt * hs2 * 8))) return e2;
        if ((e2 = ensure(c.carry2, (size_t)nt * (size_t)(nc - 1) * nseg * sizeof(uint2)))) return e2;
        fill_zero((unsigned long long *)c.hist2.p, (size_t)nt * hs2, st);
        det_forked = true;
        return det_scan_beside<TG>(tiles, nbands, nc, nt, L, nseg, 0, c.n_items_dets,
                                   (unsigned long long *)c.hist2.p, (uint2 *)c.carry2.p, st);
    };

    if (want_lam) {
        fill_zero(hist, (size_t)nt * hs, st);
        if constexpr (NORM == 1) {
            if (use_mfma<F>(1, m, tau)) {
                rqa::Bounds<float, TG> Bf;
                for (int t = 0; t < TG; t++) Bf.s[t] = (float)B.s[t];
                const int gram = tiles ? g_gram : rqa::GRAM_HILO;
                const rqa::MxParams MP = mx_params(Bf.s, nt, m, rqa::NBX_LAM, sizeof(F) == 8 ? 4 : 0, gram);
                if ((e = ensure(c.lut8, rqa::NBX_LAM))) return e;
                hipLaunchKernelGGL((rqa::k_mx_lut<TG>), dim3((MP.nb + 255) / 256), dim3(256), 0, st, Bf, nt, MP,
                                   (const unsigned int *)c.nmax.p, (uint8_t *)c.lut8.p);
                HIP_TRY(hipGetLastError());
                if (tiles) {
                    // split form: ranks and tiles, then the column runs from the tiles
                    prof_begin("k_lam_mx", st);
                    if (gram != rqa::GRAM_HILO) {
                        if constexpr (sizeof(F) == 4 && TG == 16) {
                            if ((e = launch_gram_variant<TG>(gram, dx, nc, tau, m, Bf, nt, MP, L, LW, hist, carry,
                                                             tiles, st)))
                                return e;
                        } else {
                            g_last_error = "the Gram precision comparison needs fp32 input and at most 16 "
                                           "thresholds per block";
                            return E_INVALID;
                        }
                    } else if (g_mxstats) {
                        if ((e = ensure(c.mxstats, 3 * sizeof(unsigned long long)))) return e;
                        const unsigned g = grid_for(rqa::k_lam_mx<F, TG, false, true>, LW.nitems, c.ncu);
                        hipLaunchKernelGGL((rqa::k_lam_mx<F, TG, false, true>), dim3(g), dim3(rqa::WG), 0, st,
                                           dx, (const float *)c.nrm.p, (const unsigned int *)c.nmax.p, nc, tau, m, Bf, B, nt,
                                           (const uint8_t *)c.lut8.p, MP, L, LW, hist, carry, (uint32_t *)c.priv.p,
                                           tiles, (unsigned long long *)c.mxstats.p);
                    } else {
                        const unsigned g = grid_for(rqa::k_lam_mx<F, TG, false>, LW.nitems, c.ncu);
                        hipLaunchKernelGGL((rqa::k_lam_mx<F, TG, false>), dim3(g), dim3(rqa::WG), 0, st,
                                           dx, (const float *)c.nrm.p, (const unsigned int *)c.nmax.p, nc, tau, m, Bf, B, nt,
                                           (const uint8_t *)c.lut8.p, MP, L, LW, hist, carry, (uint32_t *)c.priv.p,
                                           tiles, (unsigned long long *)nullptr);
                    }
                    prof_end(st);
                    HIP_TRY(hipGetLastError());
                    if ((e = fork_det())) return e;
                    const size_t pv = lds_cap_pad(rqa::k_lam_vscan<TG>, g_scan_waves);
                    const unsigned gv = grid_for(rqa::k_lam_vscan<TG>, LW.nitems, c.ncu, pv);
                    if ((e = ensure_priv<TG>(gv))) return e;
                    prof_begin("k_lam_vscan", st);
                    hipLaunchKernelGGL((rqa::k_lam_vscan<TG>), dim3(gv), dim3(rqa::WG), pv, st,
                                       (const uint8_t *)tiles, nbands, nc, nt, L, LW, hist, carry,
                                       (uint32_t *)c.priv.p);
                    prof_end(st);
                    HIP_TRY(hipGetLastError());
                } else {
                    const unsigned g = grid_for(rqa::k_lam_mx<F, TG, true>, LW.nitems, c.ncu);
                    if ((e = ensure_priv<TG>(g))) return e;
                    prof_begin("k_lam_mx", st);
                    hipLaunchKernelGGL((rqa::k_lam_mx<F, TG, true>), dim3(g), dim3(rqa::WG), 0, st, dx,
                                       (const float *)c.nrm.p, (const unsigned int *)c.nmax.p, nc, tau, m, Bf, B, nt, (const uint8_t *)c.lut8.p, MP, L,
                                       LW, hist, carry, (uint32_t *)c.priv.p, tiles, (unsigned long long *)nullptr);
                    prof_end(st);
                    HIP_TRY(hipGetLastError());
                }
            } else {
                const unsigned g = grid_for(rqa::k_lam<F, NORM, TG>, LW.nitems, c.ncu);
                if ((e = ensure_priv<TG>(g))) return e;
                prof_begin("k_lam", st);
                hipLaunchKernelGGL((rqa::k_lam<F, NORM, TG>), dim3(g), dim3(rqa::WG), 0, st, dx, nc, tau, m, B, nt,
                                   d_lut, P, L, LW, hist, carry, (uint32_t *)c.priv.p, tiles);
                prof_end(st);
                HIP_TRY(hipGetLastError());
            }
        } else {
            const unsigned g = grid_for(rqa::k_lam<F, NORM, TG>, LW.nitems, c.ncu);
            if ((e = ensure_priv<TG>(g))) return e;
            prof_begin("k_lam", st);
            hipLaunchKernelGGL((rqa::k_lam<F, NORM, TG>), dim3(g), dim3(rqa::WG), 0, st, dx, nc, tau, m, B, nt, d_lut,
                               P, L, LW, hist, carry, (uint32_t *)c.priv.p, tiles);
            prof_end(st);
            HIP_TRY(hipGetLastError());
        }
        if ((e = fork_det())) return e;
        if (tiles && c.n_items_lamh > 0) {
            const size_t ph = lds_cap_pad(rqa::k_lam_hscan<TG>, g_scan_waves);
            const unsigned g = grid_for(rqa::k_lam_hscan<TG>, c.n_items_lamh, c.ncu, ph);
            if ((e = ensure_priv<TG>(g))) return e;
            prof_begin("k_lam_hscan", st);
            hipLaunchKernelGGL((rqa::k_lam_hscan<TG>), dim3(g), dim3(rqa::WG), ph, st, (const uint8_t *)tiles, nbands,
                               nc, nt, L, nseg, (const int2 *)c.items_lamh.p, c.n_items_lamh, hist, carry,
                               (uint32_t *)c.priv.p);
            prof_end(st);
            HIP_TRY(hipGetLastError());
        }
        prof_begin("k_stitch", st);
        hipLaunchKernelGGL(rqa::k_stitch, dim3(stitch_blocks(nc), nt), dim3(rqa::STT), 0, st, (const uint2 *)carry,
                           nt, nc, nc, 0, tiles ? 1 : 0, L, nseg, cseg_lam, hist);
        prof_end(st);
        HIP_TRY(hipGetLastError());
        if ((e = dump_rows(hist, nt, hs, d_perm, c.dump_lam, st))) return e;
        // The LAM histogram covers every recurrent pair of the full matrix once,
        // so its sum of l*h[l] is the RR count as well.
        prof_begin("k_finalize", st);
        if ((e = finalize<F>(hist, nc, nt, 0, d_lmins, nl, d_lam, d_perm, T, counts, st))) return e;
        prof_end(st);
    }
    if (want_rr) {
        if ((e = dump_rows(counts, nt, 1, d_perm, c.dump_rr, st))) return e;
        hipLaunchKernelGGL((rqa::k_rr_out<F>), dim3(1), dim3(64), 0, st, (const unsigned long long *)counts, nt, nc,
                           d_perm, d_rr);
        HIP_TRY(hipGetLastError());
    }
    if (want_det && det_forked) {
        unsigned long long *hist2 = (unsigned long long *)c.hist2.p;
        HIP_TRY(hipStreamWaitEvent(st, c.ev_det, 0));
        prof_begin("k_stitch", st);
        hipLaunchKernelGGL(rqa::k_stitch, dim3(stitch_blocks(nc - 1), nt), dim3(rqa::STT), 0, st,
                           (const uint2 *)c.carry2.p, nt, nc - 1, nc, 1, 0, L, nseg, nseg, hist2);
        prof_end(st);
        HIP_TRY(hipGetLastError());
        prof_begin("k_finalize", st);
        if ((e = finalize<F>(hist2, nc, nt, 1, d_lmins, nl, d_det, d_perm, T, nullptr, st))) return e;
        prof_end(st);
        if ((e = dump_rows(hist2, nt, hs, d_perm, c.dump_det, st))) return e;
        return E_OK;
    }
    if (want_det) {
        fill_zero(hist, (size_t)nt * hs, st);
        if (nc > 1) {
            const int64_t ni = c.n_items_det;
            bool done = false;
            if (tiles) {
                const int64_t nig = c.n_items_dets;
                const size_t pd = lds_cap_pad(rqa::k_det_scan<TG>, g_scan_waves);
                const unsigned g = grid_for_wg(rqa::k_det_scan<TG>, nig, c.ncu, pd);
                if ((e = ensure_priv<TG>(g))) return e;
                prof_begin("k_det_scan", st);
                hipLaunchKernelGGL((rqa::k_det_scan<TG>), dim3(g), dim3(rqa::WG), pd, st, (const uint8_t *)tiles,
                                   nbands, nc, nt, L, nseg,
                                   (const int2 *)c.items_dets.p, nig, hist, carry, (uint32_t *)c.priv.p);
                prof_end(st);
                HIP_TRY(hipGetLastError());
                done = true;
            }
            if constexpr (sizeof(F) == 4 && NORM == 1) {
                if (!done && use_mfma<F>(1, m, tau)) {
                    rqa::Bounds<float, TG> Bf;
                    for (int t = 0; t < TG; t++) Bf.s[t] = (float)B.s[t];
                    const rqa::MxParams MP = mx_params((const float *)bounds, nt, m, rqa::NBX);
                    if ((e = ensure(c.lut8, rqa::NBX))) return e;
                    uint8_t *lut8d = (uint8_t *)c.lut8.p;
                    hipLaunchKernelGGL((rqa::k_mx_lut<TG>), dim3((MP.nb + 255) / 256), dim3(256), 0, st, Bf, nt, MP,
                                       (const unsigned int *)c.nmax.p, lut8d);
                    HIP_TRY(hipGetLastError());
                    const unsigned g = grid_for(rqa::k_det_mx<TG>, ni, c.ncu);
                    if ((e = ensure_priv<TG>(g))) return e;
                    prof_begin("k_det_mx", st);
                    hipLaunchKernelGGL((rqa::k_det_mx<TG>), dim3(g), dim3(rqa::WG), 0, st, (const float *)dx,
                                       (const float *)c.nrm.p, (const unsigned int *)c.nmax.p, nc, tau, m, Bf, nt,
                                       (const uint8_t *)lut8d, MP, L, nseg,
                                       (const int2 *)c.items_det.p, ni, hist, carry, (uint32_t *)c.priv.p);
                    prof_end(st);
                    HIP_TRY(hipGetLastError());
                    done = true;
                }
            }
            if (!done) {
                const unsigned g = grid_for(rqa::k_det<F, NORM, TG>, ni, c.ncu);
                if ((e = ensure_priv<TG>(g))) return e;
                prof_begin("k_det", st);
                hipLaunchKernelGGL((rqa::k_det<F, NORM, TG>), dim3(g), dim3(rqa::WG), 0, st, dx, nc, tau, m, B, nt,
                                   d_lut, P, L, nseg, (const int2 *)c.items_det.p, ni, hist, carry,
                                   (uint32_t *)c.priv.p);
                prof_end(st);
                HIP_TRY(hipGetLastError());
            }
            prof_begin("k_stitch", st);
            hipLaunchKernelGGL(rqa::k_stitch, dim3(stitch_blocks(nc - 1), nt), dim3(rqa::STT), 0, st,
                               (const uint2 *)carry, nt, nc - 1, nc, 1, 0, L, nseg, nseg, hist);
            prof_end(st);
            HIP_TRY(hipGetLastError());
        }
        prof_begin("k_finalize", st);
        if ((e = finalize<F>(hist, nc, nt, 1, d_lmins, nl, d_det, d_perm, T, nullptr, st))) return e;
        prof_end(st);
        // finalize leaves h_D doubled with the line-of-identity bin
        if ((e = dump_rows(hist, nt, hs, d_perm, c.dump_det, st))) return e;
    }
    return E_OK;
}

// Thresholds that never recur: empty histograms, measures straight from finalize.
template <typename F>
int launch_never(int64_t nc, int nt, bool want_rr, bool want_det, bool want_lam, const int *d_lmins, int nl,
                 const int *d_perm, int T, F *d_rr, F *d_det, F *d_lam, hipStream_t st) {
    Ctx &c = g_ctx;
    const size_t hs = (size_t)nc + 1;
    int e;
    if ((e = ensure(c.hist, (size_t)nt * hs * 8))) return e;
    if ((e = ensure(c.metric, (size_t)nt * hs * 8))) return e;
    if ((e = ensure(c.scan, (size_t)nt * hs * 8))) return e;
    if ((e = ensure(c.counts, std::max<size_t>(64, (size_t)nt) * sizeof(unsigned long long)))) return e;
    unsigned long long *hist = (unsigned long long *)c.hist.p;
    unsigned long long *counts = (unsigned long long *)c.counts.p;
    if (want_rr) {
        fill_zero(counts, (size_t)nt, st);
        hipLaunchKernelGGL((rqa::k_rr_out<F>), dim3((nt + 63) / 64), dim3(64), 0, st,
                           (const unsigned long long *)counts, nt, nc, d_perm, d_rr);
        HIP_TRY(hipGetLastError());
        if ((e = dump_rows(counts, nt, 1, d_perm, c.dump_rr, st))) return e;
    }
    for (int kind = 0; kind < 2; kind++) {
        const bool det = kind == 1;
        if (det ? !want_det : !want_lam) continue;
        fill_zero(hist, (size_t)nt * hs, st);
        if ((e = finalize<F>(hist, nc, nt, det ? 1 : 0, d_lmins, nl, det ? d_det : d_lam, d_perm, T, nullptr, st)))
            return e;
        if ((e = dump_rows(hist, nt, hs, d_perm, det ? c.dump_det : c.dump_lam, st))) return e;
    }
    return E_OK;
}

template <typename F, int NORM>
int launch_tg(int TG, const F *dx, int64_t nc, int tau, int m, const F *bounds, int nt, const rqa::LutEntry<F> *lut,
              const rqa::LutParams &P, bool rr, bool det, bool lam, const int *d_lmins, int nl, const int *d_perm,
              int T, F *d_rr, F *d_det, F *d_lam, hipStream_t st) {
    switch (TG) {
        case 4:
            return launch_group<F, NORM, 4>(dx, nc, tau, m, bounds, nt, lut, P, rr, det, lam, d_lmins, nl, d_perm, T,
                                            d_rr, d_det, d_lam, st);
        case 8:
            return launch_group<F, NORM, 8>(dx, nc, tau, m, bounds, nt, lut, P, rr, det, lam, d_lmins, nl, d_perm, T,
                                            d_rr, d_det, d_lam, st);
        case 16:
            return launch_group<F, NORM, 16>(dx, nc, tau, m, bounds, nt, lut, P, rr, det, lam, d_lmins, nl, d_perm,
                                             T, d_rr, d_det, d_lam, st);
        default:
            return launch_group<F, NORM, 32>(dx, nc, tau, m, bounds, nt, lut, P, rr, det, lam, d_lmins, nl, d_perm,
                                             T, d_rr, d_det, d_lam, st);
    }
}

// One (tau, emb) block: dx is the padded device copy of the series.
template <typename F>
int run_block(const F *dx, size_t n, int tau, int emb, const F *thr, int T, const int *lmins, int nl, int dist,
              F *d_rr, F *d_det, F *d_lam, hipStream_t st) {
    Ctx &c = g_ctx;
    const int64_t nc = (int64_t)n - (int64_t)(emb - 1) * tau;
    // bounds; recurring thresholds sorted ascending, then the never-recurring
    // ones; perm maps that order back to the caller's threshold index.
    std::vector<F> bnd(T);
    std::vector<int> live, dead;
    for (int t = 0; t < T; t++) {
        F b = bound_of(thr[t], dist);
        if (never_recurs(b)) {
            dead.push_back(t);
        } else {
            bnd[t] = b == F(0) ? F(0) : b;   // -0 -> +0
            live.push_back(t);
        }
    }
    std::stable_sort(live.begin(), live.end(), [&](int a, int b) { return bnd[a] < bnd[b]; });
    std::vector<int> order(live);
    order.insert(order.end(), dead.begin(), dead.end());
    const int ngroups = ((int)live.size() + 31) / 32;
    std::vector<rqa::LutEntry<F>> luts((size_t)std::max(ngroups, 1) * rqa::NB_MAX);
    std::vector<rqa::LutParams> lps(std::max(ngroups, 1));
    std::vector<F> sorted(live.size());
    for (size_t t = 0; t < live.size(); t++) sorted[t] = bnd[live[t]];
    for (int g = 0; g < ngroups; g++) {
        const int g0 = g * 32, nt = std::min(32, (int)live.size() - g0);
        build_lut(sorted.data() + g0, nt, luts.data() + (size_t)g * rqa::NB_MAX, lps[g]);
    }
    int e;
    if ((e = ensure(c.perm, (size_t)T * sizeof(int)))) return e;
    if ((e = ensure(c.lmins, (size_t)std::max(nl, 1) * sizeof(int)))) return e;
    if ((e = ensure(c.lut, luts.size() * sizeof(rqa::LutEntry<F>)))) return e;
    if ((e = ensure(c.counts, 64 * sizeof(unsigned long long)))) return e;
    if ((e = upload_once(c.perm, g_up_perm, order.data(), (size_t)T * sizeof(int), st))) return e;
    if (nl > 0 && (e = upload_once(c.lmins, g_up_lmins, lmins, (size_t)nl * sizeof(int), st))) return e;
    if (ngroups > 0 && (e = upload_once(c.lut, g_up_lut, luts.data(),
                                        (size_t)ngroups * rqa::NB_MAX * sizeof(rqa::LutEntry<F>), st)))
        return e;
    // the split form's shape: the whole tile store, slab by slab, or none
    // (launch_group decides per group whether to take it)
    c.blk_L = seg_len_for(nc);
    c.blk_slab = false;
    if (d_lam && nc > 1 && g_split == 0 && (g_slab_rows > 0 || !tile_store_fits(nc))) {
        const int Ls = slab_len_for(nc);
        if (Ls > 0) {
            c.blk_L = Ls;
            c.blk_slab = true;
        }
    }
    if ((e = build_items(nc, c.blk_L, c.blk_slab, st))) return e;
    if (use_mfma<F>(dist, emb, tau) && ngroups > 0) {
        if ((e = ensure(c.nrm, ((size_t)nc + 2 * rqa::WAVE) * sizeof(float)))) return e;
        if ((e = ensure(c.nmax, 256))) return e;
        fill_zero((unsigned long long *)c.nmax.p, 1, st);
        const int64_t nthr = nc + 2 * rqa::WAVE;
        hipLaunchKernelGGL((rqa::k_norms<F>), dim3((unsigned)((nthr + 255) / 256)), dim3(256), 0, st, dx, nc,
                           tau, emb, (float *)c.nrm.p, (unsigned int *)c.nmax.p);
        HIP_TRY(hipGetLastError());
    }

    for (int g = 0; g < ngroups; g++) {
        const int g0 = g * 32, nt = std::min(32, (int)live.size() - g0);
        const int *gperm = (const int *)c.perm.p + g0;
        const rqa::LutEntry<F> *glut = (const rqa::LutEntry<F> *)c.lut.p + (size_t)g * rqa::NB_MAX;
        if (dist == 1)
            e = launch_tg<F, 1>(g_gram != rqa::GRAM_HILO && nt <= 16 ? 16 : tg_for(nt), dx, nc, tau, emb, sorted.data() + g0, nt, glut, lps[g], d_rr != nullptr,
                                d_det != nullptr, d_lam != nullptr, (const int *)c.lmins.p, nl, gperm, T, d_rr, d_det,
                                d_lam, st);
        else
            e = launch_tg<F, 2>(tg_for(nt), dx, nc, tau, emb, sorted.data() + g0, nt, glut, lps[g], d_rr != nullptr,
                                d_det != nullptr, d_lam != nullptr, (const int *)c.lmins.p, nl, gperm, T, d_rr, d_det,
                                d_lam, st);
        if (e) return e;
    }
    if (!dead.empty()) {
        e = launch_never<F>(nc, (int)dead.size(), d_rr != nullptr, d_det != nullptr, d_lam != nullptr,
                            (const int *)c.lmins.p, nl, (const int *)c.perm.p + live.size(), T, d_rr, d_det, d_lam,
                            st);
        if (e) return e;
    }
    prof_collect(st);
    return E_OK;
}

// ---------------------------------------------------------------- validation
int check_args(size_t n, const int *tau, int ntau, const int *emb, int nemb, const int *lm, int nl, int dist) {
    if (dist != 1 && dist != 2) return E_INVALID;
    for (int a = 0; a < ntau; a++) {
        if (tau[a] < 0) return E_INVALID;
        for (int b = 0; b < nemb; b++) {
            if (emb[b] < 1) return E_INVALID;
            int64_t nc = (int64_t)n - (int64_t)(emb[b] - 1) * tau[a];
            if (nc < 1 || nc > (int64_t)INT32_MAX) return E_INVALID;
            for (int l = 0; l < nl; l++)
                if (lm[l] < 0 || (int64_t)lm[l] > nc) return E_INVALID;
        }
    }
    return E_OK;
}

template <typename F>
int upload_series(const F *src, size_t n, bool src_on_device, hipStream_t st) {
    Ctx &c = g_ctx;
    int e = ensure(c.x, (n + PAD) * sizeof(F));
    if (e) return e;
    F *dx = (F *)c.x.p;
    if (src_on_device) {
        unsigned blocks = (unsigned)std::min<size_t>((n + PAD + 255) / 256, 4096);
        hipLaunchKernelGGL((k_pad_copy<F>), dim3(blocks), dim3(256), 0, st, src, dx, n, n + PAD);
        HIP_TRY(hipGetLastError());
    } else {
        HIP_TRY(hipMemcpyAsync(dx, src, n * sizeof(F), hipMemcpyHostToDevice, st));
        std::vector<F> zeros(PAD, F(0));
        HIP_TRY(hipMemcpyAsync(dx + n, zeros.data(), PAD * sizeof(F), hipMemcpyHostToDevice, st));
        HIP_TRY(hipStreamSynchronize(st));
    }
    return E_OK;
}

// kind: 0 RR, 1 DET, 2 LAM
template <typename F>
void host_api(int kind, F *output, const F *x, size_t n, const int *tau, int ntau, const int *emb, int nemb,
              const int *lm, int nl, const F *thr, int nthr, int dist, int *error) {
    if (!error) return;
    *error = E_OK;
    const bool has_l = kind != 0;
    if (n == 0 || ntau <= 0 || nemb <= 0 || nthr <= 0 || (has_l && nl <= 0)) *error = E_INVALID;
    if (!output || !x || !tau || !emb || !thr || (has_l && !lm)) *error = E_NULL;
    if (*error) return;
    std::lock_guard<std::mutex> lk(g_mu);
    int e = check_args(n, tau, ntau, emb, nemb, lm, has_l ? nl : 0, dist);
    if (!e) e = init_ctx();
    if (e) {
        *error = e;
        return;
    }
    Ctx &c = g_ctx;
    hipStream_t st = c.stream;
    CallScope call(st);
    if ((e = call.err) || (e = upload_series(x, n, false, st))) {
        *error = e;
        return;
    }
    // every (tau, emb) block writes its own part of one device output, which
    // comes back in one copy
    const size_t per = kind == 0 ? (size_t)nthr : (size_t)nthr * nl * 5;
    const size_t total = per * (size_t)ntau * (size_t)nemb;
    DevBuf &ob = kind == 0 ? c.out_rr : (kind == 1 ? c.out_det : c.out_lam);
    if ((e = ensure(ob, total * sizeof(F)))) {
        *error = e;
        return;
    }
    for (int a = 0; a < ntau; a++)
        for (int b = 0; b < nemb; b++) {
            F *dout = (F *)ob.p + ((size_t)a * nemb + b) * per;
            e = run_block<F>((const F *)c.x.p, n, tau[a], emb[b], thr, nthr, has_l ? lm : nullptr, has_l ? nl : 0,
                             dist, kind == 0 ? dout : nullptr, kind == 1 ? dout : nullptr, kind == 2 ? dout : nullptr,
                             st);
            if (e) {
                (void)hipStreamSynchronize(st);
                *error = e;
                return;
            }
        }
    hipError_t he = hipMemcpyAsync(output, ob.p, total * sizeof(F), hipMemcpyDeviceToHost, st);
    if (he == hipSuccess) he = hipStreamSynchronize(st);
    if (he != hipSuccess) *error = map_err(he);
}

template <typename F>
void host_hist(const F *x, size_t n, int tau, int emb, const F *thr, int T, int dist, int which,
               unsigned long long *rr, unsigned long long *hd, unsigned long long *hv, int *error) {
    if (!error) return;
    *error = E_OK;
    if (n == 0 || T <= 0) *error = E_INVALID;
    if (!x || !thr || ((which & 1) && !rr) || ((which & 2) && !hd) || ((which & 4) && !hv)) *error = E_NULL;
    if (*error) return;
    std::lock_guard<std::mutex> lk(g_mu);
    int e = check_args(n, &tau, 1, &emb, 1, nullptr, 0, dist);
    if (!e) e = init_ctx();
    if (e) {
        *error = e;
        return;
    }
    Ctx &c = g_ctx;
    hipStream_t st = c.stream;
    CallScope call(st);
    if ((e = call.err) || (e = upload_series(x, n, false, st))) {
        *error = e;
        return;
    }
    const int64_t nc = (int64_t)n - (int64_t)(emb - 1) * tau;
    const size_t hs = (size_t)nc + 1;
    // The histograms come from the same kernels and launches as RR/DET/LAM,
    // up to HCH thresholds per run_block (one threshold group), each group's
    // raw rows copied to the dump sinks in the caller's order; the measures
    // finalize writes are discarded (l_min = 0 is always valid).
    constexpr int HCH = 32;
    int lm0 = 0;
    const size_t rows = (size_t)HCH * hs;
    const size_t words = HCH + ((which & 2) ? rows : 0) + ((which & 4) ? rows : 0);
    if ((e = ensure(c.out_rr, HCH * sizeof(F))) || (e = ensure(c.out_det, HCH * 5 * sizeof(F))) ||
        (e = ensure(c.out_lam, HCH * 5 * sizeof(F))) || (e = ensure(c.dump, words * 8))) {
        *error = e;
        return;
    }
    unsigned long long *d_rr = (unsigned long long *)c.dump.p;
    unsigned long long *d_hd = d_rr + HCH;
    unsigned long long *d_hv = d_hd + ((which & 2) ? rows : 0);
    for (int t0 = 0; t0 < T && !e; t0 += HCH) {
        const int nt = std::min(HCH, T - t0);
        fill_zero(d_rr, words, st);
        c.dump_rr = (which & 1) ? d_rr : nullptr;
        c.dump_det = (which & 2) ? d_hd : nullptr;
        c.dump_lam = (which & 4) ? d_hv : nullptr;
        e = run_block<F>((const F *)c.x.p, n, tau, emb, thr + t0, nt, &lm0, 1, dist,
                         (which & 1) ? (F *)c.out_rr.p : nullptr, (which & 2) ? (F *)c.out_det.p : nullptr,
                         (which & 4) ? (F *)c.out_lam.p : nullptr, st);
        c.dump_rr = c.dump_det = c.dump_lam = nullptr;
        hipError_t he = hipSuccess;
        if (!e && (which & 1)) he = hipMemcpyAsync(rr + t0, d_rr, (size_t)nt * 8, hipMemcpyDeviceToHost, st);
        if (!e && he == hipSuccess && (which & 2))
            he = hipMemcpyAsync(hd + (size_t)t0 * hs, d_hd, (size_t)nt * hs * 8, hipMemcpyDeviceToHost, st);
        if (!e && he == hipSuccess && (which & 4))
            he = hipMemcpyAsync(hv + (size_t)t0 * hs, d_hv, (size_t)nt * hs * 8, hipMemcpyDeviceToHost, st);
        if (!e && he == hipSuccess) he = hipStreamSynchronize(st);
        if (!e && he != hipSuccess) e = map_err(he);
    }
    *error = e;
}

template <typename F>
int device_api(const F *d_in, size_t n, int tau, int emb, const F *thr, int T, const int *lm, int nl, int dist,
               F *d_rr, F *d_det, F *d_lam, void *stream) {
    if (n == 0 || T <= 0 || ((d_det || d_lam) && nl <= 0)) return E_INVALID;
    if (!d_in || !thr || ((d_det || d_lam) && !lm)) return E_NULL;
    std::lock_guard<std::mutex> lk(g_mu);
    int e = check_args(n, &tau, 1, &emb, 1, lm, (d_det || d_lam) ? nl : 0, dist);
    if (!e) e = init_ctx();
    if (e) return e;
    hipStream_t st = (hipStream_t)stream;
    CallScope call(st);
    if ((e = call.err) || (e = upload_series(d_in, n, true, st))) return e;
    return run_block<F>((const F *)g_ctx.x.p, n, tau, emb, thr, T, lm, (d_det || d_lam) ? nl : 0, dist, d_rr, d_det,
                        d_lam, st);
}

// fp32 decisions against true fp64 decisions on the same samples (diagnostic):
// counts[t] = pairs of the full matrix whose decisions differ, in the
// caller's threshold order.
int device_fp64_diag(const float *d_in, size_t n, int tau, int emb, const float *thr, int T, int dist,
                     unsigned long long *counts, void *stream) {
    if (n == 0 || T <= 0) return E_INVALID;
    if (!d_in || !thr || !counts) return E_NULL;
    std::lock_guard<std::mutex> lk(g_mu);
    int e = check_args(n, &tau, 1, &emb, 1, nullptr, 0, dist);
    if (!e) e = init_ctx();
    if (e) return e;
    hipStream_t st = (hipStream_t)stream;
    CallScope call(st);
    if ((e = call.err) || (e = upload_series(d_in, n, true, st))) return e;
    Ctx &c = g_ctx;
    const int64_t nc = (int64_t)n - (int64_t)(emb - 1) * tau;
    if ((e = build_items(nc, seg_len_for(nc), false, st))) return e;
    if ((e = ensure(c.counts, std::max<size_t>(64, (size_t)T) * sizeof(unsigned long long)))) return e;
    unsigned long long *d_counts = (unsigned long long *)c.counts.p;
    fill_zero(d_counts, (size_t)T, st);
    const int L = seg_len_for(nc);
    for (int t0 = 0; t0 < T; t0 += rqa::DIAG_MAXT) {
        const int nt = std::min(rqa::DIAG_MAXT, T - t0);
        rqa::Bounds<float, rqa::DIAG_MAXT> B32;
        rqa::Bounds<double, rqa::DIAG_MAXT> B64;
        for (int t = 0; t < rqa::DIAG_MAXT; t++) {
            const float e32 = thr[t0 + std::min(t, nt - 1)];
            const float b32 = bound_of<float>(e32, dist);
            const double b64 = bound_of<double>((double)e32, dist);
            B32.s[t] = never_recurs(b32) ? -1.0f : b32;
            B64.s[t] = never_recurs(b64) ? -1.0 : b64;
        }
        const int64_t ni = c.n_items_rr;
        if (dist == 1) {
            const unsigned g = grid_for(rqa::k_fp64_diag<1>, ni, c.ncu);
            hipLaunchKernelGGL(rqa::k_fp64_diag<1>, dim3(g), dim3(rqa::WG), 0, st, (const float *)c.x.p, nc, tau, emb,
                               B32, B64, nt, L, (const int2 *)c.items_rr.p, ni, d_counts + t0);
        } else {
            const unsigned g = grid_for(rqa::k_fp64_diag<2>, ni, c.ncu);
            hipLaunchKernelGGL(rqa::k_fp64_diag<2>, dim3(g), dim3(rqa::WG), 0, st, (const float *)c.x.p, nc, tau, emb,
                               B32, B64, nt, L, (const int2 *)c.items_rr.p, ni, d_counts + t0);
        }
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipMemcpyAsync(counts, d_counts, (size_t)T * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return E_OK;
}

// The recurrence plot of one (tau, emb, threshold) into d_out (N_c x N_c
// elements of O, row-major), from the padded series dx.  A threshold that
// never recurs (negative or NaN) gets bound -1, below every s.
template <typename F, typename O>
int rp_run(const F *dx, size_t n, int tau, int emb, F thr, int dist, O *d_out, hipStream_t st) {
    Ctx &c = g_ctx;
    const int64_t nc = (int64_t)n - (int64_t)(emb - 1) * tau;
    F b = bound_of(thr, dist);
    const bool never = never_recurs(b);
    if (never) b = F(-1);
    const int64_t nblk = ((nc + rqa::BLK - 1) / rqa::BLK) * ((nc + rqa::WAVE - 1) / rqa::WAVE);
    if constexpr (sizeof(F) == 4 && sizeof(O) == 1) {
        if (!never && dist == 1 && use_mfma<F>(1, emb, tau)) {
            // MFMA form: norms, the one-bound byte table, then the plot
            int e;
            if ((e = ensure(c.nrm, ((size_t)nc + 2 * rqa::WAVE) * sizeof(float)))) return e;
            if ((e = ensure(c.nmax, 256))) return e;
            if ((e = ensure(c.lut8, rqa::NBX_LAM))) return e;
            fill_zero((unsigned long long *)c.nmax.p, 1, st);
            const int64_t nthr = nc + 2 * rqa::WAVE;
            hipLaunchKernelGGL((rqa::k_norms<F>), dim3((unsigned)((nthr + 255) / 256)), dim3(256), 0, st, dx, nc, tau,
                               emb, (float *)c.nrm.p, (unsigned int *)c.nmax.p);
            HIP_TRY(hipGetLastError());
            rqa::Bounds<float, 4> Bf;
            for (int t = 0; t < 4; t++) Bf.s[t] = (float)b;
            // the upper triangle, each square stored mirrored as well, where
            // its 16-byte stores align
            const bool sym = g_rp_sym && nc % 16 == 0 && (uintptr_t)d_out % 16 == 0;
            const rqa::MxParams MP = mx_params(Bf.s, 1, emb, sym ? rqa::NBX_RP : rqa::NBX_LAM);
            hipLaunchKernelGGL((rqa::k_mx_lut<4>), dim3((MP.nb + 255) / 256), dim3(256), 0, st, Bf, 1, MP,
                               (const unsigned int *)c.nmax.p, (uint8_t *)c.lut8.p);
            HIP_TRY(hipGetLastError());
            if (sym) {
                const int64_t nbands = (nc + rqa::WAVE - 1) / rqa::WAVE;
                const unsigned g = xcd_round(grid_for(rqa::k_rp_mx_sym, rqa::rp_items(nbands), c.ncu));
                prof_begin("k_rp_mx_sym", st);
                hipLaunchKernelGGL(rqa::k_rp_mx_sym, dim3(g), dim3(rqa::WG), 0, st, (const float *)dx,
                                   (const float *)c.nrm.p, (const unsigned int *)c.nmax.p, nc, tau, emb, (float)b,
                                   (const uint8_t *)c.lut8.p, MP, (char *)d_out);
            } else {
                const unsigned g = grid_for(rqa::k_rp_mx, nblk, c.ncu);
                prof_begin("k_rp_mx", st);
                hipLaunchKernelGGL(rqa::k_rp_mx, dim3(g), dim3(rqa::WG), 0, st, (const float *)dx,
                                   (const float *)c.nrm.p, (const unsigned int *)c.nmax.p, nc, tau, emb, (float)b,
                                   (const uint8_t *)c.lut8.p, MP, (char *)d_out);
            }
            prof_end(st);
            HIP_TRY(hipGetLastError());
            prof_collect(st);
            return E_OK;
        }
    }
    const uintptr_t oa = (uintptr_t)d_out;
    if constexpr (sizeof(O) == 1) {
        // the upper triangle, each square stored mirrored as well
        if (g_rp_sym && nc % 16 == 0 && oa % 16 == 0) {
            const int64_t nsq = rqa::rp_items((nc + rqa::WAVE - 1) / rqa::WAVE);
            prof_begin("k_rp_sym", st);
            if (dist == 1) {
                const unsigned g = xcd_round(grid_for(rqa::k_rp_sym<F, 1>, nsq, c.ncu));
                hipLaunchKernelGGL((rqa::k_rp_sym<F, 1>), dim3(g), dim3(rqa::WG), 0, st, dx, nc, tau, emb, b,
                                   (char *)d_out);
            } else {
                const unsigned g = xcd_round(grid_for(rqa::k_rp_sym<F, 2>, nsq, c.ncu));
                hipLaunchKernelGGL((rqa::k_rp_sym<F, 2>), dim3(g), dim3(rqa::WG), 0, st, dx, nc, tau, emb, b,
                                   (char *)d_out);
            }
            prof_end(st);
            HIP_TRY(hipGetLastError());
            prof_collect(st);
            return E_OK;
        }
    }
    // four columns per lane when its 4-element stores align
    if (g_rp_wide && nc % 4 == 0 && oa % (sizeof(O) == 1 ? 4 : 16) == 0) {
        constexpr int R4 = rqa::rp4_rows<F>();
        const int64_t nb4 = ((nc + R4 - 1) / R4) * ((nc + 4 * rqa::WAVE - 1) / (4 * rqa::WAVE));
        prof_begin("k_rp4", st);
        if (dist == 1) {
            const unsigned g = grid_for(rqa::k_rp4<F, 1, O>, nb4, c.ncu);
            hipLaunchKernelGGL((rqa::k_rp4<F, 1, O>), dim3(g), dim3(rqa::WG), 0, st, dx, nc, tau, emb, b, d_out);
        } else {
            const unsigned g = grid_for(rqa::k_rp4<F, 2, O>, nb4, c.ncu);
            hipLaunchKernelGGL((rqa::k_rp4<F, 2, O>), dim3(g), dim3(rqa::WG), 0, st, dx, nc, tau, emb, b, d_out);
        }
        prof_end(st);
        HIP_TRY(hipGetLastError());
        prof_collect(st);
        return E_OK;
    }
    prof_begin("k_rp", st);
    if (dist == 1) {
        const unsigned g = grid_for(rqa::k_rp<F, 1, O>, nblk, c.ncu);
        hipLaunchKernelGGL((rqa::k_rp<F, 1, O>), dim3(g), dim3(rqa::WG), 0, st, dx, nc, tau, emb, b, d_out);
    } else {
        const unsigned g = grid_for(rqa::k_rp<F, 2, O>, nblk, c.ncu);
        hipLaunchKernelGGL((rqa::k_rp<F, 2, O>), dim3(g), dim3(rqa::WG), 0, st, dx, nc, tau, emb, b, d_out);
    }
    prof_end(st);
    HIP_TRY(hipGetLastError());
    prof_collect(st);
    return E_OK;
}

template <typename F, typename O>
void host_rp(O *output, const F *x, size_t n, int tau, int emb, F thr, int dist, int *error) {
    if (!error) return;
    *error = E_OK;
    if (n == 0) *error = E_INVALID;
    if (!output || !x) *error = E_NULL;
    if (*error) return;
    std::lock_guard<std::mutex> lk(g_mu);
    int e = check_args(n, &tau, 1, &emb, 1, nullptr, 0, dist);
    if (!e) e = init_ctx();
    if (e) {
        *error = e;
        return;
    }
    Ctx &c = g_ctx;
    hipStream_t st = c.stream;
    CallScope call(st);
    e = call.err;
    if (!e) e = upload_series(x, n, false, st);
    const int64_t nc = (int64_t)n - (int64_t)(emb - 1) * tau;
    const size_t bytes = e ? 0 : (size_t)nc * (size_t)nc * sizeof(O);
    if (!e) e = ensure(c.out_rp, bytes);
    if (!e) e = rp_run<F, O>((const F *)c.x.p, n, tau, emb, thr, dist, (O *)c.out_rp.p, st);
    if (!e) {
        hipError_t he = hipMemcpyAsync(output, c.out_rp.p, bytes, hipMemcpyDeviceToHost, st);
        if (he == hipSuccess) he = hipStreamSynchronize(st);
        if (he != hipSuccess) e = map_err(he);
    }
    *error = e;
}

template <typename F, typename O>
int device_rp(const F *d_in, size_t n, int tau, int emb, F thr, int dist, O *d_out, void *stream) {
    if (n == 0) return E_INVALID;
    if (!d_in || !d_out) return E_NULL;
    std::lock_guard<std::mutex> lk(g_mu);
    int e = check_args(n, &tau, 1, &emb, 1, nullptr, 0, dist);
    if (!e) e = init_ctx();
    if (e) return e;
    hipStream_t st = (hipStream_t)stream;
    CallScope call(st);
    if ((e = call.err) || (e = upload_series(d_in, n, true, st))) return e;
    return rp_run<F, O>((const F *)g_ctx.x.p, n, tau, emb, thr, dist, d_out, st);
}

// The padded copy and the block launcher serve the batched engine's
// long series as well (rqa_batch.hip).
template int upload_series<float>(const float *, size_t, bool, hipStream_t);
template int upload_series<double>(const double *, size_t, bool, hipStream_t);
template int run_block<float>(const float *, size_t, int, int, const float *, int, const int *, int, int, float *,
                              float *, float *, hipStream_t);
template int run_block<double>(const double *, size_t, int, int, const double *, int, const int *, int, int,
                               double *, double *, double *, hipStream_t);

}  // namespace eng

using namespace eng;

// ---------------------------------------------------------------- C ABI
extern "C" {

void accrqa_amd_RR_f32(float *output, const float *input_data, size_t data_size, const int *tau_values, int nTaus,
                       const int *emb_values, int nEmbs, const float *threshold_values, int nThresholds,
                       int distance_type, int *error) {
    host_api<float>(0, output, input_data, data_size, tau_values, nTaus, emb_values, nEmbs, nullptr, 0,
                    threshold_values, nThresholds, distance_type, error);
}
void accrqa_amd_RR_f64(double *output, const double *input_data, size_t data_size, const int *tau_values, int nTaus,
                       const int *emb_values, int nEmbs, const double *threshold_values, int nThresholds,
                       int distance_type, int *error) {
    host_api<double>(0, output, input_data, data_size, tau_values, nTaus, emb_values, nEmbs, nullptr, 0,
                     threshold_values, nThresholds, distance_type, error);
}
void accrqa_amd_DET_f32(float *output, const float *input_data, size_t data_size, const int *tau_values, int nTaus,
                        const int *emb_values, int nEmbs, const int *lmin_values, int nLmins,
                        const float *threshold_values, int nThresholds, int distance_type, int calc_ENTR,
                        int *error) {
    (void)calc_ENTR;
    host_api<float>(1, output, input_data, data_size, tau_values, nTaus, emb_values, nEmbs, lmin_values, nLmins,
                    threshold_values, nThresholds, distance_type, error);
}
void accrqa_amd_DET_f64(double *output, const double *input_data, size_t data_size, const int *tau_values,
                        int nTaus, const int *emb_values, int nEmbs, const int *lmin_values, int nLmins,
                        const double *threshold_values, int nThresholds, int distance_type, int calc_ENTR,
                        int *error) {
    (void)calc_ENTR;
    host_api<double>(1, output, input_data, data_size, tau_values, nTaus, emb_values, nEmbs, lmin_values, nLmins,
                     threshold_values, nThresholds, distance_type, error);
}
void accrqa_amd_LAM_f32(float *output, const float *input_data, size_t data_size, const int *tau_values, int nTaus,
                        const int *emb_values, int nEmbs, const int *vmin_values, int nVmins,
                        const float *threshold_values, int nThresholds, int distance_type, int calc_ENTR,
                        int *error) {
    (void)calc_ENTR;
    host_api<float>(2, output, input_data, data_size, tau_values, nTaus, emb_values, nEmbs, vmin_values, nVmins,
                    threshold_values, nThresholds, distance_type, error);
}
void accrqa_amd_LAM_f64(double *output, const double *input_data, size_t data_size, const int *tau_values,
                        int nTaus, const int *emb_values, int nEmbs, const int *vmin_values, int nVmins,
                        const double *threshold_values, int nThresholds, int distance_type, int calc_ENTR,
                        int *error) {
    (void)calc_ENTR;
    host_api<double>(2, output, input_data, data_size, tau_values, nTaus, emb_values, nEmbs, vmin_values, nVmins,
                     threshold_values, nThresholds, distance_type, error);
}

void accrqa_amd_histograms_f32(const float *input_data, size_t data_size, int tau, int emb,
                               const float *threshold_values, int nThresholds, int distance_type, int which,
                               unsigned long long *rr_count, unsigned long long *hist_det,
                               unsigned long long *hist_lam, int *error) {
    host_hist<float>(input_data, data_size, tau, emb, threshold_values, nThresholds, distance_type, which, rr_count,
                     hist_det, hist_lam, error);
}
void accrqa_amd_histograms_f64(const double *input_data, size_t data_size, int tau, int emb,
                               const double *threshold_values, int nThresholds, int distance_type, int which,
                               unsigned long long *rr_count, unsigned long long *hist_det,
                               unsigned long long *hist_lam, int *error) {
    host_hist<double>(input_data, data_size, tau, emb, threshold_values, nThresholds, distance_type, which, rr_count,
                      hist_det, hist_lam, error);
}

int accrqa_amd_rqa_device_f32(const float *d_input, size_t data_size, int tau, int emb,
                              const float *threshold_values, int nThresholds, const int *lmin_values, int nL,
                              int distance_type, float *d_rr, float *d_det, float *d_lam, void *stream) {
    return device_api<float>(d_input, data_size, tau, emb, threshold_values, nThresholds, lmin_values, nL,
                             distance_type, d_rr, d_det, d_lam, stream);
}
int accrqa_amd_rqa_device_f64(const double *d_input, size_t data_size, int tau, int emb,
                              const double *threshold_values, int nThresholds, const int *lmin_values, int nL,
                              int distance_type, double *d_rr, double *d_det, double *d_lam, void *stream) {
    return device_api<double>(d_input, data_size, tau, emb, threshold_values, nThresholds, lmin_values, nL,
                              distance_type, d_rr, d_det, d_lam, stream);
}

int accrqa_amd_fp64_mismatch_f32(const float *d_input, size_t data_size, int tau, int emb,
                                 const float *threshold_values, int nThresholds, int distance_type,
                                 unsigned long long *mismatches, void *stream) {
    return device_fp64_diag(d_input, data_size, tau, emb, threshold_values, nThresholds, distance_type, mismatches,
                            stream);
}

const char *accrqa_amd_last_error(void) { return g_last_error.c_str(); }

int accrqa_amd_set_form(int form) {
    if (form < 0 || form > 2) return E_INVALID;
    std::lock_guard<std::mutex> lk(g_mu);
    g_form = form;
    return E_OK;
}

int accrqa_amd_set_slab_rows(int rows) {
    if (rows < 0 || (rows > 0 && (rows < 256 || rows % rqa::WAVE != 0 || rows > (1 << 29)))) return E_INVALID;
    std::lock_guard<std::mutex> lk(g_mu);
    g_slab_rows = rows;
    return E_OK;
}

int accrqa_amd_set_split(int mode) {
    if (mode < 0 || mode > 1) return E_INVALID;
    std::lock_guard<std::mutex> lk(g_mu);
    g_split = mode;
    return E_OK;
}

int accrqa_amd_mx_stats(int enable) {
    std::lock_guard<std::mutex> lk(g_mu);
    int e = init_ctx();
    if (e) return e;
    CallScope call(g_ctx.stream);
    if ((e = call.err)) return e;
    if ((e = ensure(g_ctx.mxstats, 3 * sizeof(unsigned long long)))) return e;
    fill_zero((unsigned long long *)g_ctx.mxstats.p, 3, g_ctx.stream);
    HIP_TRY(hipStreamSynchronize(g_ctx.stream));
    g_mxstats = enable != 0;
    return E_OK;
}

int accrqa_amd_mx_stats_read(unsigned long long *counts) {
    if (!counts) return E_NULL;
    std::lock_guard<std::mutex> lk(g_mu);
    counts[0] = counts[1] = counts[2] = 0;
    if (!g_ctx.mxstats.p) return E_OK;
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(counts, g_ctx.mxstats.p, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    return E_OK;
}

void accrqa_amd_RP_f32(char *output, const float *input_data, size_t data_size, int tau, int emb, float threshold,
                       int distance_type, int *error) {
    host_rp<float, char>(output, input_data, data_size, tau, emb, threshold, distance_type, error);
}
void accrqa_amd_RP_f64(char *output, const double *input_data, size_t data_size, int tau, int emb,
                       double threshold, int distance_type, int *error) {
    host_rp<double, char>(output, input_data, data_size, tau, emb, threshold, distance_type, error);
}
void accrqa_amd_RP_f64_i32(int *output, const double *input_data, size_t data_size, int tau, int emb,
                           double threshold, int distance_type, int *error) {
    host_rp<double, int>(output, input_data, data_size, tau, emb, threshold, distance_type, error);
}
int accrqa_amd_rp_device_f32(const float *d_input, size_t data_size, int tau, int emb, float threshold,
                             int distance_type, char *d_output, void *stream) {
    return device_rp<float, char>(d_input, data_size, tau, emb, threshold, distance_type, d_output, stream);
}
int accrqa_amd_rp_device_f64(const double *d_input, size_t data_size, int tau, int emb, double threshold,
                             int distance_type, char *d_output, void *stream) {
    return device_rp<double, char>(d_input, data_size, tau, emb, threshold, distance_type, d_output, stream);
}

int accrqa_amd_set_rp_wide(int on) {
    if (on < 0 || on > 1) return E_INVALID;
    std::lock_guard<std::mutex> lk(g_mu);
    g_rp_wide = on != 0;
    return E_OK;
}

int accrqa_amd_set_rp_sym(int on) {
    if (on < 0 || on > 1) return E_INVALID;
    std::lock_guard<std::mutex> lk(g_mu);
    g_rp_sym = on != 0;
    return E_OK;
}

int accrqa_amd_set_overlap(int on) {
    if (on < 0 || on > 1) return E_INVALID;
    std::lock_guard<std::mutex> lk(g_mu);
    g_overlap = on;
    return E_OK;
}

int accrqa_amd_set_scan_waves(int waves) {
    if (waves < 0 || waves > 8) return E_INVALID;
    std::lock_guard<std::mutex> lk(g_mu);
    g_scan_waves = waves;
    return E_OK;
}

int accrqa_amd_set_gram(int gram) {
    if (gram < rqa::GRAM_HILO || gram > rqa::GRAM_FP8) return E_INVALID;
    std::lock_guard<std::mutex> lk(g_mu);
    g_gram = gram;
    return E_OK;
}

int accrqa_amd_gram_probe(int gram, const float *a, const float *b, const float *c, float *d, float *qa,
                          float *qb) {
    if (!a || !b || !c || !d || !qa || !qb) return E_NULL;
    if (gram < rqa::GRAM_F16 || gram > rqa::GRAM_FP8) return E_INVALID;
    std::lock_guard<std::mutex> lk(g_mu);
    int e = init_ctx();
    if (e) return e;
    // a[32x16], b[16x32] at 0 and 2 KiB, c at 4 KiB, d at 8 KiB, qa at 12 KiB, qb at 14 KiB
    if ((e = ensure(g_ctx.probe, 16384))) return e;
    char *p = (char *)g_ctx.probe.p;
    hipStream_t st = g_ctx.stream;
    CallScope call(st);
    if ((e = call.err)) return e;
    HIP_TRY(hipMemcpyAsync(p, a, 2048, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(p + 2048, b, 2048, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(p + 4096, c, 4096, hipMemcpyHostToDevice, st));
    auto args = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(1), dim3(rqa::WAVE), 0, st, (const float *)p, (const float *)(p + 2048),
                           (const float *)(p + 4096), (float *)(p + 8192), (float *)(p + 12288),
                           (float *)(p + 14336));
    };
    if (gram == rqa::GRAM_F16) args(rqa::k_gram_probe<rqa::GRAM_F16>);
    else if (gram == rqa::GRAM_BF16) args(rqa::k_gram_probe<rqa::GRAM_BF16>);
    else args(rqa::k_gram_probe<rqa::GRAM_FP8>);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(d, p + 8192, 4096, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(qa, p + 12288, 2048, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(qb, p + 14336, 2048, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return E_OK;
}

int accrqa_amd_mfma16_probe(const uint16_t *a, const uint16_t *b, const float *c, float *d) {
    if (!a || !b || !c || !d) return E_NULL;
    std::lock_guard<std::mutex> lk(g_mu);
    int e = init_ctx();
    if (e) return e;
    // a[32x16] and b[16x32] (fp16 bits) at 0 and 1 KiB, c at 2 KiB, d at 6 KiB
    if ((e = ensure(g_ctx.probe, 10240))) return e;
    char *p = (char *)g_ctx.probe.p;
    hipStream_t st = g_ctx.stream;
    CallScope call(st);
    if ((e = call.err)) return e;
    HIP_TRY(hipMemcpyAsync(p, a, 1024, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(p + 1024, b, 1024, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(p + 2048, c, 4096, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(rqa::k_mfma16_probe, dim3(1), dim3(rqa::WAVE), 0, st, (const uint16_t *)p,
                       (const uint16_t *)(p + 1024), (const float *)(p + 2048), (float *)(p + 6144));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(d, p + 6144, 4096, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return E_OK;
}

void accrqa_amd_profile(int enable) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_prof.on = enable != 0;
}

void accrqa_amd_profile_reset(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    prof_fold();
    g_prof.acc.clear();
}

int accrqa_amd_profile_read(const char *kernel, double *total_ms, long long *launches) {
    std::lock_guard<std::mutex> lk(g_mu);
    prof_fold();
    for (auto &kv : g_prof.acc)
        if (kernel && kv.first == kernel) {
            if (total_ms) *total_ms = kv.second.first;
            if (launches) *launches = kv.second.second;
            return 0;
        }
    if (total_ms) *total_ms = 0.0;
    if (launches) *launches = 0;
    return 1;
}

const char *accrqa_amd_version(void) { return "0.1.0"; }

}  // extern "C"
