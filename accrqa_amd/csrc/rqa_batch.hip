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
#include <cstdio>
#include <cstring>
#include <map>
#include <numeric>
#include <string>
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
size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

// One chunk of the call, planned on the host: its descriptor blob (queue
// counter, item prefix, problems, block outputs, rows, l_min list) and what
// its launches need.
struct ChunkPlan {
    std::vector<char> blob;
    size_t o_ipre = 16, o_probs = 0, o_outs = 0, o_rows = 0, o_lm = 0;
    int np = 0;
    int64_t nitems = 0, bins = 0, nfin = 0;
    double pairs = 0;     // pair decisions of its blocks (N_c^2 per block and threshold group)
    size_t dev_off = 0;   // the blob's place in the call's descriptor buffer
};

template <typename F, int TG>
ChunkPlan plan_chunk(const std::vector<SBlock<F>> &blocks, size_t s0, size_t s1, int T, const int *lm, int nl,
                     int kinds) {
    ChunkPlan C;
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
        std::memset(&o, 0, sizeof o);   // padding too: the blob's bytes key the graph cache
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
            std::memset(&P, 0, sizeof P);
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
            C.pairs += (double)B.nc * (double)B.nc;
            probs.push_back(P);
            const int64_t ni = ((kinds & 1) ? rqa::batch_lam_items(B.nc) : 0) +
                               ((kinds & 2) ? rqa::batch_det_items(B.nc) : 0);
            ipre.push_back(ipre.back() + ni);
        }
    }
    C.np = (int)probs.size();
    C.nitems = ipre.back();
    C.bins = bins;
    C.nfin = (int64_t)outs.size() * T;
    C.o_probs = align16(C.o_ipre + ipre.size() * sizeof(int64_t));
    C.o_outs = align16(C.o_probs + probs.size() * sizeof(rqa::BatchProb<F, TG>));
    C.o_rows = align16(C.o_outs + outs.size() * sizeof(rqa::BatchBlockOut));
    C.o_lm = align16(C.o_rows + rows.size() * sizeof(longlong2));
    C.blob.assign(align16(C.o_lm + (size_t)std::max(nl, 1) * sizeof(int)), 0);
    char *h = C.blob.data();
    std::memcpy(h + C.o_ipre, ipre.data(), ipre.size() * sizeof(int64_t));
    if (C.np) std::memcpy(h + C.o_probs, probs.data(), probs.size() * sizeof(rqa::BatchProb<F, TG>));
    std::memcpy(h + C.o_outs, outs.data(), outs.size() * sizeof(rqa::BatchBlockOut));
    std::memcpy(h + C.o_rows, rows.data(), rows.size() * sizeof(longlong2));
    if (nl > 0) std::memcpy(h + C.o_lm, lm, (size_t)nl * sizeof(int));
    return C;
}

// ---------------------------------------------------------------- autotuner
// k_batch comes in instantiations that differ in the waves per SIMD their
// register allocation must allow (the compiler's choice, 4, 5, 6).  The first
// call of a (precision, norm, threshold group) with at least TUNE_MIN_PAIRS
// pair decisions in its first chunk runs that chunk once with each
// instantiation, timed with HIP events, and keeps the fastest for the rest of
// the process; every run recomputes the chunk from zeroed histograms, so the
// results are those of any one run.  accrqa_amd_batch_report() gives the
// choices.
constexpr int NVAR = 4;
constexpr int VAR_WPE[NVAR] = {1, 4, 5, 6};
constexpr double TUNE_MIN_PAIRS = 2e9;

struct TuneEntry {
    int variant = -1;
    float ms[NVAR] = {};
};
std::map<int, TuneEntry> g_tune;
bool g_autotune = true;
std::string g_tune_report;

int tune_key(int fsize, int norm, int tg) { return fsize * 10000 + norm * 100 + tg; }

// Calls f with the k_batch instantiation of variant var.
template <typename F, int NORM, int TG, typename Fn>
auto with_variant(int var, Fn &&f) {
    switch (var) {
        case 1: return f(rqa::k_batch<F, NORM, TG, 4>);
        case 2: return f(rqa::k_batch<F, NORM, TG, 5>);
        case 3: return f(rqa::k_batch<F, NORM, TG, 6>);
        default: return f(rqa::k_batch<F, NORM, TG, 1>);
    }
}

template <typename F, int NORM, int TG>
int launch_batch_variant(int var, unsigned g, const char *d, const ChunkPlan &C, int kinds, uint32_t *hist,
                         hipStream_t st) {
    const F *x = (const F *)g_ctx.x.p;
    const auto *probs = (const rqa::BatchProb<F, TG> *)(d + C.o_probs);
    const auto *ipre = (const int64_t *)(d + C.o_ipre);
    auto *queue = (unsigned long long *)d;
    with_variant<F, NORM, TG>(var, [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(g), dim3(rqa::WG), 0, st, x, probs, ipre, C.np, kinds, queue, hist);
        return 0;
    });
    HIP_TRY(hipGetLastError());
    return E_OK;
}

template <typename F, int NORM, int TG>
unsigned batch_grid(int var, int64_t nitems) {
    return with_variant<F, NORM, TG>(var, [&](auto kern) { return grid_for(kern, nitems, g_ctx.ncu); });
}

void fill_launch(uint32_t *hist, int64_t bins, unsigned long long *queue, hipStream_t st) {
    const unsigned fb = (unsigned)std::min<int64_t>((bins + 255) / 256 + 1, 8192);
    hipLaunchKernelGGL(rqa::k_fill_zero<uint32_t>, dim3(fb), dim3(256), 0, st, hist, (size_t)bins, queue);
}

// The launches of one chunk whose descriptors are at d (device).
template <typename F, int NORM, int TG>
int chunk_launches(int var, const char *d, const ChunkPlan &C, int T, int nl, int kinds, F *d_rr, F *d_det, F *d_lam,
                   hipStream_t st) {
    uint32_t *hist = (uint32_t *)g_ctx.bhist.p;
    fill_launch(hist, C.bins, (unsigned long long *)d, st);
    HIP_TRY(hipGetLastError());
    int e;
    if (C.nitems > 0 &&
        (e = launch_batch_variant<F, NORM, TG>(var, batch_grid<F, NORM, TG>(var, C.nitems), d, C, kinds, hist, st)))
        return e;
    hipLaunchKernelGGL((rqa::k_batch_fin<F>), dim3((unsigned)C.nfin), dim3(rqa::BFIN), 0, st, (const uint32_t *)hist,
                       (const rqa::BatchBlockOut *)(d + C.o_outs), (const longlong2 *)(d + C.o_rows), T,
                       (const int *)(d + C.o_lm), nl, d_rr, d_det, d_lam);
    HIP_TRY(hipGetLastError());
    return E_OK;
}

// Times every instantiation on chunk C (descriptors at d) and records the
// fastest; the histograms are left as the last run wrote them.
template <typename F, int NORM, int TG>
int autotune_chunk(TuneEntry &E, const char *d, const ChunkPlan &C, int kinds, hipStream_t st) {
    uint32_t *hist = (uint32_t *)g_ctx.bhist.p;
    hipEvent_t ev[2 * NVAR];
    for (int i = 0; i < 2 * NVAR; i++) HIP_TRY(hipEventCreate(&ev[i]));
    for (int v = 0; v < NVAR; v++) {
        fill_launch(hist, C.bins, (unsigned long long *)d, st);
        HIP_TRY(hipEventRecord(ev[2 * v], st));
        int e = launch_batch_variant<F, NORM, TG>(v, batch_grid<F, NORM, TG>(v, C.nitems), d, C, kinds, hist, st);
        if (e) return e;
        HIP_TRY(hipEventRecord(ev[2 * v + 1], st));
    }
    HIP_TRY(hipStreamSynchronize(st));
    int best = 0;
    for (int v = 0; v < NVAR; v++) {
        HIP_TRY(hipEventElapsedTime(&E.ms[v], ev[2 * v], ev[2 * v + 1]));
        if (E.ms[v] < E.ms[best]) best = v;
    }
    for (int i = 0; i < 2 * NVAR; i++) (void)hipEventDestroy(ev[i]);
    E.variant = best;
    std::string line = std::string("k_batch<") + (sizeof(F) == 4 ? "f32" : "f64") + "," +
                       (NORM == 1 ? "euclidean" : "maximal") + "," + std::to_string(TG) + ">: waves/SIMD";
    char part[64];
    for (int v = 0; v < NVAR; v++) {
        if (v == 0) std::snprintf(part, sizeof part, " auto %.2f ms", E.ms[v]);
        else std::snprintf(part, sizeof part, ", %d %.2f ms", VAR_WPE[v], E.ms[v]);
        line += part;
    }
    line += " -> " + (best == 0 ? std::string("auto") : std::to_string(VAR_WPE[best])) + "\n";
    g_tune_report += line;
    return E_OK;
}

// ---------------------------------------------------------------- graph
// The launches of a call (all chunks) are captured in a hipGraph and kept
// with a key of everything they depend on: the descriptor blobs (problems,
// bounds, offsets), the instantiation and the device buffers' addresses.  A
// later call with the same key (the same shapes and thresholds, any series
// values) replays the graph: one launch.  The graph holds only kernels (the
// histograms and the queue counter are zeroed by a kernel).  The graph reads
// its descriptors from the shared descriptor buffer, so a replay uploads them
// again first unless the buffer still holds that key's blobs (g_desc_key;
// every upload of other descriptors clears it).  Up to GRAPH_SLOTS graphs are
// kept, the least recently used replaced first.
constexpr int GRAPH_SLOTS = 4;

struct GraphSlot {
    uint64_t key = 0;
    uint64_t used = 0;
    uint64_t kernels = 0;   // kernel nodes
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};
struct GraphCache {
    GraphSlot slot[GRAPH_SLOTS];
    uint64_t tick = 0;
    uint64_t replays = 0, builds = 0;
} g_graph;
bool g_use_graphs = true;
uint64_t g_desc_key = 0;   // key whose blobs the descriptor buffer holds (0: none known)

// Counters of the batched calls (accrqa_amd_batch_report).
struct BatchCounts {
    uint64_t calls = 0, chunks = 0, kernels = 0, graph_launches = 0, desc_uploads = 0;
} g_counts;

uint64_t fnv(uint64_t h, const void *p, size_t n) {
    const unsigned char *b = (const unsigned char *)p;
    for (size_t i = 0; i < n; i++) {
        h ^= b[i];
        h *= 1099511628211ull;
    }
    return h;
}

void slot_drop(GraphSlot &S) {
    if (S.exec) (void)hipGraphExecDestroy(S.exec);
    if (S.graph) (void)hipGraphDestroy(S.graph);
    S = GraphSlot{};
}

// Drops every kept graph; the caller has synchronised the library's stream.
void graph_drop() {
    for (GraphSlot &S : g_graph.slot) slot_drop(S);
}

GraphSlot *graph_find(uint64_t key) {
    for (GraphSlot &S : g_graph.slot)
        if (S.exec && S.key == key) return &S;
    return nullptr;
}

template <typename F, int NORM, int TG>
int run_small_tg(const std::vector<SBlock<F>> &blocks, int T, const int *lm, int nl, int kinds, F *d_rr, F *d_det,
                 F *d_lam, hipStream_t st) {
    Ctx &c = g_ctx;
    // chunks: blocks up to the histogram budget (at least one)
    std::vector<ChunkPlan> plans;
    size_t s0 = 0;
    while (s0 < blocks.size()) {
        int64_t bins = 0;
        size_t s1 = s0;
        while (s1 < blocks.size()) {
            const int64_t bb = 2 * (int64_t)blocks[s1].b.size() * (blocks[s1].nc + 1);
            if (s1 > s0 && bins + bb > CHUNK_BINS) break;
            bins += bb;
            s1++;
        }
        plans.push_back(plan_chunk<F, TG>(blocks, s0, s1, T, lm, nl, kinds));
        s0 = s1;
    }
    size_t dbytes = 0;
    int64_t maxbins = 1;
    for (ChunkPlan &C : plans) {
        C.dev_off = dbytes;
        dbytes += align256(C.blob.size());
        maxbins = std::max(maxbins, C.bins);
    }
    int e;
    if ((e = ensure(c.bdesc, dbytes)) || (e = ensure(c.bhist, (size_t)maxbins * sizeof(uint32_t)))) return e;
    char *dbase = (char *)c.bdesc.p;
    static const void *desc_buf = nullptr;   // the buffer g_desc_key describes
    if (desc_buf != c.bdesc.p) {
        g_desc_key = 0;
        desc_buf = c.bdesc.p;
    }

    TuneEntry &TE = g_tune[tune_key((int)sizeof(F), NORM, TG)];
    int var = TE.variant >= 0 ? TE.variant : 0;
    g_counts.calls++;
    g_counts.chunks += plans.size();
    auto upload_all = [&](size_t from) -> int {
        int ue = E_OK;
        for (size_t i = from; i < plans.size() && !ue; i++) {
            const ChunkPlan &C = plans[i];
            ue = stage_upload(dbase + C.dev_off + 16, C.blob.data() + 16, C.blob.size() - 16, st);
        }
        g_counts.desc_uploads++;
        return ue;
    };
    auto eager = [&](size_t from) -> int {
        int le = E_OK;
        for (size_t i = from; i < plans.size() && !le; i++) {
            const ChunkPlan &C = plans[i];
            le = chunk_launches<F, NORM, TG>(var, dbase + C.dev_off, C, T, nl, kinds, d_rr, d_det, d_lam, st);
            g_counts.kernels += C.nitems > 0 ? 3 : 2;
        }
        return le;
    };
    // the first chunk with enough work tunes the instantiation (eagerly)
    if (g_autotune && TE.variant < 0) {
        const ChunkPlan &C0 = plans[0];
        if (C0.pairs >= TUNE_MIN_PAIRS && C0.nitems > 0) {
            g_desc_key = 0;
            if ((e = upload_all(0))) return e;
            if ((e = autotune_chunk<F, NORM, TG>(TE, dbase + C0.dev_off, C0, kinds, st))) return e;
            var = TE.variant;
            g_counts.kernels += 2 * NVAR;
            // finish the tuned chunk, then the rest eagerly
            const char *d0 = dbase + C0.dev_off;
            hipLaunchKernelGGL((rqa::k_batch_fin<F>), dim3((unsigned)C0.nfin), dim3(rqa::BFIN), 0, st,
                               (const uint32_t *)c.bhist.p, (const rqa::BatchBlockOut *)(d0 + C0.o_outs),
                               (const longlong2 *)(d0 + C0.o_rows), T, (const int *)(d0 + C0.o_lm), nl, d_rr, d_det,
                               d_lam);
            HIP_TRY(hipGetLastError());
            g_counts.kernels++;
            return eager(1);
        }
    }
    // key of the call's launches
    uint64_t key = 1469598103934665603ull;
    const int params[6] = {(int)sizeof(F), NORM, TG, var, T, nl};
    key = fnv(key, params, sizeof params);
    key = fnv(key, &kinds, sizeof kinds);
    const void *ptrs[6] = {c.x.p, c.bhist.p, c.bdesc.p, d_rr, d_det, d_lam};
    key = fnv(key, ptrs, sizeof ptrs);
    for (const ChunkPlan &C : plans) key = fnv(key, C.blob.data(), C.blob.size());
    if (key == 0) key = 1;
    if (g_desc_key != key) {
        g_desc_key = 0;
        if ((e = upload_all(0))) return e;
        g_desc_key = key;
    }
    if (!g_use_graphs) return eager(0);
    if (GraphSlot *S = graph_find(key)) {
        HIP_TRY(hipGraphLaunch(S->exec, st));
        S->used = ++g_graph.tick;
        g_graph.replays++;
        g_counts.graph_launches++;
        g_counts.kernels += S->kernels;
        return E_OK;
    }
    // the least recently used slot; its graph may still be running.  The
    // grid sizes are looked up (and cached) before the capture starts.
    GraphSlot *S = &g_graph.slot[0];
    for (GraphSlot &Q : g_graph.slot)
        if (Q.used < S->used) S = &Q;
    if (S->exec) HIP_TRY(hipStreamSynchronize(st));
    slot_drop(*S);
    for (const ChunkPlan &C : plans)
        if (C.nitems > 0) (void)batch_grid<F, NORM, TG>(var, C.nitems);
    HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
    int ce = E_OK;
    uint64_t kern = 0;
    for (const ChunkPlan &C : plans) {
        if (!ce) ce = chunk_launches<F, NORM, TG>(var, dbase + C.dev_off, C, T, nl, kinds, d_rr, d_det, d_lam, st);
        kern += C.nitems > 0 ? 3 : 2;
    }
    hipGraph_t graph = nullptr;
    const hipError_t he = hipStreamEndCapture(st, &graph);
    if (ce) {
        if (graph) (void)hipGraphDestroy(graph);
        return ce;
    }
    HIP_TRY(he);
    hipGraphExec_t exec = nullptr;
    const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
        (void)hipGraphDestroy(graph);
        return map_err(ie);
    }
    S->graph = graph;
    S->exec = exec;
    S->key = key;
    S->used = ++g_graph.tick;
    S->kernels = kern;
    g_graph.builds++;
    HIP_TRY(hipGraphLaunch(exec, st));
    g_counts.graph_launches++;
    g_counts.kernels += kern;
    return E_OK;
}

template <typename F, int NORM>
int run_small(const std::vector<SBlock<F>> &blocks, int T, const int *lm, int nl, int kinds, F *d_rr, F *d_det,
              F *d_lam, hipStream_t st) {
    int maxlive = 0;
    for (const SBlock<F> &B : blocks) maxlive = std::max(maxlive, (int)B.b.size());
    return maxlive <= 8 ? run_small_tg<F, NORM, 8>(blocks, T, lm, nl, kinds, d_rr, d_det, d_lam, st)
                        : run_small_tg<F, NORM, 16>(blocks, T, lm, nl, kinds, d_rr, d_det, d_lam, st);
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

int accrqa_amd_batch_options(int autotune, int graphs) {
    if (autotune < -1 || autotune > 1 || graphs < -1 || graphs > 1) return eng::E_INVALID;
    std::lock_guard<std::mutex> lk(eng::g_mu);
    if (autotune >= 0) eng::g_autotune = autotune != 0;
    if (graphs >= 0) {
        eng::g_use_graphs = graphs != 0;
        if (!eng::g_use_graphs) {
            if (eng::g_ctx.stream) (void)hipStreamSynchronize(eng::g_ctx.stream);
            eng::graph_drop();
        }
    }
    return eng::E_OK;
}

const char *accrqa_amd_batch_report(void) {
    std::lock_guard<std::mutex> lk(eng::g_mu);
    static std::string out;
    char line[320];
    const eng::BatchCounts &K = eng::g_counts;
    std::snprintf(line, sizeof line,
                  "batch: %llu calls, %llu chunks, %llu kernels run, %llu graph launches, %llu descriptor uploads\n"
                  "graph: %llu built, %llu replays\n",
                  (unsigned long long)K.calls, (unsigned long long)K.chunks, (unsigned long long)K.kernels,
                  (unsigned long long)K.graph_launches, (unsigned long long)K.desc_uploads,
                  (unsigned long long)eng::g_graph.builds, (unsigned long long)eng::g_graph.replays);
    out = eng::g_tune_report + line;
    return out.c_str();
}

}  // extern "C"

