// batch.hip -- the batching form: length-ordered chunks of reads, each run as
// encode -> MFMA search -> finalize -> scatter and replayed from a captured
// hipGraph; an autotuner over the search form, reference slices and chunk
// size.
//
// Reads are ordered by length, longest first (a counting sort of the host
// offsets), and cut into chunks.  The encoder reads chunk rows through
// (start, length) pairs in that order, so each launch holds reads of similar
// length and its workgroups (a static round-robin over sequences) finish
// together.  The chunk's rows live in batch-owned buffers; after the search,
// a scatter kernel writes each row's top-k to the read's own output row.
//
// Graph capture.  The graphs touch only batch-owned buffers and the
// database: each call first copies the caller's packed reads into the
// batch's staging buffers, and after each chunk's graph a scatter launched
// outside it writes the chunk's results to the caller's output rows.  So the
// graphs do not depend on the caller's read or output buffers, and a stream
// of calls with fresh buffers replays the same graphs.  A graph's key is the
// chunk's position and size, k, the settings, the encoder's counter width
// class of its longest read, the database pointers, and the addresses of the
// batch's buffers and of the private context's workspaces (which only move
// when they grow).  The first call that meets a key runs the chunk directly
// -- growing every workspace to its final size -- and then captures the same
// launches on the batch's stream into a hipGraph; later calls replay it.  At
// most kMaxGraphs graphs are kept, the least recently used dropped first.
// The private context is used by nothing else, so the workspace addresses
// the graphs hold stay valid.  Nothing in a chunk uses hipMemsetAsync
// (counters are cleared by kernels).
#include <algorithm>
#include <cstring>
#include <vector>

#include "search_impl.h"

namespace {

constexpr int kMinChunk = 256;
constexpr size_t kMaxGraphs = 64;

struct GraphKey {
    int64_t c0;
    int n;
    int k;
    int form, shape, splits;
    int width;          // counter width class of the chunk's longest read
    hvk_bundles db;
    const void* in;     // the batch's buffers (move when they grow)
    const void* order;
    const void* rows;
    const void* outs;
    const void* ws[5];  // the private context's search workspaces
    bool operator==(const GraphKey& o) const { return std::memcmp(this, &o, sizeof(*this)) == 0; }
};

struct Graph {
    GraphKey key;
    hipGraphExec_t exec;
};

// A descent chunk's graph: besides the chunk and the batch's buffers, every
// scalar the launches take from the plan (the candidate lists' contents are
// uploaded by each call into the same buffer and read at replay).
struct DescKey {
    int64_t c0;
    int n, width;
    int E, B, L, dense, n_cat, n_roots, Fmax, maxN, max_span;
    uint32_t thr_bits;
    int64_t max_abs_proto, tld;
    const void* tree[7];  // lo, hi, sqnorm, row maxima, child_off, child_idx, roots
    const void* in;
    const void* order;
    const void* rows;
    const void* outs;
    const void* ws;  // the descent workspace
    bool operator==(const DescKey& o) const { return std::memcmp(this, &o, sizeof(*this)) == 0; }
};

struct DescGraph {
    DescKey key;
    hipGraphExec_t exec;
};

}  // namespace

struct hvk_batch {
    hvk_ctx* ctx = nullptr;
    int form = HVK_BATCH_FORM_BOUNDED;
    int shape = 0;
    int splits = 0;
    int chunk = 0;
    hipStream_t s = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr, t0 = nullptr, t1 = nullptr;
    hvk::Workspace ws_in;     // staging copy of the caller's packed reads
    hvk::Workspace ws_rows;   // chunk rows and per-row sums
    hvk::Workspace ws_out;    // chunk outputs
    hvk::Workspace ws_order;  // perm, starts, lens of the whole batch
    std::vector<int32_t> perm, lens;
    std::vector<int64_t> starts;
    std::vector<int64_t> offsets;  // the offsets the order was computed from
    std::vector<Graph> graphs;
    std::vector<DescGraph> dgraphs;
};

namespace hvk {
namespace {

void drop_graphs(hvk_batch* b) {
    for (auto& g : b->graphs) (void)hipGraphExecDestroy(g.exec);
    b->graphs.clear();
    for (auto& g : b->dgraphs) (void)hipGraphExecDestroy(g.exec);
    b->dgraphs.clear();
}

// out[perm[i]][j] = chunk[i][j]; block 0 adds the chunk's re-scored count
__global__ void batch_scatter(const int32_t* __restrict__ perm, int n, int k,
                              const int32_t* __restrict__ ci, const float* __restrict__ cs,
                              const double* __restrict__ cc, int32_t* __restrict__ oi,
                              float* __restrict__ os, double* __restrict__ oc,
                              const int32_t* __restrict__ cnres, int32_t* __restrict__ onres) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t == 0 && onres) atomicAdd(onres, *cnres);
    if (t >= (int64_t)n * k) return;
    const int i = (int)(t / k), j = (int)(t - (int64_t)i * k);
    const int64_t o = (int64_t)perm[i] * k + j;
    oi[o] = ci[t];
    os[o] = cs[t];
    if (oc) oc[o] = cc[t];
}

__global__ void batch_zero(int32_t* p) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0;
}

struct ChunkBufs {
    int8_t* lo;
    int8_t* hi;
    int8_t* clamp8;
    int64_t* sqnorm;
    int32_t* maxabs;
    int32_t* resid1;
    int32_t* nkmers;
    int32_t* idx;
    float* score;
    double* cos64;
    int32_t* nres;
};

inline size_t al(size_t x) { return (x + 255) & ~size_t(255); }

int chunk_bufs(hvk_batch* b, int k, ChunkBufs& cb) {
    const int64_t ld = b->ctx->ld;
    const size_t m = (size_t)b->chunk;
    const size_t rows = 3 * al(m * ld) + al(m * 8) + 3 * al(m * 4);
    if (int e = b->ws_rows.ensure(rows)) return e;
    const size_t outs = al(m * k * 4) * 2 + al(m * k * 8) + 256;
    if (int e = b->ws_out.ensure(outs)) return e;
    char* p = (char*)b->ws_rows.ptr;
    auto take = [&](size_t n) { char* q = p; p += al(n); return q; };
    cb.lo = (int8_t*)take(m * ld);
    cb.hi = (int8_t*)take(m * ld);
    cb.clamp8 = (int8_t*)take(m * ld);
    cb.sqnorm = (int64_t*)take(m * 8);
    cb.maxabs = (int32_t*)take(m * 4);
    cb.resid1 = (int32_t*)take(m * 4);
    cb.nkmers = (int32_t*)take(m * 4);
    p = (char*)b->ws_out.ptr;
    cb.idx = (int32_t*)take(m * k * 4);
    cb.score = (float*)take(m * k * 4);
    cb.cos64 = (double*)take(m * k * 8);
    cb.nres = (int32_t*)take(256);
    return HVK_OK;
}

// the encoder's counter width class of a chunk whose longest read has
// max_len bases (launch_encode picks its bit-sliced counter width by it)
int width_class(int max_len, int k) {
    const int64_t nmax = (int64_t)max_len - k + 1;
    return nmax < (1 << 11) ? 0 : (nmax < (1 << 16) ? 1 : 2);
}

// the launches of one chunk up to its results in cb, on b->s (captured)
int run_chunk(hvk_batch* b, const uint32_t* packed, const uint32_t* invalid, int64_t c0, int n,
              int max_len, const hvk_bundles& db, int k, const ChunkBufs& cb,
              const int64_t* starts, const int32_t* lens) {
    hvk_ctx* ctx = b->ctx;
    const int64_t ld = ctx->ld;
    EncodeOut eo{};
    eo.lo = cb.lo;
    eo.hi = cb.hi;
    eo.clamp8 = cb.clamp8;
    eo.ld = ld;
    eo.sqnorm = cb.sqnorm;
    eo.maxabs = cb.maxabs;
    eo.nkmers = cb.nkmers;
    eo.resid1 = cb.resid1;
    if (int e = launch_encode(ctx, packed, invalid, nullptr, starts + c0, 0, n, max_len, eo, b->s,
                              lens + c0))
        return e;
    ctx->tune_splits = b->splits;
    ctx->tune_shape = b->shape;
    int e;
    if (b->form == HVK_BATCH_FORM_BOUNDED) {
        hvk_bundles q{cb.clamp8, cb.lo, cb.hi, cb.sqnorm, cb.maxabs, cb.resid1, n};
        e = hvk_search_mfma_bounded(ctx, &q, &db, ld, k, cb.idx, cb.score, cb.cos64, cb.nres, b->s);
    } else {
        e = hvk_search_mfma(ctx, cb.lo, cb.hi, cb.sqnorm, n, db.lo, db.hi, db.sqnorm, db.n, ld, k,
                            cb.idx, cb.score, cb.cos64, cb.nres, b->s);
    }
    return e;
}

// the chunk's results to the caller's rows (outside the graph)
int scatter_chunk(hvk_batch* b, int64_t c0, int n, int k, const ChunkBufs& cb,
                  const int32_t* perm, int32_t* oi, float* os, double* oc, int32_t* onres) {
    const int64_t tot = (int64_t)n * k;
    hipLaunchKernelGGL(batch_scatter, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, b->s,
                       perm + c0, n, k, cb.idx, cb.score, cb.cos64, oi, os, oc, cb.nres, onres);
    HVK_HIP(hipGetLastError());
    return HVK_OK;
}

// least recently used graphs beyond kMaxGraphs are destroyed (after the
// stream has finished any replay of them)
int trim_graphs(hvk_batch* b) {
    if (b->graphs.size() < kMaxGraphs) return HVK_OK;
    HVK_HIP(hipStreamSynchronize(b->s));
    while (b->graphs.size() >= kMaxGraphs) {
        (void)hipGraphExecDestroy(b->graphs.front().exec);
        b->graphs.erase(b->graphs.begin());
    }
    return HVK_OK;
}

// length order (longest first, ties by index) of the reads and their
// (start, length) pairs, uploaded into ws_order
int order_reads(hvk_batch* b, const int64_t* h_offsets, int n, int& max_len) {
    if ((int64_t)b->offsets.size() == (int64_t)n + 1 && b->ws_order.ptr &&
        std::memcmp(b->offsets.data(), h_offsets, ((size_t)n + 1) * 8) == 0) {
        max_len = n ? b->lens[0] : 0;  // same reads as the last call: order already uploaded
        return HVK_OK;
    }
    // the host arrays below may still be the source of the last call's copies
    HVK_HIP(hipStreamSynchronize(b->s));
    b->offsets.clear();
    b->perm.resize(n);
    b->starts.resize(n);
    b->lens.resize(n);
    max_len = 0;
    for (int i = 0; i < n; ++i) {
        const int64_t l = h_offsets[i + 1] - h_offsets[i];
        if (l < 0 || l > INT32_MAX) return HVK_ERR_ARG;
        max_len = std::max<int64_t>(max_len, l);
    }
    // counting sort by length, descending, stable
    std::vector<int64_t> cnt((size_t)max_len + 2, 0);
    for (int i = 0; i < n; ++i) ++cnt[(size_t)(max_len - (h_offsets[i + 1] - h_offsets[i]))];
    int64_t run = 0;
    for (auto& c : cnt) { const int64_t t = c; c = run; run += t; }
    for (int i = 0; i < n; ++i) {
        const int64_t l = h_offsets[i + 1] - h_offsets[i];
        const int64_t pos = cnt[(size_t)(max_len - l)]++;
        b->perm[pos] = i;
        b->starts[pos] = h_offsets[i];
        b->lens[pos] = (int32_t)l;
    }
    const size_t need = al((size_t)n * 4) * 2 + al((size_t)n * 8);
    if (int e = b->ws_order.ensure(need)) return e;
    char* p = (char*)b->ws_order.ptr;
    HVK_HIP(hipMemcpyAsync(p, b->perm.data(), (size_t)n * 4, hipMemcpyHostToDevice, b->s));
    HVK_HIP(hipMemcpyAsync(p + al((size_t)n * 4), b->lens.data(), (size_t)n * 4,
                           hipMemcpyHostToDevice, b->s));
    HVK_HIP(hipMemcpyAsync(p + 2 * al((size_t)n * 4), b->starts.data(), (size_t)n * 8,
                           hipMemcpyHostToDevice, b->s));
    b->offsets.assign(h_offsets, h_offsets + (size_t)n + 1);
    return HVK_OK;
}

int classify(hvk_batch* b, const uint32_t* caller_packed, const uint32_t* caller_invalid,
             const int64_t* h_offsets, int n, const hvk_bundles& db, int k, int32_t* oi,
             float* os, double* oc, int32_t* onres, hipStream_t caller) {
    HVK_HIP(hipEventRecord(b->ev_in, caller));
    HVK_HIP(hipStreamWaitEvent(b->s, b->ev_in, 0));
    int max_len = 0;
    if (int e = order_reads(b, h_offsets, n, max_len)) return e;
    // stage the packed reads in batch-owned buffers (the graphs read these)
    const size_t words = (size_t)hvk_pack_words(h_offsets[n]);
    if (int e = b->ws_in.ensure(2 * al(words * 4))) return e;
    uint32_t* packed = (uint32_t*)b->ws_in.ptr;
    uint32_t* invalid = (uint32_t*)((char*)b->ws_in.ptr + al(words * 4));
    HVK_HIP(hipMemcpyAsync(packed, caller_packed, words * 4, hipMemcpyDeviceToDevice, b->s));
    HVK_HIP(hipMemcpyAsync(invalid, caller_invalid, words * 4, hipMemcpyDeviceToDevice, b->s));
    const char* op = (const char*)b->ws_order.ptr;
    const int32_t* perm = (const int32_t*)op;
    const int32_t* lens = (const int32_t*)(op + al((size_t)n * 4));
    const int64_t* starts = (const int64_t*)(op + 2 * al((size_t)n * 4));
    ChunkBufs cb;
    if (int e = chunk_bufs(b, k, cb)) return e;
    if (onres) {
        hipLaunchKernelGGL(batch_zero, dim3(1), dim3(64), 0, b->s, onres);
        HVK_HIP(hipGetLastError());
    }
    for (int64_t c0 = 0; c0 < n; c0 += b->chunk) {
        const int m = (int)std::min<int64_t>(b->chunk, n - c0);
        const int chunk_max = b->lens[(size_t)c0];  // longest first
        auto make_key = [&]() {
            GraphKey key;
            std::memset(&key, 0, sizeof(key));
            key.c0 = c0;
            key.n = m;
            key.k = k;
            key.form = b->form;
            key.shape = b->shape;
            key.splits = b->splits;
            key.width = width_class(chunk_max, b->ctx->k);
            key.db.clamp8 = db.clamp8;  // field by field: the struct's padding stays zero
            key.db.lo = db.lo;
            key.db.hi = db.hi;
            key.db.sqnorm = db.sqnorm;
            key.db.maxabs = db.maxabs;
            key.db.resid1 = db.resid1;
            key.db.n = db.n;
            key.in = b->ws_in.ptr;
            key.order = b->ws_order.ptr;
            key.rows = b->ws_rows.ptr;
            key.outs = b->ws_out.ptr;
            key.ws[0] = b->ctx->ws_search.ptr;
            key.ws[1] = b->ctx->ws_norms.ptr;
            key.ws[2] = b->ctx->ws_flags.ptr;
            key.ws[3] = b->ctx->ws_search2.ptr;
            key.ws[4] = b->ctx->ws_select.ptr;
            return key;
        };
        GraphKey key = make_key();
        auto it = std::find_if(b->graphs.begin(), b->graphs.end(),
                               [&](const Graph& g) { return g.key == key; });
        if (it != b->graphs.end()) {
            HVK_HIP(hipGraphLaunch(it->exec, b->s));
            std::rotate(it, it + 1, b->graphs.end());  // most recently used last
            if (int e = scatter_chunk(b, c0, m, k, cb, perm, oi, os, oc, onres)) return e;
            continue;
        }
        // direct run (sizes the workspaces), then capture the same launches
        if (int e = run_chunk(b, packed, invalid, c0, m, chunk_max, db, k, cb, starts, lens))
            return e;
        if (int e = scatter_chunk(b, c0, m, k, cb, perm, oi, os, oc, onres)) return e;
        key = make_key();  // the run may have grown (moved) a workspace
        if (int e = trim_graphs(b)) return e;
        HVK_HIP(hipStreamBeginCapture(b->s, hipStreamCaptureModeThreadLocal));
        const int ce = run_chunk(b, packed, invalid, c0, m, chunk_max, db, k, cb, starts, lens);
        hipGraph_t g = nullptr;
        const hipError_t he = hipStreamEndCapture(b->s, &g);
        if (ce) {
            if (g) (void)hipGraphDestroy(g);
            return ce;
        }
        if (he != hipSuccess) { set_last_error("hipStreamEndCapture", he); return HVK_ERR_HIP; }
        hipGraphExec_t ex = nullptr;
        const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ie != hipSuccess) { set_last_error("hipGraphInstantiate", ie); return HVK_ERR_HIP; }
        b->graphs.push_back({key, ex});
    }
    HVK_HIP(hipEventRecord(b->ev_out, b->s));
    HVK_HIP(hipStreamWaitEvent(caller, b->ev_out, 0));
    return HVK_OK;
}

// ---- the descent in the batching form -----------------------------------
// Per length-ordered chunk: encode (int16 counts and norms) and the
// descent's launches, captured once and replayed; the chunk's paths go to
// the reads' own rows outside the graph.

// out[perm[i]] = chunk[i] for the path rows (L entries) and per-read values
__global__ void batch_desc_scatter(const int32_t* __restrict__ perm, int n, int L,
                                   const int32_t* __restrict__ cp, const float* __restrict__ cs,
                                   const int32_t* __restrict__ cl, const int32_t* __restrict__ ce,
                                   int32_t* __restrict__ op, float* __restrict__ os,
                                   int32_t* __restrict__ ol, int32_t* __restrict__ oe) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= (int64_t)n * L) return;
    const int i = (int)(t / L), j = (int)(t - (int64_t)i * L);
    const int64_t r = perm[i];
    op[r * L + j] = cp[t];
    os[r * L + j] = cs[t];
    if (j == 0) {
        ol[r] = cl[i];
        oe[r] = ce[i];
    }
}

struct DescBufs {
    int16_t* counts;
    int64_t* sqnorm;
    int32_t* maxabs;
    int32_t* nkmers;
    int32_t* path;
    float* pscore;
    int32_t* len;
    int32_t* exit;
};

int desc_bufs(hvk_batch* b, int L, DescBufs& cb) {
    const int64_t ld = b->ctx->ld;
    const size_t m = (size_t)b->chunk;
    if (int e = b->ws_rows.ensure(al(m * ld * 2) + al(m * 8) + 2 * al(m * 4))) return e;
    if (int e = b->ws_out.ensure(2 * al(m * L * 4) + 2 * al(m * 4))) return e;
    char* p = (char*)b->ws_rows.ptr;
    auto take = [&](size_t n) { char* q = p; p += al(n); return q; };
    cb.counts = (int16_t*)take(m * ld * 2);
    cb.sqnorm = (int64_t*)take(m * 8);
    cb.maxabs = (int32_t*)take(m * 4);
    cb.nkmers = (int32_t*)take(m * 4);
    p = (char*)b->ws_out.ptr;
    cb.path = (int32_t*)take(m * L * 4);
    cb.pscore = (float*)take(m * L * 4);
    cb.len = (int32_t*)take(m * 4);
    cb.exit = (int32_t*)take(m * 4);
    return HVK_OK;
}

int run_desc_chunk(hvk_batch* b, const uint32_t* packed, const uint32_t* invalid, int64_t c0,
                   int n, int max_len, const DescPlan& pl, float threshold, const DescBufs& cb,
                   const int64_t* starts, const int32_t* lens) {
    hvk_ctx* ctx = b->ctx;
    EncodeOut eo{};
    eo.counts = cb.counts;
    eo.ld = ctx->ld;
    eo.sqnorm = cb.sqnorm;
    eo.maxabs = cb.maxabs;
    eo.nkmers = cb.nkmers;
    if (int e = launch_encode(ctx, packed, invalid, nullptr, starts + c0, 0, n, max_len, eo, b->s,
                              lens + c0))
        return e;
    std::lock_guard<std::mutex> g(ctx->mu);
    return desc_launch(ctx, pl, cb.counts, cb.sqnorm, n, ctx->ld, threshold, cb.path, cb.pscore,
                       cb.len, cb.exit, b->s);
}

int descend(hvk_batch* b, const hvk_tree* tree, const uint32_t* caller_packed,
            const uint32_t* caller_invalid, const int64_t* h_offsets, int n, int beam,
            float threshold, int levels, int32_t* op, float* os, int32_t* ol, int32_t* oe,
            hipStream_t caller) {
    HVK_HIP(hipEventRecord(b->ev_in, caller));
    HVK_HIP(hipStreamWaitEvent(b->s, b->ev_in, 0));
    int max_len = 0;
    if (int e = order_reads(b, h_offsets, n, max_len)) return e;
    // |count| <= the read's valid k-mers <= its length - k + 1: the bound of
    // the longest read serves every chunk (one plan per call)
    const int q_bound = std::max(0, max_len - b->ctx->k + 1);
    DescPlan pl;
    {
        std::lock_guard<std::mutex> g(b->ctx->mu);
        if (int e = desc_prepare(b->ctx, tree, q_bound, beam, levels, pl, b->s)) return e;
    }
    const size_t words = (size_t)hvk_pack_words(h_offsets[n]);
    if (int e = b->ws_in.ensure(2 * al(words * 4))) return e;
    uint32_t* packed = (uint32_t*)b->ws_in.ptr;
    uint32_t* invalid = (uint32_t*)((char*)b->ws_in.ptr + al(words * 4));
    HVK_HIP(hipMemcpyAsync(packed, caller_packed, words * 4, hipMemcpyDeviceToDevice, b->s));
    HVK_HIP(hipMemcpyAsync(invalid, caller_invalid, words * 4, hipMemcpyDeviceToDevice, b->s));
    const char* ordp = (const char*)b->ws_order.ptr;
    const int32_t* perm = (const int32_t*)ordp;
    const int32_t* lens = (const int32_t*)(ordp + al((size_t)n * 4));
    const int64_t* starts = (const int64_t*)(ordp + 2 * al((size_t)n * 4));
    DescBufs cb;
    if (int e = desc_bufs(b, levels, cb)) return e;
    auto scatter = [&](int64_t c0, int m) -> int {
        const int64_t tot = (int64_t)m * levels;
        hipLaunchKernelGGL(batch_desc_scatter, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                           b->s, perm + c0, m, levels, cb.path, cb.pscore, cb.len, cb.exit, op,
                           os, ol, oe);
        HVK_HIP(hipGetLastError());
        return HVK_OK;
    };
    for (int64_t c0 = 0; c0 < n; c0 += b->chunk) {
        const int m = (int)std::min<int64_t>(b->chunk, n - c0);
        const int chunk_max = b->lens[(size_t)c0];  // longest first
        auto make_key = [&]() {
            DescKey key;
            std::memset(&key, 0, sizeof(key));
            key.c0 = c0;
            key.n = m;
            key.width = width_class(chunk_max, b->ctx->k);
            key.E = pl.E;
            key.B = pl.B;
            key.L = pl.L;
            key.dense = pl.dense;
            key.n_cat = pl.n_cat;
            key.n_roots = pl.n_roots;
            key.Fmax = pl.Fmax;
            key.maxN = pl.maxN;
            key.max_span = pl.max_span;
            std::memcpy(&key.thr_bits, &threshold, 4);
            key.max_abs_proto = pl.max_abs_proto;
            key.tld = pl.t.ld;
            key.tree[0] = pl.t.lo;
            key.tree[1] = pl.t.hi;
            key.tree[2] = pl.t.sqnorm;
            key.tree[3] = pl.t.row_maxabs;
            key.tree[4] = pl.t.child_off;
            key.tree[5] = pl.t.child_idx;
            key.tree[6] = pl.t.roots;
            key.in = b->ws_in.ptr;
            key.order = b->ws_order.ptr;
            key.rows = b->ws_rows.ptr;
            key.outs = b->ws_out.ptr;
            key.ws = b->ctx->ws_desc.ptr;
            return key;
        };
        DescKey key = make_key();
        auto it = std::find_if(b->dgraphs.begin(), b->dgraphs.end(),
                               [&](const DescGraph& g) { return g.key == key; });
        if (it != b->dgraphs.end()) {
            HVK_HIP(hipGraphLaunch(it->exec, b->s));
            std::rotate(it, it + 1, b->dgraphs.end());  // most recently used last
            if (int e = scatter(c0, m)) return e;
            continue;
        }
        // direct run (sizes the descent workspace), then capture the same launches
        if (int e = run_desc_chunk(b, packed, invalid, c0, m, chunk_max, pl, threshold, cb, starts,
                                   lens))
            return e;
        if (int e = scatter(c0, m)) return e;
        key = make_key();  // the run may have grown (moved) the workspace
        if (b->dgraphs.size() >= kMaxGraphs) {
            HVK_HIP(hipStreamSynchronize(b->s));
            while (b->dgraphs.size() >= kMaxGraphs) {
                (void)hipGraphExecDestroy(b->dgraphs.front().exec);
                b->dgraphs.erase(b->dgraphs.begin());
            }
        }
        HVK_HIP(hipStreamBeginCapture(b->s, hipStreamCaptureModeThreadLocal));
        const int ce = run_desc_chunk(b, packed, invalid, c0, m, chunk_max, pl, threshold, cb,
                                      starts, lens);
        hipGraph_t g = nullptr;
        const hipError_t he = hipStreamEndCapture(b->s, &g);
        if (ce) {
            if (g) (void)hipGraphDestroy(g);
            return ce;
        }
        if (he != hipSuccess) { set_last_error("hipStreamEndCapture", he); return HVK_ERR_HIP; }
        hipGraphExec_t ex = nullptr;
        const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ie != hipSuccess) { set_last_error("hipGraphInstantiate", ie); return HVK_ERR_HIP; }
        b->dgraphs.push_back({key, ex});
    }
    HVK_HIP(hipEventRecord(b->ev_out, b->s));
    HVK_HIP(hipStreamWaitEvent(caller, b->ev_out, 0));
    return HVK_OK;
}

int check_inputs(const hvk_batch* b, const uint32_t* packed, const uint32_t* invalid,
                 const int64_t* h_offsets, int n, const hvk_bundles* db, int k, const int32_t* oi,
                 const float* os) {
    if (!b || n < 0 || !db || k < 1) return HVK_ERR_ARG;
    if (n == 0) return HVK_OK;
    if (!packed || !invalid || !h_offsets || !oi || !os) return HVK_ERR_ARG;
    if (!db->clamp8 || !db->lo || !db->sqnorm || !db->maxabs || !db->resid1 || db->n <= 0)
        return HVK_ERR_ARG;
    return HVK_OK;
}

}  // namespace
}  // namespace hvk

using namespace hvk;

extern "C" int hvk_batch_create(int dim, int kmer, int chunk, hvk_batch** out) {
    if (!out || chunk < kMinChunk) return HVK_ERR_ARG;
    *out = nullptr;
    hvk_ctx* ctx = nullptr;
    if (int e = hvk_ctx_create(dim, kmer, &ctx)) return e;
    hvk_batch* b = new hvk_batch;
    b->ctx = ctx;
    b->chunk = chunk;
    hipError_t e = hipStreamCreateWithFlags(&b->s, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreate(&b->t0);
    if (e == hipSuccess) e = hipEventCreate(&b->t1);
    if (e != hipSuccess) {
        set_last_error("hvk_batch_create", e);
        hvk_batch_destroy(b);
        return HVK_ERR_HIP;
    }
    *out = b;
    return HVK_OK;
}

extern "C" int hvk_batch_destroy(hvk_batch* b) {
    if (!b) return HVK_OK;
    if (b->s) (void)hipStreamSynchronize(b->s);
    drop_graphs(b);
    for (hipEvent_t ev : {b->ev_in, b->ev_out, b->t0, b->t1})
        if (ev) (void)hipEventDestroy(ev);
    if (b->s) (void)hipStreamDestroy(b->s);
    hvk_ctx* ctx = b->ctx;
    delete b;  // frees the workspaces before the context
    return hvk_ctx_destroy(ctx);
}

extern "C" int hvk_batch_config(hvk_batch* b, int form, int shape, int splits, int chunk) {
    if (!b || (form != HVK_BATCH_FORM_BOUNDED && form != HVK_BATCH_FORM_TWO_LIMB) || shape < 0 ||
        shape >= kMfmaShapes || splits < 0 || splits > 64 || chunk < kMinChunk)
        return HVK_ERR_ARG;
    if (form != b->form || shape != b->shape || splits != b->splits || chunk != b->chunk) {
        HVK_HIP(hipStreamSynchronize(b->s));
        drop_graphs(b);
    }
    b->form = form;
    b->shape = shape;
    b->splits = splits;
    b->chunk = chunk;
    return HVK_OK;
}

extern "C" int hvk_batch_get_config(const hvk_batch* b, int* form, int* shape, int* splits,
                                    int* chunk, int* n_graphs) {
    if (!b) return HVK_ERR_ARG;
    if (form) *form = b->form;
    if (shape) *shape = b->shape;
    if (splits) *splits = b->splits;
    if (chunk) *chunk = b->chunk;
    if (n_graphs) *n_graphs = (int)b->graphs.size();
    return HVK_OK;
}

extern "C" int hvk_batch_classify(hvk_batch* b, const uint32_t* d_packed,
                                  const uint32_t* d_invalid, const int64_t* h_offsets, int n,
                                  const hvk_bundles* db, int k, int32_t* d_idx, float* d_score,
                                  double* d_cos64, int32_t* d_n_rescored, hvk_stream stream) {
    if (int e = check_inputs(b, d_packed, d_invalid, h_offsets, n, db, k, d_idx, d_score)) return e;
    if (n == 0) return HVK_OK;
    return classify(b, d_packed, d_invalid, h_offsets, n, *db, k, d_idx, d_score, d_cos64,
                    d_n_rescored, (hipStream_t)stream);
}

extern "C" int hvk_batch_autotune(hvk_batch* b, const uint32_t* d_packed,
                                  const uint32_t* d_invalid, const int64_t* h_offsets, int n,
                                  const hvk_bundles* db, int k, int32_t* d_idx, float* d_score,
                                  int sample, float* ms_per_mread, float* table, int* table_rows,
                                  hvk_stream stream) {
    if (int e = check_inputs(b, d_packed, d_invalid, h_offsets, n, db, k, d_idx, d_score)) return e;
    if (n == 0 || sample <= 0) return HVK_ERR_ARG;
    if (table && !table_rows) return HVK_ERR_ARG;
    const int m = std::min(n, sample);
    hipStream_t st = (hipStream_t)stream;
    struct Cand { int form, shape, splits, chunk; };
    const int cap = table ? *table_rows : 0;
    int rows = 0;
    // times one setting: a direct run and capture, then the faster of two replays
    auto time_one = [&](const Cand& c, float& ms) -> int {
        if (int e = hvk_batch_config(b, c.form, c.shape, c.splits, c.chunk)) return e;
        if (int e = classify(b, d_packed, d_invalid, h_offsets, m, *db, k, d_idx, d_score, nullptr,
                             nullptr, st))
            return e;
        ms = -1.0f;
        for (int rep = 0; rep < 2; ++rep) {
            HVK_HIP(hipEventRecord(b->t0, st));
            if (int e = classify(b, d_packed, d_invalid, h_offsets, m, *db, k, d_idx, d_score,
                                 nullptr, nullptr, st))
                return e;
            HVK_HIP(hipEventRecord(b->t1, st));
            HVK_HIP(hipEventSynchronize(b->t1));
            float t = 0.0f;
            HVK_HIP(hipEventElapsedTime(&t, b->t0, b->t1));
            if (ms < 0.0f || t < ms) ms = t;
        }
        if (rows < cap) {
            float* row = table + 5 * (size_t)rows;
            row[0] = (float)c.form;
            row[1] = (float)c.shape;
            row[2] = (float)c.splits;
            row[3] = (float)c.chunk;
            row[4] = ms * 1e6f / (float)m;
        }
        ++rows;
        return HVK_OK;
    };
    float best = -1.0f;
    Cand pick = {b->form, b->shape, b->splits, b->chunk};
    auto consider = [&](const Cand& c) -> int {
        float ms = 0.0f;
        if (int e = time_one(c, ms)) return e;
        if (best < 0.0f || ms < best) {
            best = ms;
            pick = c;
        }
        return HVK_OK;
    };
    // Coordinate search: the tile shape (the bounded form's shapes apply to
    // top-5, 8-entry lists; other k keep the default), then the reference
    // slices, then the chunk size, each with the others at their best so
    // far; and the two-limb form once.
    const int half = std::max(kMinChunk, m / 2);
    const int n_shapes = pick_kp(k) == 8 ? kMfmaShapes : 1;
    for (int sh = 0; sh < n_shapes; ++sh)
        if (int e = consider({HVK_BATCH_FORM_BOUNDED, sh, 0, half})) return e;
    const int best_shape = pick.shape;
    for (int s : {4, 8, 16})
        if (int e = consider({HVK_BATCH_FORM_BOUNDED, best_shape, s, half})) return e;
    const int best_splits = pick.splits;
    for (int chunk : {m / 4, m})
        if (chunk >= kMinChunk && chunk != half)
            if (int e = consider({HVK_BATCH_FORM_BOUNDED, best_shape, best_splits, chunk})) return e;
    if (int e = consider({HVK_BATCH_FORM_TWO_LIMB, 0, 4, half})) return e;
    if (int e = hvk_batch_config(b, pick.form, pick.shape, pick.splits, pick.chunk)) return e;
    if (ms_per_mread) *ms_per_mread = best * 1e6f / (float)m;
    if (table_rows) *table_rows = std::min(rows, cap);
    return HVK_OK;
}

extern "C" int hvk_batch_descend(hvk_batch* b, const hvk_tree* tree, const uint32_t* d_packed,
                                 const uint32_t* d_invalid, const int64_t* h_offsets, int n,
                                 int beam, float threshold, int levels, int32_t* d_path,
                                 float* d_path_score, int32_t* d_path_len, int32_t* d_exit_level,
                                 hvk_stream stream) {
    if (!b || !tree || n < 0) return HVK_ERR_ARG;
    if (n == 0) return HVK_OK;
    if (!d_packed || !d_invalid || !h_offsets || !d_path || !d_path_score || !d_path_len ||
        !d_exit_level)
        return HVK_ERR_ARG;
    return descend(b, tree, d_packed, d_invalid, h_offsets, n, beam, threshold, levels, d_path,
                   d_path_score, d_path_len, d_exit_level, (hipStream_t)stream);
}
