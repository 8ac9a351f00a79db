// ctx.cpp -- context (D, k, base HV tables), status strings, host packing.
#include <cstring>
#include <random>
#include <string>

#include "hvk_internal.h"

namespace hvk {

static thread_local std::string g_last_error = "no error";

void set_last_error(const char* where, hipError_t e) {
    g_last_error = std::string(where) + ": " + hipGetErrorString(e);
}

void set_last_message(const char* msg) { g_last_error = msg; }

int Workspace::ensure(size_t need) {
    if (need <= bytes) return HVK_OK;
    if (ptr) {
        hipError_t e = hipFree(ptr);
        if (e != hipSuccess) { set_last_error("hipFree", e); return HVK_ERR_HIP; }
        ptr = nullptr;
        bytes = 0;
    }
    size_t grow = need + need / 4;
    hipError_t e = hipMalloc(&ptr, grow);
    if (e != hipSuccess) { set_last_error("hipMalloc", e); ptr = nullptr; return HVK_ERR_NOMEM; }
    bytes = grow;
    return HVK_OK;
}

Workspace::~Workspace() {
    if (ptr) (void)hipFree(ptr);
}

}  // namespace hvk

extern "C" {

const char* hvk_status_string(int status) {
    switch (status) {
        case HVK_OK: return "ok";
        case HVK_ERR_ARG: return "invalid argument";
        case HVK_ERR_HIP: return "HIP runtime error";
        case HVK_ERR_RANGE: return "values outside the exact-arithmetic range";
        case HVK_ERR_NOMEM: return "device allocation failed";
        case HVK_ERR_NODEVICE: return "no usable gfx950 device";
        default: return "unknown status";
    }
}

const char* hvk_last_error(void) { return hvk::g_last_error.c_str(); }

// Base HVs: make_hv (src/HV.cpp:20-35) seeded A:0 C:1 T:2 G:3 (:37-40), held
// here in code order A C G T.  The rotated table entry (b, i) is the HV
// rotR(base_b, i) on the 64*W-bit ring (mult, src/HV.cpp:98-147), split into
// 32-bit words; the encoder XORs k of them per k-mer.
int hvk_ctx_create(int dim, int kmer, hvk_ctx** out) {
    if (!out) return HVK_ERR_ARG;
    *out = nullptr;
    if (dim < 64 || dim > hvk::kMaxDim || kmer < 1 || kmer > hvk::kMaxK) return HVK_ERR_ARG;
    int dev = 0;
    HVK_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HVK_HIP(hipGetDeviceProperties(&prop, dev));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        hvk::g_last_error = std::string("device arch ") + prop.gcnArchName + " is not gfx950";
        return HVK_ERR_NODEVICE;
    }
    auto* c = new hvk_ctx();
    c->dim = dim;
    c->k = kmer;
    c->w64 = (dim + 63) / 64;
    c->w32 = (dim + 31) / 32;
    c->ld = hvk::row_ld_for(dim);
    c->device = dev;
    c->n_cu = prop.multiProcessorCount;
    c->max_lds = (int)prop.sharedMemPerBlock;
    const int excess = c->w64 * 64 - dim;
    const uint64_t tail = excess ? (~0ULL) >> excess : ~0ULL;
    const int seeds[4] = {0, 1, 3, 2};  // A C G T
    for (int b = 0; b < 4; ++b) {
        std::mt19937_64 rng((uint64_t)seeds[b]);
        c->base[b].resize(c->w64);
        for (int w = 0; w < c->w64; ++w) c->base[b][w] = rng();
        c->base[b][c->w64 - 1] &= tail;
    }
    const int W = c->w64;
    std::vector<uint32_t> table((size_t)4 * kmer * c->w32);
    for (int b = 0; b < 4; ++b)
        for (int i = 0; i < kmer; ++i) {
            const uint64_t* r = c->base[b].data();
            for (int w = 0; w < W; ++w) {
                // output word w of rotR(r, i): ring bits [64w + i, 64w + i + 64)
                int a = w % W, n = (w + 1) % W;
                uint64_t v = i ? ((r[a] >> i) | (r[n] << (64 - i))) : r[a];
                if (w == W - 1) v &= tail;
                size_t o = ((size_t)b * kmer + i) * c->w32;
                table[o + 2 * w] = (uint32_t)v;
                if (2 * w + 1 < c->w32) table[o + 2 * w + 1] = (uint32_t)(v >> 32);
            }
        }
    hipError_t e = hipMalloc(&c->d_table, table.size() * sizeof(uint32_t));
    if (e != hipSuccess) {
        hvk::set_last_error("hipMalloc(table)", e);
        delete c;
        return HVK_ERR_NOMEM;
    }
    e = hipMemcpy(c->d_table, table.data(), table.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        hvk::set_last_error("hipMemcpy(table)", e);
        (void)hipFree(c->d_table);
        delete c;
        return HVK_ERR_HIP;
    }
    if (int rc = hvk::noalloc_setup(dev)) {
        (void)hipFree(c->d_table);
        delete c;
        return rc;
    }
    *out = c;
    return HVK_OK;
}

int hvk_ctx_destroy(hvk_ctx* ctx) {
    if (!ctx) return HVK_ERR_ARG;
    if (ctx->d_table) (void)hipFree(ctx->d_table);
    for (auto& e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    delete ctx;
    return HVK_OK;
}

int hvk_ctx_info(const hvk_ctx* ctx, int* dim, int* kmer, int* words64, int* row_ld) {
    if (!ctx) return HVK_ERR_ARG;
    if (dim) *dim = ctx->dim;
    if (kmer) *kmer = ctx->k;
    if (words64) *words64 = ctx->w64;
    if (row_ld) *row_ld = (int)ctx->ld;
    return HVK_OK;
}

int hvk_ctx_set_timing(hvk_ctx* ctx, int enable) {
    if (!ctx) return HVK_ERR_ARG;
    if (enable && !ctx->ev[0])
        for (auto& e : ctx->ev) HVK_HIP(hipEventCreate(&e));
    ctx->timing = enable != 0;
    ctx->ev_used[0] = ctx->ev_used[1] = false;
    return HVK_OK;
}

int hvk_ctx_set_tuning(hvk_ctx* ctx, int param, int value) {
    if (!ctx) return HVK_ERR_ARG;
    switch (param) {
        case HVK_TUNE_SPLITS:
            if (value < 0 || value > 64) return HVK_ERR_ARG;
            ctx->tune_splits = value;
            return HVK_OK;
        case HVK_TUNE_SHAPE:
            if (value < 0 || value > 5) return HVK_ERR_ARG;
            ctx->tune_shape = value;
            return HVK_OK;
        case HVK_TUNE_SELECT_KIB:
            if (value < 0) return HVK_ERR_ARG;
            ctx->tune_select_kib = value;
            return HVK_OK;
        case HVK_TUNE_ENCODER:
            if (value < 0 || value > 1) return HVK_ERR_ARG;
            ctx->encode_legacy = value == 1;
            return HVK_OK;
        case HVK_TUNE_GROUP_SLICES:
            if (value < 0 || value > 64) return HVK_ERR_ARG;
            ctx->tune_group_slices = value;
            return HVK_OK;
        case HVK_TUNE_DESCENT:
            if (value < 0 || value > 1) return HVK_ERR_ARG;
            ctx->tune_desc_i16 = value == 1;
            return HVK_OK;
        default:
            return HVK_ERR_ARG;
    }
}

int hvk_ctx_last_timing(hvk_ctx* ctx, float* encode_ms, float* search_ms) {
    if (!ctx) return HVK_ERR_ARG;
    float* out[2] = {encode_ms, search_ms};
    for (int w = 0; w < 2; ++w) {
        if (!out[w]) continue;
        *out[w] = -1.0f;
        if (!ctx->timing || !ctx->ev_used[w]) continue;
        HVK_HIP(hipEventSynchronize(ctx->ev[2 * w + 1]));
        HVK_HIP(hipEventElapsedTime(out[w], ctx->ev[2 * w], ctx->ev[2 * w + 1]));
    }
    return HVK_OK;
}

int hvk_base_hv(const hvk_ctx* ctx, char symbol, uint64_t* host_out) {
    if (!ctx || !host_out) return HVK_ERR_ARG;
    int b;
    switch (symbol) {
        case 'A': b = 0; break;
        case 'C': b = 1; break;
        case 'G': b = 2; break;
        case 'T': b = 3; break;
        default: b = -1;
    }
    if (b < 0) {
        const int excess = ctx->w64 * 64 - ctx->dim;
        for (int w = 0; w < ctx->w64; ++w) host_out[w] = ~0ULL;
        if (excess) host_out[ctx->w64 - 1] = (~0ULL) >> excess;
        return HVK_OK;
    }
    std::memcpy(host_out, ctx->base[b].data(), sizeof(uint64_t) * ctx->w64);
    return HVK_OK;
}

int64_t hvk_pack_words(int64_t total_bases) {
    if (total_bases < 0) return -1;
    // +2 words: the encoder reads the word after the last base's word
    return total_bases / 16 + 2;
}

int hvk_pack(const char* seqs, const int64_t* offsets, int n_seqs, uint32_t* packed,
             uint32_t* invalid) {
    if (!seqs || !offsets || !packed || !invalid || n_seqs < 0) return HVK_ERR_ARG;
    const int64_t total = offsets[n_seqs];
    if (offsets[0] < 0 || total < offsets[0]) return HVK_ERR_ARG;
    for (int i = 0; i < n_seqs; ++i)
        if (offsets[i + 1] < offsets[i]) return HVK_ERR_ARG;
    const int64_t words = hvk_pack_words(total);
    std::memset(packed, 0, sizeof(uint32_t) * (size_t)words);
    std::memset(invalid, 0, sizeof(uint32_t) * (size_t)words);
    for (int64_t p = 0; p < total; ++p) {
        uint32_t code;
        switch (seqs[p]) {
            case 'A': code = 0; break;
            case 'C': code = 1; break;
            case 'G': code = 2; break;
            case 'T': code = 3; break;
            default:
                code = 0;
                invalid[p >> 5] |= 1u << (p & 31);
        }
        packed[p >> 4] |= code << (2 * (p & 15));
    }
    return HVK_OK;
}

}  // extern "C"
