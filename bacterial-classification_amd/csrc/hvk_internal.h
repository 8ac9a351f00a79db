// hvk_internal.h -- shared definitions of the gfx950 HDC library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <mutex>
#include <vector>

#include "../../include/hdc_amd.h"

namespace hvk {

constexpr int kMaxDim = 32768;
constexpr int kMaxK = 16;
constexpr int kMaxTopK = 32;

// Row stride of the limb / count outputs: D rounded up to 128 elements so a
// K-slab of any search kernel never straddles a row end.
inline int64_t row_ld_for(int dim) { return (int64_t(dim) + 127) / 128 * 128; }

void set_last_error(const char* where, hipError_t e);
void set_last_message(const char* msg);
// Resolves, on `device` (current), the address of the allocation-free
// search's scratch array (noalloc.hip); called when a context is created.
int noalloc_setup(int device);

#define HVK_HIP(call)                                   \
    do {                                                \
        hipError_t e_ = (call);                         \
        if (e_ != hipSuccess) {                         \
            ::hvk::set_last_error(#call, e_);           \
            return HVK_ERR_HIP;                         \
        }                                               \
    } while (0)

// A device buffer that only grows; reused across calls on one context.
struct Workspace {
    void* ptr = nullptr;
    size_t bytes = 0;
    int ensure(size_t need);
    ~Workspace();
};

}  // namespace hvk

struct hvk_ctx {
    int dim = 0;
    int k = 0;
    int w64 = 0;   // words of 64 bits per binary HV
    int w32 = 0;   // words of 32 bits
    int64_t ld = 0;
    int device = 0;
    int n_cu = 0;
    int max_lds = 65536;  // bytes of LDS one workgroup may use
    std::vector<uint64_t> base[4];      // A C G T (code order), W words each
    uint32_t* d_table = nullptr;        // [4][k][w32] rotated base words
    hvk::Workspace ws_search;           // partial top lists
    hvk::Workspace ws_norms;            // inverse norms
    hvk::Workspace ws_flags;            // rescan queues
    hvk::Workspace ws_search2;          // partial lists of the second pass
    hvk::Workspace ws_desc;             // hierarchical descent: beams, scores, buckets
    hvk::Workspace ws_desc_tree;        // hierarchical descent: uploaded candidate lists
    hvk::Workspace ws_select;           // large-k selection: key matrix and sort buffers
    std::mutex mu;
    // kernel timing (hvk_ctx_set_timing)
    bool timing = false;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    bool ev_used[2] = {false, false};  // encode, search
    // tuning (hvk_ctx_set_tuning); 0 = automatic
    int tune_splits = 0;
    int tune_select_kib = 0;  // large-k selection workspace budget, 0 = 1 GiB
    int tune_shape = 0;       // tile shape of the bounded search (8-entry lists)
    bool encode_legacy = false;  // workgroup-per-sequence encoders only
    bool tune_desc_i16 = false;  // descent: int16 read rows only
    int tune_group_slices = 0;   // reference slices per XCD group of the MFMA search, 0 = auto
};

// Kernel launchers (defined in the .hip files).
namespace hvk {
// which = 0: encode, 1: search.  Records the start / stop event when timing.
inline void timing_mark(hvk_ctx* c, int which, bool stop, hipStream_t s) {
    if (!c->timing) return;
    (void)hipEventRecord(c->ev[2 * which + (stop ? 1 : 0)], s);
    if (stop) c->ev_used[which] = true;
}
// Outputs of the encode kernel (each pointer nullable; ld = row stride).
// A descent's tree on the device: prototype limbs and norms (the caller's)
// and the candidate lists (uploaded into the context).
struct TreeDev {
    const int16_t* lo;
    const int16_t* hi;
    const int64_t* sqnorm;
    int64_t ld;
    int n_cat;
    const int32_t* child_off;  // [n_cat + 1]
    const int32_t* child_idx;
    const int32_t* roots;
    int n_roots;
    const int32_t* row_maxabs;  // [n_cat] max |p|, nullable
};

// What desc_prepare checks and chooses for a tree, a read bound, beam and
// levels; desc_launch runs a descent on it.
struct DescPlan {
    TreeDev t;
    int E;  // elements per int32 partial-sum chunk (8 or 2)
    int B, L;
    int Fmax, maxN, max_span;
    int n_cat, n_roots;
    int64_t max_abs_proto;
    bool dense;  // level 0 as the int8-digit GEMM
};
int desc_prepare(hvk_ctx* ctx, const hvk_tree* tree, int q_maxabs, int beam, int levels,
                 DescPlan& pl, hipStream_t st);
int desc_launch(hvk_ctx* ctx, const DescPlan& pl, const int16_t* d_q_counts,
                const int64_t* d_q_sqnorm, int n_q, int64_t q_ld, float threshold,
                int32_t* d_path, float* d_path_score, int32_t* d_path_len,
                int32_t* d_exit_level, hipStream_t st);

struct EncodeOut {
    int16_t* counts;
    int8_t* lo;
    int8_t* hi;
    int8_t* clamp8;
    int64_t ld;
    int64_t* sqnorm;
    int32_t* maxabs;
    int32_t* nkmers;
    int32_t* resid1;
};
// Sequences are [offsets[i], offsets[i+1]) or, when starts is set, lens[i]
// (window_len when lens is null) bases starting at starts[i].
int launch_encode(hvk_ctx* ctx, const uint32_t* packed, const uint32_t* invalid,
                  const int64_t* offsets, const int64_t* starts, int window_len, int n,
                  int max_len, const EncodeOut& out, hipStream_t s,
                  const int32_t* lens = nullptr);
}
