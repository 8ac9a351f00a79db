// session.h -- the process's GPU session behind the drop-in C++ API: one
// hvk_ctx for the compile-time HV_DIM / KMER_SIZE on the current device,
// device buffers, and batched encode / search helpers over the C ABI
// (include/hdc_amd.h).  Errors throw std::runtime_error, the reference's
// convention for its GPU entry points (src/encode.cu:343-350,
// src/cosine.cu:537-540).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <string_view>
#include <vector>

#include "../../include/hdc_amd.h"
#include "config.h"

namespace hvk_host {

// Owning device allocation (hipMalloc / hipFree).
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t bytes);
    ~DeviceBuffer();
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept;
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept;
    void* get() const { return p_; }
    template <typename T> T* as() const { return static_cast<T*>(p_); }
    size_t bytes() const { return n_; }
    void upload(const void* src, size_t bytes);
    void download(void* dst, size_t bytes) const;

private:
    void* p_ = nullptr;
    size_t n_ = 0;
};

// Bundles of a set of sequences (or windows) resident on the device.
struct DeviceBundles {
    int n = 0;
    int64_t ld = 0;
    DeviceBuffer counts, lo, hi, clamp8, sqnorm, maxabs, nkmers, resid1;
    hvk_bundles view(bool two_limbs) const;
    bool needs_two_limbs() const;  // any |count| > 127 (reads maxabs back)
};

// Sequences packed on the device.
struct DevicePacked {
    DeviceBuffer packed, invalid, offsets;
    int n = 0;
    int max_len = 0;
    std::vector<int64_t> host_offsets;
};

class Session {
public:
    static Session& instance();
    hvk_ctx* ctx() const { return ctx_; }
    int64_t ld() const { return ld_; }

    DevicePacked pack(const std::vector<std::string_view>& seqs) const;
    // one bundle per sequence (HVCache::compute semantics)
    DeviceBundles encode(const DevicePacked& p) const;
    // one bundle per window [starts[i], starts[i] + len) of the packed buffer
    DeviceBundles encode_windows(const DevicePacked& p, const std::vector<int64_t>& starts,
                                 int len) const;
    // search operands from host count rows (HV_DIM values each)
    DeviceBundles upload_counts(const std::vector<const HV_16*>& rows) const;
    // count rows back to the host
    std::vector<HV_16> download_counts(const DeviceBundles& b) const;

    struct TopK {
        std::vector<int32_t> idx;   // [nq][k], -1 past the end
        std::vector<double> cos;    // c (not 2c-1)
    };
    // exact top-k (cosine desc, index asc) of q against r
    TopK search(const DeviceBundles& q, const DeviceBundles& r, int k) const;
    // the same on operand views (rows of resident bundles)
    TopK search(const hvk_bundles& q, const hvk_bundles& r, int k) const;

    // device bundles with room for n rows (b.n = 0 rows in use)
    DeviceBundles alloc(int n) const;
    // rows [i0, i1) of b as a search operand
    static hvk_bundles rows(const DeviceBundles& b, int i0, int i1, bool two_limbs);
    // appends row i of src to dst (a device-to-device copy; dst has room)
    void append_row(DeviceBundles& dst, const DeviceBundles& src, int i) const;

private:
    Session();
    ~Session();
    hvk_ctx* ctx_ = nullptr;
    int64_t ld_ = 0;
};

void check(int status, const char* what);

}  // namespace hvk_host
