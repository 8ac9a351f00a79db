// session.cpp -- device session of the drop-in C++ API.
#include "session.h"

#include <stdexcept>

namespace hvk_host {

void check(int status, const char* what) {
    if (status != HVK_OK)
        throw std::runtime_error(std::string(what) + ": " + hvk_status_string(status) + " (" +
                                 hvk_last_error() + ")");
}

static void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

DeviceBuffer::DeviceBuffer(size_t bytes) : n_(bytes) {
    if (bytes) hip_check(hipMalloc(&p_, bytes), "hipMalloc");
}
DeviceBuffer::~DeviceBuffer() {
    if (p_) (void)hipFree(p_);
}
DeviceBuffer::DeviceBuffer(DeviceBuffer&& o) noexcept : p_(o.p_), n_(o.n_) {
    o.p_ = nullptr;
    o.n_ = 0;
}
DeviceBuffer& DeviceBuffer::operator=(DeviceBuffer&& o) noexcept {
    if (this != &o) {
        if (p_) (void)hipFree(p_);
        p_ = o.p_;
        n_ = o.n_;
        o.p_ = nullptr;
        o.n_ = 0;
    }
    return *this;
}
void DeviceBuffer::upload(const void* src, size_t bytes) {
    if (bytes) hip_check(hipMemcpy(p_, src, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
}
void DeviceBuffer::download(void* dst, size_t bytes) const {
    if (bytes) hip_check(hipMemcpy(dst, p_, bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H");
}

hvk_bundles DeviceBundles::view(bool two_limbs) const {
    hvk_bundles v{};
    v.clamp8 = clamp8.as<int8_t>();
    v.lo = lo.as<int8_t>();
    v.hi = two_limbs ? hi.as<int8_t>() : nullptr;
    v.sqnorm = sqnorm.as<int64_t>();
    v.maxabs = maxabs.as<int32_t>();
    v.resid1 = resid1.as<int32_t>();
    v.n = n;
    return v;
}

bool DeviceBundles::needs_two_limbs() const {
    if (n == 0) return false;
    std::vector<int32_t> m(n);
    maxabs.download(m.data(), sizeof(int32_t) * n);
    for (int32_t x : m)
        if (x > 127) return true;
    return false;
}

Session& Session::instance() {
    static Session s;
    return s;
}

Session::Session() {
    check(hvk_ctx_create(HV_DIM, KMER_SIZE, &ctx_), "hvk_ctx_create");
    int ld = 0;
    check(hvk_ctx_info(ctx_, nullptr, nullptr, nullptr, &ld), "hvk_ctx_info");
    ld_ = ld;
}

Session::~Session() {
    if (ctx_) hvk_ctx_destroy(ctx_);
}

DevicePacked Session::pack(const std::vector<std::string_view>& seqs) const {
    DevicePacked p;
    p.n = (int)seqs.size();
    p.host_offsets.resize(seqs.size() + 1, 0);
    std::string buf;
    for (size_t i = 0; i < seqs.size(); ++i) {
        buf.append(seqs[i]);
        p.host_offsets[i + 1] = (int64_t)buf.size();
        p.max_len = std::max(p.max_len, (int)seqs[i].size());
    }
    const int64_t words = hvk_pack_words((int64_t)buf.size());
    std::vector<uint32_t> packed((size_t)words), invalid((size_t)words);
    check(hvk_pack(buf.data(), p.host_offsets.data(), p.n, packed.data(), invalid.data()),
          "hvk_pack");
    p.packed = DeviceBuffer(sizeof(uint32_t) * words);
    p.invalid = DeviceBuffer(sizeof(uint32_t) * words);
    p.offsets = DeviceBuffer(sizeof(int64_t) * p.host_offsets.size());
    p.packed.upload(packed.data(), sizeof(uint32_t) * words);
    p.invalid.upload(invalid.data(), sizeof(uint32_t) * words);
    p.offsets.upload(p.host_offsets.data(), sizeof(int64_t) * p.host_offsets.size());
    return p;
}

static DeviceBundles alloc_bundles(int n, int64_t ld) {
    DeviceBundles b;
    b.n = n;
    b.ld = ld;
    const size_t rows = (size_t)n * ld;
    b.counts = DeviceBuffer(sizeof(int16_t) * rows);
    b.lo = DeviceBuffer(rows);
    b.hi = DeviceBuffer(rows);
    b.clamp8 = DeviceBuffer(rows);
    b.sqnorm = DeviceBuffer(sizeof(int64_t) * n);
    b.maxabs = DeviceBuffer(sizeof(int32_t) * n);
    b.nkmers = DeviceBuffer(sizeof(int32_t) * n);
    b.resid1 = DeviceBuffer(sizeof(int32_t) * n);
    return b;
}

static hvk_encode_out out_of(DeviceBundles& b) {
    hvk_encode_out o{};
    o.counts = b.counts.as<int16_t>();
    o.lo = b.lo.as<int8_t>();
    o.hi = b.hi.as<int8_t>();
    o.clamp8 = b.clamp8.as<int8_t>();
    o.ld = b.ld;
    o.sqnorm = b.sqnorm.as<int64_t>();
    o.maxabs = b.maxabs.as<int32_t>();
    o.nkmers = b.nkmers.as<int32_t>();
    o.resid1 = b.resid1.as<int32_t>();
    return o;
}

DeviceBundles Session::encode(const DevicePacked& p) const {
    DeviceBundles b = alloc_bundles(p.n, ld_);
    hvk_encode_out o = out_of(b);
    check(hvk_encode_ex(ctx_, p.packed.as<uint32_t>(), p.invalid.as<uint32_t>(),
                        p.offsets.as<int64_t>(), p.n, p.max_len, &o, nullptr),
          "hvk_encode_ex");
    return b;
}

DeviceBundles Session::encode_windows(const DevicePacked& p, const std::vector<int64_t>& starts,
                                      int len) const {
    DeviceBundles b = alloc_bundles((int)starts.size(), ld_);
    if (starts.empty()) return b;
    DeviceBuffer st(sizeof(int64_t) * starts.size());
    st.upload(starts.data(), sizeof(int64_t) * starts.size());
    hvk_encode_out o = out_of(b);
    check(hvk_encode_windows(ctx_, p.packed.as<uint32_t>(), p.invalid.as<uint32_t>(),
                             st.as<int64_t>(), (int)starts.size(), len, &o, nullptr),
          "hvk_encode_windows");
    hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    return b;
}

DeviceBundles Session::upload_counts(const std::vector<const HV_16*>& rows) const {
    DeviceBundles b = alloc_bundles((int)rows.size(), ld_);
    if (rows.empty()) return b;
    std::vector<int16_t> flat((size_t)rows.size() * ld_, 0);
    for (size_t i = 0; i < rows.size(); ++i) {
        const size_t n = std::min<size_t>(rows[i]->size(), HV_DIM);
        std::copy(rows[i]->begin(), rows[i]->begin() + n, flat.begin() + i * ld_);
    }
    b.counts.upload(flat.data(), sizeof(int16_t) * flat.size());
    hvk_encode_out o = out_of(b);
    o.counts = nullptr;
    o.nkmers = nullptr;
    check(hvk_prepare_counts(ctx_, b.counts.as<int16_t>(), b.n, &o, nullptr),
          "hvk_prepare_counts");
    return b;
}

std::vector<HV_16> Session::download_counts(const DeviceBundles& b) const {
    std::vector<int16_t> flat((size_t)b.n * b.ld);
    b.counts.download(flat.data(), sizeof(int16_t) * flat.size());
    std::vector<HV_16> out((size_t)b.n);
    for (int i = 0; i < b.n; ++i)
        out[i].assign(flat.begin() + (size_t)i * b.ld, flat.begin() + (size_t)i * b.ld + HV_DIM);
    return out;
}

Session::TopK Session::search(const DeviceBundles& q, const DeviceBundles& r, int k) const {
    if (q.n == 0 || r.n == 0 || k <= 0) return search(hvk_bundles{}, hvk_bundles{}, k);
    return search(q.view(q.needs_two_limbs()), r.view(r.needs_two_limbs()), k);
}

Session::TopK Session::search(const hvk_bundles& q, const hvk_bundles& r, int k) const {
    TopK t;
    t.idx.assign((size_t)q.n * (k > 0 ? k : 0), -1);
    t.cos.assign(t.idx.size(), 0.0);
    if (q.n == 0 || r.n == 0 || k <= 0) return t;
    DeviceBuffer idx(sizeof(int32_t) * q.n * k), score(sizeof(float) * q.n * k),
        cosv(sizeof(double) * q.n * k);
    check(hvk_search_mfma_bounded(ctx_, &q, &r, ld_, k, idx.as<int32_t>(), score.as<float>(),
                                  cosv.as<double>(), nullptr, nullptr),
          "hvk_search_mfma_bounded");
    idx.download(t.idx.data(), sizeof(int32_t) * t.idx.size());
    cosv.download(t.cos.data(), sizeof(double) * t.cos.size());
    return t;
}

DeviceBundles Session::alloc(int n) const {
    DeviceBundles b = alloc_bundles(n, ld_);
    b.n = 0;
    return b;
}

hvk_bundles Session::rows(const DeviceBundles& b, int i0, int i1, bool two_limbs) {
    hvk_bundles v = b.view(two_limbs);
    const int64_t o = (int64_t)i0 * b.ld;
    v.clamp8 += o;
    v.lo += o;
    if (v.hi) v.hi += o;
    v.sqnorm += i0;
    v.maxabs += i0;
    v.resid1 += i0;
    v.n = i1 - i0;
    return v;
}

void Session::append_row(DeviceBundles& dst, const DeviceBundles& src, int i) const {
    const int j = dst.n;
    if ((size_t)(j + 1) * sizeof(int64_t) > dst.sqnorm.bytes())
        throw std::runtime_error("append_row: no room");
    const size_t ld = (size_t)ld_;
    auto cp = [&](const DeviceBuffer& s, DeviceBuffer& d, size_t row_bytes) {
        hip_check(hipMemcpy((char*)d.get() + j * row_bytes, (const char*)s.get() + i * row_bytes,
                            row_bytes, hipMemcpyDeviceToDevice),
                  "hipMemcpy D2D");
    };
    cp(src.counts, dst.counts, ld * sizeof(int16_t));
    cp(src.lo, dst.lo, ld);
    cp(src.hi, dst.hi, ld);
    cp(src.clamp8, dst.clamp8, ld);
    cp(src.sqnorm, dst.sqnorm, sizeof(int64_t));
    cp(src.maxabs, dst.maxabs, sizeof(int32_t));
    cp(src.nkmers, dst.nkmers, sizeof(int32_t));
    cp(src.resid1, dst.resid1, sizeof(int32_t));
    dst.n = j + 1;
}

}  // namespace hvk_host
