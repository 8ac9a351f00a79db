// dbfile.cpp -- the versioned prototype database file (hvk_db_*): write from
// host arrays, map read-only, validate, hand out host views, upload to the
// device.  Layout in include/hdc_amd.h.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <climits>
#include <cstdio>
#include <cstring>

#include "hvk_internal.h"

namespace {

constexpr char kMagic[8] = {'H', 'V', 'K', 'P', 'D', 'B', '0', '1'};
constexpr uint32_t kVersion = 1;
constexpr size_t kHeader = 128;
constexpr size_t kAlign = 4096;

struct Header {  // 128 bytes
    char magic[8];
    uint32_t version;
    uint32_t header_bytes;
    int32_t dim;
    int32_t kmer;
    int64_t ld;
    int64_t n;
    uint32_t flags;
    uint32_t reserved0;
    int64_t names_len;
    uint8_t reserved[64];
    uint64_t checksum;  // FNV-1a 64 of the 120 bytes above
};
static_assert(sizeof(Header) == kHeader, "header size");

uint64_t fnv1a(const void* p, size_t n) {
    uint64_t h = 1469598103934665603ull;
    const uint8_t* b = (const uint8_t*)p;
    for (size_t i = 0; i < n; ++i) {
        h ^= b[i];
        h *= 1099511628211ull;
    }
    return h;
}

size_t up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

// sections in file order: element size per row element (0 = per-row array)
struct Section {
    uint32_t flag;
    size_t row_elem;  // bytes per element of an [n][ld] array
    size_t per_row;   // bytes per row of an [n] array
};
constexpr Section kSections[] = {
    {HVK_DB_COUNTS, 2, 0}, {HVK_DB_LO, 1, 0},   {HVK_DB_HI, 1, 0},     {HVK_DB_CLAMP8, 1, 0},
    {HVK_DB_SUMS, 0, 8},   {HVK_DB_SUMS, 0, 4}, {HVK_DB_SUMS, 0, 4},   {HVK_DB_SUMS, 0, 4},
};
constexpr int kNumSections = 8;

size_t section_bytes(const Section& s, int64_t n, int64_t ld) {
    return s.row_elem ? (size_t)n * ld * s.row_elem : (size_t)n * s.per_row;
}

// file offsets of the sections (0 when absent) and of the names; total size
size_t layout(const Header& h, size_t (&off)[kNumSections], size_t& names_off) {
    size_t pos = up(kHeader);
    for (int i = 0; i < kNumSections; ++i) {
        if (h.flags & kSections[i].flag) {
            off[i] = pos;
            pos = up(pos + section_bytes(kSections[i], h.n, h.ld));
        } else {
            off[i] = 0;
        }
    }
    names_off = h.names_len ? pos : 0;
    return pos + (size_t)h.names_len;
}

const void* host_src(const hvk_encode_out* a, int i) {
    switch (i) {
        case 0: return a->counts;
        case 1: return a->lo;
        case 2: return a->hi;
        case 3: return a->clamp8;
        case 4: return a->sqnorm;
        case 5: return a->maxabs;
        case 6: return a->resid1;
        default: return a->nkmers;
    }
}

}  // namespace

struct hvk_db {
    void* map = nullptr;
    size_t size = 0;
    Header h{};
    size_t off[kNumSections] = {};
    size_t names_off = 0;
};

extern "C" int hvk_db_write(const char* path, int dim, int kmer, int64_t ld, int n, uint32_t flags,
                            const hvk_encode_out* arrays, const char* names, int64_t names_len) {
    if (!path || dim < 1 || kmer < 1 || ld < dim || n < 0 || names_len < 0 ||
        (flags & ~31u) || (names_len && !names))
        return HVK_ERR_ARG;
    if (flags && !arrays) return HVK_ERR_ARG;
    const int64_t src_ld = arrays && arrays->ld ? arrays->ld : ld;
    if (src_ld < ld) return HVK_ERR_ARG;
    for (int i = 0; i < kNumSections; ++i)
        if ((flags & kSections[i].flag) && !host_src(arrays, i)) return HVK_ERR_ARG;
    Header h{};
    std::memcpy(h.magic, kMagic, 8);
    h.version = kVersion;
    h.header_bytes = kHeader;
    h.dim = dim;
    h.kmer = kmer;
    h.ld = ld;
    h.n = n;
    h.flags = flags;
    h.names_len = names_len;
    h.checksum = fnv1a(&h, offsetof(Header, checksum));
    size_t off[kNumSections], names_off;
    const size_t total = layout(h, off, names_off);
    FILE* f = std::fopen(path, "wb");
    if (!f) return HVK_ERR_ARG;
    bool ok = std::fwrite(&h, kHeader, 1, f) == 1;
    auto pad_to = [&](size_t pos) {
        static const char zeros[kAlign] = {};
        long cur = std::ftell(f);
        while (ok && (size_t)cur < pos) {
            const size_t k = std::min(kAlign, pos - (size_t)cur);
            ok = std::fwrite(zeros, 1, k, f) == k;
            cur += (long)k;
        }
    };
    for (int i = 0; i < kNumSections && ok; ++i) {
        if (!off[i]) continue;
        pad_to(off[i]);
        const Section& s = kSections[i];
        const char* src = (const char*)host_src(arrays, i);
        if (s.row_elem && src_ld != ld) {  // repack rows to the file's stride
            for (int64_t r = 0; r < n && ok; ++r)
                ok = std::fwrite(src + (size_t)r * src_ld * s.row_elem, s.row_elem, (size_t)ld, f) ==
                     (size_t)ld;
        } else {
            const size_t b = section_bytes(s, n, ld);
            ok = b == 0 || std::fwrite(src, 1, b, f) == b;
        }
    }
    if (ok && names_len) {
        pad_to(names_off);
        ok = std::fwrite(names, 1, (size_t)names_len, f) == (size_t)names_len;
    }
    if (ok) pad_to(total);
    if (std::fclose(f) != 0) ok = false;
    return ok ? HVK_OK : HVK_ERR_ARG;
}

extern "C" int hvk_db_open(const char* path, hvk_db** out) {
    if (!path || !out) return HVK_ERR_ARG;
    *out = nullptr;
    const int fd = ::open(path, O_RDONLY);
    if (fd < 0) return HVK_ERR_ARG;
    struct stat st {};
    if (::fstat(fd, &st) != 0 || (size_t)st.st_size < kHeader) {
        ::close(fd);
        return HVK_ERR_ARG;
    }
    void* map = ::mmap(nullptr, (size_t)st.st_size, PROT_READ, MAP_PRIVATE, fd, 0);
    ::close(fd);
    if (map == MAP_FAILED) return HVK_ERR_ARG;
    hvk_db* db = new hvk_db;
    db->map = map;
    db->size = (size_t)st.st_size;
    std::memcpy(&db->h, map, kHeader);
    const Header& h = db->h;
    bool ok = std::memcmp(h.magic, kMagic, 8) == 0 && h.version == kVersion &&
              h.header_bytes == kHeader &&
              h.checksum == fnv1a(&h, offsetof(Header, checksum)) && h.dim > 0 && h.kmer > 0 &&
              h.ld >= h.dim && h.ld <= (int64_t(1) << 24) && h.n >= 0 && h.n <= INT32_MAX &&
              h.names_len >= 0 && h.names_len <= (int64_t(1) << 40) && !(h.flags & ~31u);
    if (ok) ok = layout(h, db->off, db->names_off) <= db->size;
    if (!ok) {
        hvk_db_close(db);
        return HVK_ERR_ARG;
    }
    *out = db;
    return HVK_OK;
}

extern "C" int hvk_db_close(hvk_db* db) {
    if (!db) return HVK_OK;
    if (db->map) ::munmap(db->map, db->size);
    delete db;
    return HVK_OK;
}

extern "C" int hvk_db_info(const hvk_db* db, int* dim, int* kmer, int64_t* ld, int* n,
                           uint32_t* flags, int64_t* names_len) {
    if (!db) return HVK_ERR_ARG;
    if (dim) *dim = db->h.dim;
    if (kmer) *kmer = db->h.kmer;
    if (ld) *ld = db->h.ld;
    if (n) *n = (int)db->h.n;
    if (flags) *flags = db->h.flags;
    if (names_len) *names_len = db->h.names_len;
    return HVK_OK;
}

extern "C" const void* hvk_db_array(const hvk_db* db, int which, int64_t* bytes) {
    if (bytes) *bytes = 0;
    if (!db) return nullptr;
    int i = -1;
    switch (which) {
        case (int)HVK_DB_COUNTS: i = 0; break;
        case (int)HVK_DB_LO: i = 1; break;
        case (int)HVK_DB_HI: i = 2; break;
        case (int)HVK_DB_CLAMP8: i = 3; break;
        case 32: case 33: case 34: case 35: i = 4 + (which - 32); break;
        case 64:
            if (!db->h.names_len) return nullptr;
            if (bytes) *bytes = db->h.names_len;
            return (const char*)db->map + db->names_off;
        default: return nullptr;
    }
    if (!db->off[i]) return nullptr;
    if (bytes) *bytes = (int64_t)section_bytes(kSections[i], db->h.n, db->h.ld);
    return (const char*)db->map + db->off[i];
}

extern "C" int hvk_db_upload(const hvk_db* db, const hvk_ctx* ctx, const hvk_encode_out* dst) {
    if (!db || !ctx || !dst) return HVK_ERR_ARG;
    if (db->h.dim != ctx->dim || db->h.kmer != ctx->k) return HVK_ERR_ARG;
    const int64_t dld = dst->ld ? dst->ld : db->h.ld;
    if (dld < db->h.dim) return HVK_ERR_ARG;
    const int64_t n = db->h.n, fld = db->h.ld;
    for (int i = 0; i < kNumSections; ++i) {
        void* d = const_cast<void*>(host_src(dst, i));
        if (!d || !db->off[i]) continue;
        const Section& s = kSections[i];
        const char* src = (const char*)db->map + db->off[i];
        if (s.row_elem) {
            const int64_t w = std::min(dld, fld);
            HVK_HIP(hipMemcpy2D(d, (size_t)dld * s.row_elem, src, (size_t)fld * s.row_elem,
                                (size_t)w * s.row_elem, (size_t)n, hipMemcpyHostToDevice));
            if (dld > fld)  // columns past the file's stride are zero
                HVK_HIP(hipMemset2D((char*)d + (size_t)fld * s.row_elem, (size_t)dld * s.row_elem, 0,
                                    (size_t)(dld - fld) * s.row_elem, (size_t)n));
        } else {
            HVK_HIP(hipMemcpy(d, src, section_bytes(s, n, fld), hipMemcpyHostToDevice));
        }
    }
    return HVK_OK;
}
