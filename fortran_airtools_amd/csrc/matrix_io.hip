// matrix_io.hip -- the reference's matrix dump format on the host (no device work).
//
// store_matrix / read_csr_matrix / read_csc_matrix (system_generation.f90:595-707) write and read
// six Fortran unformatted sequential records: m, n, nnz (default integers), vA (nnz reals),
// JA (nnz integers), IA (m+1 or n+1 integers).  Each record is framed by a 4-byte little-endian
// length before and after it; a record longer than 2^31-1 bytes is split into subrecords (the
// gfortran convention): a negative leading marker says the record continues in the next
// subrecord, a negative trailing marker that the subrecord continues the previous one.  These
// functions let a C, Python or Fortran caller exchange matrices with a FAIR tools program without
// going through Fortran I/O.
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <string>
#include <vector>

#include "fair_internal.hpp"

namespace fair {
namespace {

struct File {
    FILE *f = nullptr;
    File(const char *path, const char *mode) : f(std::fopen(path, mode)) {}
    ~File() {
        if (f) std::fclose(f);
    }
};

constexpr uint64_t MAX_SUB = 2147483647ull;   // largest subrecord payload

void put_marker(FILE *f, int32_t v) {
    FAIR_CHECK(std::fwrite(&v, 4, 1, f) == 1, FAIR_ERR_ARG, "write failed");
}

void put_record(FILE *f, const void *data, uint64_t bytes) {
    const unsigned char *p = (const unsigned char *)data;
    uint64_t left = bytes;
    bool first = true;
    do {
        const uint64_t chunk = left > MAX_SUB ? MAX_SUB : left;
        const bool more = left > MAX_SUB;
        put_marker(f, more ? -(int32_t)chunk : (int32_t)chunk);
        if (chunk) FAIR_CHECK(std::fwrite(p, 1, chunk, f) == chunk, FAIR_ERR_ARG, "write failed");
        put_marker(f, first ? (int32_t)chunk : -(int32_t)chunk);
        first = false;
        p += chunk;
        left -= chunk;
    } while (left > 0);
}

// Reads one record (all subrecords) into out; returns its total length.
uint64_t get_record(FILE *f, std::vector<unsigned char> *out, void *dst, uint64_t cap) {
    uint64_t total = 0;
    for (;;) {
        int32_t head = 0, tail = 0;
        FAIR_CHECK(std::fread(&head, 4, 1, f) == 1, FAIR_ERR_ARG, "truncated matrix file (record marker)");
        const bool more = head < 0;
        const uint64_t len = (uint64_t)(more ? -(int64_t)head : (int64_t)head);
        if (out) {
            out->resize(total + len);
            if (len) FAIR_CHECK(std::fread(out->data() + total, 1, len, f) == len, FAIR_ERR_ARG, "truncated matrix file");
        } else {
            FAIR_CHECK(total + len <= cap, FAIR_ERR_ARG, "record longer than expected in matrix file");
            if (len)
                FAIR_CHECK(std::fread((unsigned char *)dst + total, 1, len, f) == len, FAIR_ERR_ARG,
                           "truncated matrix file");
        }
        FAIR_CHECK(std::fread(&tail, 4, 1, f) == 1, FAIR_ERR_ARG, "truncated matrix file (record marker)");
        FAIR_CHECK(tail == (int32_t)len || tail == -(int32_t)len, FAIR_ERR_ARG, "corrupt record markers in matrix file");
        total += len;
        if (!more) return total;
    }
}

int32_t get_int_record(FILE *f) {
    int32_t v = 0;
    FAIR_CHECK(get_record(f, nullptr, &v, 4) == 4, FAIR_ERR_ARG, "expected one default integer record");
    return v;
}

}  // namespace
}  // namespace fair

using namespace fair;

// store_matrix, system_generation.f90:595-625 (IA has nia entries: m+1 for CSR, n+1 for CSC).
extern "C" int fair_matrix_store(const char *path, int32_t m, int32_t n, int64_t nnz, const double *vA,
                                 const int32_t *JA, const int32_t *IA, int32_t nia) {
    return guard([&] {
        FAIR_CHECK(path && IA && (nnz == 0 || (vA && JA)), FAIR_ERR_ARG, "null argument");
        FAIR_CHECK(nnz >= 0 && nnz <= INT32_MAX && nia >= 1, FAIR_ERR_ARG, "sizes out of range");
        File f(path, "wb");
        FAIR_CHECK(f.f != nullptr, FAIR_ERR_ARG, std::string("cannot open ") + path);
        const int32_t nz = (int32_t)nnz;
        put_record(f.f, &m, 4);
        put_record(f.f, &n, 4);
        put_record(f.f, &nz, 4);
        put_record(f.f, vA, 8ull * (uint64_t)nnz);
        put_record(f.f, JA, 4ull * (uint64_t)nnz);
        put_record(f.f, IA, 4ull * (uint64_t)nia);
        FAIR_CHECK(std::fflush(f.f) == 0, FAIR_ERR_ARG, "write failed");
    });
}

// The sizes of a dump: m, n, nnz and the length of IA (m+1 for a CSR dump, n+1 for CSC).
extern "C" int fair_matrix_read_header(const char *path, int32_t *m, int32_t *n, int64_t *nnz, int32_t *nia) {
    return guard([&] {
        FAIR_CHECK(path && m && n && nnz && nia, FAIR_ERR_ARG, "null argument");
        File f(path, "rb");
        FAIR_CHECK(f.f != nullptr, FAIR_ERR_ARG, std::string("cannot open ") + path);
        *m = get_int_record(f.f);
        *n = get_int_record(f.f);
        const int32_t nz = get_int_record(f.f);
        FAIR_CHECK(nz >= 0, FAIR_ERR_ARG, "negative nnz in matrix file");
        *nnz = nz;
        // skip vA and JA, measure IA
        int32_t head = 0;
        for (int r = 0; r < 2; ++r) {
            uint64_t seen = 0;
            for (;;) {
                FAIR_CHECK(std::fread(&head, 4, 1, f.f) == 1, FAIR_ERR_ARG, "truncated matrix file");
                const uint64_t len = (uint64_t)(head < 0 ? -(int64_t)head : (int64_t)head);
                FAIR_CHECK(std::fseek(f.f, (long)len + 4, SEEK_CUR) == 0, FAIR_ERR_ARG, "truncated matrix file");
                seen += len;
                if (head >= 0) break;
            }
            FAIR_CHECK(seen == (r == 0 ? 8ull : 4ull) * (uint64_t)nz, FAIR_ERR_ARG, "record length does not match nnz");
        }
        std::vector<unsigned char> ia;
        const uint64_t len = get_record(f.f, &ia, nullptr, 0);
        FAIR_CHECK(len % 4 == 0 && len >= 4, FAIR_ERR_ARG, "bad IA record in matrix file");
        *nia = (int32_t)(len / 4);
    });
}

// read_csr_matrix / read_csc_matrix, system_generation.f90:627-707: vA (nnz), JA (nnz), IA (nia)
// with nnz and nia as fair_matrix_read_header reported them (checked against the file).
extern "C" int fair_matrix_read(const char *path, int64_t nnz, int32_t nia, double *vA, int32_t *JA, int32_t *IA) {
    return guard([&] {
        FAIR_CHECK(path && IA && nia >= 1, FAIR_ERR_ARG, "null argument");
        File f(path, "rb");
        FAIR_CHECK(f.f != nullptr, FAIR_ERR_ARG, std::string("cannot open ") + path);
        get_int_record(f.f);
        get_int_record(f.f);
        const int32_t nz = get_int_record(f.f);
        FAIR_CHECK(nz == nnz && (nz == 0 || (vA && JA)), FAIR_ERR_ARG, "nnz does not match the matrix file");
        FAIR_CHECK(get_record(f.f, nullptr, vA, 8ull * nz) == 8ull * nz, FAIR_ERR_ARG, "bad vA record");
        FAIR_CHECK(get_record(f.f, nullptr, JA, 4ull * nz) == 4ull * nz, FAIR_ERR_ARG, "bad JA record");
        FAIR_CHECK(get_record(f.f, nullptr, IA, 4ull * (uint64_t)nia) == 4ull * (uint64_t)nia, FAIR_ERR_ARG,
                   "IA length does not match the matrix file");
    });
}
