// gf_textio.hip — the reference's per-cell text files (inoutput.cpp:11-327
// reads them, :330-864 writes them): one line per cell, "i  j  k" with
// 1-based indices, i outermost and k innermost, then each value as "%.7e",
// two spaces apart.  Host code only; no context is involved, so failures
// are reported through gf_last_error(NULL).
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "gf_internal.h"

using namespace gf;

namespace {

bool dims_ok(const int64_t dims[3], int32_t ncomp, int64_t* n) {
    if (!dims || ncomp < 1 || ncomp > 64) return false;
    if (dims[0] < 1 || dims[1] < 1 || dims[2] < 1) return false;
    if (dims[0] > (1LL << 20) || dims[1] > (1LL << 20) || dims[2] > (1LL << 20)) return false;
    *n = dims[0] * dims[1] * dims[2];
    return *n <= (1LL << 34);
}

bool is_space(char ch) { return ch == ' ' || ch == '\t' || ch == '\n' || ch == '\r' || ch == '\v' || ch == '\f'; }

}  // namespace

extern "C" {

gf_status gf_write_cells(const char* path, const int64_t dims[3], int32_t ncomp, const double* data) {
    int64_t n = 0;
    if (!path || !data || !dims_ok(dims, ncomp, &n)) return set_thread_error(GF_ERR_ARG, "gf_write_cells: bad argument");
    FILE* f = std::fopen(path, "w");
    if (!f) return set_thread_error(GF_ERR_ARG, std::string("cannot open ") + path + " for writing");
    std::vector<char> buf(1 << 20);
    size_t used = 0;
    bool ok = true;
    char line[64 * 24 + 96];
    const int64_t nyz = dims[1] * dims[2];
    for (int64_t i = 0; i < dims[0] && ok; i++)
        for (int64_t j = 0; j < dims[1] && ok; j++)
            for (int64_t k = 0; k < dims[2] && ok; k++) {
                const int64_t id = i * nyz + j * dims[2] + k;
                int len = std::snprintf(line, sizeof(line), "%lld  %lld  %lld", static_cast<long long>(i + 1),
                                        static_cast<long long>(j + 1), static_cast<long long>(k + 1));
                for (int32_t c = 0; c < ncomp; c++)
                    len += std::snprintf(line + len, sizeof(line) - static_cast<size_t>(len), "  %.7e",
                                         data[static_cast<size_t>(c) * static_cast<size_t>(n) + static_cast<size_t>(id)]);
                line[len++] = '\n';
                if (used + static_cast<size_t>(len) > buf.size()) {
                    ok = std::fwrite(buf.data(), 1, used, f) == used;
                    used = 0;
                }
                std::memcpy(buf.data() + used, line, static_cast<size_t>(len));
                used += static_cast<size_t>(len);
            }
    if (ok && used) ok = std::fwrite(buf.data(), 1, used, f) == used;
    if (std::fclose(f) != 0) ok = false;
    if (!ok) return set_thread_error(GF_ERR_ARG, std::string("writing ") + path + " failed");
    return GF_OK;
}

gf_status gf_read_cells(const char* path, const int64_t dims[3], int32_t ncomp, double* data) {
    int64_t n = 0;
    if (!path || !data || !dims_ok(dims, ncomp, &n)) return set_thread_error(GF_ERR_ARG, "gf_read_cells: bad argument");
    FILE* f = std::fopen(path, "rb");
    if (!f) return set_thread_error(GF_ERR_ARG, std::string("cannot open ") + path);
    std::vector<char> text;
    bool ok = std::fseek(f, 0, SEEK_END) == 0;
    const long len = ok ? std::ftell(f) : -1;
    ok = ok && len >= 0 && std::fseek(f, 0, SEEK_SET) == 0;
    if (ok) {
        text.resize(static_cast<size_t>(len) + 1);
        ok = std::fread(text.data(), 1, static_cast<size_t>(len), f) == static_cast<size_t>(len);
        text[static_cast<size_t>(len)] = '\0';
    }
    std::fclose(f);
    if (!ok) return set_thread_error(GF_ERR_ARG, std::string("reading ") + path + " failed");
    const std::string where = std::string(path) + ": ";
    std::vector<unsigned char> seen(static_cast<size_t>(n), 0);
    const char* p = text.data();
    const int64_t nyz = dims[1] * dims[2];
    for (int64_t rec = 0; rec < n; rec++) {
        long long idx[3];
        while (is_space(*p)) p++;
        if (!*p)
            return set_thread_error(GF_ERR_ARG, where + std::to_string(rec) + " records for " + std::to_string(n) +
                                                    " cells");
        for (int a = 0; a < 3; a++) {
            while (is_space(*p)) p++;
            char* end = nullptr;
            errno = 0;
            idx[a] = std::strtoll(p, &end, 10);
            if (end == p || errno || (*end && !is_space(*end)))
                return set_thread_error(GF_ERR_ARG, where + "record " + std::to_string(rec + 1) +
                                                        ": expected an integer cell index");
            p = end;
        }
        for (int a = 0; a < 3; a++)
            if (idx[a] < 1 || idx[a] > dims[a])
                return set_thread_error(GF_ERR_ARG, where + "record " + std::to_string(rec + 1) +
                                                        ": cell index outside the grid");
        const int64_t id = (idx[0] - 1) * nyz + (idx[1] - 1) * dims[2] + (idx[2] - 1);
        if (seen[static_cast<size_t>(id)])
            return set_thread_error(GF_ERR_ARG, where + "record " + std::to_string(rec + 1) + ": cell given twice");
        seen[static_cast<size_t>(id)] = 1;
        for (int32_t c = 0; c < ncomp; c++) {
            while (is_space(*p)) p++;
            char* end = nullptr;
            const double v = std::strtod(p, &end);
            if (end == p || (*end && !is_space(*end)))
                return set_thread_error(GF_ERR_ARG, where + "record " + std::to_string(rec + 1) +
                                                        ": expected " + std::to_string(ncomp) + " values");
            data[static_cast<size_t>(c) * static_cast<size_t>(n) + static_cast<size_t>(id)] = v;
            p = end;
        }
    }
    while (is_space(*p)) p++;
    if (*p) return set_thread_error(GF_ERR_ARG, where + "more records than the " + std::to_string(n) + " cells");
    return GF_OK;
}

}  // extern "C"
