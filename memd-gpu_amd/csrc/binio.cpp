// Raw float32 sample files, channel-major [n][T], no header: the format the
// reference's demo reads (EEGsample_32CH_30504.bin) and writes (mode_matrix.bin).
//
// Replaces getBinSize / readBin / writeBin (memReducedEpochs_dev.cu:896-932) and
// the length rule T = bytes / sizeof(float) / n (:1486).  Differences: failures
// return status codes (the reference prints and calls exit(-1)), sizes are
// size_t (the reference's int overflows past 2 GiB), and a read of the first T
// samples of every channel of a longer file is one call (runBenchmark's
// per-channel slicing, :1421-1426).
#include "../../include/memd_amd.h"
#include "binio.h"

#include <cerrno>
#include <cstdio>
#include <cstring>
#include <string>
#include <sys/types.h>

namespace memd {
namespace {
thread_local std::string g_err = "no error";
}
const char* global_error() { return g_err.c_str(); }
void set_global_error(const char* msg) { g_err = msg; }
}  // namespace memd

namespace {

memd_status io_fail(memd_status code, const char* what, const char* path) {
    char buf[768];
    std::snprintf(buf, sizeof buf, "%s '%s': %s", what, path ? path : "(null)", errno ? std::strerror(errno) : "");
    memd::set_global_error(buf);
    return code;
}

struct File {
    FILE* f = nullptr;
    ~File() { if (f) std::fclose(f); }
};

bool file_bytes(FILE* f, size_t* out) {
    if (fseeko(f, 0, SEEK_END) != 0) return false;
    long long end = ftello(f);
    if (end < 0 || fseeko(f, 0, SEEK_SET) != 0) return false;
    *out = (size_t)end;
    return true;
}

}  // namespace

extern "C" {

memd_status memd_bin_length(const char* path, size_t n, size_t* T) {
    errno = 0;
    if (!path || !T || n == 0) return io_fail(MEMD_ERR_ARG, "invalid arguments for", path);
    File F;
    F.f = std::fopen(path, "rb");
    if (!F.f) return io_fail(MEMD_ERR_IO, "cannot open", path);
    size_t bytes = 0;
    if (!file_bytes(F.f, &bytes)) return io_fail(MEMD_ERR_IO, "cannot size", path);
    *T = bytes / sizeof(float) / n;   // memReducedEpochs_dev.cu:1486 (floors like the reference)
    return MEMD_OK;
}

memd_status memd_read_bin_f32(const char* path, size_t n, size_t T_file, size_t T, float* h_y) {
    errno = 0;
    if (!path || !h_y || n == 0 || T == 0 || T > T_file) return io_fail(MEMD_ERR_ARG, "invalid arguments for", path);
    File F;
    F.f = std::fopen(path, "rb");
    if (!F.f) return io_fail(MEMD_ERR_IO, "cannot open", path);
    size_t bytes = 0;
    if (!file_bytes(F.f, &bytes)) return io_fail(MEMD_ERR_IO, "cannot size", path);
    if (bytes / sizeof(float) / n < T_file) return io_fail(MEMD_ERR_IO, "file too short for the requested shape:", path);
    for (size_t c = 0; c < n; ++c) {
        if (T != T_file && fseeko(F.f, (off_t)(c * T_file * sizeof(float)), SEEK_SET) != 0)
            return io_fail(MEMD_ERR_IO, "cannot seek in", path);
        if (std::fread(h_y + c * T, sizeof(float), T, F.f) != T) return io_fail(MEMD_ERR_IO, "short read from", path);
    }
    return MEMD_OK;
}

memd_status memd_write_bin_f32(const char* path, const float* h, size_t count) {
    errno = 0;
    if (!path || (!h && count)) return io_fail(MEMD_ERR_ARG, "invalid arguments for", path);
    File F;
    F.f = std::fopen(path, "wb");
    if (!F.f) return io_fail(MEMD_ERR_IO, "cannot create", path);
    if (std::fwrite(h, sizeof(float), count, F.f) != count) return io_fail(MEMD_ERR_IO, "short write to", path);
    if (std::fclose(F.f) != 0) {
        F.f = nullptr;
        return io_fail(MEMD_ERR_IO, "cannot close", path);
    }
    F.f = nullptr;
    return MEMD_OK;
}

}  // extern "C"
