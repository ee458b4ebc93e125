// memd_cli -- command-line driver over the C ABI, with the behaviour of the
// reference's host programs:
//
//   default   main() (memReducedEpochs_dev.cu:1466-1529): read a float32
//             [n][T] file (T = bytes / 4 / n), decompose it with n = 32, K = 64,
//             9 output slots, 10 sifts, write mode_matrix.bin [N_imf][n][T]
//             float32.
//   --bench   runBenchmark() (:1370-1464): for each length L of the list
//             2049 + 2048 i (i = 0..117) that the file holds, decompose the
//             first L samples of every channel (8 output slots unless --imfs)
//             and print "Channel: n, Length: L, DirVec: K, Time: seconds".
//             With --files A,B,C and --channel-list 32,64,128 it sweeps the
//             files in turn as runBenchmark's file table does (:1377-1383:
//             32, 64 and 128 channels with K = 64, 128, 256, i.e. K = 2n unless
//             --directions is given).
//
// The engine computes in FP64: the float32 samples are widened on the host,
// directions are the FP64 set of GenerateDirectionVector (:1029-1072), and the
// IMFs are rounded to float32 for the output file.  Failures print the
// library's message and exit with the status code.
#include "../include/memd_amd.h"

#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

struct Args {
    std::string input = "./EEGsample_32CH_30504.bin";
    std::string output = "./mode_matrix.bin";
    size_t n = 32, K = 64, imfs = 0, sifts = 10, length = 0;
    int mode = MEMD_MODE_FIX, window = MEMD_WINDOW_REFWARP32, spline = MEMD_SPLINE_NATURAL,
        boundary = MEMD_BOUNDARY_CLAMP;
    bool bench = false, K_given = false;
    std::vector<size_t> lengths, channel_list;
    std::vector<std::string> files;
};

std::vector<std::string> split(const std::string& s) {
    std::vector<std::string> out;
    size_t pos = 0;
    while (pos < s.size()) {
        size_t c = s.find(',', pos);
        if (c == std::string::npos) c = s.size();
        out.push_back(s.substr(pos, c - pos));
        pos = c + 1;
    }
    return out;
}

void usage() {
    std::fprintf(stderr,
                 "usage: memd_cli [--input FILE] [--output FILE] [--channels n] [--directions K]\n"
                 "                [--imfs N] [--sifts S] [--length T] [--mode fix|stop]\n"
                 "                [--window refwarp32|full] [--spline natural|not-a-knot]\n"
                 "                [--boundary clamp|mirror]\n"
                 "                [--bench [--lengths L1,L2,...] [--files F1,F2,... --channel-list n1,n2,...]]\n");
}

bool parse(int argc, char** argv, Args& a) {
    for (int i = 1; i < argc; ++i) {
        std::string k = argv[i];
        auto val = [&](const char** out) {
            if (i + 1 >= argc) return false;
            *out = argv[++i];
            return true;
        };
        const char* v = nullptr;
        if (k == "--bench") { a.bench = true; continue; }
        if (!val(&v)) return false;
        if (k == "--input") a.input = v;
        else if (k == "--output") a.output = v;
        else if (k == "--channels") a.n = std::strtoull(v, nullptr, 10);
        else if (k == "--directions") { a.K = std::strtoull(v, nullptr, 10); a.K_given = true; }
        else if (k == "--imfs") a.imfs = std::strtoull(v, nullptr, 10);
        else if (k == "--sifts") a.sifts = std::strtoull(v, nullptr, 10);
        else if (k == "--length") a.length = std::strtoull(v, nullptr, 10);
        else if (k == "--mode") {
            if (!std::strcmp(v, "fix")) a.mode = MEMD_MODE_FIX;
            else if (!std::strcmp(v, "stop")) a.mode = MEMD_MODE_STOP;
            else return false;
        } else if (k == "--window") {
            if (!std::strcmp(v, "refwarp32")) a.window = MEMD_WINDOW_REFWARP32;
            else if (!std::strcmp(v, "full")) a.window = MEMD_WINDOW_FULL;
            else return false;
        } else if (k == "--spline") {
            if (!std::strcmp(v, "natural")) a.spline = MEMD_SPLINE_NATURAL;
            else if (!std::strcmp(v, "not-a-knot")) a.spline = MEMD_SPLINE_NOT_A_KNOT;
            else return false;
        } else if (k == "--boundary") {
            if (!std::strcmp(v, "clamp")) a.boundary = MEMD_BOUNDARY_CLAMP;
            else if (!std::strcmp(v, "mirror")) a.boundary = MEMD_BOUNDARY_MIRROR;
            else return false;
        } else if (k == "--lengths") {
            for (const std::string& t : split(v)) a.lengths.push_back(std::strtoull(t.c_str(), nullptr, 10));
        } else if (k == "--files") {
            a.files = split(v);
        } else if (k == "--channel-list") {
            for (const std::string& t : split(v)) a.channel_list.push_back(std::strtoull(t.c_str(), nullptr, 10));
        } else {
            return false;
        }
    }
    return true;
}

int die(memd_ctx* ctx, int rc, const char* what) {
    std::fprintf(stderr, "memd_cli: %s failed (status %d): %s\n", what, rc, memd_last_error(ctx));
    return rc;
}

template <class T>
struct DevBuf {
    T* p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
    bool alloc(size_t count) { return hipMalloc((void**)&p, count * sizeof(T)) == hipSuccess; }
};

// One decomposition of h_y [n][T]; writes the float32 IMF stack to `output`.
int run_one(memd_ctx* ctx, const Args& a, size_t n, size_t K, size_t T, size_t N_imf, const std::vector<float>& h_y,
            double* secs) {
    std::vector<double> y64(h_y.begin(), h_y.end());
    std::vector<int32_t> h_x(T);
    for (size_t i = 0; i < T; ++i) h_x[i] = (int32_t)i;   // sample index 0..T-1 (:1504-1506)
    DevBuf<double> d_y, d_V, d_IMFs;
    DevBuf<int32_t> d_x;
    if (!d_y.alloc(n * T) || !d_V.alloc(K * n) || !d_IMFs.alloc(N_imf * n * T) || !d_x.alloc(T)) {
        std::fprintf(stderr, "memd_cli: device allocation failed\n");
        return MEMD_ERR_NOMEM;
    }
    if (hipMemcpy(d_y.p, y64.data(), n * T * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_x.p, h_x.data(), T * sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess) {
        std::fprintf(stderr, "memd_cli: upload failed\n");
        return MEMD_ERR_HIP;
    }
    int rc = memd_directions_f64(ctx, n, K, nullptr, d_V.p);
    if (rc != MEMD_OK) return die(ctx, rc, "memd_directions_f64");
    memd_opts o;
    memd_opts_default(&o);
    o.mode = a.mode;
    o.window = a.window;
    o.spline = a.spline;
    o.boundary = a.boundary;
    rc = memd_run_f64(ctx, n, T, K, N_imf, a.sifts, d_x.p, d_y.p, d_V.p, d_IMFs.p, &o, secs);
    if (rc != MEMD_OK) return die(ctx, rc, "memd_run_f64");
    std::vector<double> imfs(N_imf * n * T);
    if (hipMemcpy(imfs.data(), d_IMFs.p, imfs.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
        std::fprintf(stderr, "memd_cli: download failed\n");
        return MEMD_ERR_HIP;
    }
    std::vector<float> out(imfs.begin(), imfs.end());
    rc = memd_write_bin_f32(a.output.c_str(), out.data(), out.size());
    if (rc != MEMD_OK) return die(nullptr, rc, "memd_write_bin_f32");
    return MEMD_OK;
}

// Decompose the file's first L samples per channel for every length of the list
// (the bench skips lengths past the file, like runBenchmark's list).
int run_file(memd_ctx* ctx, const Args& a, const std::string& path, size_t n, size_t K) {
    size_t T_file = 0;
    int rc = memd_bin_length(path.c_str(), n, &T_file);
    if (rc != MEMD_OK) return die(nullptr, rc, "memd_bin_length");
    std::vector<size_t> lengths;
    size_t N_imf = a.imfs;
    if (a.bench) {
        if (!N_imf) N_imf = 8;   // runBenchmark's num_IMFs (:1372)
        if (!a.lengths.empty()) lengths = a.lengths;
        else
            for (size_t i = 0; i < 118; ++i) lengths.push_back(2049 + 2048 * i);   // :1383-1395
    } else {
        if (!N_imf) N_imf = 9;   // main()'s num_IMFs (:1475)
        lengths.push_back(a.length ? a.length : T_file);
    }
    for (size_t L : lengths) {
        if (L > T_file) {
            if (a.bench) continue;   // the list runs past the file: those lengths are skipped
            std::fprintf(stderr, "memd_cli: --length %zu exceeds the file's %zu samples per channel\n", L, T_file);
            return MEMD_ERR_ARG;
        }
        std::vector<float> h_y(n * L);
        rc = memd_read_bin_f32(path.c_str(), n, T_file, L, h_y.data());
        if (rc != MEMD_OK) return die(nullptr, rc, "memd_read_bin_f32");
        double secs = 0.0;
        rc = run_one(ctx, a, n, K, L, N_imf, h_y, &secs);
        if (rc != MEMD_OK) return rc;
        std::printf("Channel: %zu, Length: %zu, DirVec: %zu, Time: %f \n", n, L, K, secs);
        std::fflush(stdout);
    }
    return MEMD_OK;
}

}  // namespace

int main(int argc, char** argv) {
    Args a;
    if (!parse(argc, argv, a) || a.n == 0 || a.K == 0 || a.sifts == 0 ||
        (!a.files.empty() && (!a.bench || a.files.size() != a.channel_list.size()))) {
        usage();
        return MEMD_ERR_ARG;
    }
    for (size_t c : a.channel_list)
        if (c == 0) {
            usage();
            return MEMD_ERR_ARG;
        }
    // every input file is checked before the device is touched
    const std::vector<std::string> inputs = a.files.empty() ? std::vector<std::string>{a.input} : a.files;
    for (size_t f = 0; f < inputs.size(); ++f) {
        size_t T_file = 0;
        int rc = memd_bin_length(inputs[f].c_str(), a.files.empty() ? a.n : a.channel_list[f], &T_file);
        if (rc != MEMD_OK) return die(nullptr, rc, "memd_bin_length");
    }
    memd_ctx* ctx = nullptr;
    int rc = memd_ctx_create(0, &ctx);
    if (rc != MEMD_OK) return die(nullptr, rc, "memd_ctx_create");
    if (a.files.empty()) {
        rc = run_file(ctx, a, a.input, a.n, a.K);
    } else {   // runBenchmark's file table: K = 2n per file unless --directions
        for (size_t f = 0; f < a.files.size() && rc == MEMD_OK; ++f) {
            const size_t n = a.channel_list[f];
            rc = run_file(ctx, a, a.files[f], n, a.K_given ? a.K : 2 * n);
        }
    }
    memd_ctx_destroy(ctx);
    return rc;
}
