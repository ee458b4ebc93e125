// MATLAB gateway over the engine's C ABI, with the argument list of the
// reference's memdMatlabCuda MEX function (memdMatlabCuda.cu:996-1039) as
// assistMatlabCuda.m:3-17 calls it:
//
//   IMFs = memd_amd_mex(int32(n), int32(K), int32(N_imf), int32(S), data,
//                       int32(0:T-1), int32(primes(1:n)), zeros(T, n, N_imf, 'like', data))
//
// data is T x n (memory [n][T]), single or double; the 8th argument only gives
// the output's dimensions and class (memdMatlabCuda.cu:1031-1035); the result
// has those dimensions (memory [N_imf][n][T]).  MATLAB's gpuArray is CUDA-only,
// so on a ROCm host the arguments are ordinary MATLAB arrays (gather() any
// gpuArray first); the gateway uploads them, runs memd_mex_compat_f32/_f64
// (FP64 inside) and downloads the result.  Shape and class errors raise
// MATLAB errors (the reference does not check its arguments).
//
// Builds only where MATLAB's mex.h exists, e.g.
//   mex -R2018a memd_amd_mex.cpp -I<repo>/include -I/opt/rocm/include -D__HIP_PLATFORM_AMD__ ...
//       -L<repo>/memd-gpu_amd -lmemd_amd -L/opt/rocm/lib -lamdhip64
#include "mex.h"

#include "../../include/memd_amd.h"

#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstring>

namespace {

memd_ctx* g_ctx = nullptr;

void release_context() {
    if (g_ctx) memd_ctx_destroy(g_ctx);
    g_ctx = nullptr;
}

int32_t int32_scalar(const mxArray* a, const char* name) {
    if (!mxIsInt32(a) || mxGetNumberOfElements(a) != 1 || mxIsComplex(a))
        mexErrMsgIdAndTxt("memd_amd:arg", "%s must be an int32 scalar", name);
    return *static_cast<const int32_t*>(mxGetData(a));
}

// Device buffers released on every path before a MATLAB error is raised
// (mexErrMsgIdAndTxt does not return).
struct Buffers {
    void* data = nullptr;
    void* out = nullptr;
    void* index = nullptr;
    void* primes = nullptr;
    void release() {
        for (void** p : {&data, &out, &index, &primes})
            if (*p) (void)hipFree(*p), *p = nullptr;
    }
};

}  // namespace

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
    if (nrhs != 8) mexErrMsgIdAndTxt("memd_amd:arg", "expected 8 inputs (n, K, N_imf, S, data, index, primes, output)");
    if (nlhs > 1) mexErrMsgIdAndTxt("memd_amd:arg", "one output");
    const int32_t n = int32_scalar(prhs[0], "n");
    const int32_t K = int32_scalar(prhs[1], "K");
    const int32_t N_imf = int32_scalar(prhs[2], "N_imf");
    const int32_t S = int32_scalar(prhs[3], "S");
    const mxArray* data = prhs[4];
    const mxArray* index = prhs[5];
    const mxArray* primes = prhs[6];
    const mxArray* tmpl = prhs[7];
    const bool f32 = mxIsSingle(data);
    if ((!f32 && !mxIsDouble(data)) || mxIsComplex(data))
        mexErrMsgIdAndTxt("memd_amd:arg", "data must be a real single or double array");
    if (n < 1 || K < 1 || N_imf < 1 || S < 1) mexErrMsgIdAndTxt("memd_amd:arg", "n, K, N_imf and S must be >= 1");
    const size_t numel = mxGetNumberOfElements(data);
    if (numel % (size_t)n != 0) mexErrMsgIdAndTxt("memd_amd:arg", "numel(data) must be a multiple of n");
    const size_t T = numel / (size_t)n;   // memdMatlabCuda.cu:1038-1039
    if (!mxIsInt32(index) || mxGetNumberOfElements(index) != T)
        mexErrMsgIdAndTxt("memd_amd:arg", "index must be int32 with numel(data)/n elements");
    if (!mxIsInt32(primes) || mxGetNumberOfElements(primes) < (size_t)n)
        mexErrMsgIdAndTxt("memd_amd:arg", "primes must be int32 with at least n elements");
    if (mxGetClassID(tmpl) != mxGetClassID(data) || mxGetNumberOfElements(tmpl) != (size_t)N_imf * numel)
        mexErrMsgIdAndTxt("memd_amd:arg", "output template must match data's class and hold N_imf*numel(data) values");

    if (!g_ctx) {
        if (memd_ctx_create(-1, &g_ctx) != MEMD_OK) mexErrMsgIdAndTxt("memd_amd:hip", "%s", memd_last_error(nullptr));
        mexAtExit(release_context);
    }
    const size_t esz = f32 ? sizeof(float) : sizeof(double);
    const size_t out_numel = (size_t)N_imf * numel;
    Buffers b;
    char msg[600] = {0};
    if (hipMalloc(&b.data, numel * esz) != hipSuccess || hipMalloc(&b.out, out_numel * esz) != hipSuccess ||
        hipMalloc(&b.index, T * sizeof(int32_t)) != hipSuccess || hipMalloc(&b.primes, n * sizeof(int32_t)) != hipSuccess) {
        b.release();
        mexErrMsgIdAndTxt("memd_amd:nomem", "device allocation failed");
    }
    if (hipMemcpy(b.data, mxGetData(data), numel * esz, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b.index, mxGetData(index), T * sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b.primes, mxGetData(primes), n * sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess) {
        b.release();
        mexErrMsgIdAndTxt("memd_amd:hip", "upload failed");
    }
    memd_status rc =
        f32 ? memd_mex_compat_f32(g_ctx, n, K, N_imf, S, numel, static_cast<const float*>(b.data),
                                  static_cast<const int32_t*>(b.index), static_cast<const int32_t*>(b.primes),
                                  static_cast<float*>(b.out))
            : memd_mex_compat_f64(g_ctx, n, K, N_imf, S, numel, static_cast<const double*>(b.data),
                                  static_cast<const int32_t*>(b.index), static_cast<const int32_t*>(b.primes),
                                  static_cast<double*>(b.out));
    if (rc != MEMD_OK) {
        std::snprintf(msg, sizeof msg, "%s", memd_last_error(g_ctx));
        b.release();
        mexErrMsgIdAndTxt("memd_amd:run", "%s", msg);
    }
    plhs[0] = mxCreateNumericArray(mxGetNumberOfDimensions(tmpl), mxGetDimensions(tmpl), mxGetClassID(tmpl), mxREAL);
    hipError_t e = hipMemcpy(mxGetData(plhs[0]), b.out, out_numel * esz, hipMemcpyDeviceToHost);
    b.release();
    if (e != hipSuccess) mexErrMsgIdAndTxt("memd_amd:hip", "download failed");
}
