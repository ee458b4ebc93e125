// Hammersley direction vectors on the host, FP64.
//
// Follows GenerateDirectionVector / GenerateHammSeq / RadicalInverse /
// Sequence2Vector (memReducedEpochs_dev.cu:934-1072; MEX twin
// memdMatlabCuda.cu:915-978).  Computed once per (n, K, primes) with the host's
// libm and cached on the device by the context, so no transcendental function
// runs on the GPU.
#include "directions.h"

#include <cmath>
#include <vector>

namespace memd {

static bool composite(int v) {   // isPrimes (memReducedEpochs_dev.cu:1014-1027)
    for (int i = 2; i <= v / 2; ++i)
        if (v % i == 0) return true;
    return false;
}

void reference_primes(int n, int K, int32_t* P) {
    // memReducedEpochs_dev.cu:1033-1046: P[0] = -K, then the first n-1 primes
    int filled = 1, cand = 2;
    while (filled < n) {
        if (!composite(cand)) P[filled++] = cand;
        ++cand;
    }
    P[0] = -K;
}

double radical_inverse(int base, unsigned i) {
    // RadicalInverse (:934-946) in FP64
    double digit = 1.0 / (double)base;
    const double radical = digit;
    double inv = 0.0;
    while (i) {
        inv = inv + digit * (double)(i % (unsigned)base);
        digit = digit * radical;
        i /= (unsigned)base;
    }
    return inv;
}

bool primes_valid(int n, const int32_t* P) {
    if (P[0] == 0) return false;
    for (int j = 1; j < n; ++j)
        if (P[j] < 2) return false;
    return true;
}

void directions_host(int n, int K, const int32_t* P, double* V) {
    std::vector<double> b(n), ang(n > 1 ? n - 1 : 1);
    for (int k = 0; k < K; ++k) {
        // Hammersley coordinates (GenerateHammSeq :992-1005)
        b[0] = ((double)k + 1.5) / (double)(-P[0]);
        for (int j = 1; j < n; ++j) b[j] = radical_inverse(P[j], (unsigned)(k + 1));
        // to the unit sphere (Sequence2Vector :948-990)
        for (int i = 0; i < n; ++i) b[i] = 2.0 * b[i] - 1.0;
        for (int i = 0; i + 1 < n; ++i) {
            double tail = 0.0;
            for (int j = n - 1; j > i; --j) tail = tail + b[j] * b[j];
            ang[i] = std::atan2(std::sqrt(tail), b[i]);
        }
        double* v = V + (size_t)k * n;
        v[0] = 1.0;
        for (int i = 1; i < n; ++i) {
            double prod = 1.0;
            for (int j = 0; j < i; ++j) prod = prod * std::sin(ang[j]);
            v[i] = prod;
        }
        for (int i = 0; i + 1 < n; ++i) v[i] = v[i] * std::cos(ang[i]);
    }
}

}  // namespace memd
