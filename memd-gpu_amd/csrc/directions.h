#pragma once
#include <stdint.h>
#include <stddef.h>

namespace memd {
// P[0..n-1] = [-K, 2, 3, 5, ...] (memReducedEpochs_dev.cu:1033-1046)
void reference_primes(int n, int K, int32_t* P);
// radical inverse of i in the given base, FP64 (memReducedEpochs_dev.cu:934-946)
double radical_inverse(int base, unsigned i);
// P[0] != 0 and P[j >= 1] >= 2 (bases below 2 never terminate the digit loop)
bool primes_valid(int n, const int32_t* P);
// V [K][n] unit direction vectors from the prime table P
void directions_host(int n, int K, const int32_t* P, double* V);
}  // namespace memd
