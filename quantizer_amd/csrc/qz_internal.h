// Internal declarations shared by the translation units of libqzhip.so.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

size_t qz_sort_temp_bytes(int n, int end_bit);
hipError_t qz_sort_pairs(void *tmp, size_t bytes, const unsigned *kin, unsigned *kout, const int *vin, int *vout,
                         int n, int end_bit, hipStream_t s);
size_t qz_sort64_temp_bytes(int n, int end_bit);
hipError_t qz_sort64_pairs(void *tmp, size_t bytes, const unsigned long long *kin, unsigned long long *kout,
                           const int *vin, int *vout, int n, int end_bit, hipStream_t s);
size_t qz_scan_temp_bytes(int n);
hipError_t qz_exclusive_scan(void *tmp, size_t bytes, const int *in, int *out, int n, hipStream_t s);
