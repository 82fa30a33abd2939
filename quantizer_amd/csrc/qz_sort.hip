// Stable radix sorts and scans on the device (rocPRIM, hipCUB).
// Kept in its own translation unit: rocPRIM's radix sort is slow to compile
// and never changes with the rest of the library.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <rocprim/rocprim.hpp>

#include "qz_internal.h"

// 32-bit keys with int values: rocPRIM's onesweep radix sort with 9 bits per
// pass and 1024-thread blocks of 8 items each, ranked by warp matching.  The
// library's defaults for gfx950 fall back to an untuned configuration; this one
// sorts 2^24 pairs of 27-bit keys (the headline's head cell and one octant
// digit) in three passes, 0.39 ms on MI355X, against 0.51 ms for 8 bits per
// pass and 0.56 ms for hipCUB's default, with identical output (the sort is
// stable either way).
using SortConfig32 = rocprim::radix_sort_config<
    rocprim::default_config, rocprim::default_config,
    rocprim::radix_sort_onesweep_config<rocprim::kernel_config<1024, 8>, rocprim::kernel_config<1024, 8>, 9,
                                        rocprim::block_radix_rank_algorithm::match>>;

size_t qz_sort_temp_bytes(int n, int end_bit)
{
  size_t bytes = 0;
  (void)rocprim::radix_sort_pairs<SortConfig32>(nullptr, bytes, (const unsigned *)nullptr, (unsigned *)nullptr,
                                                (const int *)nullptr, (int *)nullptr, (size_t)n, 0, end_bit);
  return bytes;
}

hipError_t qz_sort_pairs(void *tmp, size_t bytes, const unsigned *kin, unsigned *kout, const int *vin, int *vout,
                         int n, int end_bit, hipStream_t s)
{
  return rocprim::radix_sort_pairs<SortConfig32>(tmp, bytes, kin, kout, vin, vout, (size_t)n, 0, end_bit, s);
}

size_t qz_sort64_temp_bytes(int n, int end_bit)
{
  size_t bytes = 0;
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const unsigned long long *)nullptr,
                                     (unsigned long long *)nullptr, (const int *)nullptr, (int *)nullptr, n, 0,
                                     end_bit);
  return bytes;
}

hipError_t qz_sort64_pairs(void *tmp, size_t bytes, const unsigned long long *kin, unsigned long long *kout,
                           const int *vin, int *vout, int n, int end_bit, hipStream_t s)
{
  return hipcub::DeviceRadixSort::SortPairs(tmp, bytes, kin, kout, vin, vout, n, 0, end_bit, s);
}

size_t qz_scan_temp_bytes(int n)
{
  size_t bytes = 0;
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, (const int *)nullptr, (int *)nullptr, n);
  return bytes;
}

hipError_t qz_exclusive_scan(void *tmp, size_t bytes, const int *in, int *out, int n, hipStream_t s)
{
  return hipcub::DeviceScan::ExclusiveSum(tmp, bytes, in, out, n, s);
}
