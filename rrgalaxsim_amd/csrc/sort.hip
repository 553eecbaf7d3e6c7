// Radix sort of (cell key, particle index) pairs for the PM path's spatial
// ordering of particles; hipCUB (rocPRIM) device radix sort.
#include <hipcub/hipcub.hpp>

#include "rrg_internal.hpp"

namespace rrg {
namespace pm {

void radix_sort_pairs(hipStream_t s, void* temp, size_t& temp_bytes, const unsigned* keys_in, unsigned* keys_out,
                      const unsigned* vals_in, unsigned* vals_out, int64_t n, int end_bit) {
    RRG_HIP(hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, keys_in, keys_out, vals_in, vals_out, int(n), 0,
                                               end_bit, s));
}

void exclusive_scan_i32(hipStream_t s, void* temp, size_t& temp_bytes, const int* in, int* out, int64_t n) {
    RRG_HIP(hipcub::DeviceScan::ExclusiveSum(temp, temp_bytes, in, out, int(n), s));
}

}  // namespace pm
}  // namespace rrg
