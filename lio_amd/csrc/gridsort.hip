// Stable key/value radix sort for the device grid build (hipCUB), kept in its
// own translation unit so the engine does not compile the library templates.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include "xcq.h"

namespace xcq {

// Sorts n (key, value) pairs by the low `bits` bits of the key, stably
// (equal keys keep their input order).  With temp == nullptr only sets
// *temp_bytes.  Returns the HIP status.
int radix_sort_pairs(void* temp, size_t* temp_bytes, const unsigned* keys_in,
                     unsigned* keys_out, const unsigned* vals_in, unsigned* vals_out, int n,
                     int bits, void* stream) {
  return (int)hipcub::DeviceRadixSort::SortPairs(temp, *temp_bytes, keys_in, keys_out, vals_in,
                                                 vals_out, n, 0, bits, (hipStream_t)stream);
}

}  // namespace xcq
