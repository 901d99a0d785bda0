// frg_dedupe_kernels.h -- host entry of the device sort / dedupe / merge (frg_dedupe.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace frg {

// Device scratch needed by sort_dedupe_merge_device for m old and n new roots.
size_t dedupe_temp_bytes(int64_t m, int64_t n);

// Roots are 21 complex doubles (double2).  coll: m old roots sorted on Re z_0; fpt, err: n new
// roots and their codes (duplicates become 6); out: room for m + n roots; *m_out: merged length.
// Requires m + n < 2^31.  Synchronises `st` before returning.
hipError_t sort_dedupe_merge_device(const double2 *coll, int64_t m, const double2 *fpt, int32_t *err, int64_t n,
                                    double tol, double2 *out, int64_t *m_out, void *temp, size_t temp_bytes,
                                    hipStream_t st);

}  // namespace frg
