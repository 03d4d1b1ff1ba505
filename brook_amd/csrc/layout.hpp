#pragma once
#include <algorithm>
#include <vector>

#include "stages.hpp"

namespace brook {

// out (Fortran order, `bytes` wide) <- the device volume at `base` with the given byte strides.
void gather_fortran(Context &ctx, const void *base, const long long byte_strides[3], int bytes, Dims d, void *out);
// out <- in where the value is one of `ids` (two's-complement bit patterns), else 0.
void mask_labels(Context &ctx, const void *in, int bytes, long long n, const std::vector<uint64_t> &ids, void *out);

}  // namespace brook
