// The product handle of a block-compressed operator built from block metadata whose factors are
// written by the caller (shared by pq_hmatrix_create and the device construction in aca.hip).
#pragma once
#include "pq_common.h"
#include <functional>

namespace pq {

// meta: n_blocks rows of (a_off, la, b_off, lb, rank (-1 dense), same) into idx; the factors, n_factor
// elements in block order (dense: la x lb row-major; low rank: U la x rank, then V rank x lb, both
// row-major), are written by upload(device pointer) once the buffer exists.
pq_hmatrix *hmatrix_make(int device, int64_t n, int64_t n_blocks, const int64_t *meta, const int32_t *idx,
                         int64_t n_idx, int64_t n_factor, bool is_complex,
                         const std::function<void(void *)> &upload);

}  // namespace pq
