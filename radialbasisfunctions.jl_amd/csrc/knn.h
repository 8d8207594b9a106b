// Internal interface of the grid-hashed k-NN.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rbfx {

// Grow-only device scratch for binning (owned by the library, one per device).
struct KnnWorkspace {
    int32_t *count = nullptr, *start = nullptr, *cursor = nullptr, *qcount = nullptr, *qstart = nullptr;
    int32_t *cell = nullptr, *sidx = nullptr, *qcell = nullptr, *qorder = nullptr;
    double *sx = nullptr, *bbox = nullptr;
    void* scan_tmp = nullptr;
    size_t cap_count = 0, cap_start = 0, cap_cursor = 0, cap_qcount = 0, cap_qstart = 0;
    size_t cap_cell = 0, cap_sidx = 0, cap_qcell = 0, cap_qorder = 0, cap_sx = 0, cap_bbox = 0;
    size_t scan_tmp_bytes = 0;
    hipError_t reserve(int64_t ncell, int64_t n_data, int64_t n_query, int dim);
};

hipError_t run_knn(KnnWorkspace& ws, const double* data, int64_t n_data, const double* query, int64_t n_query,
                   int dim, int k, const int64_t* q_ptr, int32_t* out, hipStream_t st);

}  // namespace rbfx
