// Internal interface of the CSR SpMV.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rbfx {

struct SpmvDev {
    int64_t n_rows;
    const int64_t* row_ptr;
    int k;
    const int32_t* col_idx;
    const double* vals;
    int64_t nnz;
    int n_valarrays;
    int n_fields;
    const double* x;
    int64_t x_ld_point, x_ld_field;
    double* y;
    int64_t y_ld_point, y_ld_field, y_ld_val;
    double alpha, beta;
    const int32_t* row_order;  // optional: processing order of the rows (spatial order)
    int rows_permuted;         // col_idx / vals / row_ptr are stored in processing order
    int64_t max_row;           // longest row (0: unknown)
    int mapping;               // 0 automatic, 1 lanes per row reading HBM, 2 staged through LDS,
                               // 3 vector loads per lane (rows of one length)
};

hipError_t launch_spmv(const SpmvDev& P, hipStream_t s);
// whether the staged (LDS) mapping serves this apply with nv value arrays per launch
bool spmv_staged_fits(const SpmvDev& P, int nv);
// entries per lane of the vector mapping for this apply (0: not served: rows of
// varying length, 8 or more fields, or rows longer than 64 lanes' worth)
int vec_entries(const SpmvDev& P, int nv);

// copy fixed-length CSR rows into processing order: out row r = in row row_order[r]
hipError_t launch_permute_rows(int64_t n_rows, int k, const int32_t* row_order, const int32_t* col, const double* vals,
                               int64_t nnz, int n_valarrays, int32_t* out_col, double* out_vals, hipStream_t s);

}  // namespace rbfx
