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
                               // 3 vector loads per lane (rows of one length), 4 source tiles
    // Source tiles (optional, rows stored in processing order): tile t holds rows
    // [t * tile_rows, (t + 1) * tile_rows); its distinct columns are
    // tile_src[tile_src_off[t] .. + tile_src_n[t]) and entry q's column is
    // tile_src[tile_src_off[t] + tile_col[q]]. Many-field applies stage a tile's X rows
    // in LDS once instead of gathering them per entry.
    int tile_rows;
    const int32_t* tile_src;
    const int64_t* tile_src_off;
    const int32_t* tile_src_n;
    const uint16_t* tile_col;
};

hipError_t launch_spmv(const SpmvDev& P, hipStream_t s);
// Source tiles of rows stored back to back (fixed k or row_ptr), tile_rows rows per tile
// with at most kTileEntries entries: phase 0 writes the number of distinct columns per
// tile (tile_n); phase 1 (tile_off = their exclusive prefix sum) writes the sorted lists
// (tile_src) and every entry's position in its tile's list (tile_col).
constexpr int kTileEntries = 4096;
hipError_t launch_source_tiles(int64_t n_rows, int k, const int64_t* row_ptr, const int32_t* col, int tile_rows,
                               int phase, int32_t* tile_n, const int64_t* tile_off, int32_t* tile_src,
                               uint16_t* tile_col, hipStream_t s);
// entries of a tile a many-field source-tile workgroup holds in LDS for nv value arrays
// (tiles with more gather X directly); with 32 KB of X rows this keeps the workgroup
// within 64 KB
inline int tile_entries_lds(int nv) { return nv <= 2 ? 1024 : 512; }
// whether the staged (LDS) mapping serves this apply with nv value arrays per launch
bool spmv_staged_fits(const SpmvDev& P, int nv);
// entries per lane of the vector mapping for this apply (0: not served: rows of
// varying length, 8 or more fields, or rows longer than 64 lanes' worth)
int vec_entries(const SpmvDev& P, int nv);

// copy fixed-length CSR rows into processing order: out row r = in row row_order[r]
hipError_t launch_permute_rows(int64_t n_rows, int k, const int32_t* row_order, const int32_t* col, const double* vals,
                               int64_t nnz, int n_valarrays, int32_t* out_col, double* out_vals, hipStream_t s);

}  // namespace rbfx
