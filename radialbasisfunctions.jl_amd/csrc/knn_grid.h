// Cell size and cell counts of the k-NN's uniform grid (host code, no HIP types, so
// the sizing can be checked on its own).
//
// Invariant: n[d] * s >= ext[d] for every dimension, so the grid covers the bounding
// box and cell_coord's clamp to the last cell only ever absorbs rounding at the upper
// face. Sizing aims at about 3 data points per cell, keeps each dimension at most
// kMaxCellsPerDim cells and the whole grid at most 4 n_data + 16 cells (degenerate
// extents, e.g. a 2-D cloud on a line, would otherwise ask for far more cells).
#pragma once
#include <math.h>
#include <stdint.h>

namespace rbfx {

constexpr int kMaxCellsPerDim = 1 << 20;

inline int64_t knn_grid_size(int dim, const double* ext_in, int64_t n_data, double* s_out, int* n)
{
    double ext[3] = {0.0, 0.0, 0.0};
    double vol = 1.0;
    for (int d = 0; d < dim; ++d) {
        ext[d] = ext_in[d] > 0.0 ? ext_in[d] : 1e-12;
        vol *= ext[d];
    }
    const double target = 3.0;
    const double np = (double)(n_data > 0 ? n_data : 1);
    double s = pow(vol * target / np, 1.0 / dim);
    // the per-dimension cap raises the cell size, so that the capped count still spans
    // the extent (a capped count at the uncapped size would cover only part of it)
    for (int d = 0; d < dim; ++d) {
        const double need = ext[d] / (double)(kMaxCellsPerDim - 1);
        if (s < need) s = need;
    }
    const int64_t max_cells = 4 * (int64_t)np + 16;
    int64_t ncell = 1;
    for (;;) {
        ncell = 1;
        for (int d = 0; d < 3; ++d) {
            n[d] = d < dim ? (int)(floor(ext[d] / s) + 1.0) : 1;
            ncell *= n[d];
        }
        if (ncell <= max_cells) break;
        s *= 1.25;
    }
    *s_out = s;
    return ncell;
}

}  // namespace rbfx
