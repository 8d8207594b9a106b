// Monomial values and derivatives for runtime exponent tables (graded order).
// Restates the reference's monomial operators (src/basis/monomial.jl:6-123,
// src/operators/monomial/monomial.jl:13-148, partial.jl:1-402) for one monomial
// x^e at a time; the exponent table replaces the hand-coded orderings, which only
// permute the polynomial block (the weights do not depend on that order).
#pragma once
#include <hip/hip_runtime.h>

#include "build.h"
#include "rbfx.h"

namespace rbfx {

// ---------------------------------------------------------------- monomials
// x^p for a runtime exponent 0 <= p <= kMaxPolyDeg, by repeated multiplication
__device__ __forceinline__ double ipow(double x, int p)
{
    double v = 1.0;
#pragma unroll
    for (int q = 0; q < kMaxPolyDeg; ++q)
        if (q < p) v *= x;
    return v;
}

// Exponent of x_d in the a-th monomial of the graded order fill_mono_table writes (total
// degree ascending; within one degree, the first exponent descending, then the second):
// a compile-time constant wherever a and d are, so a monomial's value is a fixed
// sequence of multiplies.
__host__ __device__ constexpr int graded_exp(int dim, int a, int d)
{
    int c = 0;
    for (int tot = 0; tot <= kMaxPolyDeg; ++tot) {
        if (dim == 1) {
            if (c++ == a) return d == 0 ? tot : 0;
        } else if (dim == 2) {
            for (int x = tot; x >= 0; --x)
                if (c++ == a) return d == 0 ? x : (d == 1 ? tot - x : 0);
        } else {
            for (int x = tot; x >= 0; --x)
                for (int y = tot - x; y >= 0; --y)
                    if (c++ == a) return d == 0 ? x : (d == 1 ? y : tot - x - y);
        }
    }
    return 0;
}

// value of the a-th graded monomial at x (a a compile-time constant after unrolling)
template <int DIM>
__device__ __forceinline__ double mono_graded(const int a, const double (&x)[DIM])
{
    double v = 1.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
#pragma unroll
        for (int q = 0; q < kMaxPolyDeg; ++q)
            if (q < graded_exp(DIM, a, d)) v *= x[d];
    }
    return v;
}

// operator term `kind` applied to the monomial x^e
template <int DIM>
__device__ __forceinline__ double mono_op(const int (&e)[3], int kind, int d1, int d2, const double (&x)[DIM])
{
    if (kind == RBFX_OP_MIXED && d1 == d2) kind = RBFX_OP_D2;
    if (kind == RBFX_OP_LAP) {
        double acc = 0.0;
#pragma unroll
        for (int dd = 0; dd < DIM; ++dd) {
            if (e[dd] >= 2) {
                double t = (double)(e[dd] * (e[dd] - 1));
#pragma unroll
                for (int d = 0; d < DIM; ++d) t *= ipow(x[d], d == dd ? e[d] - 2 : e[d]);
                acc += t;
            }
        }
        return acc;
    }
    double coef = 1.0, val = 1.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        int o = 0;
        if (kind == RBFX_OP_D1 && d == d1) o = 1;
        if (kind == RBFX_OP_D2 && d == d1) o = 2;
        if (kind == RBFX_OP_MIXED && (d == d1 || d == d2)) o = 1;
        if (o > e[d]) return 0.0;
        if (o == 1) coef *= (double)e[d];
        if (o == 2) coef *= (double)(e[d] * (e[d] - 1));
        val *= ipow(x[d], e[d] - o);
    }
    return coef * val;
}

template <int DIM>
__device__ __forceinline__ double mono_normal(const int (&e)[3], const double (&n)[DIM], const double (&x)[DIM])
{
    double acc = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) acc += n[d] * mono_op<DIM>(e, RBFX_OP_D1, d, 0, x);
    return acc;
}

}  // namespace rbfx
