// raster_span.hpp -- exact covered column range of one rasterizer piece in one row of a tile.
// Host and device code (the same function runs in the fine raster and in host checks).
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>

// Edge i passes at tile column x (pixel centre X = X0 + 256 x) iff E_i(x) = K_i + dA_i x > 0,
// with K_i = A_i X0 + B_i Y + C_i + tl_i (the integer form of "E > 0, or E == 0 on a top-left
// edge") and dA_i = 256 A_i.  Inside the guard band |E| < 2^53 (|A|, |B| < 2^26 and the
// offsets of a pixel centre from a vertex below 2^25.3 in 1/256 pixel), so K, dA and E are exact
// in double and fma(dA, x, K) > 0 is the exact test.  t_i estimates the edge's crossing
// -K_i / dA_i within 1/16 column; the boundary column next to it is settled by at most two
// exact tests; the three edges are evaluated without branches (selects), since neighbouring
// pieces take different paths.
struct SpanEdges {
    double K[3];    // E_i at column 0 of the current row
    double dA[3];
    double t[3];    // estimate of -K_i / dA_i (any value for dA_i == 0)
};

__host__ __device__ __forceinline__ void row_span_exact(const SpanEdges& e, int& lo, int& hi) {
    bool empty = lo > hi;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const double K = e.K[i], dA = e.dA[i];
        const bool pos = dA > 0.0, neg = dA < 0.0;
        // dA > 0, failing ... passing: the first passing column is floor(-K / dA) + 1;
        // dA < 0, passing ... failing: the last passing column is ceil(-K / dA) - 1
        const double cp = fmin(fmax(floor(e.t[i]) + 1.0, (double)lo), (double)(hi + 1));
        const double cn = fmin(fmax(ceil(e.t[i]) - 1.0, (double)(lo - 1)), (double)hi);
        const int x = (int)(pos ? cp : cn);
        // the column one step toward the passing side, and whether it is inside [lo, hi]
        const int x1 = pos ? x - 1 : x + 1;
        const bool in1 = pos ? x - 1 >= lo : x + 1 <= hi;
        const bool in0 = pos ? x <= hi : x >= lo;
        const bool p1 = fma(dA, (double)x1, K) > 0.0, p0 = fma(dA, (double)x, K) > 0.0;
        const int xb = (in1 && p1) ? x1 : (in0 && !p0) ? (pos ? x + 1 : x - 1) : x;
        lo = pos ? xb : lo;
        hi = neg ? xb : hi;
        // dA == 0: the edge passes everywhere or nowhere
        empty = empty || (!pos && !neg && !(K > 0.0)) || lo > hi;
    }
    if (empty) { lo = 1; hi = 0; }
}
