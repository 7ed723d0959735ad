// Hierarchy construction on the device: the fine operator from the
// permeability, Galerkin coarsening, FP32 copies for the mixed-precision
// cycle, and input validation.  All of it is FP64 and bit-exact with the
// reference (include/darcy/level.hpp:79-97 and :171-211).
#pragma once

#include "device.hpp"

namespace darcy_hip {

// level.hpp:57: 2xy / (x + y), evaluated as (2x)y / (x + y).
__device__ inline double hmean(double x, double y) { return 2.0 * x * y / (x + y); }

// Fine faces, level.hpp:83-91 in face form: the interior face between cells
// k-1 and k is the e of cell k-1, hmean(a[k-1], a[k]); boundary faces carry
// the cell's own permeability.
__device__ inline void build_fine_faces(const double* __restrict__ a, int n, double* __restrict__ ex,
                                        double* __restrict__ ny, Exec e) {
    const int fx = n * (n + 1);
    for (int q = e.tid; q < fx; q += e.nthr) {
        const int i = q / (n + 1), j = q - i * (n + 1);
        const double* row = a + static_cast<size_t>(i) * n;
        ex[q] = (j == 0) ? row[0] : (j == n) ? row[n - 1] : hmean(row[j - 1], row[j]);
    }
    for (int q = e.tid; q < fx; q += e.nthr) {
        const int i = q / n, j = q - i * n;
        ny[q] = (i == 0) ? a[j]
              : (i == n) ? a[static_cast<size_t>(n - 1) * n + j]
                         : hmean(a[static_cast<size_t>(i - 1) * n + j], a[static_cast<size_t>(i) * n + j]);
    }
}

// inv_diag of the fine level (level.hpp:88-94), recomputed from a so that it
// needs no barrier after the faces.
__device__ inline void build_fine_invdiag(const double* __restrict__ a, int n, double* __restrict__ id, Exec e) {
    const int N = n * n;
    for (int k = e.tid; k < N; k += e.nthr) {
        const int i = k / n, j = k - i * n;
        const double c = a[k];
        const double fe = (j + 1 < n) ? hmean(c, a[k + 1]) : c;
        const double fw = (j > 0) ? hmean(c, a[k - 1]) : c;
        const double fn = (i + 1 < n) ? hmean(c, a[k + n]) : c;
        const double fs = (i > 0) ? hmean(c, a[k - n]) : c;
        id[k] = 1.0 / (((fe + fw) + fn) + fs);
    }
}

// Coarse faces (level.hpp:171-211 in face form).  Coarse ex column J sums
// the fine ex column min(2J, n) over the aggregate's rows; coarse ny row I
// sums the fine ny row min(2I, n) over the aggregate's columns.
__device__ inline void galerkin_faces(int n, const double* __restrict__ ex, const double* __restrict__ ny,
                                      double* __restrict__ cex, double* __restrict__ cny, Exec e) {
    const int m = (n + 1) >> 1, fx = m * (m + 1);
    for (int q = e.tid; q < fx; q += e.nthr) {
        const int I = q / (m + 1), J = q - I * (m + 1);
        const int i0 = 2 * I, i1 = (2 * I + 1 < n) ? 2 * I + 1 : 2 * I;
        const int col = (2 * J < n) ? 2 * J : n;
        double s = 0.0;
        for (int i = i0; i <= i1; ++i) s += ex[i * (n + 1) + col];
        cex[q] = s;
    }
    for (int q = e.tid; q < fx; q += e.nthr) {
        const int I = q / m, J = q - I * m;
        const int j0 = 2 * J, j1 = (2 * J + 1 < n) ? 2 * J + 1 : 2 * J;
        const int row = (2 * I < n) ? 2 * I : n;
        double s = 0.0;
        for (int j = j0; j <= j1; ++j) s += ny[row * n + j];
        cny[q] = s;
    }
}

// Coarse inv_diag from the fine faces directly (no barrier after
// galerkin_faces): d = ((ce + cw) + cn) + cs, level.hpp:205-207.
__device__ inline void galerkin_invdiag(int n, const double* __restrict__ ex, const double* __restrict__ ny,
                                        double* __restrict__ cid, Exec e) {
    const int m = (n + 1) >> 1, M = m * m;
    for (int q = e.tid; q < M; q += e.nthr) {
        const int I = q / m, J = q - I * m;
        const int i0 = 2 * I, i1 = (2 * I + 1 < n) ? 2 * I + 1 : 2 * I;
        const int j0 = 2 * J, j1 = (2 * J + 1 < n) ? 2 * J + 1 : 2 * J;
        double ce = 0.0, cw = 0.0, cn = 0.0, cs = 0.0;
        for (int i = i0; i <= i1; ++i) {
            ce += ex[i * (n + 1) + j1 + 1];
            cw += ex[i * (n + 1) + j0];
        }
        for (int j = j0; j <= j1; ++j) {
            cn += ny[(i1 + 1) * n + j];
            cs += ny[i0 * n + j];
        }
        cid[q] = 1.0 / (((ce + cw) + cn) + cs);
    }
}

// FP32 copies of one level's operator for the mixed-precision cycle.  The
// reciprocal diagonal is rounded from its FP64 value; the cycle's diagonal
// is summed in FP32 from the rounded faces where a kernel needs it.
__device__ inline void to_fp32(int n, const double* __restrict__ ex, const double* __restrict__ ny,
                               const double* __restrict__ id, float* __restrict__ ex32, float* __restrict__ ny32,
                               float* __restrict__ id32, Exec e) {
    const int fx = n * (n + 1), N = n * n;
    for (int q = e.tid; q < fx; q += e.nthr) {
        ex32[q] = static_cast<float>(ex[q]);
        ny32[q] = static_cast<float>(ny[q]);
    }
    for (int q = e.tid; q < N; q += e.nthr) id32[q] = static_cast<float>(id[q]);
}

// Smallest flat index with !(a > 0) || !isfinite(a) among this thread's
// cells (solver.hpp:68-77), or INT64_MAX.
__device__ inline long long first_invalid(const double* __restrict__ a, int N, Exec e) {
    long long bad = INT64_MAX;
    for (int k = e.tid; k < N; k += e.nthr) {
        const double v = a[k];
        if (!(v > 0.0) || !isfinite(v)) {
            bad = k;
            break;
        }
    }
    return bad;
}

}  // namespace darcy_hip
