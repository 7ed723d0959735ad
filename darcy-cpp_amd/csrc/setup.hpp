// Hierarchy construction on the device: the fine operator from the
// permeability, Galerkin coarsening, FP32 copies for the mixed-precision
// cycle, and input validation.  All of it is FP64 and bit-exact with the
// reference (include/darcy/level.hpp:79-97 and :171-211).
#pragma once

#include "device.hpp"

namespace darcy_hip {

// level.hpp:57: 2xy / (x + y), evaluated as (2x)y / (x + y).
__device__ inline double hmean(double x, double y) { return 2.0 * x * y / (x + y); }

// out[q] = value(q) for q in [0, count): four items per thread and step,
// the loads of all four issued before the first store (`value` sees only
// in-range indices: the items past `count` recompute the last one and are not
// stored).  The values read by the setup loops below are clamped to the grid
// and their boundary cases selected afterwards, so no load is conditional.
// With `out32`, the value rounded to FP32 is stored there too (the operator
// copies of the FP32 cycle, written where the FP64 values are formed).
template <class T, class F>
DARCY_INLINE void map4(int count, T* __restrict__ out, Exec e, F&& value, float* __restrict__ out32 = nullptr) {
    for (int q0 = e.tid; q0 < count; q0 += 4 * e.nthr) {
        T v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int q = q0 + u * e.nthr;
            v[u] = value(q < count ? q : count - 1);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (q0 + u * e.nthr < count) {
                out[q0 + u * e.nthr] = v[u];
                if (out32) out32[q0 + u * e.nthr] = static_cast<float>(v[u]);
            }
    }
}

// Fine faces, level.hpp:83-91 in face form: the interior face between cells
// k-1 and k is the e of cell k-1, hmean(a[k-1], a[k]); boundary faces carry
// the cell's own permeability.
__device__ inline void build_fine_faces(const double* __restrict__ a, int n, double* __restrict__ ex,
                                        double* __restrict__ ny, Exec e, float* ex32 = nullptr,
                                        float* ny32 = nullptr) {
    const int fx = n * (n + 1);
    map4(fx, ex, e, [&](int q) {
        const int i = q / (n + 1), j = q - i * (n + 1);
        const double* row = a + static_cast<size_t>(i) * n;
        const double xl = row[j > 0 ? j - 1 : 0], xr = row[j < n ? j : n - 1];
        return (j == 0) ? xr : (j == n) ? xl : hmean(xl, xr);
    }, ex32);
    map4(fx, ny, e, [&](int q) {
        const int i = q / n, j = q - i * n;
        const double up = a[static_cast<size_t>(i > 0 ? i - 1 : 0) * n + j];
        const double dn = a[static_cast<size_t>(i < n ? i : n - 1) * n + j];
        return (i == 0) ? dn : (i == n) ? up : hmean(up, dn);
    }, ny32);
}

// The FP32 cycle's reciprocal diagonal: 1 / d with d summed in FP32 from the
// rounded faces in the order of level.hpp:92.
DARCY_INLINE float inv_diag32(double e, double w, double n, double s) {
    return 1.0f / (((static_cast<float>(e) + static_cast<float>(w)) + static_cast<float>(n)) + static_cast<float>(s));
}

// An FP64 value together with an FP32 one stored beside it.
struct DF {
    double d;
    float f;
};

// map4 for a DF value: d to out, f to out32 (when given).
template <class F>
DARCY_INLINE void map4_df(int count, double* __restrict__ out, float* __restrict__ out32, Exec e, F&& value) {
    for (int q0 = e.tid; q0 < count; q0 += 4 * e.nthr) {
        DF v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int q = q0 + u * e.nthr;
            v[u] = value(q < count ? q : count - 1);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (q0 + u * e.nthr < count) {
                out[q0 + u * e.nthr] = v[u].d;
                if (out32) out32[q0 + u * e.nthr] = v[u].f;
            }
    }
}

// inv_diag of the fine level (level.hpp:88-94), recomputed from a so that it
// needs no barrier after the faces; with id32, the FP32 cycle's too.
__device__ inline void build_fine_invdiag(const double* __restrict__ a, int n, double* __restrict__ id, Exec e,
                                          float* id32 = nullptr) {
    map4_df(n * n, id, id32, e, [&](int k) {
        const int i = k / n, j = k - i * n;
        const bool hE = j + 1 < n, hW = j > 0, hN = i + 1 < n, hS = i > 0;
        const double c = a[k], aE = a[hE ? k + 1 : k], aW = a[hW ? k - 1 : k], aN = a[hN ? k + n : k],
                     aS = a[hS ? k - n : k];
        const double fe = hE ? hmean(c, aE) : c;
        const double fw = hW ? hmean(c, aW) : c;
        const double fn = hN ? hmean(c, aN) : c;
        const double fs = hS ? hmean(c, aS) : c;
        return DF{1.0 / (((fe + fw) + fn) + fs), inv_diag32(fe, fw, fn, fs)};
    });
}

// Coarse faces (level.hpp:171-211 in face form).  Coarse ex column J sums
// the fine ex column min(2J, n) over the aggregate's rows; coarse ny row I
// sums the fine ny row min(2I, n) over the aggregate's columns.  An
// aggregate has one or two rows (columns); the sum starts from 0.0 and adds
// them in order.
__device__ inline void galerkin_faces(int n, const double* __restrict__ ex, const double* __restrict__ ny,
                                      double* __restrict__ cex, double* __restrict__ cny, Exec e,
                                      float* cex32 = nullptr, float* cny32 = nullptr) {
    const int m = (n + 1) >> 1, fx = m * (m + 1);
    map4(fx, cex, e, [&](int q) {
        const int I = q / (m + 1), J = q - I * (m + 1);
        const int i0 = 2 * I, i1 = (2 * I + 1 < n) ? 2 * I + 1 : 2 * I;
        const int col = (2 * J < n) ? 2 * J : n;
        const double v0 = ex[i0 * (n + 1) + col], v1 = ex[i1 * (n + 1) + col];
        const double s = 0.0 + v0;
        return i1 != i0 ? s + v1 : s;
    }, cex32);
    map4(fx, cny, e, [&](int q) {
        const int I = q / m, J = q - I * m;
        const int j0 = 2 * J, j1 = (2 * J + 1 < n) ? 2 * J + 1 : 2 * J;
        const int row = (2 * I < n) ? 2 * I : n;
        const double v0 = ny[row * n + j0], v1 = ny[row * n + j1];
        const double s = 0.0 + v0;
        return j1 != j0 ? s + v1 : s;
    }, cny32);
}

// Coarse inv_diag from the fine faces directly (no barrier after
// galerkin_faces): d = ((ce + cw) + cn) + cs, level.hpp:205-207.
__device__ inline void galerkin_invdiag(int n, const double* __restrict__ ex, const double* __restrict__ ny,
                                        double* __restrict__ cid, Exec e, float* cid32 = nullptr) {
    const int m = (n + 1) >> 1;
    map4_df(m * m, cid, cid32, e, [&](int q) {
        const int I = q / m, J = q - I * m;
        const int i0 = 2 * I, i1 = (2 * I + 1 < n) ? 2 * I + 1 : 2 * I;
        const int j0 = 2 * J, j1 = (2 * J + 1 < n) ? 2 * J + 1 : 2 * J;
        const bool two_i = i1 != i0, two_j = j1 != j0;
        const double e0 = ex[i0 * (n + 1) + j1 + 1], e1 = ex[i1 * (n + 1) + j1 + 1];
        const double w0 = ex[i0 * (n + 1) + j0], w1 = ex[i1 * (n + 1) + j0];
        const double n0 = ny[(i1 + 1) * n + j0], n1 = ny[(i1 + 1) * n + j1];
        const double s0 = ny[i0 * n + j0], s1 = ny[i0 * n + j1];
        double ce = 0.0 + e0, cw = 0.0 + w0, cn = 0.0 + n0, cs = 0.0 + s0;
        ce = two_i ? ce + e1 : ce;
        cw = two_i ? cw + w1 : cw;
        cn = two_j ? cn + n1 : cn;
        cs = two_j ? cs + s1 : cs;
        return DF{1.0 / (((ce + cw) + cn) + cs), inv_diag32(ce, cw, cn, cs)};
    });
}

// FP32 copies of one level's operator for the mixed-precision cycle: the
// rounded faces and inv_diag32 of them.
__device__ inline void to_fp32(int n, const double* __restrict__ ex, const double* __restrict__ ny,
                               float* __restrict__ ex32, float* __restrict__ ny32, float* __restrict__ id32, Exec e) {
    const int fx = n * (n + 1), N = n * n;
    map4(fx, ex32, e, [&](int q) { return static_cast<float>(ex[q]); });
    map4(fx, ny32, e, [&](int q) { return static_cast<float>(ny[q]); });
    map4(N, id32, e, [&](int k) {
        const int i = k / n, j = k - i * n, f = i * (n + 1) + j;
        return inv_diag32(ex[f + 1], ex[f], ny[(i + 1) * n + j], ny[i * n + j]);
    });
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
