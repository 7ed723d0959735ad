// Gaussian random permeability fields on the GPU: the input generator of
// every workload (reference python/darcy/fields.py:31-83) for batches of
// media, so that `darcy dataset` builds its media where it solves them.
//
// For normals xi (drawn on the host by numpy's generator, which fixes the
// media draw for draw) and the spectral amplitude s[k1][k2] =
// tau^(alpha-1) (pi^2 (k1^2 + k2^2) + tau^2)^(-alpha/2), s[0][0] = 0, the
// field is the orthonormal 2-D DCT-III of s * xi:
//   g = C^T (s * xi) C,   C[k][m] = w_k cos(pi k (2m + 1) / (2n)),
//   w_0 = sqrt(1/n), w_k = sqrt(2/n),
// and the medium is a = (g >= 0 ? hi : lo).  Both products are FP64
// GEMM-shaped (2 n^3 flops per sample and product; the DCT over n points is
// n cosines per output): 64 x 64 output tiles, 16-deep K slices in LDS, a
// 4 x 4 register tile per thread, one z-slice of the grid per sample.
#include <hip/hip_runtime.h>

#include <stdint.h>

namespace darcy_fields {

constexpr int kTile = 64;
constexpr int kK = 16;
constexpr int kThreads = 256;

// C[k][m] with the angle reduced exactly in integers first: k (2m + 1) mod 4n.
__global__ __launch_bounds__(kThreads) void k_dct_matrix(int n, double* c) {
    const int64_t total = static_cast<int64_t>(n) * n;
    const double s0 = sqrt(1.0 / n), s1 = sqrt(2.0 / n);
    for (int64_t q = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; q < total;
         q += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t k = q / n, m = q - k * n;
        const int64_t r = (k * (2 * m + 1)) % (4 * static_cast<int64_t>(n));
        const double angle = 3.14159265358979323846 * static_cast<double>(r) / (2.0 * n);
        c[q] = (k == 0 ? s0 : s1) * cos(angle);
    }
}

// The spectral amplitude, formed as the reference forms it (the same
// operations in the same order, so the same doubles).
__global__ __launch_bounds__(kThreads) void k_amplitude(int n, double alpha, double tau, double* s) {
    const int64_t total = static_cast<int64_t>(n) * n;
    const double pi2 = 3.14159265358979323846 * 3.14159265358979323846;
    const double lead = pow(tau, alpha - 1.0), expo = -alpha / 2.0, tau2 = tau * tau;
    for (int64_t q = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; q < total;
         q += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t k1 = q / n, k2 = q - k1 * n;
        const double coef = pi2 * static_cast<double>(k1 * k1 + k2 * k2) + tau2;
        s[q] = q == 0 ? 0.0 : lead * pow(coef, expo);
    }
}

// out_b = L_b R_b for n x n row-major operands; L and R are read through the
// loaders (row, col) -> value, 0 outside the matrix.  Each sum runs over k in
// increasing order.  STAGE 0: L = s * xi_b, R = C, out = T_b.  STAGE 1:
// L = C^T, R = T_b, out = g_b and, if a is given, the thresholded medium.
template <int STAGE>
__global__ __launch_bounds__(kThreads) void k_dct_gemm(int n, const double* __restrict__ amp,
                                                       const double* __restrict__ xi, const double* __restrict__ c,
                                                       double* __restrict__ t, double* __restrict__ g,
                                                       double* __restrict__ a, double hi, double lo) {
    __shared__ double ls[kK][kTile + 1];
    __shared__ double rs[kK][kTile + 1];
    const int64_t plane = static_cast<int64_t>(n) * n, b = blockIdx.z;
    const int row0 = blockIdx.y * kTile, col0 = blockIdx.x * kTile;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const double* xs = xi + b * plane;
    const double* ts = t + b * plane;
    double acc[4][4] = {};
    for (int k0 = 0; k0 < n; k0 += kK) {
        // 64 x 16 of L and 16 x 64 of R: 4 values of each per thread.
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int q = threadIdx.x + u * kThreads;  // 0..1023
            const int lr = q >> 4, lk = q & 15;        // L tile: row lr, column lk
            const int i = row0 + lr, k = k0 + lk;
            double lv = 0.0;
            if (i < n && k < n) {
                if constexpr (STAGE == 0)
                    lv = amp[static_cast<int64_t>(i) * n + k] * xs[static_cast<int64_t>(i) * n + k];
                else
                    lv = c[static_cast<int64_t>(k) * n + i];  // C^T
            }
            ls[lk][lr] = lv;
            const int rk = q >> 6, rc = q & 63;  // R tile: row rk, column rc
            const int kr = k0 + rk, m = col0 + rc;
            double rv = 0.0;
            if (kr < n && m < n) rv = STAGE == 0 ? c[static_cast<int64_t>(kr) * n + m] : ts[static_cast<int64_t>(kr) * n + m];
            rs[rk][rc] = rv;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kK; ++kk) {
            double lv[4], rv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                lv[u] = ls[kk][ty * 4 + u];
                rv[u] = rs[kk][tx * 4 + u];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[u][v] += lv[u] * rv[v];
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int i = row0 + ty * 4 + u;
        if (i >= n) continue;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int m = col0 + tx * 4 + v;
            if (m >= n) continue;
            const int64_t q = b * plane + static_cast<int64_t>(i) * n + m;
            if constexpr (STAGE == 0) {
                t[q] = acc[u][v];
            } else {
                if (g) g[q] = acc[u][v];
                if (a) a[q] = acc[u][v] >= 0.0 ? hi : lo;
            }
        }
    }
}

}  // namespace darcy_fields

using namespace darcy_fields;

// Fields for `count` samples: xi, t, g, a hold count * n * n doubles (device
// pointers; g or a may be null); cmat and amp n * n doubles of scratch.
extern "C" hipError_t darcy_hip_launch_grf(int n, int count, double alpha, double tau, const double* xi,
                                           double* cmat, double* amp, double* t, double* g, double* a, double hi,
                                           double lo, hipStream_t s) {
    const int64_t cells = static_cast<int64_t>(n) * n;
    const int fill = static_cast<int>((cells + kThreads - 1) / kThreads < 65536 ? (cells + kThreads - 1) / kThreads
                                                                                  : 65536);
    hipLaunchKernelGGL(k_dct_matrix, dim3(fill), dim3(kThreads), 0, s, n, cmat);
    hipLaunchKernelGGL(k_amplitude, dim3(fill), dim3(kThreads), 0, s, n, alpha, tau, amp);
    const int tiles = (n + kTile - 1) / kTile;
    const dim3 grid(tiles, tiles, count);
    hipLaunchKernelGGL(k_dct_gemm<0>, grid, dim3(kThreads), 0, s, n, amp, xi, cmat, t, nullptr, nullptr, 0.0, 0.0);
    hipLaunchKernelGGL(k_dct_gemm<1>, grid, dim3(kThreads), 0, s, n, amp, xi, cmat, t, g, a, hi, lo);
    return hipGetLastError();
}
