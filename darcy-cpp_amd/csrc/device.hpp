// Device building blocks of the solve path for gfx950 (CDNA4, wave64).
//
// Everything here is written once and used both by the persistent batch
// kernel (one workgroup owns a whole problem) and by the grid kernels of the
// large single solves (many workgroups share one level), so the per-cell
// arithmetic -- and therefore every result -- is identical on both paths.
//
// Per-cell floating-point operations follow the reference's operation order
// exactly (include/darcy/level.hpp), so that with -ffp-contract=off the FP64
// cycle reproduces the reference bit for bit:
//   apply/residual  v = d*x_k - e*x_E - w*x_W - n*x_N - s*x_S   (level.hpp:111-116, 129-134)
//   GS update       x_k = (((b_k + e*x_E) + w*x_W) + n*x_N + s*x_S) * inv_d   (level.hpp:153-158)
//   restriction     sum over (i0,j0),(i0,j1),(i1,j0),(i1,j1)   (level.hpp:221-224)
//
// Reductions (the CG dot products) use one canonical order that does not
// depend on launch shape: the flat vector is cut into chunks of 256 cells;
// lane l of the wave that owns a chunk adds cells l, l+64, l+128, l+192 in
// turn, the 64 lane sums are combined by a fixed xor butterfly, and the chunk
// sums are combined the same way by one wave.  Any number of workgroups or
// waves yields the same bits.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "layout.hpp"

namespace darcy_hip {

constexpr int kWave = 64;

// Operator view of one level in the cycle's arithmetic type T.
template <class T>
struct Lv {
    int n;
    const T* __restrict__ ex;
    const T* __restrict__ ny;
    const T* __restrict__ id;
};

template <class T>
__device__ inline Lv<T> level_view(const Hier& h, int l);
template <>
__device__ inline Lv<double> level_view<double>(const Hier& h, int l) {
    return {h.n[l], h.ex[l], h.ny[l], h.id[l]};
}
template <>
__device__ inline Lv<float> level_view<float>(const Hier& h, int l) {
    return {h.n[l], h.ex32[l], h.ny32[l], h.id32[l]};
}

// ---------------------------------------------------------------- per cell

// Diagonal ((e + w) + n) + s, exactly as level.hpp:92 / :205 build it.
template <class T>
__device__ inline T diag_at(const Lv<T>& L, int i, int j) {
    const int n = L.n, f = i * (n + 1) + j;
    return ((L.ex[f + 1] + L.ex[f]) + L.ny[(i + 1) * n + j]) + L.ny[i * n + j];
}

// (A x)_k with off-grid neighbours dropped (level.hpp:105-119).  `X` maps a
// flat index to the value of x there, so callers can form x on the fly.
template <class T, class X>
__device__ inline T apply_at(const Lv<T>& L, int i, int j, X&& x) {
    const int n = L.n, k = i * n + j, f = i * (n + 1) + j;
    const T e = L.ex[f + 1], w = L.ex[f], nn = L.ny[(i + 1) * n + j], s = L.ny[i * n + j];
    const T d = ((e + w) + nn) + s;
    T v = d * x(k);
    if (j + 1 < n) v -= e * x(k + 1);
    if (j > 0) v -= w * x(k - 1);
    if (i + 1 < n) v -= nn * x(k + n);
    if (i > 0) v -= s * x(k - n);
    return v;
}

// Gauss-Seidel value of cell (i, j) from its neighbours (level.hpp:153-158).
template <class T, class X>
__device__ inline T gs_at(const Lv<T>& L, int i, int j, T bk, X&& x) {
    const int n = L.n, k = i * n + j, f = i * (n + 1) + j;
    T v = bk;
    if (j + 1 < n) v += L.ex[f + 1] * x(k + 1);
    if (j > 0) v += L.ex[f] * x(k - 1);
    if (i + 1 < n) v += L.ny[(i + 1) * n + j] * x(k + n);
    if (i > 0) v += L.ny[i * n + j] * x(k - n);
    return v * L.id[k];
}

// ------------------------------------------------------------------ phases
// A phase is executed by `nthr` cooperating threads; `tid` is this thread's
// rank among them.  Callers separate dependent phases with a barrier.

struct Exec {
    int tid;
    int nthr;
};

// One colour of red-black GS (level.hpp:145-161).  Cells of row i with
// (i + j) % 2 == colour are enumerated as j = 2q + ((i + colour) & 1).
template <class T>
__device__ inline void gs_colour(const Lv<T>& L, T* __restrict__ x, const T* __restrict__ b, int colour,
                                 Exec e) {
    const int n = L.n, h = (n + 1) >> 1, total = n * h;
    for (int q = e.tid; q < total; q += e.nthr) {
        const int i = q / h;
        const int j = 2 * (q - i * h) + ((i + colour) & 1);
        if (j >= n) continue;
        const int k = i * n + j;
        x[k] = gs_at(L, i, j, b[k], [&](int m) { return x[m]; });
    }
}

// The first red sweep on a level whose x is zero: every neighbour term is
// e * 0, so the update is b * inv_d (the reference's value, up to the sign of
// a zero).  Black cells are left untouched: they are still zero and the black
// sweep that follows reads only red cells.
template <class T>
__device__ inline void gs_red_from_zero(const Lv<T>& L, T* __restrict__ x, const T* __restrict__ b, Exec e) {
    const int n = L.n, h = (n + 1) >> 1, total = n * h;
    for (int q = e.tid; q < total; q += e.nthr) {
        const int i = q / h;
        const int j = 2 * (q - i * h) + (i & 1);
        if (j >= n) continue;
        const int k = i * n + j;
        x[k] = b[k] * L.id[k];
    }
}

// Black sweep right after gs_red_from_zero: black cells were zero and are
// never read, so only red neighbours contribute -- the general update.
template <class T>
__device__ inline void zero_colour(int n, T* __restrict__ x, int colour, Exec e) {
    const int h = (n + 1) >> 1, total = n * h;
    for (int q = e.tid; q < total; q += e.nthr) {
        const int i = q / h;
        const int j = 2 * (q - i * h) + ((i + colour) & 1);
        if (j < n) x[i * n + j] = T(0);
    }
}

// b_c = P^T (b - A x) over each <= 2x2 aggregate (level.hpp:122-137 then
// :214-227, fused: the fine residual is never stored).
template <class T>
__device__ inline void residual_restrict(const Lv<T>& L, const T* __restrict__ x, const T* __restrict__ b,
                                         T* __restrict__ bc, Exec e) {
    const int n = L.n, m = (n + 1) >> 1, total = m * m;
    for (int q = e.tid; q < total; q += e.nthr) {
        const int I = q / m, J = q - I * m;
        const int i0 = 2 * I, i1 = (2 * I + 1 < n) ? 2 * I + 1 : 2 * I;
        const int j0 = 2 * J, j1 = (2 * J + 1 < n) ? 2 * J + 1 : 2 * J;
        T s = T(0);
        for (int i = i0; i <= i1; ++i)
            for (int j = j0; j <= j1; ++j) {
                const int k = i * n + j;
                s += b[k] - apply_at(L, i, j, [&](int t) { return x[t]; });
            }
        bc[q] = s;
    }
}

// x += P x_c (level.hpp:230-237).
template <class T>
__device__ inline void prolong_add(int n, T* __restrict__ x, const T* __restrict__ xc, Exec e) {
    const int m = (n + 1) >> 1, total = n * n;
    for (int k = e.tid; k < total; k += e.nthr) {
        const int i = k / n, j = k - i * n;
        x[k] += xc[(i >> 1) * m + (j >> 1)];
    }
}

// ------------------------------------------------------ workgroup W/V cycle

__device__ inline void block_sync() { __syncthreads(); }

// Coarsest solve (multigrid.hpp:129-144): exact for 1x1, otherwise 8 + 8
// symmetric sweeps (reached only when the 32-level cap stops coarsening).
template <class T>
__device__ inline void coarsest_solve(const Lv<T>& L, T* x, const T* b, bool fresh, Exec e) {
    if (L.n == 1) {
        if (e.tid == 0) x[0] = b[0] * L.id[0];
        block_sync();
        return;
    }
    if (fresh) {
        for (int k = e.tid; k < L.n * L.n; k += e.nthr) x[k] = T(0);
        block_sync();
    }
    for (int s = 0; s < 8; ++s) {
        gs_colour(L, x, b, 0, e);
        block_sync();
        gs_colour(L, x, b, 1, e);
        block_sync();
    }
    for (int s = 0; s < 8; ++s) {
        gs_colour(L, x, b, 1, e);
        block_sync();
        gs_colour(L, x, b, 0, e);
        block_sync();
    }
}

template <class T>
__device__ inline void pre_smooth(const Lv<T>& L, T* x, const T* b, int sweeps, bool fresh, Exec e) {
    for (int s = 0; s < sweeps; ++s) {
        if (s == 0 && fresh)
            gs_red_from_zero(L, x, b, e);
        else
            gs_colour(L, x, b, 0, e);
        block_sync();
        gs_colour(L, x, b, 1, e);
        block_sync();
    }
}

template <class T>
__device__ inline void post_smooth(const Lv<T>& L, T* x, const T* b, int sweeps, Exec e) {
    for (int s = 0; s < sweeps; ++s) {
        gs_colour(L, x, b, 1, e);
        block_sync();
        gs_colour(L, x, b, 0, e);
        block_sync();
    }
}

struct CycleOpts {
    int pre;
    int post;
    int gamma;  // 1 = V-cycle, 2 = W-cycle
};

// The recursive cycle of multigrid.hpp:97-127, run iteratively by one
// workgroup from level l0 down to the coarsest level L-1.  `fresh0` says
// whether x_{l0} is zero on entry.  Bit k of `done` counts the completed
// sub-cycles of level k (gamma <= 2).
template <class T>
__device__ void block_cycle(const Hier& h, int l0, bool fresh0, const CycleOpts& o, Exec e) {
    const int L = h.L;
    uint32_t done = 0;
    int l = l0;
    bool fresh = fresh0;
    for (;;) {
        while (l + 1 < L) {
            const Lv<T> A = level_view<T>(h, l);
            T* x = static_cast<T*>(h.x[l]);
            const T* b = static_cast<const T*>(h.b[l]);
            pre_smooth(A, x, b, o.pre, fresh, e);
            residual_restrict(A, x, b, static_cast<T*>(h.b[l + 1]), e);
            block_sync();
            done &= ~(1u << l);
            ++l;
            fresh = true;
        }
        coarsest_solve(level_view<T>(h, l), static_cast<T*>(h.x[l]), static_cast<const T*>(h.b[l]), fresh, e);
        for (;;) {
            if (l == l0) return;
            --l;
            const bool first_done = (done >> l) & 1u;
            if (!first_done && o.gamma == 2) {
                done |= 1u << l;
                ++l;
                fresh = false;
                break;
            }
            const Lv<T> A = level_view<T>(h, l);
            T* x = static_cast<T*>(h.x[l]);
            prolong_add(A.n, x, static_cast<const T*>(h.x[l + 1]), e);
            block_sync();
            post_smooth(A, x, static_cast<const T*>(h.b[l]), o.post, e);
        }
    }
}

// ------------------------------------------------------------ reductions

__device__ inline double wave_sum(double s) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, kWave);
    return __shfl(s, 0, kWave);
}

// Runs `cell(k, acc)` over the chunks owned by wave `w` of `nw` waves and
// stores each chunk's canonical sums in part[d][chunk].
template <int ND, class F>
__device__ inline void chunk_sums(int N, int w, int nw, double* const* part, F&& cell) {
    const int lane = threadIdx.x & (kWave - 1);
    const int C = chunks_of(N);
    for (int c = w; c < C; c += nw) {
        double acc[ND];
#pragma unroll
        for (int d = 0; d < ND; ++d) acc[d] = 0.0;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const int k = c * kChunk + m * kWave + lane;
            if (k < N) cell(k, acc);
        }
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            const double s = wave_sum(acc[d]);
            if (lane == 0) part[d][c] = s;
        }
    }
}

// Combine C chunk sums (one full wave must call this).
__device__ inline double final_sum(const double* part, int C) {
    const int lane = threadIdx.x & (kWave - 1);
    double s = 0.0;
    for (int c = lane; c < C; c += kWave) s += part[c];
    return wave_sum(s);
}

}  // namespace darcy_hip
