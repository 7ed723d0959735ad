// Grid path for large single solves (1024^2, 4096^2): each phase of a level
// larger than the tail side T is one launch over the whole GPU; levels of
// side <= T run inside one workgroup (the block cycle of block.hpp, with the
// LDS levels and the dense FP32 tail); the CG vector operations are flat
// launches storing the chunk sums of the canonical reductions, each followed
// by a one-workgroup kernel that adds them and advances the device-resident
// CG scalars.
//
// The per-cell arithmetic, the cycle and the reduction order are those of
// the persistent batch kernel, so a solve on either path gives the same
// bits.  The host drives the W/V recursion over the large levels
// (multigrid.hpp:97-127) and the PCG loop (solver.hpp:139-172).
#include "block.hpp"

namespace darcy_hip {

DARCY_INLINE Exec grid_exec() {
    return Exec{static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x), static_cast<int>(gridDim.x * blockDim.x)};
}

// ------------------------------------------------------------------ setup

__global__ __launch_bounds__(kBlock) void kg_fine(const double* a, const Hier* hd, unsigned long long* bad) {
    const Exec e = grid_exec();
    const int n = hd->n[0];
    const long long b = first_invalid(a, n * n, e);
    if (b != INT64_MAX) atomicMin(bad, static_cast<unsigned long long>(b));
    build_fine_faces(a, n, hd->ex[0], hd->ny[0], e);
    build_fine_invdiag(a, n, hd->id[0], e);
}

__global__ __launch_bounds__(kBlock) void kg_galerkin(const Hier* hd, int l) {
    const Exec e = grid_exec();
    galerkin_faces(hd->n[l], hd->ex[l], hd->ny[l], hd->ex[l + 1], hd->ny[l + 1], e);
    galerkin_invdiag(hd->n[l], hd->ex[l], hd->ny[l], hd->id[l + 1], e);
}

__global__ __launch_bounds__(kBlock) void kg_fp32(const Hier* hd, int l) {
    const Exec e = grid_exec();
    to_fp32(hd->n[l], hd->ex[l], hd->ny[l], hd->ex32[l], hd->ny32[l], hd->id32[l], e);
}

// One workgroup: the tail matrix into global memory, in FP32 (the FP32
// cycle's) or FP64 (the FP64 cycle's dense tail).
template <class OT>
__global__ __launch_bounds__(kBlock) void kg_tail_matrix(const Hier* hd, CycleOpts o, OT* tail_out) {
    __shared__ Shared s;
    __shared__ __align__(16) float pool[pool_elems<float>()];
    if (threadIdx.x == 0) {
        s.h = *hd;
        plan_lds<float, LdsSide<float>::value>(s.h, s.t);
    }
    __syncthreads();
    if (s.t.wave < s.h.L) build_tail_matrix(s.h, s.t.wave, o, pool, tail_out);
}

// ----------------------------------------------------------- cycle phases

template <class T>
__global__ __launch_bounds__(kBlock) void kg_presmooth(const Hier* hd, int l, int kind) {
    const GlobalLv<T> L = global_level<T>(*hd, l);
    const Exec e = grid_exec();
    if (kind == 0)
        presmooth_walk(L, e);  // red + black from x = 0
    else
        gs_colour(L, kind == 1 ? 0 : 1, e);  // 1: red, 2: black
}

template <class T>
__global__ __launch_bounds__(kBlock) void kg_restrict(const Hier* hd, int l) {
    const GlobalLv<T> L = global_level<T>(*hd, l);
    restrict_walk(L, to_global(static_cast<T*>(hd->b[l + 1])), grid_exec());
}

template <class T>
__global__ __launch_bounds__(kBlock) void kg_postsmooth(const Hier* hd, int l, int kind) {
    const GlobalLv<T> L = global_level<T>(*hd, l);
    const Exec e = grid_exec();
    if (kind == 0)
        black_with_prolong(L, to_global(static_cast<const T*>(hd->x[l + 1])), e);
    else
        gs_colour(L, kind == 1 ? 0 : 1, e);  // 1: red, 2: black
}

// The levels from lb (8 < side <= 16) down as one dense matrix for the tail
// kernel (column k at out + k * ld), in FP32 (the FP32 cycle) or FP64: each
// wave of the launch builds columns with the FP64 recursive cycle; with the
// W-cycle, the matrix of both visits of lb (build_tail_matrix's pair), which
// the tail then applies once per pair.  Scratch:
// the FP64 operators of levels lb.. (<= 1085 doubles) and x, b of those
// levels per wave (<= 682 doubles each).
constexpr int kBigScratchDoubles = 1088 + kWaves * 682;
template <class OT>
__global__ __launch_bounds__(kBlock) void kg_big_matrix(const Hier* hd, CycleOpts o, int lb, OT* out, int ld) {
    __shared__ Hier h;
    __shared__ __align__(16) double scratch[kBigScratchDoubles];
    if (threadIdx.x == 0) h = *hd;
    __syncthreads();
    build_tail_matrix(h, lb, o, scratch, out, ld, static_cast<int>(blockIdx.x) * kWaves,
                      static_cast<int>(gridDim.x) * kWaves, o.gamma == 2);
}

// Largest side of the levels the tail kernel keeps in LDS.  It runs alone on
// its CU, so it holds far more than a batch workgroup: sides <= 64 in FP32
// (111 KB with the tail matrix beside it), <= 32 in FP64 (57 KB).  The tail
// then reads the grid only once per call -- the operators, b of its first
// level (and x on a repeated visit) -- instead of once per phase.
template <class T>
struct TailLdsSide {
    static constexpr int value = sizeof(T) == 4 ? 64 : 32;
};

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
using gvec = __attribute__((address_space(1))) const u32x4;
using gvec_out = __attribute__((address_space(1))) u32x4;

// count elements global -> LDS (both 16-byte aligned): 16 bytes per thread
// and step, four steps' loads issued before their stores.  Loads past the
// end re-read the last vector (kept in registers: no per-thread arrays) and
// are not stored.
template <class T, class P>
DARCY_INLINE void stage_in(T* lds, P src, int count, Exec e) {
    constexpr int per = 16 / static_cast<int>(sizeof(T));
    const int nv = count / per, st = e.nthr;
    gvec* g = (gvec*)src;  // global address space
    u32x4* d = reinterpret_cast<u32x4*>(lds);
    for (int v0 = e.tid; v0 < nv; v0 += 4 * st) {
        const int v1 = v0 + st, v2 = v0 + 2 * st, v3 = v0 + 3 * st;
        const u32x4 t0 = g[v0], t1 = g[v1 < nv ? v1 : nv - 1], t2 = g[v2 < nv ? v2 : nv - 1],
                    t3 = g[v3 < nv ? v3 : nv - 1];
        d[v0] = t0;
        if (v1 < nv) d[v1] = t1;
        if (v2 < nv) d[v2] = t2;
        if (v3 < nv) d[v3] = t3;
    }
    for (int k = nv * per + e.tid; k < count; k += e.nthr) lds[k] = src[k];
}

// count elements LDS -> global (both 16-byte aligned).
template <class P, class T>
DARCY_INLINE void stage_out(P dst, const T* lds, int count, Exec e) {
    constexpr int per = 16 / static_cast<int>(sizeof(T));
    const int nv = count / per;
    gvec_out* g = (gvec_out*)dst;
    const u32x4* s = reinterpret_cast<const u32x4*>(lds);
    for (int v = e.tid; v < nv; v += e.nthr) g[v] = s[v];
    for (int k = nv * per + e.tid; k < count; k += e.nthr) dst[k] = lds[k];
}

// The cycle from level l0 (side <= T) down, in one workgroup.  When l0 is an
// LDS level its b (and x unless fresh) is staged in and its x staged out.
// Threads of the tail workgroup.  It runs alone on its CU, and 512 (two waves
// per SIMD) hide more of the LDS phases' latency than 256: 4096^2 FP32-cycle
// solve 226 -> 213 ms, FP64 373 -> 358 ms; 1024 measured no better.  Each
// phase computes every cell once, so the count does not change any result.
constexpr int kTailThreads = 512;

// With `big` (levels from big_l down as one dense matrix in global memory,
// columns big_ld apart; with the W-cycle, the matrix of both visits of big_l,
// big_l > l0) the levels below big_l are neither staged nor run.
// mode 1: the cycle from x = 0; 0: a repeated visit from x as stored; 2:
// both, one after the other (the W-cycle's two visits of l0, in one launch:
// the operators are staged once, and x of an LDS level stays there).
template <class T>
__global__ __launch_bounds__(kTailThreads) void kg_tail(const Hier* hd, int l0, int mode, CycleOpts o,
                                                        const float* tail_in, const double* tail64_in, const T* big,
                                                        int big_l, int big_ld) {
    constexpr int NL = TailLdsSide<T>::value;
    __shared__ Shared s;
    __shared__ __align__(16) T pool[LdsPool<T, NL>::kElems];
    __shared__ __align__(16) float tail[sizeof(T) == 4 ? kTailLds : 1];
    __shared__ __align__(16) double tail64[sizeof(T) == 8 ? kTailLds : 1];
    const Exec e{static_cast<int>(threadIdx.x), kTailThreads};
    // The hierarchy (about 2.2 KB) is copied by the whole workgroup, 8 bytes
    // per thread, rather than by one thread: this prologue runs on every one
    // of the tail's calls (64 per CG iteration at 4096^2 with the W-cycle).
    static_assert(sizeof(Hier) % 8 == 0 && sizeof(Hier) / 8 <= kTailThreads, "hierarchy copy");
    if (threadIdx.x < sizeof(Hier) / 8)
        reinterpret_cast<unsigned long long*>(&s.h)[threadIdx.x] =
            reinterpret_cast<const unsigned long long*>(hd)[threadIdx.x];
    __syncthreads();
    if (threadIdx.x == 0) plan_lds<T, NL>(s.h, s.t, l0);
    __syncthreads();
    const Hier& h = s.h;
    const int last = big ? big_l + 1 : h.L;  // levels staged into LDS
    for (int l = s.t.first; l < last; ++l) {
        const GlobalLv<T> g = global_level<T>(h, l);
        const int* f = s.t.off[l];
        const int nf = g.n * (g.n + 1), nc = g.n * g.n;
        stage_in(pool + f[0], g.ex, nf, e);
        stage_in(pool + f[1], g.ny, nf, e);
        stage_in(pool + f[2], g.id, nc, e);
        if (l == l0) {
            stage_in(pool + f[4], g.b, nc, e);
            if (mode == 0) stage_in(pool + f[3], g.x, nc, e);
        }
    }
    if (big) {
        // the dense big matrix replaces the small tail matrix
    } else if constexpr (sizeof(T) == 4) {
        if (s.t.wave < h.L) {
            const int dim = h.n[s.t.wave] * h.n[s.t.wave];
            stage_in(tail, tail_in, dim * dim, e);
        }
    } else {
        if (tail64_in && s.t.wave < h.L) {
            const int dim = h.n[s.t.wave] * h.n[s.t.wave];
            stage_in(tail64, tail64_in, dim * dim, e);
        }
    }
    __syncthreads();
    T* scratch = nullptr;  // the small tail matrix's LDS, free when big is used
    if constexpr (sizeof(T) == 4)
        scratch = tail;
    else
        scratch = tail64;
    const Cycle<T> cyc{h,      s.t,   pool,   tail, o, kStepCellsAlone, sizeof(T) == 8 && tail64_in ? tail64 : nullptr,
                       big,    big ? big_l : -1, big_ld, scratch, big != nullptr && o.gamma == 2};
    cyc.run(l0, mode != 0, e);
    if (mode == 2) cyc.run(l0, false, e);  // the cycle ends with a barrier
    if (s.t.first <= l0) {  // the cycle ends with a barrier
        const GlobalLv<T> g = global_level<T>(h, l0);
        stage_out(g.x, pool + s.t.off[l0][3], g.n * g.n, e);
    }
}

// ------------------------------------------------------------ CG kernels
// Each streaming kernel stores its chunk sums; kg_final, one workgroup
// launched after it on the same stream, adds them in the canonical order and
// advances the device-resident CG scalars.

constexpr int kFinalThreads = 1024;
constexpr int kFinalPer = 16;                           // chunk sums per thread and slice
constexpr int kFinalSlice = kFinalThreads * kFinalPer;  // 16384 doubles: 128 KB of LDS

// The value of final_sum (lane l adds chunk sums l, l+64, ... in order, then
// the butterfly) with the chunk sums staged through LDS by the whole
// workgroup a slice at a time, so that a slice's loads are in flight
// together instead of one wave's sixteen; wave 0 adds.  Valid in wave 0;
// every thread must call it.
__device__ double staged_sum(const double* part, int C, double* buf) {
    const int lane = threadIdx.x & (kWave - 1);
    double s = 0.0;
    for (int base = 0; base < C; base += kFinalSlice) {
        const int cnt = C - base < kFinalSlice ? C - base : kFinalSlice;
        double v[kFinalPer];
#pragma unroll
        for (int u = 0; u < kFinalPer; ++u) {
            const int q = static_cast<int>(threadIdx.x) + u * kFinalThreads;
            v[u] = part[base + (q < cnt ? q : cnt - 1)];
        }
#pragma unroll
        for (int u = 0; u < kFinalPer; ++u) {
            const int q = static_cast<int>(threadIdx.x) + u * kFinalThreads;
            if (q < cnt) buf[q] = v[u];
        }
        __syncthreads();
        if (threadIdx.x < kWave) {  // slices hold whole rounds of 64: lane l owns q = l mod 64
            int q = lane;
            for (; q + 7 * kWave < cnt; q += 8 * kWave) {
                double t[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) t[u] = buf[q + u * kWave];
#pragma unroll
                for (int u = 0; u < 8; ++u) s += t[u];
            }
            for (; q < cnt; q += kWave) s += buf[q];
        }
        __syncthreads();
    }
    return wave_sum(s);
}

// The scalar step after CG kernel KIND: 0 init (<b,b>, <r,r>; solver.hpp:
// 92-108), 1 rz (<r,z> and, flexible, <z,Ap>; sets beta and rz), 2 apply
// (<p,Ap>; alpha or breakdown, solver.hpp:147-155), 3 update (<r,r>;
// residual, history, convergence, solver.hpp:156-166).
template <int KIND>
__global__ __launch_bounds__(kFinalThreads) void kg_final(GridCg g) {
    __shared__ double buf[kFinalSlice];
    const Hier& h = *g.hd;
    const int C = chunks_of(h.n[0] * h.n[0]);
    const double s0 = staged_sum(h.part[0], C, buf);
    const bool two = KIND == 0 || (KIND == 1 && g.mixed);
    const double s1 = two ? staged_sum(h.part[1], C, buf) : 0.0;
    if (threadIdx.x != 0) return;
    GridState& s = *g.st;
    if constexpr (KIND == 0) {
        s.bnorm = sqrt(s0);
        s.res = g.u0 ? sqrt(s1) / s.bnorm : 1.0;
        s.it = 0;
        s.first = 1;
        s.beta = 0.0;
        s.status = kRunning;
        if (s.bnorm == 0.0)
            s.status = kConverged;
        else if (g.u0 && s.res < g.tol)
            s.status = kConverged;
        if (g.history && g.history_cap > 0) g.history[0] = s.res;
    } else if constexpr (KIND == 1) {
        if (!s.first) s.beta = g.mixed ? (-s.alpha * s1) / s.rz : s0 / s.rz;
        s.rz = s0;
    } else if constexpr (KIND == 2) {
        s.pap = s0;
        if (!(fabs(s0) > 0.0) || !isfinite(s0) || s.it + 1 == g.breakdown_at)
            s.status = kBreakdown;
        else
            s.alpha = s.rz / s0;
        s.first = 0;
    } else {
        if (s.status != kRunning) return;  // a breakdown in the apply step ended the solve
        s.res = sqrt(s0) / s.bnorm;
        s.it += 1;
        if (g.history && s.it < g.history_cap) g.history[s.it] = s.res;
        if (s.res < g.tol)
            s.status = kConverged;
        else if (s.it >= g.max_iter)
            s.status = kMaxIter;
    }
}

DARCY_INLINE int grid_wave() { return static_cast<int>(blockIdx.x * kWaves + threadIdx.x / kWave); }
DARCY_INLINE int grid_waves() { return static_cast<int>(gridDim.x * kWaves); }

// r = b - A u0 or b; u = u0 or 0; chunk sums of <b, b> and <r, r>.
__global__ __launch_bounds__(kBlock) void kg_init(GridCg g) {
    const Hier& h = *g.hd;
    const int n = h.n[0], N = n * n;
    const double hh = 1.0 / static_cast<double>(n + 1), h2 = hh * hh;
    const GlobalLv<double> A0 = global_level<double>(h, 0);
    float* b0f = g.mixed ? static_cast<float*>(h.b[0]) : nullptr;
    struct Init {
        double bk, rk, u0;
    };
    chunk_pass<2>(
        N, grid_wave(), grid_waves(), h.part,
        [&](int k) {
            Init v;
            v.bk = (g.rhs ? g.rhs[k] : g.f) * h2;
            v.rk = v.bk;
            v.u0 = 0.0;
            if (g.u0) {
                const int i = k / n, j = k - i * n;
                v.rk = v.bk - apply_at(n, A0.ex, A0.ny, i, j, [&](int t) { return g.u0[t]; });
                v.u0 = g.u0[k];
            }
            return v;
        },
        [&](int k, const Init& v, double* acc) {
            acc[0] += v.bk * v.bk;
            g.u[k] = v.u0;
            h.r[k] = v.rk;
            acc[1] += v.rk * v.rk;
            if (b0f) b0f[k] = static_cast<float>(v.rk);
        });
}

// The streaming CG kernels (kg_rz, kg_update) take two chunks per round so
// each wave has twice the loads in flight (kg_apply, whose stencil loads need
// more waves, takes one); the chunk sums, and so the result, do not depend on
// it.
constexpr int kCgChunks = 2;

// Chunk sums of <r, z> (and <z, A p> in the flexible variant).
// The precision test stays outside the cell loops, as in cg_direction.
__global__ __launch_bounds__(kBlock) void kg_rz(GridCg g) {
    const Hier& h = *g.hd;
    const int N = h.n[0] * h.n[0];
    if (g.mixed) {
        struct Rz {
            double r;
            float z;  // widened after the loads
            double ap;
        };
        const float* z = static_cast<const float*>(h.x[0]);
        chunk_pass<2, kCgChunks>(
            N, grid_wave(), grid_waves(), h.part, [&](int k) { return Rz{h.r[k], z[k], h.ap[k]}; },
            [&](int, const Rz& v, double* acc) {
                const double zk = static_cast<double>(v.z);
                acc[0] += v.r * zk;
                acc[1] += zk * v.ap;
            });
    } else {
        struct Rz {
            double r, z;
        };
        chunk_pass<1, kCgChunks>(
            N, grid_wave(), grid_waves(), h.part, [&](int k) { return Rz{h.r[k], h.z[k]}; },
            [&](int, const Rz& v, double* acc) { acc[0] += v.r * v.z; });
    }
}

// p = z + beta p (first: p = z), ap = A p, chunk sums of <p, ap>.
__global__ __launch_bounds__(kBlock) void kg_apply(GridCg g, const double* pold, double* pnew) {
    const Hier& h = *g.hd;
    const GlobalLv<double> A0 = global_level<double>(h, 0);
    const bool first = g.st->first != 0;
    const double beta = g.st->beta;
    if (g.mixed)
        cg_direction<1, false>(A0.n, A0, static_cast<const float*>(h.x[0]), pold, beta, first, pnew, h.ap, grid_wave(),
                        grid_waves(), h.part);
    else
        cg_direction<1, false>(A0.n, A0, static_cast<const double*>(h.z), pold, beta, first, pnew, h.ap, grid_wave(),
                        grid_waves(), h.part);
}

// z = r * inv_d (Jacobi, solver.hpp:127-132) or z = r (no preconditioner,
// :124-126), as the persistent kernel forms it.
__global__ __launch_bounds__(kBlock) void kg_diag(GridCg g, int jacobi) {
    const Hier& h = *g.hd;
    const int N = h.n[0] * h.n[0];
    const double* id = h.id[0];
    for (int k = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x); k < N;
         k += static_cast<int>(gridDim.x * blockDim.x))
        h.z[k] = jacobi ? h.r[k] * id[k] : h.r[k];
}

// u += alpha p, r -= alpha ap, chunk sums of <r, r>.  Nothing after a
// breakdown (solver.hpp:149-155 returns before the update): the host
// launches this kernel before it reads the state of the apply step.
__global__ __launch_bounds__(kBlock) void kg_update(GridCg g, const double* p) {
    if (g.st->status != kRunning) return;
    const Hier& h = *g.hd;
    const int N = h.n[0] * h.n[0];
    const double alpha = g.st->alpha;
    float* b0f = g.mixed ? static_cast<float*>(h.b[0]) : nullptr;
    struct Upd {
        double u, p, r, ap;
    };
    chunk_pass<1, kCgChunks>(
        N, grid_wave(), grid_waves(), h.part, [&](int k) { return Upd{g.u[k], p[k], h.r[k], h.ap[k]}; },
        [&](int k, const Upd& v, double* acc) {
            g.u[k] = v.u + alpha * v.p;
            const double ri = v.r - alpha * v.ap;
            h.r[k] = ri;
            acc[0] += ri * ri;
            if (b0f) b0f[k] = static_cast<float>(ri);
        });
}

template __global__ void kg_presmooth<double>(const Hier*, int, int);
template __global__ void kg_presmooth<float>(const Hier*, int, int);
template __global__ void kg_restrict<double>(const Hier*, int);
template __global__ void kg_restrict<float>(const Hier*, int);
template __global__ void kg_postsmooth<double>(const Hier*, int, int);
template __global__ void kg_postsmooth<float>(const Hier*, int, int);
template __global__ void kg_tail<double>(const Hier*, int, int, CycleOpts, const float*, const double*, const double*,
                                         int, int);
template __global__ void kg_tail<float>(const Hier*, int, int, CycleOpts, const float*, const double*, const float*, int,
                                        int);

}  // namespace darcy_hip

// ------------------------------------------------------------ launchers
// Plain C entry points for the host code (capi.cpp); `blocks` is chosen by
// the caller.
using namespace darcy_hip;

extern "C" hipError_t darcy_grid_fine(const double* a, const Hier* hd, unsigned long long* bad, int blocks,
                                      hipStream_t s) {
    hipLaunchKernelGGL(kg_fine, dim3(blocks), dim3(kBlock), 0, s, a, hd, bad);
    return hipGetLastError();
}
extern "C" hipError_t darcy_grid_galerkin(const Hier* hd, int l, int blocks, hipStream_t s) {
    hipLaunchKernelGGL(kg_galerkin, dim3(blocks), dim3(kBlock), 0, s, hd, l);
    return hipGetLastError();
}
extern "C" hipError_t darcy_grid_fp32(const Hier* hd, int l, int blocks, hipStream_t s) {
    hipLaunchKernelGGL(kg_fp32, dim3(blocks), dim3(kBlock), 0, s, hd, l);
    return hipGetLastError();
}
extern "C" hipError_t darcy_grid_tail_matrix(const Hier* hd, CycleOpts o, float* out, hipStream_t s) {
    hipLaunchKernelGGL(kg_tail_matrix<float>, dim3(1), dim3(kBlock), 0, s, hd, o, out);
    return hipGetLastError();
}
extern "C" hipError_t darcy_grid_tail_matrix64(const Hier* hd, CycleOpts o, double* out, hipStream_t s) {
    hipLaunchKernelGGL(kg_tail_matrix<double>, dim3(1), dim3(kBlock), 0, s, hd, o, out);
    return hipGetLastError();
}
extern "C" hipError_t darcy_grid_presmooth(int mixed, const Hier* hd, int l, int kind, int blocks, hipStream_t s) {
    if (mixed)
        hipLaunchKernelGGL(kg_presmooth<float>, dim3(blocks), dim3(kBlock), 0, s, hd, l, kind);
    else
        hipLaunchKernelGGL(kg_presmooth<double>, dim3(blocks), dim3(kBlock), 0, s, hd, l, kind);
    return hipGetLastError();
}
extern "C" hipError_t darcy_grid_restrict(int mixed, const Hier* hd, int l, int blocks, hipStream_t s) {
    if (mixed)
        hipLaunchKernelGGL(kg_restrict<float>, dim3(blocks), dim3(kBlock), 0, s, hd, l);
    else
        hipLaunchKernelGGL(kg_restrict<double>, dim3(blocks), dim3(kBlock), 0, s, hd, l);
    return hipGetLastError();
}
extern "C" hipError_t darcy_grid_postsmooth(int mixed, const Hier* hd, int l, int kind, int blocks,
                                            hipStream_t s) {
    if (mixed)
        hipLaunchKernelGGL(kg_postsmooth<float>, dim3(blocks), dim3(kBlock), 0, s, hd, l, kind);
    else
        hipLaunchKernelGGL(kg_postsmooth<double>, dim3(blocks), dim3(kBlock), 0, s, hd, l, kind);
    return hipGetLastError();
}
// big (null: none): the levels from big_l down as one dense matrix, FP32
// for mixed, FP64 otherwise, columns big_ld elements apart.
// mode: kg_tail's (1 fresh, 0 repeated, 2 both visits of the W-cycle).
extern "C" hipError_t darcy_grid_tail(int mixed, const Hier* hd, int l0, int mode, CycleOpts o, const float* tail,
                                      const double* tail64, const void* big, int big_l, int big_ld, hipStream_t s) {
    if (mode < 0 || mode > 2 || (big && big_l <= l0)) return hipErrorInvalidValue;
    if (mixed)
        hipLaunchKernelGGL(kg_tail<float>, dim3(1), dim3(kTailThreads), 0, s, hd, l0, mode, o, tail, nullptr,
                           static_cast<const float*>(big), big_l, big_ld);
    else
        hipLaunchKernelGGL(kg_tail<double>, dim3(1), dim3(kTailThreads), 0, s, hd, l0, mode, o, nullptr, tail64,
                           static_cast<const double*>(big), big_l, big_ld);
    return hipGetLastError();
}

// The big tail matrix of levels lb.. (side n_lb <= 16), columns ld apart.
extern "C" hipError_t darcy_grid_big_matrix(int mixed, const Hier* hd, CycleOpts o, int lb, int dim, void* out, int ld,
                                            hipStream_t s) {
    const int blocks = (dim + kWaves - 1) / kWaves;
    if (mixed)
        hipLaunchKernelGGL(kg_big_matrix<float>, dim3(blocks), dim3(kBlock), 0, s, hd, o, lb, static_cast<float*>(out), ld);
    else
        hipLaunchKernelGGL(kg_big_matrix<double>, dim3(blocks), dim3(kBlock), 0, s, hd, o, lb,
                           static_cast<double*>(out), ld);
    return hipGetLastError();
}
extern "C" hipError_t darcy_grid_cg(int which, const GridCg* g, const double* pold, double* pnew, int blocks,
                                    hipStream_t s) {
    switch (which) {
        case 0:
            hipLaunchKernelGGL(kg_init, dim3(blocks), dim3(kBlock), 0, s, *g);
            hipLaunchKernelGGL(kg_final<0>, dim3(1), dim3(kFinalThreads), 0, s, *g);
            break;
        case 1:
            hipLaunchKernelGGL(kg_rz, dim3(blocks), dim3(kBlock), 0, s, *g);
            hipLaunchKernelGGL(kg_final<1>, dim3(1), dim3(kFinalThreads), 0, s, *g);
            break;
        case 2:
            hipLaunchKernelGGL(kg_apply, dim3(blocks), dim3(kBlock), 0, s, *g, pold, pnew);
            hipLaunchKernelGGL(kg_final<2>, dim3(1), dim3(kFinalThreads), 0, s, *g);
            break;
        case 3:
            hipLaunchKernelGGL(kg_update, dim3(blocks), dim3(kBlock), 0, s, *g, pold);
            hipLaunchKernelGGL(kg_final<3>, dim3(1), dim3(kFinalThreads), 0, s, *g);
            break;
        default:  // 4: Jacobi, 5: no preconditioner
            hipLaunchKernelGGL(kg_diag, dim3(blocks), dim3(kBlock), 0, s, *g, which == 4 ? 1 : 0);
            break;
    }
    return hipGetLastError();
}
extern "C" size_t darcy_grid_state_bytes() { return sizeof(GridState); }

// Launch limits of every kernel the grid path captures into its cycle graph,
// checked on the host before the first capture: block size within the
// kernel's and the device's thread limit, static LDS within the device's
// per-workgroup limit (no kernel here takes dynamic LDS).  Returns the first
// violation as an error code and names the kernel in *what.
extern "C" hipError_t darcy_grid_check_kernels(int device, const char** what) {
    int max_lds = 0, max_threads = 0;
    hipError_t e = hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&max_threads, hipDeviceAttributeMaxThreadsPerBlock, device);
    if (e != hipSuccess) return e;
    struct K {
        const void* fn;
        int block;
        const char* name;
    };
    const K ks[] = {
        {reinterpret_cast<const void*>(&kg_presmooth<double>), kBlock, "kg_presmooth<double>"},
        {reinterpret_cast<const void*>(&kg_presmooth<float>), kBlock, "kg_presmooth<float>"},
        {reinterpret_cast<const void*>(&kg_restrict<double>), kBlock, "kg_restrict<double>"},
        {reinterpret_cast<const void*>(&kg_restrict<float>), kBlock, "kg_restrict<float>"},
        {reinterpret_cast<const void*>(&kg_postsmooth<double>), kBlock, "kg_postsmooth<double>"},
        {reinterpret_cast<const void*>(&kg_postsmooth<float>), kBlock, "kg_postsmooth<float>"},
        {reinterpret_cast<const void*>(&kg_tail<double>), kTailThreads, "kg_tail<double>"},
        {reinterpret_cast<const void*>(&kg_tail<float>), kTailThreads, "kg_tail<float>"},
        {reinterpret_cast<const void*>(&kg_rz), kBlock, "kg_rz"},
        {reinterpret_cast<const void*>(&kg_final<1>), kFinalThreads, "kg_final<1>"},
    };
    for (const K& k : ks) {
        hipFuncAttributes at{};
        e = hipFuncGetAttributes(&at, k.fn);
        if (e != hipSuccess) return e;
        if (k.block > at.maxThreadsPerBlock || k.block > max_threads ||
            static_cast<int>(at.sharedSizeBytes) > max_lds) {
            if (what) *what = k.name;
            return hipErrorInvalidConfiguration;
        }
    }
    return hipSuccess;
}
