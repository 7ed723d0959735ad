// Persistent solve kernel: each workgroup takes problems from a device work
// queue and runs the whole solve of one problem -- validation, hierarchy
// construction, and the preconditioned CG loop with its multigrid cycles --
// without returning to the host.  Problems of different sizes share one
// launch (uneven batches); a single problem is a batch of one.
//
// Mirrors include/darcy/solver.hpp:83-174 (the PCG driver) per problem, with
// the batch semantics of src/bindings.cpp:140-197 (independent samples,
// per-sample diagnostics, invalid samples reported rather than solved).
// Two additions have no reference counterpart: a warm start u0, and the FP32
// cycle inside an FP64 flexible CG (beta = -alpha <z_new, A p> / <z, r>,
// Notay's FCG(1), since r_new - r_old = -alpha A p).
#include <mutex>

#include "block.hpp"

namespace darcy_hip {

// B threads per workgroup: kBlock, or kWideBlock for the large problems of a
// batch (one workgroup per CU, all its wave slots on one problem).  Every
// phase takes the thread count from Exec and the reductions keep their
// canonical order, so both widths give the same bits.
template <class T, int B>
DARCY_INLINE void run_problem(Shared& s, T* pool, float* tail, const Problem& pb, ProblemStats& st,
                              const SolveParams& P, char* slot) {
    constexpr int kWavesB = B / kWave;  // this workgroup's waves
    const Exec e{static_cast<int>(threadIdx.x), B};
    const int wave = threadIdx.x / kWave;
    const int n = pb.n, N = n * n, C = chunks_of(N);
    constexpr bool mixed = sizeof(T) == 4;

    // Validation (solver.hpp:68-77): the first offending flat index.
    if (threadIdx.x == 0) s.bad = ~0ull;
    __syncthreads();
    const long long bad = first_invalid(pb.a, N, e);
    if (bad != INT64_MAX) atomicMin(&s.bad, static_cast<unsigned long long>(bad));
    __syncthreads();
    if (s.bad != ~0ull) {
        if (threadIdx.x == 0) st = ProblemStats{0, kInvalid, 1, 0, 0.0, static_cast<long long>(s.bad)};
        return;
    }

    const bool mg = P.mode == kModeApplyPreconditioner || (P.mode == kModeSolve && P.precond == kPcMultigrid);
    if (threadIdx.x == 0) {
        plan_slot(slot, n, mixed, &s.h, P.min_coarse, P.max_levels);
        plan_lds<T, LdsSide<T, B>::value>(s.h, s.t);
    }
    __syncthreads();
    const Hier& h = s.h;

    // Hierarchy (level.hpp:79-97, multigrid.hpp:55-71, level.hpp:171-211),
    // with the FP32 cycle's operator copies written where each level's FP64
    // values are formed.
    const bool c32 = mixed && mg;
    build_fine_faces(pb.a, n, h.ex[0], h.ny[0], e, c32 ? h.ex32[0] : nullptr, c32 ? h.ny32[0] : nullptr);
    build_fine_invdiag(pb.a, n, h.id[0], e, c32 ? h.id32[0] : nullptr);
    __syncthreads();
    if (mg) {
        for (int l = 0; l + 1 < h.L; ++l) {
            galerkin_faces(h.n[l], h.ex[l], h.ny[l], h.ex[l + 1], h.ny[l + 1], e, c32 ? h.ex32[l + 1] : nullptr,
                           c32 ? h.ny32[l + 1] : nullptr);
            galerkin_invdiag(h.n[l], h.ex[l], h.ny[l], h.id[l + 1], e, c32 ? h.id32[l + 1] : nullptr);
            __syncthreads();
        }
        if constexpr (mixed) {
            if (s.t.wave < h.L) build_tail_matrix(h, s.t.wave, CycleOpts{P.pre, P.post, P.gamma}, pool, tail);
        }
        // Coarse operators into LDS for the whole solve.
        for (int l = s.t.first; l < h.L; ++l) {
            const GlobalLv<T> g = global_level<T>(h, l);
            const int* f = s.t.off[l];
            const int nf = g.n * (g.n + 1), nc = g.n * g.n;
            for (int q = e.tid; q < nf; q += e.nthr) {
                pool[f[0] + q] = g.ex[q];
                pool[f[1] + q] = g.ny[q];
            }
            for (int q = e.tid; q < nc; q += e.nthr) pool[f[2] + q] = g.id[q];
        }
        __syncthreads();
    }
    const Cycle<T> cyc{h, s.t, pool, tail, CycleOpts{P.pre, P.post, P.gamma},
                       P.step_cells > 0 ? P.step_cells : kStepCellsShared};
    const GlobalLv<double> A0 = global_level<double>(h, 0);
    float* const b0f = mixed ? static_cast<float*>(h.b[0]) : nullptr;
    const float* const x0f = mixed ? static_cast<const float*>(h.x[0]) : nullptr;
    auto zval = [&](int k) -> double { return mixed ? static_cast<double>(x0f[k]) : h.z[k]; };

    if (P.mode == kModeApplyOperator) {  // bindings.cpp:199-216
        for (int k = e.tid; k < N; k += e.nthr) {
            const int i = k / n, j = k - i * n;
            pb.out[k] = apply_at(n, A0.ex, A0.ny, i, j, [&](int t) { return pb.in[t]; });
        }
        if (threadIdx.x == 0) st = ProblemStats{0, kConverged, 1, 0, 0.0, -1};
        return;
    }
    if (P.mode == kModeApplyPreconditioner) {  // bindings.cpp:218-245
        for (int k = e.tid; k < N; k += e.nthr) {
            h.r[k] = pb.in[k];
            if (mixed) b0f[k] = static_cast<float>(pb.in[k]);
        }
        __syncthreads();
        precondition(cyc, h, P, N, e);
        for (int k = e.tid; k < N; k += e.nthr) pb.out[k] = zval(k);
        if (threadIdx.x == 0) st = ProblemStats{0, kConverged, h.L, 0, 0.0, -1};
        return;
    }

    // b = f h^2 (solver.hpp:92-102); r = b - A u0 (warm start) or b.
    const double hh = 1.0 / static_cast<double>(n + 1);
    const double h2 = hh * hh;
    double* const u = pb.out;
    struct Init {
        double bk, rk, u0;
    };
    // The warm-start test stays outside the cell loop (a conditional stencil
    // would branch around its loads).
    auto init = [&](auto warm) {
        chunk_pass<2>(
            N, wave, kWavesB, h.part,
            [&](int k) {
                Init v;
                v.bk = (pb.rhs ? pb.rhs[k] : pb.f) * h2;
                v.rk = v.bk;
                v.u0 = 0.0;
                if constexpr (decltype(warm)::value) {
                    const int i = k / n, j = k - i * n;
                    v.rk = v.bk - apply_at(n, A0.ex, A0.ny, i, j, [&](int t) { return pb.in[t]; });
                    v.u0 = pb.in[k];
                }
                return v;
            },
            [&](int k, const Init& v, double* acc) {
                acc[0] += v.bk * v.bk;
                u[k] = v.u0;
                h.r[k] = v.rk;
                acc[1] += v.rk * v.rk;
                if (mixed) b0f[k] = static_cast<float>(v.rk);
            });
    };
    if (pb.in)
        init(Flag<true>{});
    else
        init(Flag<false>{});
    reduce_parts(s, 2, C);
    const double bnorm = sqrt(s.red[0]);
    const int levels = mg ? h.L : 1;
    if (bnorm == 0.0) {  // solver.hpp:105-108: f == 0 gives u == 0 exactly
        for (int k = e.tid; k < N; k += e.nthr) u[k] = 0.0;
        if (threadIdx.x == 0) st = ProblemStats{0, kConverged, 1, 0, 0.0, -1};
        return;
    }
    double res = pb.in ? sqrt(s.red[1]) / bnorm : 1.0;
    // History entries past history_cap are not stored (the host reports at
    // most that many).
    const bool hist = P.record_history && pb.history;
    if (hist && threadIdx.x == 0 && P.history_cap > 0) pb.history[0] = res;
    if (pb.in && res < P.tol) {
        if (threadIdx.x == 0) st = ProblemStats{0, kConverged, levels, hist ? 1 : 0, res, -1};
        return;
    }

    precondition(cyc, h, P, N, e);
    chunk_pass<1>(
        N, wave, kWavesB, h.part, [&](int k) { return h.r[k] * zval(k); },
        [&](int, double v, double* acc) { acc[0] += v; });
    reduce_parts(s, 1, C);
    double rz = s.red[0];

    double* pold = h.p;
    double* pnew = h.u;  // the slot's u is free: the solution lives in pb.out
    double beta = 0.0;
    bool first = true;
    int status = kMaxIter, iters = 0;
    // With the FP32 cycle, u += alpha p of iteration it is applied by the
    // direction pass of it + 1 (or after the loop), where p is read anyway.
    // The FP64 kernel, short of registers, keeps it in the update pass.
    constexpr bool defer = mixed;
    struct Upd {
        double u, p, r, ap;
    };
    // z is read as stored (float for the FP32 cycle) and widened after the
    // loads, so that no conversion waits between them.
    struct Rz {
        double r;
        T z;
        double ap;
    };
    double alpha = 0.0;
    bool pending = false;  // a deferred u += alpha p
    for (int it = 0; it < P.max_iter; ++it) {
        // p = z + beta p, ap = A p, <p, ap>; the pending u update.
        if constexpr (mixed)
            cg_direction<1, defer>(n, A0, x0f, pold, beta, first, pnew, h.ap, wave, kWavesB, h.part,
                                   pending ? u : nullptr, alpha);
        else
            cg_direction<1, defer>(n, A0, static_cast<const double*>(h.z), pold, beta, first, pnew, h.ap, wave,
                                   kWavesB, h.part);
        pending = false;
        reduce_parts(s, 1, C);
        const double pap = s.red[0];
        // solver.hpp:149-155; breakdown_at forces the exit for tests.
        if (!(fabs(pap) > 0.0) || !isfinite(pap) || it + 1 == P.breakdown_at) {
            status = kBreakdown;
            iters = it;
            break;
        }
        alpha = rz / pap;
        pending = defer;
        // u += alpha p (unless deferred); r -= alpha ap; <r, r>  (blas.hpp:62-77)
        chunk_pass<1>(
            N, wave, kWavesB, h.part,
            [&](int k) { return defer ? Upd{0.0, 0.0, h.r[k], h.ap[k]} : Upd{u[k], pnew[k], h.r[k], h.ap[k]}; },
            [&](int k, const Upd& v, double* acc) {
                if (!defer) u[k] = v.u + alpha * v.p;
                const double ri = v.r - alpha * v.ap;
                h.r[k] = ri;
                acc[0] += ri * ri;
                if (mixed) b0f[k] = static_cast<float>(ri);
            });
        reduce_parts(s, 1, C);
        res = sqrt(s.red[0]) / bnorm;
        iters = it + 1;
        if (hist && threadIdx.x == 0 && iters < P.history_cap) pb.history[iters] = res;
        if (res < P.tol) {
            status = kConverged;
            break;
        }
        if (iters == P.max_iter) break;
        precondition(cyc, h, P, N, e);
        const T* const z0 = mixed ? reinterpret_cast<const T*>(h.x[0]) : reinterpret_cast<const T*>(h.z);
        chunk_pass<2>(
            N, wave, kWavesB, h.part, [&](int k) { return Rz{h.r[k], z0[k], mixed ? h.ap[k] : 0.0}; },
            [&](int, const Rz& v, double* acc) {
                const double z = static_cast<double>(v.z);
                acc[0] += v.r * z;
                if (mixed) acc[1] += z * v.ap;
            });
        reduce_parts(s, mixed ? 2 : 1, C);
        beta = mixed ? (-alpha * s.red[1]) / rz : s.red[0] / rz;
        rz = s.red[0];
        double* t = pold;
        pold = pnew;
        pnew = t;
        first = false;
    }
    if (pending) {  // the last u += alpha p; reduce_parts ended with a barrier
        for (int k = e.tid; k < N; k += e.nthr) u[k] = u[k] + alpha * pnew[k];
    }
    if (threadIdx.x == 0) st = ProblemStats{iters, status, levels, hist ? iters + 1 : 0, res, -1};
}

template <class T, int B>
__global__ __launch_bounds__(B, B == kBlock ? 4 : 1) void k_batch(BatchArgs A) {
    __shared__ Shared s;
    __shared__ __align__(16) T pool[pool_elems<T, B>()];
    __shared__ __align__(16) float tail[sizeof(T) == 4 ? kTailLds : 1];
    char* slot = A.slots + static_cast<size_t>(blockIdx.x) * A.slot_bytes;
    for (;;) {
        if (threadIdx.x == 0) s.prob = static_cast<int>(atomicAdd(A.next, 1u));
        __syncthreads();
        const int t = s.prob;
        if (t >= A.count) return;
        const int p = A.order ? A.order[t] : t;
        run_problem<T, B>(s, pool, tail, A.probs[p], A.stats[p], A.prm, slot);
        __syncthreads();
    }
}

template __global__ void k_batch<double, kBlock>(BatchArgs);
template __global__ void k_batch<float, kBlock>(BatchArgs);
template __global__ void k_batch<double, kWideBlock>(BatchArgs);
template __global__ void k_batch<float, kWideBlock>(BatchArgs);

// Scratch of the allocation-free entry point (darcy_solve_batch_noalloc):
// its problem tables, work-queue counter and solve slots.  Each device the
// library runs on holds one copy until the process ends; calls on one device
// share it.
constexpr size_t kScratchBytes = size_t{64} << 20;
static __device__ __align__(256) char scratch[kScratchBytes];

// Zeroes `words` 32-bit words (the work-queue counter): a kernel rather than
// a memset, so that the sequence stays valid inside a captured graph.
__global__ void k_zero_words(uint32_t* p, int words) {
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < words) p[i] = 0u;
}

// Streaming copy of `n` doubles: the reference kernel for achievable HBM
// bandwidth and for scaling the byte counters of the profiler.
__global__ __launch_bounds__(256) void k_stream_copy(const double2* __restrict__ src, double2* __restrict__ dst,
                                                      size_t n2) {
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n2;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        dst[i] = src[i];
}

}  // namespace darcy_hip

// Launch entry points for the host code of this library.
// wide: the kWideBlock-thread kernel.
extern "C" hipError_t darcy_hip_launch_batch(const darcy_hip::BatchArgs* args, int grid, int mixed, int wide,
                                             hipStream_t stream) {
    using namespace darcy_hip;
    if (wide) {
        if (mixed)
            hipLaunchKernelGGL((k_batch<float, kWideBlock>), dim3(grid), dim3(kWideBlock), 0, stream, *args);
        else
            hipLaunchKernelGGL((k_batch<double, kWideBlock>), dim3(grid), dim3(kWideBlock), 0, stream, *args);
    } else if (mixed) {
        hipLaunchKernelGGL((k_batch<float, kBlock>), dim3(grid), dim3(kBlock), 0, stream, *args);
    } else {
        hipLaunchKernelGGL((k_batch<double, kBlock>), dim3(grid), dim3(kBlock), 0, stream, *args);
    }
    return hipGetLastError();
}

extern "C" hipError_t darcy_hip_batch_occupancy(int mixed, int wide, int* blocks_per_cu) {
    using namespace darcy_hip;
    if (wide)
        return mixed ? hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_batch<float, kWideBlock>,
                                                                    kWideBlock, 0)
                     : hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_batch<double, kWideBlock>,
                                                                    kWideBlock, 0);
    return mixed ? hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_batch<float, kBlock>, kBlock, 0)
                 : hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_batch<double, kBlock>, kBlock, 0);
}

// Device address of `scratch` on the current device.  The array's name is a
// host-side stand-in, so every host use of its memory goes through the
// address looked up here, once per device.
extern "C" hipError_t darcy_hip_scratch(void** ptr, size_t* bytes) {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static void* addr[kMaxDevices] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(darcy_hip::scratch));
        if (e != hipSuccess) return e;
        addr[dev] = p;
    }
    *ptr = addr[dev];
    *bytes = darcy_hip::kScratchBytes;
    return hipSuccess;
}

extern "C" hipError_t darcy_hip_zero_words(uint32_t* p, int words, hipStream_t stream) {
    hipLaunchKernelGGL(darcy_hip::k_zero_words, dim3((words + 255) / 256), dim3(256), 0, stream, p, words);
    return hipGetLastError();
}

extern "C" hipError_t darcy_hip_launch_copy(const void* src, void* dst, size_t bytes, int grid, hipStream_t stream) {
    hipLaunchKernelGGL(darcy_hip::k_stream_copy, dim3(grid), dim3(256), 0, stream, static_cast<const double2*>(src),
                       static_cast<double2*>(dst), bytes / sizeof(double2));
    return hipGetLastError();
}
