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
#include "abi_types.hpp"
#include "setup.hpp"

namespace darcy_hip {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / kWave;

// Largest side of the levels kept in LDS: 16 (14 KB in FP64, 7.5 KB in FP32;
// the FP32 kernel also keeps the 16 KB tail matrix), so four workgroups fit
// on a CU.
template <class T>
struct LdsSide {
    static constexpr int value = 16;
};

// LDS elements of the coarse-level pool.  The FP32 pool doubles as FP64
// scratch while the tail matrix is built: FP64 operators of the tail levels
// (<= 285 doubles) and x, b for one unit vector per wave (4 x 170 doubles).
constexpr int kTailScratchDoubles = 285 + 4 * 170;
template <class T>
constexpr int pool_elems() {
    constexpr int need = LdsPool<T, LdsSide<T>::value>::kElems;
    constexpr int scratch = sizeof(T) == 4 ? 2 * kTailScratchDoubles : 0;
    return need > scratch ? need : scratch;
}
// FP32 tail matrix, stored by columns (64 x 64 at most), and a residual.
constexpr int kTailMat = 64 * 64;
constexpr int kTailLds = kTailMat + 64;

struct Shared {
    Hier h;
    LdsTable t;
    double red[4];
    unsigned long long bad;
    int prob;
};

// --------------------------------------------------------------- reductions

__device__ inline double wave_sum(double s) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, kWave);
    return __shfl(s, 0, kWave);
}

// Canonical chunked reduction.  For each chunk owned by this wave, every
// lane loads its four cells first (`load(k)` returns a value bundle), then
// commits them in order (`commit(k, bundle, acc)` adds to acc[0..ND)).
// A wave takes CPR chunks per round, issuing all their loads first.
template <int ND, int CPR = 1, class Load, class Commit>
DARCY_INLINE void chunk_pass(int N, int w, int nw, double* const* part, Load&& load, Commit&& commit) {
    const int lane = threadIdx.x & (kWave - 1);
    const int C = chunks_of(N);
    for (int c0 = w * CPR; c0 < C; c0 += nw * CPR) {
        decltype(load(0)) v[4 * CPR];
        bool ok[4 * CPR];
#pragma unroll
        for (int u = 0; u < 4 * CPR; ++u) {
            const int k = (c0 + u / 4) * kChunk + (u % 4) * kWave + lane;
            ok[u] = k < N;
            if (ok[u]) v[u] = load(k);
        }
#pragma unroll
        for (int cc = 0; cc < CPR; ++cc) {
            double acc[ND];
#pragma unroll
            for (int d = 0; d < ND; ++d) acc[d] = 0.0;
#pragma unroll
            for (int m = 0; m < 4; ++m)
                if (ok[4 * cc + m]) commit((c0 + cc) * kChunk + m * kWave + lane, v[4 * cc + m], acc);
            if (c0 + cc < C) {
#pragma unroll
                for (int d = 0; d < ND; ++d) {
                    const double s = wave_sum(acc[d]);
                    if (lane == 0) part[d][c0 + cc] = s;
                }
            }
        }
    }
}

__device__ inline double final_sum(const double* part, int C) {
    const int lane = threadIdx.x & (kWave - 1);
    double s = 0.0;
    for (int c = lane; c < C; c += kWave) s += part[c];
    return wave_sum(s);
}

// Barrier, then one wave combines the chunk sums of `nd` reductions.
__device__ inline void reduce_parts(Shared& s, int nd, int C) {
    __syncthreads();
    if (threadIdx.x < kWave) {
        for (int d = 0; d < nd; ++d) {
            const double v = final_sum(s.h.part[d], C);
            if (threadIdx.x == 0) s.red[d] = v;
        }
    }
    __syncthreads();
}

// ------------------------------------------------------------ the cycle

template <class T>
DARCY_INLINE LvP<T> global_level(const Hier& h, int l);
template <>
DARCY_INLINE LvP<double> global_level<double>(const Hier& h, int l) {
    return {h.n[l], h.ex[l], h.ny[l], h.id[l], static_cast<double*>(h.x[l]), static_cast<double*>(h.b[l])};
}
template <>
DARCY_INLINE LvP<float> global_level<float>(const Hier& h, int l) {
    return {h.n[l], h.ex32[l], h.ny32[l], h.id32[l], static_cast<float*>(h.x[l]), static_cast<float*>(h.b[l])};
}

// The cycle of multigrid.hpp:97-127 run by one workgroup, with the coarse
// levels in the LDS pool and the smallest ones on a single wave.
template <class T>
struct Cycle {
    const Hier& h;
    const LdsTable& t;
    T* pool;
    float* tail;  // FP32 kernel: tail matrix by columns, then a residual vector
    CycleOpts o;

    DARCY_INLINE LvP<T> G(int l) const { return global_level<T>(h, l); }
    DARCY_INLINE LvP<T> S(int l) const {
        const int* f = t.off[l];
        return {h.n[l], pool + f[0], pool + f[1], pool + f[2], pool + f[3], pool + f[4]};
    }

    // --- one wave, everything in LDS (levels >= t.wave): the recursive cycle
    DARCY_INLINE void wave_cycle(int l0, bool fresh0, Exec e) const {
        cycle_levels<T>(l0, h.L, fresh0, o, e, WaveSync{}, [&](int l) { return S(l); });
    }

    // --- FP32 cycle: the tail as its matrix.  Fresh: x = M b; repeated
    // W-cycle visit: x += M (b - A x).  Lane i owns row i; the row sum runs
    // over k in order.
    DARCY_INLINE void dense_tail(int l, bool fresh, Exec e) const {
        const LvP<T> A = S(l);
        const int n = A.n, dim = n * n, i = e.tid;
        const WaveSync sync;
        float* res = tail + kTailMat;
        const T* rhs = A.b;
        if (!fresh) {
            if (i < dim) {
                const int r = i / n, c = i - r * n;
                res[i] = A.b[i] - apply_at(n, A.ex, A.ny, r, c, [&](int q) { return A.x[q]; });
            }
            sync();
            rhs = res;
        }
        if (i < dim) {
            float s = 0.0f;
            for (int k = 0; k < dim; ++k) s += tail[k * dim + i] * rhs[k];
            A.x[i] = fresh ? s : A.x[i] + s;
        }
        sync();
    }

    // Global-memory levels stream one pass per visit when one sweep is
    // asked for (the default); otherwise they run phase by phase.
    DARCY_INLINE void down(int l, bool fresh, Exec e) const {
        const BlockSync sync;
        if (l >= t.first)
            visit_down(S(l), S(l + 1).b, fresh, o, e, sync);
        else if (l + 1 >= t.first)
            o.pre == 1 ? down_stream(G(l), S(l + 1).b, fresh, e) : visit_down(G(l), S(l + 1).b, fresh, o, e, sync);
        else
            o.pre == 1 ? down_stream(G(l), G(l + 1).b, fresh, e) : visit_down(G(l), G(l + 1).b, fresh, o, e, sync);
    }

    DARCY_INLINE void up(int l, Exec e) const {
        const BlockSync sync;
        if (l >= t.first)
            visit_up(S(l), S(l + 1).x, o, e, sync);
        else if (l + 1 >= t.first)
            o.post == 1 ? up_stream(G(l), static_cast<const T*>(S(l + 1).x), e)
                        : visit_up(G(l), static_cast<const T*>(S(l + 1).x), o, e, sync);
        else
            o.post == 1 ? up_stream(G(l), static_cast<const T*>(G(l + 1).x), e)
                        : visit_up(G(l), static_cast<const T*>(G(l + 1).x), o, e, sync);
    }

    // z = M r from level l0 (x_{l0} zero on entry when fresh0).
    DARCY_INLINE void run(int l0, bool fresh0, Exec e) const {
        const int L = h.L;
        uint32_t done = 0;
        int l = l0;
        bool fresh = fresh0;
        for (;;) {
            while (l + 1 < L && l < t.wave) {
                down(l, fresh, e);
                done &= ~(1u << l);
                ++l;
                fresh = true;
            }
            if (l >= t.wave) {
                if (threadIdx.x < kWave) {
                    const Exec we{static_cast<int>(threadIdx.x), kWave};
                    if constexpr (sizeof(T) == 4)
                        dense_tail(l, fresh, we);
                    else
                        wave_cycle(l, fresh, we);
                }
                __syncthreads();
            } else if (l >= t.first) {
                coarsest_solve(S(l), fresh, e, BlockSync{});
            } else {
                coarsest_solve(G(l), fresh, e, BlockSync{});
            }
            for (;;) {
                if (l == l0) return;
                --l;
                if (o.gamma == 2 && !((done >> l) & 1u)) {
                    done |= 1u << l;
                    ++l;
                    fresh = false;
                    break;
                }
                up(l, e);
            }
        }
    }
};

// Tail matrix of the FP32 cycle.  Column k is the FP64 recursive cycle from
// level tl -- the reference's arithmetic, bit-exact with the oracle -- applied
// to the unit vector e_k, rounded to FP32 and stored by columns.  The four
// waves take columns in turn; `scratch` (the LDS pool, before it receives the
// FP32 operators) holds the FP64 operators of the tail levels and each wave's
// x and b.
DARCY_INLINE void build_tail_matrix(const Hier& h, int tl, const CycleOpts& o, void* scratch, float* tail) {
    double* dp = static_cast<double*>(scratch);
    const int L = h.L, lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    auto op_off = [&](int l) {
        int off = 0;
        for (int q = tl; q < l; ++q) off += 2 * h.n[q] * (h.n[q] + 1) + h.n[q] * h.n[q];
        return off;
    };
    const int ops_end = op_off(L);
    for (int l = tl; l < L; ++l) {
        const int n = h.n[l], nf = n * (n + 1), base = op_off(l);
        for (int q = threadIdx.x; q < nf; q += kBlock) {
            dp[base + q] = h.ex[l][q];
            dp[base + nf + q] = h.ny[l][q];
        }
        for (int q = threadIdx.x; q < n * n; q += kBlock) dp[base + 2 * nf + q] = h.id[l][q];
    }
    __syncthreads();
    double* vec = dp + ops_end + wave * 170;
    auto view = [&](int l) {
        int off = 0;
        for (int q = tl; q < l; ++q) off += 2 * h.n[q] * h.n[q];
        const int n = h.n[l], base = op_off(l);
        return LvP<double>{n, dp + base, dp + base + n * (n + 1), dp + base + 2 * n * (n + 1), vec + off,
                           vec + off + n * n};
    };
    const int dim = h.n[tl] * h.n[tl];
    const WaveSync sync;
    for (int k = wave; k < dim; k += kWaves) {
        const LvP<double> top = view(tl);
        for (int q = lane; q < dim; q += kWave) top.b[q] = q == k ? 1.0 : 0.0;
        sync();
        cycle_levels<double>(tl, L, true, o, Exec{lane, kWave}, sync, view);
        for (int q = lane; q < dim; q += kWave) tail[k * dim + q] = static_cast<float>(top.x[q]);
        sync();
    }
    __syncthreads();
}

// --------------------------------------------------------------- a problem

template <class T>
DARCY_INLINE void precondition(const Cycle<T>& cyc, const Hier& h, const SolveParams& P, int N, Exec e) {
    if (P.precond == kPcMultigrid || P.mode == kModeApplyPreconditioner) {
        cyc.run(0, true, e);
    } else if (P.precond == kPcJacobi) {  // solver.hpp:127-132
        for (int k = e.tid; k < N; k += e.nthr) h.z[k] = h.r[k] * h.id[0][k];
        __syncthreads();
    } else {  // solver.hpp:124-126
        for (int k = e.tid; k < N; k += e.nthr) h.z[k] = h.r[k];
        __syncthreads();
    }
}

template <class T>
DARCY_INLINE void run_problem(Shared& s, T* pool, float* tail, const Problem& pb, ProblemStats& st,
                              const SolveParams& P, char* slot) {
    const Exec e{static_cast<int>(threadIdx.x), kBlock};
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
        plan_slot(slot, n, mixed, &s.h);
        plan_lds<T, LdsSide<T>::value>(s.h, s.t);
    }
    __syncthreads();
    const Hier& h = s.h;

    // Hierarchy (level.hpp:79-97, multigrid.hpp:55-71, level.hpp:171-211).
    build_fine_faces(pb.a, n, h.ex[0], h.ny[0], e);
    build_fine_invdiag(pb.a, n, h.id[0], e);
    __syncthreads();
    if (mg) {
        for (int l = 0; l + 1 < h.L; ++l) {
            galerkin_faces(h.n[l], h.ex[l], h.ny[l], h.ex[l + 1], h.ny[l + 1], e);
            galerkin_invdiag(h.n[l], h.ex[l], h.ny[l], h.id[l + 1], e);
            if (mixed) to_fp32(h.n[l], h.ex[l], h.ny[l], h.id[l], h.ex32[l], h.ny32[l], h.id32[l], e);
            __syncthreads();
        }
        if (mixed) {
            const int l = h.L - 1;
            to_fp32(h.n[l], h.ex[l], h.ny[l], h.id[l], h.ex32[l], h.ny32[l], h.id32[l], e);
            __syncthreads();
        }
        if constexpr (mixed) {
            if (s.t.wave < h.L) build_tail_matrix(h, s.t.wave, CycleOpts{P.pre, P.post, P.gamma}, pool, tail);
        }
        // Coarse operators into LDS for the whole solve.
        for (int l = s.t.first; l < h.L; ++l) {
            const LvP<T> g = global_level<T>(h, l);
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
    const Cycle<T> cyc{h, s.t, pool, tail, CycleOpts{P.pre, P.post, P.gamma}};
    const LvP<double> A0 = global_level<double>(h, 0);
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
    chunk_pass<2>(
        N, wave, kWaves, h.part,
        [&](int k) {
            Init v;
            v.bk = (pb.rhs ? pb.rhs[k] : pb.f) * h2;
            v.rk = v.bk;
            v.u0 = 0.0;
            if (pb.in) {
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
    reduce_parts(s, 2, C);
    const double bnorm = sqrt(s.red[0]);
    const int levels = mg ? h.L : 1;
    if (bnorm == 0.0) {  // solver.hpp:105-108: f == 0 gives u == 0 exactly
        for (int k = e.tid; k < N; k += e.nthr) u[k] = 0.0;
        if (threadIdx.x == 0) st = ProblemStats{0, kConverged, 1, 0, 0.0, -1};
        return;
    }
    double res = pb.in ? sqrt(s.red[1]) / bnorm : 1.0;
    const bool hist = P.record_history && pb.history;
    if (hist && threadIdx.x == 0) pb.history[0] = res;
    if (pb.in && res < P.tol) {
        if (threadIdx.x == 0) st = ProblemStats{0, kConverged, levels, hist ? 1 : 0, res, -1};
        return;
    }

    precondition(cyc, h, P, N, e);
    chunk_pass<1>(
        N, wave, kWaves, h.part, [&](int k) { return h.r[k] * zval(k); },
        [&](int, double v, double* acc) { acc[0] += v; });
    reduce_parts(s, 1, C);
    double rz = s.red[0];

    double* pold = h.p;
    double* pnew = h.u;  // the slot's u is free: the solution lives in pb.out
    double beta = 0.0;
    bool first = true;
    int status = kMaxIter, iters = 0;
    struct PAp {
        double p, ap;
    };
    struct Upd {
        double u, p, r, ap;
    };
    struct Rz {
        double r, z, ap;
    };
    for (int it = 0; it < P.max_iter; ++it) {
        // p = z + beta p (solver.hpp:171, blas.hpp:50-56), fused with
        // ap = A p (solver.hpp:147) and <p, ap> (solver.hpp:148).
        chunk_pass<1>(
            N, wave, kWaves, h.part,
            [&](int k) {
                const int i = k / n, j = k - i * n;
                auto pv = [&](int t) -> double { return first ? zval(t) : zval(t) + beta * pold[t]; };
                PAp v;
                v.ap = apply_at(n, A0.ex, A0.ny, i, j, pv);
                v.p = pv(k);
                return v;
            },
            [&](int k, const PAp& v, double* acc) {
                pnew[k] = v.p;
                h.ap[k] = v.ap;
                acc[0] += v.p * v.ap;
            });
        reduce_parts(s, 1, C);
        const double pap = s.red[0];
        if (!(fabs(pap) > 0.0) || !isfinite(pap)) {  // solver.hpp:149-155
            status = kBreakdown;
            iters = it;
            break;
        }
        const double alpha = rz / pap;
        // u += alpha p; r -= alpha ap; <r, r>  (blas.hpp:62-77)
        chunk_pass<1>(
            N, wave, kWaves, h.part, [&](int k) { return Upd{u[k], pnew[k], h.r[k], h.ap[k]}; },
            [&](int k, const Upd& v, double* acc) {
                u[k] = v.u + alpha * v.p;
                const double ri = v.r - alpha * v.ap;
                h.r[k] = ri;
                acc[0] += ri * ri;
                if (mixed) b0f[k] = static_cast<float>(ri);
            });
        reduce_parts(s, 1, C);
        res = sqrt(s.red[0]) / bnorm;
        iters = it + 1;
        if (hist && threadIdx.x == 0) pb.history[iters] = res;
        if (res < P.tol) {
            status = kConverged;
            break;
        }
        if (iters == P.max_iter) break;
        precondition(cyc, h, P, N, e);
        chunk_pass<2>(
            N, wave, kWaves, h.part, [&](int k) { return Rz{h.r[k], zval(k), mixed ? h.ap[k] : 0.0}; },
            [&](int, const Rz& v, double* acc) {
                acc[0] += v.r * v.z;
                if (mixed) acc[1] += v.z * v.ap;
            });
        reduce_parts(s, mixed ? 2 : 1, C);
        beta = mixed ? (-alpha * s.red[1]) / rz : s.red[0] / rz;
        rz = s.red[0];
        double* t = pold;
        pold = pnew;
        pnew = t;
        first = false;
    }
    if (threadIdx.x == 0) st = ProblemStats{iters, status, levels, hist ? iters + 1 : 0, res, -1};
}

template <class T>
__global__ __launch_bounds__(kBlock, 4) void k_batch(BatchArgs A) {
    __shared__ Shared s;
    __shared__ __align__(16) T pool[pool_elems<T>()];
    __shared__ __align__(16) float tail[sizeof(T) == 4 ? kTailLds : 1];
    char* slot = A.slots + static_cast<size_t>(blockIdx.x) * A.slot_bytes;
    for (;;) {
        if (threadIdx.x == 0) s.prob = static_cast<int>(atomicAdd(A.next, 1u));
        __syncthreads();
        const int t = s.prob;
        if (t >= A.count) return;
        const int p = A.order ? A.order[t] : t;
        run_problem<T>(s, pool, tail, A.probs[p], A.stats[p], A.prm, slot);
        __syncthreads();
    }
}

template __global__ void k_batch<double>(BatchArgs);
template __global__ void k_batch<float>(BatchArgs);

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
extern "C" hipError_t darcy_hip_launch_batch(const darcy_hip::BatchArgs* args, int grid, int mixed,
                                             hipStream_t stream) {
    if (mixed)
        hipLaunchKernelGGL(darcy_hip::k_batch<float>, dim3(grid), dim3(darcy_hip::kBlock), 0, stream, *args);
    else
        hipLaunchKernelGGL(darcy_hip::k_batch<double>, dim3(grid), dim3(darcy_hip::kBlock), 0, stream, *args);
    return hipGetLastError();
}

extern "C" hipError_t darcy_hip_batch_occupancy(int mixed, int* blocks_per_cu) {
    if (mixed)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, darcy_hip::k_batch<float>,
                                                            darcy_hip::kBlock, 0);
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, darcy_hip::k_batch<double>,
                                                        darcy_hip::kBlock, 0);
}

extern "C" hipError_t darcy_hip_launch_copy(const void* src, void* dst, size_t bytes, int grid, hipStream_t stream) {
    hipLaunchKernelGGL(darcy_hip::k_stream_copy, dim3(grid), dim3(256), 0, stream, static_cast<const double2*>(src),
                       static_cast<double2*>(dst), bytes / sizeof(double2));
    return hipGetLastError();
}
