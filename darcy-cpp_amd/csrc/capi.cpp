// Host side of the C ABI (include/darcy_hip.h): argument checks, device
// workspace, staging of host buffers, kernel launches and result readback.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <mutex>
#include <numeric>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/darcy_hip.h"
#include "abi_types.hpp"
#include "ctx.hpp"

using namespace darcy_hip;

using CycleOptsC = darcy_hip::CycleOpts;
extern "C" hipError_t darcy_hip_launch_batch(const BatchArgs* args, int grid, int mixed, int wide, hipStream_t stream);
extern "C" hipError_t darcy_grid_fine(const double* a, const Hier* hd, unsigned long long* bad, int blocks,
                                      hipStream_t s);
extern "C" hipError_t darcy_grid_galerkin(const Hier* hd, int l, int blocks, hipStream_t s);
extern "C" hipError_t darcy_grid_fp32(const Hier* hd, int l, int blocks, hipStream_t s);
extern "C" hipError_t darcy_grid_tail_matrix(const Hier* hd, CycleOptsC o, float* out, hipStream_t s);
extern "C" hipError_t darcy_grid_tail_matrix64(const Hier* hd, CycleOptsC o, double* out, hipStream_t s);
extern "C" hipError_t darcy_grid_presmooth(int mixed, const Hier* hd, int l, int kind, int blocks, hipStream_t s);
extern "C" hipError_t darcy_grid_restrict(int mixed, const Hier* hd, int l, int blocks, hipStream_t s);
extern "C" hipError_t darcy_grid_postsmooth(int mixed, const Hier* hd, int l, int kind, int blocks, hipStream_t s);
extern "C" hipError_t darcy_grid_tail(int mixed, const Hier* hd, int l0, int mode, CycleOptsC o, const float* tail,
                                      const double* tail64, const void* big, int big_l, int big_ld, hipStream_t s);
extern "C" hipError_t darcy_grid_big_matrix(int mixed, const Hier* hd, CycleOptsC o, int lb, int dim, void* out, int ld,
                                            hipStream_t s);
extern "C" hipError_t darcy_grid_cg(int which, const GridCg* g, const double* pold, double* pnew, int blocks,
                                    hipStream_t s);
extern "C" hipError_t darcy_hip_batch_occupancy(int mixed, int wide, int* blocks_per_cu);
extern "C" hipError_t darcy_grid_check_kernels(int device, const char** what);
extern "C" hipError_t darcy_hip_launch_copy(const void* src, void* dst, size_t bytes, int grid, hipStream_t stream);
extern "C" hipError_t darcy_hip_scratch(void** ptr, size_t* bytes);
extern "C" hipError_t darcy_hip_zero_words(uint32_t* p, int words, hipStream_t stream);
extern "C" hipError_t darcy_hip_launch_grf(int n, int count, double alpha, double tau, const double* xi,
                                           double* cmat, double* amp, double* t, double* g, double* a, double hi,
                                           double lo, hipStream_t s);

namespace {

int fail(darcy_ctx* c, int code, const std::string& msg) {
    if (c) c->error = msg;
    return code;
}

int hip_fail(darcy_ctx* c, hipError_t e, const char* what) {
    const int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? DARCY_OUT_OF_MEMORY : DARCY_HIP_ERROR;
    return fail(c, code, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(expr, what)                           \
    do {                                              \
        const hipError_t e_ = (expr);                 \
        if (e_ != hipSuccess) return hip_fail(ctx, e_, what); \
    } while (0)

// Option checks in the order and wording of the reference
// (solver.hpp:85-87, bindings.cpp:43-46).
int check_opts(darcy_ctx* ctx, const darcy_opts& o) {
    if (!(o.tol > 0.0)) return fail(ctx, DARCY_INVALID_INPUT, "tol must be positive");
    if (o.max_iter < 1) return fail(ctx, DARCY_INVALID_INPUT, "max_iter must be at least 1");
    if (o.preconditioner < DARCY_PC_NONE || o.preconditioner > DARCY_PC_MULTIGRID)
        return fail(ctx, DARCY_INVALID_INPUT, "unknown preconditioner; expected one of: multigrid, jacobi, none");
    if (o.pre_sweeps < 1 || o.post_sweeps < 1)
        return fail(ctx, DARCY_INVALID_INPUT, "pre_sweeps and post_sweeps must be at least 1");
    if (o.cycle_index < 1 || o.cycle_index > 2)
        return fail(ctx, DARCY_INVALID_INPUT, "cycle_index must be 1 (V-cycle) or 2 (W-cycle)");
    if (o.precision != DARCY_FP64 && o.precision != DARCY_MIXED)
        return fail(ctx, DARCY_INVALID_INPUT, "precision must be 0 (fp64) or 1 (mixed)");
    if (o.breakdown_at < 0) return fail(ctx, DARCY_INVALID_INPUT, "breakdown_at must be 0 (off) or positive");
    return DARCY_OK;
}

// Largest grid side one call may use: flat indices and face counts are
// 32-bit inside the kernels.
constexpr int64_t kMaxSide = 32768;
// Pipeline step of a wide (1024-thread) batch workgroup: the tuned step of a
// 256-thread one times 4, the same cells per thread and per L2 share (a
// quarter as many workgroups share each L2).
constexpr int kWideStepFactor = 4;

// Checks every entry point makes before any device work: the side, and the
// arrays every solve reads and writes.
int check_problem(darcy_ctx* ctx, int64_t n, const double* a, const double* out) {
    if (n < 1) return fail(ctx, DARCY_INVALID_INPUT, "grid size must be at least 1");
    if (n > kMaxSide) return fail(ctx, DARCY_INVALID_INPUT, "grid size exceeds 32768");
    if (!a || !out) return fail(ctx, DARCY_INVALID_INPUT, "null array argument");
    return DARCY_OK;
}

// Entries of a history buffer: max_iter + 1 are all a solve can write, and
// the caller's capacity bounds them.
size_t history_entries(const darcy_opts& o, int32_t cap) {
    const size_t want = static_cast<size_t>(o.max_iter) + 1;
    return std::min<size_t>(want, static_cast<size_t>(std::max<int32_t>(cap, 0)));
}

struct Job {
    int mode;
    const darcy_opts* opts;
    int64_t count;
    const int64_t* n;
    const double* const* a;
    const double* const* rhs;
    const double* f;
    const double* const* in;
    double* const* out;
    darcy_stats* stats;
    double* history;  // single-problem history (host or device per opts)
    int32_t history_cap;
    bool device_ptrs;
};

// Per-problem stats for the caller, and the reference's message for the
// lowest problem with an invalid permeability (solver.hpp:68-77).
int report(darcy_ctx* ctx, const std::vector<ProblemStats>& st, darcy_stats* out, int32_t history_cap) {
    const int64_t count = static_cast<int64_t>(st.size());
    int64_t first_bad_problem = -1;
    for (int64_t i = 0; i < count; ++i) {
        const ProblemStats& s = st[i];
        if (out) {
            darcy_stats& d = out[i];
            d.iterations = s.iterations;
            d.converged = s.status == kConverged ? 1 : 0;
            d.levels = s.levels;
            d.status = s.status;
            d.residual = s.residual;
            d.bad_index = s.status == kInvalid ? s.bad_index : -1;
            d.history_len = std::min<int32_t>(s.history_len, history_cap);
            d.reserved = 0;
        }
        if (s.status == kInvalid && first_bad_problem < 0) first_bad_problem = i;
    }
    if (first_bad_problem >= 0) {
        char msg[256];
        std::snprintf(msg, sizeof msg,
                      "permeability must be finite and strictly positive everywhere; offending value at flat index "
                      "%lld%s",
                      static_cast<long long>(st[first_bad_problem].bad_index),
                      count > 1 ? (" of problem " + std::to_string(first_bad_problem)).c_str() : "");
        return fail(ctx, DARCY_INVALID_INPUT, msg);
    }
    ctx->error.clear();
    return DARCY_OK;
}

// Solve parameters of the persistent kernel from the options.
SolveParams solve_params(const darcy_ctx* ctx, const darcy_opts& o, int mode, bool mixed, int32_t history_cap) {
    SolveParams p{};
    p.tol = o.tol;
    p.max_iter = o.max_iter;
    p.precond = o.preconditioner;
    p.pre = o.pre_sweeps;
    p.post = o.post_sweeps;
    p.gamma = o.cycle_index;
    p.mixed = mixed ? 1 : 0;
    p.mode = mode;
    p.record_history = history_cap > 0 ? 1 : 0;
    p.history_cap = history_cap;
    p.step_cells = mixed ? ctx->batch_step_cells_fp32 : ctx->batch_step_cells;
    p.min_coarse = o.min_coarse_n;
    p.max_levels = o.max_levels;
    p.breakdown_at = o.breakdown_at;
    return p;
}

int run(darcy_ctx* ctx, const Job& job) {
    const darcy_opts& o = *job.opts;
    const bool mg = job.mode == kModeApplyPreconditioner || o.preconditioner == DARCY_PC_MULTIGRID;
    const bool mixed = mg && o.precision == DARCY_MIXED;
    const int64_t count = job.count;
    if (count == 0) return DARCY_OK;
    int nmax = 0;
    for (int64_t i = 0; i < count; ++i) {
        if (const int rc = check_problem(ctx, job.n[i], job.a[i], job.out[i])) return rc;
        nmax = std::max<int>(nmax, static_cast<int>(job.n[i]));
    }
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");

    // Device copies of host inputs; identical host pointers (a shared rhs)
    // are uploaded once, with the largest size any problem reads through
    // them (a pointer shared by problems of different sides).
    std::vector<Problem> probs(static_cast<size_t>(count));
    const bool hist = job.mode == kModeSolve && o.record_history && job.history && count == 1;
    const size_t hist_n = hist ? history_entries(o, job.history_cap) : 0;
    size_t stage_bytes = 0;
    std::unordered_map<const void*, size_t> want;  // host ptr -> bytes
    auto note = [&](const double* p, size_t bytes) {
        size_t& w = want[p];
        w = std::max(w, bytes);
    };
    std::unordered_map<const void*, size_t> staged;
    std::vector<std::pair<const void*, size_t>> uploads;  // host ptr, bytes
    auto plan_upload = [&](const double* p) -> size_t {
        auto it = staged.find(p);
        if (it != staged.end()) return it->second;
        const size_t off = stage_bytes, bytes = want.at(p);
        stage_bytes += align_up(bytes);
        staged.emplace(p, off);
        uploads.emplace_back(p, bytes);
        return off;
    };
    std::vector<size_t> off_a(count), off_rhs(count, SIZE_MAX), off_in(count, SIZE_MAX), off_out(count);
    if (!job.device_ptrs) {
        for (int64_t i = 0; i < count; ++i) {
            const size_t bytes = static_cast<size_t>(job.n[i] * job.n[i]) * 8;
            note(job.a[i], bytes);
            if (job.rhs && job.rhs[i]) note(job.rhs[i], bytes);
            if (job.in && job.in[i]) note(job.in[i], bytes);
        }
        for (int64_t i = 0; i < count; ++i) {
            off_a[i] = plan_upload(job.a[i]);
            if (job.rhs && job.rhs[i]) off_rhs[i] = plan_upload(job.rhs[i]);
            if (job.in && job.in[i]) off_in[i] = plan_upload(job.in[i]);
        }
        for (int64_t i = 0; i < count; ++i) {
            off_out[i] = stage_bytes;
            stage_bytes += align_up(static_cast<size_t>(job.n[i] * job.n[i]) * 8);
        }
    }
    size_t hist_off = stage_bytes;
    if (hist && !job.device_ptrs) stage_bytes += align_up(hist_n * 8);
    HIP_TRY(ctx->staging.reserve(std::max<size_t>(stage_bytes, 256)), "staging allocation");
    char* stage = ctx->staging.as<char>();
    for (const auto& up : uploads)
        HIP_TRY(hipMemcpyAsync(stage + staged[up.first], up.first, up.second, hipMemcpyHostToDevice, ctx->stream),
                "upload");

    for (int64_t i = 0; i < count; ++i) {
        Problem& p = probs[i];
        p.n = static_cast<int32_t>(job.n[i]);
        p.pad = 0;
        p.f = job.f ? job.f[i] : 1.0;
        if (job.device_ptrs) {
            p.a = job.a[i];
            p.rhs = job.rhs ? job.rhs[i] : nullptr;
            p.in = job.in ? job.in[i] : nullptr;
            p.out = job.out[i];
            p.history = hist ? job.history : nullptr;
        } else {
            p.a = reinterpret_cast<const double*>(stage + off_a[i]);
            p.rhs = off_rhs[i] != SIZE_MAX ? reinterpret_cast<const double*>(stage + off_rhs[i]) : nullptr;
            p.in = off_in[i] != SIZE_MAX ? reinterpret_cast<const double*>(stage + off_in[i]) : nullptr;
            p.out = reinterpret_cast<double*>(stage + off_out[i]);
            p.history = hist ? reinterpret_cast<double*>(stage + hist_off) : nullptr;
        }
    }

    // Largest problems first, so the tail of the queue holds the short jobs.
    std::vector<int32_t> order(static_cast<size_t>(count));
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return job.n[x] > job.n[y]; });

    // A batch whose largest side is >= batch_wide_side runs on kWideBlock-
    // thread workgroups, one per CU, any other on kBlock-thread ones, four to
    // a CU.  One workgroup per problem either way, largest first: a large
    // problem on a narrow workgroup runs only as fast as its few waves'
    // memory requests in flight allow and ends the launch, while a batch of
    // equal mid-size problems is fastest four to a CU.
    const int wide = ctx->batch_wide_side > 0 && nmax >= ctx->batch_wide_side ? 1 : 0;
    int per_cu = 1;
    HIP_TRY(darcy_hip_batch_occupancy(mixed, wide, &per_cu), "occupancy query");
    per_cu = std::max(per_cu, 1);
    Hier scratch_h;
    const size_t slot_bytes = plan_slot(nullptr, nmax, mixed, &scratch_h, o.min_coarse_n, o.max_levels);
    int64_t grid = std::min<int64_t>(count, static_cast<int64_t>(ctx->cus) * per_cu);
    // Keep the workspace within a quarter of free device memory.
    size_t free_b = 0, total_b = 0;
    HIP_TRY(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
    const size_t budget = std::max<size_t>(free_b / 4 + ctx->slots.bytes, slot_bytes);
    grid = std::max<int64_t>(1, std::min<int64_t>(grid, static_cast<int64_t>(budget / slot_bytes)));

    HIP_TRY(ctx->slots.reserve(static_cast<size_t>(grid) * slot_bytes), "workspace allocation");
    HIP_TRY(ctx->probs.reserve(sizeof(Problem) * count), "problem table");
    HIP_TRY(ctx->stats.reserve(sizeof(ProblemStats) * count), "stats table");
    HIP_TRY(ctx->order.reserve(sizeof(int32_t) * count), "order table");
    HIP_TRY(ctx->counter.reserve(256), "counter");
    HIP_TRY(hipMemcpyAsync(ctx->probs.ptr, probs.data(), sizeof(Problem) * count, hipMemcpyHostToDevice, ctx->stream),
            "problem upload");
    HIP_TRY(hipMemcpyAsync(ctx->order.ptr, order.data(), sizeof(int32_t) * count, hipMemcpyHostToDevice, ctx->stream),
            "order upload");
    HIP_TRY(hipMemsetAsync(ctx->counter.ptr, 0, 256, ctx->stream), "counter reset");

    BatchArgs args{};
    args.probs = ctx->probs.as<Problem>();
    args.stats = ctx->stats.as<ProblemStats>();
    args.order = ctx->order.as<int32_t>();
    args.next = ctx->counter.as<uint32_t>();
    args.slots = ctx->slots.as<char>();
    args.slot_bytes = slot_bytes;
    args.count = static_cast<int32_t>(count);
    args.prm = solve_params(ctx, o, job.mode, mixed, static_cast<int32_t>(hist_n));
    if (wide) args.prm.step_cells *= kWideStepFactor;

    HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream), "event record");
    HIP_TRY(darcy_hip_launch_batch(&args, static_cast<int>(grid), mixed, wide, ctx->stream), "solve kernel launch");
    HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream), "event record");

    std::vector<ProblemStats> st(static_cast<size_t>(count));
    HIP_TRY(hipMemcpyAsync(st.data(), ctx->stats.ptr, sizeof(ProblemStats) * count, hipMemcpyDeviceToHost, ctx->stream),
            "stats readback");
    if (!job.device_ptrs) {
        for (int64_t i = 0; i < count; ++i)
            HIP_TRY(hipMemcpyAsync(job.out[i], stage + off_out[i], static_cast<size_t>(job.n[i] * job.n[i]) * 8,
                                   hipMemcpyDeviceToHost, ctx->stream),
                    "result readback");
    }
    HIP_TRY(hipStreamSynchronize(ctx->stream), "solve");
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1), "event timing");
    ctx->last_ms = ms;
    if (hist && !job.device_ptrs && st[0].status != kInvalid) {
        const int32_t len = std::min<int32_t>(st[0].history_len, job.history_cap);
        if (len > 0)
            HIP_TRY(hipMemcpy(job.history, stage + hist_off, static_cast<size_t>(len) * 8, hipMemcpyDeviceToHost),
                    "history readback");
    }

    return report(ctx, st, job.stats, job.history_cap);
}

// ------------------------------------------------------------ grid path

int blocks_for(const darcy_ctx* ctx, int64_t work, int per_block) {
    const int64_t b = (work + per_block - 1) / per_block;
    return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(b, static_cast<int64_t>(ctx->cus) * 8)));
}

// One multigrid solve on the grid path (solver.hpp:83-174 with the cycle of
// multigrid.hpp:97-127 recursed on the host over the large levels).
int run_grid(darcy_ctx* ctx, const Job& job) {
    const darcy_opts& o = *job.opts;
    // Jacobi and no preconditioner need level 0 only and run in FP64, as on
    // the persistent kernel.
    const bool mg = o.preconditioner == DARCY_PC_MULTIGRID;
    const bool mixed = mg && o.precision == DARCY_MIXED;
    const int n = static_cast<int>(job.n[0]);
    const int64_t N = static_cast<int64_t>(n) * n;
    const size_t cells = static_cast<size_t>(N) * 8;
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    hipStream_t s = ctx->stream;

    // Workspace: slot, device Hier, state, bad-index word, tail matrix,
    // history, staging of host inputs/outputs.
    Hier h;
    const size_t slot_bytes = plan_slot(nullptr, n, mixed, &h, o.min_coarse_n, o.max_levels);
    const bool hist = o.record_history && job.history;
    const size_t hist_n = hist ? history_entries(o, job.history_cap) : 0;
    size_t off = 0;
    auto take = [&](size_t b) {
        const size_t at = off;
        off += align_up(b);
        return at;
    };
    const size_t o_slot = take(slot_bytes), o_h = take(sizeof(Hier)), o_st = take(sizeof(GridState)),
                 o_bad = take(8), o_tail = take(64 * 64 * 8), o_big = take(256 * 256 * 8),
                 o_hist = take(hist ? hist_n * 8 : 8);
    const bool host = !job.device_ptrs;
    const size_t o_a = host ? take(cells) : 0, o_rhs = host && job.rhs && job.rhs[0] ? take(cells) : 0,
                 o_u0 = host && job.in && job.in[0] ? take(cells) : 0, o_u = host ? take(cells) : 0;
    HIP_TRY(ctx->grid.reserve(off), "workspace allocation");
    char* base = ctx->grid.as<char>();
    plan_slot(base + o_slot, n, mixed, &h, o.min_coarse_n, o.max_levels);
    Hier* hd = reinterpret_cast<Hier*>(base + o_h);
    GridState* st = reinterpret_cast<GridState*>(base + o_st);
    auto* bad = reinterpret_cast<unsigned long long*>(base + o_bad);
    float* tail = reinterpret_cast<float*>(base + o_tail);
    // The FP64 cycle's tail as a dense FP64 matrix (grid_fp64_dense_tail),
    // in the same workspace.
    double* tail64 = reinterpret_cast<double*>(base + o_tail);
    // The levels from the first of side <= 16 down as one dense matrix
    // (grid_dense_side 16), read by the tail from L2.
    void* big = base + o_big;
    double* history = hist ? (host ? reinterpret_cast<double*>(base + o_hist) : job.history) : nullptr;

    const double* a = job.a[0];
    const double* rhs = job.rhs ? job.rhs[0] : nullptr;
    const double* u0 = job.in ? job.in[0] : nullptr;
    double* u = job.out[0];
    if (host) {
        HIP_TRY(hipMemcpyAsync(base + o_a, a, cells, hipMemcpyHostToDevice, s), "upload");
        a = reinterpret_cast<const double*>(base + o_a);
        if (rhs) {
            HIP_TRY(hipMemcpyAsync(base + o_rhs, rhs, cells, hipMemcpyHostToDevice, s), "upload");
            rhs = reinterpret_cast<const double*>(base + o_rhs);
        }
        if (u0) {
            HIP_TRY(hipMemcpyAsync(base + o_u0, u0, cells, hipMemcpyHostToDevice, s), "upload");
            u0 = reinterpret_cast<const double*>(base + o_u0);
        }
        u = reinterpret_cast<double*>(base + o_u);
    }
    const unsigned long long none = ~0ull;
    HIP_TRY(hipMemcpyAsync(hd, &h, sizeof(Hier), hipMemcpyHostToDevice, s), "hierarchy table");
    HIP_TRY(hipMemcpyAsync(bad, &none, 8, hipMemcpyHostToDevice, s), "bad word");
    HIP_TRY(hipMemsetAsync(st, 0, sizeof(GridState), s), "state reset");

    HIP_TRY(hipEventRecord(ctx->ev0, s), "event record");
    // Hierarchy (level.hpp:79-97, :171-211).
    HIP_TRY(darcy_grid_fine(a, hd, bad, blocks_for(ctx, N, 1024), s), "fine level");
    for (int l = 0; mg && l + 1 < h.L; ++l) {
        const int64_t nc = static_cast<int64_t>(h.n[l + 1]) * h.n[l + 1];
        HIP_TRY(darcy_grid_galerkin(hd, l, blocks_for(ctx, nc, 256), s), "coarsening");
        if (mixed) HIP_TRY(darcy_grid_fp32(hd, l, blocks_for(ctx, static_cast<int64_t>(h.n[l]) * h.n[l], 1024), s),
                           "fp32 copy");
    }
    if (mixed) HIP_TRY(darcy_grid_fp32(hd, h.L - 1, 1, s), "fp32 copy");
    const int T = std::max(32, ctx->grid_tail_side);
    int lT = 0;
    while (lT < h.L - 1 && h.n[lT] > T) ++lT;
    const CycleOptsC co{o.pre_sweeps, o.post_sweeps, o.cycle_index};
    const bool dense64 = mg && !mixed && h.L > 1 && ctx->grid_fp64_dense_tail;
    int big_l = -1, big_ld = 0;
    if ((mixed || dense64) && ctx->grid_dense_side == 16) {
        for (int l = lT + 1; l < h.L; ++l)
            if (h.n[l] <= 16) {
                big_l = l;
                break;
            }
    }
    if (big_l > 0) {
        const int dim = h.n[big_l] * h.n[big_l], vec = mixed ? 4 : 2;  // elements per 16 bytes
        big_ld = (dim + vec - 1) / vec * vec;
        HIP_TRY(darcy_grid_big_matrix(mixed ? 1 : 0, hd, co, big_l, dim, big, big_ld, s), "tail matrix");
    } else {
        if (mixed && h.L > 1) HIP_TRY(darcy_grid_tail_matrix(hd, co, tail, s), "tail matrix");
        if (dense64) HIP_TRY(darcy_grid_tail_matrix64(hd, co, tail64, s), "tail matrix");
    }

    GridCg g{hd,    st,         a,          rhs,          job.f ? job.f[0] : 1.0,           u0, u, history,
             o.tol, o.max_iter, mixed ? 1 : 0, static_cast<int>(hist_n), o.breakdown_at};
    // CG kernels: the stencil apply on up to 8 blocks per CU; the streaming
    // rz and update kernels on grid_cg_per_cu (each wave takes several rounds
    // of two chunks).
    const int apply_blocks = blocks_for(ctx, chunks_of(N), 4);
    const int cg_blocks = std::min(blocks_for(ctx, chunks_of(N), 8), ctx->grid_cg_per_cu * std::max(1, ctx->cus));
    GridState hs{};
    auto read_state = [&]() -> int {
        HIP_TRY(hipMemcpyAsync(&hs, st, sizeof(GridState), hipMemcpyDeviceToHost, s), "state readback");
        HIP_TRY(hipStreamSynchronize(s), "solve");
        return DARCY_OK;
    };

    // The cycle: large levels phase by phase, the tail in one workgroup.
    int rc = DARCY_OK;
    // mode 1: from x = 0; 0: a repeated visit; 2 (the tail only): both
    // visits of the W-cycle in one launch.
    std::function<void(int, int)> cycle = [&](int l, int mode) {
        if (rc) return;
        if (l >= lT) {
            const hipError_t e = darcy_grid_tail(mixed, hd, l, mode, co, tail, dense64 ? tail64 : nullptr,
                                                 big_l > 0 ? big : nullptr, big_l, big_ld, s);
            if (e != hipSuccess) rc = hip_fail(ctx, e, "tail cycle");
            return;
        }
        const bool fresh = mode == 1;
        const int64_t nl = static_cast<int64_t>(h.n[l]) * h.n[l];
        // 512 cells per workgroup: the small levels are latency-bound, and
        // the pre-smoothing walk takes as many rows per lane as the launch
        // leaves it (each a memory round trip in turn), so the levels below
        // 2048^2 get one row per lane and more CUs; the largest are capped.
        const int bl = blocks_for(ctx, nl, 512), bc = blocks_for(ctx, nl / 4 + 1, 256), bw = bl;
        hipError_t e = hipSuccess;
        for (int sw = 0; sw < o.pre_sweeps && e == hipSuccess; ++sw) {
            if (sw == 0 && fresh) {
                e = darcy_grid_presmooth(mixed, hd, l, 0, bw, s);
            } else {
                e = darcy_grid_presmooth(mixed, hd, l, 1, bl, s);
                if (e == hipSuccess) e = darcy_grid_presmooth(mixed, hd, l, 2, bl, s);
            }
        }
        if (e == hipSuccess) e = darcy_grid_restrict(mixed, hd, l, bc, s);
        if (e != hipSuccess) {
            rc = hip_fail(ctx, e, "cycle launch");
            return;
        }
        if (o.cycle_index == 2 && l + 1 >= lT) {
            cycle(l + 1, 2);
        } else {
            cycle(l + 1, 1);
            if (o.cycle_index == 2) cycle(l + 1, 0);
        }
        if (rc) return;
        e = darcy_grid_postsmooth(mixed, hd, l, 0, bl, s);
        if (e == hipSuccess) e = darcy_grid_postsmooth(mixed, hd, l, 1, bl, s);
        for (int sw = 1; sw < o.post_sweeps && e == hipSuccess; ++sw) {
            e = darcy_grid_postsmooth(mixed, hd, l, 2, bl, s);
            if (e == hipSuccess) e = darcy_grid_postsmooth(mixed, hd, l, 1, bl, s);
        }
        if (e != hipSuccess) rc = hip_fail(ctx, e, "cycle launch");
    };

    // One preconditioner application and the rz kernel, launched eagerly or
    // replayed from a graph captured on first use (the sequence depends only
    // on the hierarchy, the options and the workspace address).
    auto precondition = [&]() -> int {
        if (!mg) {
            HIP_TRY(darcy_grid_cg(o.preconditioner == DARCY_PC_JACOBI ? 4 : 5, &g, nullptr, nullptr,
                                  blocks_for(ctx, N, 1024), s),
                    "diagonal preconditioner");
            HIP_TRY(darcy_grid_cg(1, &g, nullptr, nullptr, cg_blocks, s), "rz");
            return DARCY_OK;
        }
        if (!ctx->grid_graph || lT == 0) {
            cycle(0, 1);
            if (rc) return rc;
            HIP_TRY(darcy_grid_cg(1, &g, nullptr, nullptr, cg_blocks, s), "rz");
            return DARCY_OK;
        }
        const int key[9] = {n, mixed ? 1 : 0, o.pre_sweeps, o.post_sweeps, o.cycle_index, lT, cg_blocks, dense64 ? 1 : 0,
                            big_l};
        const bool hit = ctx->cycle_exec && ctx->cycle_base == base && std::equal(key, key + 9, ctx->cycle_key);
        if (!hit) {
            if (!ctx->graph_checked) {  // every captured kernel within the device's launch limits
                const char* what = "";
                const hipError_t ce = darcy_grid_check_kernels(ctx->device, &what);
                if (ce != hipSuccess) return hip_fail(ctx, ce, (std::string("graph kernel limits: ") + what).c_str());
                ctx->graph_checked = true;
            }
            if (ctx->cycle_exec) (void)hipGraphExecDestroy(ctx->cycle_exec);
            if (ctx->cycle_graph) (void)hipGraphDestroy(ctx->cycle_graph);
            ctx->cycle_exec = nullptr;
            ctx->cycle_graph = nullptr;
            HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "graph capture");
            cycle(0, 1);
            const hipError_t e1 = rc ? hipErrorUnknown : darcy_grid_cg(1, &g, nullptr, nullptr, cg_blocks, s);
            hipGraph_t graph = nullptr;
            const hipError_t e2 = hipStreamEndCapture(s, &graph);
            if (rc) return rc;
            HIP_TRY(e1, "rz");
            HIP_TRY(e2, "graph capture");
            ctx->cycle_graph = graph;
            HIP_TRY(hipGraphInstantiate(&ctx->cycle_exec, graph, nullptr, nullptr, 0), "graph instantiate");
            ctx->cycle_base = base;
            std::copy(key, key + 9, ctx->cycle_key);
        }
        HIP_TRY(hipGraphLaunch(ctx->cycle_exec, s), "graph launch");
        return DARCY_OK;
    };

    HIP_TRY(darcy_grid_cg(0, &g, nullptr, nullptr, cg_blocks, s), "init");
    if (const int r = read_state()) return r;
    unsigned long long bad_h = 0;
    HIP_TRY(hipMemcpy(&bad_h, bad, 8, hipMemcpyDeviceToHost), "bad word");
    if (bad_h != ~0ull) {
        if (job.stats) {
            darcy_stats& d = job.stats[0];
            d = darcy_stats{0, 0, 1, kInvalid, 0.0, static_cast<int64_t>(bad_h), 0, 0};
        }
        char msg[160];
        std::snprintf(msg, sizeof msg,
                      "permeability must be finite and strictly positive everywhere; offending value at flat index %llu",
                      bad_h);
        return fail(ctx, DARCY_INVALID_INPUT, msg);
    }
    const bool zero_rhs = hs.status == kConverged && hs.bnorm == 0.0;
    double* pold = h.p;
    double* pnew = h.u;
    if (hs.status == kRunning) {
        if (const int r = precondition()) return r;
        for (;;) {
            HIP_TRY(darcy_grid_cg(2, &g, pold, pnew, apply_blocks, s), "apply");
            HIP_TRY(darcy_grid_cg(3, &g, pnew, nullptr, cg_blocks, s), "update");
            if (const int r = read_state()) return r;
            if (hs.status != kRunning) break;
            if (const int r = precondition()) return r;
            std::swap(pold, pnew);
        }
    }
    if (zero_rhs) HIP_TRY(hipMemsetAsync(u, 0, cells, s), "zero solution");
    HIP_TRY(hipEventRecord(ctx->ev1, s), "event record");
    if (host) HIP_TRY(hipMemcpyAsync(job.out[0], u, cells, hipMemcpyDeviceToHost, s), "result readback");
    HIP_TRY(hipStreamSynchronize(s), "solve");
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1), "event timing");
    ctx->last_ms = ms;
    // A breakdown reports the iterations completed before it (solver.hpp:149-155).
    const int iters = hs.it;
    const int32_t hist_len = hist && !zero_rhs ? std::min<int32_t>(iters + 1, static_cast<int32_t>(hist_n)) : 0;
    if (hist && host && hist_len > 0)
        HIP_TRY(hipMemcpy(job.history, history, static_cast<size_t>(hist_len) * 8, hipMemcpyDeviceToHost),
                "history readback");
    if (job.stats) {
        darcy_stats& d = job.stats[0];
        d.iterations = iters;
        d.converged = hs.status == kConverged ? 1 : 0;
        d.levels = zero_rhs || !mg ? 1 : h.L;
        d.status = hs.status;
        d.residual = zero_rhs ? 0.0 : hs.res;
        d.bad_index = -1;
        d.history_len = hist_len;
        d.reserved = 0;
    }
    ctx->error.clear();
    return DARCY_OK;
}

}  // namespace

extern "C" void darcy_default_opts(darcy_opts* o) {
    if (!o) return;
    o->tol = 1e-10;
    o->max_iter = 10000;
    o->preconditioner = DARCY_PC_MULTIGRID;
    o->pre_sweeps = 1;
    o->post_sweeps = 1;
    o->cycle_index = 2;
    o->record_history = 0;
    o->precision = DARCY_FP64;
    o->device_pointers = 0;
    o->min_coarse_n = 1;
    o->max_levels = 32;
    o->breakdown_at = 0;
}

extern "C" int darcy_ctx_create(int device, darcy_ctx** out) {
    if (!out) return DARCY_INVALID_INPUT;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return DARCY_HIP_ERROR;
    if (device < 0 || device >= ndev) return DARCY_INVALID_INPUT;
    auto* c = new darcy_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    if (hipSetDevice(device) != hipSuccess || hipGetDeviceProperties(&prop, device) != hipSuccess ||
        hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        darcy_ctx_destroy(c);
        return DARCY_HIP_ERROR;
    }
    c->cus = prop.multiProcessorCount;
    c->arch = prop.gcnArchName;
    // Once per context, with its device current: the scratch array's device
    // address (an error is kept and reported by the entry point that needs it).
    c->scratch_status = darcy_hip_scratch(&c->scratch, &c->scratch_bytes);
    if (c->scratch_status != hipSuccess) c->scratch = nullptr;
    *out = c;
    return DARCY_OK;
}

extern "C" void darcy_ctx_destroy(darcy_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->cycle_exec) (void)hipGraphExecDestroy(c->cycle_exec);
    if (c->cycle_graph) (void)hipGraphDestroy(c->cycle_graph);
    for (DevBuf* b : {&c->slots, &c->probs, &c->stats, &c->order, &c->counter, &c->staging, &c->grid, &c->fields})
        b->release();
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" const char* darcy_last_error(const darcy_ctx* c) { return c ? c->error.c_str() : "null context"; }

extern "C" int darcy_solve(darcy_ctx* ctx, const double* a, const double* rhs, double f, int64_t n,
                           const darcy_opts* opts, const double* u0, double* u_out, darcy_stats* stats,
                           double* history, int32_t history_cap) {
    if (!ctx) return DARCY_INVALID_INPUT;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (const int rc = check_problem(ctx, n, a, u_out)) return rc;
    darcy_opts o;
    darcy_default_opts(&o);
    if (opts) o = *opts;
    if (const int rc = check_opts(ctx, o)) return rc;
    Job job{kModeSolve, &o, 1, &n, &a, &rhs, &f, &u0, &u_out, stats, history, history ? history_cap : 0,
            o.device_pointers != 0};
    if (n > ctx->grid_min_n) return run_grid(ctx, job);
    return run(ctx, job);
}

extern "C" int darcy_solve_batch(darcy_ctx* ctx, int64_t count, const int64_t* n, const double* const* a,
                                 const double* const* rhs, const double* f, const darcy_opts* opts,
                                 const double* const* u0, double* const* u_out, darcy_stats* stats) {
    if (!ctx) return DARCY_INVALID_INPUT;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (count < 0 || (count > 0 && (!n || !a || !u_out)))
        return fail(ctx, DARCY_INVALID_INPUT, "a must have shape (batch, n, n)");
    if (count > INT32_MAX) return fail(ctx, DARCY_INVALID_INPUT, "batch too large");
    darcy_opts o;
    darcy_default_opts(&o);
    if (opts) o = *opts;
    if (const int rc = check_opts(ctx, o)) return rc;
    Job job{kModeSolve, &o, count, n, a, rhs, f, u0, u_out, stats, nullptr, 0, o.device_pointers != 0};
    return run(ctx, job);
}

extern "C" int darcy_apply_operator(darcy_ctx* ctx, const double* a, int64_t n, const double* x, double* y,
                                    int32_t device_pointers) {
    if (!ctx) return DARCY_INVALID_INPUT;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (!x) return fail(ctx, DARCY_INVALID_INPUT, "x must have the same shape as a");
    darcy_opts o;
    darcy_default_opts(&o);
    o.device_pointers = device_pointers;
    Job job{kModeApplyOperator, &o, 1, &n, &a, nullptr, nullptr, &x, &y, nullptr, nullptr, 0, device_pointers != 0};
    return run(ctx, job);
}

extern "C" int darcy_apply_preconditioner(darcy_ctx* ctx, const double* a, int64_t n, const double* r, double* z,
                                          int32_t pre_sweeps, int32_t post_sweeps, int32_t cycle_index,
                                          int32_t precision, int32_t device_pointers) {
    if (!ctx) return DARCY_INVALID_INPUT;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (!r) return fail(ctx, DARCY_INVALID_INPUT, "r must have the same shape as a");
    darcy_opts o;
    darcy_default_opts(&o);
    o.pre_sweeps = pre_sweeps;
    o.post_sweeps = post_sweeps;
    o.cycle_index = cycle_index;
    o.precision = precision;
    o.device_pointers = device_pointers;
    if (pre_sweeps < 1 || post_sweeps < 1) return fail(ctx, DARCY_INVALID_INPUT, "sweeps must be at least 1");
    if (cycle_index < 1 || cycle_index > 2) return fail(ctx, DARCY_INVALID_INPUT, "cycle_index must be 1 or 2");
    if (const int rc = check_opts(ctx, o)) return rc;
    Job job{kModeApplyPreconditioner, &o, 1, &n, &a, nullptr, nullptr, &r, &z, nullptr, nullptr, 0,
            device_pointers != 0};
    return run(ctx, job);
}

extern "C" int darcy_apply_preconditioner_opts(darcy_ctx* ctx, const double* a, int64_t n, const double* r,
                                               double* z, const darcy_opts* opts) {
    if (!ctx) return DARCY_INVALID_INPUT;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (!r) return fail(ctx, DARCY_INVALID_INPUT, "r must have the same shape as a");
    darcy_opts o;
    darcy_default_opts(&o);
    if (opts) o = *opts;
    o.preconditioner = DARCY_PC_MULTIGRID;
    if (const int rc = check_opts(ctx, o)) return rc;
    Job job{kModeApplyPreconditioner, &o, 1, &n, &a, nullptr, nullptr, &r, &z, nullptr, nullptr, 0,
            o.device_pointers != 0};
    return run(ctx, job);
}

extern "C" int32_t darcy_num_levels(int64_t n) {
    if (n < 1) return 0;
    return level_count(static_cast<int>(std::min<int64_t>(n, INT32_MAX)));
}

extern "C" int darcy_device_alloc(darcy_ctx* ctx, size_t bytes, void** out) {
    if (!ctx || !out) return DARCY_INVALID_INPUT;
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    HIP_TRY(hipMalloc(out, bytes), "hipMalloc");
    return DARCY_OK;
}

extern "C" int darcy_device_free(darcy_ctx* ctx, void* p) {
    if (!ctx) return DARCY_INVALID_INPUT;
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    HIP_TRY(hipFree(p), "hipFree");
    return DARCY_OK;
}

extern "C" int darcy_copy_to_device(darcy_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx) return DARCY_INVALID_INPUT;
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
    return DARCY_OK;
}

extern "C" int darcy_copy_to_host(darcy_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx) return DARCY_INVALID_INPUT;
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H");
    return DARCY_OK;
}

extern "C" int darcy_last_kernel_ms(const darcy_ctx* ctx, double* ms) {
    if (!ctx || !ms) return DARCY_INVALID_INPUT;
    *ms = ctx->last_ms;
    return DARCY_OK;
}

extern "C" int darcy_device_info(const darcy_ctx* ctx, char* buf, int32_t cap) {
    if (!ctx || !buf || cap <= 0) return DARCY_INVALID_INPUT;
    int rt = 0;
    (void)hipRuntimeGetVersion(&rt);
    std::snprintf(buf, static_cast<size_t>(cap), "%s, %d CUs, HIP runtime %d", ctx->arch.c_str(), ctx->cus, rt);
    return DARCY_OK;
}

extern "C" int darcy_bandwidth_probe(darcy_ctx* ctx, size_t bytes, int32_t repeats, double* gbps) {
    if (!ctx || !gbps || repeats < 1 || bytes < 16) return DARCY_INVALID_INPUT;
    std::lock_guard<std::mutex> lock(ctx->mu);
    bytes &= ~static_cast<size_t>(15);
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    void* src = nullptr;
    void* dst = nullptr;
    HIP_TRY(hipMalloc(&src, bytes), "probe allocation");
    const hipError_t e = hipMalloc(&dst, bytes);
    if (e != hipSuccess) {
        (void)hipFree(src);
        return hip_fail(ctx, e, "probe allocation");
    }
    const int grid = ctx->cus * 8;
    hipError_t err = hipMemsetAsync(src, 0, bytes, ctx->stream);
    if (err == hipSuccess) err = darcy_hip_launch_copy(src, dst, bytes, grid, ctx->stream);  // warm-up
    if (err == hipSuccess) err = hipEventRecord(ctx->ev0, ctx->stream);
    for (int32_t r = 0; r < repeats && err == hipSuccess; ++r) err = darcy_hip_launch_copy(src, dst, bytes, grid, ctx->stream);
    if (err == hipSuccess) err = hipEventRecord(ctx->ev1, ctx->stream);
    if (err == hipSuccess) err = hipStreamSynchronize(ctx->stream);
    float ms = 0.f;
    if (err == hipSuccess) err = hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    (void)hipFree(src);
    (void)hipFree(dst);
    if (err != hipSuccess) return hip_fail(ctx, err, "bandwidth probe");
    *gbps = 2.0 * static_cast<double>(bytes) * repeats / (ms * 1e-3) / 1e9;
    return DARCY_OK;
}

extern "C" int darcy_ctx_set_tuning(darcy_ctx* ctx, const char* key, int32_t value) {
    if (!ctx || !key) return DARCY_INVALID_INPUT;
    std::lock_guard<std::mutex> lock(ctx->mu);
    const std::string k = key;
    if (k == "grid_min_n" && value >= 1) {
        ctx->grid_min_n = value;
    } else if (k == "grid_tail_side" && value >= 32) {
        ctx->grid_tail_side = value;
    } else if (k == "grid_graph" && (value == 0 || value == 1)) {
        ctx->grid_graph = value;
    } else if (k == "grid_cg_per_cu" && value >= 1 && value <= 8) {
        ctx->grid_cg_per_cu = value;
    } else if (k == "batch_step_cells" && value >= 128 && value <= 16384) {
        ctx->batch_step_cells = value;
    } else if (k == "batch_step_cells_fp32" && value >= 128 && value <= 16384) {
        ctx->batch_step_cells_fp32 = value;
    } else if (k == "grid_fp64_dense_tail" && (value == 0 || value == 1)) {
        ctx->grid_fp64_dense_tail = value;
    } else if (k == "grid_dense_side" && (value == 8 || value == 16)) {
        ctx->grid_dense_side = value;
    } else if (k == "batch_wide_side" && (value == 0 || (value >= 64 && value <= 32768))) {
        ctx->batch_wide_side = value;
    } else {
        return fail(ctx, DARCY_INVALID_INPUT, "unknown tuning key or value out of range: " + k);
    }
    return DARCY_OK;
}

extern "C" int darcy_ctx_get_tuning(const darcy_ctx* ctx, const char* key, int32_t* value) {
    if (!ctx || !key || !value) return DARCY_INVALID_INPUT;
    const std::string k = key;
    if (k == "grid_min_n")
        *value = ctx->grid_min_n;
    else if (k == "grid_tail_side")
        *value = ctx->grid_tail_side;
    else if (k == "grid_graph")
        *value = ctx->grid_graph;
    else if (k == "grid_cg_per_cu")
        *value = ctx->grid_cg_per_cu;
    else if (k == "batch_step_cells")
        *value = ctx->batch_step_cells;
    else if (k == "batch_step_cells_fp32")
        *value = ctx->batch_step_cells_fp32;
    else if (k == "grid_fp64_dense_tail")
        *value = ctx->grid_fp64_dense_tail;
    else if (k == "grid_dense_side")
        *value = ctx->grid_dense_side;
    else if (k == "batch_wide_side")
        *value = ctx->batch_wide_side;
    else
        return DARCY_INVALID_INPUT;
    return DARCY_OK;
}

extern "C" int darcy_grf_fields(darcy_ctx* ctx, int64_t count, int64_t n, double alpha, double tau, const double* xi,
                                double hi, double lo, double* grf_out, double* a_out, int32_t device_pointers) {
    if (!ctx) return DARCY_INVALID_INPUT;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (n < 1) return fail(ctx, DARCY_INVALID_INPUT, "n must be positive");
    if (n > kMaxSide) return fail(ctx, DARCY_INVALID_INPUT, "grid size exceeds 32768");
    if (count < 0) return fail(ctx, DARCY_INVALID_INPUT, "count must not be negative");
    if (!(std::isfinite(alpha) && std::isfinite(tau) && tau > 0.0))
        return fail(ctx, DARCY_INVALID_INPUT, "alpha must be finite and tau finite and positive");
    if (a_out && !(hi > 0.0 && lo > 0.0))
        return fail(ctx, DARCY_INVALID_INPUT, "permeability phases must be strictly positive");
    if (count > 0 && (!xi || (!grf_out && !a_out))) return fail(ctx, DARCY_INVALID_INPUT, "null array argument");
    if (count == 0) return DARCY_OK;
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    hipStream_t s = ctx->stream;
    const size_t plane = static_cast<size_t>(n) * static_cast<size_t>(n), pb = plane * 8;
    // Samples per launch: the whole call with device pointers; otherwise
    // chunks of about 2^27 cells (1 GiB per staged array), at most 4096.
    const bool dev = device_pointers != 0;
    int64_t chunk = dev ? count : std::max<int64_t>(1, static_cast<int64_t>((size_t{1} << 27) / plane));
    chunk = std::min<int64_t>(std::min<int64_t>(chunk, count), 4096);
    const size_t per = static_cast<size_t>(chunk) * pb;
    const int nout = (grf_out ? 1 : 0) + (a_out ? 1 : 0);
    const size_t o_c = 0, o_amp = align_up(pb), o_t = o_amp + align_up(pb), o_x = o_t + align_up(per),
                 o_g = o_x + (dev ? 0 : align_up(per)), total = o_g + (dev ? 0 : nout * align_up(per));
    HIP_TRY(ctx->fields.reserve(total), "workspace allocation");
    char* base = ctx->fields.as<char>();
    double* cmat = reinterpret_cast<double*>(base + o_c);
    double* amp = reinterpret_cast<double*>(base + o_amp);
    double* t = reinterpret_cast<double*>(base + o_t);
    HIP_TRY(hipEventRecord(ctx->ev0, s), "event record");
    for (int64_t b0 = 0; b0 < count; b0 += chunk) {
        const int64_t cnt = std::min<int64_t>(chunk, count - b0);
        const size_t off = static_cast<size_t>(b0) * plane, bytes = static_cast<size_t>(cnt) * pb;
        const double* x = xi + off;
        double* g = grf_out ? grf_out + off : nullptr;
        double* a = a_out ? a_out + off : nullptr;
        if (!dev) {
            double* xs = reinterpret_cast<double*>(base + o_x);
            HIP_TRY(hipMemcpyAsync(xs, x, bytes, hipMemcpyHostToDevice, s), "upload");
            x = xs;
            double* out = reinterpret_cast<double*>(base + o_g);
            g = grf_out ? out : nullptr;
            a = a_out ? out + (grf_out ? per / 8 : 0) : nullptr;
        }
        HIP_TRY(darcy_hip_launch_grf(static_cast<int>(n), static_cast<int>(cnt), alpha, tau, x, cmat, amp, t, g, a,
                                     hi, lo, s),
                "field kernels");
        if (!dev) {
            if (grf_out) HIP_TRY(hipMemcpyAsync(grf_out + off, g, bytes, hipMemcpyDeviceToHost, s), "readback");
            if (a_out) HIP_TRY(hipMemcpyAsync(a_out + off, a, bytes, hipMemcpyDeviceToHost, s), "readback");
            HIP_TRY(hipStreamSynchronize(s), "fields");
        }
    }
    HIP_TRY(hipEventRecord(ctx->ev1, s), "event record");
    HIP_TRY(hipStreamSynchronize(s), "fields");
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1), "event timing");
    ctx->last_ms = ms;
    ctx->error.clear();
    return DARCY_OK;
}

// darcy_solve_batch on the library's static scratch array: no device
// allocation, no free.  The caller's arrays are device pointers.  The tables
// (problems, stats, order, counter) and then as many solve slots as fit take
// the scratch in that order; a batch whose first slot does not fit is
// rejected.
extern "C" int darcy_solve_batch_noalloc(darcy_ctx* ctx, int64_t count, const int64_t* n, const double* const* a,
                                         const double* const* rhs, const double* f, const darcy_opts* opts,
                                         const double* const* u0, double* const* u_out, darcy_stats* stats) {
    if (!ctx) return DARCY_INVALID_INPUT;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (count < 1 || !n || !a || !u_out) return fail(ctx, DARCY_INVALID_INPUT, "a must have shape (batch, n, n)");
    if (count > INT32_MAX) return fail(ctx, DARCY_INVALID_INPUT, "batch too large");
    darcy_opts o;
    darcy_default_opts(&o);
    if (opts) o = *opts;
    if (!o.device_pointers)
        return fail(ctx, DARCY_INVALID_INPUT, "the allocation-free entry point takes device pointers only");
    o.record_history = 0;
    if (const int rc = check_opts(ctx, o)) return rc;
    int nmax = 0;
    for (int64_t i = 0; i < count; ++i) {
        if (const int rc = check_problem(ctx, n[i], a[i], u_out[i])) return rc;
        nmax = std::max<int>(nmax, static_cast<int>(n[i]));
    }
    if (!ctx->scratch) return hip_fail(ctx, ctx->scratch_status, "scratch address lookup");
    const bool mixed = o.preconditioner == DARCY_PC_MULTIGRID && o.precision == DARCY_MIXED;
    Hier sizing;
    const size_t slot_bytes = plan_slot(nullptr, nmax, mixed, &sizing, o.min_coarse_n, o.max_levels);
    const size_t c = static_cast<size_t>(count);
    const size_t o_probs = 0, o_stats = align_up(sizeof(Problem) * c), o_order = o_stats + align_up(sizeof(ProblemStats) * c),
                 o_counter = o_order + align_up(sizeof(int32_t) * c), o_slots = o_counter + kAlign;
    if (o_slots + slot_bytes > ctx->scratch_bytes) {
        char msg[200];
        std::snprintf(msg, sizeof msg,
                      "batch exceeds the allocation-free scratch (%zu bytes): %zu bytes of tables and %zu bytes per "
                      "solve slot",
                      ctx->scratch_bytes, o_slots, slot_bytes);
        return fail(ctx, DARCY_INVALID_INPUT, msg);
    }
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    char* base = static_cast<char*>(ctx->scratch);
    std::vector<Problem> probs(c);
    for (size_t i = 0; i < c; ++i) {
        Problem& p = probs[i];
        p.n = static_cast<int32_t>(n[i]);
        p.pad = 0;
        p.a = a[i];
        p.rhs = rhs ? rhs[i] : nullptr;
        p.f = f ? f[i] : 1.0;
        p.in = u0 ? u0[i] : nullptr;
        p.out = u_out[i];
        p.history = nullptr;
    }
    std::vector<int32_t> order(c);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return n[x] > n[y]; });
    // The workgroup width as in darcy_solve_batch (batch_wide_side).
    const int wide = ctx->batch_wide_side > 0 && nmax >= ctx->batch_wide_side ? 1 : 0;
    int per_cu = 1;
    HIP_TRY(darcy_hip_batch_occupancy(mixed, wide, &per_cu), "occupancy query");
    const int64_t fit = static_cast<int64_t>((ctx->scratch_bytes - o_slots) / slot_bytes);
    const int64_t grid = std::max<int64_t>(
        1, std::min<int64_t>({count, static_cast<int64_t>(ctx->cus) * std::max(per_cu, 1), fit}));

    hipStream_t s = ctx->stream;
    HIP_TRY(hipMemcpyAsync(base + o_probs, probs.data(), sizeof(Problem) * c, hipMemcpyHostToDevice, s),
            "problem upload");
    HIP_TRY(hipMemcpyAsync(base + o_order, order.data(), sizeof(int32_t) * c, hipMemcpyHostToDevice, s),
            "order upload");
    HIP_TRY(darcy_hip_zero_words(reinterpret_cast<uint32_t*>(base + o_counter), 1, s), "counter reset");
    BatchArgs args{};
    args.probs = reinterpret_cast<const Problem*>(base + o_probs);
    args.stats = reinterpret_cast<ProblemStats*>(base + o_stats);
    args.order = reinterpret_cast<const int32_t*>(base + o_order);
    args.next = reinterpret_cast<uint32_t*>(base + o_counter);
    args.slots = base + o_slots;
    args.slot_bytes = slot_bytes;
    args.count = static_cast<int32_t>(count);
    args.prm = solve_params(ctx, o, kModeSolve, mixed, 0);
    if (wide) args.prm.step_cells *= kWideStepFactor;
    HIP_TRY(hipEventRecord(ctx->ev0, s), "event record");
    HIP_TRY(darcy_hip_launch_batch(&args, static_cast<int>(grid), mixed, wide, s), "solve kernel launch");
    HIP_TRY(hipEventRecord(ctx->ev1, s), "event record");
    std::vector<ProblemStats> st(c);
    HIP_TRY(hipMemcpyAsync(st.data(), base + o_stats, sizeof(ProblemStats) * c, hipMemcpyDeviceToHost, s),
            "stats readback");
    HIP_TRY(hipStreamSynchronize(s), "solve");
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1), "event timing");
    ctx->last_ms = ms;
    return report(ctx, st, stats, 0);
}

extern "C" int darcy_scratch_bytes(const darcy_ctx* ctx, size_t* bytes) {
    if (!ctx || !bytes) return DARCY_INVALID_INPUT;
    *bytes = ctx->scratch ? ctx->scratch_bytes : 0;
    return DARCY_OK;
}
