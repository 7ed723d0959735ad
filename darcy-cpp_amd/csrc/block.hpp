// Workgroup-level building blocks shared by the persistent batch kernel
// (batch.hip) and the coarse-tail kernel of the grid path (grid.hip):
// canonical reductions, the LDS-resident coarse levels, the dense FP32 tail,
// and the W/V cycle run by one workgroup (multigrid.hpp:97-127 of the
// reference).
#pragma once

#include "abi_types.hpp"
#include "setup.hpp"

namespace darcy_hip {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / kWave;
// The persistent kernel's workgroup for large problems: a whole CU's wave
// slots (16 waves at the 128-VGPR budget of four kBlock workgroups).
constexpr int kWideBlock = 1024;

// Largest side of the levels kept in LDS by a B-thread batch workgroup: 16
// for kBlock (14 KB in FP64, 7.5 KB in FP32; the FP32 kernel also keeps the
// 16 KB tail matrix), so four workgroups fit on a CU; for kWideBlock, alone
// on its CU, 64 in FP32 (about 111 KB with the tail matrix) and 32 in FP64
// (57 KB: a 64^2 FP64 level alone would need 163 KB).
template <class T, int B = kBlock>
struct LdsSide {
    static constexpr int value = B == kBlock ? 16 : (sizeof(T) == 4 ? 64 : 32);
};

// LDS elements of the coarse-level pool.  The FP32 pool doubles as FP64
// scratch while the tail matrix is built: FP64 operators of the tail levels
// (<= 285 doubles) and x, b for one unit vector per wave (4 x 170 doubles).
constexpr int kTailScratchDoubles = 285 + 4 * 170;
template <class T, int B = kBlock>
constexpr int pool_elems() {
    constexpr int need = LdsPool<T, LdsSide<T, B>::value>::kElems;
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

// ----------------------------------------------- register-resident wave levels
// The FP64 levels of side <= 8 on one wave, one cell per lane: lane q works
// on cell (q / n, q % n) of each such level.  A visit loads its cell's x, b,
// faces and reciprocal diagonal from the LDS pool once, runs its phases in
// registers, reading neighbours and children with cross-lane permutes, and
// stores x (and the next level's b) back: no barrier and no LDS round trip
// per phase.  Each per-cell expression has the operations and order of
// gs_at, apply_at, restricted_residual and black_with_prolong, so the cycle
// gives the bits of the LDS version.  All 64 lanes execute every permute;
// results are selected afterwards.  LDS accesses of one wave complete in
// order, so a lane reads what another lane of the wave stored before.

constexpr int kRegLevels = 4;  // sides 8, 4, 2, 1 at most

// Read-only description of one level for this lane.
struct RegLevel {
    LvP<double> A;  // the level in the LDS pool
    int i, j;       // this lane's cell
    bool on;        // lane < n * n
};

struct RegCell {
    double e, w, nn, s, inv, b;
    bool hE, hW, hN, hS;
};

DARCY_INLINE RegCell reg_cell(const RegLevel& L) {
    const int n = L.A.n, i = L.i, j = L.j, f = i * (n + 1) + j;
    RegCell c{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, false, false, false, false};
    if (L.on) {
        c.e = L.A.ex[f + 1];
        c.w = L.A.ex[f];
        c.nn = L.A.ny[(i + 1) * n + j];
        c.s = L.A.ny[i * n + j];
        c.inv = L.A.id[i * n + j];
        c.b = L.A.b[i * n + j];
        c.hE = j + 1 < n;
        c.hW = j > 0;
        c.hN = i + 1 < n;
        c.hS = i > 0;
    }
    return c;
}

DARCY_INLINE double lane_value(double v, int src) { return __shfl(v, src, kWave); }

// Lane exchange within a row of 16 lanes (DPP row shift, no LDS): lane l
// reads lane l + k (row_shl:k) or l - k (row_shr:k); 0 from outside the row.
template <int CTRL>
DARCY_INLINE double row_value(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}
constexpr int kRowShl1 = 0x101, kRowShr1 = 0x111, kRowShl4 = 0x104, kRowShr4 = 0x114;

struct RegNeighbours {
    double xE, xW, xN, xS;
};

DARCY_INLINE RegNeighbours reg_neighbours(const RegLevel& L, const RegCell& c, double v) {
    const int q = static_cast<int>(threadIdx.x & (kWave - 1)), n = L.A.n;
    // The 4 x 4 level fills lanes 0..15, one DPP row: east / west are lanes
    // q +- 1 and north / south q +- 4, read by row shifts.  Where a neighbour
    // is off the grid the lane read is another cell (or 0), and reg_gs /
    // reg_residual select its product away, as for the clamped permute.
    if (n == 4) return {row_value<kRowShl1>(v), row_value<kRowShr1>(v), row_value<kRowShl4>(v), row_value<kRowShr4>(v)};
    return {lane_value(v, c.hE ? q + 1 : q), lane_value(v, c.hW ? q - 1 : q), lane_value(v, c.hN ? q + n : q),
            lane_value(v, c.hS ? q - n : q)};
}

// gs_at with the neighbour values gathered.
DARCY_INLINE double reg_gs(const RegCell& c, const RegNeighbours& x) {
    double v = c.b;
    v += c.hE ? c.e * x.xE : 0.0;
    v += c.hW ? c.w * x.xW : 0.0;
    v += c.hN ? c.nn * x.xN : 0.0;
    v += c.hS ? c.s * x.xS : 0.0;
    return v * c.inv;
}

// b - A x at this lane's cell (apply_at with the neighbour values gathered).
DARCY_INLINE double reg_residual(const RegCell& c, double xk, const RegNeighbours& x) {
    const double d = ((c.e + c.w) + c.nn) + c.s;
    double v = d * xk;
    v -= c.hE ? c.e * x.xE : 0.0;
    v -= c.hW ? c.w * x.xW : 0.0;
    v -= c.hN ? c.nn * x.xN : 0.0;
    v -= c.hS ? c.s * x.xS : 0.0;
    return c.b - v;
}

DARCY_INLINE bool reg_red(const RegLevel& L) { return ((L.i + L.j) & 1) == 0; }

// One colour of red-black GS on x (0: red, 1: black).
DARCY_INLINE double reg_colour(const RegLevel& L, const RegCell& c, double x, int colour) {
    const double v = reg_gs(c, reg_neighbours(L, c, x));
    return L.on && reg_red(L) == (colour == 0) ? v : x;
}

// Going down through L (visit_down): pre-smoothing, the first sweep from
// x = 0 when fresh (presmooth_from_zero), then the restricted residual into
// the b of the next level C (residual_restrict).
DARCY_INLINE void reg_down(const RegLevel& L, const RegLevel& C, bool fresh, const CycleOpts& o) {
    const int q = static_cast<int>(threadIdx.x & (kWave - 1));
    const RegCell c = reg_cell(L);
    double x = !fresh && L.on ? L.A.x[q] : 0.0;
    for (int s = 0; s < o.pre; ++s) {
        if (s == 0 && fresh) {
            const double red = c.b * c.inv;
            const double black = reg_gs(c, reg_neighbours(L, c, red));
            x = reg_red(L) ? red : black;
        } else {
            x = reg_colour(L, c, x, 0);
            x = reg_colour(L, c, x, 1);
        }
    }
    const double r = reg_residual(c, x, reg_neighbours(L, c, x));
    const int n = L.A.n;
    double s = 0.0;
#pragma unroll
    for (int d = 0; d < 4; ++d) {
        const int i = 2 * C.i + (d >> 1), j = 2 * C.j + (d & 1);
        const bool ok = C.on && i < n && j < n;
        const double rd = lane_value(r, ok ? i * n + j : 0);
        if (ok) s += rd;
    }
    if (L.on) L.A.x[q] = x;
    if (C.on) C.A.b[q] = s;
    WaveSync{}();
}

// Coming up through L from C (visit_up): the black sweep reads its red
// neighbours as x + x_c of their aggregate, then red, then further sweeps.
DARCY_INLINE void reg_up(const RegLevel& L, const RegLevel& C, const CycleOpts& o) {
    const int q = static_cast<int>(threadIdx.x & (kWave - 1));
    const RegCell c = reg_cell(L);
    double x = L.on ? L.A.x[q] : 0.0;
    const double xc = L.on ? C.A.x[(L.i >> 1) * C.A.n + (L.j >> 1)] : 0.0;
    const double v = reg_gs(c, reg_neighbours(L, c, x + xc));
    if (L.on && !reg_red(L)) x = v;
    x = reg_colour(L, c, x, 0);
    for (int s = 1; s < o.post; ++s) {
        x = reg_colour(L, c, x, 1);
        x = reg_colour(L, c, x, 0);
    }
    if (L.on) L.A.x[q] = x;
    WaveSync{}();
}

// Lane exchange within a quad (DPP quad_perm, no LDS): CTRL packs the
// source lane of each quad lane in 2 bits.
template <int CTRL>
DARCY_INLINE double quad_value(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}
constexpr int kQuadSwap1 = 0xB1;  // lanes (1, 0, 3, 2): east / west in a 2 x 2 level
constexpr int kQuadSwap2 = 0x4E;  // lanes (2, 3, 0, 1): north / south

// A visit of the 2 x 2 level L together with the 1 x 1 level C below it, on
// lane quads: quad lane q holds cell (q / 2, q % 2) and reads its neighbours
// (q ^ 1 east / west, q ^ 2 north / south) by quad permutes; the restricted
// residual gathers the four residuals by quad broadcasts and adds them in
// restricted_residual's order.  The phases of reg_down, the coarsest solve
// and reg_up in their order; every quad computes the same values and lanes
// 0..3 store x.
DARCY_INLINE void reg_visit2(const RegLevel& L, const RegLevel& C, bool fresh, const CycleOpts& o) {
    const int lane = static_cast<int>(threadIdx.x & (kWave - 1)), q = lane & 3, i = q >> 1, j = q & 1;
    const int f = 3 * i + j;
    const RegCell c{L.A.ex[f + 1], L.A.ex[f], L.A.ny[2 * (i + 1) + j], L.A.ny[2 * i + j], L.A.id[q], L.A.b[q],
                    j == 0, j == 1, i == 0, i == 1};
    const bool red = q == 0 || q == 3;
    auto nb = [](double v) {
        const double ew = quad_value<kQuadSwap1>(v), ns = quad_value<kQuadSwap2>(v);
        return RegNeighbours{ew, ew, ns, ns};
    };
    double x = fresh ? 0.0 : L.A.x[q];
    for (int s = 0; s < o.pre; ++s) {
        if (s == 0 && fresh) {
            const double r = c.b * c.inv;
            const double v = reg_gs(c, nb(r));
            x = red ? r : v;
        } else {
            const double vr = reg_gs(c, nb(x));
            x = red ? vr : x;
            const double vb = reg_gs(c, nb(x));
            x = red ? x : vb;
        }
    }
    const double r = reg_residual(c, x, nb(x));
    double sum = 0.0;
    sum += quad_value<0x00>(r);
    sum += quad_value<0x55>(r);
    sum += quad_value<0xAA>(r);
    sum += quad_value<0xFF>(r);
    const double xc = sum * C.A.id[0];  // the 1 x 1 level: x = b * inv_d
    {
        const double v = reg_gs(c, nb(x + xc));
        x = red ? x : v;
    }
    {
        const double v = reg_gs(c, nb(x));
        x = red ? v : x;
    }
    for (int s = 1; s < o.post; ++s) {
        const double vb = reg_gs(c, nb(x));
        x = red ? x : vb;
        const double vr = reg_gs(c, nb(x));
        x = red ? vr : x;
    }
    if (lane < 4) L.A.x[q] = x;
    WaveSync{}();
}

// Calls f(lv[k], lv[k + 1]) for a wave-uniform k < kRegLevels - 1.  The pair
// is selected first and f is called once, so its body is compiled once: the
// kernels that run this cycle are large, and one copy per level pair made
// the code the single wave steps through three times longer.
// Field by field, so that lv is only ever indexed by constants and stays in
// registers.
DARCY_INLINE RegLevel reg_pick(int k, const RegLevel& a, const RegLevel& b, const RegLevel& c) {
    auto sel = [k](auto x, auto y, auto z) { return k == 0 ? x : (k == 1 ? y : z); };
    RegLevel r;
    r.A.n = sel(a.A.n, b.A.n, c.A.n);
    r.A.ex = sel(a.A.ex, b.A.ex, c.A.ex);
    r.A.ny = sel(a.A.ny, b.A.ny, c.A.ny);
    r.A.id = sel(a.A.id, b.A.id, c.A.id);
    r.A.x = sel(a.A.x, b.A.x, c.A.x);
    r.A.b = sel(a.A.b, b.A.b, c.A.b);
    r.i = sel(a.i, b.i, c.i);
    r.j = sel(a.j, b.j, c.j);
    r.on = sel(a.on, b.on, c.on);
    return r;
}

template <class F>
DARCY_INLINE void reg_pair(const RegLevel* lv, int k, F&& f) {
    f(reg_pick(k, lv[0], lv[1], lv[2]), reg_pick(k, lv[1], lv[2], lv[3]));
}

// The cycle over levels 0..nl-1 of lv, the last of side 1 (cycle_levels
// with the visits above).  The last two levels (sides 2 and 1) are visited
// together by reg_visit2; a repeated visit of the 1 x 1 level recomputes
// x = b * inv_d from an unchanged b, so it is not repeated.
DARCY_INLINE void reg_cycle_levels(const RegLevel* lv, int nl, bool fresh0, const CycleOpts& o) {
    if (nl == 1) {  // a lone 1 x 1 level
        if (threadIdx.x == 0) lv[0].A.x[0] = lv[0].A.b[0] * lv[0].A.id[0];
        WaveSync{}();
        return;
    }
    const int bottom = nl - 2;
    uint32_t done = 0;
    int l = 0;
    bool fresh = fresh0;
    for (;;) {
        while (l < bottom) {
            reg_pair(lv, l, [&](const RegLevel& L, const RegLevel& C) { reg_down(L, C, fresh, o); });
            done &= ~(1u << l);
            ++l;
            fresh = true;
        }
        reg_pair(lv, l, [&](const RegLevel& L, const RegLevel& C) { reg_visit2(L, C, fresh, o); });
        for (;;) {
            if (l == 0) return;
            --l;
            if (o.gamma == 2 && !((done >> l) & 1u)) {
                done |= 1u << l;
                ++l;
                fresh = false;
                break;
            }
            reg_pair(lv, l, [&](const RegLevel& L, const RegLevel& C) { reg_up(L, C, o); });
        }
    }
}

// --------------------------------------------------------------- reductions

__device__ inline double wave_sum(double s) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, kWave);
    return __shfl(s, 0, kWave);
}

template <bool B>
struct Flag {
    static constexpr bool value = B;
};

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
                    if (lane == 0) to_global(part[d])[c0 + cc] = s;
                }
            }
        }
    }
}

// The CG direction pass: p = z + beta p_old, or p = z on the first iteration
// (solver.hpp:171, blas.hpp:50-56), fused with ap = A p (solver.hpp:147) and
// the chunk sums of <p, ap> (solver.hpp:148).  z is stored as Zt (float for
// the FP32 cycle) and widened to double.  With DEFER and a non-null `u`, the
// previous iteration's u += alpha_u p_old (blas.hpp:62-67) is applied here
// too, where p_old is read anyway, instead of in the update pass: the same
// operation on the same values, one read of p fewer per iteration.
template <int CPR, bool DEFER, class Zt>
DARCY_INLINE void cg_direction(int n, const GlobalLv<double>& A0, const Zt* zf, const double* poldf, double beta,
                               bool first, double* pnewf, double* apf, int w, int nw, double* const* part,
                               double* uf = nullptr, double alpha_u = 0.0) {
    // All of these are device allocations.
    const auto z = to_global(zf);
    const auto pold = to_global(poldf);
    const auto pnew = to_global(pnewf);
    const auto ap = to_global(apf);
    const auto u = to_global(uf);
    struct PAp {
        double p, ap, u;
    };
    // `load(t)` reads what p at t is formed from, `load3(k)` the same at
    // k - 1, k, k + 1 as one access per array (a value off the row is read
    // and its product selected away), and `pv` forms p.  All of a cell's
    // loads are issued before any arithmetic (the scheduling barrier keeps
    // them together): left to itself, the scheduler of this register-bound
    // kernel waits for each stencil point's loads before issuing the next
    // point's.
    auto pass = [&](auto&& load, auto&& load3, auto&& pv, auto upd) {
        constexpr bool U = decltype(upd)::value;
        chunk_pass<1, CPR>(
            n * n, w, nw, part,
            [&](int k) {
                const int i = k / n, j = k - i * n, f = i * (n + 1) + j;
                const bool hE = j + 1 < n, hW = j > 0, hN = i + 1 < n, hS = i > 0;
                const double e = A0.ex[f + 1], wf = A0.ex[f], nn = A0.ny[(i + 1) * n + j], s = A0.ny[i * n + j];
                decltype(load(0)) l[5];
                const auto r3 = load3(k);
                l[0] = r3.c;
                l[1] = r3.e;
                l[2] = r3.w;
                l[3] = load(hN ? k + n : k);
                l[4] = load(hS ? k - n : k);
                double uk = 0.0;
                if constexpr (U) uk = u[k];
                __builtin_amdgcn_sched_barrier(0);
                // apply_at's operations, in its order.
                const double d = ((e + wf) + nn) + s;
                const double x0 = pv(l[0]), xE = pv(l[1]), xW = pv(l[2]), xN = pv(l[3]), xS = pv(l[4]);
                double a = d * x0;
                a -= hE ? e * xE : 0.0;
                a -= hW ? wf * xW : 0.0;
                a -= hN ? nn * xN : 0.0;
                a -= hS ? s * xS : 0.0;
                if constexpr (U) uk = uk + alpha_u * l[0].p;
                return PAp{x0, a, uk};
            },
            [&](int k, const PAp& v, double* acc) {
                pnew[k] = v.p;
                ap[k] = v.ap;
                if constexpr (U) u[k] = v.u;
                acc[0] += v.p * v.ap;
            });
    };
    struct ZP {
        Zt z;
        double p;
    };
    auto zp = [&](int t) { return ZP{z[t], pold[t]}; };
    auto zp3 = [&](int t) {
        const Row3<Zt> a = load_row3(z + t);
        const Row3<double> b = load_row3(pold + t);
        return Row3<ZP>{ZP{a.w, b.w}, ZP{a.c, b.c}, ZP{a.e, b.e}};
    };
    auto zpv = [&](const ZP& v) { return static_cast<double>(v.z) + beta * v.p; };
    if (first)
        pass([&](int t) { return z[t]; }, [&](int t) { return load_row3(z + t); },
             [](Zt v) { return static_cast<double>(v); }, Flag<false>{});
    else if constexpr (DEFER) {
        if (uf)
            pass(zp, zp3, zpv, Flag<true>{});
        else
            pass(zp, zp3, zpv, Flag<false>{});
    } else
        pass(zp, zp3, zpv, Flag<false>{});
}

// Lane l adds chunk sums l, l+64, l+128, ... in that order; the loads of
// sixteen terms are issued before their adds.
__device__ inline double final_sum(const double* partf, int C) {
    const auto part = to_global(partf);
    const int lane = threadIdx.x & (kWave - 1);
    double s = 0.0;
    int c = lane;
    for (; c + 15 * kWave < C; c += 16 * kWave) {
        double v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) v[u] = part[c + u * kWave];
#pragma unroll
        for (int u = 0; u < 16; ++u) s += v[u];
    }
    for (; c < C; c += kWave) s += part[c];
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

// Level l of the hierarchy in global memory, as a global view.
template <class T>
DARCY_INLINE GlobalLv<T> global_level(const Hier& h, int l);
template <>
DARCY_INLINE GlobalLv<double> global_level<double>(const Hier& h, int l) {
    return {h.n[l], to_global<const double>(h.ex[l]), to_global<const double>(h.ny[l]),
            to_global<const double>(h.id[l]), to_global(static_cast<double*>(h.x[l])),
            to_global(static_cast<double*>(h.b[l]))};
}
template <>
DARCY_INLINE GlobalLv<float> global_level<float>(const Hier& h, int l) {
    return {h.n[l], to_global<const float>(h.ex32[l]), to_global<const float>(h.ny32[l]),
            to_global<const float>(h.id32[l]), to_global(static_cast<float*>(h.x[l])),
            to_global(static_cast<float*>(h.b[l]))};
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
    // Cells per step of the row-pipelined global levels (rows_per_step):
    // kStepCellsShared where many workgroups share each L2, kStepCellsAlone
    // for a workgroup that runs alone.
    int step_cells;
    // FP64 cycle only: the tail as an FP64 matrix (then a 64-entry residual),
    // or null for the recursive register-resident cycle (the reference's
    // operations, bit for bit).
    double* tail64 = nullptr;
    // Grid-path tail only: the levels from big_l (side <= 16) down as one
    // dense matrix in global memory, column k at big + k * big_ld (big_ld a
    // multiple of 16 bytes), applied by the whole workgroup; big_scratch is
    // LDS for the residual (256 entries) and the partial sums (16 bytes per
    // thread).  big_l = -1: not used.
    const T* big = nullptr;
    int big_l = -1, big_ld = 0;
    T* big_scratch = nullptr;
    // big holds the W-cycle's pair of visits of big_l (build_tail_matrix's
    // pair): the first visit applies it from zero and the second is skipped.
    // Valid only with big_l > the level run() starts from.
    bool big_pair = false;

    DARCY_INLINE GlobalLv<T> G(int l) const { return global_level<T>(h, l); }
    DARCY_INLINE LvP<T> S(int l) const {
        const int* f = t.off[l];
        return {h.n[l], pool + f[0], pool + f[1], pool + f[2], pool + f[3], pool + f[4]};
    }

    // --- one wave, everything in LDS (levels >= t.wave): the recursive cycle,
    // in FP64 on the register-resident levels when they end at side 1.
    DARCY_INLINE void wave_cycle(int l0, bool fresh0, Exec e) const {
        if constexpr (sizeof(T) == 8) {
            if (h.L - l0 <= kRegLevels && h.n[h.L - 1] == 1) {
                reg_cycle(l0, fresh0);
                return;
            }
        }
        cycle_levels<T>(l0, h.L, fresh0, o, e, WaveSync{}, [&](int l) { return S(l); });
    }

    // The register-resident cycle over levels l0..L-1 (all in LDS).
    DARCY_INLINE void reg_cycle(int l0, bool fresh0) const {
        const int q = static_cast<int>(threadIdx.x & (kWave - 1)), nl = h.L - l0;
        RegLevel lv[kRegLevels];
#pragma unroll
        for (int k = 0; k < kRegLevels; ++k) {
            const LvP<T> A = S(k < nl ? l0 + k : l0);
            const int n = A.n;
            const bool on = k < nl && q < n * n;
            const int i = on ? q / n : 0;
            lv[k] = RegLevel{A, i, on ? q - i * n : 0, on};
        }
        reg_cycle_levels(lv, nl, fresh0, o);
    }

    // --- The tail as its matrix M (by columns, then a residual vector):
    // the FP32 cycle's always, the FP64 cycle's when tail64 is set.  Fresh:
    // x = M b; repeated W-cycle visit: x += M (b - A x).  Lane i owns row i;
    // the row sum runs over k in order.
    template <class MT>
    DARCY_INLINE void dense_tail(MT* mat, int l, bool fresh, Exec e) const {
        const LvP<T> A = S(l);
        const int n = A.n, dim = n * n, i = e.tid;
        const WaveSync sync;
        MT* res = mat + kTailMat;
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
            MT s = MT(0);
            for (int k = 0; k < dim; ++k) s += mat[k * dim + i] * rhs[k];
            A.x[i] = fresh ? s : A.x[i] + s;
        }
        sync();
    }

    // The levels from big_l down as the dense matrix M: x = M b fresh, x +=
    // M (b - A x) on a repeated visit.  Threads own V consecutive rows (one
    // 16-byte load of a column) and groups of threads ranges of k, summed in
    // order; the groups' partial sums are then added in group order.  M
    // (256 KB in FP32, 512 KB in FP64 at side 16) is read from L2.
    DARCY_INLINE void dense_big(int l, bool fresh, Exec e) const {
        constexpr int V = 16 / static_cast<int>(sizeof(T));
        typedef T vec __attribute__((ext_vector_type(V)));
        using gvecT = __attribute__((address_space(1))) const vec;
        const LvP<T> A = S(l);
        const int n = A.n, dim = n * n, ld = big_ld;
        T* res = big_scratch;
        T* part = big_scratch + 256;
        const T* rhs = A.b;
        if (!fresh) {
            for (int i = e.tid; i < dim; i += e.nthr) {
                const int r = i / n, c = i - r * n;
                res[i] = A.b[i] - apply_at(n, A.ex, A.ny, r, c, [&](int q) { return A.x[q]; });
            }
            __syncthreads();
            rhs = res;
        }
        const int rows = ld / V, groups = e.nthr / rows, kper = (dim + groups - 1) / groups;
        const int rg = e.tid % rows, g = e.tid / rows;
        if (g < groups) {
            const int k0 = g * kper, k1 = k0 + kper < dim ? k0 + kper : dim;
            gvecT* col = (gvecT*)(big + rg * V);
            const int stride = ld / V;
            vec acc = {};
#pragma unroll 8
            for (int k = k0; k < k1; ++k) acc += col[k * stride] * rhs[k];
#pragma unroll
            for (int v = 0; v < V; ++v) part[g * ld + rg * V + v] = acc[v];
        }
        __syncthreads();
        for (int i = e.tid; i < dim; i += e.nthr) {
            T sum = part[i];
            for (int q = 1; q < groups; ++q) sum += part[q * ld + i];
            A.x[i] = fresh ? sum : A.x[i] + sum;
        }
        __syncthreads();
    }

    // Global-memory levels stream one pass per visit when one sweep is
    // asked for (the default); with more sweeps they run phase by phase.
    // In the FP32 cycle a fresh visit runs as walks: 1000 x 256^2 launch
    // 95.0 -> 86.3 ms.  The FP64 cycle keeps the cell-by-cell pass: the
    // walks' registers raised its spills from 76 to 292 bytes and the FP64
    // launch from 126.4 to 156.3 ms.
    template <class BP>
    DARCY_INLINE void down_global(int l, BP bc, bool fresh, Exec e) const {
        if (o.pre != 1)
            visit_down(G(l), bc, fresh, o, e, BlockSync{});
        else if (sizeof(T) == 4 && fresh)
            down_walk(G(l), bc, step_cells, e);
        else
            down_stream(G(l), bc, fresh, step_cells, e);
    }

    DARCY_INLINE void down(int l, bool fresh, Exec e) const {
        if (l >= t.first)
            visit_down(S(l), S(l + 1).b, fresh, o, e, BlockSync{});
        else if (l + 1 >= t.first)
            down_global(l, S(l + 1).b, fresh, e);
        else
            down_global(l, G(l + 1).b, fresh, e);
    }

    // The up visit of a global level.  Walks were measured here too (the
    // black sweep with prolongation, then the red one a row behind, one
    // updated cell per row and column pair) and dropped: the FP32 launch
    // went from 86.3 to 91.5 ms.
    template <class PC>
    DARCY_INLINE void up_global(int l, PC xc, Exec e) const {
        if (o.post == 1)
            up_stream(G(l), xc, step_cells, e);
        else
            visit_up(G(l), xc, o, e, BlockSync{});
    }

    DARCY_INLINE void up(int l, Exec e) const {
        if (l >= t.first)
            visit_up(S(l), S(l + 1).x, o, e, BlockSync{});
        else if (l + 1 >= t.first)
            up_global(l, static_cast<const T*>(S(l + 1).x), e);
        else
            up_global(l, G(l + 1).x, e);
    }

    // z = M r from level l0 (x_{l0} zero on entry when fresh0).
    DARCY_INLINE void run(int l0, bool fresh0, Exec e) const {
        const int L = h.L;
        uint32_t done = 0;
        int l = l0;
        bool fresh = fresh0;
        for (;;) {
            while (l + 1 < L && l < t.wave && l != big_l) {
                down(l, fresh, e);
                done &= ~(1u << l);
                ++l;
                fresh = true;
            }
            if (l == big_l) {
                dense_big(l, fresh, e);
                if (big_pair) done |= 1u << (l - 1);  // both visits applied
            } else if (l >= t.wave) {
                if (threadIdx.x < kWave) {
                    const Exec we{static_cast<int>(threadIdx.x), kWave};
                    if constexpr (sizeof(T) == 4)
                        dense_tail(tail, l, fresh, we);
                    else if (tail64)
                        dense_tail(tail64, l, fresh, we);
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

// Tail matrix of the cycle from level tl.  Column k is the FP64 recursive
// cycle from level tl -- the reference's arithmetic, bit-exact with the
// oracle -- applied to the unit vector e_k, rounded to OT (FP32 for the FP32
// cycle, FP64 for the FP64 cycle's dense tail) and stored by columns, column
// k at tail + k * ld (ld >= dim; rows dim..ld-1 are zero).  The waves of the
// launch take columns in turn (this workgroup's first is col0, then every
// col_step); `scratch` (the LDS pool, before it receives the operators, or
// an array of its own) holds the FP64 operators of the tail levels and each
// wave's x and b; waves past the first kWaves (a wider workgroup) take no
// columns, so the scratch holds kWaves waves' vectors whatever the width.
// pair: column k is the W-cycle's two visits of level tl instead -- the
// cycle from zero, then the cycle again from its result with the same
// right-hand side -- which is the matrix 2M - M A M of the single cycle M.
template <class OT>
DARCY_INLINE void build_tail_matrix(const Hier& h, int tl, const CycleOpts& o, void* scratch, OT* tail, int ld = 0,
                                    int col0 = 0, int col_step = kWaves, bool pair = false) {
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
        for (int q = threadIdx.x; q < nf; q += blockDim.x) {
            dp[base + q] = h.ex[l][q];
            dp[base + nf + q] = h.ny[l][q];
        }
        for (int q = threadIdx.x; q < n * n; q += blockDim.x) dp[base + 2 * nf + q] = h.id[l][q];
    }
    __syncthreads();
    int vsz = 0;  // x and b of every level from tl, per wave
    for (int q = tl; q < L; ++q) vsz += 2 * h.n[q] * h.n[q];
    double* vec = dp + ops_end + wave * vsz;
    auto view = [&](int l) {
        int off = 0;
        for (int q = tl; q < l; ++q) off += 2 * h.n[q] * h.n[q];
        const int n = h.n[l], base = op_off(l);
        return LvP<double>{n, dp + base, dp + base + n * (n + 1), dp + base + 2 * n * (n + 1), vec + off,
                           vec + off + n * n};
    };
    const int dim = h.n[tl] * h.n[tl];
    if (ld == 0) ld = dim;
    const WaveSync sync;
    for (int k = col0 + wave; wave < kWaves && k < dim; k += col_step) {
        const LvP<double> top = view(tl);
        for (int q = lane; q < dim; q += kWave) top.b[q] = q == k ? 1.0 : 0.0;
        sync();
        cycle_levels<double>(tl, L, true, o, Exec{lane, kWave}, sync, view);
        if (pair) cycle_levels<double>(tl, L, false, o, Exec{lane, kWave}, sync, view);
        for (int q = lane; q < ld; q += kWave) tail[static_cast<size_t>(k) * ld + q] = q < dim ? static_cast<OT>(top.x[q]) : OT(0);
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

}  // namespace darcy_hip
