// Device building blocks of the solve path for gfx950 (CDNA4, wave64).
//
// Written once, used by the persistent batch kernel (one workgroup owns a
// whole problem), so the per-cell arithmetic -- and every result -- is the
// same whichever way a problem is scheduled.
//
// Per-cell floating-point operations follow the reference's operation order
// exactly (include/darcy/level.hpp), so that with -ffp-contract=off the FP64
// cycle reproduces the reference bit for bit:
//   apply/residual  v = d*x_k - e*x_E - w*x_W - n*x_N - s*x_S   (level.hpp:111-116, 129-134)
//   GS update       x_k = ((((b_k + e*x_E) + w*x_W) + n*x_N) + s*x_S) * inv_d   (level.hpp:153-158)
//   restriction     sum over (i0,j0),(i0,j1),(i1,j0),(i1,j1)   (level.hpp:221-224)
// Fusions keep those values: a value the reference stores and reloads (the
// first red sweep from zero, a prolongated cell) is recomputed in the same
// operation order where a fused phase needs it.
//
// Memory placement of the hierarchy inside one workgroup:
//   levels with side > NL ........ global memory (L2/HBM), whole workgroup
//   levels with side <= NL ....... LDS, whole workgroup
//   levels with side <= 8 ........ LDS, the first wave alone, with no
//                                  workgroup barrier (the W-cycle visits the
//                                  coarse levels 2^l times: this is where the
//                                  barriers were)
//
// Reductions (the CG dot products) use one canonical order that does not
// depend on launch shape: the flat vector is cut into chunks of 256 cells;
// lane l of the wave that owns a chunk adds cells l, l+64, l+128, l+192 in
// turn, the 64 lane sums are combined by a fixed xor butterfly, and the chunk
// sums are combined the same way by one wave.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <type_traits>

#include "layout.hpp"

#define DARCY_INLINE __device__ __forceinline__

namespace darcy_hip {

constexpr int kWave = 64;
constexpr int kWaveTail = 8;  // levels with side <= 8 run on one wave

// ------------------------------------------------------------ level views

// Where a level's arrays live, as the type of its pointers.  FlatPtr: any
// memory (the LDS levels; host-side checks of the walks).  GlobalPtr: device
// allocations, in the global address space, so that their accesses are
// global instructions -- an SGPR base and a 32-bit VGPR offset, waited for
// on the vector-memory counter alone -- rather than flat ones, which take a
// 64-bit VGPR address and count against the LDS counter too.  A pointer the
// kernel reads from LDS (the batch kernel's slot header) or from a buffer is
// otherwise flat.
template <class T>
struct FlatPtr {
    using type = T*;
};
template <class T>
struct GlobalPtr {
    using type = __attribute__((address_space(1))) T*;
};

// The element type of a pointer of either kind, without qualifiers.
template <class P>
struct Pointee;
template <class T>
struct Pointee<T*> {
    using type = std::remove_cv_t<T>;
};
template <class T>
struct Pointee<__attribute__((address_space(1))) T*> {
    using type = std::remove_cv_t<T>;
};
template <class P>
using pointee_t = typename Pointee<std::remove_cv_t<P>>::type;

template <class P>
struct IsGlobalPtr : std::false_type {};
template <class T>
struct IsGlobalPtr<__attribute__((address_space(1))) T*> : std::true_type {};

// p viewed as a U pointer in the same memory.
template <class U, class T>
DARCY_HD __forceinline__ U* retype(T* p) {
    return reinterpret_cast<U*>(p);
}
template <class U, class T>
DARCY_HD __forceinline__ __attribute__((address_space(1))) U* retype(__attribute__((address_space(1))) T* p) {
    return reinterpret_cast<__attribute__((address_space(1))) U*>(p);
}

// p, a pointer into a device allocation (never LDS), as a global pointer.
template <class T>
DARCY_INLINE typename GlobalPtr<T>::type to_global(T* p) {
    return (typename GlobalPtr<T>::type)p;
}

// Operator and vectors of one level in the cycle's arithmetic type T.
template <class T, template <class> class Space = FlatPtr>
struct LvP {
    using ptr = typename Space<T>::type;
    using cptr = typename Space<const T>::type;
    int n;
    cptr ex;  // n x (n+1) faces
    cptr ny;  // (n+1) x n faces
    cptr id;  // n x n reciprocal diagonal
    ptr x;
    ptr b;
};
template <class T>
using GlobalLv = LvP<T, GlobalPtr>;

// A cell's west, own and east values.
template <class T>
struct Row3 {
    T w, c, e;
};

// Vectors of 2 and 3 elements aligned as one element: loaded as one access
// (dwordx2 / dwordx3, or dwordx4 + dwordx2 for three doubles) wherever they
// start.
template <class T>
struct Vec;
template <>
struct Vec<float> {
    typedef float v2 __attribute__((ext_vector_type(2), aligned(4)));
    typedef float v3 __attribute__((ext_vector_type(3), aligned(4)));
};
template <>
struct Vec<double> {
    typedef double v2 __attribute__((ext_vector_type(2), aligned(8)));
    typedef double v3 __attribute__((ext_vector_type(3), aligned(8)));
};

// p[-1], p[0], p[1] as one access.  At a row's first or last cell the
// element beyond it belongs to the neighbouring row, to the alignment
// padding between arrays or to the guard bytes of the slot (or of the LDS
// pool): it is read and not used.
template <class P>
DARCY_HD __forceinline__ Row3<pointee_t<P>> load_row3(P p) {
    using T = pointee_t<P>;
    const typename Vec<T>::v3 v = *retype<const typename Vec<T>::v3>(p - 1);
    return Row3<T>{v.x, v.y, v.z};
}

// p[0], p[1] as one access.
template <class P>
DARCY_HD __forceinline__ Row3<pointee_t<P>> load_two(P p) {
    using T = pointee_t<P>;
    const typename Vec<T>::v2 v = *retype<const typename Vec<T>::v2>(p);
    return Row3<T>{v.x, v.x, v.y};
}

// x as stored, read by grid coordinates; row3 reads a cell's west, own and
// east values as one access (gs_at and apply_at take it where UsesRow3).
template <class P>
struct XRow {
    static constexpr bool kRow3 = IsGlobalPtr<std::remove_cv_t<P>>::value;
    P x;
    int n;
    DARCY_INLINE pointee_t<P> operator()(int a, int c) const { return x[a * n + c]; }
    DARCY_INLINE Row3<pointee_t<P>> row3(int a, int c) const { return load_row3(x + a * n + c); }
};

// x as stored, read by flat index.
template <class P>
struct XFlat {
    static constexpr bool kRow3 = IsGlobalPtr<std::remove_cv_t<P>>::value;
    P x;
    DARCY_INLINE pointee_t<P> operator()(int t) const { return x[t]; }
    DARCY_INLINE Row3<pointee_t<P>> row3(int t) const { return load_row3(x + t); }
};

// Whether a loader's row3 is used: only for arrays in global memory.  In LDS
// the three values are three reads either way (the wide read would need
// 16-byte alignment), one more than the two neighbours alone.
template <class Ld, class = void>
struct UsesRow3 : std::false_type {};
template <class Ld>
struct UsesRow3<Ld, std::void_t<decltype(Ld::kRow3)>> : std::integral_constant<bool, Ld::kRow3> {};

// ------------------------------------------------------------------ sync

struct BlockSync {
    DARCY_INLINE void operator()() const { __syncthreads(); }
};

// One wave: LDS operations of a wave complete in order; the fences keep the
// compiler from moving memory operations across the barrier.
struct WaveSync {
    DARCY_INLINE void operator()() const {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
};

struct Exec {
    int tid;
    int nthr;
};

// ---------------------------------------------------------------- per cell

// (A x)_k with off-grid neighbours dropped (level.hpp:105-119).  `X` maps a
// flat index to the value of x there, so callers can form x on the fly.
//
// The stencil is branch-free so that every load of a cell issues at once:
// an off-grid neighbour is read at the cell itself and its product is
// selected away for an exact zero (the reference skips the term; results
// differ at most in the sign of a zero).  The product, not the coefficient,
// is selected: the value read in place of a missing neighbour may be stale
// or uninitialised.
// (A x)_k from the cell's faces and x values: d x_k - e x_E - w x_W - n x_N
// - s x_S with d = ((e + w) + n) + s, an off-grid neighbour's product
// selected away.
template <class T>
DARCY_HD __forceinline__ T apply_combine(T e, T w, T nn, T s, T xk, T xE, T xW, T xN, T xS, bool hE, bool hW,
                                         bool hN, bool hS) {
    const T d = ((e + w) + nn) + s;
    T v = d * xk;
    v -= hE ? e * xE : T(0);
    v -= hW ? w * xW : T(0);
    v -= hN ? nn * xN : T(0);
    v -= hS ? s * xS : T(0);
    return v;
}

template <class P, class X>
DARCY_INLINE pointee_t<P> apply_at(int n, P ex, P ny, int i, int j, X&& x) {
    using T = pointee_t<P>;
    const int k = i * n + j, f = i * (n + 1) + j;
    const T e = ex[f + 1], w = ex[f], nn = ny[(i + 1) * n + j], s = ny[i * n + j];
    const bool hE = j + 1 < n, hW = j > 0, hN = i + 1 < n, hS = i > 0;
    T xk, xE, xW;
    if constexpr (UsesRow3<std::decay_t<X>>::value) {
        const Row3<T> r = x.row3(k);  // off-row values are read and not used
        xk = r.c;
        xE = r.e;
        xW = r.w;
    } else {
        xk = x(k);
        xE = x(hE ? k + 1 : k);
        xW = x(hW ? k - 1 : k);
    }
    const T xN = x(hN ? k + n : k), xS = x(hS ? k - n : k);
    __builtin_amdgcn_sched_barrier(0);  // all loads in flight together, as in gs_at
    const T d = ((e + w) + nn) + s;
    T v = d * xk;
    v -= hE ? e * xE : T(0);
    v -= hW ? w * xW : T(0);
    v -= hN ? nn * xN : T(0);
    v -= hS ? s * xS : T(0);
    return v;
}

// The reciprocal diagonal from the cell's four faces: ((e + w) + n) + s
// (level.hpp:92), then 1 / d in T.  In FP64 this is the stored inv_diag bit
// for bit (the faces are the ones it was formed from, and the harmonic mean
// is symmetric bit for bit); in FP32 it is the FP32 cycle's definition, which
// its stored copy follows too.  Forming it saves the kernels its load.
template <class T>
DARCY_HD __forceinline__ T inv_from_faces(T e, T w, T n, T s) {
    return T(1) / (((e + w) + n) + s);
}

template <class T, template <class> class Sp>
DARCY_INLINE T inv_diag_at(const LvP<T, Sp>& L, int i, int j) {
    const int n = L.n, f = i * (n + 1) + j;
    return inv_from_faces(L.ex[f + 1], L.ex[f], L.ny[(i + 1) * n + j], L.ny[i * n + j]);
}

struct Ident {
    template <class V>
    DARCY_INLINE V operator()(const V& v) const {
        return v;
    }
};

// The GS update of one cell (level.hpp:153-158) from its faces, b and the
// neighbour values: ((((b + e x_E) + w x_W) + n x_N) + s x_S) * inv_d, an
// off-grid neighbour's product selected away.  gs_at keeps its own copy of
// this expression (and apply_at of apply_combine's): routing them through
// these functions changed the persistent kernel's schedule and made the
// 2000-problem uneven batch 5-13% slower.  The copies must stay identical;
// the grid path's walks, which use these, are checked against them bit for
// bit.
template <class T>
DARCY_HD __forceinline__ T gs_combine(T e, T w, T nn, T s, T bk, T xE, T xW, T xN, T xS, bool hE, bool hW, bool hN, bool hS) {
    const T inv = inv_from_faces(e, w, nn, s);
    T v = bk;
    v += hE ? e * xE : T(0);
    v += hW ? w * xW : T(0);
    v += hN ? nn * xN : T(0);
    v += hS ? s * xS : T(0);
    return v * inv;
}

// Gauss-Seidel value of cell (i, j) from its neighbours (level.hpp:153-158).
// `ld(i, j)` reads what x at grid coordinates (i, j) is formed from and
// `cv` forms it (by default x is read as stored).  All of the cell's loads
// are issued before its arithmetic -- the scheduling barrier keeps them
// together: left to itself, the scheduler of a register-bound kernel waits
// for each neighbour's loads before issuing the next neighbour's.
template <class T, template <class> class Sp, class Ld, class Cv = Ident>
DARCY_INLINE T gs_at(const LvP<T, Sp>& L, int i, int j, T bk, Ld&& ld, Cv&& cv = Cv{}) {
    const int n = L.n, f = i * (n + 1) + j;
    const bool hE = j + 1 < n, hW = j > 0, hN = i + 1 < n, hS = i > 0;
    const T e = L.ex[f + 1], w = L.ex[f], nn = L.ny[(i + 1) * n + j], s = L.ny[i * n + j];
    decltype(ld(i, j)) rE, rW;
    if constexpr (UsesRow3<std::decay_t<Ld>>::value) {
        const auto r = ld.row3(i, j);  // off-row values are read and not used
        rE = r.e;
        rW = r.w;
    } else {
        rE = ld(i, hE ? j + 1 : j);
        rW = ld(i, hW ? j - 1 : j);
    }
    const auto rN = ld(hN ? i + 1 : i, j), rS = ld(hS ? i - 1 : i, j);
    __builtin_amdgcn_sched_barrier(0);
    const T inv = inv_from_faces(e, w, nn, s);
    const T xE = cv(rE), xW = cv(rW), xN = cv(rN), xS = cv(rS);
    T v = bk;
    v += hE ? e * xE : T(0);
    v += hW ? w * xW : T(0);
    v += hN ? nn * xN : T(0);
    v += hS ? s * xS : T(0);
    return v * inv;
}

// ------------------------------------------------------------------ phases
// A phase is run by `nthr` cooperating threads; `tid` is this thread's rank.
// Threads handle up to four cells per step with all loads issued before the
// stores, so each thread keeps several memory requests in flight.
// Callers separate dependent phases with a barrier.

// Colour `c` cell number q of a level: row i, column 2*(q mod h) + ((i+c)&1).
DARCY_INLINE bool colour_cell(int q, int n, int h, int c, int& i, int& j) {
    i = q / h;
    j = 2 * (q - i * h) + ((i + c) & 1);
    return j < n;
}

// One colour of red-black GS (level.hpp:145-161).
template <class T, template <class> class Sp>
DARCY_INLINE void gs_colour(const LvP<T, Sp>& L, int colour, Exec e) {
    const int n = L.n, h = (n + 1) >> 1, total = n * h;
    const auto x = L.x;
    for (int q0 = e.tid; q0 < total; q0 += 4 * e.nthr) {
        T v[4];
        int kk[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            int i, j;
            const int q = q0 + u * e.nthr;
            kk[u] = -1;
            if (q < total && colour_cell(q, n, h, colour, i, j)) {
                kk[u] = i * n + j;
                v[u] = gs_at(L, i, j, L.b[kk[u]], XRow<decltype(x)>{x, n});
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (kk[u] >= 0) x[kk[u]] = v[u];
    }
}

// Pre-smoothing red + black on a level whose x is zero, in one phase.  The
// red sweep from zero gives b * inv_d at every red cell (each neighbour term
// is e * 0); the black update reads those red values, recomputed here in the
// same operations instead of being reloaded.
template <class T, template <class> class Sp>
DARCY_INLINE void presmooth_from_zero(const LvP<T, Sp>& L, Exec e) {
    const int n = L.n, h = (n + 1) >> 1, total = n * h;
    const auto x = L.x;
    const auto b = L.b;
    const auto id = L.id;
    struct BI {
        T b, id;
    };
    auto red_ld = [&](int a, int c) { return BI{b[a * n + c], id[a * n + c]}; };
    auto red_cv = [](const BI& v) { return v.b * v.id; };
    auto red = [&](int a, int c) { return b[a * n + c] * id[a * n + c]; };
    for (int q0 = e.tid; q0 < total; q0 += 4 * e.nthr) {
        T vr[4], vb[4];
        int kr[4], kb[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            int i, j;
            const int q = q0 + u * e.nthr;
            kr[u] = kb[u] = -1;
            if (q < total && colour_cell(q, n, h, 0, i, j)) {
                kr[u] = i * n + j;
                vr[u] = red(i, j);
            }
            if (q < total && colour_cell(q, n, h, 1, i, j)) {
                kb[u] = i * n + j;
                vb[u] = gs_at(L, i, j, b[kb[u]], red_ld, red_cv);
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (kr[u] >= 0) x[kr[u]] = vr[u];
            if (kb[u] >= 0) x[kb[u]] = vb[u];
        }
    }
}

// Pre-smoothing red + black from x = 0 as a vertical walk, for levels that
// have the whole GPU (the grid path).  Work item (k, rows i0 .. i1-1) owns
// the column pair (2k, 2k+1): row a's red cell is column 2k + (a & 1), its
// black cell the other one.  Each row's red value b * inv_d is formed once
// and kept in registers for the black cells of the rows below and above; the
// red neighbour in the adjacent pair (east on even rows, west on odd rows) is
// formed from its own loads.  The loads of the next row are issued before
// the arithmetic of the current one.  inv_d is formed from the faces (the stored
// value bit for bit), so the walk reads b and the faces and writes x, and
// every value is the one presmooth_from_zero stores, in the same operations.
constexpr int kWalkRows = 16;  // most rows per work item

template <class T>
struct WalkRow {
    T b0, b1;  // b of the pair's two cells
    T w0, m, e1;  // ex: west of cell 0, between the cells, east of cell 1
    T s0, s1;  // ny: south faces of the two cells
};

template <class T>
struct WalkCell {
    T b, e, w, nn, s;  // b and the four faces of one cell
};

// Two consecutive values at an even offset, loaded or stored as one access.
template <class T>
struct alignas(2 * sizeof(T)) Pair2 {
    T lo, hi;
};

// The pair at p (aligned), in either memory; member by member, which the
// compiler merges into one access.
template <class T, class P>
DARCY_HD __forceinline__ Pair2<T> load_pair(P p) {
    const auto q = retype<const Pair2<T>>(p);
    return Pair2<T>{q->lo, q->hi};
}
template <class T, class P>
DARCY_HD __forceinline__ void store_pair(P p, Pair2<T> v) {
    const auto q = retype<Pair2<T>>(p);
    q->lo = v.lo;
    q->hi = v.hi;
}

// kPairs: n is even, so every pair starts at an even offset and both of its
// cells are on the grid; b, ny and x move as pairs.
// Values at (a, c0) and (a, c1) of an n-wide array (c1 = c0 + 1, or c0 when
// the pair's second cell is off the grid).
template <class T, bool kPairs, class P>
DARCY_HD __forceinline__ Pair2<T> walk_pair(P p, int n, int c0, int c1, int a) {
    if constexpr (kPairs) {
        return load_pair<T>(p + a * n + c0);
    } else {
        return Pair2<T>{p[a * n + c0], p[a * n + c1]};
    }
}

// b, the ex faces and the south faces of the pair's cells on row a.  b and
// ny pairs start at even offsets; the ex row (n + 1 long) holds an aligned
// pair at f on even rows and at f + 1 on odd ones.  `ss` is the pair's
// south faces (ny of row a), which a walk has already loaded as the north
// faces of the row below: one load per row fewer.
template <class T, bool kPairs, template <class> class Sp = FlatPtr>
DARCY_HD __forceinline__ WalkRow<T> walk_row(const LvP<T, Sp>& L, int c0, int c1, int a, Pair2<T> ss) {
    const int n = L.n, f = a * (n + 1) + c0;
    const auto ex = L.ex;
    const Pair2<T> bb = walk_pair<T, kPairs>(L.b, n, c0, c1, a);
    if constexpr (kPairs) {
        if (a & 1) {
            const Pair2<T> q = load_pair<T>(ex + f + 1);
            return WalkRow<T>{bb.lo, bb.hi, ex[f], q.lo, q.hi, ss.lo, ss.hi};
        }
        const Pair2<T> q = load_pair<T>(ex + f);
        return WalkRow<T>{bb.lo, bb.hi, q.lo, q.hi, ex[f + 2], ss.lo, ss.hi};
    } else {
        return WalkRow<T>{bb.lo, bb.hi, ex[f], ex[f + 1], ex[f + 1 + (c1 - c0)], ss.lo, ss.hi};
    }
}
template <class T, bool kPairs, template <class> class Sp = FlatPtr>
DARCY_HD __forceinline__ WalkRow<T> walk_row(const LvP<T, Sp>& L, int c0, int c1, int a) {
    return walk_row<T, kPairs>(L, c0, c1, a, walk_pair<T, kPairs>(L.ny, L.n, c0, c1, a));
}

// Where the walk takes the adjacent pair's red value from.  NoShare: from
// its own loads.  The device walk passes LaneShare: lane l + 1 (l - 1) of the
// wave walks pair k + 1 (k - 1) of the same rows whenever `east` (`west`)
// holds, so that value is its current red value, taken by a lane shuffle.
struct NoShare {
    bool east = false, west = false;
    template <class T>
    DARCY_HD T from_east(T) const {
        return T(0);
    }
    template <class T>
    DARCY_HD T from_west(T) const {
        return T(0);
    }
};

template <class T, bool kPairs, class Share = NoShare, template <class> class Sp = FlatPtr>
DARCY_HD __forceinline__ void presmooth_walk_item(const LvP<T, Sp>& L, int k, int i0, int i1, Share sh = Share{}) {
    const int n = L.n, c0 = 2 * k;
    const bool two = c0 + 1 < n;     // the pair's second cell is on the grid
    const int c1 = two ? c0 + 1 : c0;  // its column, clamped for the loads
    const auto ex = L.ex;
    const auto ny = L.ny;
    const auto b = L.b;
    const auto x = L.x;
    auto row = [&](int a) { return walk_row<T, kPairs>(L, c0, c1, a); };
    // North faces of row a - 1 (the south faces of row a), into (s0, s1).
    auto south = [&](int a, T& s0, T& s1) {
        const Pair2<T> q = walk_pair<T, kPairs>(ny, n, c0, c1, a);
        s0 = q.lo;
        s1 = q.hi;
    };
    // Red value of row a from its row data and north faces (n0, n1).
    auto red_of = [&](int a, const WalkRow<T>& R, T n0, T n1) {
        return (a & 1) ? R.b1 * inv_from_faces(R.e1, R.m, n1, R.s1) : R.b0 * inv_from_faces(R.m, R.w0, n0, R.s0);
    };
    // b and faces of cell (a, c), and the adjacent pair's red cell on row a
    // (east on even rows, west on odd rows; clamped to c0 where it is off the
    // grid, its product then unused).
    auto cell = [&](int a, int c) {
        const int f = a * (n + 1) + c;
        return WalkCell<T>{b[a * n + c], ex[f + 1], ex[f], ny[(a + 1) * n + c], ny[a * n + c]};
    };
    auto cross_on = [&](int a) { return (a & 1) ? c0 > 0 : c0 + 2 < n; };
    auto shared = [&](int a) { return (a & 1) ? sh.west : sh.east; };
    auto cross_col = [&](int a) { return cross_on(a) ? ((a & 1) ? c0 - 1 : c0 + 2) : c0; };

    WalkRow<T> cur = row(i0);
    T cn0, cn1;  // north faces of row i0
    south(i0 + 1, cn0, cn1);
    T rcur = red_of(i0, cur, cn0, cn1), rprev = T(0);
    if (i0 > 0) rprev = red_of(i0 - 1, row(i0 - 1), cur.s0, cur.s1);
    // One row ahead: row i0 + 1 with its north faces, and row i0's cross cell.
    WalkRow<T> nx = cur;
    T nn0 = T(0), nn1 = T(0);
    if (i0 + 1 < n) {
        nx = walk_row<T, kPairs>(L, c0, c1, i0 + 1, Pair2<T>{cn0, cn1});
        south(i0 + 2, nn0, nn1);
    }
    WalkCell<T> cx{};
    if (!shared(i0)) cx = cell(i0, cross_col(i0));
    for (int a = i0; a < i1; ++a) {
        const bool hN = a + 1 < n, hS = a > 0, odd = (a & 1) != 0, more = a + 1 < i1;
        // Loads for the next row, issued before this row's arithmetic: row
        // a + 2 with its north faces and row a + 1's cross cell.
        WalkRow<T> nx2 = nx;
        T mm0 = T(0), mm1 = T(0);
        if (more && a + 2 < n) {
            nx2 = walk_row<T, kPairs>(L, c0, c1, a + 2, Pair2<T>{nn0, nn1});  // nn: ny of row a + 2
            south(a + 3, mm0, mm1);
        }
        WalkCell<T> cx2 = cx;
        if (more && !shared(a + 1)) cx2 = cell(a + 1, cross_col(a + 1));
        const T rnext = hN ? red_of(a + 1, nx, nn0, nn1) : T(0);
        const bool hX = cross_on(a);
        // Every lane takes part in both shuffles.
        const T ve = sh.from_east(rcur), vw = sh.from_west(rcur);
        const T cross = shared(a) ? (odd ? vw : ve) : cx.b * inv_from_faces(cx.e, cx.w, cx.nn, cx.s);
        const T blk = odd ? gs_combine(cur.m, cur.w0, cn0, cur.s0, cur.b0, rcur, cross, rnext, rprev, two, hX, hN, hS)
                          : gs_combine(cur.e1, cur.m, cn1, cur.s1, cur.b1, cross, rcur, rnext, rprev, hX, true, hN, hS);
        const auto xr = x + a * n + c0;
        if constexpr (kPairs) {
            store_pair<T>(xr, odd ? Pair2<T>{blk, rcur} : Pair2<T>{rcur, blk});
        } else if (odd) {
            xr[0] = blk;
            if (two) xr[1] = rcur;
        } else {
            xr[0] = rcur;
            if (two) xr[1] = blk;
        }
        rprev = rcur;
        rcur = rnext;
        cur = nx;
        cn0 = nn0;
        cn1 = nn1;
        nx = nx2;
        nn0 = mm0;
        nn1 = mm1;
        cx = cx2;
    }
}

// Residual and restriction (level.hpp:122-137, :214-227) as a vertical walk
// for the grid path.  Work item (J, coarse rows I0 .. I1-1) owns the fine
// column pair (2J, 2J+1) and walks its fine rows, holding x of the rows
// below, at and above in registers; x of the columns beside the pair comes
// from the neighbouring lanes' current rows through `sh` (from loads at wave
// and strip edges).  The next row's loads are issued before the current
// row's arithmetic.  Each residual is apply_at's expression and each coarse
// value sums its aggregate from zero in the order (i0,j0),(i0,j1),(i1,j0),
// (i1,j1), as restricted_residual does: the same bits.
template <class T, bool kPairs, class Share = NoShare, template <class> class Sp = FlatPtr, class BP = T*>
DARCY_HD __forceinline__ void restrict_walk_item(const LvP<T, Sp>& L, BP bc, int J, int I0, int I1, Share sh = Share{}) {
    const int n = L.n, m = (n + 1) >> 1, c0 = 2 * J;
    const bool two = c0 + 1 < n, west_on = c0 > 0, east_on = c0 + 2 < n;
    const int c1 = two ? c0 + 1 : c0;
    const auto x = L.x;
    const auto ny = L.ny;
    const int a0 = 2 * I0, a1 = 2 * I1 < n ? 2 * I1 : n;
    auto xpair = [&](int a) { return walk_pair<T, kPairs>(x, n, c0, c1, a); };
    Pair2<T> xp{T(0), T(0)}, xc = xpair(a0), xn{T(0), T(0)};
    if (a0 > 0) xp = xpair(a0 - 1);
    if (a0 + 1 < n) xn = xpair(a0 + 1);
    WalkRow<T> cur = walk_row<T, kPairs>(L, c0, c1, a0);
    Pair2<T> north = walk_pair<T, kPairs>(ny, n, c0, c1, a0 + 1);
    T xw = T(0), xe = T(0);  // x beside the pair, where no lane shares it
    if (!sh.west && west_on) xw = x[a0 * n + c0 - 1];
    if (!sh.east && east_on) xe = x[a0 * n + c0 + 2];
    T acc = T(0);
    for (int a = a0; a < a1; ++a) {
        const bool hN = a + 1 < n, hS = a > 0, more = a + 1 < a1;
        Pair2<T> xn2 = xn, north2 = north;
        WalkRow<T> nx = cur;
        T xw2 = T(0), xe2 = T(0);
        if (more) {
            nx = walk_row<T, kPairs>(L, c0, c1, a + 1, north);  // north: ny of row a + 1
            north2 = walk_pair<T, kPairs>(ny, n, c0, c1, a + 2);
            if (a + 2 < n) xn2 = xpair(a + 2);
            if (!sh.west && west_on) xw2 = x[(a + 1) * n + c0 - 1];
            if (!sh.east && east_on) xe2 = x[(a + 1) * n + c0 + 2];
        }
        // Every lane takes part in both shuffles.
        const T vw = sh.from_west(xc.hi), ve = sh.from_east(xc.lo);
        const T XW = sh.west ? vw : xw, XE = sh.east ? ve : xe;
        const T r0 =
            cur.b0 - apply_combine(cur.m, cur.w0, north.lo, cur.s0, xc.lo, xc.hi, XW, xn.lo, xp.lo, two, west_on, hN, hS);
        const T r1 =
            cur.b1 - apply_combine(cur.e1, cur.m, north.hi, cur.s1, xc.hi, XE, xc.lo, xn.hi, xp.hi, east_on, true, hN, hS);
        if (!(a & 1)) acc = T(0);
        acc += r0;
        if (two) acc += r1;
        if ((a & 1) || !hN) bc[(a >> 1) * m + J] = acc;
        xp = xc;
        xc = xn;
        xn = xn2;
        cur = nx;
        north = north2;
        xw = xw2;
        xe = xe2;
    }
}

// The red half of a GS sweep over rows [i0, i1) of column pair k as a
// vertical walk (level.hpp:145-161), the black cells holding the values the
// sweep reads (final for the post-smoothing's second colour, as stored for a
// repeated visit's first): each row's red cell from b, its faces
// and its four black neighbours -- the pair's other cell, the rows above and
// below (held in registers as the walk goes down), and the cell beside the
// pair (from the neighbouring lane through `sh`, else loaded).  A red cell
// reads no red cell, so items need no order among themselves.  gs_at's
// operations in its order: the same bits.
template <class T, bool kPairs, class Share = NoShare, template <class> class Sp = FlatPtr>
DARCY_HD __forceinline__ void red_walk_item(const LvP<T, Sp>& L, int k, int i0, int i1, Share sh = Share{}) {
    const int n = L.n, c0 = 2 * k;
    const bool two = c0 + 1 < n, west_on = c0 > 0, east_on = c0 + 2 < n;
    const int c1 = two ? c0 + 1 : c0;
    const auto x = L.x;
    const auto ny = L.ny;
    auto xpair = [&](int a) { return walk_pair<T, kPairs>(x, n, c0, c1, a); };
    // the black cell beside the red one where no lane shares it: west of c0
    // on even rows, east of c1 on odd rows
    auto side = [&](int a) -> T {
        if (a & 1) return (!sh.east && east_on) ? x[a * n + c1 + 1] : T(0);
        return (!sh.west && west_on) ? x[a * n + c0 - 1] : T(0);
    };
    Pair2<T> xp{T(0), T(0)}, xc = xpair(i0), xn{T(0), T(0)};
    if (i0 > 0) xp = xpair(i0 - 1);
    if (i0 + 1 < n) xn = xpair(i0 + 1);
    WalkRow<T> cur = walk_row<T, kPairs>(L, c0, c1, i0);
    Pair2<T> north = walk_pair<T, kPairs>(ny, n, c0, c1, i0 + 1);
    T xs = side(i0);
    for (int a = i0; a < i1; ++a) {
        const bool hN = a + 1 < n, hS = a > 0, odd = (a & 1) != 0, more = a + 1 < i1;
        Pair2<T> xn2 = xn, north2 = north;
        WalkRow<T> nx = cur;
        T xs2 = T(0);
        if (more) {
            nx = walk_row<T, kPairs>(L, c0, c1, a + 1, north);  // north: ny of row a + 1
            north2 = walk_pair<T, kPairs>(ny, n, c0, c1, a + 2);
            if (a + 2 < n) xn2 = xpair(a + 2);
            xs2 = side(a + 1);
        }
        // Every lane takes part in both shuffles.
        const T vw = sh.from_west(xc.hi), ve = sh.from_east(xc.lo);
        if (!odd) {
            const T XW = sh.west ? vw : xs;
            x[a * n + c0] =
                gs_combine(cur.m, cur.w0, north.lo, cur.s0, cur.b0, xc.hi, XW, xn.lo, xp.lo, two, west_on, hN, hS);
        } else if (two) {
            const T XE = sh.east ? ve : xs;
            x[a * n + c1] =
                gs_combine(cur.e1, cur.m, north.hi, cur.s1, cur.b1, XE, xc.lo, xn.hi, xp.hi, east_on, true, hN, hS);
        }
        xp = xc;
        xc = xn;
        xn = xn2;
        cur = nx;
        north = north2;
        xs = xs2;
    }
}

// The black half of the first post-smoothing sweep fused with the
// prolongation (black_with_prolong's values) over rows [i0, i1) of column
// pair k as a vertical walk: each row's black cell from b, its faces and its
// four red neighbours read as x + x_c of their aggregates -- the pair's other
// cell, the rows above and below, and the cell beside the pair (from the
// neighbouring lane, which forms the same value for its own red cell, else
// loaded).  Black cells are written, red cells and x_c only read: items need
// no order among themselves.  gs_at's operations with XPlusC, in its order.
// kProlong false: the black half of a plain GS sweep (red cells final, read
// as stored; x_c unused), the second colour of a repeated visit's
// pre-smoothing.
template <class T, bool kPairs, bool kProlong = true, class Share = NoShare, template <class> class Sp = FlatPtr,
          class PC = const T*>
DARCY_HD __forceinline__ void black_prolong_walk_item(const LvP<T, Sp>& L, PC xcc, int k, int i0, int i1,
                                                      Share sh = Share{}) {
    const int n = L.n, m = (n + 1) >> 1, c0 = 2 * k;
    const bool two = c0 + 1 < n, west_on = c0 > 0, east_on = c0 + 2 < n;
    const int c1 = two ? c0 + 1 : c0;
    const auto x = L.x;
    const auto ny = L.ny;
    auto xpair = [&](int a) { return walk_pair<T, kPairs>(x, n, c0, c1, a); };
    auto coarse = [&](int a) -> T {
        if constexpr (kProlong)
            return xcc[(a >> 1) * m + k];
        else
            return T(0);
    };
    // x + x_c with the prolongation, x without it
    auto plus = [](T v, T c) -> T {
        if constexpr (kProlong)
            return T(v + c);
        else
            return v;
    };
    // the red cell beside the black one, prolongated, where no lane shares
    // it: east of c1 on even rows, west of c0 on odd rows
    auto side = [&](int a) -> T {
        if (a & 1) {
            if (sh.west || !west_on) return T(0);
            if constexpr (kProlong)
                return T(x[a * n + c0 - 1] + xcc[(a >> 1) * m + k - 1]);
            else
                return x[a * n + c0 - 1];
        }
        if (sh.east || !east_on) return T(0);
        if constexpr (kProlong)
            return T(x[a * n + c1 + 1] + xcc[(a >> 1) * m + k + 1]);
        else
            return x[a * n + c1 + 1];
    };
    Pair2<T> xp{T(0), T(0)}, xc = xpair(i0), xn{T(0), T(0)};
    T cp = T(0), ca = coarse(i0), cn = T(0);  // x_c of the aggregates of rows a - 1, a, a + 1
    if (i0 > 0) {
        xp = xpair(i0 - 1);
        cp = coarse(i0 - 1);
    }
    if (i0 + 1 < n) {
        xn = xpair(i0 + 1);
        cn = coarse(i0 + 1);
    }
    WalkRow<T> cur = walk_row<T, kPairs>(L, c0, c1, i0);
    Pair2<T> north = walk_pair<T, kPairs>(ny, n, c0, c1, i0 + 1);
    T xs = side(i0);
    for (int a = i0; a < i1; ++a) {
        const bool hN = a + 1 < n, hS = a > 0, odd = (a & 1) != 0, more = a + 1 < i1;
        Pair2<T> xn2 = xn, north2 = north;
        WalkRow<T> nx = cur;
        T cn2 = cn, xs2 = T(0);
        if (more) {
            nx = walk_row<T, kPairs>(L, c0, c1, a + 1, north);  // north: ny of row a + 1
            north2 = walk_pair<T, kPairs>(ny, n, c0, c1, a + 2);
            if (a + 2 < n) {
                xn2 = xpair(a + 2);
                cn2 = coarse(a + 2);
            }
            xs2 = side(a + 1);
        }
        // this pair's red cell on row a, prolongated (c0 on even rows, c1 on
        // odd ones); every lane forms it and takes part in both shuffles
        const T pv = plus(odd ? xc.hi : xc.lo, ca);
        const T vw = sh.from_west(pv), ve = sh.from_east(pv);
        if (!odd) {
            if (two) {
                const T XE = sh.east ? ve : xs;
                x[a * n + c1] = gs_combine(cur.e1, cur.m, north.hi, cur.s1, cur.b1, XE, pv, plus(xn.hi, cn),
                                           plus(xp.hi, cp), east_on, true, hN, hS);
            }
        } else {
            const T XW = sh.west ? vw : xs;
            x[a * n + c0] = gs_combine(cur.m, cur.w0, north.lo, cur.s0, cur.b0, pv, XW, plus(xn.lo, cn),
                                       plus(xp.lo, cp), two, west_on, hN, hS);
        }
        xp = xc;
        xc = xn;
        xn = xn2;
        cp = ca;
        ca = cn;
        cn = cn2;
        cur = nx;
        north = north2;
        xs = xs2;
    }
}

struct LaneShare {
    bool east, west;
    template <class T>
    DARCY_INLINE T from_east(T v) const {
        return __shfl_down(v, 1, kWave);
    }
    template <class T>
    DARCY_INLINE T from_west(T v) const {
        return __shfl_up(v, 1, kWave);
    }
};

// Rows per walk item for `rows` x `cols` items' worth of work over `nthr`
// threads: about one item per thread, at most `most` rows.
DARCY_HD __forceinline__ int walk_strip(int rows, int cols, int nthr, int most) {
    const long long per = (static_cast<long long>(rows) * cols) / nthr;
    return per < 1 ? 1 : (per > most ? most : static_cast<int>(per));
}

// Pre-smoothing from zero of rows [ra, rb) as walks of `strip` rows.  Every
// item re-forms the red row above it and the red row below it from b and the
// faces, so the items of a range depend on nothing outside it.
template <class T, template <class> class Sp>
DARCY_INLINE void presmooth_walk_rows(const LvP<T, Sp>& L, int ra, int rb, int strip, Exec e) {
    const int n = L.n, h = (n + 1) >> 1;
    const int strips = (rb - ra + strip - 1) / strip, total = strips * h;
    const int lane = threadIdx.x & (kWave - 1);
    for (int w = e.tid; w < total; w += e.nthr) {
        const int st = w / h, k = w - st * h, i0 = ra + st * strip, i1 = i0 + strip < rb ? i0 + strip : rb;
        // Consecutive lanes hold consecutive items, so lane l + 1 walks pair
        // k + 1 of the same strip when k + 1 < h (and l - 1 pair k - 1 when k > 0).
        const LaneShare sh{lane + 1 < kWave && k + 1 < h, lane > 0 && k > 0};
        if (n & 1)
            presmooth_walk_item<T, false>(L, k, i0, i1, sh);
        else
            presmooth_walk_item<T, true>(L, k, i0, i1, sh);
    }
}

// The walk over a level: rows per item chosen so that the items about fill
// the launch (at most kWalkRows).
template <class T, template <class> class Sp>
DARCY_INLINE void presmooth_walk(const LvP<T, Sp>& L, Exec e) {
    const int n = L.n;
    presmooth_walk_rows(L, 0, n, walk_strip(n, (n + 1) >> 1, e.nthr, kWalkRows), e);
}

// Residual and restriction of coarse rows [Ia, Ib) as walks of `strip`
// coarse rows; x of fine rows 2 Ia - 1 .. 2 Ib must be final.
template <class T, template <class> class Sp, class BP>
DARCY_INLINE void restrict_walk_rows(const LvP<T, Sp>& L, BP bc, int Ia, int Ib, int strip, Exec e) {
    const int m = (L.n + 1) >> 1;
    const int strips = (Ib - Ia + strip - 1) / strip, total = strips * m;
    const int lane = threadIdx.x & (kWave - 1);
    for (int w = e.tid; w < total; w += e.nthr) {
        const int st = w / m, J = w - st * m, I0 = Ia + st * strip, I1 = I0 + strip < Ib ? I0 + strip : Ib;
        const LaneShare sh{lane + 1 < kWave && J + 1 < m, lane > 0 && J > 0};
        if (L.n & 1)
            restrict_walk_item<T, false>(L, bc, J, I0, I1, sh);
        else
            restrict_walk_item<T, true>(L, bc, J, I0, I1, sh);
    }
}

// The restriction walk over a level, coarse rows per item chosen as in
// presmooth_walk (at most kWalkRows / 2).
template <class T, template <class> class Sp, class BP>
DARCY_INLINE void restrict_walk(const LvP<T, Sp>& L, BP bc, Exec e) {
    const int m = (L.n + 1) >> 1;
    restrict_walk_rows(L, bc, 0, m, walk_strip(m, m, e.nthr, kWalkRows / 2), e);
}

// The red half of a GS sweep over rows [ra, rb) as walks of `strip` rows;
// the black cells of rows ra - 1 .. rb must hold the values the sweep reads.
template <class T, template <class> class Sp>
DARCY_INLINE void red_walk_rows(const LvP<T, Sp>& L, int ra, int rb, int strip, Exec e) {
    const int n = L.n, h = (n + 1) >> 1;
    const int strips = (rb - ra + strip - 1) / strip, total = strips * h;
    const int lane = threadIdx.x & (kWave - 1);
    for (int w = e.tid; w < total; w += e.nthr) {
        const int st = w / h, k = w - st * h, i0 = ra + st * strip, i1 = i0 + strip < rb ? i0 + strip : rb;
        const LaneShare sh{lane + 1 < kWave && k + 1 < h, lane > 0 && k > 0};
        if (n & 1)
            red_walk_item<T, false>(L, k, i0, i1, sh);
        else
            red_walk_item<T, true>(L, k, i0, i1, sh);
    }
}

// The black half of the first post-smoothing sweep with the prolongation
// over rows [ra, rb) as walks of `strip` rows (x_c: the coarse correction,
// m = (n + 1) / 2 per row); kProlong false: of a plain GS sweep (x_c unused).
template <class T, bool kProlong = true, template <class> class Sp, class PC>
DARCY_INLINE void black_prolong_walk_rows(const LvP<T, Sp>& L, PC xcc, int ra, int rb, int strip, Exec e) {
    const int n = L.n, h = (n + 1) >> 1;
    const int strips = (rb - ra + strip - 1) / strip, total = strips * h;
    const int lane = threadIdx.x & (kWave - 1);
    for (int w = e.tid; w < total; w += e.nthr) {
        const int st = w / h, k = w - st * h, i0 = ra + st * strip, i1 = i0 + strip < rb ? i0 + strip : rb;
        const LaneShare sh{lane + 1 < kWave && k + 1 < h, lane > 0 && k > 0};
        if (n & 1)
            black_prolong_walk_item<T, false, kProlong>(L, xcc, k, i0, i1, sh);
        else
            black_prolong_walk_item<T, true, kProlong>(L, xcc, k, i0, i1, sh);
    }
}

// x and the coarse correction x_c of its aggregate, read together; XPlusC
// forms the prolongated value x + x_c.
template <class T>
struct XC {
    T x, c;
};
template <class T, class PX, class PC>
struct Prolonged {
    static constexpr bool kRow3 = IsGlobalPtr<std::remove_cv_t<PX>>::value && IsGlobalPtr<std::remove_cv_t<PC>>::value;
    PX x;
    PC xc;
    int n, m;
    DARCY_INLINE XC<T> operator()(int a, int c) const { return XC<T>{x[a * n + c], xc[(a >> 1) * m + (c >> 1)]}; }
    // West and east neighbours of (a, c): x as one access, and their
    // aggregates' x_c, always adjacent ((c - 1) >> 1 and (c + 1) >> 1), as
    // another.  Off-grid values are read and not used.
    DARCY_INLINE Row3<XC<T>> row3(int a, int c) const {
        const Row3<T> v = load_row3(x + a * n + c);
        const Row3<T> q = load_two(xc + (a >> 1) * m + ((c - 1) >> 1));
        return Row3<XC<T>>{XC<T>{v.w, q.w}, XC<T>{v.c, q.c}, XC<T>{v.e, q.e}};
    }
};
struct XPlusC {
    template <class T>
    DARCY_INLINE T operator()(const XC<T>& v) const {
        return v.x + v.c;
    }
};

// First post-smoothing sweep (black) fused with the prolongation
// x += P x_c (level.hpp:230-237): red neighbours are read as x + x_c of their
// aggregate, the value the reference stores before smoothing.  Red cells keep
// their unprolongated values: the red sweep that follows overwrites them
// without reading them.
template <class T, template <class> class Sp, class PC>
DARCY_INLINE void black_with_prolong(const LvP<T, Sp>& L, PC xc, Exec e) {
    const int n = L.n, h = (n + 1) >> 1, total = n * h, m = (n + 1) >> 1;
    const auto x = L.x;
    const Prolonged<T, decltype(x), PC> pr{x, xc, n, m};
    for (int q0 = e.tid; q0 < total; q0 += 4 * e.nthr) {
        T v[4];
        int kk[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            int i, j;
            const int q = q0 + u * e.nthr;
            kk[u] = -1;
            if (q < total && colour_cell(q, n, h, 1, i, j)) {
                kk[u] = i * n + j;
                v[u] = gs_at(L, i, j, L.b[kk[u]], pr, XPlusC{});
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (kk[u] >= 0) x[kk[u]] = v[u];
    }
}

// b_c = P^T (b - A x) over each <= 2x2 aggregate (level.hpp:122-137 then
// :214-227, fused: the fine residual is never stored).
template <class T, template <class> class Sp>
DARCY_INLINE T restricted_residual(const LvP<T, Sp>& L, int q, int m);

template <class T, template <class> class Sp, class BP>
DARCY_INLINE void residual_restrict(const LvP<T, Sp>& L, BP bc, Exec e) {
    const int m = (L.n + 1) >> 1, total = m * m;
    for (int q = e.tid; q < total; q += e.nthr) bc[q] = restricted_residual(L, q, m);
}

// Coarsest solve (multigrid.hpp:129-144): exact for 1x1, otherwise 8 + 8
// symmetric sweeps (reached only when the 32-level cap stops coarsening).
template <class T, template <class> class Sp, class S>
DARCY_INLINE void coarsest_solve(const LvP<T, Sp>& L, bool fresh, Exec e, S sync) {
    if (L.n == 1) {
        if (e.tid == 0) L.x[0] = L.b[0] * L.id[0];
        sync();
        return;
    }
    if (fresh) {
        for (int k = e.tid; k < L.n * L.n; k += e.nthr) L.x[k] = T(0);
        sync();
    }
    for (int s = 0; s < 8; ++s) {
        gs_colour(L, 0, e);
        sync();
        gs_colour(L, 1, e);
        sync();
    }
    for (int s = 0; s < 8; ++s) {
        gs_colour(L, 1, e);
        sync();
        gs_colour(L, 0, e);
        sync();
    }
}


// Going down through level l: pre-smoothing (red, black) then the restricted
// residual into the next level's b (multigrid.hpp:107-115).
template <class T, template <class> class Sp, class BP, class S>
DARCY_INLINE void visit_down(const LvP<T, Sp>& L, BP bc, bool fresh, const CycleOpts& o, Exec e, S sync) {
    for (int s = 0; s < o.pre; ++s) {
        if (s == 0 && fresh) {
            presmooth_from_zero(L, e);
        } else {
            gs_colour(L, 0, e);
            sync();
            gs_colour(L, 1, e);
        }
        sync();
    }
    residual_restrict(L, bc, e);
    sync();
}

// Coming up through level l: prolongation and post-smoothing (black, red)
// (multigrid.hpp:120-126).
template <class T, template <class> class Sp, class PC, class S>
DARCY_INLINE void visit_up(const LvP<T, Sp>& L, PC xc, const CycleOpts& o, Exec e, S sync) {
    black_with_prolong(L, xc, e);
    sync();
    gs_colour(L, 0, e);
    sync();
    for (int s = 1; s < o.post; ++s) {
        gs_colour(L, 1, e);
        sync();
        gs_colour(L, 0, e);
        sync();
    }
}

// The cycle of multigrid.hpp:97-127 from level l0 to the coarsest level
// L-1, run iteratively by the threads of `e` with barrier `sync`; `view(l)`
// gives the LvP of level l.  Bit k of `done` marks that level k has had its
// first coarse-grid visit (gamma <= 2).
template <class T, class S, class V>
DARCY_INLINE void cycle_levels(int l0, int L, bool fresh0, const CycleOpts& o, Exec e, S sync, V&& view) {
    uint32_t done = 0;
    int l = l0;
    bool fresh = fresh0;
    for (;;) {
        while (l + 1 < L) {
            visit_down(view(l), view(l + 1).b, fresh, o, e, sync);
            done &= ~(1u << l);
            ++l;
            fresh = true;
        }
        coarsest_solve(view(l), fresh, e, sync);
        for (;;) {
            if (l == l0) return;
            --l;
            if (o.gamma == 2 && !((done >> l) & 1u)) {
                done |= 1u << l;
                ++l;
                fresh = false;
                break;
            }
            visit_up(view(l), static_cast<const T*>(view(l + 1).x), o, e, sync);
        }
    }
}

// ------------------------------------------- row-pipelined (streamed) visits
// For the levels in global memory a visit is one pass over the rows instead
// of one pass per phase: the workgroup walks down the grid R rows at a time,
// the black update trails the red one by a row and the residual/restriction
// trails the black one, so every row is finished while it is still in cache.
// Values and operation order are those of the phase-by-phase visit: a black
// cell is updated only after both red neighbours rows are, a residual only
// after its whole 3-row neighbourhood is final.

// Rows per pipeline step: about `cells` cells, even, at least 2.  A step's
// rows are read two or three times (red, black, residual), so they should
// stay in L2 between those reads: the persistent batch kernel, 128
// workgroups to each 4 MB L2, takes kStepCellsShared-cell steps unless its
// launch sets another size (a tuning value; 1000 x 256^2: FP32 cycle
// 100.1 -> 99.2 ms, FP64 141.1 -> 135.4 ms at 1024 against 2048); the tail
// workgroup of the grid path, alone on its CU, takes 4096 (4096^2 FP64
// 352 -> 345 ms).  The step size does not change any result.
constexpr int kStepCellsShared = 1024;
constexpr int kStepCellsAlone = 4096;
DARCY_HD __forceinline__ int rows_per_step(int n, int cells) {
    int r = cells / n;
    r &= ~1;
    return r < 2 ? 2 : r;
}

// x[k] = value(i, j) for the cells of colour `colour` in rows [ra, rb); four
// cells per thread and step, loads before stores.
template <class P, class F>
DARCY_INLINE void update_rows(P x, int n, int colour, int ra, int rb, Exec e, F&& value) {
    using T = pointee_t<P>;
    const int h = (n + 1) >> 1, qa = ra * h, qb = rb * h;
    for (int q0 = qa + e.tid; q0 < qb; q0 += 4 * e.nthr) {
        T v[4];
        int kk[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int q = q0 + u * e.nthr;
            kk[u] = -1;
            if (q < qb) {
                const int i = q / h;
                const int j = 2 * (q - i * h) + ((i + colour) & 1);
                if (j < n) {
                    kk[u] = i * n + j;
                    v[u] = value(i, j);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (kk[u] >= 0) x[kk[u]] = v[u];
    }
}

// Restricted residual of coarse cell q: the fine residuals of its <= 2x2
// aggregate summed in the order (i0,j0),(i0,j1),(i1,j0),(i1,j1)
// (level.hpp:122-137, :214-227).  The aggregate is unrolled with predicates
// so that all loads issue before the sums.
template <class T, template <class> class Sp>
DARCY_INLINE T restricted_residual(const LvP<T, Sp>& L, int q, int m) {
    const int n = L.n, I = q / m, J = q - I * m;
    const auto x = L.x;
    T r[4];
    bool ok[4];
#pragma unroll
    for (int d = 0; d < 4; ++d) {
        const int i = 2 * I + (d >> 1), j = 2 * J + (d & 1);
        ok[d] = i < n && j < n;
        if (ok[d]) r[d] = L.b[i * n + j] - apply_at(n, L.ex, L.ny, i, j, XFlat<decltype(x)>{x});
    }
    T s = T(0);
#pragma unroll
    for (int d = 0; d < 4; ++d)
        if (ok[d]) s += r[d];
    return s;
}

// Restricted residual of the coarse rows [Ia, Ib), two coarse cells per
// thread and step.
template <class T, template <class> class Sp, class BP>
DARCY_INLINE void restrict_rows(const LvP<T, Sp>& L, BP bc, int Ia, int Ib, Exec e) {
    const int n = L.n, m = (n + 1) >> 1, qb = Ib * m;
    for (int q0 = Ia * m + e.tid; q0 < qb; q0 += 2 * e.nthr) {
        T v[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int q = q0 + u * e.nthr;
            if (q < qb) v[u] = restricted_residual(L, q, m);
        }
#pragma unroll
        for (int u = 0; u < 2; ++u)
            if (q0 + u * e.nthr < qb) bc[q0 + u * e.nthr] = v[u];
    }
    (void)n;
}

// A fresh down visit (pre-smoothing from x = 0, residual, restriction;
// multigrid.hpp:107-115 with pre_sweeps = 1) by one workgroup as walks,
// R rows per step: the pre-smoothing walks over the step's rows, a barrier,
// then the restriction walks over the coarse rows whose three fine rows
// around them are final.  The walks hold a column pair's rows in registers
// and take the neighbouring pair's values by lane shuffles, so a cell costs
// a few loads where the per-cell phases of down_stream re-read every
// neighbour.  The restriction of step t reads rows below r1 only, and the
// pre-smoothing of step t + 1 reads no x, so no barrier separates them.
// Same values and operation order as down_stream (the walks restate
// presmooth_from_zero and restricted_residual).
template <class T, template <class> class Sp, class BP>
DARCY_INLINE void down_walk(const LvP<T, Sp>& L, BP bc, int step_cells, Exec e) {
    const int n = L.n, m = (n + 1) >> 1, R = rows_per_step(n, step_cells);
    int coarse_done = 0;
    for (int r0 = 0; r0 < n; r0 += R) {
        const int r1 = r0 + R < n ? r0 + R : n;
        presmooth_walk_rows(L, r0, r1, walk_strip(r1 - r0, m, e.nthr, kWalkRows), e);
        __syncthreads();
        // rows < r1 are final; coarse row I needs fine rows 2I - 1 .. 2I + 2
        const int coarse_hi = r1 == n ? m : (r1 - 1) / 2;
        if (coarse_hi > coarse_done)
            restrict_walk_rows(L, bc, coarse_done, coarse_hi, walk_strip(coarse_hi - coarse_done, m, e.nthr, kWalkRows / 2), e);
        coarse_done = coarse_hi;
    }
    __syncthreads();
}

// One pre-smoothing sweep (red, black) and the restricted residual in a
// single pass (multigrid.hpp:107-115 with pre_sweeps = 1).
template <class T, template <class> class Sp, class BP>
DARCY_INLINE void down_stream(const LvP<T, Sp>& L, BP bc, bool fresh, int step_cells, Exec e) {
    const int n = L.n, m = (n + 1) >> 1, R = rows_per_step(n, step_cells);
    const auto x = L.x;
    const auto b = L.b;
    const XRow<decltype(x)> xv{x, n};
    int coarse_done = 0;
    for (int r0 = 0;; r0 += R) {
        const int red_hi = r0 + R < n ? r0 + R : n;
        if (r0 < n) {
            if (fresh)  // the red sweep from zero: b * inv_d (each neighbour term is e * 0)
                update_rows(x, n, 0, r0, red_hi, e,
                            [&](int i, int j) { return b[i * n + j] * inv_diag_at(L, i, j); });
            else if constexpr (sizeof(T) == 4)  // FP32: walks, as in up_stream
                red_walk_rows(L, r0, red_hi, walk_strip(red_hi - r0, m, e.nthr, kWalkRows), e);
            else
                update_rows(x, n, 0, r0, red_hi, e, [&](int i, int j) { return gs_at(L, i, j, b[i * n + j], xv); });
        }
        __syncthreads();
        // black row i needs red rows i-1 and i+1
        const int black_lo = r0 > 0 ? r0 - 1 : 0;
        const int black_hi = red_hi == n ? n : red_hi - 1;
        if constexpr (sizeof(T) == 4) {
            if (black_hi > black_lo)
                black_prolong_walk_rows<T, false>(L, static_cast<const T*>(nullptr), black_lo, black_hi,
                                                  walk_strip(black_hi - black_lo, m, e.nthr, kWalkRows), e);
        } else {
            update_rows(x, n, 1, black_lo, black_hi, e, [&](int i, int j) { return gs_at(L, i, j, b[i * n + j], xv); });
        }
        __syncthreads();
        // residual row i needs final rows i-1..i+1; coarse row I needs fine rows 2I, 2I+1
        const int ready = black_hi == n ? n : black_hi - 1;
        const int coarse_hi = ready == n ? m : ready / 2;
        // FP32: the restriction walks (a column pair's rows held in
        // registers: a few loads per cell where the per-cell residuals
        // re-read every neighbour); FP64 keeps the per-cell pass, which
        // needs fewer registers.  Same values either way.
        if constexpr (sizeof(T) == 4) {
            if (coarse_hi > coarse_done)
                restrict_walk_rows(L, bc, coarse_done, coarse_hi,
                                   walk_strip(coarse_hi - coarse_done, m, e.nthr, kWalkRows / 2), e);
        } else {
            restrict_rows(L, bc, coarse_done, coarse_hi, e);
        }
        coarse_done = coarse_hi;
        if (coarse_done == m) break;
    }
    __syncthreads();
}

// Prolongation and one post-smoothing sweep (black, red) in a single pass
// (multigrid.hpp:120-126 with post_sweeps = 1).  Black cells read their red
// neighbours as x + x_c (the prolongated values); the red update trails the
// black one by a row.
template <class T, template <class> class Sp, class PC>
DARCY_INLINE void up_stream(const LvP<T, Sp>& L, PC xc, int step_cells, Exec e) {
    const int n = L.n, m = (n + 1) >> 1, R = rows_per_step(n, step_cells);
    const auto x = L.x;
    const auto b = L.b;
    const XRow<decltype(x)> xv{x, n};
    const Prolonged<T, decltype(x), PC> pr{x, xc, n, m};
    for (int r0 = 0;; r0 += R) {
        const int black_hi = r0 + R < n ? r0 + R : n;
        if (r0 < n) {
            // FP32: walks, as for the red half below
            if constexpr (sizeof(T) == 4)
                black_prolong_walk_rows(L, xc, r0, black_hi, walk_strip(black_hi - r0, m, e.nthr, kWalkRows), e);
            else
                update_rows(x, n, 1, r0, black_hi, e,
                            [&](int i, int j) { return gs_at(L, i, j, b[i * n + j], pr, XPlusC{}); });
        }
        __syncthreads();
        // red row i needs black rows i-1 and i+1; red rows of step t never
        // meet the black rows of step t+1, so no barrier follows.
        const int red_lo = r0 > 0 ? r0 - 1 : 0;
        const int red_hi = black_hi == n ? n : black_hi - 1;
        // FP32: the red half as walks (the black rows around a red cell held
        // in registers); FP64 keeps the per-cell pass, which needs fewer
        // registers.  Same values either way.
        if constexpr (sizeof(T) == 4) {
            if (red_hi > red_lo)
                red_walk_rows(L, red_lo, red_hi, walk_strip(red_hi - red_lo, m, e.nthr, kWalkRows), e);
        } else {
            update_rows(x, n, 0, red_lo, red_hi, e, [&](int i, int j) { return gs_at(L, i, j, b[i * n + j], xv); });
        }
        if (red_hi == n) break;
    }
    __syncthreads();
}

// ---------------------------------------------------------- LDS hierarchy

// LDS pool of the coarse levels: operators and vectors of every level with
// side <= NL.  Sized for the worst case (sides NL, NL/2, ..., 1).
template <class T, int NL>
struct LdsPool {
    static constexpr int elems_of(int s) { return s > 0 ? 2 * s * (s + 1) + 3 * s * s + 20 : 0; }
    static constexpr int total_elems() {
        int t = 0;
        for (int s = NL; s >= 1; s = (s == 1) ? 0 : (s + 1) / 2) t += elems_of(s);
        return t;
    }
    static constexpr int kElems = total_elems();
};

// Offsets (in elements of T) of the LDS levels, filled by one thread.
struct LdsTable {
    int first;  // first level held in LDS (L when none)
    int wave;   // first level run by one wave
    int off[kMaxLevels][5];
};

DARCY_INLINE int align4(int v) { return (v + 3) & ~3; }

// Levels from `from` on whose side is <= NL go to LDS.
template <class T, int NL>
DARCY_INLINE void plan_lds(const Hier& h, LdsTable& t, int from = 1) {
    int first = h.L;
    for (int l = from; l < h.L; ++l)
        if (h.n[l] <= NL) {
            first = l;
            break;
        }
    t.first = first;
    int wave = h.L;
    for (int l = first; l < h.L; ++l)
        if (h.n[l] <= kWaveTail) {
            wave = l;
            break;
        }
    t.wave = wave;
    int o = 0;
    for (int l = first; l < h.L; ++l) {
        const int s = h.n[l];
        t.off[l][0] = o;
        o = align4(o + s * (s + 1));
        t.off[l][1] = o;
        o = align4(o + s * (s + 1));
        t.off[l][2] = o;
        o = align4(o + s * s);
        t.off[l][3] = o;
        o = align4(o + s * s);
        t.off[l][4] = o;
        o = align4(o + s * s);
    }
}

}  // namespace darcy_hip
