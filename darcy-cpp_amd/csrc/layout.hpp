// Device-memory layout of one solve "slot": the FP64 CG vectors of the fine
// grid and the whole multigrid hierarchy, carved out of one allocation.
//
// Shared by host code (sizing, the grid-path launches) and device code (the
// persistent batch kernel computes the same table for each problem it takes).
//
// Operator storage, per level of side n:
//   ex   n x (n+1)  face between cells (i, j-1) and (i, j) at ex[i*(n+1)+j];
//                   ex[i*(n+1)] is the west boundary face of (i, 0) and
//                   ex[i*(n+1)+n] the east boundary face of (i, n-1).
//   ny   (n+1) x n  face between cells (i-1, j) and (i, j) at ny[i*n+j];
//                   rows 0 and n hold the south/north boundary faces.
//   id   n x n      1 / diagonal, computed in FP64.
// Each interior face is stored once: the reference's six arrays
// (include/darcy/level.hpp:253) hold it twice, as e of the left cell and w of
// the right one, and the two copies are bit-identical (harmonic_mean is
// commutative bit for bit; coarse faces sum the same fine faces in the same
// order).  The diagonal ((e + w) + n) + s is recomputed from the faces where a
// kernel needs it, which reproduces the stored value of the reference exactly.
#pragma once

#include <stddef.h>
#include <stdint.h>

#ifndef DARCY_HD
#if defined(__HIPCC__)
#define DARCY_HD __host__ __device__
#else
#define DARCY_HD
#endif
#endif

namespace darcy_hip {

constexpr int kMaxLevels = 32;   // multigrid.hpp:46
constexpr int kChunk = 256;      // cells per canonical reduction chunk
constexpr size_t kAlign = 256;   // byte alignment of every array

// Options of the multigrid cycle (multigrid.hpp:29-47).
struct CycleOpts {
    int pre;
    int post;
    int gamma;  // 1 = V-cycle, 2 = W-cycle
};

DARCY_HD inline size_t align_up(size_t v) { return (v + kAlign - 1) & ~(kAlign - 1); }

// Level count for side n: coarsen with (n + 1) / 2 while the level is wider
// than min_coarse (and than 1) and fewer than max_levels exist
// (multigrid.hpp:55-71, level.hpp:164).  max_levels is capped at kMaxLevels,
// which no side below 2^31 reaches.
DARCY_HD inline int level_count(int n, int min_coarse = 1, int max_levels = kMaxLevels) {
    const int cap = max_levels < kMaxLevels ? max_levels : kMaxLevels;
    int l = 1;
    while (l < cap && n > min_coarse && n > 1) {
        n = (n + 1) / 2;
        ++l;
    }
    return l;
}

DARCY_HD inline int chunks_of(int64_t cells) { return static_cast<int>((cells + kChunk - 1) / kChunk); }

// Pointers of one slot.  `mixed` adds FP32 copies of the operators and FP32
// cycle vectors; otherwise the cycle runs on the FP64 arrays, with b_0
// aliasing r and x_0 aliasing z.
struct Hier {
    int L;
    int n[kMaxLevels];
    double* ex[kMaxLevels];
    double* ny[kMaxLevels];
    double* id[kMaxLevels];
    float* ex32[kMaxLevels];
    float* ny32[kMaxLevels];
    float* id32[kMaxLevels];
    void* x[kMaxLevels];
    void* b[kMaxLevels];
    // fine-grid CG vectors (FP64)
    double* u;
    double* r;
    double* p;
    double* ap;
    double* z;
    // canonical-reduction partials: three arrays of chunks_of(n0^2) doubles
    double* part[3];
};

// Fill `*h` for a slot starting at `base` (nullptr to size only); returns the
// slot's size in bytes.  `h` must not be null (device code passes LDS).
// min_coarse / max_levels are the hierarchy options of level_count.
DARCY_HD inline size_t plan_slot(char* base, int n0, bool mixed, Hier* h, int min_coarse = 1,
                                 int max_levels = kMaxLevels) {
    size_t off = 0;
    auto take = [&](size_t bytes) -> char* {
        char* p = base ? base + off : nullptr;
        off += align_up(bytes);
        return p;
    };
    const size_t N0 = static_cast<size_t>(n0) * n0;
    Hier& t = *h;
    t.L = level_count(n0, min_coarse, max_levels);
    // Guard bytes before the first array (and after the last): the kernels
    // read a row's west, own and east values as one access, one element
    // beyond an array's ends at its first and last cell (the value is then
    // unused).  Between arrays the alignment padding or the next array is
    // read.
    take(kAlign);
    t.u = reinterpret_cast<double*>(take(N0 * 8));
    t.r = reinterpret_cast<double*>(take(N0 * 8));
    t.p = reinterpret_cast<double*>(take(N0 * 8));
    t.ap = reinterpret_cast<double*>(take(N0 * 8));
    t.z = reinterpret_cast<double*>(take(N0 * 8));
    for (int k = 0; k < 3; ++k) t.part[k] = reinterpret_cast<double*>(take(static_cast<size_t>(chunks_of(N0)) * 8));
    int n = n0;
    for (int l = 0; l < kMaxLevels; ++l) {
        if (l >= t.L) {
            t.n[l] = 0;
            t.ex[l] = t.ny[l] = t.id[l] = nullptr;
            t.ex32[l] = t.ny32[l] = t.id32[l] = nullptr;
            t.x[l] = t.b[l] = nullptr;
            continue;
        }
        const size_t c = static_cast<size_t>(n) * n, f = static_cast<size_t>(n) * (n + 1);
        t.n[l] = n;
        t.ex[l] = reinterpret_cast<double*>(take(f * 8));
        t.ny[l] = reinterpret_cast<double*>(take(f * 8));
        t.id[l] = reinterpret_cast<double*>(take(c * 8));
        if (mixed) {
            t.ex32[l] = reinterpret_cast<float*>(take(f * 4));
            t.ny32[l] = reinterpret_cast<float*>(take(f * 4));
            t.id32[l] = reinterpret_cast<float*>(take(c * 4));
            t.x[l] = take(c * 4);
            t.b[l] = take(c * 4);
        } else {
            t.ex32[l] = t.ny32[l] = t.id32[l] = nullptr;
            if (l == 0) {
                t.x[l] = t.z;
                t.b[l] = t.r;
            } else {
                t.x[l] = take(c * 8);
                t.b[l] = take(c * 8);
            }
        }
        n = (n + 1) / 2;
    }
    take(kAlign);
    return off;
}

}  // namespace darcy_hip
