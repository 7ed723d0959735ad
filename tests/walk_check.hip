// Host build of the vertical walks (presmooth_walk_item, restrict_walk_item,
// red_walk_item, black_prolong_walk_item in device.hpp), checked bit for bit
// against plain loops:
// the red sweep from zero then the black sweep with the stored inv_diag
// (level.hpp:145-161), and the residual summed over each 2x2 aggregate
// (level.hpp:122-137, :214-227).  Every size from 1 cell up, odd and even
// sides, scalar and pair loads, and several rows-per-item choices.  Built
// with AddressSanitizer on the host, so an index outside any array fails.
// The lane shuffles of the device walk are not exercised here (NoShare
// forms the side values from loads); the GPU tests cover them.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <random>
#include "device.hpp"
using namespace darcy_hip;

template <class T>
int check(int n, long long nthr, unsigned seed, bool pairs) {
    std::mt19937 g(seed);
    std::uniform_real_distribution<double> U(0.1, 12.0), V(-1.0, 1.0);
    // exact-size heap arrays so ASan catches any out-of-bounds access
    T* ex = new T[(size_t)n * (n + 1)];
    T* ny = new T[(size_t)(n + 1) * n];
    T* id = new T[(size_t)n * n];
    T* b = new T[(size_t)n * n];
    T* x = new T[(size_t)n * n];
    T* xr = new T[(size_t)n * n];
    for (size_t i = 0; i < (size_t)n * (n + 1); ++i) ex[i] = (T)U(g), ny[i] = (T)U(g);
    for (size_t i = 0; i < (size_t)n * n; ++i) b[i] = (T)V(g), x[i] = (T)12345, xr[i] = 0;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            const int f = i * (n + 1) + j;
            id[i * n + j] = inv_from_faces(ex[f + 1], ex[f], ny[(i + 1) * n + j], ny[i * n + j]);
        }
    // reference: red sweep from zero then black sweep (level.hpp order)
    for (int i = 0; i < n; ++i)
        for (int j = (i & 1); j < n; j += 2) xr[i * n + j] = b[i * n + j] * id[i * n + j];
    for (int i = 0; i < n; ++i)
        for (int j = 1 - (i & 1); j < n; j += 2) {
            const int f = i * (n + 1) + j, k = i * n + j;
            T v = b[k];
            if (j + 1 < n) v += ex[f + 1] * xr[k + 1];
            if (j > 0) v += ex[f] * xr[k - 1];
            if (i + 1 < n) v += ny[(i + 1) * n + j] * xr[k + n];
            if (i > 0) v += ny[i * n + j] * xr[k - n];
            xr[k] = v * id[k];
        }
    LvP<T> L{n, ex, ny, id, x, b};
    // emulate presmooth_walk's item loop with nthr threads
    const int h = (n + 1) >> 1;
    long long per = ((long long)n * h) / nthr;
    const int rows = per < 1 ? 1 : (per > kWalkRows ? kWalkRows : (int)per);
    const int strips = (n + rows - 1) / rows, total = strips * h;
    for (int w = 0; w < total; ++w) {
        const int st = w / h, k = w - st * h, i0 = st * rows;
        if (n & 1) presmooth_walk_item<T, false>(L, k, i0, i0 + rows < n ? i0 + rows : n);
        else if (pairs) presmooth_walk_item<T, true>(L, k, i0, i0 + rows < n ? i0 + rows : n);
        else presmooth_walk_item<T, false>(L, k, i0, i0 + rows < n ? i0 + rows : n);
    }
    int bad = 0;
    for (size_t i = 0; i < (size_t)n * n; ++i)
        if (std::memcmp(&x[i], &xr[i], sizeof(T))) {
            if (bad < 3) printf("  n=%d i=%zu got %.17g want %.17g\n", n, i, (double)x[i], (double)xr[i]);
            ++bad;
        }
    printf("%s n=%d nthr=%lld rows=%d pairs=%d: %s\n", sizeof(T) == 8 ? "f64" : "f32", n, nthr, rows, (int)pairs, bad ? "MISMATCH" : "ok");
    delete[] ex; delete[] ny; delete[] id; delete[] b; delete[] x; delete[] xr;
    return bad;
}

template <class T>
int check_restrict(int n, long long nthr, unsigned seed, bool pairs) {
    std::mt19937 g(seed);
    std::uniform_real_distribution<double> U(0.1, 12.0), V(-1.0, 1.0);
    const int m = (n + 1) / 2;
    T* ex = new T[(size_t)n * (n + 1)];
    T* ny = new T[(size_t)(n + 1) * n];
    T* b = new T[(size_t)n * n];
    T* x = new T[(size_t)n * n];
    T* bc = new T[(size_t)m * m];
    T* bcr = new T[(size_t)m * m];
    for (size_t i = 0; i < (size_t)n * (n + 1); ++i) ex[i] = (T)U(g), ny[i] = (T)U(g);
    for (size_t i = 0; i < (size_t)n * n; ++i) b[i] = (T)V(g), x[i] = (T)V(g);
    for (size_t i = 0; i < (size_t)m * m; ++i) bc[i] = (T)777, bcr[i] = (T)0;
    auto res = [&](int i, int j) {
        const int f = i * (n + 1) + j, k = i * n + j;
        const T e = ex[f + 1], w = ex[f], nn = ny[(i + 1) * n + j], s = ny[i * n + j];
        const T d = ((e + w) + nn) + s;
        T v = d * x[k];
        if (j + 1 < n) v -= e * x[k + 1];
        if (j > 0) v -= w * x[k - 1];
        if (i + 1 < n) v -= nn * x[k + n];
        if (i > 0) v -= s * x[k - n];
        return b[k] - v;
    };
    for (int I = 0; I < m; ++I)
        for (int J = 0; J < m; ++J) {
            T s = T(0);
            for (int d = 0; d < 4; ++d) {
                const int i = 2 * I + (d >> 1), j = 2 * J + (d & 1);
                if (i < n && j < n) s += res(i, j);
            }
            bcr[I * m + J] = s;
        }
    LvP<T> L{n, ex, ny, nullptr, x, b};
    long long per = ((long long)m * m) / nthr;
    const int rows = per < 1 ? 1 : (per > kWalkRows / 2 ? kWalkRows / 2 : (int)per);
    const int strips = (m + rows - 1) / rows, total = strips * m;
    for (int w = 0; w < total; ++w) {
        const int st = w / m, J = w - st * m, I0 = st * rows, I1 = I0 + rows < m ? I0 + rows : m;
        if (n & 1) restrict_walk_item<T, false>(L, bc, J, I0, I1);
        else if (pairs) restrict_walk_item<T, true>(L, bc, J, I0, I1);
        else restrict_walk_item<T, false>(L, bc, J, I0, I1);
    }
    int bad = 0;
    for (size_t i = 0; i < (size_t)m * m; ++i)
        if (std::memcmp(&bc[i], &bcr[i], sizeof(T))) {
            if (bad < 3) printf("  restrict n=%d i=%zu got %.17g want %.17g\n", n, i, (double)bc[i], (double)bcr[i]);
            ++bad;
        }
    printf("restrict %s n=%d nthr=%lld rows=%d pairs=%d: %s\n", sizeof(T) == 8 ? "f64" : "f32", n, nthr, rows, (int)pairs, bad ? "MISMATCH" : "ok");
    delete[] ex; delete[] ny; delete[] b; delete[] x; delete[] bc; delete[] bcr;
    return bad;
}

// The fresh down visit as the persistent kernel runs it (down_walk): steps
// of R rows, the pre-smoothing walks over a step's rows, then the
// restriction walks over the coarse rows whose fine rows around them are
// final.  Executed in that order here, a coarse row taken too early reads
// the sentinel x of a row not yet smoothed and fails the comparison.
template <class T>
int check_down(int n, int step_cells, int nthr, unsigned seed, bool pairs) {
    std::mt19937 g(seed);
    std::uniform_real_distribution<double> U(0.1, 12.0), V(-1.0, 1.0);
    const int m = (n + 1) / 2, h = m;
    std::vector<T> ex((size_t)n * (n + 1)), ny((size_t)(n + 1) * n), b((size_t)n * n);
    T* x = new T[(size_t)n * n];
    T* bc = new T[(size_t)m * m];
    std::vector<T> xr((size_t)n * n), bcr((size_t)m * m);
    for (auto& v : ex) v = (T)U(g);
    for (auto& v : ny) v = (T)U(g);
    for (auto& v : b) v = (T)V(g);
    for (size_t i = 0; i < (size_t)n * n; ++i) x[i] = (T)12345;
    for (size_t i = 0; i < (size_t)m * m; ++i) bc[i] = (T)777;
    auto inv = [&](int i, int j) {
        const int f = i * (n + 1) + j;
        return inv_from_faces(ex[f + 1], ex[f], ny[(i + 1) * n + j], ny[i * n + j]);
    };
    for (int i = 0; i < n; ++i)
        for (int j = (i & 1); j < n; j += 2) xr[i * n + j] = b[i * n + j] * inv(i, j);
    for (int i = 0; i < n; ++i)
        for (int j = 1 - (i & 1); j < n; j += 2) {
            const int f = i * (n + 1) + j, k = i * n + j;
            T v = b[k];
            if (j + 1 < n) v += ex[f + 1] * xr[k + 1];
            if (j > 0) v += ex[f] * xr[k - 1];
            if (i + 1 < n) v += ny[(i + 1) * n + j] * xr[k + n];
            if (i > 0) v += ny[i * n + j] * xr[k - n];
            xr[k] = v * inv(i, j);
        }
    for (int I = 0; I < m; ++I)
        for (int J = 0; J < m; ++J) {
            T s = T(0);
            for (int d = 0; d < 4; ++d) {
                const int i = 2 * I + (d >> 1), j = 2 * J + (d & 1);
                if (i >= n || j >= n) continue;
                const int f = i * (n + 1) + j, k = i * n + j;
                const T e = ex[f + 1], w = ex[f], nn = ny[(i + 1) * n + j], so = ny[i * n + j];
                const T dd = ((e + w) + nn) + so;
                T v = dd * xr[k];
                if (j + 1 < n) v -= e * xr[k + 1];
                if (j > 0) v -= w * xr[k - 1];
                if (i + 1 < n) v -= nn * xr[k + n];
                if (i > 0) v -= so * xr[k - n];
                s += b[k] - v;
            }
            bcr[I * m + J] = s;
        }
    LvP<T> L{n, ex.data(), ny.data(), nullptr, x, b.data()};
    const int R = rows_per_step(n, step_cells);
    int coarse_done = 0;
    for (int r0 = 0; r0 < n; r0 += R) {
        const int r1 = r0 + R < n ? r0 + R : n;
        const int strip = walk_strip(r1 - r0, h, nthr, kWalkRows);
        const int strips = (r1 - r0 + strip - 1) / strip;
        for (int w = 0; w < strips * h; ++w) {
            const int st = w / h, k = w - st * h, i0 = r0 + st * strip, i1 = i0 + strip < r1 ? i0 + strip : r1;
            if ((n & 1) || !pairs) presmooth_walk_item<T, false>(L, k, i0, i1);
            else presmooth_walk_item<T, true>(L, k, i0, i1);
        }
        const int coarse_hi = r1 == n ? m : (r1 - 1) / 2;
        if (coarse_hi > coarse_done) {
            const int cs = walk_strip(coarse_hi - coarse_done, m, nthr, kWalkRows / 2);
            const int cstrips = (coarse_hi - coarse_done + cs - 1) / cs;
            for (int w = 0; w < cstrips * m; ++w) {
                const int st = w / m, J = w - st * m, I0 = coarse_done + st * cs, I1 = I0 + cs < coarse_hi ? I0 + cs : coarse_hi;
                if ((n & 1) || !pairs) restrict_walk_item<T, false>(L, bc, J, I0, I1);
                else restrict_walk_item<T, true>(L, bc, J, I0, I1);
            }
        }
        coarse_done = coarse_hi;
    }
    int bad = 0;
    for (size_t i = 0; i < (size_t)n * n; ++i) bad += std::memcmp(&x[i], &xr[i], sizeof(T)) != 0;
    for (size_t i = 0; i < (size_t)m * m; ++i) bad += std::memcmp(&bc[i], &bcr[i], sizeof(T)) != 0;
    printf("down %s n=%d step=%d nthr=%d pairs=%d: %s\n", sizeof(T) == 8 ? "f64" : "f32", n, step_cells, nthr, (int)pairs,
           bad ? "MISMATCH" : "ok");
    delete[] x;
    delete[] bc;
    return bad;
}

// The red half of the post-smoothing as a walk (red_walk_item): black cells
// hold final values, red cells are updated from them; items run one after
// the other (a red cell reads no red cell), strips of several heights.
template <class T>
int check_red(int n, int strip, unsigned seed, bool pairs) {
    std::mt19937 g(seed);
    std::uniform_real_distribution<double> U(0.1, 12.0), V(-1.0, 1.0);
    const int h = (n + 1) / 2;
    T* ex = new T[(size_t)n * (n + 1)];
    T* ny = new T[(size_t)(n + 1) * n];
    T* b = new T[(size_t)n * n];
    T* x = new T[(size_t)n * n];
    std::vector<T> xr((size_t)n * n);
    for (size_t i = 0; i < (size_t)n * (n + 1); ++i) ex[i] = (T)U(g), ny[i] = (T)U(g);
    for (size_t i = 0; i < (size_t)n * n; ++i) b[i] = (T)V(g), x[i] = (T)V(g), xr[i] = x[i];
    for (int i = 0; i < n; ++i)
        for (int j = (i & 1); j < n; j += 2) {
            const int f = i * (n + 1) + j, kk = i * n + j;
            const T e = ex[f + 1], w = ex[f], nn = ny[(i + 1) * n + j], so = ny[i * n + j];
            const T inv = inv_from_faces(e, w, nn, so);
            T v = b[kk];
            if (j + 1 < n) v += e * x[kk + 1];
            if (j > 0) v += w * x[kk - 1];
            if (i + 1 < n) v += nn * x[kk + n];
            if (i > 0) v += so * x[kk - n];
            xr[kk] = v * inv;
        }
    LvP<T> L{n, ex, ny, nullptr, x, b};
    for (int i0 = 0; i0 < n; i0 += strip)
        for (int k = 0; k < h; ++k) {
            const int i1 = i0 + strip < n ? i0 + strip : n;
            if ((n & 1) || !pairs)
                red_walk_item<T, false>(L, k, i0, i1);
            else
                red_walk_item<T, true>(L, k, i0, i1);
        }
    int bad = 0;
    for (size_t i = 0; i < (size_t)n * n; ++i) bad += std::memcmp(&x[i], &xr[i], sizeof(T)) != 0;
    printf("red %s n=%d strip=%d pairs=%d: %s\n", sizeof(T) == 8 ? "f64" : "f32", n, strip, (int)pairs,
           bad ? "MISMATCH" : "ok");
    delete[] ex;
    delete[] ny;
    delete[] b;
    delete[] x;
    return bad;
}

// The black half of the first post-smoothing sweep with the prolongation as
// a walk (black_prolong_walk_item): red neighbours read as x + x_c of their
// aggregates, black cells written; items one after the other.
template <class T, bool kProlong = true>
int check_black_prolong(int n, int strip, unsigned seed, bool pairs) {
    std::mt19937 g(seed);
    std::uniform_real_distribution<double> U(0.1, 12.0), V(-1.0, 1.0);
    const int h = (n + 1) / 2, m = h;
    T* ex = new T[(size_t)n * (n + 1)];
    T* ny = new T[(size_t)(n + 1) * n];
    T* b = new T[(size_t)n * n];
    T* x = new T[(size_t)n * n];
    T* xc = new T[(size_t)m * m];
    std::vector<T> xr((size_t)n * n);
    for (size_t i = 0; i < (size_t)n * (n + 1); ++i) ex[i] = (T)U(g), ny[i] = (T)U(g);
    for (size_t i = 0; i < (size_t)n * n; ++i) b[i] = (T)V(g), x[i] = (T)V(g), xr[i] = x[i];
    for (size_t i = 0; i < (size_t)m * m; ++i) xc[i] = (T)V(g);
    auto pv = [&](int i, int j) {
        if constexpr (kProlong)
            return T(x[(size_t)i * n + j] + xc[(size_t)(i >> 1) * m + (j >> 1)]);
        else
            return x[(size_t)i * n + j];
    };
    for (int i = 0; i < n; ++i)
        for (int j = 1 - (i & 1); j < n; j += 2) {
            const int f = i * (n + 1) + j, kk = i * n + j;
            const T e = ex[f + 1], w = ex[f], nn = ny[(i + 1) * n + j], so = ny[i * n + j];
            const T inv = inv_from_faces(e, w, nn, so);
            T v = b[kk];
            if (j + 1 < n) v += e * pv(i, j + 1);
            if (j > 0) v += w * pv(i, j - 1);
            if (i + 1 < n) v += nn * pv(i + 1, j);
            if (i > 0) v += so * pv(i - 1, j);
            xr[kk] = v * inv;
        }
    LvP<T> L{n, ex, ny, nullptr, x, b};
    for (int i0 = 0; i0 < n; i0 += strip)
        for (int k = h - 1; k >= 0; --k) {
            const int i1 = i0 + strip < n ? i0 + strip : n;
            if ((n & 1) || !pairs)
                black_prolong_walk_item<T, false, kProlong>(L, (const T*)xc, k, i0, i1);
            else
                black_prolong_walk_item<T, true, kProlong>(L, (const T*)xc, k, i0, i1);
        }
    int bad = 0;
    for (size_t i = 0; i < (size_t)n * n; ++i) bad += std::memcmp(&x[i], &xr[i], sizeof(T)) != 0;
    printf("black%s %s n=%d strip=%d pairs=%d: %s\n", kProlong ? "+prolong" : "", sizeof(T) == 8 ? "f64" : "f32", n,
           strip, (int)pairs,
           bad ? "MISMATCH" : "ok");
    delete[] ex;
    delete[] ny;
    delete[] b;
    delete[] x;
    delete[] xc;
    return bad;
}

int main() {
    int bad = 0;
    const int sizes[] = {1, 2, 3, 5, 8, 65, 66, 67, 100, 125, 128, 129, 257, 512};
    const long long thr[] = {1, 7, 64, 1000, 1 << 20};
    unsigned seed = 1;
    for (int n : sizes)
        for (long long t : thr) {
            for (int p = 0; p < 2; ++p) {
                bad += check<double>(n, t, seed++, p);
                bad += check<float>(n, t, seed++, p);
                bad += check_restrict<double>(n, t, seed++, p);
                bad += check_restrict<float>(n, t, seed++, p);
            }
        }
    // the persistent kernel's pipelined visits: 256 threads, step sizes of
    // the tuning range and beyond, odd and even sides
    const int vsizes[] = {17, 18, 33, 64, 65, 100, 128, 129, 256, 257};
    const int steps[] = {64, 512, 1024, 2048, 4096, 1 << 20};
    for (int n : vsizes)
        for (int st : steps)
            for (int p = 0; p < 2; ++p) {
                bad += check_down<double>(n, st, 256, seed++, p);
                bad += check_down<float>(n, st, 256, seed++, p);
            }
    // the red half of the post-smoothing as walks
    const int rsizes[] = {1, 2, 3, 4, 5, 8, 17, 18, 33, 64, 65, 100, 128, 129, 257};
    const int rstrips[] = {1, 2, 3, 8, 16, 1 << 20};
    for (int n : rsizes)
        for (int st : rstrips)
            for (int p = 0; p < 2; ++p) {
                bad += check_red<double>(n, st, seed++, p);
                bad += check_red<float>(n, st, seed++, p);
            }
    for (int n : rsizes)
        for (int st : rstrips)
            for (int p = 0; p < 2; ++p) {
                bad += check_black_prolong<double>(n, st, seed++, p);
                bad += check_black_prolong<float>(n, st, seed++, p);
                bad += check_black_prolong<double, false>(n, st, seed++, p);
                bad += check_black_prolong<float, false>(n, st, seed++, p);
            }
    printf(bad ? "FAILED\n" : "ALL OK\n");
    return bad != 0;
}
