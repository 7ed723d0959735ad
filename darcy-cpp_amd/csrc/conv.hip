// Convolution kernels of the learned warm start: a BF16 encoder-decoder
// with FP32 accumulation, on the matrix cores of gfx950.
//
// Forward convolutions with a power-of-two input channel count (and the
// stride-1 backward data, which is the forward conv of dZ with the flipped,
// transposed weights) run in k_conv, a direct implicit GEMM on
// v_mfma_f32_16x16x32_bf16: a wave owns a 64-pixel x 64-channel tile, the K
// loop walks (tap, ci) in steps of 32 with the next step's operands loaded
// while the current step's MFMAs run, and the pixel coordinates are decoded
// once per lane.  Every other shape (strided backward data, the weight
// gradient, channel counts that are not powers of two) runs in k_gemm, a
// generic 64 x 64 tile whose `gather` functors read the 8 consecutive K
// values a lane needs straight from the NHWC tensors (one 16-byte load when
// they are contiguous).  No operand is ever materialised.
//
//   forward        Z[pixel][co]      = sum_(tap,ci) X[pixel@tap][ci] * W[co][tap][ci] + bias[co]
//   backward data  dX[pixel][ci]     = sum_(tap,co) dZ[pixel@-tap][co] * W[co][tap][ci]
//   backward wt.   dW[co][(tap,ci)]  = sum_pixel    dZT[co][pixel] * X[pixel@tap][ci]
//                  db[co]            = sum_pixel    dZT[co][pixel]            (a column of ones)
//
// The weight gradient splits K (pixels) over blockIdx.z and sums the FP32
// partials in split order, so gradients do not depend on scheduling.  The
// head keeps FP32 weights and runs on the matrix cores as two bf16 parts
// (k_head_mfma); its gradients are FP32 kernels.
// Activations are NHWC bf16; weights W[co][tap][ci] bf16 from FP32 masters.
// No reference counterpart: the reference solve always starts from u = 0.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace darcy_nn {

typedef __bf16 bf16;
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kT = 256;  // threads per workgroup
constexpr int kTile = 64;

__device__ __forceinline__ float bf2f(bf16 v) { return static_cast<float>(v); }
__device__ __forceinline__ bf16 f2bf(float v) { return static_cast<bf16>(v); }

// GELU, tanh form: 0.5 z (1 + tanh(u)), u = c (z + 0.044715 z^3), c = sqrt(2/pi),
// evaluated as z * s with s = 1 / (1 + exp(-2u)) (= (1 + tanh u) / 2): no
// cancellation in 1 + tanh u for negative z.  Derivative:
// s + 2 z s (1 - s) c (1 + 3 * 0.044715 z^2).  exp and the reciprocal use the
// hardware v_exp_f32 / v_rcp_f32 (about 1 ulp): every GELU value is rounded
// to bf16, and the network evaluates about 8 million of them per 256^2 sample.
__device__ __forceinline__ float gelu_s(float z) {
    const float k = -2.0f * 0.7978845608028654f * 1.4426950408889634f;  // -2 c log2(e)
    const float e = __builtin_amdgcn_exp2f(k * z * (1.0f + 0.044715f * z * z));
    return __builtin_amdgcn_rcpf(1.0f + e);
}
__device__ __forceinline__ float gelu(float z) { return z * gelu_s(z); }
__device__ __forceinline__ float gelu_grad(float z) {
    const float c = 0.7978845608028654f;
    const float s = gelu_s(z);
    return s + 2.0f * z * s * (1.0f - s) * c * (1.0f + 3.0f * 0.044715f * z * z);
}

// Activation tensor geometry (NHWC).
struct Geo {
    int b, h, w, c;
};

__device__ __forceinline__ size_t at(const Geo& g, int b, int y, int x) {
    return ((static_cast<size_t>(b) * g.h + y) * g.w + x) * g.c;
}

// ------------------------------------------------------------ gathers
// load8(row, k) returns the operand values (row, k .. k+7); zero outside.

// im2col of the forward conv: row = output pixel, k = tap * Cin + ci.
struct FwdA {
    const bf16* x;
    Geo in;
    int oh, ow, stride, M;
    __device__ v8bf load8(int m, int k) const {
        v8bf r = {};
        if (m >= M) return r;
        const int ox = m % ow, t = m / ow, oy = t % oh, b = t / oh;
        const int tap = k / in.c, ci = k - tap * in.c;
        if (tap >= 9) return r;
        const int iy = oy * stride + tap / 3 - 1, ix = ox * stride + tap % 3 - 1;
        if (iy < 0 || iy >= in.h || ix < 0 || ix >= in.w) return r;
        return *reinterpret_cast<const v8bf*>(x + at(in, b, iy, ix) + ci);
    }
};

// Row-major [N][K] matrix used as B (B[k][n] = w[n][k]): weights in the
// forward and backward-data GEMMs.
struct RowsB {
    const bf16* w;
    int K, N;
    __device__ v8bf load8(int k, int n) const {
        v8bf r = {};
        if (n >= N || k >= K) return r;
        return *reinterpret_cast<const v8bf*>(w + static_cast<size_t>(n) * K + k);
    }
};

// Backward data against the flipped, transposed weights
// WF[ci][t][co] = W[co][8 - t][ci]: row = input pixel (y, x), k = t * Cout + co.
// The output pixel that reads (y, x) through the original tap 8 - t is
// ((y + t/3 - 1) / s, (x + t%3 - 1) / s) when the division is exact and the
// pixel lies on the output grid.
struct BwdDataA {
    const bf16* dz;
    Geo out;  // geometry of the conv output (and of dz)
    int ih, iw, stride, M;
    __device__ v8bf load8(int m, int k) const {
        v8bf r = {};
        if (m >= M) return r;
        const int x = m % iw, t = m / iw, y = t % ih, b = t / ih;
        const int tap = k / out.c, co = k - tap * out.c;
        if (tap >= 9) return r;
        int oy = y + tap / 3 - 1, ox = x + tap % 3 - 1;
        if (stride == 2) {
            if ((oy & 1) || (ox & 1)) return r;
            oy >>= 1;
            ox >>= 1;
        }
        if (oy < 0 || oy >= out.h || ox < 0 || ox >= out.w) return r;
        return *reinterpret_cast<const v8bf*>(dz + at(out, b, oy, ox) + co);
    }
};

// Row-major [M][K] matrix used as A: dZ^T in the weight gradient (K, the
// pixel count, need not be a multiple of 8 on the coarsest levels).
struct RowsA {
    const bf16* a;
    int K, M;
    __device__ v8bf load8(int m, int k) const {
        v8bf r = {};
        if (m >= M || k >= K) return r;
        const bf16* row = a + static_cast<size_t>(m) * K;
        if ((K & 7) == 0) return *reinterpret_cast<const v8bf*>(row + k);
        for (int t = 0; t < 8 && k + t < K; ++t) r[t] = row[k + t];
        return r;
    }
};

// Weight gradient B: B[pixel][n] = X at the input pixel under tap n / Cin for
// n < 9 Cin, and 1 in column 9 Cin (the bias).
struct BwdWtB {
    const bf16* x;
    Geo in;
    int oh, ow, stride, P, NW;  // NW = 9 * Cin
    __device__ v8bf load8(int p0, int n) const {
        v8bf r = {};
        if (n > NW || p0 >= P) return r;
        if (n == NW) {
#pragma unroll
            for (int t = 0; t < 8; ++t) r[t] = f2bf(p0 + t < P ? 1.0f : 0.0f);
            return r;
        }
        const int tap = n / in.c, ci = n - tap * in.c, ky = tap / 3 - 1, kx = tap % 3 - 1;
        if ((ow & 7) == 0) {  // the 8 pixels share one output row
            const int ox0 = p0 % ow, q = p0 / ow, oy = q % oh, b = q / oh;
            const int iy = oy * stride + ky;
            if (iy < 0 || iy >= in.h) return r;
            const bf16* row = x + at(in, b, iy, 0) + ci;
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const int ix = (ox0 + t) * stride + kx;
                if (ix >= 0 && ix < in.w) r[t] = row[static_cast<size_t>(ix) * in.c];
            }
            return r;
        }
        for (int t = 0; t < 8; ++t) {
            const int p = p0 + t, ox = p % ow, q = p / ow, oy = q % oh, b = q / oh;
            if (p >= P) break;
            const int iy = oy * stride + ky, ix = ox * stride + kx;
            if (iy >= 0 && iy < in.h && ix >= 0 && ix < in.w) r[t] = x[at(in, b, iy, ix) + ci];
        }
        return r;
    }
};

// ------------------------------------------------------------ epilogues

// Epilogues take a row handle from row(m), computed once per output row, and
// then put(row, n, acc) for each column the lane holds.
struct RowOff {
    size_t off;
    bool ok;
};

// Forward: z = acc + bias (bias may be null); stores z and gelu(z) where wanted.
struct EpiFwd {
    bf16* z;
    bf16* act;
    const float* bias;
    int M, N;
    __device__ RowOff row(int m) const { return {static_cast<size_t>(m) * N, m < M}; }
    __device__ void put(const RowOff& r, int n, float acc) const {
        if (!r.ok || n >= N) return;
        const bf16 zb = f2bf(bias ? acc + bias[n] : acc);
        if (z) z[r.off + n] = zb;
        if (act) act[r.off + n] = f2bf(gelu(bf2f(zb)));
    }
    __device__ void operator()(int m, int n, float acc) const { put(row(m), n, acc); }
    // Two-phase form (k_conv): pre() is the value staged as bf16 z, put8()
    // stores z and act = bf16(gelu(z)) for channels n .. n+7 of row m.
    __device__ float pre(int n, float acc) const { return bias && n < N ? acc + bias[n] : acc; }
    __device__ void put8(int m, int n, const v8bf& zv) const {
        if (m >= M || n >= N) return;
        const size_t i = static_cast<size_t>(m) * N + n;
        if (z) *reinterpret_cast<v8bf*>(z + i) = zv;
        if (act) {
            v8bf av;
#pragma unroll
            for (int e = 0; e < 8; ++e) av[e] = f2bf(gelu(bf2f(zv[e])));
            *reinterpret_cast<v8bf*>(act + i) = av;
        }
    }
};

// Decoder conv with its merge: c = bf16(acc + bias) at the coarse pixel
// (oy, ox) is upsampled x2 (nearest) and added to the encoder activation
// `skip`; stores z = c + skip (if wanted) and gelu(z) at the 4 fine pixels.
// The same values as the conv followed by k_up_add_gelu, without the c tensor.
struct EpiMerge {
    const float* bias;
    const bf16* skip;
    bf16* z;
    bf16* act;
    int M, N, oh, ow;  // conv output (coarse) geometry
    __device__ RowOff row(int m) const {
        const int ox = m % ow, q = m / ow, oy = q % oh, b = q / oh;
        const size_t W = 2 * static_cast<size_t>(ow);
        return {((static_cast<size_t>(b) * 2 * oh + 2 * oy) * W + 2 * ox) * N, m < M};
    }
    __device__ float pre(int n, float acc) const { return n < N ? acc + bias[n] : acc; }
    // c = the staged bf16 conv output of coarse row m, channels n .. n+7.
    // The 4 fine pixels' skip values are loaded before any store, so the
    // loads are in flight together (skip and the outputs may alias as far as
    // the compiler knows).
    __device__ void put8(int m, int n, const v8bf& cv) const {
        if (m >= M || n >= N) return;
        const RowOff r = row(m);
        const size_t wn = 2 * static_cast<size_t>(ow) * N;
        v8bf sk[4];
#pragma unroll
        for (int d = 0; d < 4; ++d) sk[d] = *reinterpret_cast<const v8bf*>(skip + r.off + n + (d >> 1) * wn + (d & 1) * N);
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            const size_t i = r.off + n + (d >> 1) * wn + (d & 1) * N;
            const v8bf sv = sk[d];
            v8bf zv, av;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                zv[e] = f2bf(bf2f(cv[e]) + bf2f(sv[e]));
                av[e] = f2bf(gelu(bf2f(zv[e])));
            }
            if (z) *reinterpret_cast<v8bf*>(z + i) = zv;
            *reinterpret_cast<v8bf*>(act + i) = av;
        }
    }
    __device__ void put(const RowOff& r, int n, float acc) const {
        if (!r.ok || n >= N) return;
        const float vc = bf2f(f2bf(acc + bias[n]));
        const size_t wn = 2 * static_cast<size_t>(ow) * N;
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            const size_t i = r.off + n + (d >> 1) * wn + (d & 1) * N;
            const bf16 zb = f2bf(vc + bf2f(skip[i]));
            if (z) z[i] = zb;
            act[i] = f2bf(gelu(bf2f(zb)));
        }
    }
};

// Backward data: dX = acc.
struct EpiStore {
    bf16* out;
    int M, N;
    __device__ void operator()(int m, int n, float acc) const {
        if (m >= M || n >= N) return;
        out[static_cast<size_t>(m) * N + n] = f2bf(acc);
    }
};

// Weight gradient: FP32 partial of split blockIdx.z.
struct EpiPartial {
    float* part;
    int M, N;
    __device__ void operator()(int m, int n, float acc) const {
        if (m >= M || n >= N) return;
        part[(static_cast<size_t>(blockIdx.z) * M + m) * N + n] = acc;
    }
};

// ------------------------------------------------------------ the GEMM
// v_mfma_f32_16x16x32_bf16 operand layout: lane l holds
//   A[m = l % 16][k = 8 (l / 16) + t] and B[k = 8 (l / 16) + t][n = l % 16], t = 0..7,
// and receives D[m = 4 (l / 16) + t][n = l % 16], t = 0..3.
template <class GA, class GB, class Epi>
__global__ __launch_bounds__(kT) void k_gemm(int kchunk, int K, GA ga, GB gb, Epi epi) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m0 = blockIdx.x * kTile + (wave & 1) * 32, n0 = blockIdx.y * kTile + (wave >> 1) * 32;
    const int r = lane & 15, kq = (lane >> 4) * 8;
    const int kbeg = blockIdx.z * kchunk, kend = min(K, kbeg + kchunk);
    v4f acc[2][2] = {};
    if (kbeg < kend) {
        v8bf a0 = ga.load8(m0 + r, kbeg + kq), a1 = ga.load8(m0 + 16 + r, kbeg + kq);
        v8bf b0 = gb.load8(kbeg + kq, n0 + r), b1 = gb.load8(kbeg + kq, n0 + 16 + r);
        for (int k0 = kbeg; k0 < kend; k0 += 32) {
            // Operands of the next step (past kend they are not used).
            const int kn = k0 + 32 < kend ? k0 + 32 + kq : kend;
            const v8bf na0 = ga.load8(m0 + r, kn), na1 = ga.load8(m0 + 16 + r, kn);
            const v8bf nb0 = gb.load8(kn, n0 + r), nb1 = gb.load8(kn, n0 + 16 + r);
            acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1, acc[1][1], 0, 0, 0);
            a0 = na0;
            a1 = na1;
            b0 = nb0;
            b1 = nb1;
        }
    }
    const int rq = (lane >> 4) * 4;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int t = 0; t < 4; ++t) epi(m0 + 16 * i + rq + t, n0 + 16 * j + r, acc[i][j][t]);
}

// ------------------------------------------------- the direct conv kernel
// Forward 3x3 conv (and stride-1 backward data, which is a forward conv of dZ
// with the flipped, transposed weights) for a power-of-two input channel
// count >= 8.  Rows are output pixels, columns output channels, and
// K = (tap, ci) in steps of 32.  A workgroup owns a BM x BN tile (waves
// WAVES_M x WAVES_N, each 64 x WN = 4 x WN/16 MFMA tiles).  Each K step's A
// (BM pixels x 32) and B (BN channels x 32) tiles go through LDS, double
// buffered, with one barrier per step; the 16-byte chunks of step s + 2 are
// loaded into registers while step s is multiplied and stored to LDS after
// step s + 1.  A thread stages one chunk column of BM / 64 pixel rows, each
// decoded once; per step it adds the tap offset (tap / 3 as
// (tap * 11) >> 5).  Workgroups are dealt to the 8 XCDs
// round-robin; the tile order gives each XCD a contiguous band of pixel rows,
// so the taps' overlapping input rows stay in one L2.
struct ConvArgs {
    const bf16* x;  // NHWC input, in.c = 1 << lg_cin channels
    const bf16* w;  // [cout][9][cin]
    Geo in;
    int oh, ow, stride, lg_cin, cout, M, tiles_m, tiles_n;
};

template <int WN, int WAVES_N, class Epi>
__global__ __launch_bounds__(kT) void k_conv(ConvArgs c, Epi epi) {
    constexpr int WAVES_M = 4 / WAVES_N, TN = WN / 16;
    constexpr int BM = 64 * WAVES_M, BN = WN * WAVES_N;
    constexpr int A_PER = BM * 4 / kT;                     // A chunks per thread per step
    constexpr int B_PER = BN * 4 >= kT ? BN * 4 / kT : 1;  // B chunks per thread per step
    // A K step (32 values of k) of an operand row is 4 chunks of 8 bf16.  Four
    // consecutive lanes stage the 4 chunks of one row, so a wave's load reads
    // 64 contiguous bytes from each of 16 rows (pixels or output channels)
    // instead of 16 bytes from each of 64: the address unit and the L1 then
    // serve a quarter of the cache lines per load.  LDS holds each operand
    // row-major, 80 bytes a row (64 + 16 of padding).  (Measured on the
    // encoder: 18.0 -> 16.0 ms of inference per 1000 samples; an XOR-swizzled
    // 64-byte row was no faster for the 128 x 128 tiles and slower for the
    // 256 x 64 ones.)
    constexpr int RB = 80;  // bytes per LDS row
    constexpr int kBuf = (BM + BN) * RB;
    constexpr int kStage = 2 * kBuf, kEpi = 4 * 64 * (WN + 8) * 2;
    __shared__ __align__(16) char smem[kStage > kEpi ? kStage : kEpi];
    auto lds_a = [&](int buf, int row, int kc) -> v8bf& {
        return *reinterpret_cast<v8bf*>(smem + buf * kBuf + row * RB + kc * 16);
    };
    auto lds_b = [&](int buf, int col, int kc) -> v8bf& {
        return *reinterpret_cast<v8bf*>(smem + buf * kBuf + (BM + col) * RB + kc * 16);
    };

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nb = gridDim.x;
    int t = blockIdx.x;
    if ((nb & 7) == 0) t = (t & 7) * (nb >> 3) + (t >> 3);
    const int tm = t / c.tiles_n, tn = t - tm * c.tiles_n;
    const int m_wg = tm * BM, n_wg = tn * BN;
    const int wm = wave / WAVES_N, wn = wave % WAVES_N;
    const int r = lane & 15, g = lane >> 4;
    const int cin = 1 << c.lg_cin, K = 9 << c.lg_cin, steps = (K + 31) / 32;

    // This thread stages chunk kc of A rows (tid >> 2) + 64 s and of B rows
    // (tid >> 2) + 64 s; the pixel of each A row is decoded once.
    // Offsets within one image are 32-bit (an NHWC image is far below 2^31
    // elements); only the image base is a 64-bit pointer.
    const int kc = tid & 3, row0 = tid >> 2;
    int iy0[A_PER], ix0[A_PER], pix[A_PER];
    const bf16* img[A_PER];
#pragma unroll
    for (int s = 0; s < A_PER; ++s) {
        const int m = m_wg + row0 + 64 * s;
        const int ox = m % c.ow, q = m / c.ow, oy = q % c.oh, b = q / c.oh;
        iy0[s] = m < c.M ? oy * c.stride - 1 : -1000000;
        ix0[s] = ox * c.stride - 1;
        img[s] = c.x + static_cast<size_t>(m < c.M ? b : 0) * c.in.h * c.in.w * cin;
        pix[s] = m < c.M ? (iy0[s] * c.in.w + ix0[s]) * cin : 0;
    }
    const bf16* wrow = c.w + static_cast<size_t>(n_wg + row0) * K;
    struct Regs {
        v8bf a[A_PER], b[B_PER];
    };
    auto load = [&](int k0, Regs& rg) {
        const int k = k0 + 8 * kc;
        const int tap = k >> c.lg_cin, ci = k & (cin - 1);
        const int ky = (tap * 11) >> 5, kx = tap - 3 * ky;
        const int toff = (ky * c.in.w + kx) * cin + ci;
#pragma unroll
        for (int s = 0; s < A_PER; ++s) {
            const int iy = iy0[s] + ky, ix = ix0[s] + kx;
            v8bf v = {};
            if (tap < 9 && static_cast<unsigned>(iy) < static_cast<unsigned>(c.in.h) &&
                static_cast<unsigned>(ix) < static_cast<unsigned>(c.in.w))
                v = *reinterpret_cast<const v8bf*>(img[s] + (pix[s] + toff));
            rg.a[s] = v;
        }
#pragma unroll
        for (int s = 0; s < B_PER; ++s) {
            const int col = row0 + 64 * s, n = n_wg + col;
            v8bf v = {};
            if (col < BN && n < c.cout && k < K) v = *reinterpret_cast<const v8bf*>(wrow + (64 * s * K + k));
            rg.b[s] = v;
        }
    };
    auto store = [&](int buf, const Regs& rg) {
#pragma unroll
        for (int s = 0; s < A_PER; ++s) lds_a(buf, row0 + 64 * s, kc) = rg.a[s];
#pragma unroll
        for (int s = 0; s < B_PER; ++s)
            if (row0 + 64 * s < BN) lds_b(buf, row0 + 64 * s, kc) = rg.b[s];
    };

    v4f acc[4][TN] = {};
    auto mma = [&](int buf) {
        v8bf fa[4], fb[TN];
#pragma unroll
        for (int i = 0; i < 4; ++i) fa[i] = lds_a(buf, wm * 64 + 16 * i + r, g);
#pragma unroll
        for (int j = 0; j < TN; ++j) fb[j] = lds_b(buf, wn * WN + 16 * j + r, g);
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
    };
    // Step s computes buffer s & 1.  Register sets r0 / r1 alternate: the
    // loads of step s + 2 go to r[s & 1] at the start of step s, and
    // r[(s + 1) & 1] (the loads of step s + 1, issued a step earlier) is
    // stored to buffer (s + 1) & 1 at its end; that buffer was last read by
    // step s - 1, which every wave finished at the previous barrier.
    Regs r0, r1;
    load(0, r0);
    store(0, r0);
    if (steps > 1) load(32, r1);
    __syncthreads();
    for (int step = 0; step < steps; ++step) {
        const int odd = step & 1;
        if (step + 2 < steps) {
            if (odd)
                load(32 * (step + 2), r1);
            else
                load(32 * (step + 2), r0);
        }
        mma(odd);
        if (step + 1 < steps) {
            if (odd)
                store(0, r0);
            else
                store(1, r1);
        }
        __syncthreads();
    }
    // Epilogue: each wave writes bf16(acc + bias) of its 64 x WN tile to LDS
    // (row stride WN + 8), then stores 8 consecutive channels per lane.
    const int m_base = m_wg + wm * 64, n_base = n_wg + wn * WN;
    constexpr int RS = WN + 8;
    bf16* tile = reinterpret_cast<bf16*>(smem) + wave * 64 * RS;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int j = 0; j < TN; ++j)
                tile[(16 * i + 4 * g + q) * RS + 16 * j + r] = f2bf(epi.pre(n_base + 16 * j + r, acc[i][j][q]));
    __syncthreads();
    constexpr int C8 = WN / 8;
#pragma unroll
    for (int s = 0; s < 64 * C8 / 64; ++s) {
        const int idx = s * 64 + lane, rr = idx / C8, c8 = idx % C8;
        epi.put8(m_base + rr, n_base + 8 * c8, *reinterpret_cast<const v8bf*>(tile + rr * RS + 8 * c8));
    }
}

// Stages COUNT 16-byte items into LDS: item q is read by load(q) and stored
// by put(q, v).  Each thread issues up to kStageBatch loads before their
// stores, so they are in flight together: the loads go through generic
// pointers, which the compiler cannot tell apart from the LDS stores, and a
// load-store loop would wait for each load in turn.  The batch bounds the
// registers held (occupancy of the MFMA kernels).
constexpr int kStageBatch = 6;
template <int COUNT, class Load, class Put>
__device__ __forceinline__ void stage_items(Load&& load, Put&& put) {
    constexpr int PER = (COUNT + kT - 1) / kT;
    for (int s0 = 0; s0 < PER; s0 += kStageBatch) {
        v8bf v[kStageBatch];
#pragma unroll
        for (int s = 0; s < kStageBatch; ++s) {
            const int q = static_cast<int>(threadIdx.x) + (s0 + s) * kT;
            v[s] = s0 + s < PER && q < COUNT ? load(q) : v8bf{};
        }
#pragma unroll
        for (int s = 0; s < kStageBatch; ++s) {
            const int q = static_cast<int>(threadIdx.x) + (s0 + s) * kT;
            if (s0 + s < PER && q < COUNT) put(q, v[s]);
        }
    }
}

// ------------------------------------------ the halo-tiled conv (stride 1)
// For stride-1 convs on wide images with few output channels (the decoder's
// full- and half-resolution convs), the im2col operand dominates: each input
// pixel is read once per tap.  Here a workgroup owns a TH x TW pixel tile
// (TH * TW = 256, a wave per 64 pixels) and all BN output channels, and
// stages the (TH+2) x (TW+2) input patch once per 32-channel chunk, together
// with that chunk's weights for all 9 taps; the 9 K steps of the chunk read
// their A fragments from the patch at the tap's offset.  The two-phase
// epilogue of k_conv follows.
struct HaloArgs {
    const bf16* x;  // NHWC, in.c channels (a multiple of 32)
    const bf16* w;  // [cout][9][cin]
    Geo in;
    int cout, tiles_x, tiles_y;
};

// Epilogue of the halo kernels, as in k_conv: bf16(acc + bias) through LDS,
// then 8 consecutive channels per lane; tile pixel p -> global output row m.
template <int TW, int BN, class Epi>
__device__ __forceinline__ void halo_epilogue(char* smem, const v4f (&acc)[4][BN / 16], const Epi& epi, const Geo& in,
                                              int b, int y0, int x0) {
    constexpr int RS = BN + 8, TN = BN / 16, C8 = BN / 8;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, r = lane & 15, g = lane >> 4;
    bf16* tile = reinterpret_cast<bf16*>(smem) + wave * 64 * RS;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int qq = 0; qq < 4; ++qq)
#pragma unroll
            for (int j = 0; j < TN; ++j)
                tile[(16 * i + 4 * g + qq) * RS + 16 * j + r] = f2bf(epi.pre(16 * j + r, acc[i][j][qq]));
    __syncthreads();
#pragma unroll
    for (int s = 0; s < C8; ++s) {
        const int idx = s * 64 + lane, rr = idx / C8, c8 = idx % C8;
        const int p = 64 * wave + rr;
        const int m = (b * in.h + y0 + p / TW) * in.w + x0 + p % TW;
        epi.put8(m, 8 * c8, *reinterpret_cast<const v8bf*>(tile + rr * RS + 8 * c8));
    }
}

template <int TW, int BN, class Epi>
__global__ __launch_bounds__(kT) void k_conv_halo(HaloArgs c, Epi epi) {
    constexpr int TH = 256 / TW, PW = TW + 2, PP = PW * (TH + 2), TN = BN / 16;
    // The patch and the weights are row-major in LDS, 80 bytes a row (the 4
    // chunks of a pixel's or an output channel's 32 channels, then padding),
    // and four consecutive lanes stage the 4 chunks of one row: each load
    // reads 64 contiguous bytes per row.
    constexpr int RB = 80;
    constexpr int kPatch = PP * RB, kW = 9 * BN * RB, kEpi = 4 * 64 * (BN + 8) * 2;
    constexpr int kBytes = kPatch + kW > kEpi ? kPatch + kW : kEpi;
    __shared__ __align__(16) char smem[kBytes];
    auto lp = [&](int pix, int gg) -> v8bf& { return *reinterpret_cast<v8bf*>(smem + pix * RB + 16 * gg); };
    auto lw = [&](int tap, int n, int gg) -> v8bf& {
        return *reinterpret_cast<v8bf*>(smem + kPatch + (tap * BN + n) * RB + 16 * gg);
    };

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, r = lane & 15, g = lane >> 4;
    const int nb = gridDim.x;
    int t = blockIdx.x;
    if ((nb & 7) == 0) t = (t & 7) * (nb >> 3) + (t >> 3);
    const int tx = t % c.tiles_x, q = t / c.tiles_x, ty = q % c.tiles_y, b = q / c.tiles_y;
    const int x0 = tx * TW, y0 = ty * TH, cin = c.in.c, K = 9 * cin;

    v4f acc[4][TN] = {};
    for (int c0 = 0; c0 < cin; c0 += 32) {
        // Stage the patch (zero outside the image) and the chunk's weights.
        // items [0, 4 PP): the patch, (pixel, chunk); then the weights,
        // (tap, n, chunk).
        stage_items<4 * PP + 4 * 9 * BN>(
            [&](int qq) {
                v8bf v = {};
                const int gg = qq & 3, row = qq >> 2;
                if (row < PP) {
                    const int iy = y0 - 1 + row / PW, ix = x0 - 1 + row % PW;
                    if (iy >= 0 && iy < c.in.h && ix >= 0 && ix < c.in.w)
                        v = *reinterpret_cast<const v8bf*>(c.x + at(c.in, b, iy, ix) + c0 + 8 * gg);
                } else {
                    const int rest = row - PP, tap = rest / BN, n = rest - tap * BN;
                    if (n < c.cout)
                        v = *reinterpret_cast<const v8bf*>(c.w + static_cast<size_t>(n) * K + tap * cin + c0 + 8 * gg);
                }
                return v;
            },
            [&](int qq, const v8bf& v) {
                const int gg = qq & 3, row = qq >> 2;
                if (row < PP) {
                    lp(row, gg) = v;
                } else {
                    const int rest = row - PP, tap = rest / BN;
                    lw(tap, rest - tap * BN, gg) = v;
                }
            });
        __syncthreads();
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) {
            const int ky = tap / 3, kx = tap % 3;
            v8bf fa[4], fb[TN];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int p = 64 * wave + 16 * i + r;  // pixel of the tile
                fa[i] = lp((p / TW + ky) * PW + p % TW + kx, g);
            }
#pragma unroll
            for (int j = 0; j < TN; ++j) fb[j] = lw(tap, 16 * j + r, g);
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    halo_epilogue<TW, BN>(smem, acc, epi, c.in, b, y0, x0);
}

// Cin = 8 (the network's first conv): the patch is one 16-byte chunk per
// pixel, staged once, and the 72 K values are 3 steps of 32 in which lane
// group g takes tap 4 s + g (nothing past tap 8) with its 8 channels, in A
// and in B alike.
template <int TW, int BN, class Epi>
__global__ __launch_bounds__(kT) void k_conv_halo8(HaloArgs c, Epi epi) {
    constexpr int TH = 256 / TW, PW = TW + 2, PP = PW * (TH + 2), TN = BN / 16;
    constexpr int kPatch = PP * 16, kW = 9 * BN * 16, kEpi = 4 * 64 * (BN + 8) * 2;
    constexpr int kBytes = kPatch + kW > kEpi ? kPatch + kW : kEpi;
    __shared__ __align__(16) char smem[kBytes];
    v8bf* lp = reinterpret_cast<v8bf*>(smem);
    v8bf(*lw)[BN] = reinterpret_cast<v8bf(*)[BN]>(smem + kPatch);

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, r = lane & 15, g = lane >> 4;
    const int nb = gridDim.x;
    int t = blockIdx.x;
    if ((nb & 7) == 0) t = (t & 7) * (nb >> 3) + (t >> 3);
    const int tx = t % c.tiles_x, q = t / c.tiles_x, ty = q % c.tiles_y, b = q / c.tiles_y;
    const int x0 = tx * TW, y0 = ty * TH;
    stage_items<PP + 9 * BN>(  // the patch, then the weights
        [&](int qq) {
            v8bf v = {};
            if (qq < PP) {
                const int iy = y0 - 1 + qq / PW, ix = x0 - 1 + qq % PW;
                if (iy >= 0 && iy < c.in.h && ix >= 0 && ix < c.in.w)
                    v = *reinterpret_cast<const v8bf*>(c.x + at(c.in, b, iy, ix));
            } else {
                const int q = qq - PP, tap = q / BN, n = q - tap * BN;
                if (n < c.cout) v = *reinterpret_cast<const v8bf*>(c.w + static_cast<size_t>(n) * 72 + tap * 8);
            }
            return v;
        },
        [&](int qq, const v8bf& v) {
            if (qq < PP) {
                lp[qq] = v;
            } else {
                const int q = qq - PP, tap = q / BN;
                lw[tap][q - tap * BN] = v;
            }
        });
    __syncthreads();
    v4f acc[4][TN] = {};
#pragma unroll
    for (int step = 0; step < 3; ++step) {
        const int tap = 4 * step + g, ky = tap / 3, kx = tap % 3;
        v8bf fa[4] = {}, fb[TN] = {};
        if (tap < 9) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int p = 64 * wave + 16 * i + r;
                fa[i] = lp[(p / TW + ky) * PW + p % TW + kx];
            }
#pragma unroll
            for (int j = 0; j < TN; ++j) fb[j] = lw[tap][16 * j + r];
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
    halo_epilogue<TW, BN>(smem, acc, epi, c.in, b, y0, x0);
}

// The halo-tiled kernels apply to stride-1 convs with Cin 8 or a multiple of
// 32, Cout 32 or 64, and image sides divisible by the tile.
template <class Epi>
bool launch_conv_halo(const bf16* x, Geo in, const bf16* w, int cout, Epi epi, hipStream_t s, hipError_t* err) {
    if ((in.c != 8 && in.c % 32 != 0) || (cout != 32 && cout != 64)) return false;
    const int tw = in.w % 64 == 0 && in.h % 4 == 0 ? 64 : (in.w % 32 == 0 && in.h % 8 == 0 ? 32 : 0);
    if (tw == 0) return false;
    HaloArgs c{x, w, in, cout, in.w / tw, in.h / (256 / tw)};
    const unsigned grid = static_cast<unsigned>(c.tiles_x) * c.tiles_y * in.b;
    if (in.c == 8) {
        if (tw == 64 && cout == 32) k_conv_halo8<64, 32, Epi><<<grid, kT, 0, s>>>(c, epi);
        if (tw == 64 && cout == 64) k_conv_halo8<64, 64, Epi><<<grid, kT, 0, s>>>(c, epi);
        if (tw == 32 && cout == 32) k_conv_halo8<32, 32, Epi><<<grid, kT, 0, s>>>(c, epi);
        if (tw == 32 && cout == 64) k_conv_halo8<32, 64, Epi><<<grid, kT, 0, s>>>(c, epi);
        *err = hipGetLastError();
        return true;
    }
    if (tw == 64 && cout == 32) k_conv_halo<64, 32, Epi><<<grid, kT, 0, s>>>(c, epi);
    if (tw == 64 && cout == 64) k_conv_halo<64, 64, Epi><<<grid, kT, 0, s>>>(c, epi);
    if (tw == 32 && cout == 32) k_conv_halo<32, 32, Epi><<<grid, kT, 0, s>>>(c, epi);
    if (tw == 32 && cout == 64) k_conv_halo<32, 64, Epi><<<grid, kT, 0, s>>>(c, epi);
    *err = hipGetLastError();
    return true;
}

template <class Epi>
hipError_t launch_conv(const bf16* x, Geo in, int stride, const bf16* w, int cout, Epi epi, hipStream_t s) {
    hipError_t err = hipSuccess;
    if (stride == 1 && launch_conv_halo(x, in, w, cout, epi, s, &err)) return err;
    int lg = 0;
    while ((1 << lg) < in.c) ++lg;
    const int oh = in.h / stride, ow = in.w / stride, M = in.b * oh * ow;
    ConvArgs c{x, w, in, oh, ow, stride, lg, cout, M, 0, 0};
    auto go = [&](auto kernel, int wn, int waves_n) {
        c.tiles_m = (M + 64 * (4 / waves_n) - 1) / (64 * (4 / waves_n));
        c.tiles_n = (cout + wn * waves_n - 1) / (wn * waves_n);
        kernel<<<c.tiles_m * c.tiles_n, kT, 0, s>>>(c, epi);
        return hipGetLastError();
    };
    if (cout >= 128) return go(k_conv<64, 2, Epi>, 64, 2);
    if (cout >= 64) return go(k_conv<64, 1, Epi>, 64, 1);
    return go(k_conv<32, 1, Epi>, 32, 1);
}

inline bool pow2_at_least_8(int c) { return c >= 8 && (c & (c - 1)) == 0; }

// Sum of the split-K partials in split order: column NW of each row is the
// bias gradient, the others the weight gradient [M][NW].
__global__ __launch_bounds__(kT) void k_reduce_splits(const float* part, int splits, int M, int NW, float* dw,
                                                      float* db) {
    const int N = NW + 1;
    const size_t count = static_cast<size_t>(M) * N;
    for (size_t i = blockIdx.x * static_cast<size_t>(kT) + threadIdx.x; i < count; i += static_cast<size_t>(gridDim.x) * kT) {
        float s = 0.0f;
        for (int z = 0; z < splits; ++z) s += part[static_cast<size_t>(z) * count + i];
        const int m = static_cast<int>(i / N), n = static_cast<int>(i % N);
        if (n < NW)
            dw[static_cast<size_t>(m) * NW + n] = s;
        else
            db[m] = s;
    }
}

// ------------------------------------------ the head (Cin -> C, FP32)
// The head is a 3x3 conv from the finest network level (side m = n / sub)
// to C channels, yc: [b][m][m][C] FP32; w: [C][9][cin]; bias: [C].  With
// sub = 1, C = 1 and yc is the output y itself.  With sub = 2, C = 5 and
// the output map below turns yc into the n x n field y.

// Where the head (or the half-resolution output map) writes the warm start
// itself, u0 = f scale pos y in FP64, instead of y.
struct HeadOut {
    const float* pos;
    const double* f;
    double scale;
    double* const* u0;
};
constexpr int kHeadC = 32;

// dX[pixel][ci] = sum_tap sum_r w[r][tap][ci] dyc[pixel - tap][r].
__global__ __launch_bounds__(kT) void k_head_bwd_data(const float* dyc, Geo in, int C, const float* w, bf16* dx) {
    const size_t P = static_cast<size_t>(in.b) * in.h * in.w;
    for (size_t q = blockIdx.x * static_cast<size_t>(kT) + threadIdx.x; q < P * in.c; q += static_cast<size_t>(gridDim.x) * kT) {
        const int ci = static_cast<int>(q % in.c);
        const size_t p = q / in.c;
        const int x = static_cast<int>(p % in.w), t = static_cast<int>(p / in.w), y = t % in.h, b = t / in.h;
        float s = 0.0f;
        for (int tap = 0; tap < 9; ++tap) {
            const int oy = y + 1 - tap / 3, ox = x + 1 - tap % 3;
            if (oy < 0 || oy >= in.h || ox < 0 || ox >= in.w) continue;
            const float* d = dyc + ((static_cast<size_t>(b) * in.h + oy) * in.w + ox) * C;
            for (int r = 0; r < C; ++r) s += w[(r * 9 + tap) * in.c + ci] * d[r];
        }
        dx[q] = f2bf(s);
    }
}

// Head weight gradient, stage 1: workgroup g sums pixels [g per, (g+1) per)
// for all C (9 Cin + 1) weights and biases (thread i owns entries i,
// i + 256, ...); part[g][r][9 Cin + 1], the bias last.
__global__ __launch_bounds__(kT) void k_head_dw_part(const bf16* x, Geo in, int C, const float* dyc, int per,
                                                     float* part) {
    const int nw = 9 * in.c;
    const size_t P = static_cast<size_t>(in.b) * in.h * in.w;
    const size_t p0 = static_cast<size_t>(blockIdx.x) * per;
    for (int e = threadIdx.x; e < C * (nw + 1); e += kT) {
        const int r = e / (nw + 1), widx = e % (nw + 1);
        const int ci = widx % in.c, tap = widx / in.c;
        float s = 0.0f;
        for (size_t p = p0; p < p0 + per && p < P; ++p) {
            const float d = dyc[p * C + r];
            if (widx == nw) {
                s += d;
                continue;
            }
            const int ox = static_cast<int>(p % in.w), t = static_cast<int>(p / in.w), oy = t % in.h, b = t / in.h;
            const int iy = oy + tap / 3 - 1, ix = ox + tap % 3 - 1;
            if (iy < 0 || iy >= in.h || ix < 0 || ix >= in.w) continue;
            s += d * bf2f(x[at(in, b, iy, ix) + ci]);
        }
        part[(static_cast<size_t>(blockIdx.x) * C + r) * (nw + 1) + widx] = s;
    }
}

// The head on the matrix cores: the FP32 weights are split as w = hi + lo
// with hi = bf16(w), lo = bf16(w - hi) (|w - hi - lo| <= 2^-18 |w|), and
// each K step multiplies the bf16 activations by hi and by lo into the same
// FP32 accumulator (bf16 x bf16 products are exact in FP32).  A workgroup
// owns a TH x TW pixel tile (TH TW = 256) and stages its (TH+2) x (TW+2)
// patch of 32 channels once, as k_conv_halo does; the B operand is one
// 16-column tile of which the first C columns are the outputs.  Output:
// yc [b][h][w][C] FP32 = acc + bias, or with out.u0 (C = 1) the warm start.
// whl: [2][C][9][32] bf16 (hi, then lo).
template <int TW, int C>
__global__ __launch_bounds__(kT) void k_head_mfma(const bf16* x, Geo in, const bf16* whl, const float* bias,
                                                  float* yc, HeadOut out) {
    constexpr int TH = 256 / TW, PW = TW + 2, PP = PW * (TH + 2), RB = 80;
    constexpr int kPatch = PP * RB, kW = 2 * 9 * C * RB, kOut = 256 * C * 4;
    constexpr int kBytes = kPatch + kW > kOut ? kPatch + kW : kOut;
    __shared__ __align__(16) char smem[kBytes];
    auto lp = [&](int pix, int gg) -> v8bf& { return *reinterpret_cast<v8bf*>(smem + pix * RB + 16 * gg); };
    auto lw = [&](int part, int tap, int n, int gg) -> v8bf& {
        return *reinterpret_cast<v8bf*>(smem + kPatch + ((part * 9 + tap) * C + n) * RB + 16 * gg);
    };
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, r = lane & 15, g = lane >> 4;
    const int tiles_x = in.w / TW, tiles_y = in.h / TH, nb = gridDim.x;
    int t = blockIdx.x;
    if ((nb & 7) == 0) t = (t & 7) * (nb >> 3) + (t >> 3);
    const int tx = t % tiles_x, q = t / tiles_x, ty = q % tiles_y, b = q / tiles_y;
    const int x0 = tx * TW, y0 = ty * TH;
    // items [0, 4 PP): the patch, (pixel, chunk); then the weights,
    // (part, tap, n, chunk) for the C live columns.
    stage_items<4 * PP + 4 * 2 * 9 * C>(
        [&](int qq) {
            v8bf v = {};
            const int gg = qq & 3, row = qq >> 2;
            if (row < PP) {
                const int iy = y0 - 1 + row / PW, ix = x0 - 1 + row % PW;
                if (iy >= 0 && iy < in.h && ix >= 0 && ix < in.w)
                    v = *reinterpret_cast<const v8bf*>(x + at(in, b, iy, ix) + 8 * gg);
            } else {
                const int rest = row - PP, n = rest % C, pt = rest / C;  // pt = part * 9 + tap
                v = *reinterpret_cast<const v8bf*>(whl + ((pt / 9) * C + n) * 288 + (pt % 9) * 32 + 8 * gg);
            }
            return v;
        },
        [&](int qq, const v8bf& v) {
            const int gg = qq & 3, row = qq >> 2;
            if (row < PP) {
                lp(row, gg) = v;
            } else {
                const int rest = row - PP, pt = rest / C;
                lw(pt / 9, pt % 9, rest % C, gg) = v;
            }
        });
    __syncthreads();
    v4f acc[4] = {};
#pragma unroll
    for (int tap = 0; tap < 9; ++tap) {
        const int ky = tap / 3, kx = tap % 3;
        // columns r >= C of the B tile are zero
        const v8bf bh = r < C ? lw(0, tap, r, g) : v8bf{}, bl = r < C ? lw(1, tap, r, g) : v8bf{};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int p = 64 * wave + 16 * i + r;
            const v8bf a = lp((p / TW + ky) * PW + p % TW + kx, g);
            acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bh, acc[i], 0, 0, 0);
            acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bl, acc[i], 0, 0, 0);
        }
    }
    // Lane (r, g) holds column r of tile rows 4 g .. 4 g + 3: staged as
    // [pixel][C] in LDS, then stored row by row of the tile, consecutive
    // threads on consecutive floats.
    __syncthreads();
    float* st = reinterpret_cast<float*>(smem);
    if (r < C) {
        const float bn = bias[r];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int e = 0; e < 4; ++e) st[(64 * wave + 16 * i + 4 * g + e) * C + r] = acc[i][e] + bn;
    }
    __syncthreads();
    if (C == 1 && out.u0) {
        const int p = tid;
        const size_t k = static_cast<size_t>(y0 + p / TW) * in.w + x0 + p % TW;
        out.u0[b][k] = out.f[b] * out.scale * static_cast<double>(out.pos[k] * st[p]);
        return;
    }
    for (int e = tid; e < 256 * C; e += kT) {
        const int p = e / C, rr = e % C;
        const size_t k = static_cast<size_t>(y0 + p / TW) * in.w + x0 + p % TW;
        yc[(static_cast<size_t>(b) * in.h * in.w + k) * C + rr] = st[e];
    }
}

// hi / lo bf16 parts of the head's FP32 weights: whl[0][i] = bf16(w[i]),
// whl[1][i] = bf16(w[i] - hi).
__global__ __launch_bounds__(kT) void k_head_split(const float* w, int count, bf16* whl) {
    for (int i = blockIdx.x * kT + threadIdx.x; i < count; i += gridDim.x * kT) {
        const float v = w[i];
        const bf16 hi = f2bf(v);
        whl[i] = hi;
        whl[count + i] = f2bf(v - bf2f(hi));
    }
}

// ------------------------------------ the half-resolution output map (sub 2)
// yc: [b][m][m][5] -> y: [b][n][n], n = 2m.  Cell (i, j) = (2I + di, 2J + dj)
// takes the bilinear interpolation of channel 0 at its centre (cell-centred
// grids: weights 3/4 on its own pixel's row / column and 1/4 on the
// neighbour's on its side, the neighbour index clamped to the grid) plus
// its own sub-pixel channel 1 + 2 di + dj.  The interpolation makes a
// smooth output by construction; the sub-pixel channels carry the detail
// inside a block.

// Neighbouring pixel index on the side of fine index i, and the weight pair.
__device__ __forceinline__ int sub2_side(int i, int m) {
    const int I = i >> 1, J = (i & 1) ? I + 1 : I - 1;
    return J < 0 ? 0 : (J >= m ? m - 1 : J);
}

// y (training) or u0 = f scale pos y (inference, with out.u0): one thread
// per pixel (I, J) forms its block's four cells from channel 0 of the 3 x 3
// pixels around it (indices clamped) and its four detail channels, and
// stores each row's two cells as one 16-byte (u0) or 8-byte (y) store.
__global__ __launch_bounds__(kT) void k_sub2_out(const float* yc, int n, int batch, float* y, HeadOut out) {
    const int m = n / 2;
    const size_t per = static_cast<size_t>(m) * m;
    for (size_t q = blockIdx.x * static_cast<size_t>(kT) + threadIdx.x; q < per * batch; q += static_cast<size_t>(gridDim.x) * kT) {
        const size_t b = q / per, k = q % per;
        const int I = static_cast<int>(k / m), J = static_cast<int>(k % m);
        const float* ycb = yc + b * per * 5;
        const int rows[3] = {I > 0 ? I - 1 : 0, I, I + 1 < m ? I + 1 : m - 1};
        const int cols[3] = {J > 0 ? J - 1 : 0, J, J + 1 < m ? J + 1 : m - 1};
        float c[3][3];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int e = 0; e < 3; ++e) c[a][e] = ycb[(static_cast<size_t>(rows[a]) * m + cols[e]) * 5];
        const float* det = ycb + k * 5 + 1;
        const float d0 = det[0], d1 = det[1], d2 = det[2], d3 = det[3];
#pragma unroll
        for (int di = 0; di < 2; ++di) {
            // cell row 2I + di: own row I, side row I - 1 (di 0) or I + 1 (di 1)
            const int side = di ? 2 : 0;
            float v[2];
#pragma unroll
            for (int dj = 0; dj < 2; ++dj) {
                const int cs = dj ? 2 : 0;
                const float row_own = 0.75f * c[1][1] + 0.25f * c[1][cs];
                const float row_side = 0.75f * c[side][1] + 0.25f * c[side][cs];
                v[dj] = (0.75f * row_own + 0.25f * row_side) + (di ? (dj ? d3 : d2) : (dj ? d1 : d0));
            }
            const size_t kf = static_cast<size_t>(2 * I + di) * n + 2 * J;
            if (out.u0) {
                const float2 p2 = *reinterpret_cast<const float2*>(out.pos + kf);
                const double fs = out.f[b] * out.scale;
                double2 u;
                u.x = fs * static_cast<double>(p2.x * v[0]);
                u.y = fs * static_cast<double>(p2.y * v[1]);
                double* dst = out.u0[b] + kf;
                if ((reinterpret_cast<uintptr_t>(dst) & 15) == 0) {
                    *reinterpret_cast<double2*>(dst) = u;
                } else {  // a caller's array that is only 8-byte aligned
                    dst[0] = u.x;
                    dst[1] = u.y;
                }
            } else {
                *reinterpret_cast<float2*>(y + b * static_cast<size_t>(n) * n + kf) = make_float2(v[0], v[1]);
            }
        }
    }
}

// The adjoint: dyc[I][J][1 + r] = dy of the pixel's cell r; dyc[I][J][0] =
// sum over the cells whose interpolation reads pixel (I, J) of their
// weight times dy (fine rows 2I - 1 .. 2I + 2 and the same columns).
__device__ __forceinline__ float sub2_weight(int i, int I, int m) {
    // Weight of coarse index I in fine index i's interpolation (0 if unused).
    float w = 0.0f;
    if ((i >> 1) == I) w += 0.75f;
    if (sub2_side(i, m) == I) w += 0.25f;
    return w;
}

__global__ __launch_bounds__(kT) void k_sub2_adjoint(const float* dy, int n, int batch, float* dyc) {
    const int m = n / 2;
    const size_t per = static_cast<size_t>(m) * m;
    for (size_t q = blockIdx.x * static_cast<size_t>(kT) + threadIdx.x; q < per * batch; q += static_cast<size_t>(gridDim.x) * kT) {
        const size_t b = q / per, k = q % per;
        const int I = static_cast<int>(k / m), J = static_cast<int>(k % m);
        const float* d = dy + b * static_cast<size_t>(n) * n;
        float s = 0.0f;
        for (int i = 2 * I - 1; i <= 2 * I + 2; ++i) {
            if (i < 0 || i >= n) continue;
            const float wi = sub2_weight(i, I, m);
            if (wi == 0.0f) continue;
            float r = 0.0f;
            for (int j = 2 * J - 1; j <= 2 * J + 2; ++j) {
                if (j < 0 || j >= n) continue;
                r += sub2_weight(j, J, m) * d[static_cast<size_t>(i) * n + j];
            }
            s += wi * r;
        }
        float* o = dyc + q * 5;
        o[0] = s;
        for (int r = 0; r < 4; ++r) o[1 + r] = d[static_cast<size_t>(2 * I + (r >> 1)) * n + 2 * J + (r & 1)];
    }
}

// ------------------------------------------------------------ elementwise

// pos[i n + j] = sin(pi (i + 1/2) / n) sin(pi (j + 1/2) / n), rounded to float:
// the position channel of the input and the factor of the output.
__global__ __launch_bounds__(kT) void k_position(int n, float* pos) {
    const double pi = 3.14159265358979323846;
    for (int q = blockIdx.x * kT + threadIdx.x; q < n * n; q += gridDim.x * kT) {
        const int i = q / n, j = q % n;
        pos[q] = static_cast<float>(sin(pi * (i + 0.5) / n) * sin(pi * (j + 0.5) / n));
    }
}

// Input features, NHWC bf16 with 8 channels per pixel of the network's
// finest level (side m = n / sub, sub 1 or 2): the sub x sub cells' (ln a -
// mu) / sd in row-major order, then 1 (the forcing, which enters the output
// linearly), then the position (the mean of the cells' pos, summed in
// row-major order in FP32); the rest is 0.  mu, sd map the two phases ln 3
// and ln 12 to -1 and +1.
__global__ __launch_bounds__(kT) void k_features(const double* const* a, const float* pos, int n, int sub, int batch,
                                                 bf16* x) {
    const int m = n / sub;
    const size_t per = static_cast<size_t>(m) * m;
    const double mu = 0.5 * (log(3.0) + log(12.0)), sd = 0.5 * (log(12.0) - log(3.0));
    for (size_t q = blockIdx.x * static_cast<size_t>(kT) + threadIdx.x; q < per * batch; q += static_cast<size_t>(gridDim.x) * kT) {
        const size_t b = q / per, k = q % per;
        const int I = static_cast<int>(k / m), J = static_cast<int>(k % m);
        v8bf v = {};
        float ps = 0.0f;
        const double* ab = a[b];
        if (sub == 2 && (reinterpret_cast<uintptr_t>(ab) & 15) == 0) {
            // the block's two rows as one 16-byte load each (pos rows as 8)
            const size_t k0 = static_cast<size_t>(2 * I) * n + 2 * J;
            const double2 r0 = *reinterpret_cast<const double2*>(ab + k0);
            const double2 r1 = *reinterpret_cast<const double2*>(ab + k0 + n);
            const float2 p0 = *reinterpret_cast<const float2*>(pos + k0);
            const float2 p1 = *reinterpret_cast<const float2*>(pos + k0 + n);
            v[0] = f2bf(static_cast<float>((log(r0.x) - mu) / sd));
            v[1] = f2bf(static_cast<float>((log(r0.y) - mu) / sd));
            v[2] = f2bf(static_cast<float>((log(r1.x) - mu) / sd));
            v[3] = f2bf(static_cast<float>((log(r1.y) - mu) / sd));
            ps = ((p0.x + p0.y) + p1.x) + p1.y;
        } else {
            for (int r = 0; r < sub * sub; ++r) {
                const size_t kf = static_cast<size_t>(sub * I + r / sub) * n + sub * J + r % sub;
                v[r] = f2bf(static_cast<float>((log(ab[kf]) - mu) / sd));
                ps += pos[kf];
            }
        }
        v[sub * sub] = f2bf(1.0f);
        v[sub * sub + 1] = f2bf(ps * (1.0f / static_cast<float>(sub * sub)));
        *reinterpret_cast<v8bf*>(x + 8 * q) = v;
    }
}

// Decoder merge: z = up2(c) + skip, act = gelu(z) (nearest upsampling), 8
// channels per thread.
__global__ __launch_bounds__(kT) void k_up_add_gelu(const bf16* c, Geo cg, const bf16* skip, bf16* z, bf16* act) {
    const int H = cg.h * 2, W = cg.w * 2, c8 = cg.c / 8;
    const size_t count = static_cast<size_t>(cg.b) * H * W * c8;
    for (size_t q = blockIdx.x * static_cast<size_t>(kT) + threadIdx.x; q < count; q += static_cast<size_t>(gridDim.x) * kT) {
        const int ch = static_cast<int>(q % c8) * 8;
        const size_t p = q / c8;
        const int x = static_cast<int>(p % W), t = static_cast<int>(p / W), y = t % H, b = t / H;
        const v8bf cv = *reinterpret_cast<const v8bf*>(c + at(cg, b, y / 2, x / 2) + ch);
        const v8bf sv = *reinterpret_cast<const v8bf*>(skip + 8 * q);
        v8bf zv, av;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            zv[e] = f2bf(bf2f(cv[e]) + bf2f(sv[e]));
            av[e] = f2bf(gelu(bf2f(zv[e])));
        }
        if (z) *reinterpret_cast<v8bf*>(z + 8 * q) = zv;
        *reinterpret_cast<v8bf*>(act + 8 * q) = av;
    }
}

// d <- d * gelu'(z), in place.
__global__ __launch_bounds__(kT) void k_gelu_bwd(bf16* d, const bf16* z, size_t count) {
    for (size_t q = blockIdx.x * static_cast<size_t>(kT) + threadIdx.x; q < count; q += static_cast<size_t>(gridDim.x) * kT)
        d[q] = f2bf(bf2f(d[q]) * gelu_grad(bf2f(z[q])));
}

// Gradient of the upsampled branch: dc = sum of dz over each 2x2 block.
__global__ __launch_bounds__(kT) void k_down_sum(const bf16* dz, Geo cg, bf16* dc) {
    const size_t count = static_cast<size_t>(cg.b) * cg.h * cg.w * cg.c;
    const Geo fine = {cg.b, cg.h * 2, cg.w * 2, cg.c};
    for (size_t q = blockIdx.x * static_cast<size_t>(kT) + threadIdx.x; q < count; q += static_cast<size_t>(gridDim.x) * kT) {
        const int ch = static_cast<int>(q % cg.c);
        const size_t p = q / cg.c;
        const int x = static_cast<int>(p % cg.w), t = static_cast<int>(p / cg.w), y = t % cg.h, b = t / cg.h;
        float s = 0.0f;
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) s += bf2f(dz[at(fine, b, 2 * y + dy, 2 * x + dx) + ch]);
        dc[q] = f2bf(s);
    }
}

// dst <- dst + src.
__global__ __launch_bounds__(kT) void k_add(bf16* dst, const bf16* src, size_t count) {
    for (size_t q = blockIdx.x * static_cast<size_t>(kT) + threadIdx.x; q < count; q += static_cast<size_t>(gridDim.x) * kT)
        dst[q] = f2bf(bf2f(dst[q]) + bf2f(src[q]));
}

// [P][C] -> [C][P].
__global__ __launch_bounds__(kT) void k_transpose(const bf16* in, int P, int C, bf16* out) {
    __shared__ bf16 tile[32][33];
    const int p0 = blockIdx.x * 32, c0 = blockIdx.y * 32, tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int y = ty; y < 32; y += 8)
        if (p0 + y < P && c0 + tx < C) tile[y][tx] = in[static_cast<size_t>(p0 + y) * C + c0 + tx];
    __syncthreads();
    for (int y = ty; y < 32; y += 8)
        if (c0 + y < C && p0 + tx < P) out[static_cast<size_t>(c0 + y) * P + p0 + tx] = tile[tx][y];
}

// Training loss against t = u / scale with e = pos y - t:
//   L = (sum_cells e^2 + lam e^T A e) / cells,
// A the finite-volume operator of the problem's permeability (harmonic
// interior faces, boundary faces carrying the cell's own a), so the second
// term is the energy norm in which CG converges; dy = pos dL/de with
// dL/de = 2 (e + lam A e) / cells.  Each workgroup writes its partial loss.
__global__ __launch_bounds__(kT) void k_loss(const float* y, const double* const* u, const double* const* a,
                                             const float* pos, int n, int batch, double scale, float lam, float* dy,
                                             float* partial) {
    __shared__ float red[kT];
    const size_t per = static_cast<size_t>(n) * n, count = per * batch;
    const float inv_c = 1.0f / static_cast<float>(count);
    auto err = [&](size_t b, size_t k, size_t q) { return pos[k] * y[q] - static_cast<float>(u[b][k] / scale); };
    auto hm = [](float x, float z) { return 2.0f * x * z / (x + z); };
    float s = 0.0f;
    for (size_t q = blockIdx.x * static_cast<size_t>(kT) + threadIdx.x; q < count; q += static_cast<size_t>(gridDim.x) * kT) {
        const size_t b = q / per, k = q % per;
        const int i = static_cast<int>(k / n), j = static_cast<int>(k % n);
        const double* ab = a[b];
        const float e = err(b, k, q), ac = static_cast<float>(ab[k]);
        float ae = 0.0f, energy = 0.0f;  // (A e)_c, and the faces this cell owns
        if (j + 1 < n) {
            const float d = e - err(b, k + 1, q + 1), f = hm(ac, static_cast<float>(ab[k + 1]));
            ae += f * d;
            energy += f * d * d;
        } else {
            ae += ac * e;
            energy += ac * e * e;
        }
        if (j > 0)
            ae += hm(ac, static_cast<float>(ab[k - 1])) * (e - err(b, k - 1, q - 1));
        else {
            ae += ac * e;
            energy += ac * e * e;
        }
        if (i + 1 < n) {
            const float d = e - err(b, k + n, q + n), f = hm(ac, static_cast<float>(ab[k + n]));
            ae += f * d;
            energy += f * d * d;
        } else {
            ae += ac * e;
            energy += ac * e * e;
        }
        if (i > 0)
            ae += hm(ac, static_cast<float>(ab[k - n])) * (e - err(b, k - n, q - n));
        else {
            ae += ac * e;
            energy += ac * e * e;
        }
        dy[q] = 2.0f * (e + lam * ae) * inv_c * pos[k];
        s += (e * e + lam * energy) * inv_c;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = kT / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Adam on the FP32 masters; refreshes the bf16 copy (wb may be null).
__global__ __launch_bounds__(kT) void k_adam(float* w, const float* g, float* m, float* v, bf16* wb, size_t count,
                                             float lr, float b1, float b2, float eps, float c1, float c2) {
    for (size_t i = blockIdx.x * static_cast<size_t>(kT) + threadIdx.x; i < count; i += static_cast<size_t>(gridDim.x) * kT) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi, vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        // The step ratio m^ / (sqrt(v^) + eps) is at most about 1 while the
        // gradient's scale holds; it is bounded to [-1, 1] so that a
        // gradient arriving where v^ is still tiny cannot move a weight far
        // in one step (the blow-ups that stalled half-resolution training).
        const float ratio = fminf(fmaxf((mi / c1) / (sqrtf(vi / c2) + eps), -1.0f), 1.0f);
        const float wi = w[i] - lr * ratio;
        w[i] = wi;
        if (wb) wb[i] = f2bf(wi);
    }
}

__global__ __launch_bounds__(kT) void k_to_bf16(const float* in, bf16* out, size_t count) {
    for (size_t i = blockIdx.x * static_cast<size_t>(kT) + threadIdx.x; i < count; i += static_cast<size_t>(gridDim.x) * kT)
        out[i] = f2bf(in[i]);
}

// WF[ci][8 - tap][co] = W[co][tap][ci]: the weights of the backward-data conv.
__global__ __launch_bounds__(kT) void k_weight_t(const bf16* w, int cout, int cin, bf16* wt) {
    const size_t count = static_cast<size_t>(cout) * 9 * cin;
    for (size_t i = blockIdx.x * static_cast<size_t>(kT) + threadIdx.x; i < count; i += static_cast<size_t>(gridDim.x) * kT) {
        const int ci = static_cast<int>(i % cin), tap = static_cast<int>((i / cin) % 9), co = static_cast<int>(i / (9 * cin));
        wt[(static_cast<size_t>(ci) * 9 + (8 - tap)) * cout + co] = w[i];
    }
}

// Lane-layout check of the MFMA: C (16 x 16) = A (16 x 32) * B (32 x 16), row-major.
__global__ void k_mfma_probe(const bf16* A, const bf16* B, float* C) {
    const int lane = threadIdx.x, r = lane & 15, kq = (lane >> 4) * 8;
    v8bf a, b;
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        a[t] = A[r * 32 + kq + t];
        b[t] = B[(kq + t) * 16 + r];
    }
    v4f acc = {};
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
#pragma unroll
    for (int t = 0; t < 4; ++t) C[((lane >> 4) * 4 + t) * 16 + r] = acc[t];
}

inline unsigned ew_grid(size_t count) {
    const size_t g = (count + kT - 1) / kT;
    return static_cast<unsigned>(g < 8192 ? (g ? g : 1) : 8192);
}

}  // namespace darcy_nn

using namespace darcy_nn;

// ------------------------------------------------------------ launchers
// All pointers are device pointers; shapes were checked by the caller.
// The head conv: x [b][h][w][cin] -> yc [b][h][w][C], C 1 or 5, on the
// matrix cores (k_head_mfma) with the hi / lo weights whl
// (darcy_nn_head_split); 32 channels and a side that is a multiple of 32.
// With u0 (C = 1 only) the head stores the warm start itself.
template <int C>
static bool launch_head_mfma(const bf16* x, Geo in, const bf16* whl, const float* bias, float* yc, HeadOut out,
                      hipStream_t s) {
    if (!whl || in.c != kHeadC) return false;
    const int tw = in.w % 64 == 0 && in.h % 4 == 0 ? 64 : (in.w % 32 == 0 && in.h % 8 == 0 ? 32 : 0);
    if (tw == 0) return false;
    const unsigned grid = static_cast<unsigned>(in.w / tw) * (in.h / (256 / tw)) * in.b;
    if (tw == 64)
        k_head_mfma<64, C><<<grid, kT, 0, s>>>(x, in, whl, bias, yc, out);
    else
        k_head_mfma<32, C><<<grid, kT, 0, s>>>(x, in, whl, bias, yc, out);
    return true;
}

static hipError_t head_conv(const void* x, Geo in, int C, const void* whl, const float* bias, float* yc, HeadOut out,
                            hipStream_t s) {
    if (C != 1 && C != 5) return hipErrorInvalidValue;
    const bf16* xb = static_cast<const bf16*>(x);
    const bf16* hl = static_cast<const bf16*>(whl);
    const bool done = C == 1 ? launch_head_mfma<1>(xb, in, hl, bias, yc, out, s)
                             : launch_head_mfma<5>(xb, in, hl, bias, yc, out, s);
    // The network's finest level always has 32 channels and a side that is a
    // multiple of 64, so every head it launches fits the tiles.
    return done ? hipGetLastError() : hipErrorInvalidValue;
}

extern "C" {

hipError_t darcy_nn_mfma_probe(const void* A, const void* B, float* C, hipStream_t s) {
    k_mfma_probe<<<1, 64, 0, s>>>(static_cast<const bf16*>(A), static_cast<const bf16*>(B), C);
    return hipGetLastError();
}

// z / act: [b][h/stride][w/stride][cout] (either may be null).
hipError_t darcy_nn_conv_fwd(const void* x, int b, int h, int w, int cin, const void* wgt, const float* bias, int cout,
                             int stride, void* z, void* act, hipStream_t s) {
    const int oh = h / stride, ow = w / stride, M = b * oh * ow;
    EpiFwd epi{static_cast<bf16*>(z), static_cast<bf16*>(act), bias, M, cout};
    if (pow2_at_least_8(cin))
        return launch_conv(static_cast<const bf16*>(x), Geo{b, h, w, cin}, stride, static_cast<const bf16*>(wgt), cout,
                           epi, s);
    FwdA ga{static_cast<const bf16*>(x), Geo{b, h, w, cin}, oh, ow, stride, M};
    RowsB gb{static_cast<const bf16*>(wgt), 9 * cin, cout};
    const dim3 grid((M + kTile - 1) / kTile, (cout + kTile - 1) / kTile, 1);
    k_gemm<<<grid, kT, 0, s>>>(9 * cin, 9 * cin, ga, gb, epi);
    return hipGetLastError();
}

// Decoder conv (stride 1) fused with its merge: act / z (z may be null) are
// [b][2h][2w][cout] = gelu / value of up2(conv(x) + bias) + skip.
hipError_t darcy_nn_conv_fwd_merge(const void* x, int b, int h, int w, int cin, const void* wgt, const float* bias,
                                   int cout, const void* skip, void* z, void* act, hipStream_t s) {
    if (!pow2_at_least_8(cin)) return hipErrorInvalidValue;
    EpiMerge epi{bias, static_cast<const bf16*>(skip), static_cast<bf16*>(z), static_cast<bf16*>(act), b * h * w, cout,
                 h, w};
    return launch_conv(static_cast<const bf16*>(x), Geo{b, h, w, cin}, 1, static_cast<const bf16*>(wgt), cout, epi, s);
}

// dx: [b][h][w][cin]; dz: [b][h/stride][w/stride][cout]; wt: WF[cin][9][cout]
// (flipped taps, from darcy_nn_weight_t).  Stride 1 is the forward conv of
// dz with WF.
hipError_t darcy_nn_conv_bwd_data(const void* dz, int b, int h, int w, int cin, const void* wt, int cout, int stride,
                                  void* dx, hipStream_t s) {
    const int M = b * h * w;
    if (stride == 1 && pow2_at_least_8(cout)) {
        EpiFwd epi{static_cast<bf16*>(dx), nullptr, nullptr, M, cin};
        return launch_conv(static_cast<const bf16*>(dz), Geo{b, h, w, cout}, 1, static_cast<const bf16*>(wt), cin, epi,
                           s);
    }
    BwdDataA ga{static_cast<const bf16*>(dz), Geo{b, h / stride, w / stride, cout}, h, w, stride, M};
    RowsB gb{static_cast<const bf16*>(wt), 9 * cout, cin};
    EpiStore epi{static_cast<bf16*>(dx), M, cin};
    const dim3 grid((M + kTile - 1) / kTile, (cin + kTile - 1) / kTile, 1);
    k_gemm<<<grid, kT, 0, s>>>(9 * cout, 9 * cout, ga, gb, epi);
    return hipGetLastError();
}

// Weight and bias gradient.  dzt: [cout][P] (P = b * oh * ow); part:
// splits * cout * (9 cin + 1) floats; dw: [cout][9][cin]; db: [cout].
hipError_t darcy_nn_conv_bwd_weight(const void* x, int b, int h, int w, int cin, const void* dzt, int cout, int stride,
                                    int splits, float* part, float* dw, float* db, hipStream_t s) {
    const int oh = h / stride, ow = w / stride, P = b * oh * ow, NW = 9 * cin;
    RowsA ga{static_cast<const bf16*>(dzt), P, cout};
    BwdWtB gb{static_cast<const bf16*>(x), Geo{b, h, w, cin}, oh, ow, stride, P, NW};
    EpiPartial epi{part, cout, NW + 1};
    int kchunk = (P + splits - 1) / splits;
    kchunk = (kchunk + 31) / 32 * 32;
    const dim3 grid((cout + kTile - 1) / kTile, (NW + 1 + kTile - 1) / kTile, splits);
    k_gemm<<<grid, kT, 0, s>>>(kchunk, P, ga, gb, epi);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_reduce_splits<<<ew_grid(static_cast<size_t>(cout) * (NW + 1)), kT, 0, s>>>(part, splits, cout, NW, dw, db);
    return hipGetLastError();
}

// x: [b][h][w][cin] -> yc [b][h][w][C] (head_conv).
hipError_t darcy_nn_head_fwd(const void* x, int b, int h, int w, int cin, int C, const void* whl, const float* bias,
                             float* yc, hipStream_t s) {
    return head_conv(x, Geo{b, h, w, cin}, C, whl, bias, yc, HeadOut{nullptr, nullptr, 0.0, nullptr}, s);
}

// hi / lo parts of the head's count FP32 weights: whl [2][count] bf16.
hipError_t darcy_nn_head_split(const float* w, int count, void* whl, hipStream_t s) {
    k_head_split<<<ew_grid(static_cast<size_t>(count)), kT, 0, s>>>(w, count, static_cast<bf16*>(whl));
    return hipGetLastError();
}

hipError_t darcy_nn_head_bwd_data(const float* dyc, int b, int h, int w, int cin, int C, const float* wgt, void* dx,
                                  hipStream_t s) {
    k_head_bwd_data<<<ew_grid(static_cast<size_t>(b) * h * w * cin), kT, 0, s>>>(dyc, Geo{b, h, w, cin}, C, wgt,
                                                                                  static_cast<bf16*>(dx));
    return hipGetLastError();
}

// part: groups * C * (9 cin + 1) floats with groups = ceil(b h w / per);
// summed in group order into dw [C][9 cin] and db [C].
hipError_t darcy_nn_head_bwd_weight(const void* x, int b, int h, int w, int cin, int C, const float* dyc, int per,
                                    float* part, float* dw, float* db, hipStream_t s) {
    const size_t P = static_cast<size_t>(b) * h * w;
    const int groups = static_cast<int>((P + per - 1) / per);
    k_head_dw_part<<<groups, kT, 0, s>>>(static_cast<const bf16*>(x), Geo{b, h, w, cin}, C, dyc, per, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_reduce_splits<<<ew_grid(static_cast<size_t>(C) * (9 * cin + 1)), kT, 0, s>>>(part, groups, C, 9 * cin, dw, db);
    return hipGetLastError();
}

// Half-resolution output map: y [b][n][n] from yc [b][n/2][n/2][5].
hipError_t darcy_nn_sub2_out(const float* yc, int n, int batch, float* y, hipStream_t s) {
    k_sub2_out<<<ew_grid(static_cast<size_t>(n / 2) * (n / 2) * batch), kT, 0, s>>>(yc, n, batch, y,
                                                                         HeadOut{nullptr, nullptr, 0.0, nullptr});
    return hipGetLastError();
}

// Its adjoint: dyc [b][n/2][n/2][5] from dy [b][n][n].
hipError_t darcy_nn_sub2_adjoint(const float* dy, int n, int batch, float* dyc, hipStream_t s) {
    k_sub2_adjoint<<<ew_grid(static_cast<size_t>(n / 2) * (n / 2) * batch), kT, 0, s>>>(dy, n, batch, dyc);
    return hipGetLastError();
}

hipError_t darcy_nn_position(int n, float* pos, hipStream_t s) {
    k_position<<<ew_grid(static_cast<size_t>(n) * n), kT, 0, s>>>(n, pos);
    return hipGetLastError();
}

// x: [b][n/sub][n/sub][8].
hipError_t darcy_nn_features(const double* const* a, const float* pos, int n, int sub, int batch, void* x,
                             hipStream_t s) {
    const size_t px = static_cast<size_t>(n / sub) * (n / sub) * batch;
    k_features<<<ew_grid(px), kT, 0, s>>>(a, pos, n, sub, batch, static_cast<bf16*>(x));
    return hipGetLastError();
}

// c: [b][h][w][ch] (coarse); skip, z, act: [b][2h][2w][ch]; z may be null.
hipError_t darcy_nn_up_add_gelu(const void* c, int b, int h, int w, int ch, const void* skip, void* z, void* act,
                                hipStream_t s) {
    k_up_add_gelu<<<ew_grid(static_cast<size_t>(b) * h * w * ch / 2), kT, 0, s>>>(
        static_cast<const bf16*>(c), Geo{b, h, w, ch}, static_cast<const bf16*>(skip), static_cast<bf16*>(z),
        static_cast<bf16*>(act));
    return hipGetLastError();
}

hipError_t darcy_nn_gelu_bwd(void* d, const void* z, size_t count, hipStream_t s) {
    k_gelu_bwd<<<ew_grid(count), kT, 0, s>>>(static_cast<bf16*>(d), static_cast<const bf16*>(z), count);
    return hipGetLastError();
}

// dc: [b][h][w][ch] (coarse); dz: [b][2h][2w][ch].
hipError_t darcy_nn_down_sum(const void* dz, int b, int h, int w, int ch, void* dc, hipStream_t s) {
    k_down_sum<<<ew_grid(static_cast<size_t>(b) * h * w * ch), kT, 0, s>>>(static_cast<const bf16*>(dz), Geo{b, h, w, ch},
                                                                            static_cast<bf16*>(dc));
    return hipGetLastError();
}

hipError_t darcy_nn_add(void* dst, const void* src, size_t count, hipStream_t s) {
    k_add<<<ew_grid(count), kT, 0, s>>>(static_cast<bf16*>(dst), static_cast<const bf16*>(src), count);
    return hipGetLastError();
}

hipError_t darcy_nn_transpose(const void* in, int P, int C, void* out, hipStream_t s) {
    const dim3 grid((P + 31) / 32, (C + 31) / 32);
    k_transpose<<<grid, kT, 0, s>>>(static_cast<const bf16*>(in), P, C, static_cast<bf16*>(out));
    return hipGetLastError();
}

// The head and the output in one: u0 = f scale pos y for n x n problems, x
// the finest network level (side n / sub).  sub 1: the head stores u0
// itself; sub 2: the 5-channel head writes yc, then the output map forms u0.
hipError_t darcy_nn_head_output(const void* x, int b, int n, int cin, int sub, const void* whl, const float* bias,
                                float* yc, const float* pos, const double* f, double scale, double* const* u0,
                                hipStream_t s) {
    if (sub != 1 && sub != 2) return hipErrorInvalidValue;
    if (sub == 2) {
        hipError_t e = darcy_nn_head_fwd(x, b, n / 2, n / 2, cin, 5, whl, bias, yc, s);
        if (e != hipSuccess) return e;
        k_sub2_out<<<ew_grid(static_cast<size_t>(n / 2) * (n / 2) * b), kT, 0, s>>>(yc, n, b, nullptr,
                                                                                    HeadOut{pos, f, scale, u0});
        return hipGetLastError();
    }
    return head_conv(x, Geo{b, n, n, cin}, 1, whl, bias, yc, HeadOut{pos, f, scale, u0}, s);
}

// partial: `blocks` floats.
hipError_t darcy_nn_loss(const float* y, const double* const* u, const double* const* a, const float* pos, int n,
                         int batch, double scale, float lam, float* dy, float* partial, int blocks, hipStream_t s) {
    k_loss<<<blocks, kT, 0, s>>>(y, u, a, pos, n, batch, scale, lam, dy, partial);
    return hipGetLastError();
}

hipError_t darcy_nn_adam(float* w, const float* g, float* m, float* v, void* wb, size_t count, float lr, float b1,
                         float b2, float eps, float c1, float c2, hipStream_t s) {
    k_adam<<<ew_grid(count), kT, 0, s>>>(w, g, m, v, static_cast<bf16*>(wb), count, lr, b1, b2, eps, c1, c2);
    return hipGetLastError();
}

hipError_t darcy_nn_to_bf16(const float* in, void* out, size_t count, hipStream_t s) {
    k_to_bf16<<<ew_grid(count), kT, 0, s>>>(in, static_cast<bf16*>(out), count);
    return hipGetLastError();
}

hipError_t darcy_nn_weight_t(const void* w, int cout, int cin, void* wt, hipStream_t s) {
    k_weight_t<<<ew_grid(static_cast<size_t>(cout) * 9 * cin), kT, 0, s>>>(static_cast<const bf16*>(w), cout, cin,
                                                                            static_cast<bf16*>(wt));
    return hipGetLastError();
}

}  // extern "C"
