// Host side of the learned warm start (include/darcy_hip.h, darcy_net_*):
// parameter layout, activation and gradient buffers, the forward and
// backward sequence of the encoder-decoder, Adam, and the single-layer
// entry points used to check the conv kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/darcy_hip.h"
#include "ctx.hpp"

using darcy_hip::DevBuf;

extern "C" {
hipError_t darcy_nn_mfma_probe(const void* A, const void* B, float* C, hipStream_t s);
hipError_t darcy_nn_conv_fwd(const void* x, int b, int h, int w, int cin, const void* wgt, const float* bias, int cout,
                             int stride, void* z, void* act, hipStream_t s);
hipError_t darcy_nn_conv_bwd_data(const void* dz, int b, int h, int w, int cin, const void* wt, int cout, int stride,
                                  void* dx, hipStream_t s);
hipError_t darcy_nn_conv_bwd_weight(const void* x, int b, int h, int w, int cin, const void* dzt, int cout, int stride,
                                    int splits, float* part, float* dw, float* db, hipStream_t s);
hipError_t darcy_nn_conv_fwd_merge(const void* x, int b, int h, int w, int cin, const void* wgt, const float* bias,
                                   int cout, const void* skip, void* z, void* act, hipStream_t s);
hipError_t darcy_nn_head_fwd(const void* x, int b, int h, int w, int cin, int C, const void* whl, const float* bias,
                             float* yc, hipStream_t s);
hipError_t darcy_nn_head_split(const float* w, int count, void* whl, hipStream_t s);
hipError_t darcy_nn_head_bwd_data(const float* dyc, int b, int h, int w, int cin, int C, const float* wgt, void* dx,
                                  hipStream_t s);
hipError_t darcy_nn_head_bwd_weight(const void* x, int b, int h, int w, int cin, int C, const float* dyc, int per,
                                    float* part, float* dw, float* db, hipStream_t s);
hipError_t darcy_nn_sub2_out(const float* yc, int n, int batch, float* y, hipStream_t s);
hipError_t darcy_nn_sub2_adjoint(const float* dy, int n, int batch, float* dyc, hipStream_t s);
hipError_t darcy_nn_position(int n, float* pos, hipStream_t s);
hipError_t darcy_nn_features(const double* const* a, const float* pos, int n, int sub, int batch, void* x,
                             hipStream_t s);
hipError_t darcy_nn_up_add_gelu(const void* c, int b, int h, int w, int ch, const void* skip, void* z, void* act,
                                hipStream_t s);
hipError_t darcy_nn_gelu_bwd(void* d, const void* z, size_t count, hipStream_t s);
hipError_t darcy_nn_down_sum(const void* dz, int b, int h, int w, int ch, void* dc, hipStream_t s);
hipError_t darcy_nn_add(void* dst, const void* src, size_t count, hipStream_t s);
hipError_t darcy_nn_transpose(const void* in, int P, int C, void* out, hipStream_t s);
hipError_t darcy_nn_head_output(const void* x, int b, int n, int cin, int sub, const void* whl, const float* bias,
                                float* yc, const float* pos, const double* f, double scale, double* const* u0,
                                hipStream_t s);
hipError_t darcy_nn_loss(const float* y, const double* const* u, const double* const* a, const float* pos, int n,
                         int batch, double scale, float lam, float* dy, float* partial, int blocks, hipStream_t s);
hipError_t darcy_nn_adam(float* w, const float* g, float* m, float* v, void* wb, size_t count, float lr, float b1,
                         float b2, float eps, float c1, float c2, hipStream_t s);
hipError_t darcy_nn_to_bf16(const float* in, void* out, size_t count, hipStream_t s);
hipError_t darcy_nn_weight_t(const void* w, int cout, int cin, void* wt, hipStream_t s);
}

namespace {

// Levels of the U-Net: level l has side m / 2^l and kChan[l] channels, m =
// n / sub the side of the finest network level (sub 1: the grid itself; sub
// 2: half resolution, each network pixel standing for a 2 x 2 block of
// cells, whose four ln a values are its input features; the head then has
// 5 channels: one interpolated bilinearly to the cells, four adding each
// cell's own detail).  The encoder conv D_l produces
// level l (stride 2 for l > 0, from level l - 1; D_0 reads the 8-channel
// features).  The decoder conv U_i maps level i + 1
// (kChan[i+1] channels) to kChan[i] channels at that size; its output is
// upsampled x2 (nearest), added to the encoder activation of level i and
// passed through GELU.  Levels 0..3 (32-64-128-256) are the encoder-decoder
// proper; levels 4..6 (256 channels, n/16 .. n/64) give every cell a
// receptive field covering the whole grid for 18% more FLOPs.
constexpr int kLevels = 7;
constexpr int kChan[kLevels] = {32, 64, 128, 256, 256, 256, 256};
constexpr int kInC = 8;
constexpr int kConvs = 2 * kLevels - 1;  // D_0 .. D_6, then U_5 .. U_0
constexpr int kHeadCin = 32;
constexpr double kScale = 1.0 / 64.0;  // u0 = f * pos * y * kScale
constexpr float kEnergyWeight = 1000.0f;  // lam of the energy term of the loss
constexpr int kLossBlocks = 1024;
constexpr int kHeadPer = 2048;  // pixels per head-gradient partial

struct ConvSpec {
    int cin, cout, stride, level_in;
};
constexpr int u_index(int i) { return 2 * kLevels - 2 - i; }  // conv index of U_i
constexpr ConvSpec spec(int c) {
    return c < kLevels ? ConvSpec{c ? kChan[c - 1] : kInC, kChan[c], c ? 2 : 1, c ? c - 1 : 0}
                       : ConvSpec{kChan[2 * kLevels - 1 - c], kChan[2 * kLevels - 2 - c], 1, 2 * kLevels - 1 - c};
}

// Output channels of the head: 1, or 5 at half resolution.
constexpr int head_channels(int sub) { return sub == 1 ? 1 : 5; }

// Parameter offsets.
struct Layout {
    size_t w[kConvs + 1], b[kConvs + 1], total;
    explicit Layout(int sub) {
        size_t o = 0;
        for (int c = 0; c <= kConvs; ++c) {
            const int cin = c < kConvs ? spec(c).cin : kHeadCin, cout = c < kConvs ? spec(c).cout : head_channels(sub);
            w[c] = o;
            o += static_cast<size_t>(cout) * 9 * cin;
            b[c] = o;
            o += cout;
        }
        total = o;
    }
};
const Layout kLayouts[2] = {Layout(1), Layout(2)};

// Sub-pixel factor of the creation flags (DARCY_NET_HALF_RES: 2, else 1).
int sub_of(int32_t flags) { return (flags & DARCY_NET_HALF_RES) ? 2 : 1; }

int fail(darcy_ctx* c, int code, const std::string& msg) {
    if (c) c->error = msg;
    return code;
}

int hip_fail(darcy_ctx* c, hipError_t e, const char* what) {
    const int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? DARCY_OUT_OF_MEMORY : DARCY_HIP_ERROR;
    return fail(c, code, std::string(what) + ": " + hipGetErrorString(e));
}

#define NN_TRY(expr, what)                                    \
    do {                                                      \
        const hipError_t e_ = (expr);                         \
        if (e_ != hipSuccess) return hip_fail(ctx, e_, what); \
    } while (0)

int splits_for(int cout, int cin, size_t P) {
    const int tiles = ((cout + 63) / 64) * ((9 * cin + 1 + 63) / 64);
    long s = 2048 / tiles;
    const long cap = static_cast<long>(P / 1024);
    if (s > cap) s = cap;
    return s < 1 ? 1 : static_cast<int>(s);
}

}  // namespace

struct darcy_net {
    darcy_ctx* ctx = nullptr;
    int n = 0, max_batch = 0, sub = 1;
    bool train = false;
    const Layout* lay = &kLayouts[0];
    long step = 0;
    DevBuf params, wb, wt, whl, grad, m, v, pos, feat, y, yc;
    DevBuf act[kLevels], z[kLevels];               // encoder outputs (after / before GELU)
    DevBuf zm[kLevels - 1], am[kLevels - 1];       // merges at level i (before / after GELU)
    DevBuf dA[kLevels], dM[kLevels - 1], dC[kLevels - 1];
    DevBuf dy, dyc, tmp, wpart, hpart, lossp, tab_a, tab_u, tab_f;

    int side(int level) const { return (n / sub) >> level; }
    size_t px(int level, int b) const { return static_cast<size_t>(b) * side(level) * side(level); }
    size_t cells(int b) const { return static_cast<size_t>(b) * n * n; }  // the output y
    int hc() const { return head_channels(sub); }
    // The head conv's output (and its gradient): y itself at full resolution.
    float* head_out() const { return (sub == 1 ? y : yc).as<float>(); }
    float* head_grad() const { return (sub == 1 ? dy : dyc).as<float>(); }
    const uint16_t* wbf(int c) const { return wb.as<uint16_t>() + lay->w[c]; }
    const uint16_t* wtr(int c) const { return wt.as<uint16_t>() + lay->w[c]; }
    const float* bias(int c) const { return params.as<float>() + lay->b[c]; }

    int alloc();
    int refresh_weights();
    int upload_tables(const void* const* a, const void* const* u, const double* f, int64_t count);
    int forward(int b, int64_t offset, bool to_u0 = false);
    int backward(int b, int64_t offset, double* loss);
    int wgrad(int c, const void* x, const void* dz, int b);
    void release();
};

int darcy_net::alloc() {
    const int B = max_batch;
    const size_t P = lay->total;
    auto bf = [](DevBuf& buf, size_t elems) { return buf.reserve(elems * 2); };
    NN_TRY(params.reserve(P * 4), "net parameters");
    NN_TRY(wb.reserve(P * 2), "net parameters");
    NN_TRY(wt.reserve(P * 2), "net parameters");
    NN_TRY(whl.reserve(static_cast<size_t>(2) * hc() * 9 * kHeadCin * 2), "net parameters");
    NN_TRY(hipMemsetAsync(params.ptr, 0, P * 4, ctx->stream), "net parameters");
    NN_TRY(pos.reserve(static_cast<size_t>(n) * n * 4), "net position");
    NN_TRY(darcy_nn_position(n, pos.as<float>(), ctx->stream), "net position");
    NN_TRY(bf(feat, px(0, B) * kInC), "net activations");
    NN_TRY(y.reserve(cells(B) * 4), "net activations");
    if (sub != 1) NN_TRY(yc.reserve(px(0, B) * hc() * 4), "net activations");
    for (int l = 0; l < kLevels; ++l) {
        NN_TRY(bf(act[l], px(l, B) * kChan[l]), "net activations");
        if (train) NN_TRY(bf(z[l], px(l, B) * kChan[l]), "net activations");
    }
    for (int i = 0; i + 1 < kLevels; ++i) {
        NN_TRY(bf(am[i], px(i, B) * kChan[i]), "net activations");
        if (train) NN_TRY(bf(zm[i], px(i, B) * kChan[i]), "net activations");
    }
    if (train) {
        NN_TRY(grad.reserve(P * 4), "net gradients");
        NN_TRY(m.reserve(P * 4), "optimizer state");
        NN_TRY(v.reserve(P * 4), "optimizer state");
        NN_TRY(hipMemsetAsync(m.ptr, 0, P * 4, ctx->stream), "optimizer state");
        NN_TRY(hipMemsetAsync(v.ptr, 0, P * 4, ctx->stream), "optimizer state");
        NN_TRY(dy.reserve(cells(B) * 4), "net gradients");
        if (sub != 1) NN_TRY(dyc.reserve(px(0, B) * hc() * 4), "net gradients");
        for (int l = 0; l < kLevels; ++l) NN_TRY(bf(dA[l], px(l, B) * kChan[l]), "net gradients");
        for (int i = 0; i + 1 < kLevels; ++i) {
            NN_TRY(bf(dM[i], px(i, B) * kChan[i]), "net gradients");
            NN_TRY(bf(dC[i], px(i + 1, B) * kChan[i]), "net gradients");
        }
        size_t tmax = 0, part = 0;
        for (int c = 0; c < kConvs; ++c) {
            const ConvSpec s = spec(c);
            const size_t out_px = px(s.level_in, B) / (s.stride * s.stride);
            tmax = std::max(tmax, out_px * s.cout);
            part = std::max(part, static_cast<size_t>(splits_for(s.cout, s.cin, out_px)) * s.cout * (9 * s.cin + 1));
        }
        NN_TRY(bf(tmp, tmax), "net gradients");
        NN_TRY(wpart.reserve(part * 4), "net gradients");
        NN_TRY(hpart.reserve(((px(0, B) + kHeadPer - 1) / kHeadPer) * hc() * (9 * kHeadCin + 1) * 4),
               "net gradients");
        NN_TRY(lossp.reserve(kLossBlocks * 4), "net gradients");
    }
    return refresh_weights();
}

void darcy_net::release() {
    DevBuf* all[] = {&params, &wb, &wt, &whl, &grad, &m, &v, &pos, &feat, &y, &yc, &dy, &dyc, &tmp, &wpart, &hpart,
                     &lossp, &tab_a, &tab_u, &tab_f};
    for (DevBuf* b : all) b->release();
    for (int l = 0; l < kLevels; ++l) {
        act[l].release();
        z[l].release();
        dA[l].release();
    }
    for (int i = 0; i + 1 < kLevels; ++i) {
        zm[i].release();
        am[i].release();
        dM[i].release();
        dC[i].release();
    }
}

// bf16 copies of the weights, the flipped transposes used by backward data
// and the head's hi / lo split.
int darcy_net::refresh_weights() {
    NN_TRY(darcy_nn_to_bf16(params.as<float>(), wb.ptr, lay->total, ctx->stream), "net weights");
    NN_TRY(darcy_nn_head_split(params.as<float>() + lay->w[kConvs], hc() * 9 * kHeadCin, whl.ptr, ctx->stream),
           "net weights");
    for (int c = 1; c < kConvs; ++c)
        NN_TRY(darcy_nn_weight_t(wb.as<uint16_t>() + lay->w[c], spec(c).cout, spec(c).cin,
                                 wt.as<uint16_t>() + lay->w[c], ctx->stream),
               "net weights");
    return DARCY_OK;
}

int darcy_net::upload_tables(const void* const* a, const void* const* u, const double* f, int64_t count) {
    // A null entry would reach the kernels as an address: rejected here.
    for (int64_t i = 0; i < count; ++i)
        if (!a[i] || (u && !u[i])) return fail(ctx, DARCY_INVALID_INPUT, "null pointer in the a or u table");
    const size_t bytes = static_cast<size_t>(count) * sizeof(void*);
    NN_TRY(tab_a.reserve(bytes), "net tables");
    NN_TRY(hipMemcpyAsync(tab_a.ptr, a, bytes, hipMemcpyHostToDevice, ctx->stream), "net tables");
    if (u) {
        NN_TRY(tab_u.reserve(bytes), "net tables");
        NN_TRY(hipMemcpyAsync(tab_u.ptr, u, bytes, hipMemcpyHostToDevice, ctx->stream), "net tables");
    }
    if (f) {
        NN_TRY(tab_f.reserve(count * sizeof(double)), "net tables");
        NN_TRY(hipMemcpyAsync(tab_f.ptr, f, count * sizeof(double), hipMemcpyHostToDevice, ctx->stream), "net tables");
    }
    // The host tables may be released when the call returns.
    NN_TRY(hipStreamSynchronize(ctx->stream), "net tables");
    return DARCY_OK;
}

// Forward pass of problems [offset, offset + b) of the uploaded tables: the
// head's output y, or with `to_u0` the warm starts u0 themselves.
int darcy_net::forward(int b, int64_t offset, bool to_u0) {
    hipStream_t s = ctx->stream;
    NN_TRY(darcy_nn_features(tab_a.as<const double*>() + offset, pos.as<float>(), n, sub, b, feat.ptr, s),
           "net features");
    const void* x = feat.ptr;
    for (int l = 0; l < kLevels; ++l) {
        const ConvSpec c = spec(l);
        const int h = side(c.level_in);
        NN_TRY(darcy_nn_conv_fwd(x, b, h, h, c.cin, wbf(l), bias(l), c.cout, c.stride, train ? z[l].ptr : nullptr,
                                 act[l].ptr, s),
               "net encoder");
        x = act[l].ptr;
    }
    for (int i = kLevels - 2; i >= 0; --i) {
        const int c = u_index(i), h = side(i + 1);
        NN_TRY(darcy_nn_conv_fwd_merge(x, b, h, h, kChan[i + 1], wbf(c), bias(c), kChan[i], act[i].ptr,
                                       train ? zm[i].ptr : nullptr, am[i].ptr, s),
               "net decoder");
        x = am[i].ptr;
    }
    if (to_u0) {
        NN_TRY(darcy_nn_head_output(x, b, n, kHeadCin, sub, whl.ptr, params.as<float>() + lay->b[kConvs],
                                    head_out(), pos.as<float>(),
                                    tab_f.as<double>() + offset, kScale, tab_u.as<double*>() + offset, s),
               "net head");
    } else {
        NN_TRY(darcy_nn_head_fwd(x, b, side(0), side(0), kHeadCin, hc(), whl.ptr, params.as<float>() + lay->b[kConvs],
                                 head_out(), s),
               "net head");
        if (sub != 1) NN_TRY(darcy_nn_sub2_out(yc.as<float>(), n, b, y.as<float>(), s), "net head");
    }
    return DARCY_OK;
}

// Weight and bias gradient of conv c from its input x and output gradient dz.
int darcy_net::wgrad(int c, const void* x, const void* dz, int b) {
    const ConvSpec sp = spec(c);
    const int h = side(sp.level_in), oh = h / sp.stride;
    const size_t P = static_cast<size_t>(b) * oh * oh;
    NN_TRY(darcy_nn_transpose(dz, static_cast<int>(P), sp.cout, tmp.ptr, ctx->stream), "net backward");
    const int splits = splits_for(sp.cout, sp.cin, P);
    NN_TRY(darcy_nn_conv_bwd_weight(x, b, h, h, sp.cin, tmp.ptr, sp.cout, sp.stride, splits, wpart.as<float>(),
                                    grad.as<float>() + lay->w[c], grad.as<float>() + lay->b[c], ctx->stream),
           "net backward");
    return DARCY_OK;
}

// Loss and gradient for the batch whose forward pass just ran.
int darcy_net::backward(int b, int64_t offset, double* loss) {
    hipStream_t s = ctx->stream;
    NN_TRY(darcy_nn_loss(y.as<float>(), tab_u.as<const double*>() + offset, tab_a.as<const double*>() + offset,
                         pos.as<float>(), n, b, kScale, kEnergyWeight, dy.as<float>(), lossp.as<float>(), kLossBlocks,
                         s),
           "net loss");
    const float* wh = params.as<float>() + lay->w[kConvs];
    const int m = side(0);
    if (sub != 1) NN_TRY(darcy_nn_sub2_adjoint(dy.as<float>(), n, b, dyc.as<float>(), s), "net backward");
    NN_TRY(darcy_nn_head_bwd_data(head_grad(), b, m, m, kHeadCin, hc(), wh, dM[0].ptr, s), "net backward");
    NN_TRY(darcy_nn_head_bwd_weight(am[0].ptr, b, m, m, kHeadCin, hc(), head_grad(), kHeadPer, hpart.as<float>(),
                                    grad.as<float>() + lay->w[kConvs], grad.as<float>() + lay->b[kConvs], s),
           "net backward");
    // Decoder, finest merge first: dM[i] holds the gradient of merge i's
    // activation; after the GELU step it is the gradient of the merge's sum,
    // which is also the skip gradient of the encoder activation act[i].
    for (int i = 0; i + 1 < kLevels; ++i) {
        const int h = side(i + 1), ch = kChan[i], c = u_index(i);
        NN_TRY(darcy_nn_gelu_bwd(dM[i].ptr, zm[i].ptr, px(i, b) * ch, s), "net backward");
        NN_TRY(darcy_nn_down_sum(dM[i].ptr, b, h, h, ch, dC[i].ptr, s), "net backward");
        const bool last = i + 2 == kLevels;
        if (int r = wgrad(c, last ? act[kLevels - 1].ptr : am[i + 1].ptr, dC[i].ptr, b)) return r;
        NN_TRY(darcy_nn_conv_bwd_data(dC[i].ptr, b, h, h, kChan[i + 1], wtr(c), ch, 1,
                                      last ? dA[kLevels - 1].ptr : dM[i + 1].ptr, s),
               "net backward");
    }
    // Encoder, deepest first.
    NN_TRY(darcy_nn_gelu_bwd(dA[kLevels - 1].ptr, z[kLevels - 1].ptr, px(kLevels - 1, b) * kChan[kLevels - 1], s),
           "net backward");
    for (int l = kLevels - 1; l >= 0; --l) {
        if (int r = wgrad(l, l ? act[l - 1].ptr : feat.ptr, dA[l].ptr, b)) return r;
        if (l == 0) break;
        const int h = side(l - 1);
        const size_t cnt = px(l - 1, b) * kChan[l - 1];
        NN_TRY(darcy_nn_conv_bwd_data(dA[l].ptr, b, h, h, kChan[l - 1], wtr(l), kChan[l], 2, dA[l - 1].ptr, s),
               "net backward");
        NN_TRY(darcy_nn_add(dA[l - 1].ptr, dM[l - 1].ptr, cnt, s), "net backward");
        NN_TRY(darcy_nn_gelu_bwd(dA[l - 1].ptr, z[l - 1].ptr, cnt, s), "net backward");
    }
    if (loss) {
        std::vector<float> part(kLossBlocks);
        NN_TRY(hipMemcpyAsync(part.data(), lossp.ptr, kLossBlocks * 4, hipMemcpyDeviceToHost, s), "net loss");
        NN_TRY(hipStreamSynchronize(s), "net loss");
        double sum = 0.0;
        for (float p : part) sum += p;
        *loss = sum;
    }
    return DARCY_OK;
}

// ------------------------------------------------------------ C ABI

namespace {

int check_net(darcy_net* net) { return (!net || !net->ctx) ? DARCY_INVALID_INPUT : DARCY_OK; }

hipError_t set_device(darcy_ctx* ctx) { return hipSetDevice(ctx->device); }

}  // namespace

extern "C" int64_t darcy_net_param_count(int32_t flags) {
    return static_cast<int64_t>(kLayouts[sub_of(flags) - 1].total);
}

extern "C" void darcy_net_destroy(darcy_net* net) {
    if (!net) return;
    if (net->ctx) {
        std::lock_guard<std::mutex> lock(net->ctx->mu);
        (void)hipSetDevice(net->ctx->device);
        (void)hipStreamSynchronize(net->ctx->stream);
        net->release();
    }
    delete net;
}

extern "C" int darcy_net_create(darcy_ctx* ctx, int64_t n, int64_t max_batch, int32_t flags, darcy_net** out) {
    if (!ctx || !out) return DARCY_INVALID_INPUT;
    *out = nullptr;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (flags & ~(DARCY_NET_TRAIN | DARCY_NET_HALF_RES)) return fail(ctx, DARCY_INVALID_INPUT, "unknown network flags");
    const int sub = sub_of(flags);
    const int64_t unit = int64_t(sub) << (kLevels - 1);
    if (n < unit || n % unit != 0 || n > 4096)
        return fail(ctx, DARCY_INVALID_INPUT,
                    sub == 1 ? "the warm-start network needs n a multiple of 64 in [64, 4096]"
                             : "the half-resolution warm-start network needs n a multiple of 128 in [128, 4096]");
    if (max_batch < 1 || static_cast<double>(max_batch) * n * n > 1.34e8)
        return fail(ctx, DARCY_INVALID_INPUT, "max_batch must be at least 1 and max_batch * n * n at most 1.34e8");
    NN_TRY(set_device(ctx), "hipSetDevice");
    darcy_net* net = new darcy_net;
    net->ctx = ctx;
    net->n = static_cast<int>(n);
    net->max_batch = static_cast<int>(max_batch);
    net->train = (flags & DARCY_NET_TRAIN) != 0;
    net->sub = sub;
    net->lay = &kLayouts[sub - 1];
    const int rc = net->alloc();
    if (rc != DARCY_OK) {
        (void)hipStreamSynchronize(ctx->stream);
        net->release();
        delete net;
        return rc;
    }
    *out = net;
    return DARCY_OK;
}

extern "C" int darcy_net_set_params(darcy_net* net, const float* params, int64_t count) {
    if (check_net(net) || !params) return DARCY_INVALID_INPUT;
    darcy_ctx* ctx = net->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (count != static_cast<int64_t>(net->lay->total))
        return fail(ctx, DARCY_INVALID_INPUT, "parameter vector has the wrong length");
    NN_TRY(set_device(ctx), "hipSetDevice");
    NN_TRY(hipMemcpyAsync(net->params.ptr, params, net->lay->total * 4, hipMemcpyHostToDevice, ctx->stream),
           "net parameters");
    if (net->train) {
        NN_TRY(hipMemsetAsync(net->m.ptr, 0, net->lay->total * 4, ctx->stream), "optimizer state");
        NN_TRY(hipMemsetAsync(net->v.ptr, 0, net->lay->total * 4, ctx->stream), "optimizer state");
        net->step = 0;
    }
    if (int r = net->refresh_weights()) return r;
    NN_TRY(hipStreamSynchronize(ctx->stream), "net parameters");
    return DARCY_OK;
}

extern "C" int darcy_net_get_params(darcy_net* net, float* params, int64_t count) {
    if (check_net(net) || !params) return DARCY_INVALID_INPUT;
    darcy_ctx* ctx = net->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (count != static_cast<int64_t>(net->lay->total))
        return fail(ctx, DARCY_INVALID_INPUT, "parameter vector has the wrong length");
    NN_TRY(set_device(ctx), "hipSetDevice");
    NN_TRY(hipMemcpyAsync(params, net->params.ptr, net->lay->total * 4, hipMemcpyDeviceToHost, ctx->stream),
           "net parameters");
    NN_TRY(hipStreamSynchronize(ctx->stream), "net parameters");
    return DARCY_OK;
}

extern "C" int darcy_net_predict(darcy_net* net, int64_t count, const double* const* a, const double* f,
                                 double* const* u0) {
    if (check_net(net)) return DARCY_INVALID_INPUT;
    darcy_ctx* ctx = net->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (count < 0 || (count > 0 && (!a || !f || !u0)))
        return fail(ctx, DARCY_INVALID_INPUT, "predict needs a, f and u0 tables");
    if (count == 0) return DARCY_OK;
    NN_TRY(set_device(ctx), "hipSetDevice");
    if (int r = net->upload_tables(reinterpret_cast<const void* const*>(a), reinterpret_cast<const void* const*>(u0),
                                   f, count))
        return r;
    NN_TRY(hipEventRecord(ctx->ev0, ctx->stream), "hipEventRecord");
    for (int64_t off = 0; off < count; off += net->max_batch) {
        const int b = static_cast<int>(count - off < net->max_batch ? count - off : net->max_batch);
        if (int r = net->forward(b, off, true)) return r;
    }
    NN_TRY(hipEventRecord(ctx->ev1, ctx->stream), "hipEventRecord");
    NN_TRY(hipEventSynchronize(ctx->ev1), "net predict");
    float ms = 0.0f;
    NN_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1), "hipEventElapsedTime");
    ctx->last_ms = ms;
    return DARCY_OK;
}

namespace {

int train_or_grad(darcy_net* net, int64_t batch, const double* const* a, const double* const* u, double* loss,
                  float* grad_out, const float* lr) {
    if (check_net(net)) return DARCY_INVALID_INPUT;
    darcy_ctx* ctx = net->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (!net->train) return fail(ctx, DARCY_INVALID_INPUT, "the network was created without training buffers");
    if (batch < 1 || batch > net->max_batch || !a || !u)
        return fail(ctx, DARCY_INVALID_INPUT, "batch must be in [1, max_batch] with a and u tables");
    NN_TRY(set_device(ctx), "hipSetDevice");
    if (int r = net->upload_tables(reinterpret_cast<const void* const*>(a), reinterpret_cast<const void* const*>(u),
                                   nullptr, batch))
        return r;
    const int b = static_cast<int>(batch);
    if (int r = net->forward(b, 0)) return r;
    if (int r = net->backward(b, 0, loss)) return r;
    if (grad_out) {
        NN_TRY(hipMemcpyAsync(grad_out, net->grad.ptr, net->lay->total * 4, hipMemcpyDeviceToHost, ctx->stream),
               "net gradient");
        NN_TRY(hipStreamSynchronize(ctx->stream), "net gradient");
    }
    if (lr) {
        const float b1 = 0.9f, b2 = 0.999f;
        net->step += 1;
        const float c1 = 1.0f - std::pow(b1, static_cast<float>(net->step));
        const float c2 = 1.0f - std::pow(b2, static_cast<float>(net->step));
        NN_TRY(darcy_nn_adam(net->params.as<float>(), net->grad.as<float>(), net->m.as<float>(), net->v.as<float>(),
                             net->wb.ptr, net->lay->total, *lr, b1, b2, 1e-8f, c1, c2, ctx->stream),
               "net update");
        for (int c = 1; c < kConvs; ++c)
            NN_TRY(darcy_nn_weight_t(net->wb.as<uint16_t>() + net->lay->w[c], spec(c).cout, spec(c).cin,
                                     net->wt.as<uint16_t>() + net->lay->w[c], ctx->stream),
                   "net update");
        NN_TRY(darcy_nn_head_split(net->params.as<float>() + net->lay->w[kConvs], net->hc() * 9 * kHeadCin,
                                   net->whl.ptr, ctx->stream),
               "net update");
    }
    return DARCY_OK;
}

}  // namespace

extern "C" int darcy_net_train_step(darcy_net* net, int64_t batch, const double* const* a, const double* const* u,
                                    float lr, double* loss) {
    return train_or_grad(net, batch, a, u, loss, nullptr, &lr);
}

extern "C" int darcy_net_gradient(darcy_net* net, int64_t batch, const double* const* a, const double* const* u,
                                  double* loss, float* grad) {
    if (!grad) return DARCY_INVALID_INPUT;
    return train_or_grad(net, batch, a, u, loss, grad, nullptr);
}

// ------------------------------------------------ single-layer entry points

namespace {

struct Scratch {
    std::vector<void*> ptrs;
    ~Scratch() {
        for (void* p : ptrs) (void)hipFree(p);
    }
    hipError_t get(size_t bytes, void** out) {
        const hipError_t e = hipMalloc(out, bytes ? bytes : 1);
        if (e == hipSuccess) ptrs.push_back(*out);
        return e;
    }
};

int check_layer(darcy_ctx* ctx, int b, int h, int w, int cin, int cout, int stride) {
    if (b < 1 || h < 1 || w < 1 || cin < 8 || cout < 8 || cin % 8 || cout % 8 || (stride != 1 && stride != 2) ||
        h % stride || w % stride || (static_cast<int64_t>(b) * (h / stride) * (w / stride)) % 8)
        return fail(ctx, DARCY_INVALID_INPUT,
                    "conv shape: stride 1 or 2, channels multiples of 8, h and w multiples of the stride, "
                    "b * (h/s) * (w/s) a multiple of 8");
    if (static_cast<double>(b) * h * w * (cin > cout ? cin : cout) > 1.0e9)
        return fail(ctx, DARCY_INVALID_INPUT, "conv tensors too large for the single-layer entry points");
    return DARCY_OK;
}

}  // namespace

extern "C" int darcy_nn_conv_forward(darcy_ctx* ctx, int32_t b, int32_t h, int32_t w, int32_t cin, int32_t cout,
                                     int32_t stride, const uint16_t* x, const uint16_t* wgt, const float* bias,
                                     uint16_t* z, uint16_t* act) {
    if (!ctx) return DARCY_INVALID_INPUT;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (int r = check_layer(ctx, b, h, w, cin, cout, stride)) return r;
    if (!x || !wgt || !bias || !z) return fail(ctx, DARCY_INVALID_INPUT, "conv forward needs x, wgt, bias and z");
    NN_TRY(set_device(ctx), "hipSetDevice");
    const size_t nx = static_cast<size_t>(b) * h * w * cin, nw = static_cast<size_t>(cout) * 9 * cin;
    const size_t nz = static_cast<size_t>(b) * (h / stride) * (w / stride) * cout;
    Scratch sc;
    void *dx, *dw, *db, *dz, *da = nullptr;
    NN_TRY(sc.get(nx * 2, &dx), "hipMalloc");
    NN_TRY(sc.get(nw * 2, &dw), "hipMalloc");
    NN_TRY(sc.get(cout * 4, &db), "hipMalloc");
    NN_TRY(sc.get(nz * 2, &dz), "hipMalloc");
    if (act) NN_TRY(sc.get(nz * 2, &da), "hipMalloc");
    hipStream_t s = ctx->stream;
    NN_TRY(hipMemcpyAsync(dx, x, nx * 2, hipMemcpyHostToDevice, s), "upload");
    NN_TRY(hipMemcpyAsync(dw, wgt, nw * 2, hipMemcpyHostToDevice, s), "upload");
    NN_TRY(hipMemcpyAsync(db, bias, cout * 4, hipMemcpyHostToDevice, s), "upload");
    NN_TRY(darcy_nn_conv_fwd(dx, b, h, w, cin, dw, static_cast<float*>(db), cout, stride, dz, da, s), "conv forward");
    NN_TRY(hipMemcpyAsync(z, dz, nz * 2, hipMemcpyDeviceToHost, s), "download");
    if (act) NN_TRY(hipMemcpyAsync(act, da, nz * 2, hipMemcpyDeviceToHost, s), "download");
    NN_TRY(hipStreamSynchronize(s), "conv forward");
    return DARCY_OK;
}

extern "C" int darcy_nn_conv_backward_data(darcy_ctx* ctx, int32_t b, int32_t h, int32_t w, int32_t cin, int32_t cout,
                                           int32_t stride, const uint16_t* dz, const uint16_t* wgt, uint16_t* dx) {
    if (!ctx) return DARCY_INVALID_INPUT;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (int r = check_layer(ctx, b, h, w, cin, cout, stride)) return r;
    if (!dz || !wgt || !dx) return fail(ctx, DARCY_INVALID_INPUT, "conv backward data needs dz, wgt and dx");
    NN_TRY(set_device(ctx), "hipSetDevice");
    const size_t nx = static_cast<size_t>(b) * h * w * cin, nw = static_cast<size_t>(cout) * 9 * cin;
    const size_t nz = static_cast<size_t>(b) * (h / stride) * (w / stride) * cout;
    Scratch sc;
    void *ddz, *dw, *dwt, *ddx;
    NN_TRY(sc.get(nz * 2, &ddz), "hipMalloc");
    NN_TRY(sc.get(nw * 2, &dw), "hipMalloc");
    NN_TRY(sc.get(nw * 2, &dwt), "hipMalloc");
    NN_TRY(sc.get(nx * 2, &ddx), "hipMalloc");
    hipStream_t s = ctx->stream;
    NN_TRY(hipMemcpyAsync(ddz, dz, nz * 2, hipMemcpyHostToDevice, s), "upload");
    NN_TRY(hipMemcpyAsync(dw, wgt, nw * 2, hipMemcpyHostToDevice, s), "upload");
    NN_TRY(darcy_nn_weight_t(dw, cout, cin, dwt, s), "weight transpose");
    NN_TRY(darcy_nn_conv_bwd_data(ddz, b, h, w, cin, dwt, cout, stride, ddx, s), "conv backward data");
    NN_TRY(hipMemcpyAsync(dx, ddx, nx * 2, hipMemcpyDeviceToHost, s), "download");
    NN_TRY(hipStreamSynchronize(s), "conv backward data");
    return DARCY_OK;
}

extern "C" int darcy_nn_conv_backward_weight(darcy_ctx* ctx, int32_t b, int32_t h, int32_t w, int32_t cin,
                                             int32_t cout, int32_t stride, int32_t splits, const uint16_t* x,
                                             const uint16_t* dz, float* dw, float* db) {
    if (!ctx) return DARCY_INVALID_INPUT;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (int r = check_layer(ctx, b, h, w, cin, cout, stride)) return r;
    if (!x || !dz || !dw || !db) return fail(ctx, DARCY_INVALID_INPUT, "conv backward weight needs x, dz, dw and db");
    const size_t P = static_cast<size_t>(b) * (h / stride) * (w / stride);
    if (splits < 1 || static_cast<size_t>(splits) > P)
        return fail(ctx, DARCY_INVALID_INPUT, "splits must be in [1, pixels]");
    NN_TRY(set_device(ctx), "hipSetDevice");
    const size_t nx = static_cast<size_t>(b) * h * w * cin, nw = static_cast<size_t>(cout) * 9 * cin;
    const size_t nz = P * cout;
    Scratch sc;
    void *dx, *ddz, *dzt, *part, *ddw, *ddb;
    NN_TRY(sc.get(nx * 2, &dx), "hipMalloc");
    NN_TRY(sc.get(nz * 2, &ddz), "hipMalloc");
    NN_TRY(sc.get(nz * 2, &dzt), "hipMalloc");
    NN_TRY(sc.get(static_cast<size_t>(splits) * cout * (9 * cin + 1) * 4, &part), "hipMalloc");
    NN_TRY(sc.get(nw * 4, &ddw), "hipMalloc");
    NN_TRY(sc.get(cout * 4, &ddb), "hipMalloc");
    hipStream_t s = ctx->stream;
    NN_TRY(hipMemcpyAsync(dx, x, nx * 2, hipMemcpyHostToDevice, s), "upload");
    NN_TRY(hipMemcpyAsync(ddz, dz, nz * 2, hipMemcpyHostToDevice, s), "upload");
    NN_TRY(darcy_nn_transpose(ddz, static_cast<int>(P), cout, dzt, s), "transpose");
    NN_TRY(darcy_nn_conv_bwd_weight(dx, b, h, w, cin, dzt, cout, stride, splits, static_cast<float*>(part),
                                    static_cast<float*>(ddw), static_cast<float*>(ddb), s),
           "conv backward weight");
    NN_TRY(hipMemcpyAsync(dw, ddw, nw * 4, hipMemcpyDeviceToHost, s), "download");
    NN_TRY(hipMemcpyAsync(db, ddb, cout * 4, hipMemcpyDeviceToHost, s), "download");
    NN_TRY(hipStreamSynchronize(s), "conv backward weight");
    return DARCY_OK;
}

extern "C" int darcy_nn_mfma_check(darcy_ctx* ctx, const uint16_t* a, const uint16_t* b, float* c) {
    if (!ctx || !a || !b || !c) return DARCY_INVALID_INPUT;
    std::lock_guard<std::mutex> lock(ctx->mu);
    NN_TRY(set_device(ctx), "hipSetDevice");
    Scratch sc;
    void *da, *db, *dc;
    NN_TRY(sc.get(16 * 32 * 2, &da), "hipMalloc");
    NN_TRY(sc.get(32 * 16 * 2, &db), "hipMalloc");
    NN_TRY(sc.get(16 * 16 * 4, &dc), "hipMalloc");
    hipStream_t s = ctx->stream;
    NN_TRY(hipMemcpyAsync(da, a, 16 * 32 * 2, hipMemcpyHostToDevice, s), "upload");
    NN_TRY(hipMemcpyAsync(db, b, 32 * 16 * 2, hipMemcpyHostToDevice, s), "upload");
    NN_TRY(darcy_nn_mfma_probe(da, db, static_cast<float*>(dc), s), "mfma probe");
    NN_TRY(hipMemcpyAsync(c, dc, 16 * 16 * 4, hipMemcpyDeviceToHost, s), "download");
    NN_TRY(hipStreamSynchronize(s), "mfma probe");
    return DARCY_OK;
}
