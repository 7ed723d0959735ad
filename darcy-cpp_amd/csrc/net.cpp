// Host side of the learned warm start (include/darcy_hip.h, darcy_net_*):
// parameter layout, activation and gradient buffers, the forward and
// backward sequence of the encoder-decoder, Adam, and the single-layer
// entry points used to check the conv kernels.
#include <hip/hip_runtime.h>

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
hipError_t darcy_nn_head_fwd(const void* x, int b, int h, int w, int cin, const float* wgt, const float* bias, float* y,
                             hipStream_t s);
hipError_t darcy_nn_head_bwd_data(const float* dy, int b, int h, int w, int cin, const float* wgt, void* dx,
                                  hipStream_t s);
hipError_t darcy_nn_head_bwd_weight(const void* x, int b, int h, int w, int cin, const float* dy, int per, float* part,
                                    float* dw, float* db, hipStream_t s);
hipError_t darcy_nn_features(const double* const* a, int n, int batch, void* x, hipStream_t s);
hipError_t darcy_nn_up_add_gelu(const void* c, int b, int h, int w, int ch, const void* skip, void* z, void* act,
                                hipStream_t s);
hipError_t darcy_nn_gelu_bwd(void* d, const void* z, size_t count, hipStream_t s);
hipError_t darcy_nn_down_sum(const void* dz, int b, int h, int w, int ch, void* dc, hipStream_t s);
hipError_t darcy_nn_add(void* dst, const void* src, size_t count, hipStream_t s);
hipError_t darcy_nn_transpose(const void* in, int P, int C, void* out, hipStream_t s);
hipError_t darcy_nn_output(const float* y, int n, int batch, const double* f, double scale, double* const* u0,
                           hipStream_t s);
hipError_t darcy_nn_mse(const float* y, const double* const* u, int n, int batch, double scale, float* dy,
                        float* partial, int blocks, hipStream_t s);
hipError_t darcy_nn_adam(float* w, const float* g, float* m, float* v, void* wb, size_t count, float lr, float b1,
                         float b2, float eps, float c1, float c2, hipStream_t s);
hipError_t darcy_nn_to_bf16(const float* in, void* out, size_t count, hipStream_t s);
hipError_t darcy_nn_weight_t(const void* w, int cout, int cin, void* wt, hipStream_t s);
}

namespace {

// The convolutions in parameter order: input channels, output channels,
// stride, and the side of the input as a divisor of n.
struct ConvSpec {
    int cin, cout, stride, div;
};
constexpr int kConvs = 7;
constexpr ConvSpec kSpec[kConvs] = {
    {8, 32, 1, 1},    {32, 64, 2, 1},  {64, 128, 2, 2}, {128, 256, 2, 4},  // encoder
    {256, 128, 1, 8}, {128, 64, 1, 4}, {64, 32, 1, 2},                     // decoder
};
constexpr int kHeadCin = 32;
constexpr double kScale = 1.0 / 64.0;  // u0 = f * pos * y * kScale
constexpr int kMseBlocks = 1024;
constexpr int kHeadPer = 2048;  // pixels per head-gradient partial

struct Layout {
    size_t w[kConvs + 1], b[kConvs + 1], total;
    Layout() {
        size_t o = 0;
        for (int l = 0; l <= kConvs; ++l) {
            const int cin = l < kConvs ? kSpec[l].cin : kHeadCin, cout = l < kConvs ? kSpec[l].cout : 1;
            w[l] = o;
            o += static_cast<size_t>(cout) * 9 * cin;
            b[l] = o;
            o += cout;
        }
        total = o;
    }
};
const Layout kLayout;

int fail(darcy_ctx* c, int code, const std::string& msg) {
    if (c) c->error = msg;
    return code;
}

int hip_fail(darcy_ctx* c, hipError_t e, const char* what) {
    const int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? DARCY_OUT_OF_MEMORY : DARCY_HIP_ERROR;
    return fail(c, code, std::string(what) + ": " + hipGetErrorString(e));
}

#define NN_TRY(expr, what)                                         \
    do {                                                           \
        const hipError_t e_ = (expr);                              \
        if (e_ != hipSuccess) return hip_fail(ctx, e_, what);      \
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
    int n = 0, max_batch = 0;
    bool train = false;
    long step = 0;
    DevBuf params, wb, wt, grad, m, v;
    DevBuf feat, y;
    DevBuf z[kConvs], act[kConvs];   // conv outputs (decoder: the conv result c)
    DevBuf zm[3], am[3];             // decoder merges: up2(c) + skip, and its GELU
    DevBuf dy, d[7], dc[3], tmp, wpart, hpart, mse, tab_a, tab_u, tab_f;

    int side(int l) const { return n / kSpec[l].div; }
    size_t px(int div, int b) const { return static_cast<size_t>(b) * (n / div) * (n / div); }
    const uint16_t* wbf(int l) const { return wb.as<uint16_t>() + kLayout.w[l]; }
    const uint16_t* wtr(int l) const { return wt.as<uint16_t>() + kLayout.w[l]; }
    const float* bias(int l) const { return params.as<float>() + kLayout.b[l]; }

    int alloc();
    int refresh_weights();
    int upload_tables(const void* const* a, const void* const* u, const double* f, int64_t count);
    int forward(int b, int64_t offset);
    int backward(int b, int64_t offset, double* loss);
    int wgrad(int l, const void* x, const void* dz, int b);
};

int darcy_net::alloc() {
    const size_t B = static_cast<size_t>(max_batch);
    const size_t P = kLayout.total;
    auto bf = [&](DevBuf& buf, size_t elems) { return buf.reserve(elems * 2); };
    NN_TRY(params.reserve(P * 4), "net parameters");
    NN_TRY(wb.reserve(P * 2), "net parameters");
    NN_TRY(wt.reserve(P * 2), "net parameters");
    NN_TRY(hipMemsetAsync(params.ptr, 0, P * 4, ctx->stream), "net parameters");
    NN_TRY(bf(feat, px(1, B) * 8), "net activations");
    NN_TRY(y.reserve(px(1, B) * 4), "net activations");
    // Encoder: outputs at n, n/2, n/4, n/8.  Decoder convs at n/8, n/4, n/2.
    const int out_div[kConvs] = {1, 2, 4, 8, 8, 4, 2};
    for (int l = 0; l < kConvs; ++l) {
        const size_t e = px(out_div[l], B) * kSpec[l].cout;
        if (l < 4) {
            NN_TRY(bf(act[l], e), "net activations");
            if (train) NN_TRY(bf(z[l], e), "net activations");
        } else {
            NN_TRY(bf(z[l], e), "net activations");
        }
    }
    const int merge_div[3] = {4, 2, 1}, merge_c[3] = {128, 64, 32};
    for (int k = 0; k < 3; ++k) {
        const size_t e = px(merge_div[k], B) * merge_c[k];
        NN_TRY(bf(am[k], e), "net activations");
        if (train) NN_TRY(bf(zm[k], e), "net activations");
    }
    if (train) {
        NN_TRY(grad.reserve(P * 4), "net gradients");
        NN_TRY(m.reserve(P * 4), "optimizer state");
        NN_TRY(v.reserve(P * 4), "optimizer state");
        NN_TRY(hipMemsetAsync(m.ptr, 0, P * 4, ctx->stream), "optimizer state");
        NN_TRY(hipMemsetAsync(v.ptr, 0, P * 4, ctx->stream), "optimizer state");
        NN_TRY(dy.reserve(px(1, B) * 4), "net gradients");
        // d[0..6]: gradients of the activations a0 a1 a2 a3 (encoder) and of
        // the merges (n/4, 128), (n/2, 64), (n, 32).
        const int d_div[7] = {1, 2, 4, 8, 4, 2, 1}, d_c[7] = {32, 64, 128, 256, 128, 64, 32};
        for (int k = 0; k < 7; ++k) NN_TRY(bf(d[k], px(d_div[k], B) * d_c[k]), "net gradients");
        const int dc_div[3] = {8, 4, 2}, dc_c[3] = {128, 64, 32};
        for (int k = 0; k < 3; ++k) NN_TRY(bf(dc[k], px(dc_div[k], B) * dc_c[k]), "net gradients");
        NN_TRY(bf(tmp, px(1, B) * 32), "net gradients");
        size_t part = 0;
        for (int l = 0; l < kConvs; ++l) {
            const int od = out_div[l];
            const int s = splits_for(kSpec[l].cout, kSpec[l].cin, px(od, B));
            const size_t e = static_cast<size_t>(s) * kSpec[l].cout * (9 * kSpec[l].cin + 1);
            if (e > part) part = e;
        }
        NN_TRY(wpart.reserve(part * 4), "net gradients");
        NN_TRY(hpart.reserve(((px(1, B) + kHeadPer - 1) / kHeadPer) * (9 * kHeadCin + 1) * 4), "net gradients");
        NN_TRY(mse.reserve(kMseBlocks * 4), "net gradients");
    }
    return refresh_weights();
}

// bf16 copies of the weights and, for training, their transposes.
int darcy_net::refresh_weights() {
    NN_TRY(darcy_nn_to_bf16(params.as<float>(), wb.ptr, kLayout.total, ctx->stream), "net weights");
    for (int l = 1; l < kConvs; ++l)
        NN_TRY(darcy_nn_weight_t(wb.as<uint16_t>() + kLayout.w[l], kSpec[l].cout, kSpec[l].cin,
                                 wt.as<uint16_t>() + kLayout.w[l], ctx->stream),
               "net weights");
    return DARCY_OK;
}

int darcy_net::upload_tables(const void* const* a, const void* const* u, const double* f, int64_t count) {
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

// Forward pass of problems [offset, offset + b) of the uploaded tables.
int darcy_net::forward(int b, int64_t offset) {
    hipStream_t s = ctx->stream;
    const double* const* a = tab_a.as<const double*>() + offset;
    NN_TRY(darcy_nn_features(a, n, b, feat.ptr, s), "net features");
    const void* x = feat.ptr;
    for (int l = 0; l < 4; ++l) {
        const int h = side(l);
        NN_TRY(darcy_nn_conv_fwd(x, b, h, h, kSpec[l].cin, wbf(l), bias(l), kSpec[l].cout, kSpec[l].stride,
                                 train ? z[l].ptr : nullptr, act[l].ptr, s),
               "net encoder");
        x = act[l].ptr;
    }
    const void* skip[3] = {act[2].ptr, act[1].ptr, act[0].ptr};
    for (int k = 0; k < 3; ++k) {
        const int l = 4 + k, h = side(l);
        NN_TRY(darcy_nn_conv_fwd(x, b, h, h, kSpec[l].cin, wbf(l), bias(l), kSpec[l].cout, 1, z[l].ptr, nullptr, s),
               "net decoder");
        NN_TRY(darcy_nn_up_add_gelu(z[l].ptr, b, h, h, kSpec[l].cout, skip[k], train ? zm[k].ptr : nullptr, am[k].ptr,
                                    s),
               "net decoder");
        x = am[k].ptr;
    }
    NN_TRY(darcy_nn_head_fwd(x, b, n, n, kHeadCin, params.as<float>() + kLayout.w[kConvs],
                             params.as<float>() + kLayout.b[kConvs], y.as<float>(), s),
           "net head");
    return DARCY_OK;
}

// Weight and bias gradient of conv l from its input x and output gradient dz.
int darcy_net::wgrad(int l, const void* x, const void* dz, int b) {
    const ConvSpec& c = kSpec[l];
    const int h = side(l), oh = h / c.stride;
    const size_t P = static_cast<size_t>(b) * oh * oh;
    NN_TRY(darcy_nn_transpose(dz, static_cast<int>(P), c.cout, tmp.ptr, ctx->stream), "net backward");
    const int splits = splits_for(c.cout, c.cin, P);
    NN_TRY(darcy_nn_conv_bwd_weight(x, b, h, h, c.cin, tmp.ptr, c.cout, c.stride, splits, wpart.as<float>(),
                                    grad.as<float>() + kLayout.w[l], grad.as<float>() + kLayout.b[l], ctx->stream),
           "net backward");
    return DARCY_OK;
}

// Loss and gradient for the batch whose forward pass just ran.
int darcy_net::backward(int b, int64_t offset, double* loss) {
    hipStream_t s = ctx->stream;
    const double* const* u = tab_u.as<const double*>() + offset;
    NN_TRY(darcy_nn_mse(y.as<float>(), u, n, b, kScale, dy.as<float>(), mse.as<float>(), kMseBlocks, s), "net loss");
    // Head.
    const float* wh = params.as<float>() + kLayout.w[kConvs];
    NN_TRY(darcy_nn_head_bwd_data(dy.as<float>(), b, n, n, kHeadCin, wh, d[6].ptr, s), "net backward");
    NN_TRY(darcy_nn_head_bwd_weight(am[2].ptr, b, n, n, kHeadCin, dy.as<float>(), kHeadPer, hpart.as<float>(),
                                    grad.as<float>() + kLayout.w[kConvs], grad.as<float>() + kLayout.b[kConvs], s),
           "net backward");
    // Decoder: merge k (sizes n/4, n/2, n) follows conv 4 + k; walk back.
    const void* conv_in[3] = {act[3].ptr, am[0].ptr, am[1].ptr};
    DevBuf* dmerge[3] = {&d[4], &d[5], &d[6]};
    DevBuf* dconv_in[3] = {&d[3], &d[4], &d[5]};
    for (int k = 2; k >= 0; --k) {
        const int l = 4 + k, h = side(l), ch = kSpec[l].cout;
        const size_t cnt = static_cast<size_t>(b) * (2 * h) * (2 * h) * ch;
        NN_TRY(darcy_nn_gelu_bwd(dmerge[k]->ptr, zm[k].ptr, cnt, s), "net backward");
        NN_TRY(darcy_nn_down_sum(dmerge[k]->ptr, b, h, h, ch, dc[k].ptr, s), "net backward");
        if (int r = wgrad(l, conv_in[k], dc[k].ptr, b)) return r;
        NN_TRY(darcy_nn_conv_bwd_data(dc[k].ptr, b, h, h, kSpec[l].cin, wtr(l), ch, 1, dconv_in[k]->ptr, s),
               "net backward");
    }
    // Encoder: conv l reads act[l-1] (or the features) and feeds act[l];
    // act[0..2] also receive the skip gradients held by merges 2, 1, 0.
    const size_t cnt3 = px(8, b) * 256;
    NN_TRY(darcy_nn_gelu_bwd(d[3].ptr, z[3].ptr, cnt3, s), "net backward");
    for (int l = 3; l >= 0; --l) {
        const ConvSpec& c = kSpec[l];
        const int h = side(l);
        const void* x = l > 0 ? act[l - 1].ptr : feat.ptr;
        if (int r = wgrad(l, x, d[l].ptr, b)) return r;
        if (l == 0) break;
        const size_t cnt = static_cast<size_t>(b) * h * h * c.cin;
        NN_TRY(darcy_nn_conv_bwd_data(d[l].ptr, b, h, h, c.cin, wtr(l), c.cout, c.stride, d[l - 1].ptr, s),
               "net backward");
        NN_TRY(darcy_nn_add(d[l - 1].ptr, d[7 - l].ptr, cnt, s), "net backward");  // skip of merge 3 - l
        NN_TRY(darcy_nn_gelu_bwd(d[l - 1].ptr, z[l - 1].ptr, cnt, s), "net backward");
    }
    if (loss) {
        std::vector<float> part(kMseBlocks);
        NN_TRY(hipMemcpyAsync(part.data(), mse.ptr, kMseBlocks * 4, hipMemcpyDeviceToHost, s), "net loss");
        NN_TRY(hipStreamSynchronize(s), "net loss");
        double sum = 0.0;
        for (float p : part) sum += p;
        *loss = sum;
    }
    return DARCY_OK;
}

// ------------------------------------------------------------ C ABI

namespace {

int check_net(darcy_net* net) {
    if (!net || !net->ctx) return DARCY_INVALID_INPUT;
    return DARCY_OK;
}

hipError_t set_device(darcy_ctx* ctx) { return hipSetDevice(ctx->device); }

}  // namespace

extern "C" int64_t darcy_net_param_count(void) { return static_cast<int64_t>(kLayout.total); }

namespace {

void release_all(darcy_net* net) {
    DevBuf* all[] = {&net->params, &net->wb,  &net->wt,   &net->grad,  &net->m,     &net->v,   &net->feat,  &net->y,
                     &net->dy,     &net->tmp, &net->wpart, &net->hpart, &net->mse, &net->tab_a, &net->tab_u, &net->tab_f};
    for (DevBuf* b : all) b->release();
    for (int l = 0; l < kConvs; ++l) {
        net->z[l].release();
        net->act[l].release();
    }
    for (int k = 0; k < 3; ++k) {
        net->zm[k].release();
        net->am[k].release();
        net->dc[k].release();
    }
    for (int k = 0; k < 7; ++k) net->d[k].release();
}

}  // namespace

extern "C" void darcy_net_destroy(darcy_net* net) {
    if (!net) return;
    if (net->ctx) {
        std::lock_guard<std::mutex> lock(net->ctx->mu);
        (void)hipSetDevice(net->ctx->device);
        (void)hipStreamSynchronize(net->ctx->stream);
        release_all(net);
    }
    delete net;
}

extern "C" int darcy_net_create(darcy_ctx* ctx, int64_t n, int64_t max_batch, int32_t train, darcy_net** out) {
    if (!ctx || !out) return DARCY_INVALID_INPUT;
    *out = nullptr;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (n < 32 || n % 32 != 0 || n > 4096)
        return fail(ctx, DARCY_INVALID_INPUT, "the warm-start network needs n a multiple of 32 in [32, 4096]");
    if (max_batch < 1 || static_cast<double>(max_batch) * n * n > 2.0e9 / 32)
        return fail(ctx, DARCY_INVALID_INPUT, "max_batch must be at least 1 and max_batch * n * n at most 6.25e7");
    NN_TRY(set_device(ctx), "hipSetDevice");
    darcy_net* net = new darcy_net;
    net->ctx = ctx;
    net->n = static_cast<int>(n);
    net->max_batch = static_cast<int>(max_batch);
    net->train = train != 0;
    const int rc = net->alloc();
    if (rc != DARCY_OK) {
        (void)hipStreamSynchronize(ctx->stream);
        release_all(net);
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
    if (count != static_cast<int64_t>(kLayout.total))
        return fail(ctx, DARCY_INVALID_INPUT, "parameter vector has the wrong length");
    NN_TRY(set_device(ctx), "hipSetDevice");
    NN_TRY(hipMemcpyAsync(net->params.ptr, params, kLayout.total * 4, hipMemcpyHostToDevice, ctx->stream),
           "net parameters");
    if (net->train) {
        NN_TRY(hipMemsetAsync(net->m.ptr, 0, kLayout.total * 4, ctx->stream), "optimizer state");
        NN_TRY(hipMemsetAsync(net->v.ptr, 0, kLayout.total * 4, ctx->stream), "optimizer state");
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
    if (count != static_cast<int64_t>(kLayout.total))
        return fail(ctx, DARCY_INVALID_INPUT, "parameter vector has the wrong length");
    NN_TRY(set_device(ctx), "hipSetDevice");
    NN_TRY(hipMemcpyAsync(params, net->params.ptr, kLayout.total * 4, hipMemcpyDeviceToHost, ctx->stream),
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
    NN_TRY(hipEventRecord(ctx->ev0, ctx->stream), "hipEventRecord");
    if (int r = net->upload_tables(reinterpret_cast<const void* const*>(a), reinterpret_cast<const void* const*>(u0),
                                   f, count))
        return r;
    for (int64_t off = 0; off < count; off += net->max_batch) {
        const int b = static_cast<int>(count - off < net->max_batch ? count - off : net->max_batch);
        if (int r = net->forward(b, off)) return r;
        NN_TRY(darcy_nn_output(net->y.as<float>(), net->n, b, net->tab_f.as<double>() + off, kScale,
                               net->tab_u.as<double*>() + off, ctx->stream),
               "net output");
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
        NN_TRY(hipMemcpyAsync(grad_out, net->grad.ptr, kLayout.total * 4, hipMemcpyDeviceToHost, ctx->stream),
               "net gradient");
        NN_TRY(hipStreamSynchronize(ctx->stream), "net gradient");
    }
    if (lr) {
        const float b1 = 0.9f, b2 = 0.999f;
        net->step += 1;
        const float c1 = 1.0f - std::pow(b1, static_cast<float>(net->step));
        const float c2 = 1.0f - std::pow(b2, static_cast<float>(net->step));
        NN_TRY(darcy_nn_adam(net->params.as<float>(), net->grad.as<float>(), net->m.as<float>(), net->v.as<float>(),
                             net->wb.ptr, kLayout.total, *lr, b1, b2, 1e-8f, c1, c2, ctx->stream),
               "net update");
        for (int l = 1; l < kConvs; ++l)
            NN_TRY(darcy_nn_weight_t(net->wb.as<uint16_t>() + kLayout.w[l], kSpec[l].cout, kSpec[l].cin,
                                     net->wt.as<uint16_t>() + kLayout.w[l], ctx->stream),
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
