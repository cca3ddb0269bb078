// spt_guide.hip - the learned guide's kernels (batched inference, fused
// training step, Adam) and their C ABI (include/spt.h, spt_guide_*).
#include <chrono>
#include <cmath>
#include <cstring>

#include "spt_guide.h"

namespace spt {

using namespace guide;

namespace {

// Raw outputs for queries (p, n, wo) as 9 floats with a stride: each wave
// encodes 64 queries (one per lane) into its stage, then runs the
// register-resident MLP on 4 groups of 16.
__global__ __launch_bounds__(kBlock) void k_guide_eval(const float* __restrict__ theta,
                                                       const unsigned short* __restrict__ wbf,
                                                       const float* __restrict__ q, uint32_t stride, uint32_t n,
                                                       float* __restrict__ out, GuideBox box) {
    extern __shared__ float dyn[];
    SharedWeights& sw = *reinterpret_cast<SharedWeights*>(dyn);
    WaveStage* stages = reinterpret_cast<WaveStage*>(reinterpret_cast<char*>(dyn) + sizeof(SharedWeights));
    load_weights(sw, wbf, theta);
    __syncthreads();
    const int wave = threadIdx.x >> 6, r = __lane_id() & 15;
    WaveStage& st = stages[wave];
    for (uint32_t base = blockIdx.x * kBlock; base < n; base += gridDim.x * kBlock) {
        const uint32_t row = base + threadIdx.x;
        float f[kEnc];
        if (row < n) {
            const float* p = q + (size_t)row * stride;
            encode(V3{p[0], p[1], p[2]}, V3{p[3], p[4], p[5]}, V3{p[6], p[7], p[8]}, box.bmin, box.bext, f);
        } else {
            for (int i = 0; i < kEnc; ++i) f[i] = 0.0f;
        }
        stage_row(st, f);
        __syncthreads();
        const uint32_t wbase = base + (uint32_t)wave * 64;
        if (wbase < n)
            mlp_stage_fwd(sw, st, [&](int g16, const f32x4 y[3]) {
                const uint32_t smp = wbase + 16 * g16 + r;
                if (smp < n) store_outputs(out + (size_t)smp * kOut, y);
            });
        __syncthreads();
    }
}

// Sum of the record weights (for the batch-mean normalisation).
__global__ __launch_bounds__(kBlock) void k_guide_wsum(const float* __restrict__ rec, uint32_t n,
                                                       float* __restrict__ acc) {
    float s = 0.0f;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) s += rec[(size_t)i * kRecord + 12];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (__lane_id() == 0) atomicAdd(acc, s);
}

// One training step's gradient: forward, per-row KL-loss gradient (weights
// normalised by the batch mean, averaged over the batch, times the loss
// scale), backward; weight gradients summed in LDS then added to grad (fp32).
// Reads ctl->wsum and ctl->scale, adds the step's loss to ctl->loss.
__global__ __launch_bounds__(kBlock) void k_guide_train(const float* __restrict__ theta,
                                                        const unsigned short* __restrict__ wbf,
                                                        const float* __restrict__ rec, uint32_t n,
                                                        float* __restrict__ grad, GuideCtl* __restrict__ ctl,
                                                        GuideBox box) {
    extern __shared__ float dyn[];
    SharedWeights& sw = *reinterpret_cast<SharedWeights*>(dyn);
    float* dW = reinterpret_cast<float*>(reinterpret_cast<char*>(dyn) + sizeof(SharedWeights));
    WaveTiles* tiles = reinterpret_cast<WaveTiles*>(dW + kParams);
    load_weights(sw, wbf, theta);
    for (int i = threadIdx.x; i < kParams; i += blockDim.x) dW[i] = 0.0f;
    __syncthreads();
    const int wave = threadIdx.x >> 6, l = __lane_id();
    WaveTiles& t = tiles[wave];
    const float wmean = fmaxf(ctl->wsum / (float)n, 1e-30f);
    const float scale = ctl->scale / (float)n;
    float loss = 0.0f;
    const uint32_t ntiles = (n + 15) / 16;
    for (uint32_t base = blockIdx.x * (kBlock / 64); base < ntiles; base += gridDim.x * (kBlock / 64)) {
        const uint32_t tile = base + wave;
        const bool act = tile < ntiles;
        if (act && l < 16) {
            const uint32_t row = tile * 16 + l;
            float f[kEnc];
            if (row < n) {
                const float* r = rec + (size_t)row * kRecord;
                encode(V3{r[0], r[1], r[2]}, V3{r[3], r[4], r[5]}, V3{r[6], r[7], r[8]}, box.bmin, box.bext, f);
            } else {
                for (int i = 0; i < kEnc; ++i) f[i] = 0.0f;
            }
            for (int i = 0; i < kEnc; ++i) t.x[l * kEnc + i] = f2bf(f[i]);
        }
        __syncthreads();
        if (act) mlp_tile_fwd(sw, t);
        __syncthreads();
        if (act && l < 16) {
            const uint32_t row = tile * 16 + l;
            float dy[kOutP];
            if (row < n) {
                const float* r = rec + (size_t)row * kRecord;
                const float W = r[12] / wmean;
                loss += loss_grad(&t.y[l * kOutP], V3{r[9], r[10], r[11]}, W, dy);
            } else {
                for (int i = 0; i < kOutP; ++i) dy[i] = 0.0f;
            }
            for (int i = 0; i < kOutP; ++i) {
                const float v = dy[i] * scale;
                t.dy[l * kOutP + i] = f2bf(v);
                if (v != 0.0f) atomicAdd(&dW[kB4 + i], v);
            }
        }
        __syncthreads();
        if (act) {
            dense_wgrad<kOutP, kHid>(t.dy, t.h3, dW + kW4);
            dense_igrad<kOutP, kHid>(t.dy, sw.w4, t.h3, t.da, dW + kB3);
        }
        __syncthreads();
        if (act) {
            dense_wgrad<kHid, kHid>(t.da, t.h2, dW + kW3);
            dense_igrad<kHid, kHid>(t.da, sw.w3, t.h2, t.db, dW + kB2);
        }
        __syncthreads();
        if (act) {
            dense_wgrad<kHid, kHid>(t.db, t.h1, dW + kW2);
            dense_igrad<kHid, kHid>(t.db, sw.w2, t.h1, t.da, dW + kB1);
        }
        __syncthreads();
        if (act) dense_wgrad<kHid, kEnc>(t.da, t.x, dW + kW1);
        __syncthreads();
    }
    for (int o = 32; o > 0; o >>= 1) loss += __shfl_xor(loss, o);
    if (l == 0) atomicAdd(&ctl->loss, loss / (float)n);
    __syncthreads();
    for (int i = threadIdx.x; i < kParams; i += blockDim.x)
        if (dW[i] != 0.0f) atomicAdd(&grad[i], dW[i]);
}

// Start of a step: clears the per-step fields of the control block.
__global__ void k_guide_begin(GuideCtl* ctl) {
    if (threadIdx.x == 0) {
        ctl->wsum = 0.0f;
        ctl->loss = 0.0f;
        ctl->bad = 0u;
    }
}

// ctl->bad = 1 when any gradient entry is not finite (loss-scale overflow)
__global__ __launch_bounds__(kBlock) void k_guide_check(const float* __restrict__ grad, GuideCtl* ctl) {
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < (uint32_t)kParams; i += gridDim.x * kBlock)
        if (!isfinite(grad[i])) atomicOr(&ctl->bad, 1u);
}

// FP32 Adam on the unscaled gradient, skipped when the gradient overflowed or
// the batch carried no weight; refreshes the bf16 copy and clears grad. The
// step number (for the bias corrections) is the count of successful steps + 1.
__global__ __launch_bounds__(kBlock) void k_guide_adam(float* __restrict__ theta, float* __restrict__ m,
                                                       float* __restrict__ v, float* __restrict__ grad,
                                                       unsigned short* __restrict__ wbf, const GuideCtl* ctl,
                                                       float lr) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= (uint32_t)kParams) return;
    if (ctl->bad == 0u && ctl->wsum > 0.0f) {
        const float t = (float)(ctl->steps + 1u);
        const float bc1 = 1.0f - powf(0.9f, t), bc2 = 1.0f - powf(0.99f, t);
        const float g = grad[i] * (1.0f / ctl->scale);
        const float mi = 0.9f * m[i] + 0.1f * g;
        const float vi = 0.99f * v[i] + 0.01f * g * g;
        m[i] = mi;
        v[i] = vi;
        theta[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + 1e-8f);
    }
    wbf[i] = f2bf(theta[i]);
    grad[i] = 0.0f;
}

// End of a step: counters, last loss, and the loss scale halved on overflow.
__global__ void k_guide_post(GuideCtl* ctl, uint32_t n) {
    if (threadIdx.x != 0) return;
    if (ctl->bad == 0u && ctl->wsum > 0.0f) {
        ctl->steps += 1u;
        ctl->trained += n;
        ctl->last_loss = ctl->loss;
    } else {
        ctl->skipped += 1u;
        if (ctl->bad) ctl->scale *= 0.5f;
    }
}

__global__ void k_guide_tobf16(const float* __restrict__ theta, unsigned short* __restrict__ wbf) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < (uint32_t)kParams) wbf[i] = f2bf(theta[i]);
}

size_t eval_lds_bytes() { return sizeof(SharedWeights) + sizeof(WaveStage) * (kBlock / 64); }
size_t train_lds_bytes() { return sizeof(SharedWeights) + sizeof(float) * kParams + sizeof(WaveTiles) * (kBlock / 64); }

// The training and evaluation kernels use more than the default 64 KiB of
// dynamic LDS (gfx950 has 160 KiB per CU).
spt_status allow_lds(spt_ctx* ctx) {
    static bool done = false;
    if (done) return SPT_OK;
    spt_status s = check_hip(ctx,
                             hipFuncSetAttribute(reinterpret_cast<const void*>(&k_guide_train),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)train_lds_bytes()),
                             "train LDS size");
    if (s == SPT_OK)
        s = check_hip(ctx,
                      hipFuncSetAttribute(reinterpret_cast<const void*>(&k_guide_eval),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)eval_lds_bytes()),
                      "eval LDS size");
    done = (s == SPT_OK);
    return s;
}

}  // namespace

GuideBox guide_box(const spt_ctx* ctx) {
    GuideBox b;
    for (int i = 0; i < 3; ++i) {
        b.bmin[i] = ctx->root_box[i];
        b.bext[i] = fmaxf(ctx->root_box[3 + i] - ctx->root_box[i], 1e-6f);
    }
    return b;
}

spt_status guide_eval_launch(spt_ctx* ctx, const float* d_q, uint32_t stride, uint32_t n, float* d_out) {
    GuideState& gs = ctx->guide;
    spt_status ls = allow_lds(ctx);
    if (ls != SPT_OK) return ls;
    const uint32_t grid = std::min<uint32_t>((n + kBlock - 1) / kBlock, (uint32_t)ctx->num_cus * 4);
    hipLaunchKernelGGL(k_guide_eval, dim3(grid), dim3(kBlock), eval_lds_bytes(), ctx->stream, gs.d_theta, gs.d_wbf,
                       d_q, stride, n, d_out, guide_box(ctx));
    return check_hip(ctx, hipGetLastError(), "guide eval launch");
}

spt_status guide_train_launch(spt_ctx* ctx, const float* d_rec, uint32_t n, float* loss_out) {
    GuideState& gs = ctx->guide;
    spt_status ls = allow_lds(ctx);
    if (ls != SPT_OK) return ls;
    auto t0 = std::chrono::steady_clock::now();
    hipStream_t st = ctx->stream;
    GuideCtl* ctl = gs.ctl();
    hipLaunchKernelGGL(k_guide_begin, dim3(1), dim3(64), 0, st, ctl);
    hipLaunchKernelGGL(k_guide_wsum, dim3(256), dim3(kBlock), 0, st, d_rec, n, &ctl->wsum);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (gs.event_pair(e0, e1)) (void)hipEventRecord(e0, st);
    hipLaunchKernelGGL(k_guide_train, dim3((uint32_t)ctx->num_cus), dim3(kBlock), train_lds_bytes(), st, gs.d_theta,
                       gs.d_wbf, d_rec, n, gs.d_grad, ctl, guide_box(ctx));
    if (e1) (void)hipEventRecord(e1, st);
    hipLaunchKernelGGL(k_guide_check, dim3(64), dim3(kBlock), 0, st, gs.d_grad, ctl);
    hipLaunchKernelGGL(k_guide_adam, dim3((kParams + kBlock - 1) / kBlock), dim3(kBlock), 0, st, gs.d_theta, gs.d_m,
                       gs.d_v, gs.d_grad, gs.d_wbf, ctl, gs.lr);
    hipLaunchKernelGGL(k_guide_post, dim3(1), dim3(64), 0, st, ctl, n);
    spt_status s = check_hip(ctx, hipGetLastError(), "guide train launch");
    if (s == SPT_OK && loss_out) {  // only a caller that wants the loss waits for it
        GuideCtl h;
        s = read_small(ctx, &h, ctl, sizeof(h), "guide step");
        if (s == SPT_OK) *loss_out = h.loss;
    }
    gs.train_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return s;
}

spt_status guide_grad_launch(spt_ctx* ctx, const float* d_rec, uint32_t n, float* grad_out, float* loss_out) {
    GuideState& gs = ctx->guide;
    spt_status ls = allow_lds(ctx);
    if (ls != SPT_OK) return ls;
    hipStream_t st = ctx->stream;
    GuideCtl* ctl = gs.ctl();
    hipLaunchKernelGGL(k_guide_begin, dim3(1), dim3(64), 0, st, ctl);
    spt_status s = check_hip(ctx, hipMemsetAsync(gs.d_grad, 0, kParams * sizeof(float), st), "guide grad");
    if (s != SPT_OK) return s;
    hipLaunchKernelGGL(k_guide_wsum, dim3(256), dim3(kBlock), 0, st, d_rec, n, &ctl->wsum);
    hipLaunchKernelGGL(k_guide_train, dim3((uint32_t)ctx->num_cus), dim3(kBlock), train_lds_bytes(), st, gs.d_theta,
                       gs.d_wbf, d_rec, n, gs.d_grad, ctl, guide_box(ctx));
    s = check_hip(ctx, hipMemcpyAsync(grad_out, gs.d_grad, kParams * sizeof(float), hipMemcpyDeviceToHost, st),
                  "guide grad read");
    GuideCtl h;
    if (s == SPT_OK) s = read_small(ctx, &h, ctl, sizeof(h), "guide grad");
    if (s != SPT_OK) return s;
    for (int i = 0; i < kParams; ++i) grad_out[i] /= h.scale;
    if (loss_out) *loss_out = h.loss;
    return check_hip(ctx, hipMemsetAsync(gs.d_grad, 0, kParams * sizeof(float), st), "guide grad clear");
}

}  // namespace spt

using namespace spt;

namespace {

uint64_t sm64(uint64_t& s) {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

spt_status upload_theta(spt_ctx* ctx, const float* theta) {
    GuideState& gs = ctx->guide;
    spt_status s = check_hip(ctx, hipMemcpy(gs.d_theta, theta, kParams * sizeof(float), hipMemcpyHostToDevice), "theta");
    if (s != SPT_OK) return s;
    (void)hipMemset(gs.d_m, 0, kParams * sizeof(float));
    (void)hipMemset(gs.d_v, 0, kParams * sizeof(float));
    (void)hipMemset(gs.d_grad, 0, kParams * sizeof(float));
    hipLaunchKernelGGL(k_guide_tobf16, dim3((kParams + 255) / 256), dim3(256), 0, ctx->stream, gs.d_theta, gs.d_wbf);
    // fresh moments: step count and statistics restart; the loss scale is the configured one
    GuideCtl c{};
    c.scale = gs.loss_scale;
    s = check_hip(ctx, hipMemcpyAsync(gs.ctl(), &c, sizeof(c), hipMemcpyHostToDevice, ctx->stream), "guide control");
    if (s == SPT_OK) s = stream_wait(ctx, "guide upload");
    return s;
}

}  // namespace

extern "C" {

spt_status spt_guide_init(spt_ctx* ctx, const spt_guide_config* cfg) {
    if (!ctx) return SPT_ERR_INVALID;
    spt_guide_config c{};
    c.seed = 1;
    c.learning_rate = 5e-3f;
    c.loss_scale = 65536.0f;
    c.bsdf_fraction = 0.5f;
    c.batch = 262144;
    c.steps_per_pass = 1;
    if (cfg) c = *cfg;
    if (c.steps_per_pass == 0) c.steps_per_pass = 1;
    if (!(c.bsdf_fraction > 0.0f && c.bsdf_fraction < 1.0f) || c.batch == 0 || !(c.learning_rate > 0.0f) ||
        !(c.loss_scale > 0.0f) || c.steps_per_pass > 64 || (uint64_t)c.batch * c.steps_per_pass >= (1ull << 31))
        return set_error(ctx, SPT_ERR_INVALID, "guide config out of range");
    const uint32_t cap = c.batch * c.steps_per_pass;
    GuideState& gs = ctx->guide;
    if (!gs.d_theta) {
        void* p = nullptr;
        const size_t bytes = (size_t)kParams * (4 * 4 + 2) + 256;
        spt_status s = check_hip(ctx, hipMalloc(&p, bytes), "guide state");
        if (s != SPT_OK) return s;
        char* b = reinterpret_cast<char*>(p);
        gs.d_theta = reinterpret_cast<float*>(b);
        gs.d_m = gs.d_theta + kParams;
        gs.d_v = gs.d_m + kParams;
        gs.d_grad = gs.d_v + kParams;
        gs.d_acc = gs.d_grad + kParams;
        gs.d_wbf = reinterpret_cast<unsigned short*>(gs.d_acc + 64);
    }
    if (gs.records_alloc < cap) {
        if (gs.d_records) (void)hipFree(gs.d_records);
        gs.d_records = nullptr;
        gs.records_alloc = 0;
        spt_status s = check_hip(ctx, hipMalloc(&gs.d_records, (size_t)cap * kRecord * sizeof(float) + 256),
                                 "guide records");
        if (s != SPT_OK) return s;
        gs.records_alloc = cap;
    }
    gs.records_cap = cap;
    gs.lr = c.learning_rate;
    gs.loss_scale = c.loss_scale;
    gs.bsdf_fraction = c.bsdf_fraction;
    gs.batch = c.batch;
    gs.train_ms = 0.0;
    gs.kernel_ms = 0.0;
    gs.candidates = 0;
    gs.rec_stride = 1;
    gs.cand_per_path = 0.0;
    gs.collected = 0;
    gs.settle_events();
    // Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights, zero biases; output
    // layer scaled by 0.1 so the first mixtures are broad.
    std::vector<float> th(kParams, 0.0f);
    uint64_t s = c.seed;
    auto unif = [&](float lim) { return (float)(((sm64(s) >> 11) * (1.0 / 9007199254740992.0)) * 2.0 - 1.0) * lim; };
    for (int i = 0; i < kHid * kEnc; ++i) th[kW1 + i] = unif(1.0f / std::sqrt((float)kEnc));
    for (int i = 0; i < kHid * kHid; ++i) th[kW2 + i] = unif(1.0f / 8.0f);
    for (int i = 0; i < kHid * kHid; ++i) th[kW3 + i] = unif(1.0f / 8.0f);
    for (int i = 0; i < kOut * kHid; ++i) th[kW4 + i] = 0.1f * unif(1.0f / 8.0f);
    gs.ready = true;
    return upload_theta(ctx, th.data());
}

spt_status spt_guide_set_params(spt_ctx* ctx, const float* theta) {
    if (!ctx || !theta) return SPT_ERR_INVALID;
    if (!ctx->guide.ready) return set_error(ctx, SPT_ERR_STATE, "spt_guide_init first");
    std::vector<float> th(theta, theta + kParams);
    for (int i = kOut * kHid; i < kOutP * kHid; ++i) th[kW4 + i] = 0.0f;  // padding rows stay zero
    for (int i = kOut; i < kOutP; ++i) th[kB4 + i] = 0.0f;
    return upload_theta(ctx, th.data());
}

spt_status spt_guide_get_params(spt_ctx* ctx, float* theta) {
    if (!ctx || !theta) return SPT_ERR_INVALID;
    if (!ctx->guide.ready) return set_error(ctx, SPT_ERR_STATE, "spt_guide_init first");
    return check_hip(ctx, hipMemcpy(theta, ctx->guide.d_theta, kParams * sizeof(float), hipMemcpyDeviceToHost),
                     "theta read");
}

spt_status spt_guide_eval(spt_ctx* ctx, const float* d_queries, uint64_t n, float* d_out) {
    if (!ctx || (n && (!d_queries || !d_out)) || n >= (1ull << 32)) return SPT_ERR_INVALID;
    if (!ctx->guide.ready) return set_error(ctx, SPT_ERR_STATE, "spt_guide_init first");
    if (!ctx->has_scene) return set_error(ctx, SPT_ERR_STATE, "the encoding needs the scene box: upload a scene");
    if (n == 0) return SPT_OK;
    return guide_eval_launch(ctx, d_queries, 9, (uint32_t)n, d_out);
}

spt_status spt_guide_train(spt_ctx* ctx, const float* d_records, uint32_t n, float* loss) {
    if (!ctx || !d_records || n == 0) return SPT_ERR_INVALID;
    if (!ctx->guide.ready) return set_error(ctx, SPT_ERR_STATE, "spt_guide_init first");
    if (!ctx->has_scene) return set_error(ctx, SPT_ERR_STATE, "the encoding needs the scene box: upload a scene");
    return guide_train_launch(ctx, d_records, n, loss);
}

spt_status spt_guide_gradient(spt_ctx* ctx, const float* d_records, uint32_t n, float* grad, float* loss) {
    if (!ctx || !d_records || n == 0 || !grad) return SPT_ERR_INVALID;
    if (!ctx->guide.ready) return set_error(ctx, SPT_ERR_STATE, "spt_guide_init first");
    if (!ctx->has_scene) return set_error(ctx, SPT_ERR_STATE, "the encoding needs the scene box: upload a scene");
    return guide_grad_launch(ctx, d_records, n, grad, loss);
}

spt_status spt_guide_train_collected(spt_ctx* ctx, float* loss) {
    if (!ctx) return SPT_ERR_INVALID;
    GuideState& gs = ctx->guide;
    if (!gs.ready) return set_error(ctx, SPT_ERR_STATE, "spt_guide_init first");
    if (gs.collected == 0) {
        if (loss) *loss = 0.0f;
        return SPT_OK;
    }
    const uint32_t n = (uint32_t)std::min<uint64_t>(gs.collected, gs.records_cap);
    spt_status s = SPT_OK;
    // one step per batch; a remainder shorter than half a batch joins the last step
    for (uint32_t b0 = 0; b0 < n && s == SPT_OK;) {
        uint32_t len = std::min(gs.batch, n - b0);
        if (n - b0 - len < gs.batch / 2) len = n - b0;
        const bool last = b0 + len >= n;
        s = guide_train_launch(ctx, gs.d_records + (size_t)b0 * kRecord, len, last ? loss : nullptr);
        b0 += len;
    }
    gs.collected = 0;
    return s;
}

spt_status spt_guide_get_stats(spt_ctx* ctx, spt_guide_stats* out) {
    if (!ctx || !out) return SPT_ERR_INVALID;
    GuideState& gs = ctx->guide;
    GuideCtl h{};
    if (gs.ready) {
        spt_status s = read_small(ctx, &h, gs.ctl(), sizeof(h), "guide stats");
        if (s != SPT_OK) return s;
        gs.settle_events();
    } else {
        h.scale = gs.loss_scale;
    }
    out->steps = h.steps;
    out->skipped_steps = h.skipped;
    out->records = gs.collected;
    out->last_loss = h.last_loss;
    out->loss_scale = h.scale;
    out->train_ms = gs.train_ms;
    out->train_kernel_ms = gs.kernel_ms;
    out->trained_records = h.trained;
    out->candidates = gs.candidates;
    out->record_stride = gs.rec_stride;
    out->pad = 0;
    return SPT_OK;
}

spt_status spt_guide_records(spt_ctx* ctx, float** d_records, uint32_t* n) {
    if (!ctx || !d_records || !n) return SPT_ERR_INVALID;
    *d_records = ctx->guide.d_records;
    *n = (uint32_t)std::min<uint64_t>(ctx->guide.collected, ctx->guide.records_cap);
    return SPT_OK;
}

}  // extern "C"
