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
    FwdWeights& sw = *reinterpret_cast<FwdWeights*>(dyn);
    WaveStage* stages = reinterpret_cast<WaveStage*>(reinterpret_cast<char*>(dyn) + sizeof(FwdWeights));
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
__global__ __launch_bounds__(kBlock) void k_guide_wsum(const float* __restrict__ rec, uint32_t n, uint32_t rs,
                                                       float* __restrict__ acc) {
    float s = 0.0f;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        s += rec[(size_t)i * rs * kRecord + 12];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (__lane_id() == 0) atomicAdd(acc, s);
}

// ---------------------------------------------------------------------------
// Training step in four stages over a feature-major scratch (feature f of
// sample s at f * rows + s): forward + loss gradient, backward through the
// activations, weight gradients as split-K MFMA products, reduction.
// ---------------------------------------------------------------------------
struct TrainScratch {
    unsigned short* xT;   // [32][rows]
    unsigned short* h1T;  // [64][rows]
    unsigned short* h2T;
    unsigned short* h3T;
    float* dyT;           // [48][rows], rows 40..47 stay zero
    unsigned short* dz3T; // [64][rows]
    unsigned short* dz2T;
    unsigned short* dz1T;
    float* partial;       // [splits][kParams]
    uint32_t rows;
};

__device__ inline bf16x4 cvt4(float a, float b, float c, float d) {
    bf16x4 r;
    r[0] = (short)f2bf(a);
    r[1] = (short)f2bf(b);
    r[2] = (short)f2bf(c);
    r[3] = (short)f2bf(d);
    return r;
}

// feature-major bf16 column: features f0 + i (i = 0..3) of sample s
__device__ inline bf16x4 ld_col4(const unsigned short* T, uint32_t rows, uint32_t f0, uint32_t s) {
    bf16x4 r;
    r[0] = (short)T[(size_t)f0 * rows + s];
    r[1] = (short)T[(size_t)(f0 + 1) * rows + s];
    r[2] = (short)T[(size_t)(f0 + 2) * rows + s];
    r[3] = (short)T[(size_t)(f0 + 3) * rows + s];
    return r;
}

__device__ inline void st_col4(unsigned short* T, uint32_t rows, uint32_t f0, uint32_t s, bf16x4 v) {
    T[(size_t)f0 * rows + s] = (unsigned short)v[0];
    T[(size_t)(f0 + 1) * rows + s] = (unsigned short)v[1];
    T[(size_t)(f0 + 2) * rows + s] = (unsigned short)v[2];
    T[(size_t)(f0 + 3) * rows + s] = (unsigned short)v[3];
}

// A operand of W^T for the backward pass: A[r][4g + j] = W[16 ot + 4g + j][16 kt + r]
__device__ inline bf16x4 ld_wT(const unsigned short* W, int K, int ot, int kt) {
    const int l = __lane_id(), r = l & 15, g = l >> 4;
    return ld4s(W + (16 * ot + 4 * g) * K + 16 * kt + r, K);
}

// Stage 1: encode, forward (activations stored feature-major), and each
// record's loss gradient w.r.t. the 48 raw outputs, scaled by loss_scale / n
// (rows n..rows16: zero gradient). One block = 256 records, lane = record.
__global__ __launch_bounds__(kBlock) void k_train_fwd(const float* __restrict__ theta,
                                                      const unsigned short* __restrict__ wbf,
                                                      const float* __restrict__ rec, uint32_t n, uint32_t rs,
                                                      uint32_t n16, GuideCtl* __restrict__ ctl, GuideBox box,
                                                      TrainScratch ts) {
    extern __shared__ float dyn[];
    FwdWeights& sw = *reinterpret_cast<FwdWeights*>(dyn);
    WaveStage* stages = reinterpret_cast<WaveStage*>(reinterpret_cast<char*>(dyn) + sizeof(FwdWeights));
    float* ystage = reinterpret_cast<float*>(stages + kBlock / 64);  // [4 waves][64][kOut]
    load_weights(sw, wbf, theta);
    const int wave = threadIdx.x >> 6, l = __lane_id(), r = l & 15, g = l >> 4;
    WaveStage& st = stages[wave];
    float* ys = ystage + wave * 64 * kOut;
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t wbase = blockIdx.x * kBlock + wave * 64;
    float f[kEnc];
    if (s < n) {
        const float* p = rec + (size_t)s * rs * kRecord;
        encode(V3{p[0], p[1], p[2]}, V3{p[3], p[4], p[5]}, V3{p[6], p[7], p[8]}, box.bmin, box.bext, f);
    } else {
        for (int i = 0; i < kEnc; ++i) f[i] = 0.0f;
    }
    stage_row(st, f);
    if (s < n16)
        for (int i = 0; i < kEnc; ++i) ts.xT[(size_t)i * ts.rows + s] = f2bf(f[i]);
    __syncthreads();
    // forward, column form, 16 records at a time; activations out to scratch
    for (int q = 0; q < 4; ++q) {
        const uint32_t sq = wbase + 16 * q + r;
        const bool keep = sq < n16;
        const unsigned short* row = st.x + (16 * q + r) * kStageRow;
        const bf16x4 x[2] = {ld4(row + 4 * g), ld4(row + 16 + 4 * g)};
        bf16x4 ha[4], hb[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            ha[t] = pack_relu(mfma_w(sw.w1 + 64 * t, x[0], x[1], bias4(sw.b1 + 16 * t + 4 * g)));
            if (keep) st_col4(ts.h1T, ts.rows, 16 * t + 4 * g, sq, ha[t]);
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            f32x4 acc = bias4(sw.b2 + 16 * t + 4 * g);
#pragma unroll
            for (int p = 0; p < 2; ++p) acc = mfma_w(sw.w2 + 64 * (2 * t + p), ha[2 * p], ha[2 * p + 1], acc);
            hb[t] = pack_relu(acc);
            if (keep) st_col4(ts.h2T, ts.rows, 16 * t + 4 * g, sq, hb[t]);
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            f32x4 acc = bias4(sw.b3 + 16 * t + 4 * g);
#pragma unroll
            for (int p = 0; p < 2; ++p) acc = mfma_w(sw.w3 + 64 * (2 * t + p), hb[2 * p], hb[2 * p + 1], acc);
            ha[t] = pack_relu(acc);
            if (keep) st_col4(ts.h3T, ts.rows, 16 * t + 4 * g, sq, ha[t]);
        }
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            f32x4 acc = bias4(sw.b4 + 16 * t + 4 * g);
#pragma unroll
            for (int p = 0; p < 2; ++p) acc = mfma_w(sw.w4 + 64 * (2 * t + p), ha[2 * p], ha[2 * p + 1], acc);
            if (16 * t + 4 * g < kOut)
                for (int i = 0; i < 4; ++i) ys[(16 * q + r) * kOut + 16 * t + 4 * g + i] = acc[i];
        }
    }
    __syncthreads();
    // loss gradient of this lane's record
    const float wmean = fmaxf(ctl->wsum / (float)n, 1e-30f);
    const float scale = ctl->scale / (float)n;
    float loss = 0.0f;
    if (s < n16) {
        float dy[kOutP];
        if (s < n) {
            const float* p = rec + (size_t)s * rs * kRecord;
            loss = loss_grad(ys + l * kOut, V3{p[9], p[10], p[11]}, p[12] / wmean, dy);
        } else {
            for (int i = 0; i < kOutP; ++i) dy[i] = 0.0f;
        }
        for (int o = 0; o < kOut; ++o) ts.dyT[(size_t)o * ts.rows + s] = dy[o] * scale;
    }
    for (int o = 32; o > 0; o >>= 1) loss += __shfl_xor(loss, o);
    if (l == 0) atomicAdd(&ctl->loss, loss / (float)n);
}

// Stage 2: gradients w.r.t. the hidden pre-activations, dz3 = (W4^T dy) . [h3 > 0],
// dz2 = (W3^T dz3) . [h2 > 0], dz1 = (W2^T dz2) . [h1 > 0], column form, 16
// records per wave per iteration; stored feature-major as bf16.
__global__ __launch_bounds__(kBlock) void k_train_bwd(const float* __restrict__ theta,
                                                      const unsigned short* __restrict__ wbf, uint32_t n16,
                                                      TrainScratch ts) {
    extern __shared__ float dyn[];
    SharedWeights& sw = *reinterpret_cast<SharedWeights*>(dyn);
    load_weights(sw, wbf, theta);
    __syncthreads();
    const int l = __lane_id(), r = l & 15, g = l >> 4;
    const uint32_t wave = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6), nw = gridDim.x * (kBlock / 64);
    for (uint32_t base = wave * 16; base < n16; base += nw * 16) {
        const uint32_t s = base + r;
        bf16x4 dy[3];
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            const float* c = ts.dyT + (size_t)(16 * t + 4 * g) * ts.rows + s;
            dy[t] = cvt4(c[0], c[ts.rows], c[2 * (size_t)ts.rows], c[3 * (size_t)ts.rows]);
        }
        bf16x4 za[4], zb[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {  // dz3: K = 48 outputs
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
            acc = mfma_k32(ld_wT(sw.w4, kHid, 0, t), ld_wT(sw.w4, kHid, 1, t), dy[0], dy[1], acc);
            acc = mfma_k16(ld_wT(sw.w4, kHid, 2, t), dy[2], acc);
            const bf16x4 h = ld_col4(ts.h3T, ts.rows, 16 * t + 4 * g, s);
            for (int i = 0; i < 4; ++i)
                if (bf2f((unsigned short)h[i]) <= 0.0f) acc[i] = 0.0f;
            za[t] = cvt4(acc[0], acc[1], acc[2], acc[3]);
            st_col4(ts.dz3T, ts.rows, 16 * t + 4 * g, s, za[t]);
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) {  // dz2
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int k = 0; k < 4; k += 2)
                acc = mfma_k32(ld_wT(sw.w3, kHid, k, t), ld_wT(sw.w3, kHid, k + 1, t), za[k], za[k + 1], acc);
            const bf16x4 h = ld_col4(ts.h2T, ts.rows, 16 * t + 4 * g, s);
            for (int i = 0; i < 4; ++i)
                if (bf2f((unsigned short)h[i]) <= 0.0f) acc[i] = 0.0f;
            zb[t] = cvt4(acc[0], acc[1], acc[2], acc[3]);
            st_col4(ts.dz2T, ts.rows, 16 * t + 4 * g, s, zb[t]);
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) {  // dz1
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int k = 0; k < 4; k += 2)
                acc = mfma_k32(ld_wT(sw.w2, kHid, k, t), ld_wT(sw.w2, kHid, k + 1, t), zb[k], zb[k + 1], acc);
            const bf16x4 h = ld_col4(ts.h1T, ts.rows, 16 * t + 4 * g, s);
            for (int i = 0; i < 4; ++i)
                if (bf2f((unsigned short)h[i]) <= 0.0f) acc[i] = 0.0f;
            st_col4(ts.dz1T, ts.rows, 16 * t + 4 * g, s, cvt4(acc[0], acc[1], acc[2], acc[3]));
        }
    }
}

// 4 consecutive samples s0 .. s0 + 3 of feature f (feature-major bf16 / f32)
__device__ inline bf16x4 ld_row4(const unsigned short* T, uint32_t rows, uint32_t f, uint32_t s0) {
    const uint2 v = *reinterpret_cast<const uint2*>(T + (size_t)f * rows + s0);
    bf16x4 r;
    r[0] = (short)(v.x & 0xFFFFu);
    r[1] = (short)(v.x >> 16);
    r[2] = (short)(v.y & 0xFFFFu);
    r[3] = (short)(v.y >> 16);
    return r;
}

// Stage 3: weight and bias gradients of one layer over one slice of samples:
// dW[o][k] = sum_s dz[o][s] in[k][s] as MFMA with the samples as the
// reduction axis (K), db[o] = sum_s dz[o][s]; written to partial[split].
// Block = (layer, split); wave w takes output tile w.
__global__ __launch_bounds__(kBlock) void k_train_wgrad(TrainScratch ts, uint32_t n16, uint32_t span) {
    const int layer = blockIdx.x & 3;  // 0: W1, 1: W2, 2: W3, 3: W4
    const uint32_t split = blockIdx.x >> 2;
    const int ot = threadIdx.x >> 6, l = __lane_id(), r = l & 15, g = l >> 4;
    const int n_ot = layer == 3 ? 3 : 4, n_kt = layer == 0 ? 2 : 4, K = layer == 0 ? kEnc : kHid;
    if (ot >= n_ot) return;
    const unsigned short* dz = layer == 0 ? ts.dz1T : layer == 1 ? ts.dz2T : ts.dz3T;
    const unsigned short* in = layer == 0 ? ts.xT : layer == 1 ? ts.h1T : layer == 2 ? ts.h2T : ts.h3T;
    const uint32_t s_begin = split * span, s_end = min(n16, s_begin + span);
    f32x4 acc[4];
    for (int k = 0; k < 4; ++k) acc[k] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    float bsum = 0.0f;
    // dz of 4 samples s0 + 4 g .. + 3 for output row 16 ot + r (and their bias sum)
    auto dz4 = [&](uint32_t s0) {
        bf16x4 a;
        if (layer == 3) {
            const float4 v = *reinterpret_cast<const float4*>(ts.dyT + (size_t)(16 * ot + r) * ts.rows + s0 + 4 * g);
            a = cvt4(v.x, v.y, v.z, v.w);
            bsum += (v.x + v.y) + (v.z + v.w);
        } else {
            a = ld_row4(dz, ts.rows, 16 * ot + r, s0 + 4 * g);
            bsum += (bf2f((unsigned short)a[0]) + bf2f((unsigned short)a[1])) +
                    (bf2f((unsigned short)a[2]) + bf2f((unsigned short)a[3]));
        }
        return a;
    };
    // samples are the reduction axis: 32 per MFMA, a last odd 16 alone
    uint32_t s0 = s_begin;
    for (; s0 + 32 <= s_end; s0 += 32) {
        const bf16x4 a0 = dz4(s0), a1 = dz4(s0 + 16);
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (k < n_kt)
                acc[k] = mfma_k32(a0, a1, ld_row4(in, ts.rows, 16 * k + r, s0 + 4 * g),
                                  ld_row4(in, ts.rows, 16 * k + r, s0 + 16 + 4 * g), acc[k]);
    }
    if (s0 < s_end) {
        const bf16x4 a = dz4(s0);
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (k < n_kt) acc[k] = mfma_k16(a, ld_row4(in, ts.rows, 16 * k + r, s0 + 4 * g), acc[k]);
    }
    float* out = ts.partial + (size_t)split * kParams;
    const int wofs = layer == 0 ? kW1 : layer == 1 ? kW2 : layer == 2 ? kW3 : kW4;
    const int bofs = layer == 0 ? kB1 : layer == 1 ? kB2 : layer == 2 ? kB3 : kB4;
    for (int k = 0; k < n_kt; ++k)
        for (int i = 0; i < 4; ++i) out[wofs + (16 * ot + 4 * g + i) * K + 16 * k + r] = acc[k][i];
    bsum += __shfl_xor(bsum, 16);
    bsum += __shfl_xor(bsum, 32);
    if (g == 0) out[bofs + 16 * ot + r] = bsum;
}

// Stage 4: grad = sum of the split partials (grad holds the scaled gradient).
__global__ __launch_bounds__(kBlock) void k_train_reduce(const float* __restrict__ partial, uint32_t splits,
                                                         float* __restrict__ grad) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= (uint32_t)kParams) return;
    float acc = 0.0f;
    for (uint32_t sp = 0; sp < splits; ++sp) acc += partial[(size_t)sp * kParams + i];
    grad[i] = acc;
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

size_t eval_lds_bytes() { return sizeof(FwdWeights) + sizeof(WaveStage) * (kBlock / 64); }
size_t fwd_lds_bytes() {
    return sizeof(FwdWeights) + sizeof(WaveStage) * (kBlock / 64) + sizeof(float) * 64 * kOut * (kBlock / 64);
}

// The inference and forward kernels use more than the default 64 KiB of
// dynamic LDS (gfx950 has 160 KiB per CU).
spt_status allow_lds(spt_ctx* ctx) {
    static bool done = false;
    if (done) return SPT_OK;
    spt_status s = check_hip(ctx,
                             hipFuncSetAttribute(reinterpret_cast<const void*>(&k_train_fwd),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)fwd_lds_bytes()),
                             "forward LDS size");
    if (s == SPT_OK)
        s = check_hip(ctx,
                      hipFuncSetAttribute(reinterpret_cast<const void*>(&k_guide_eval),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)eval_lds_bytes()),
                      "eval LDS size");
    done = (s == SPT_OK);
    return s;
}

// Scratch for n16 records (grown on demand; the dy rows 40..47 stay zero).
spt_status train_scratch(spt_ctx* ctx, uint32_t n16, uint32_t splits, TrainScratch& ts) {
    GuideState& gs = ctx->guide;
    const uint32_t rows = (n16 + 63) & ~63u;
    const size_t per_row = 2 * kEnc + 3 * 2 * kHid + 4 * kOutP + 3 * 2 * kHid;
    const size_t need = (size_t)rows * per_row + (size_t)splits * kParams * sizeof(float) + 1024;
    if (rows > gs.scratch_rows || need > gs.scratch_bytes) {
        spt_status s = stream_wait(ctx, "guide scratch");
        if (s != SPT_OK) return s;
        if (gs.d_scratch) (void)hipFree(gs.d_scratch);
        gs.d_scratch = nullptr;
        gs.scratch_bytes = 0;
        gs.scratch_rows = 0;
        s = check_hip(ctx, hipMalloc(&gs.d_scratch, need), "guide scratch");
        if (s != SPT_OK) return s;
        s = check_hip(ctx, hipMemsetAsync(gs.d_scratch, 0, need, ctx->stream), "guide scratch clear");
        if (s != SPT_OK) return s;
        gs.scratch_bytes = need;
        gs.scratch_rows = rows;
    }
    const uint32_t R = gs.scratch_rows;
    char* p = gs.d_scratch;
    auto take = [&](size_t bytes) {
        char* q = p;
        p += (bytes + 255) & ~size_t(255);
        return q;
    };
    ts.xT = reinterpret_cast<unsigned short*>(take((size_t)R * kEnc * 2));
    ts.h1T = reinterpret_cast<unsigned short*>(take((size_t)R * kHid * 2));
    ts.h2T = reinterpret_cast<unsigned short*>(take((size_t)R * kHid * 2));
    ts.h3T = reinterpret_cast<unsigned short*>(take((size_t)R * kHid * 2));
    ts.dyT = reinterpret_cast<float*>(take((size_t)R * kOutP * 4));
    ts.dz3T = reinterpret_cast<unsigned short*>(take((size_t)R * kHid * 2));
    ts.dz2T = reinterpret_cast<unsigned short*>(take((size_t)R * kHid * 2));
    ts.dz1T = reinterpret_cast<unsigned short*>(take((size_t)R * kHid * 2));
    ts.partial = reinterpret_cast<float*>(take((size_t)splits * kParams * 4));
    ts.rows = R;
    if ((size_t)(p - gs.d_scratch) > gs.scratch_bytes) return set_error(ctx, SPT_ERR_NOMEM, "guide scratch layout");
    return SPT_OK;
}

// Queues the gradient of one step (ctl->wsum and ctl->loss, grad = scaled
// gradient) on the context stream.
spt_status queue_gradient(spt_ctx* ctx, const float* d_rec, uint32_t n, uint32_t rs, hipEvent_t e0, hipEvent_t e1) {
    GuideState& gs = ctx->guide;
    hipStream_t st = ctx->stream;
    GuideCtl* ctl = gs.ctl();
    const uint32_t n16 = (n + 15) & ~15u;
    const uint32_t span = std::max<uint32_t>(1024, (((n16 + 511) / 512) + 15) & ~15u);
    const uint32_t splits = (n16 + span - 1) / span;
    TrainScratch ts;
    spt_status s = train_scratch(ctx, n16, splits, ts);
    if (s != SPT_OK) return s;
    hipLaunchKernelGGL(k_guide_wsum, dim3(256), dim3(kBlock), 0, st, d_rec, n, rs, &ctl->wsum);
    if (e0) (void)hipEventRecord(e0, st);
    hipLaunchKernelGGL(k_train_fwd, dim3((n16 + kBlock - 1) / kBlock), dim3(kBlock), fwd_lds_bytes(), st, gs.d_theta,
                       gs.d_wbf, d_rec, n, rs, n16, ctl, guide_box(ctx), ts);
    hipLaunchKernelGGL(k_train_bwd, dim3(std::min<uint32_t>((n16 + 63) / 64, (uint32_t)ctx->num_cus * 4)), dim3(kBlock),
                       sizeof(SharedWeights), st, gs.d_theta, gs.d_wbf, n16, ts);
    hipLaunchKernelGGL(k_train_wgrad, dim3(4 * splits), dim3(kBlock), 0, st, ts, n16, span);
    hipLaunchKernelGGL(k_train_reduce, dim3((kParams + kBlock - 1) / kBlock), dim3(kBlock), 0, st, ts.partial, splits,
                       gs.d_grad);
    if (e1) (void)hipEventRecord(e1, st);
    return check_hip(ctx, hipGetLastError(), "guide gradient launch");
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

spt_status guide_train_launch(spt_ctx* ctx, const float* d_rec, uint32_t n, float* loss_out, uint32_t rs) {
    GuideState& gs = ctx->guide;
    spt_status ls = allow_lds(ctx);
    if (ls != SPT_OK) return ls;
    auto t0 = std::chrono::steady_clock::now();
    hipStream_t st = ctx->stream;
    GuideCtl* ctl = gs.ctl();
    hipLaunchKernelGGL(k_guide_begin, dim3(1), dim3(64), 0, st, ctl);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!gs.event_pair(e0, e1)) e0 = e1 = nullptr;
    spt_status qs = queue_gradient(ctx, d_rec, n, rs, e0, e1);
    if (qs != SPT_OK) return qs;
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
    spt_status s = queue_gradient(ctx, d_rec, n, 1u, nullptr, nullptr);
    if (s != SPT_OK) return s;
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
        !(c.loss_scale > 0.0f) || c.steps_per_pass > 64 || (uint64_t)c.batch * c.steps_per_pass >= (1ull << 31) ||
        (c.flags & ~(SPT_GUIDE_COATED | SPT_GUIDE_PRODUCT)))
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
    gs.flags = c.flags;
    gs.train_every = c.train_every;
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
    return guide_train_launch(ctx, d_records, n, loss, 1u);
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
    // one step per batch (a remainder shorter than half a batch is spread over
    // the others); step k takes records k, k + steps, k + 2 steps, ...: the
    // records were appended in shading order, grouped by material and region,
    // and interleaved batches sample the whole pass
    const uint32_t steps = std::max<uint32_t>(1, n / gs.batch + ((n % gs.batch) >= gs.batch / 2 ? 1u : 0u));
    for (uint32_t k = 0; k < steps && s == SPT_OK; ++k) {
        const uint32_t len = (n - k + steps - 1) / steps;
        s = guide_train_launch(ctx, gs.d_records + (size_t)k * kRecord, len, k + 1 == steps ? loss : nullptr, steps);
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
