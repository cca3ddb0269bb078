// spt_guide.h - learned path guide on the device: input encoding, a 32-64-64-
// 64-40 MLP evaluated 16 rows per wave with gfx950 bf16 MFMA
// (v_mfma_f32_16x16x32_bf16, fp32 accumulation), the 8-lobe von Mises-Fisher
// mixture it outputs (density,
// sampling, KL-loss gradient) and the fused backward pass.
//
// Parameter vector (fp32 master, 13,552 floats): W1 64x32 | b1 64 | W2 64x64 |
// b2 64 | W3 64x64 | b3 64 | W4 48x64 (rows 40..47 zero) | b4 48. Weights are
// row-major (out, in). Methods: vMF sampling after W. Jakob (2012); KL training
// of a sampling density on Monte Carlo targets after Mueller et al. (2019).
#pragma once

#include <hip/hip_fp16.h>

#include "spt_device.h"

namespace spt {
namespace guide {

constexpr int kEnc = 32, kHid = 64, kOut = 40, kOutP = 48, kLobes = 8;
constexpr int kW1 = 0, kB1 = kW1 + kHid * kEnc, kW2 = kB1 + kHid, kB2 = kW2 + kHid * kHid, kW3 = kB2 + kHid,
              kB3 = kW3 + kHid * kHid, kW4 = kB3 + kHid, kB4 = kW4 + kOutP * kHid, kParams = kB4 + kOutP;
static_assert(kParams == 13552, "parameter layout");
constexpr int kRecord = 24;  // floats per training record: p n wo wi W g rad_after thr_next pad
constexpr float kKappaMin = 1e-2f, kKappaMax = 1e4f;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef short bf16x8s __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// round-to-nearest-even float -> bfloat16 bits, and back
__device__ inline unsigned short f2bf(float x) {
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ inline float bf2f(unsigned short b) { return __uint_as_float((uint32_t)b << 16); }

// Input encoding (32): x = (p - bmin) / bext clamped to [0,1]; sin/cos of
// 2 pi 2^k x_c for k = 0..2; n; wo; one-blob (4 Gaussian bins, sigma 1/4) of
// (n.wo + 1)/2; constant 1.
__device__ inline void encode(V3 p, V3 n, V3 wo, const float* bmin, const float* bext, float f[kEnc]) {
    const float x[3] = {fminf(fmaxf((p.x - bmin[0]) / bext[0], 0.0f), 1.0f),
                        fminf(fmaxf((p.y - bmin[1]) / bext[1], 0.0f), 1.0f),
                        fminf(fmaxf((p.z - bmin[2]) / bext[2], 0.0f), 1.0f)};
    f[0] = x[0];
    f[1] = x[1];
    f[2] = x[2];
    int o = 3;
    for (int k = 0; k < 3; ++k) {
        for (int c = 0; c < 3; ++c) {
            // v_sin / v_cos take the angle in revolutions: sin(2 pi 2^k x)
            const float a = (float)(1 << k) * x[c];
            f[o++] = __builtin_amdgcn_sinf(a);
            f[o++] = __builtin_amdgcn_cosf(a);
        }
    }
    f[21] = n.x;
    f[22] = n.y;
    f[23] = n.z;
    f[24] = wo.x;
    f[25] = wo.y;
    f[26] = wo.z;
    const float c = 0.5f * (dot(n, wo) + 1.0f);
    for (int b = 0; b < 4; ++b) {
        const float dc = c - (b + 0.5f) * 0.25f;
        f[27 + b] = __expf(-dc * dc * 8.0f);  // sigma 1/4: exp(-dc^2 / (2 sigma^2))
    }
    f[31] = 1.0f;
}

// Block-shared weights: bf16 copies of W1..W4 and fp32 biases.
struct SharedWeights {
    unsigned short w1[kHid * kEnc];
    unsigned short w2[kHid * kHid];
    unsigned short w3[kHid * kHid];
    unsigned short w4[kOutP * kHid];
    float b1[kHid], b2[kHid], b3[kHid], b4[kOutP];
};

__device__ inline void load_weights(SharedWeights& s, const unsigned short* wbf, const float* theta) {
    for (int i = threadIdx.x; i < kHid * kEnc; i += blockDim.x) s.w1[i] = wbf[kW1 + i];
    for (int i = threadIdx.x; i < kHid * kHid; i += blockDim.x) {
        s.w2[i] = wbf[kW2 + i];
        s.w3[i] = wbf[kW3 + i];
    }
    for (int i = threadIdx.x; i < kOutP * kHid; i += blockDim.x) s.w4[i] = wbf[kW4 + i];
    for (int i = threadIdx.x; i < kHid; i += blockDim.x) {
        s.b1[i] = theta[kB1 + i];
        s.b2[i] = theta[kB2 + i];
        s.b3[i] = theta[kB3 + i];
    }
    for (int i = threadIdx.x; i < kOutP; i += blockDim.x) s.b4[i] = theta[kB4 + i];
}

// 4 consecutive bf16 of an LDS row as an MFMA operand
__device__ inline bf16x4 ld4(const unsigned short* p) {
    const uint2 v = *reinterpret_cast<const uint2*>(p);
    bf16x4 r;
    r[0] = (short)(v.x & 0xFFFFu);
    r[1] = (short)(v.x >> 16);
    r[2] = (short)(v.y & 0xFFFFu);
    r[3] = (short)(v.y >> 16);
    return r;
}
// 4 bf16 with a stride
__device__ inline bf16x4 ld4s(const unsigned short* p, int stride) {
    bf16x4 r;
    r[0] = (short)p[0];
    r[1] = (short)p[stride];
    r[2] = (short)p[2 * stride];
    r[3] = (short)p[3 * stride];
    return r;
}

// Operand layout of the data (wave64), lane l, r = l % 16, g = l / 16, per
// 16-wide k-tile:
//   A (16 rows x 16 k):  A[r][4g + j]         j = 0..3
//   B (16 k x 16 cols):  B[4g + j][r]
//   D (16 x 16, f32):    D[4g + i][r]         i = 0..3
// v_mfma_f32_16x16x32_bf16 takes 8 bf16 of A and of B per lane: element e of
// lane (r, g) multiplies element e of lane (c, g) for every e, and all 32
// products per (row, col) are summed. k is only a summation index, so two
// k-tiles are fed as one by concatenating the lane's 4 values of tile k and
// of tile k + 1 in A and in B alike (mfma_k32).
// ---------------------------------------------------------------------------
// Register-resident forward pass ("column" form): the layer is computed as
// out^T = W * in^T, weights as the MFMA A operand (from LDS) and 16 samples as
// the B operand. A lane (r, g) then holds features 16 t + 4 g + i (i = 0..3)
// of sample r in each output tile t, which is exactly the B-operand layout of
// the next layer's k-tile t: activations never leave registers.
// ---------------------------------------------------------------------------
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef short i16x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

// Two floats as packed bfloat16 (a in the low half) by gfx950's
// v_cvt_pk_bf16_f32, round to nearest even: the same bits as f2bf for every
// non-NaN float.
__device__ inline uint32_t pack_bf2(float a, float b) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){a, b}, bf16x2));
}
// ReLU of two packed bfloat16: a negative value (sign bit set) is a negative
// int16, so an int16 max with 0 gives +0 (the bits of f2bf(fmaxf(x, 0)) for
// every non-NaN x).
__device__ inline uint32_t relu_bf2(uint32_t v) {
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(i16x2, v), (i16x2){0, 0}));
}

__device__ inline bf16x4 pack_relu(f32x4 v) {
    return __builtin_bit_cast(bf16x4, (u32x2){relu_bf2(pack_bf2(v[0], v[1])), relu_bf2(pack_bf2(v[2], v[3]))});
}

__device__ inline f32x4 bias4(const float* b) { return f32x4{b[0], b[1], b[2], b[3]}; }

// D += A B over two 16-wide k-tiles (a0, b0: tile k; a1, b1: tile k + 1).
__device__ inline f32x4 mfma_k32(bf16x4 a0, bf16x4 a1, bf16x4 b0, bf16x4 b1, f32x4 c) {
    const bf16x8s a = {a0[0], a0[1], a0[2], a0[3], a1[0], a1[1], a1[2], a1[3]};
    const bf16x8s b = {b0[0], b0[1], b0[2], b0[3], b1[0], b1[1], b1[2], b1[3]};
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                   0, 0, 0);
}
// D += A B over one 16-wide k-tile (the odd tile of a 48-wide reduction).
__device__ inline f32x4 mfma_k16(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Inference weights in MFMA A-operand order. For a layer of IN inputs, output
// tile t and k-tile pair p (tiles 2p and 2p + 1) one 1-KB block holds, for
// lane r + 16 g, the 8 bf16 W[16 t + r][32 p + 4 g + 0..3] and
// W[16 t + r][32 p + 16 + 4 g + 0..3]: each MFMA's weight operand is one
// 16-B LDS read per lane at consecutive addresses (the row-major layout
// puts the 16 rows of a read 64 or 128 B apart, on the same banks).
struct FwdWeights {
    uint4 w1[4 * (kEnc / 32) * 64];
    uint4 w2[4 * (kHid / 32) * 64];
    uint4 w3[4 * (kHid / 32) * 64];
    uint4 w4[(kOutP / 16) * (kHid / 32) * 64];
    float b1[kHid], b2[kHid], b3[kHid], b4[kOutP];
};

template <int IN>
__device__ inline void swizzle_layer(uint4* dst, const unsigned short* w, int tiles) {
    for (int i = threadIdx.x; i < tiles * (IN / 32) * 64; i += blockDim.x) {
        const int l = i & 63, p = (i >> 6) % (IN / 32), t = (i >> 6) / (IN / 32);
        const unsigned short* row = w + (16 * t + (l & 15)) * IN + 32 * p + 4 * (l >> 4);
        dst[i] = make_uint4((uint32_t)row[0] | ((uint32_t)row[1] << 16), (uint32_t)row[2] | ((uint32_t)row[3] << 16),
                            (uint32_t)row[16] | ((uint32_t)row[17] << 16),
                            (uint32_t)row[18] | ((uint32_t)row[19] << 16));
    }
}

__device__ inline void load_weights(FwdWeights& s, const unsigned short* wbf, const float* theta) {
    swizzle_layer<kEnc>(s.w1, wbf + kW1, kHid / 16);
    swizzle_layer<kHid>(s.w2, wbf + kW2, kHid / 16);
    swizzle_layer<kHid>(s.w3, wbf + kW3, kHid / 16);
    swizzle_layer<kHid>(s.w4, wbf + kW4, kOutP / 16);
    for (int i = threadIdx.x; i < kHid; i += blockDim.x) {
        s.b1[i] = theta[kB1 + i];
        s.b2[i] = theta[kB2 + i];
        s.b3[i] = theta[kB3 + i];
    }
    for (int i = threadIdx.x; i < kOutP; i += blockDim.x) s.b4[i] = theta[kB4 + i];
}

// D += A B with A the lane's 8 weights of block blk, B two k-tiles of data.
__device__ inline f32x4 mfma_w(const uint4* blk, bf16x4 b0, bf16x4 b1, f32x4 c) {
    const uint4 v = blk[__lane_id()];
    const bf16x8s b = {b0[0], b0[1], b0[2], b0[3], b1[0], b1[1], b1[2], b1[3]};
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, v), __builtin_bit_cast(bf16x8, b), c,
                                                   0, 0, 0);
}

// x[kt]: features 16 kt + 4 g + j of sample r (bf16); y[t]: raw outputs
// 16 t + 4 g + i of sample r (f32), t = 0..2 (outputs 40..47 are padding).
__device__ inline void mlp_cols_fwd(const FwdWeights& s, const bf16x4 x[2], f32x4 y[3]) {
    const int g = __lane_id() >> 4;
    bf16x4 ha[4], hb[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) ha[t] = pack_relu(mfma_w(s.w1 + 64 * t, x[0], x[1], bias4(s.b1 + 16 * t + 4 * g)));
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        f32x4 acc = bias4(s.b2 + 16 * t + 4 * g);
#pragma unroll
        for (int p = 0; p < 2; ++p) acc = mfma_w(s.w2 + 64 * (2 * t + p), ha[2 * p], ha[2 * p + 1], acc);
        hb[t] = pack_relu(acc);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        f32x4 acc = bias4(s.b3 + 16 * t + 4 * g);
#pragma unroll
        for (int p = 0; p < 2; ++p) acc = mfma_w(s.w3 + 64 * (2 * t + p), hb[2 * p], hb[2 * p + 1], acc);
        ha[t] = pack_relu(acc);
    }
#pragma unroll
    for (int t = 0; t < 3; ++t) {
        f32x4 acc = bias4(s.b4 + 16 * t + 4 * g);
#pragma unroll
        for (int p = 0; p < 2; ++p) acc = mfma_w(s.w4 + 64 * (2 * t + p), ha[2 * p], ha[2 * p + 1], acc);
        y[t] = acc;
    }
}

// Per-wave staging of 64 encoded samples (bf16 rows of 32 features, padded
// to 36 so the B-operand reads of 16 rows spread over the LDS banks).
constexpr int kStageRow = kEnc + 4;
struct WaveStage {
    unsigned short x[64 * kStageRow];
};

// Lane l stores its sample's 32 features (bf16) as row l of the stage.
__device__ inline void stage_row(WaveStage& st, const float f[kEnc]) {
    uint2* row = reinterpret_cast<uint2*>(st.x + __lane_id() * kStageRow);
#pragma unroll
    for (int i = 0; i < kEnc / 4; ++i)
        row[i] = make_uint2(pack_bf2(f[4 * i], f[4 * i + 1]), pack_bf2(f[4 * i + 2], f[4 * i + 3]));
}

// Writes the 40 raw outputs of one sample from the column-form result.
__device__ inline void store_outputs(float* dst, const f32x4 y[3]) {
    const int g = __lane_id() >> 4;
    *reinterpret_cast<float4*>(dst + 4 * g) = make_float4(y[0][0], y[0][1], y[0][2], y[0][3]);
    *reinterpret_cast<float4*>(dst + 16 + 4 * g) = make_float4(y[1][0], y[1][1], y[1][2], y[1][3]);
    if (g < 2) *reinterpret_cast<float4*>(dst + 32 + 4 * g) = make_float4(y[2][0], y[2][1], y[2][2], y[2][3]);
}

// The same outputs as IEEE half floats (round to nearest), 80 B per sample:
// the per-path copy a guided render keeps between inference and shading.
__device__ inline uint32_t pack_h2(float a, float b) {
    return (uint32_t)__half_as_ushort(__float2half_rn(a)) | ((uint32_t)__half_as_ushort(__float2half_rn(b)) << 16);
}
__device__ inline void store_outputs_h(unsigned short* dst, const f32x4 y[3]) {
    const int g = __lane_id() >> 4;
    *reinterpret_cast<uint2*>(dst + 4 * g) = make_uint2(pack_h2(y[0][0], y[0][1]), pack_h2(y[0][2], y[0][3]));
    *reinterpret_cast<uint2*>(dst + 16 + 4 * g) = make_uint2(pack_h2(y[1][0], y[1][1]), pack_h2(y[1][2], y[1][3]));
    if (g < 2)
        *reinterpret_cast<uint2*>(dst + 32 + 4 * g) = make_uint2(pack_h2(y[2][0], y[2][1]), pack_h2(y[2][2], y[2][3]));
}

// Output k of a half-float row as float.
struct HalfRow {
    const unsigned short* p;
    __device__ float operator[](int k) const { return __half2float(__ushort_as_half(p[k])); }
};

// Forward of the 64 samples staged by this wave: for sample group q (16
// samples 16 q .. 16 q + 15) calls out(q, y) with lane (r, g) holding outputs
// 16 t + 4 g + i of sample 16 q + r.
// The weights are read from LDS again for every group: the offset z is
// hidden from the compiler by an empty asm statement, which keeps it from
// holding all the weights in registers across the groups (244 registers and
// two waves per SIMD against 93 and five; the same results, and guided
// passes 1-3% faster while other wavefronts share the GPU).
template <typename Out>
__device__ inline void mlp_stage_fwd(const FwdWeights& s, const WaveStage& st, Out out) {
    const int l = __lane_id(), r = l & 15, g = l >> 4;
#pragma unroll 1
    for (int q = 0; q < 4; ++q) {
        const unsigned short* row = st.x + (16 * q + r) * kStageRow;
        const bf16x4 x[2] = {ld4(row + 4 * g), ld4(row + 16 + 4 * g)};
        f32x4 y[3];
        uint32_t z = 0;
        asm volatile("" : "+v"(z));
        mlp_cols_fwd(*reinterpret_cast<const FwdWeights*>(reinterpret_cast<const char*>(&s) + z), x, y);
        out(q, y);
    }
}


// Mixture parameters of one raw output row.
struct Lobe {
    V3 mu;
    float kappa, logw, norm, s;
    bool inside;  // log kappa inside its clamp range (gradient flows)
};

// y: the 40 raw outputs (a float pointer, or a HalfRow).
template <typename Row>
__device__ inline void mixture_params(const Row& y, Lobe L[kLobes]) {
    float amax = -3.0e38f;
    for (int i = 0; i < kLobes; ++i) amax = fmaxf(amax, y[5 * i + 4]);
    float sum = 0.0f;
    for (int i = 0; i < kLobes; ++i) sum += expf(y[5 * i + 4] - amax);
    const float lse = amax + logf(sum);
    for (int i = 0; i < kLobes; ++i) {
        const V3 raw{y[5 * i], y[5 * i + 1], y[5 * i + 2]};
        const float nr = sqrtf(dot(raw, raw) + 1e-12f);
        L[i].mu = raw * (1.0f / nr);
        L[i].norm = nr;
        const float s = y[5 * i + 3];
        const float sc = fminf(fmaxf(s, logf(kKappaMin)), logf(kKappaMax));
        L[i].inside = (s == sc);
        L[i].s = sc;
        L[i].kappa = expf(sc);
        L[i].logw = y[5 * i + 4] - lse;
    }
}

// log of kappa / (2 pi (1 - e^{-2 kappa})) exp(kappa (mu.d - 1))
__device__ inline float log_vmf(const Lobe& L, V3 d) {
    return logf(L.kappa) - 1.8378770664093453f - logf(-expm1f(-2.0f * L.kappa)) + L.kappa * (dot(L.mu, d) - 1.0f);
}

// log of the vMF normalisation kappa / (2 pi (1 - e^{-2 kappa}))
__device__ inline float log_vmf_norm(float k) { return logf(k) - 1.8378770664093453f - logf(-expm1f(-2.0f * k)); }

// Product of the mixture with one vMF lobe (mb, kb), renormalised (product
// importance sampling with vMF lobes, Herholz et al. 2016):
// vMF(mu_i, k_i) vMF(mb, kb) = C_i C_b / C' e^{k' - k_i - kb} vMF(mu', k')
// with k' mu' = k_i mu_i + kb mb. The result is again a mixture, so sampling
// and density stay exact.
__device__ inline void mixture_product(Lobe L[kLobes], V3 mb, float kb) {
    const float lcb = log_vmf_norm(kb);
    float lw[kLobes];
    float m = -3.0e38f;
    for (int i = 0; i < kLobes; ++i) {
        const V3 v = L[i].mu * L[i].kappa + mb * kb;
        const float nv = sqrtf(dot(v, v));
        const float kp = fmaxf(nv, kKappaMin);
        lw[i] = L[i].logw + log_vmf_norm(L[i].kappa) + lcb - log_vmf_norm(kp) + (nv - L[i].kappa - kb);
        L[i].mu = (nv > 1e-6f) ? v * (1.0f / nv) : mb;
        L[i].kappa = kp;
        m = fmaxf(m, lw[i]);
    }
    float sum = 0.0f;
    for (int i = 0; i < kLobes; ++i) sum += expf(lw[i] - m);
    const float lse = m + logf(sum);
    for (int i = 0; i < kLobes; ++i) L[i].logw = lw[i] - lse;
}

__device__ inline float mixture_pdf(const Lobe L[kLobes], V3 d) {
    float q = 0.0f;
    for (int i = 0; i < kLobes; ++i) q += expf(L[i].logw + log_vmf(L[i], d));
    return q;
}

// Orthonormal basis around n (Duff et al. 2017).
__device__ inline void basis(V3 n, V3& t, V3& b) {
    const float s = n.z >= 0.0f ? 1.0f : -1.0f;
    const float a = -1.0f / (s + n.z);
    const float bb = n.x * n.y * a;
    t = V3{1.0f + s * n.x * n.x * a, s * bb, -s * n.x};
    b = V3{bb, s + n.y * n.y * a, -n.y};
}

// Sample the mixture: lobe by weight (u0), then Jakob's stable vMF sampler.
__device__ inline V3 mixture_sample(const Lobe L[kLobes], float u0, float u1, float u2) {
    int k = kLobes - 1;
    float c = 0.0f;
    for (int i = 0; i < kLobes; ++i) {
        c += expf(L[i].logw);
        if (u0 < c) {
            k = i;
            break;
        }
    }
    const float kap = L[k].kappa;
    // for kappa above ~44, e^{-2 kappa} underflows and u1 = 0 gives log(0):
    // the exact limit is w = -1, so clamp (a NaN direction would make the
    // next ray visit every node of the tree)
    const float w = fmaxf(1.0f + logf(u1 + (1.0f - u1) * expf(-2.0f * kap)) / kap, -1.0f);
    const float st = sqrtf(fmaxf(0.0f, 1.0f - w * w));
    const float ph = 2.0f * kPi * u2;
    V3 t, b;
    basis(L[k].mu, t, b);
    return t * (st * cosf(ph)) + b * (st * sinf(ph)) + L[k].mu * w;
}

// Gradient of -W log q(d) w.r.t. the 40 raw outputs (dy[40..47] = 0).
__device__ inline float loss_grad(const float* y, V3 d, float W, float* dy) {
    Lobe L[kLobes];
    mixture_params(y, L);
    float lv[kLobes];
    float m = -3.0e38f;
    for (int i = 0; i < kLobes; ++i) {
        lv[i] = L[i].logw + log_vmf(L[i], d);
        m = fmaxf(m, lv[i]);
    }
    float sum = 0.0f;
    for (int i = 0; i < kLobes; ++i) sum += expf(lv[i] - m);
    const float logq = m + logf(sum);
    for (int i = 0; i < kLobes; ++i) {
        const float r = expf(lv[i] - logq);
        const float w = expf(L[i].logw);
        const float c = dot(L[i].mu, d);
        dy[5 * i + 4] = -W * (r - w);
        const float kap = L[i].kappa;
        const float dlv = 1.0f - 2.0f * kap / expm1f(2.0f * kap) + kap * (c - 1.0f);
        dy[5 * i + 3] = L[i].inside ? -W * r * dlv : 0.0f;
        const V3 gmu = d * (-W * r * kap);
        const V3 proj = (gmu - L[i].mu * dot(gmu, L[i].mu)) * (1.0f / L[i].norm);
        dy[5 * i + 0] = proj.x;
        dy[5 * i + 1] = proj.y;
        dy[5 * i + 2] = proj.z;
    }
    for (int i = kOut; i < kOutP; ++i) dy[i] = 0.0f;
    return -W * logq;
}

}  // namespace guide
}  // namespace spt
