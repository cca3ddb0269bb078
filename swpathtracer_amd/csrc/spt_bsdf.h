// spt_bsdf.h - device BSDF layers and the layered random walk
// (materials/, bsdf_samplers/ of the reference), plus the GGX conductor with a
// complex-IOR Fresnel term that the reference lacks.
#pragma once

#include "spt_device.h"

namespace spt {

// Draw source of one path: counter-based (include/spt.h) or a replay stream.
struct Draws {
    uint64_t key;
    const float* stream;  // replay: this path's row, nullptr for counter mode
    uint32_t stride;
    uint32_t k;
    bool overflow;
    __device__ float next() {
        if (stream) {
            if (k >= stride) {
                overflow = true;
                return 0.0f;
            }
            return stream[k++];
        }
        const float u = rng_u(key, k);
        ++k;
        return u;
    }
};

struct BSample {
    V3 wi;
    float pdf;
};

// FresnelSchlick with reflectance at normal incidence (fresnel_schlick.cpp:11-18)
__device__ inline V3 schlick_r0(const float* r0, float c) {
    float c5 = 1.0f - c;
    c5 = c5 * c5 * c5 * c5 * c5;
    return V3{r0[0] + (1.0f - r0[0]) * c5, r0[1] + (1.0f - r0[1]) * c5, r0[2] + (1.0f - r0[2]) * c5};
}

// FresnelSchlick with two IORs (fresnel_schlick.cpp:20-38)
__device__ inline float schlick_ior(float ior_i_, float ior_t_, float c) {
    float ior_i = ior_i_, ior_t = ior_t_;
    if (c < 0.0f) {
        ior_i = ior_t_;
        ior_t = ior_i_;
        const float eta = ior_i / ior_t;
        const float a = 1 - eta * eta * (1.0f - c * c);
        c = sqrt_rn(a);
    }
    float r0 = (ior_i - ior_t) / (ior_i + ior_t);
    r0 = r0 * r0;
    float c5 = 1.0f - c;
    c5 = c5 * c5 * c5 * c5 * c5;
    return r0 + (1.0f - r0) * c5;
}

// Exact conductor Fresnel for complex IOR eta + i k from a medium of IOR 1
// (textbook form, Pharr/Jakob/Humphreys PBR 3rd ed. sec. 8.2.1).
__device__ inline float fresnel_conductor(float ci, float eta, float k) {
    ci = gmin(gmax(ci, -1.0f), 1.0f);
    const float c2 = ci * ci, s2 = 1.0f - c2;
    const float e2 = eta * eta, k2 = k * k;
    const float t0 = e2 - k2 - s2;
    const float a2b2 = sqrt_rn(t0 * t0 + 4.0f * e2 * k2);
    const float t1 = a2b2 + c2;
    const float a = sqrt_rn(0.5f * (a2b2 + t0));
    const float t2 = 2.0f * ci * a;
    const float rs = (t1 - t2) / (t1 + t2);
    const float t3 = c2 * a2b2 + s2 * s2;
    const float t4 = t2 * s2;
    const float rp = rs * (t3 - t4) / (t3 + t4);
    return 0.5f * (rp + rs);
}
__device__ inline V3 fresnel_conductor3(float ci, const spt_layer& l) {
    return V3{fresnel_conductor(ci, l.eta[0], l.k[0]), fresnel_conductor(ci, l.eta[1], l.k[1]),
              fresnel_conductor(ci, l.eta[2], l.k[2])};
}

// GGX D and Smith G1 (Walter, Marschner, Li, Torrance 2007)
__device__ inline float ggx_d(float ch, float alpha) {
    if (ch <= 0.0f) return 0.0f;
    const float a2 = alpha * alpha, c2 = ch * ch;
    const float den = c2 * (a2 - 1.0f) + 1.0f;
    return a2 / (kPi * den * den);
}
__device__ inline float ggx_g1(float cv, float alpha) {
    if (cv <= 0.0f) return 0.0f;
    const float a2 = alpha * alpha, c2 = cv * cv;
    return 2.0f * cv / (cv + sqrt_rn(a2 + (1.0f - a2) * c2));
}

__device__ inline bool layer_opaque(const spt_layer& l) { return l.type != SPT_BSDF_SMOOTH_DIELECTRIC; }

__device__ inline V3 layer_f(const spt_layer& l, V3 wo, V3 wi) {
    switch (l.type) {
        case SPT_BSDF_LAMBERTIAN: {  // lambertian.cpp:9-16 (brdf_ = kd / pi)
            if (wi.y > 0.0f) return V3{l.kd[0] / kPi, l.kd[1] / kPi, l.kd[2] / kPi};
            return V3{0.0f, 0.0f, 0.0f};
        }
        case SPT_BSDF_COOK_TORRANCE: {  // cook_torrance.cpp:9-44 (Beckmann, V-cavity G, Schlick)
            if (wi.y <= 0.0f) return V3{0.0f, 0.0f, 0.0f};
            const V3 h = normalize(wo + wi);
            const float nv = gmax(0.0f, wo.y), nl = gmax(0.0f, wi.y);
            const float vh = gmax(0.0f, dot(wo, h)), nh = gmax(0.0f, h.y);
            const float m = l.roughness;
            if ((nv == 0.0f) || (nl == 0.0f) || (vh == 0.0f) || (nh == 0.0f) || (m == 0.0f))
                return V3{0.0f, 0.0f, 0.0f};
            const float lh = gmax(0.0f, dot(wi, h));
            const float nh2 = nh * nh, m2 = m * m;
            const float d1 = 1.0f / (kPi * m2 * nh2 * nh2);
            const float d2 = (nh2 - 1.0f) / (m2 * nh2);
            const float d = d1 * f_exp(d2);
            const float g1 = 2.0f * nh / vh;
            const float g = gmin(1.0f, gmin(g1 * nv, g1 * nl));
            const V3 F = schlick_r0(l.r0, lh);
            return (F * d * g) / (4.0f * nv * nl);
        }
        case SPT_BSDF_SMOOTH_CONDUCTOR:  // smooth_conductor.cpp:6-14
            if (wi.y > 0.0f) return schlick_r0(l.r0, wi.y) / wi.y;
            return V3{0.0f, 0.0f, 0.0f};
        case SPT_BSDF_SMOOTH_DIELECTRIC:  // smooth_dielectric.cpp:15-21
            if (wi.y == 0.0f) return V3{0.0f, 0.0f, 0.0f};
            return V3{1.0f / wi.y, 1.0f / wi.y, 1.0f / wi.y};
        case SPT_BSDF_GGX_CONDUCTOR: {
            if (wi.y <= 0.0f || wo.y <= 0.0f) return V3{0.0f, 0.0f, 0.0f};
            const V3 h = normalize(wo + wi);
            const float D = ggx_d(h.y, l.roughness);
            const float G = ggx_g1(wo.y, l.roughness) * ggx_g1(wi.y, l.roughness);
            const V3 F = fresnel_conductor3(dot(wi, h), l);
            return F * (D * G / (4.0f * wo.y * wi.y));
        }
    }
    return V3{0.0f, 0.0f, 0.0f};
}

__device__ inline V3 layer_reflectance(const spt_layer& l, V3 wo, V3 wi) {
    switch (l.type) {
        case SPT_BSDF_LAMBERTIAN:
            return V3{l.kd[0], l.kd[1], l.kd[2]};
        case SPT_BSDF_COOK_TORRANCE: {  // cook_torrance.cpp:46-51
            const V3 h = normalize(wo + wi);
            return schlick_r0(l.r0, gmax(0.0f, dot(wi, h)));
        }
        case SPT_BSDF_SMOOTH_CONDUCTOR:
            return schlick_r0(l.r0, wi.y);
        case SPT_BSDF_SMOOTH_DIELECTRIC:
            return V3{1.0f, 1.0f, 1.0f};
        case SPT_BSDF_GGX_CONDUCTOR: {
            const V3 h = normalize(wo + wi);
            return fresnel_conductor3(gmax(0.0f, dot(wi, h)), l);
        }
    }
    return V3{0.0f, 0.0f, 0.0f};
}

// UniformBSDFSampler (uniform_bsdf_sampler.cpp:5-14)
__device__ inline BSample sample_uniform(Draws& r) {
    const float r1 = r.next();
    const float r2 = r.next();
    const float phi = 2.0f * kPi * r2;
    const float s = sqrt_rn(1.0f - r1 * r1);
    float sp, cp;
    f_sincos(phi, sp, cp);
    return BSample{V3{cp * s, r1, sp * s}, 1.0f / (2.0f * kPi)};
}

// CookTorranceBSDFSampler::getProbability (cook_torrance_bsdf_sampler.cpp:22-42)
__device__ inline float beckmann_pdf(const spt_layer& l, V3 wo, V3 wi) {
    if (wi.y <= 0.0f) return 1.0f;
    const V3 h = normalize(wo + wi);
    const float nh = gmax(0.0f, h.y);
    if (nh == 0.0f) return 1e-5f;
    const float nh_2 = nh * nh;
    const float m2 = l.roughness * l.roughness;
    const float d1 = 1.0f / (kPi * m2 * nh * nh * nh * nh);
    const float d2 = (nh_2 - 1.0f) / (m2 * nh_2);
    const float om = dot(wi, h);
    if (om <= 0.0f) return 1e+5f;
    return (d1 * f_exp(d2) * nh) / (4.0f * om);
}

__device__ inline float ggx_pdf(const spt_layer& l, V3 wo, V3 wi) {
    if (wi.y <= 0.0f || wo.y <= 0.0f) return 0.0f;
    const V3 h = normalize(wo + wi);
    const float om = dot(wi, h);
    if (om <= 0.0f) return 0.0f;
    return ggx_d(h.y, l.roughness) * h.y / (4.0f * om);
}

__device__ inline BSample layer_sample(const spt_layer& l, V3 wo, Draws& r) {
    switch (l.type) {
        case SPT_BSDF_LAMBERTIAN: {
            if (l.sampler == SPT_SAMPLER_UNIFORM) return sample_uniform(r);
            // CosineBSDFSampler (cosine_bsdf_sampler.cpp:5-17)
            const float r1 = r.next();
            const float r2 = r.next();
            const float phi = 2.0f * kPi * r2;
            const float ct = sqrt_rn(r1);
            const float s = sqrt_rn(1.0f - ct * ct);
            float sp, cp;
            f_sincos(phi, sp, cp);
            return BSample{V3{cp * s, ct, sp * s}, ct / kPi};
        }
        case SPT_BSDF_COOK_TORRANCE: {
            if (l.sampler == SPT_SAMPLER_UNIFORM) return sample_uniform(r);
            // Beckmann sampling (cook_torrance_bsdf_sampler.cpp:7-20)
            const float r1 = r.next();
            const float r2 = r.next();
            const float phi = 2.0f * kPi * r2;
            const float theta = f_atan(sqrt_rn(-(l.roughness * l.roughness) * f_log(1.0f - r1)));
            float sp, cp, st, ct;
            f_sincos(phi, sp, cp);
            f_sincos(theta, st, ct);
            const V3 m{cp * st, ct, sp * st};
            const V3 wi = 2.0f * m * dot(m, wo) - wo;
            return BSample{wi, beckmann_pdf(l, wo, wi)};
        }
        case SPT_BSDF_SMOOTH_CONDUCTOR:
            if (l.sampler == SPT_SAMPLER_UNIFORM) return sample_uniform(r);
            return BSample{V3{-wo.x, wo.y, -wo.z}, 1.0f};  // smooth_specular_bsdf_sampler.cpp:5-7
        case SPT_BSDF_SMOOTH_DIELECTRIC: {                 // smooth_dielectric.cpp:30-51
            float ior_i = 1.0f, ior_t = l.ior;
            if (wo.y < 0.0f) {
                ior_i = l.ior;
                ior_t = 1.0f;
            }
            const float a = ((ior_i * ior_i) / (ior_t * ior_t)) * (1.0f - wo.y * wo.y);
            if (a > 1.0f) return BSample{V3{-wo.x, wo.y, -wo.z}, 1.0f};
            const float F = schlick_ior(1.0f, l.ior, wo.y);  // grey: max over RGB is F
            if (r.next() < F) return BSample{V3{-wo.x, wo.y, -wo.z}, 1.0f};
            // RefractionBSDFSampler (refraction_bsdf_sampler.cpp:5-24)
            const V3 n{0.0f, 1.0f, 0.0f};
            float ri = 1.0f, rt = l.ior, signal = 1.0f;
            if (wo.y < 0.0f) {
                ri = l.ior;
                rt = 1.0f;
                signal = -1.0f;
            }
            const float b = 1.0f - ((ri * ri) / (rt * rt)) * (1.0f - wo.y * wo.y);
            const V3 wi = normalize((ri / rt) * (n * wo.y - wo) - signal * n * sqrt_rn(b));
            return BSample{wi, 1.0f};
        }
        case SPT_BSDF_GGX_CONDUCTOR: {
            if (l.sampler == SPT_SAMPLER_UNIFORM) return sample_uniform(r);
            // D(m) cos(theta_m) sampling (Walter et al. 2007, eq. 35-36)
            const float r1 = r.next();
            const float r2 = r.next();
            const float phi = 2.0f * kPi * r2;
            const float a2 = l.roughness * l.roughness;
            const float tan2 = a2 * r1 / (1.0f - r1);
            const float cm = 1.0f / sqrt_rn(1.0f + tan2);
            const float sm = sqrt_rn(gmax(0.0f, 1.0f - cm * cm));
            float sp, cp;
            f_sincos(phi, sp, cp);
            const V3 m{cp * sm, cm, sp * sm};
            const V3 wi = 2.0f * m * dot(m, wo) - wo;
            float pdf = 1.0f;
            if (wi.y > 0.0f) {
                const float om = dot(wi, m);
                pdf = (om > 0.0f) ? ggx_d(cm, l.roughness) * cm / (4.0f * om) : 1.0f;
            }
            return BSample{wi, pdf};
        }
    }
    return BSample{V3{0.0f, 1.0f, 0.0f}, 1.0f};
}

// LayeredBSDF::evalBSDF (materials/layered_bsdf.cpp:10-74): a random walk
// through the stack, at most 500 inner steps.
__device__ inline void layered_eval(const spt_material& m, V3 wo, V3& bsdf, V3& refl, float& pdf, V3& wi,
                                    Draws& r) {
    const int n = (int)m.n_layers;
    V3 fb{1.0f, 1.0f, 1.0f}, fr{1.0f, 1.0f, 1.0f};
    float fp = 1.0f;
    V3 nwo = wo;
    bool reached_inf = false;
    int bounce = 0;
    int id = (nwo.y > 0.0f) ? 0 : n - 1;
    const float inf = __builtin_inff();
    while ((id > -1) && (id < n) && (fp > 0.0f) && (!reached_inf) && (bounce < 500)) {
        const spt_layer& l = m.layers[id];
        const BSample s = layer_sample(l, nwo, r);
        wi = s.wi;
        fp *= s.pdf;
        if (!(fp < inf)) reached_inf = true;
        if (wi.y > 0.0f) {
            fb = (id == 0) ? fb * layer_f(l, nwo, wi) : fb * (layer_f(l, nwo, wi) * wi.y);
            fr = fr * layer_reflectance(l, nwo, wi);
            --id;
        } else if ((wi.y < 0.0f) && !layer_opaque(l)) {
            fb = (id == n - 1) ? fb * layer_f(l, nwo, wi) : fb * (layer_f(l, nwo, wi) * wi.y);
            fr = fr * layer_reflectance(l, nwo, wi);
            ++id;
        } else {
            fp = 0.0f;
        }
        if (!((fb.x < inf) && (fb.y < inf) && (fb.z < inf))) reached_inf = true;
        nwo = -wi;
        ++bounce;
    }
    if ((bounce < 500) && !reached_inf && (fp > 0.0f)) {
        pdf = fp;
        bsdf = fb;
        refl = fr;
    } else {
        pdf = 1.0f;
        bsdf = V3{0.0f, 0.0f, 0.0f};
        refl = V3{0.0f, 0.0f, 0.0f};
    }
}

// Pixel-sampler offset of sample s (pixel_samplers/*.cpp). getSample() is
// called twice per camera sample (pathtracer.cpp:110-111), x from the first
// call and y from the second; each random call draws two numbers, the first
// going to x.
__device__ inline void pixel_offset(uint32_t type, uint32_t n, float w, uint32_t s, Draws& r, float& sx,
                                    float& sy) {
    if (type == SPT_PIXEL_REGULAR) {
        const float off = w * 0.5f;
        sx = w * (s / n) + off;
        sy = w * (s % n) + off;
    } else if (type == SPT_PIXEL_JITTERED) {
        const float bx = w * (s / n), by = w * (s % n);
        const float u0 = r.next();
        (void)r.next();
        sx = bx + w * u0;
        (void)r.next();
        const float u3 = r.next();
        sy = by + w * u3;
    } else {
        const float u0 = r.next();
        (void)r.next();
        (void)r.next();
        const float u3 = r.next();
        sx = u0;
        sy = u3;
    }
}

}  // namespace spt
