// PT-NEE-MIS shading math on the device (one path vertex per lane).
//
// Restates, with the reference's float/double roundings:
//   hit record          src/shape/triangle.h:79-113
//   first-hit frame     src/shape/intersection.h:25-33, src/bxdf/standard_material.h:64-71
//   same_hemisphere     src/bxdf/bxdf.h:42-49
//   Lambertian          src/bxdf/diffuse_bxdf.h:25-49, src/math/math.h:647-688
//   GGX glossy          src/bxdf/glossy_bxdf.h:29-79, src/math/microfacet_distribution.h:13-165
//   dielectric          src/bxdf/refraction_bxdf.h:26-52
//   light sampling      src/light/light_group.h:25-55, src/light/shape_light.h:24-48,
//                       src/shape/triangle.h:197-269
//   camera              src/camera/camera_perspective.h:30-39
#pragma once
#include "pt_device.h"

namespace cpt {

struct Isect {
    V3 normal, g_normal, bitangent, tangent, hitpoint;
};

struct Mat {
    int type;
    V3 albedo;
    float rough;
    V3 emission;
    float ior;
    V3 nmap;
    bool has_nmap;
};

__device__ __forceinline__ Mat load_mat(const float* m) {
    const float4* p = reinterpret_cast<const float4*>(m);
    float4 a = p[0], b = p[1], c = p[2], e = p[3];
    Mat r;
    r.type = (int)a.x;
    r.albedo = {a.y, a.z, a.w};
    r.rough = b.x;
    r.emission = {b.y, b.z, b.w};
    r.ior = c.x;
    r.nmap = {c.y, c.z, c.w};
    r.has_nmap = e.x != 0.0f;
    return r;
}

// Image::handle_border, REPEAT (rgb_image.h:215-280): a negative multiple of
// the size maps to the size itself, which the reference then reads as the
// first texel of the next row; the index is clamped only past the last texel
// (where the reference reads outside the image)
__device__ __forceinline__ int tex_wrap(int i, int n) {
    if (i >= n) return i % n;
    if (i < 0) return i % n + n;
    return i;
}

// ImageTexture::sample with point filtering (image_texture.h:24-30)
__device__ __forceinline__ const float* tex_texel(const DevScene& sc, int t, float u, float v) {
    const int4 d = sc.tex[t];  // width, height, channels, first texel
    const float x = u * (float)d.x;
    const float y = (float)((1.0 - (double)v) * (double)d.y);
    const int i = tex_wrap((int)floorf(x), d.x), j = tex_wrap((int)floorf(y), d.y);
    const long long n = (long long)d.x * d.y;
    long long idx = (long long)i + (long long)j * d.x;
    if (idx >= n) idx = n - 1;
    return sc.texels + d.w + idx * d.z;
}

// triangle.h:90-95: uv = uv1*alpha + uv2*beta + uv3*gamma
__device__ __forceinline__ float2 hit_uv(const DevScene& sc, const Hit& h) {
    const int4 ti = sc.tri_info[h.tri];
    const float alpha = 1.0f - (h.gamma + h.beta);
    const float2 a = sc.uv[ti.x], b = sc.uv[ti.y], c = sc.uv[ti.z];
    return make_float2(a.x * alpha + b.x * h.beta + c.x * h.gamma, a.y * alpha + b.y * h.beta + c.y * h.gamma);
}

// StandardMaterial::get_bxdf / get_normal_mapped inputs at the hit
// (standard_material.h:46-71): image textures replace the record's constants
__device__ __forceinline__ Mat hit_mat(const DevScene& sc, int matid, const Hit& h) {
    Mat m = load_mat(sc.mats + 16 * matid);
    const int4 mt = sc.mtex[matid];
    if (mt.x < 0 && mt.y < 0 && mt.z < 0 && mt.w < 0) return m;  // constant textures only
    const float2 uv = hit_uv(sc, h);
    if (mt.x >= 0) {
        const float* t = tex_texel(sc, mt.x, uv.x, uv.y);
        m.albedo = sc.tex[mt.x].z == 3 ? v3(t[0], t[1], t[2]) : v3s(t[0]);
    }
    if (mt.y >= 0) m.rough = tex_texel(sc, mt.y, uv.x, uv.y)[0];
    if (mt.z >= 0) m.type = tex_texel(sc, mt.z, uv.x, uv.y)[0] < 0.1f ? CPT_MAT_DIFFUSE : CPT_MAT_GLOSSY;
    if (mt.w >= 0) {
        const float* t = tex_texel(sc, mt.w, uv.x, uv.y);
        m.nmap = sc.tex[mt.w].z == 3 ? v3(t[0], t[1], t[2]) : v3s(t[0]);
    }
    return m;
}

// triangle.h:79-113.  The per-triangle shading record holds the geometric normal
// normalized(cross(b-a, c-a)) (computed on the host with the same float
// operations) and the three vertex normals and bitangents, so one hit costs one
// 96-byte read instead of ten scattered ones.
__device__ __forceinline__ void fill_isect(const DevScene& sc, V3 o, V3 d, const Hit& h, Isect& it, int& matid) {
    const float4* sr = sc.shade + 6 * (size_t)h.tri;
    float4 r0 = sr[0], r1 = sr[1], r2 = sr[2], r3 = sr[3], r4 = sr[4], r5 = sr[5];
    matid = __float_as_int(r0.w);
    float alpha = 1.0f - (h.gamma + h.beta);
    V3 n0 = {r1.x, r1.y, r1.z}, n1 = {r1.w, r2.x, r2.y}, n2 = {r2.z, r2.w, r3.x};
    V3 b0 = {r3.y, r3.z, r3.w}, b1 = {r4.x, r4.y, r4.z}, b2 = {r4.w, r5.x, r5.y};
    it.normal = normalized(add(add(mulf(n0, alpha), mulf(n1, h.beta)), mulf(n2, h.gamma)));
    V3 bt = normalized(add(add(mulf(b0, alpha), mulf(b1, h.beta)), mulf(b2, h.gamma)));
    it.g_normal = {r0.x, r0.y, r0.z};
    it.bitangent = bt;
    it.tangent = normalized(cross(bt, it.normal));
    it.hitpoint = add(o, mulf(d, h.t));
}

// intersection.h:25-33: normal map and frame re-orthogonalisation, first hit only
__device__ __forceinline__ void first_hit_frame(const Mat& m, Isect& it) {
    if (m.has_nmap) {
        V3 nn = m.nmap;
        nn.x = (float)((double)nn.x * 2.0 - 1.0);
        nn.y = (float)((double)nn.y * 2.0 - 1.0);
        it.normal = normalized(mat_cols_mul(it.tangent, it.bitangent, it.normal, nn));
    }
    it.bitangent = normalized(cross(it.normal, it.tangent));
    it.tangent = normalized(cross(it.bitangent, it.normal));
}

__device__ __forceinline__ bool same_hemisphere(const Isect& i, V3 wo, V3 wi) {
    float sign = dot(wo, i.g_normal);
    return sign * dot(wi, i.normal) > 0.0f && sign * dot(wi, i.g_normal) > 0.0f;
}

// math.h:647-653 (cos/sin evaluated in double, as ::cos(double) on libstdc++)
__device__ __forceinline__ V3 sample_hemisphere_cos(float r1, float r2) {
    float theta = (float)((double)r1 * 2.0 * CPT_M_PI);
    float sin_phi = (float)sqrt((double)r2);
    float cos_phi = (float)sqrt(1.0 - (double)(sin_phi * sin_phi));
    double s, c;
    sincos((double)theta, &s, &c);
    return {(float)(c * (double)sin_phi), (float)(s * (double)sin_phi), cos_phi};
}

// math.h:663-688
__device__ __forceinline__ void orthogonal(V3 v1, V3& x, V3& y, V3& z) {
    float ax = fabsf(v1.x), ay = fabsf(v1.y), az = fabsf(v1.z);
    V3 aux;
    if (ax < ay) aux = ax < az ? v3(1, 0, 0) : v3(0, 0, 1);
    else aux = ay < az ? v3(0, 1, 0) : v3(0, 0, 1);
    V3 aux2 = cross(v1, aux);
    x = normalized(v1);
    y = normalized(cross(aux2, v1));
    z = normalized(aux2);
}

__device__ __forceinline__ V3 reflect(V3 wo, V3 wh) {  // math.h:735-737
    return normalized(add(mulf(wo, -1.0f), mulf(mulf(wh, dot(wh, wo)), 2.0f)));
}

// ---- GGX (microfacet_distribution.h) ----
__device__ __forceinline__ float cos2T(V3 w) { return w.z * w.z; }
__device__ __forceinline__ float sinT(V3 w) { return (float)sqrt(1.0 - (double)cos2T(w)); }
__device__ __forceinline__ float cosP(V3 w) {
    float st = sinT(w);
    if (st == 0.0f) return 1.0f;
    return w.x / st;
}
__device__ __forceinline__ float sinP(V3 w) {
    float st = sinT(w);
    if (st == 0.0f) return 1.0f;
    return w.y / st;
}

__device__ __forceinline__ void ggx_sample11(float cosT, float U1, float U2, float& sx, float& sy) {
    if ((double)cosT > .9999) {
        float r = (float)sqrt((double)(U1 / (1 - U1)));
        float phi = (float)(6.28318530718 * (double)U2);
        double s, c;
        sincos((double)phi, &s, &c);
        sx = (float)((double)r * c);
        sy = (float)((double)r * s);
        return;
    }
    float sinTh = (float)sqrt((double)fmaxf(0.0f, 1.0f - cosT * cosT));
    float tanT = sinTh / cosT;
    float a = 1.0f / tanT;
    float G1 = (float)(2.0 / (double)(1.0f + (float)sqrt((double)(1.f + 1.f / (a * a)))));
    float A = 2 * U1 / G1 - 1;
    float tmp = 1.f / (A * A - 1.f);
    if ((double)tmp > 1e10) tmp = 1e10f;
    float B = tanT;
    float dd = B * B * tmp * tmp - (A * A - B * B) * tmp;
    float D = (float)sqrt((double)(dd < 0.0f ? 0.0f : dd));  // std::max(x, 0): NaN stays NaN
    float s1 = B * tmp - D;
    float s2 = B * tmp + D;
    sx = (A < 0 || s2 > 1.f / tanT) ? s1 : s2;
    float S;
    if (U2 > 0.5f) {
        S = 1.f;
        U2 = 2.f * (U2 - .5f);
    } else {
        S = -1.f;
        U2 = 2.f * (.5f - U2);
    }
    float z = (U2 * (U2 * (U2 * 0.27385f - 0.73369f) + 0.46341f)) /
              (U2 * (U2 * (U2 * 0.093073f + 0.309420f) - 1.000000f) + 0.597999f);
    sy = S * z * (float)sqrt((double)(1.f + sx * sx));
    if (isinf(sy) || isnan(sy)) sy = 0.0f;
}

__device__ __forceinline__ V3 ggx_sample(float al, V3 wi, float U1, float U2) {
    V3 ws = normalized(v3(al * wi.x, al * wi.y, wi.z));
    float sx, sy;
    ggx_sample11(ws.z, U1, U2, sx, sy);
    float cp = cosP(ws), spv = sinP(ws);
    float tmp = cp * sx - spv * sy;
    sy = spv * sx + cp * sy;
    sx = tmp;
    sx = al * sx;
    sy = al * sy;
    return normalized(v3(-sx, -sy, 1.0f));
}

__device__ __forceinline__ float ggx_d(float al, V3 wh) {
    float t2 = sinT(wh) / wh.z;
    t2 = t2 * t2;
    if (isinf(t2)) return 0.0f;
    float c4 = cos2T(wh) * cos2T(wh);
    float cp = cosP(wh), spv = sinP(wh);
    float e = (cp * cp / (al * al) + spv * spv / (al * al)) * t2;
    return (float)(1.0 / (CPT_M_PI * (double)al * (double)al * (double)c4 * (double)(1 + e) * (double)(1 + e)));
}

__device__ __forceinline__ float ggx_lambda(float al, V3 w) {
    float at = fabsf(w.z / sinT(w));
    if (isinf(at)) return 0.0f;
    float cp = cosP(w), spv = sinP(w);
    float alpha = (float)sqrt((double)(cp * cp * al * al + spv * spv * al * al));
    float a2t2 = (alpha * at) * (alpha * at);
    return (float)((-1.0 + sqrt(1.0 + (double)a2t2)) / 2.0);
}

__device__ __forceinline__ float ggx_pdf(float al, V3 wo, V3 wh) {
    float ad = ggx_d(al, wh);
    float ag1 = (float)(1.0 / (1.0 + (double)ggx_lambda(al, wo)));
    float adot = fabsf(dot(wo, wh));
    float aot = fabsf(wo.z);
    return ad * ag1 * adot / aot;
}

__device__ __forceinline__ float rough2alpha(float r) { return fmaxf(r * r, 0.01f); }

__device__ __forceinline__ float glossy_eval(float rough, const Isect& it, V3 wo, V3 wi) {
    float al = rough2alpha(rough);
    V3 wol = normalized(mat_cols_tmul(it.tangent, it.bitangent, it.normal, wo));
    V3 wil = normalized(mat_cols_tmul(it.tangent, it.bitangent, it.normal, wi));
    V3 wh = normalized(add(wol, wil));
    float g = (float)(1.0 / (1.0 + (double)ggx_lambda(al, wol) + (double)ggx_lambda(al, wil)));
    float gd = g * ggx_d(al, wh);
    return (float)((double)gd / (4.0 * (double)dot(wo, it.normal)));
}

__device__ __forceinline__ float glossy_pdf(float rough, const Isect& it, V3 wo, V3 wi) {
    float al = rough2alpha(rough);
    V3 wol = normalized(mat_cols_tmul(it.tangent, it.bitangent, it.normal, wo));
    V3 wh = normalized(mat_cols_tmul(it.tangent, it.bitangent, it.normal, add(wo, wi)));
    return (float)((double)ggx_pdf(al, wol, wh) / (4.0 * fabs((double)dot(wol, wh))));
}

struct DSample {
    V3 wi;
    float pdf;
    bool delta;
};

__device__ __forceinline__ DSample bsdf_sample(const Mat& m, const Isect& it, V3 wo, uint64_t& rng) {
    DSample r;
    if (m.type == 0) {
        float r1 = pcg_float(rng);
        float r2 = pcg_float(rng);
        V3 s = sample_hemisphere_cos(r1, r2);
        r.pdf = (float)((double)s.z / CPT_M_PI);
        V3 n = it.normal;
        if (dot(n, wo) < 0.0f) n = negv(n);
        V3 x, y, z;
        orthogonal(n, x, y, z);
        r.wi = mat_cols_mul(y, z, x, s);
        r.delta = false;
        return r;
    }
    if (m.type == 1) {
        if (m.rough < 0.01f) {
            r.wi = reflect(wo, it.normal);
            r.pdf = 1.0f;
            r.delta = true;
            return r;
        }
        float al = rough2alpha(m.rough);
        V3 wol = normalized(mat_cols_tmul(it.tangent, it.bitangent, it.normal, wo));
        float u1 = pcg_float(rng);
        float u2 = pcg_float(rng);
        V3 whl = normalized(ggx_sample(al, wol, u1, u2));
        V3 wh = normalized(mat_cols_mul(it.tangent, it.bitangent, it.normal, whl));
        r.wi = reflect(wo, wh);
        r.pdf = glossy_pdf(m.rough, it, wo, r.wi);
        r.delta = false;
        return r;
    }
    // dielectric (refraction_bxdf.h:26-52)
    V3 n = it.normal;
    float n2 = 1.0f / m.ior;
    if (dot(n, wo) < 0.0f) {
        n2 = 1.0f / n2;
        n = negv(n);
    }
    float cti = dot(wo, n);
    float ctt2 = (float)(1.0 - (double)(n2 * n2) * (1.0 - (double)(cti * cti)));
    V3 rr = normalized(add(mulf(mulf(wo, -1.0f), n2), mulf(n, (float)((double)(n2 * cti) - sqrt((double)ctt2)))));
    if (ctt2 < 0) rr = normalized(add(mulf(wo, -1.0f), mulf(mulf(it.normal, dot(it.normal, wo)), 2.0f)));
    r.wi = rr;
    r.pdf = 1.0f;
    r.delta = true;
    return r;
}

__device__ __forceinline__ V3 bsdf_eval(const Mat& m, const Isect& it, V3 wo, V3 wi) {
    if (m.type == 0) return divs(mulf(m.albedo, fabsf(dot(it.normal, wi))), CPT_PI_F);
    if (m.type == 1) return v3s(glossy_eval(m.rough, it, wo, wi));
    return v3s(1.0f);
}

__device__ __forceinline__ float bsdf_pdf(const Mat& m, const Isect& it, V3 wo, V3 wi) {
    if (m.type == 0) return (float)(fabs((double)dot(it.normal, wi)) / CPT_M_PI);
    if (m.type == 1) return glossy_pdf(m.rough, it, wo, wi);
    return 1.0f;
}

__device__ __forceinline__ bool bsdf_non_zero(const Mat& m, const Isect& it, V3 wo, V3 wi) {
    if (m.type == 3) return true;  // transmission allowed for the dielectric
    return same_hemisphere(it, wo, wi);
}

__device__ __forceinline__ float tri_area(const DevScene& sc, int id) {  // triangle.h:197-207
    int4 ti = sc.tri_info[id];
    V3 a = ld3(sc.vpos[ti.x]), b = ld3(sc.vpos[ti.y]), c = ld3(sc.vpos[ti.z]);
    return (float)((double)len3(cross(sub(b, a), sub(c, a))) * 0.5);
}

struct LSample {
    float pdf;
    V3 position, intensity, normal;
    bool delta;  // a point light, which BSDF sampling never reaches
};

__device__ __forceinline__ LSample light_sample(const DevScene& sc, uint64_t& rng, V3 from) {
    unsigned NL = (unsigned)sc.n_lights;
    int pick = (int)((unsigned)(int)(pcg_float(rng) * (float)NL) % NL);
    int id = sc.lights[pick];
    LSample ls;
    if (id < 0) {  // PointLight::connect_eye_path (point_light.h:20-28): no random numbers
        const float4 p = sc.plights[2 * (-1 - id)], e = sc.plights[2 * (-1 - id) + 1];
        ls.position = {p.x, p.y, p.z};
        ls.intensity = {e.x, e.y, e.z};
        ls.normal = {1.0f, 0.0f, 0.0f};
        ls.pdf = (float)(1.0 * (1.0 / (double)(float)NL));
        ls.delta = true;
        return ls;
    }
    float r1 = pcg_float(rng);
    float r2 = pcg_float(rng);
    float alpha = r1, beta = r2;
    if ((double)(alpha + beta) > 1.0) {
        alpha = 1.0f - alpha;
        beta = 1.0f - beta;
    }
    float gamma = (float)(1.0 - (double)alpha - (double)beta);
    int4 ti = sc.tri_info[id];
    V3 a = ld3(sc.vpos[ti.x]), b = ld3(sc.vpos[ti.y]), c = ld3(sc.vpos[ti.z]);
    V3 n0 = ld3(sc.vnrm[ti.x]), n1 = ld3(sc.vnrm[ti.y]), n2 = ld3(sc.vnrm[ti.z]);
    ls.position = add(add(mulf(a, alpha), mulf(b, beta)), mulf(c, gamma));
    ls.normal = normalized(add(add(mulf(n0, alpha), mulf(n1, beta)), mulf(n2, gamma)));
    const float4* mp = reinterpret_cast<const float4*>(sc.mats + 16 * ti.w);
    float4 mb = mp[1];
    V3 em = {mb.y, mb.z, mb.w};
    ls.intensity = mulf(em, fabsf(dot(ls.normal, normalized(sub(from, ls.position)))));
    ls.pdf = 1.0f / tri_area(sc, id);
    ls.pdf = (float)((double)ls.pdf * (1.0 / (double)(float)NL));
    ls.delta = false;
    return ls;
}

__device__ __forceinline__ float light_pdf(const DevScene& sc, int id) {
    float pdf = 1.0f / (float)sc.n_lights;
    pdf *= 1.0f / tri_area(sc, id);
    return pdf;
}

__device__ __forceinline__ float russian_roulette(V3 c) {  // pathtracerMIS.h:24-27
    return (float)(fabs((double)c.x * 0.2 + (double)c.y * 0.5 + (double)c.z * 0.3) * 0.5 + 0.4);
}

// survival probability; the corrected mode clamps it at 1 (the reference's can exceed 1)
__device__ __forceinline__ float rr_prob(V3 c, int corrected) {
    const float q = russian_roulette(c);
    return corrected ? fminf(q, 1.0f) : q;
}

// Start of a light path (Lighttracer::integrate, lighttracer.h:27-35):
// LightGroup::sample (light_group.h:33-39) picks a light; an emitter triangle
// gives ShapeLight::sample (shape_light.h:50-68): Triangle::sample's point,
// normal, bitangent and tangent (triangle.h:209-269), then
// EmissiveBxDF::emit_sample (emissive_bxdf.h:36-55), a cosine direction about the
// normal or, with probability 1/2, about the flipped normal and tangent, pdf
// 0.5 cos / pi; a point light gives PointLight::sample (point_light.h:30-37) with
// sample_sphere (math.h:655-660), whose acos argument leaves [-1, 1] for most
// draws (those paths carry no light), pdf 1 / (4 pi^2).  The two draws of
// sample_sphere's arguments are taken in source order.  Returns false when the
// direction is not a number or the scene has no light.
__device__ __forceinline__ bool light_start(const DevScene& sc, uint64_t& rng, V3& o, V3& d, V3& rad, float& pdf) {
    const unsigned NL = (unsigned)sc.n_lights;
    if (NL == 0) return false;
    const int pick = (int)((unsigned)(int)(pcg_float(rng) * (float)NL) % NL);
    const int id = sc.lights[pick];
    if (id < 0) {
        const float4 p = sc.plights[2 * (-1 - id)], e = sc.plights[2 * (-1 - id) + 1];
        const float r1 = pcg_float(rng);
        const float r2 = pcg_float(rng);
        const float theta = (float)(2.0 * CPT_M_PI * (double)r1);
        const float phi = (float)acos(1.0 - 2.0 * CPT_M_PI * (double)r2);
        double st, ct, sp, cp;
        sincos((double)theta, &st, &ct);
        sincos((double)phi, &sp, &cp);
        d = {(float)(sp * ct), (float)(sp * st), (float)cp};
        o = {p.x, p.y, p.z};
        rad = {e.x, e.y, e.z};
        pdf = (float)(1.0 / (4.0 * CPT_M_PI * CPT_M_PI));
    } else {
        const float r1 = pcg_float(rng);
        const float r2 = pcg_float(rng);
        float alpha = r1, beta = r2;
        if ((double)(alpha + beta) > 1.0) {
            alpha = 1.0f - alpha;
            beta = 1.0f - beta;
        }
        const float gamma = (float)(1.0 - (double)alpha - (double)beta);
        const int4 ti = sc.tri_info[id];
        const V3 a = ld3(sc.vpos[ti.x]), b = ld3(sc.vpos[ti.y]), c = ld3(sc.vpos[ti.z]);
        const V3 n0 = ld3(sc.vnrm[ti.x]), n1 = ld3(sc.vnrm[ti.y]), n2 = ld3(sc.vnrm[ti.z]);
        const V3 b0 = ld3(sc.vbit[ti.x]), b1 = ld3(sc.vbit[ti.y]), b2 = ld3(sc.vbit[ti.z]);
        o = add(add(mulf(a, alpha), mulf(b, beta)), mulf(c, gamma));
        const V3 n = normalized(add(add(mulf(n0, alpha), mulf(n1, beta)), mulf(n2, gamma)));
        const V3 bt = normalized(add(add(mulf(b0, alpha), mulf(b1, beta)), mulf(b2, gamma)));
        const V3 t = normalized(cross(bt, n));
        const float e1 = pcg_float(rng);
        const float e2 = pcg_float(rng);
        const V3 hs = sample_hemisphere_cos(e1, e2);
        const float p = (float)(0.5 * (double)hs.z / CPT_M_PI);
        const float side = pcg_float(rng);
        V3 nn = n, tt = t;
        if (side < 0.5f) {
            nn = negv(n);
            tt = negv(t);
        }
        d = mat_cols_mul(tt, bt, nn, hs);
        const float4 mb = reinterpret_cast<const float4*>(sc.mats + 16 * ti.w)[1];
        rad = mulf(v3(mb.y, mb.z, mb.w), fabsf(dot(n, d)));
        pdf = (float)((1.0 / (double)tri_area(sc, id)) * (double)p);
    }
    pdf = (float)((double)pdf * (1.0 / (double)(float)NL));
    return !(isnan(d.x) || isnan(d.y) || isnan(d.z));
}

// CameraPerspective::connect_light_path (camera_perspective.h:46-82): the pixel a
// point projects to and the importance factor 1 / (4 |cos^3| d^2 tan^2) (the
// reference leaves the aspect ratio out).  ci = -1 when the point is behind the
// camera, closer than EPS or outside the image.  with_aspect adds the aspect ratio
// to the image-plane area (the corrected mode).
__device__ __forceinline__ void camera_connect(const float* cam, int W, int H, V3 p, int& ci, int& cj, float& factor,
                                               bool with_aspect) {
    ci = cj = -1;
    factor = 0.0f;
    const V3 x = {cam[3], cam[4], cam[5]}, y = {cam[6], cam[7], cam[8]}, z = {cam[9], cam[10], cam[11]};
    const V3 v = sub(p, v3(cam[0], cam[1], cam[2]));
    V3 dir = {dot(x, v), dot(y, v), dot(z, v)};  // coords.transpose() * v
    const float ndl = len3(dir);
    if ((double)ndl < 1e-4 || dir.z >= 0.0f) return;
    const V3 nd = normalized(divs(dir, ndl));
    dir = mulf(divs(dir, dir.z), -1.0f);
    dir.x = dir.x / (cam[12] * cam[13]);
    dir.y = dir.y / cam[12];
    const double fx = ((double)dir.x * 0.5 + 0.5) * (double)W;
    const double fy = ((double)dir.y * 0.5 + 0.5) * (double)H;
    if (!(fx >= 0.0 && fx < (double)W && fy >= 0.0 && fy < (double)H)) return;
    ci = (int)fx;
    cj = H - (int)fy - 1;
    const float c3 = fabsf(nd.z * nd.z * nd.z);
    double den = 4.0 * (double)c3 * (double)ndl * (double)ndl * (double)cam[12] * (double)cam[12];
    if (with_aspect) den *= (double)cam[13];
    factor = (float)(1.0 / den);
}

// camera_perspective.h:30-39
__device__ __forceinline__ void camera_ray(const float* cam, float cx, float cy, V3& o, V3& d) {
    V3 dir;
    dir.z = -1.0f;
    dir.y = cy * cam[12];
    dir.x = cx * cam[12] * cam[13];
    V3 x = {cam[3], cam[4], cam[5]}, y = {cam[6], cam[7], cam[8]}, z = {cam[9], cam[10], cam[11]};
    d = normalized(mat_cols_mul(x, y, z, dir));
    o = {cam[0], cam[1], cam[2]};
}

// pathtracerMIS.h:150-153
__device__ __forceinline__ void film_coords(int i, int j, int k, int l, int S, int W, int H, float r1, float r2,
                                            float& cx, float& cy) {
    float qx = ((float)i + (float)((float)k + r1) / (float)S) / (float)W;
    float qy = ((float)j + (float)((float)l + r2) / (float)S) / (float)H;
    cx = qx * 2.0f - 1.0f;
    cy = -(qy * 2.0f - 1.0f);
}

}  // namespace cpt
