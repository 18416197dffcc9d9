// shade.h -- camera rays, Disney BRDF, lobe sampling and path events on the device.
#pragma once

#include "scene_dev.h"

#define FH_SC_QUAL __device__ __forceinline__
#define FH_SC_TABLE static __constant__ const
#include "sincos_poly.h"
#include "sincos_hard.h"

namespace fh
{
// Shading arithmetic (BRDF, lobe sampling and pdfs, MIS weights, Russian
// roulette, transmission) performs the reference's float operations in the
// reference's order: correctly rounded division and square root (the device
// code is compiled with -fhip-fp32-correctly-rounded-divide-sqrt and without
// contraction), `v / s` of a vector as a multiply by 1 / s and `v /= s` as
// three divisions (fresnel/common/VectorMath.h:508-542), `fast::rsqrt(x)` as
// 1 / sqrtf(x) (VectorMath.h:46-53).  cosf and sinf are the correctly rounded
// values, as the CPU checker computes them: C libraries' cosf/sinf are not
// specified to the last bit (glibc's differs from the correctly rounded value
// for about 1.3 % of angles in [0, 2 pi]), and the correctly rounded value is
// what they return for all but near-midpoint arguments.  So a path's every
// bounce is bit-identical to the checker's.

// Correctly rounded sinf(x), cosf(x) for 0 <= x <= 8 (the angles 2 pi u of
// lobe and aperture samples, u in (0, 1]): the double evaluation of
// sincos_poly rounded to float, or, for the 101 floats in that range whose
// rounding it cannot decide, the tabulated correctly rounded values.
__device__ __forceinline__ void cr_sincosf(float x, float& s, float& c)
    {
    double sd, cd;
    const bool hard = sincos_poly(x, sd, cd);
    s = (float)sd;
    c = (float)cd;
    if (hard)
        {
        int lo = 0, hi = SINCOS_HARD_N - 1;
        while (lo < hi)
            {
            const int mid = (lo + hi) >> 1;
            if (SINCOS_HARD[mid][0] < x)
                lo = mid + 1;
            else
                hi = mid;
            }
        if (SINCOS_HARD[lo][0] == x)
            {
            s = SINCOS_HARD[lo][1];
            c = SINCOS_HARD[lo][2];
            }
        }
    }

__device__ __forceinline__ v3 mat_color(const Mat& m)
    {
    return mk3(m.cr, m.cg, m.cb);
    }
// Material::getColor, fresnel/common/Material.h:225-228
__device__ __forceinline__ v3 get_color(const Mat& m, v3 sc)
    {
    return lerp(m.mix, mat_color(m), sc);
    }
// schlick, Material.h:26-33
__device__ __forceinline__ float schlick(float x)
    {
    const float v = 1.0f - x;
    const float v_sq = v * v;
    return v_sq * v_sq * v;
    }

// Disney diffuse term shared by brdf and brdf_diffuse (Material.h:88-93)
__device__ __forceinline__ v3 disney_diffuse(const Mat& m, v3 base_color, float ndotl, float ndotv, float ldoth)
    {
    const float FL = schlick(ndotl);
    const float FV = schlick(ndotv);
    const float RR = 2 * m.roughness * ldoth * ldoth;
    return div_rcp(base_color, PI_F)
           * ((1.0f - 0.5f * FL) * (1.0f - 0.5f * FV) + RR * (FL + FV + FL * FV * (RR - 1.0f)));
    }

// Material::brdf, Material.h:67-122
__device__ inline v3 brdf(const Mat& m, v3 l, v3 v, v3 n, v3 sc)
    {
    const float ndotv = dot(n, v);
    if (ndotv <= 0)
        return mk3(0, 0, 0);
    v3 h = l + v;
    h = div_true(h, sqrtf(dot(h, h)));
    const float ndotl = dot(n, l);
    const float ldoth = dot(l, h);
    const float ndoth = dot(n, h);
    const v3 base_color = get_color(m, sc);
    const v3 f_d = disney_diffuse(m, base_color, ndotl, ndotv, ldoth);
    const float alpha = m.roughness * m.roughness;
    const float alpha_sq = alpha * alpha;
    const float denom_rt = (1.0f + (alpha_sq - 1.0f) * ndoth * ndoth);
    const float D = alpha_sq / (PI_F * denom_rt * denom_rt);
    const float f0 = 0.08f * m.specular;
    const v3 F0 = lerp(m.metal, mk3(f0, f0, f0), base_color);
    const v3 F = F0 + (mk3(1.0f, 1.0f, 1.0f) - F0) * schlick(ldoth);
    const float ndotv_sq = ndotv * ndotv;
    const float ndotl_sq = ndotl * ndotl;
    const float V1 = 1.0f / (ndotv + sqrtf(alpha_sq + ndotv_sq - alpha_sq * ndotv_sq));
    const float V2 = 1.0f / (ndotl + sqrtf(alpha_sq + ndotl_sq - alpha_sq * ndotl_sq));
    const float V = V1 * V2;
    const v3 f_s = D * F * V;
    return f_d * (1.0f - m.metal) + f_s;
    }

// Material::brdf_diffuse, Material.h:124-153
__device__ inline v3 brdf_diffuse(const Mat& m, v3 l, v3 v, v3 n, v3 sc)
    {
    const float ndotv = dot(n, v);
    if (ndotv <= 0 || m.metal == 1.0f)
        return mk3(0, 0, 0);
    v3 h = l + v;
    h = div_true(h, sqrtf(dot(h, h)));
    const float ndotl = dot(n, l);
    const float ldoth = dot(l, h);
    return disney_diffuse(m, get_color(m, sc), ndotl, ndotv, ldoth) * (1.0f - m.metal);
    }

// Material::brdf_specular, Material.h:155-218; tan_clamped is tanf of the
// light half angle after the 0.99 pi/2 clamp (a per-light host constant).
__device__ inline v3 brdf_specular(const Mat& m, v3 l, v3 v, v3 n, v3 sc, float light_half_angle,
                                   float tan_clamped)
    {
    const float ndotv = dot(n, v);
    if (ndotv <= 0)
        return mk3(0, 0, 0);
    v3 h = l + v;
    h = div_true(h, sqrtf(dot(h, h)));
    const float ndotl = dot(n, l);
    const float ldoth = dot(l, h);
    const float ndoth = dot(n, h);
    const v3 base_color = get_color(m, sc);
    const float alpha = m.roughness * m.roughness;
    const float alpha_sq = alpha * alpha;
    const float denom_rt = (1.0f + (alpha_sq - 1.0f) * ndoth * ndoth);
    float D = alpha_sq / (PI_F * denom_rt * denom_rt);
    if (light_half_angle > 0.99f * PI_F / 2.0f)
        light_half_angle = 0.99f * PI_F / 2.0f;
    if (light_half_angle > 0.0f)
        {
        float alpha_prime = alpha + 0.5f * tan_clamped;
        if (alpha_prime > 1.0f)
            alpha_prime = 1.0f;
        D = D * alpha_sq / (alpha_prime * alpha_prime);
        }
    const float f0 = 0.08f * m.specular;
    const v3 F0 = lerp(m.metal, mk3(f0, f0, f0), base_color);
    const v3 F = F0 + (mk3(1.0f, 1.0f, 1.0f) - F0) * schlick(ldoth);
    const float ndotv_sq = ndotv * ndotv;
    const float ndotl_sq = ndotl * ndotl;
    const float V1 = 1.0f / (ndotv + sqrtf(alpha_sq + ndotv_sq - alpha_sq * ndotv_sq));
    const float V2 = 1.0f / (ndotl + sqrtf(alpha_sq + ndotl_sq - alpha_sq * ndotl_sq));
    return D * F * (V1 * V2);
    }

// tangent frame of Material.h:243-249 / 293-299.  `fabs(n.z) > 0.999` is a
// double comparison; float(0.999) lies above 0.999, so it is `>= 0.999f`.
__device__ __forceinline__ void tangent_frame(v3 n, v3& t_x, v3& t_y)
    {
    v3 up = mk3(0, 0, 1.0f);
    if (fabsf(n.z) >= 0.999f)
        up = mk3(1.0f, 0, 0);
    t_x = cross(up, n);
    t_x = div_rcp(t_x, sqrtf(dot(t_x, t_x)));
    t_y = cross(n, t_x);
    }

// Material::pdfGGX, Material.h:258-280
__device__ inline float pdf_ggx(const Mat& m, v3 l, v3 v, v3 n)
    {
    v3 h = l + v;
    h = div_true(h, sqrtf(dot(h, h)));
    const float ldoth = dot(l, h);
    const float ndoth = dot(n, h);
    const float alpha = m.roughness * m.roughness;
    const float alpha_sq = alpha * alpha;
    const float denom_rt = (1.0f + (alpha_sq - 1.0f) * ndoth * ndoth);
    const float D = alpha_sq / (PI_F * denom_rt * denom_rt);
    const float pdf_l = D * ndoth / (4.0f * ldoth);
    return (pdf_l > 0.0f) ? pdf_l : 0.0f;
    }
// Material::pdfDiffuse, Material.h:306-313
__device__ __forceinline__ float pdf_diffuse(v3 l, v3 n)
    {
    const float ndotl = dot(n, l);
    return (ndotl > 0.0f) ? ndotl / PI_F : 0.0f;
    }

// ---------------------------------------------------------------- camera
// Camera::importanceSampleAA + generateRay, fresnel/common/Camera.h:169-269
// xs_off = 0.5f * float(W) / float(H), evaluated once per image by the host
// (the same IEEE float operations, so the same value); inv_h = 1 / H in double.
//
// x / H as (float)(x * inv_h) in double is the correctly rounded float
// quotient for H <= 2^20: with x = A 2^j (2^23 <= |A| < 2^24) a float
// midpoint M 2^k (M odd, 25 bits) near x / H satisfies x - H M 2^k != 0 (the
// odd part of H M has 25 or more bits, that of A fewer), so
// |x / H - M 2^k| >= 2^k / H and, as j - k <= log2 H + 2, the quotient lies at
// least 2^-46 (relative) from every midpoint, while the double product errs by
// at most 2^-52.  Larger H take the float division.
__device__ inline void camera_ray(const CamDev& c, uint32_t W, uint32_t H, float xs_off, double inv_h, uint32_t seed,
                                  uint32_t i, uint32_t j, uint32_t sample, bool aa, v3& o, v3& d)
    {
    const uint32_t key0 = j * W + i;
    float i_f, j_f;
    if (aa)
        {
        const u4 r = philox(0, 0, sample, RNG_AA, key0, seed);
        const float r1 = u01(r.v[0]) * 2.0f;
        const float r2 = u01(r.v[1]) * 2.0f;
        // tent filter: r < 1 ? sqrt(r) - 1 : 1 - sqrt(2 - r), one square root
        const bool low1 = r1 < 1.0f, low2 = r2 < 1.0f;
        const float s1 = sqrtf(low1 ? r1 : 2.0f - r1);
        const float s2 = sqrtf(low2 ? r2 : 2.0f - r2);
        const float dx = low1 ? (s1 - 1.0f) : (1.0f - s1);
        const float dy = low2 ? (s2 - 1.0f) : (1.0f - s2);
        const float aa_w = 0.707106781f;
        i_f = float(i) + 0.5f + dx * aa_w;
        j_f = float(j) + 0.5f + dy * aa_w;
        }
    else
        {
        i_f = float(i) + 0.5f;
        j_f = float(j) + 0.5f;
        }
    float i_h, j_h;
    if (H <= (1u << 20))
        {
        i_h = (float)((double)i_f * inv_h);
        j_h = (float)((double)j_f * inv_h);
        }
    else
        {
        i_h = i_f / float(H);
        j_h = j_f / float(H);
        }
    const float ys = -1.0f * (j_h - 0.5f);
    const float xs = i_h - xs_off;
    if (c.model == 0)
        {
        o = c.p + (ys * c.v + xs * c.u) * c.focal_h;
        d = -c.w;
        return;
        }
    const v3 F = c.p - c.w * c.focal_d;
    const v3 C = F + (ys * c.v + xs * c.u) * c.focal_h;
    if (c.pinhole)
        o = c.p; // the lens offset below is (+-0, +-0, +-0)
    else
        {
        const u4 r = philox(0, 0, sample, RNG_APERTURE, key0, seed);
        const float theta = u01(r.v[0]) * 2.0f * PI_F;
        const float rr = u01(r.v[1]) * c.a * 0.5f;
        float st, ct;
        cr_sincosf(theta, st, ct);
        const v3 offset = rr * ct * c.u + rr * st * c.v;
        o = c.p + offset;
        }
    d = C - o;
    d = d * (1.0f / sqrtf(dot(d, d)));
    }

// ---------------------------------------------------------------- path events
// path_tracer_hit after the outline choice: fresnel/common/TracerPathMethods.h:79-168
// with RayGen::MISReflectionTransmission and shouldTerminatePath
// (fresnel/common/RayGen.h:110-186).  Returns true when the path continues;
// `add` receives a solid emitter's contribution (add_valid).
__device__ inline bool path_hit(const Mat& m, const Attr& at, v3 ro, v3 rd, float t_hit, uint32_t key0, uint32_t seed,
                                uint32_t depth, uint32_t sidx, v3& att, v3& next_o, v3& next_d, v3& add,
                                bool& add_valid)
    {
    add_valid = false;
    const v3 Ng = at.Ng;
    const v3 n = Ng * (1.0f / sqrtf(dot(Ng, Ng)));
    const v3 v = -rd;
    if (m.solid > 0.5f)
        {
        add = att * get_color(m, at.color);
        add_valid = true;
        return false;
        }
    bool done = false;
    const u4 ru = philox(0, depth, sidx, RNG_MIS, key0, seed);
    const float xi_x = u01(ru.v[0]);
    const float xi_y = u01(ru.v[1]);
    const float choice_mis = u01(ru.v[2]);
    const float choice_trans = u01(ru.v[3]);
    float factor = 1.0f;
    v3 l;
    const bool transmit = (choice_trans <= m.spec_trans);
    if (transmit)
        l = -v;
    else
        {
        // Material::importanceSampleDiffuse (Material.h:282-304) or
        // importanceSampleGGX (Material.h:230-256) in one code path, so a
        // wavefront whose lanes pick both lobes runs the frame, the angle and
        // the two pdfs once; each lane performs its own lobe's operations.
        const bool diffuse = choice_mis <= 0.5f;
        const float alpha = m.roughness * m.roughness;
        const float cos_theta = sqrtf((1.0f - xi_y) / (1.0f + (alpha * alpha - 1.0f) * xi_y));
        const float sin_theta = sqrtf(1.0f - cos_theta * cos_theta);
        const float radial = diffuse ? sqrtf(xi_x) : sin_theta;
        const float angle = diffuse ? xi_y : xi_x;
        const float zc = diffuse ? sqrtf(1.0f - xi_x) : cos_theta;
        float sn, cs;
        cr_sincosf(2.0f * PI_F * angle, sn, cs);
        const v3 s_t = mk3(radial * cs, radial * sn, zc);
        v3 t_x, t_y;
        tangent_frame(n, t_x, t_y);
        const v3 w = t_x * s_t.x + t_y * s_t.y + n * s_t.z;
        l = diffuse ? w : (2.0f * dot(v, w) * w - v);
        const float pd = pdf_diffuse(l, n);
        const float pg = pdf_ggx(m, l, v, n);
        const float p_lobe = diffuse ? pd : pg;
        factor = (p_lobe / (pd + pg)) / (0.5f * p_lobe);
        }
    if (transmit)
        {
        const v3 tc = get_color(m, at.color);
        att = att * mk3(sqrtf(tc.x), sqrtf(tc.y), sqrtf(tc.z));
        }
    else
        {
        const float ndotl = dot(n, l);
        if (dot(n, v) > 0.0f && ndotl > 0.0f)
            att = att * (brdf(m, l, v, n, at.color) * ndotl * factor);
        else
            {
            att = mk3(0, 0, 0);
            done = true;
            }
        }
    next_o = ro + rd * t_hit;
    next_d = l;
    const u4 rr = philox(0, depth, sidx, RNG_RR, key0, seed);
    float p_continue = fmaxf(att.x, fmaxf(att.y, att.z));
    p_continue = fminf(p_continue, 1.0f);
    if (u01(rr.v[0]) > p_continue)
        done = true;
    else
        att = div_true(att, p_continue);
    return !done;
    }

// Lights whose cone contains an escaping direction (TracerPathMethods.h:55-73).
__device__ __forceinline__ uint32_t light_mask(const LightsDev& L, v3 rd)
    {
    uint32_t mask = 0;
    for (uint32_t li = 0; li < L.N; li++)
        if (dot(L.dir[li], rd) >= L.cos_theta[li])
            mask |= 1u << li;
    return mask;
    }

// sRGB of fresnel/common/ColorMath.h:514-546
__device__ __forceinline__ unsigned char srgb_channel(float c)
    {
    float t = (c < 0.0031308f) ? 12.92f * c : (1.0f + 0.055f) * powf(c, 1.0f / 2.4f) - 0.055f;
    if (t > 1.0f)
        t = 1.0f;
    return (unsigned char)(t * 255.0f + 0.5f);
    }
__device__ __forceinline__ uchar4 to_srgb(float4 c)
    {
    return make_uchar4(srgb_channel(c.x), srgb_channel(c.y), srgb_channel(c.z), (unsigned char)(c.w * 255.0f + 0.5f));
    }
    } // namespace fh
