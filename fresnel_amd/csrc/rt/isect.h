// isect.h -- per-primitive bounds and ray intersectors on the device.
//
// Each intersector evaluates one primitive exactly as the reference's Embree
// callback does (float operation for float operation) and reports the
// candidate distance t; the callback's `tnear < t` test is applied here and
// its `t < tfar` test is the caller's closest-hit comparison in the total
// order (t, geometry id, primitive id).  With ATTR the shading attributes of
// the hit (geometric normal, distance to the outline edge, color) are filled
// in as well.
#pragma once

#include "scene_dev.h"

namespace fh
{
struct Attr
    {
    v3 Ng;
    float d;
    v3 color;
    };

// The reference compares floats against the double 1e-5; float(1e-5) lies
// below 1e-5 and the next float above it, so `x < 1e-5` is `x <= EPS5F`.
constexpr float EPS5F = 9.99999974737875163555e-06f;

__device__ __forceinline__ v3 ld3(const float* p, uint32_t i)
    {
    return mk3(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
    }
__device__ __forceinline__ v3 xyz(float4 f)
    {
    return mk3(f.x, f.y, f.z);
    }
__device__ __forceinline__ quat q4(float4 f)
    {
    quat q;
    q.s = f.x;
    q.v = mk3(f.y, f.z, f.w);
    return q;
    }

// ---------------------------------------------------------------- bounds
// Per-primitive boxes of the reference bounds callbacks.
__device__ inline void prim_bounds(const GeomDev& g, uint32_t prim, v3& lo, v3& hi)
    {
    switch (g.type)
        {
    case G_SPHERE: // fresnel/cpu/GeometrySphere.cc:60-73
        {
        const float4 s = g.a[prim];
        lo = mk3(s.x - s.w, s.y - s.w, s.z - s.w);
        hi = mk3(s.x + s.w, s.y + s.w, s.z + s.w);
        break;
        }
    case G_CYLINDER: // fresnel/cpu/GeometryCylinder.cc:55-70
        {
        const float4 A = g.a[prim];
        const float4 B = g.b[prim];
        const float r = A.w;
        lo = mk3(smin(A.x - r, B.x - r), smin(A.y - r, B.y - r), smin(A.z - r, B.z - r));
        hi = mk3(smax(A.x + r, B.x + r), smax(A.y + r, B.y + r), smax(A.z + r, B.z + r));
        break;
        }
    case G_POLYHEDRON: // fresnel/cpu/GeometryConvexPolyhedron.cc:119-132
        {
        const float4 p = g.a[prim];
        const float r = g.xy_radius;
        lo = mk3(p.x - r, p.y - r, p.z - r);
        hi = mk3(p.x + r, p.y + r, p.z + r);
        break;
        }
    case G_MESH: // fresnel/cpu/GeometryMesh.cc:81-110
        {
        const uint32_t inst = prim / g.n_aux, face = prim - inst * g.n_aux;
        const v3 p3 = xyz(g.a[inst]);
        const quat q = q4(g.b[inst]);
        const v3 a = rotate(q, xyz(g.aux0[3 * face + 0])) + p3;
        const v3 b = rotate(q, xyz(g.aux0[3 * face + 1])) + p3;
        const v3 c = rotate(q, xyz(g.aux0[3 * face + 2])) + p3;
        lo = mk3(smin(a.x, smin(b.x, c.x)), smin(a.y, smin(b.y, c.y)), smin(a.z, smin(b.z, c.z)));
        hi = mk3(smax(a.x, smax(b.x, c.x)), smax(a.y, smax(b.y, c.y)), smax(a.z, smax(b.z, c.z)));
        break;
        }
    default: // G_POLYGON, fresnel/cpu/GeometryPolygon.cc:89-103 (z pad rounded from double)
        {
        const float4 p = g.a[prim];
        lo = mk3(p.x - g.xy_radius, p.y - g.xy_radius, -9.99999974737875163555e-06f);
        hi = mk3(p.x + g.xy_radius, p.y + g.xy_radius, 9.99999974737875163555e-06f);
        break;
        }
        }
    }

// ---------------------------------------------------------------- sphere
// fresnel/common/IntersectSphere.h:36-85 (first root past 1e-4, else the second)
// `first` tells which root was taken (the normal flips for the second).
// Cull margin of the closest-hit traversal.  The reference's sphere solve
// (b*b - v.v + r*r) cancels for distant grazing hits: in float arithmetic
// t = b - sqrt(det) can fall below the true entry distance by up to about
// sqrt(14 u)|v| ~ 9.2e-4 |v| (u = 2^-24, |v| = distance from the ray origin to
// the centre), and the spherocylinder caps use the same solve.  For a
// primitive inside a box with slab interval [tmin, tmax], tmin >= 0, a computed
// hit needs the ray to pass within about its radius of the centre, so
// |v| <= tmax + 3 R with R the scene's largest primitive half-extent; no
// primitive of the box has a computed distance below
// tmin - (CULL_K max(tmax, 0) + CULL_K 3 R).  CULL_K = 2^-9 keeps a factor two.
constexpr float CULL_K = 0.001953125f; // 2^-9

__device__ __forceinline__ bool sphere_t(v3 o, v3 d, v3 p, float r, float& t, float& d_edge, bool& first)
    {
    const v3 v = p - o;
    const float b = dot(v, d);
    float det = b * b - dot(v, v) + r * r;
    if (det < 0)
        return false;
    const v3 w = cross(v, d);
    d_edge = r - sqrtf(dot(w, w));
    det = sqrtf(det);
    t = b - det;
    first = true;
    if (t > 1e-4f)
        return true;
    t = b + det;
    first = false;
    return t > 1e-4f;
    }

// ---------------------------------------------------------------- spherocylinder
// fresnel/common/IntersectCylinder.h:126-204
__device__ inline bool cylinder_body(v3 o, v3 d, v3 A, v3 B, float r, float& t, v3& N)
    {
    const v3 Oa = o - A;
    const v3 C = B - A;
    const float cdotc_inv = 1.0f / dot(C, C);
    const float cdotd = dot(C, d);
    const float cdotOa = dot(C, Oa);
    const float a = (1.0f - cdotd * cdotd * cdotc_inv);
    const float minus_b = -2.0f * (dot(Oa, d) - cdotOa * cdotd * cdotc_inv);
    const float c = dot(Oa, Oa) - cdotOa * cdotOa * cdotc_inv - r * r;
    float det = minus_b * minus_b - 4 * a * c;
    if (det < 0)
        return false;
    det = sqrtf(det);
    bool hit = false;
    v3 P = mk3(0, 0, 0);
    float pdotc = 0;
    t = (minus_b - det) / (2.0f * a);
    if (t > 1e-4f)
        {
        P = Oa + t * d;
        pdotc = dot(P, C);
        if (pdotc >= 0 && pdotc <= dot(C, C))
            hit = true;
        }
    if (!hit)
        {
        t = (minus_b + det) / (2.0f * a);
        if (t > 1e-4f)
            {
            P = Oa + t * d;
            pdotc = dot(P, C);
            if (pdotc >= 0 && pdotc <= dot(C, C))
                hit = true;
            }
        }
    if (hit)
        {
        N = (P - pdotc * cdotc_inv * C);
        if (dot(N, d) > 0)
            N = -N;
        }
    return hit;
    }

// fresnel/common/IntersectCylinder.h:222-318
template<bool ATTR>
__device__ inline bool spherocylinder(v3 o, v3 d, v3 A, v3 B, float r, float& t, float& d_edge, v3& N,
                                      uint32_t& color_index)
    {
    t = __builtin_huge_valf();
    float t_sc = 0, d_sc = 0;
    v3 N_sc = mk3(0, 0, 0);
    const v3 Oa = o - A;
    const v3 C = B - A;
    const float cdotc = dot(C, C);
    bool hit_cyl = cylinder_body(o, d, A, B, r, t_sc, N_sc);
    if (hit_cyl && t_sc < t)
        {
        t = t_sc;
        N = N_sc;
        }
    bool first = true;
    bool hit_A = sphere_t(o, d, A, r, t_sc, d_sc, first);
    if (hit_A && t_sc < t)
        {
        const v3 P_A = Oa + t_sc * d;
        if (dot(P_A, C) <= 0)
            {
            t = t_sc;
            if (ATTR)
                N = first ? (o + t_sc * d - A) : -(o + t_sc * d - A);
            }
        else
            hit_A = false;
        }
    bool hit_B = sphere_t(o, d, B, r, t_sc, d_sc, first);
    if (hit_B && t_sc < t)
        {
        const v3 P_A = Oa + t_sc * d;
        if (dot(P_A, C) >= cdotc)
            {
            t = t_sc;
            if (ATTR)
                N = first ? (o + t_sc * d - B) : -(o + t_sc * d - B);
            }
        else
            hit_B = false;
        }
    if (ATTR)
        {
        const v3 P = Oa + t * d;
        color_index = (dot(P, C) < cdotc / 2.0f) ? 0u : 1u;
        const v3 C_vp = C - dot(C, d) * d;
        const v3 Oa_vp = Oa - dot(Oa, d) * d;
        const float oc = dot(Oa_vp, C_vp);
        const float cc = dot(C_vp, C_vp);
        if (oc < 0)
            d_edge = r - sqrtf(dot(Oa_vp, Oa_vp));
        else if (oc > cc)
            d_edge = r - sqrtf(dot(Oa_vp - C_vp, Oa_vp - C_vp));
        else
            {
            const v3 n_vp = Oa_vp - oc / cc * C_vp;
            d_edge = r - sqrtf(dot(n_vp, n_vp));
            }
        }
    return hit_A || hit_B || hit_cyl;
    }

// ---------------------------------------------------------------- triangle
// fresnel/common/IntersectTriangle.h:43-132 (Ericson's segment-triangle test)
template<bool ATTR>
__device__ inline bool triangle(float& u, float& v, float& w, float& t, float& d_edge, v3& n, v3 p, v3 q,
                                v3 a, v3 b, v3 c)
    {
    const v3 ab = b - a;
    const v3 ac = c - a;
    const v3 qp = p - q;
    n = cross(ab, ac);
    const float d = dot(qp, n);
    if (d <= 0.0f)
        return false;
    const v3 ap = p - a;
    t = dot(ap, n);
    if (t < 0.0f)
        return false;
    const v3 e = cross(qp, ap);
    v = dot(ac, e);
    if (v < 0.0f || v > d)
        return false;
    w = -dot(ab, e);
    if (w < 0.0f || v + w > d)
        return false;
    const float ood = 1.0f / d;
    t *= ood;
    if (ATTR)
        {
        v *= ood;
        w *= ood;
        u = 1.0f - v - w;
        v3 edge, pt;
        if (u < v)
            {
            if (u < w)
                {
                edge = c - b;
                pt = b;
                }
            else
                {
                edge = b - a;
                pt = a;
                }
            }
        else
            {
            if (v < w)
                {
                edge = a - c;
                pt = c;
                }
            else
                {
                edge = b - a;
                pt = a;
                }
            }
        const v3 r_hit = p + t * (q - p);
        const v3 vec = cross(edge, r_hit - pt);
        const float dsq = dot(vec, vec) / dot(edge, edge);
        d_edge = sqrtf(dsq);
        }
    return true;
    }

// p[i]; with UNI (p and i the same in every lane of the wavefront) through the
// constant address space, so that the load is a scalar one.
template<bool UNI>
__device__ __forceinline__ float4 ld4(const float4* p, uint32_t i)
    {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (UNI)
        return ((const __attribute__((address_space(4))) float4*)p)[i];
#endif
    return p[i];
    }

// The fields of geometry record `slot` that intersect_prim reads without
// attributes, loaded through the constant address space (slot the same in
// every lane: scalar loads).
__device__ __forceinline__ GeomDev uniform_geom(const GeomDev* geoms, uint32_t slot)
    {
#if defined(__HIP_DEVICE_COMPILE__)
    const GeomDev __attribute__((address_space(4)))* gp = (const GeomDev __attribute__((address_space(4)))*)geoms + slot;
#else
    const GeomDev* gp = geoms + slot;
#endif
    GeomDev g;
    g.type = gp->type;
    g.n_aux = gp->n_aux;
    g.rounding = gp->rounding;
    g.xy_radius = gp->xy_radius;
    g.a = gp->a;
    g.b = gp->b;
    g.aux0 = gp->aux0;
    g.aux1 = gp->aux1;
    return g;
    }

// ---------------------------------------------------------------- intersect one primitive
// The callbacks of fresnel/cpu/Geometry*.cc with `t < tfar` left to the caller.
// Geometry-type sets a kernel is compiled for: a spheres-only scene uses
// kernels without the other intersectors (fewer registers, more wavefronts).
// Instantiated for one type, intersect_prim holds only that intersector (the
// other cases return false), for callers that group tests by type.
constexpr uint32_t T_SPHERES = 1u << G_SPHERE;
constexpr uint32_t T_ALL = 0x1fu;

template<bool ATTR, uint32_t TYPES = T_ALL, bool UNI = false>
__device__ inline bool intersect_prim(const GeomDev& g, uint32_t prim, v3 o, v3 d, float tnear, float& t_out,
                                      Attr* attr)
    {
    if constexpr (TYPES == T_SPHERES)
        {
        const float4 s = ld4<UNI>(g.a, prim);
        const v3 p = xyz(s);
        float t, de;
        bool first;
        if (!sphere_t(o, d, p, s.w, t, de, first) || !(tnear < t))
            return false;
        t_out = t;
        if (ATTR)
            {
            attr->Ng = mk3(o.x + t * d.x - p.x, o.y + t * d.y - p.y, o.z + t * d.z - p.z);
            attr->d = de;
            attr->color = ld3(g.color, prim);
            }
        return true;
        }
    switch (g.type)
        {
    case G_SPHERE: // fresnel/cpu/GeometrySphere.cc:81-117
        {
        if constexpr (!(TYPES & (1u << G_SPHERE)))
            return false;
        const float4 s = ld4<UNI>(g.a, prim);
        const v3 p = xyz(s);
        float t, de;
        bool first;
        if (!sphere_t(o, d, p, s.w, t, de, first) || !(tnear < t))
            return false;
        t_out = t;
        if (ATTR)
            {
            attr->Ng = mk3(o.x + t * d.x - p.x, o.y + t * d.y - p.y, o.z + t * d.z - p.z);
            attr->d = de;
            attr->color = ld3(g.color, prim);
            }
        return true;
        }
    case G_CYLINDER: // fresnel/cpu/GeometryCylinder.cc:76-118
        {
        if constexpr (!(TYPES & (1u << G_CYLINDER)))
            return false;
        const float4 A = ld4<UNI>(g.a, prim);
        const float4 B = ld4<UNI>(g.b, prim);
        float t, de = 0;
        v3 N = mk3(0, 0, 0);
        uint32_t ci = 0;
        if (!spherocylinder<ATTR>(o, d, xyz(A), xyz(B), A.w, t, de, N, ci) || !(tnear < t))
            return false;
        t_out = t;
        if (ATTR)
            {
            attr->Ng = N;
            attr->d = de;
            attr->color = ld3(g.color, prim * 2 + ci);
            }
        return true;
        }
    case G_POLYHEDRON: // fresnel/cpu/GeometryConvexPolyhedron.cc:140-335
        {
        if constexpr (!(TYPES & (1u << G_POLYHEDRON)))
            return false;
        const v3 pos = xyz(ld4<UNI>(g.a, prim));
        const quat q = q4(ld4<UNI>(g.b, prim));
        const v3 dl = rotate(conj(q), d);
        const v3 ol = rotate(conj(q), o - pos);
        float t0 = -3.40282346638528859812e+38f, t1 = 3.40282346638528859812e+38f;
        uint32_t i0 = 0, i1 = 0;
        const uint32_t P = g.n_aux;
        for (uint32_t i = 0; i < P && t0 < t1; ++i)
            {
            const float4 pl = ld4<UNI>(g.aux0, i);
            const v3 n = xyz(pl);
            const float denom = dot(n, dl);
            const float t = -(pl.w + dot(n, ol)) / denom;
            if (fabsf(denom) <= EPS5F)
                {
                if (dot(ol - xyz(ld4<UNI>(g.aux1, i)), n) > 0)
                    return false;
                }
            else if (denom < 0)
                {
                if (t > t0)
                    {
                    t0 = t;
                    i0 = i;
                    }
                }
            else if (t < t1)
                {
                t1 = t;
                i1 = i;
                }
            }
        if (t0 > t1)
            return false;
        float th;
        uint32_t ih;
        if (tnear < t0)
            {
            th = t0;
            ih = i0;
            }
        else if (tnear < t1)
            {
            th = t1;
            ih = i1;
            }
        else
            return false;
        t_out = th;
        if (ATTR)
            {
            // the entry/exit plane's normal and origin; a plane index that was
            // never assigned keeps the reference's (0,0,0) initial normal
            // an exit plane that was never found keeps the reference's zero
            // normal and origin and plane 0 (only for degenerate plane sets)
            const bool set = (tnear < t0) ? true : (t1 != 3.40282346638528859812e+38f);
            const v3 n_hit = set ? xyz(ld4<UNI>(g.aux0, ih)) : mk3(0, 0, 0);
            const v3 p_hit = set ? xyz(ld4<UNI>(g.aux1, ih)) : mk3(0, 0, 0);
            attr->Ng = rotate(q, n_hit);
            attr->color = lerp(g.color_by_face, ld3(g.color, prim), ld3(g.aux_color, ih));
            float min_d = 3.40282346638528859812e+38f;
            const v3 r_hit = ol + th * dl;
            for (uint32_t i = 0; i < P; ++i)
                {
                const v3 n = xyz(ld4<UNI>(g.aux0, i));
                const v3 p = xyz(ld4<UNI>(g.aux1, i));
                v3 u = cross(n, n_hit);
                if (fabsf(dot(u, u)) > EPS5F)
                    {
                    int maxc;
                    if (fabsf(u.x) > fabsf(u.y))
                        maxc = (fabsf(u.x) > fabsf(u.z)) ? 1 : 3;
                    else
                        maxc = (fabsf(u.y) > fabsf(u.z)) ? 2 : 3;
                    v3 x0;
                    const float d1 = -dot(n, p);
                    const float d2 = -dot(n_hit, p_hit);
                    if (maxc == 1)
                        {
                        x0.x = 0;
                        x0.y = (d2 * n.z - d1 * n_hit.z) / u.x;
                        x0.z = (d1 * n_hit.y - d2 * n.y) / u.x;
                        }
                    else if (maxc == 2)
                        {
                        x0.x = (d1 * n_hit.z - d2 * n.z) / u.y;
                        x0.y = 0;
                        x0.z = (d2 * n.x - d1 * n_hit.x) / u.y;
                        }
                    else
                        {
                        x0.x = (d2 * n.y - d1 * n_hit.y) / u.z;
                        x0.y = (d1 * n_hit.x - d2 * n.x) / u.z;
                        x0.z = 0;
                        }
                    const v3 view = div_rcp(-dl, sqrtf(dot(dl, dl)));
                    u = u - dot(u, view) * view;
                    const v3 w = x0 - r_hit;
                    const v3 w_perp = w - dot(w, view) * view;
                    x0 = r_hit + w_perp;
                    const v3 vv = cross(u, x0 - r_hit);
                    const float dsq = dot(vv, vv) / dot(u, u);
                    const float dist = sqrtf(dsq);
                    if (dist < min_d)
                        min_d = dist;
                    }
                }
            attr->d = min_d;
            }
        return true;
        }
    case G_MESH: // fresnel/cpu/GeometryMesh.cc:115-192
        {
        if constexpr (!(TYPES & (1u << G_MESH)))
            return false;
        const uint32_t inst = prim / g.n_aux, face = prim - inst * g.n_aux;
        const v3 p3 = xyz(ld4<UNI>(g.a, inst));
        const quat q = q4(ld4<UNI>(g.b, inst));
        const v3 dl = rotate(conj(q), d);
        const v3 ol = rotate(conj(q), o - p3);
        const v3 v0 = xyz(ld4<UNI>(g.aux0, 3 * face)), v1 = xyz(ld4<UNI>(g.aux0, 3 * face + 1)), v2_ = xyz(ld4<UNI>(g.aux0, 3 * face + 2));
        const v3 qend = ol + dl;
        float u = 0, v = 0, w = 0, t = 0, de = 0;
        v3 n = mk3(0, 0, 0);
        if (!triangle<ATTR>(u, v, w, t, de, n, ol, qend, v0, v1, v2_)
            && !triangle<ATTR>(v, u, w, t, de, n, ol, qend, v1, v0, v2_))
            return false;
        if (!(tnear < t))
            return false;
        t_out = t;
        if (ATTR)
            {
            attr->Ng = rotate(q, n);
            attr->color = ld3(g.color, 3 * face + 0) * u + ld3(g.color, 3 * face + 1) * v
                          + ld3(g.color, 3 * face + 2) * w;
            attr->d = de;
            }
        return true;
        }
    default: // G_POLYGON, fresnel/cpu/GeometryPolygon.cc:150-232
        {
        if constexpr (!(TYPES & (1u << G_POLYGON)))
            return false;
        const float4 pa = ld4<UNI>(g.a, prim);
        const quat q = q4(ld4<UNI>(g.b, prim));
        const v3 pw = mk3(pa.x, pa.y, 0.0f);
        const v3 dl = rotate(conj(q), d);
        const v3 ol = rotate(conj(q), o - pw);
        const v3 n = mk3(0, 0, 1);
        const float dd = -dot(n, mk3(0, 0, 0));
        const float denom = dot(n, dl);
        const float t = -(dd + dot(n, ol)) / denom;
        if (fabsf(denom) <= EPS5F)
            return false;
        const v3 rh = ol + t * dl;
        v2 p2;
        p2.x = rh.x;
        p2.y = rh.y;
        // even-odd test with distance to the nearest edge (GeometryPolygon.cc:116-142)
        const uint32_t nv = g.n_aux;
        float min_d = 3.40282346638528859812e+38f;
        bool odd = false;
        uint32_t j = nv - 1;
        for (uint32_t i = 0; i < nv; i++)
            {
            const float4 fi = ld4<UNI>(g.aux0, i), fj = ld4<UNI>(g.aux0, j);
            v2 vi, vj;
            vi.x = fi.x;
            vi.y = fi.y;
            vj.x = fj.x;
            vj.y = fj.y;
            // point_line_segment_distance, fresnel/common/GeometryMath.h:28-39
            const v2 ab = vj - vi;
            const float Lab = dot(ab, ab);
            const float s = smax(0.0f, smin(1.0f, dot(p2 - vi, ab) / Lab));
            const v2 proj = vi + s * ab;
            const v2 dv = p2 - proj;
            min_d = smin(min_d, sqrtf(dot(dv, dv)));
            if ((vi.y < p2.y && vj.y >= p2.y) || (vj.y < p2.y && vi.y >= p2.y))
                {
                if (vi.x + (p2.y - vi.y) / (vj.y - vi.y) * (vj.x - vi.x) < p2.x)
                    odd = !odd;
                }
            j = i;
            }
        float d_edge = odd ? -min_d : min_d;
        if (d_edge > g.rounding)
            return false;
        if (!(tnear < t))
            return false;
        t_out = t;
        if (ATTR)
            {
            const v3 n_flip = (dot(n, dl) < 0.0f) ? n : -n;
            attr->Ng = rotate(q, n_flip);
            attr->color = ld3(g.color, prim);
            attr->d = g.rounding - d_edge;
            }
        return true;
        }
        }
    }
    } // namespace fh
