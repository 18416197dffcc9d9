// fresnel_oracle.cpp -- CPU restatement of fresnel's path-tracer arithmetic.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the fresnel_amd package links, loads or
// calls this file; only tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline use it, and only as the checker.
//
// It restates, in plain scalar C++, the semantics of the reference's Embree CPU
// path (fresnel 0.13.5):
//   common/VectorMath.h, ColorMath.h, Camera.h, Light.h, Material.h, RayGen.h,
//   TracerPathMethods.h, IntersectSphere.h, IntersectCylinder.h,
//   IntersectTriangle.h, GeometryMath.h,
//   cpu/TracerPath.cc, cpu/TracerDirect.cc, cpu/Geometry*.cc (bounds and
//   intersect callbacks).
// Each function cites the reference file:line it follows.
//
// Two deliberate, documented choices replace what Embree decides internally:
//   * closest hit is taken in the total order (t, geomID, primID) instead of
//     Embree's unspecified visit order for exact ties;
//   * a primitive's intersector runs only when the ray's float slab test over
//     that primitive's bounds box passes (Embree also culls by the bounds the
//     callbacks report; its exact box arithmetic is not public).
// Both rules are shared with the GPU path, so results do not depend on BVH
// shape.
//
// Build: g++ -O2 -ffp-contract=off (no FMA contraction, IEEE single precision
// on SSE), matching the reference's host build flags (CMakeLists.txt:28-43).

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <thread>
#include <vector>

namespace orc
{
// ---------------------------------------------------------------- vectors
// VectorMath.h:330-600.  Every operator keeps the reference's evaluation order.
struct v3
    {
    float x, y, z;
    };
struct v2
    {
    float x, y;
    };
static inline v3 mk(float x, float y, float z)
    {
    return v3 {x, y, z};
    }
static inline v3 operator+(v3 a, v3 b)
    {
    return mk(a.x + b.x, a.y + b.y, a.z + b.z);
    }
static inline v3 operator-(v3 a, v3 b)
    {
    return mk(a.x - b.x, a.y - b.y, a.z - b.z);
    }
static inline v3 operator-(v3 a)
    {
    return mk(-a.x, -a.y, -a.z);
    }
static inline v3 operator*(v3 a, v3 b)
    {
    return mk(a.x * b.x, a.y * b.y, a.z * b.z);
    }
static inline v3 operator*(v3 a, float s)
    {
    return mk(a.x * s, a.y * s, a.z * s);
    }
static inline v3 operator*(float s, v3 a)
    {
    return mk(a.x * s, a.y * s, a.z * s);
    }
// `vec / scalar` is a reciprocal multiply (VectorMath.h:508-512)
static inline v3 div_rcp(v3 a, float s)
    {
    float q = 1.0f / s;
    return a * q;
    }
// `vec /= scalar` is a true division (VectorMath.h:536-542)
static inline v3 div_true(v3 a, float s)
    {
    return mk(a.x / s, a.y / s, a.z / s);
    }
// dot summed left to right (VectorMath.h:570-573)
static inline float dot(v3 a, v3 b)
    {
    return a.x * b.x + a.y * b.y + a.z * b.z;
    }
static inline v3 cross(v3 a, v3 b)
    {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
    }
static inline v2 operator-(v2 a, v2 b)
    {
    return v2 {a.x - b.x, a.y - b.y};
    }
static inline v2 operator+(v2 a, v2 b)
    {
    return v2 {a.x + b.x, a.y + b.y};
    }
static inline v2 operator*(float s, v2 a)
    {
    return v2 {a.x * s, a.y * s};
    }
static inline float dot(v2 a, v2 b)
    {
    return a.x * b.x + a.y * b.y;
    }

// quaternion (s, v), VectorMath.h:900-1102
struct quat
    {
    float s;
    v3 v;
    };
static inline quat conj(quat a)
    {
    return quat {a.s, -a.v};
    }
// VectorMath.h:1080-1092 (the bkph form)
static inline v3 rotate(quat a, v3 b)
    {
    return (a.s * a.s - dot(a.v, a.v)) * b + 2 * a.s * cross(a.v, b) + 2 * dot(a.v, b) * a.v;
    }
// VectorMath.h:920-925 (fast::cos / fast::sin are cosf / sinf on the host)
static inline quat from_axis_angle(v3 axis, float theta)
    {
    return quat {cosf(theta / 2.0f), sinf(theta / 2.0f) * axis};
    }

// ---------------------------------------------------------------- colors
// ColorMath.h: RGB uses the same operators as vec3.
typedef v3 rgb;
struct rgba
    {
    float r, g, b, a;
    };
// lerp (ColorMath.h:504-507): (1-x)*a + x*b
static inline v3 lerp(float x, v3 a, v3 b)
    {
    return (1.0f - x) * a + x * b;
    }

// sRGB conversion, ColorMath.h:514-546
static inline void srgb(const rgba& c, uint8_t* out)
    {
    float t[3] = {c.r, c.g, c.b};
    for (int k = 0; k < 3; k++)
        {
        if (t[k] < 0.0031308f)
            t[k] = 12.92f * t[k];
        else
            t[k] = (1.0f + 0.055f) * powf(t[k], 1.0f / 2.4f) - 0.055f;
        if (t[k] > 1.0f)
            t[k] = 1.0f;
        }
    out[0] = (unsigned char)(t[0] * 255.0f + 0.5f);
    out[1] = (unsigned char)(t[1] * 255.0f + 0.5f);
    out[2] = (unsigned char)(t[2] * 255.0f + 0.5f);
    out[3] = (unsigned char)(c.a * 255.0f + 0.5f);
    }

// ---------------------------------------------------------------- Philox4x32-10
// Published algorithm: Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3" (SC'11).  Random123 is absent from the reference checkout
// (extern/random123 is an empty submodule); call sites RayGen.h:77-79,118-125,
// 171-177 and Camera.h:194-199,235-240.
struct u4
    {
    uint32_t v[4];
    };
static inline u4 philox(u4 c, uint32_t k0, uint32_t k1)
    {
    for (int r = 0; r < 10; r++)
        {
        if (r)
            {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
            }
        uint64_t p0 = (uint64_t)0xD2511F53u * c.v[0];
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c.v[2];
        u4 n;
        n.v[0] = (uint32_t)(p1 >> 32) ^ c.v[1] ^ k0;
        n.v[1] = (uint32_t)p1;
        n.v[2] = (uint32_t)(p0 >> 32) ^ c.v[3] ^ k1;
        n.v[3] = (uint32_t)p0;
        c = n;
        }
    return c;
    }
// Random123 u01<float>: float(u) * 2^-32 + 2^-33, in (0, 1]
static inline float u01(uint32_t u)
    {
    return (float)u * 0x1p-32f + 0x1p-33f;
    }

static const uint32_t RNG_MIS = 0x8754abcd; // RayGen.h:34
static const uint32_t RNG_RR = 0x54abf853; // RayGen.h:37
static const uint32_t RNG_AA = 0x22ab5871; // Camera.h:305
static const uint32_t RNG_APERTURE = 0x983abc12; // Camera.h:302
static const float PI_F = float(M_PI);

// ---------------------------------------------------------------- material
// Material.h:47-314
struct Material
    {
    float solid;
    rgb color;
    float primitive_color_mix;
    float roughness;
    float specular;
    float metal;
    float spec_trans;
    };

static inline float schlick(float x) // Material.h:26-33
    {
    float v = 1.0f - x;
    float v_sq = v * v;
    float v_fifth = v_sq * v_sq * v;
    return v_fifth;
    }
static inline rgb get_color(const Material& m, rgb sc) // Material.h:225-228
    {
    return lerp(m.primitive_color_mix, m.color, sc);
    }
static inline bool is_solid(const Material& m) // Material.h:220-223
    {
    return m.solid > 0.5f;
    }
static rgb brdf(const Material& m, v3 l, v3 v, v3 n, rgb sc) // Material.h:67-122
    {
    float ndotv = dot(n, v);
    if (ndotv <= 0)
        return mk(0, 0, 0);
    v3 h = l + v;
    h = div_true(h, sqrtf(dot(h, h)));
    float ndotl = dot(n, l);
    float ldoth = dot(l, h);
    float ndoth = dot(n, h);
    rgb base_color = get_color(m, sc);
    float FL = schlick(ndotl);
    float FV = schlick(ndotv);
    float RR = 2 * m.roughness * ldoth * ldoth;
    rgb f_d = div_rcp(base_color, PI_F)
              * ((1.0f - 0.5f * FL) * (1.0f - 0.5f * FV) + RR * (FL + FV + FL * FV * (RR - 1.0f)));
    float alpha = m.roughness * m.roughness;
    float alpha_sq = alpha * alpha;
    float denom_rt = (1.0f + (alpha_sq - 1.0f) * ndoth * ndoth);
    float D = alpha_sq / (PI_F * denom_rt * denom_rt);
    rgb F0_dielectric = mk(0.08f * m.specular, 0.08f * m.specular, 0.08f * m.specular);
    rgb F0 = lerp(m.metal, F0_dielectric, base_color);
    rgb F = F0 + (mk(1.0f, 1.0f, 1.0f) - F0) * schlick(ldoth);
    float ndotv_sq = ndotv * ndotv;
    float ndotl_sq = ndotl * ndotl;
    float V1 = 1.0f / (ndotv + sqrtf(alpha_sq + ndotv_sq - alpha_sq * ndotv_sq));
    float V2 = 1.0f / (ndotl + sqrtf(alpha_sq + ndotl_sq - alpha_sq * ndotl_sq));
    float V = V1 * V2;
    rgb f_s = D * F * V;
    return f_d * (1.0f - m.metal) + f_s;
    }
static rgb brdf_diffuse(const Material& m, v3 l, v3 v, v3 n, rgb sc) // Material.h:124-153
    {
    float ndotv = dot(n, v);
    if (ndotv <= 0 || m.metal == 1.0f)
        return mk(0, 0, 0);
    v3 h = l + v;
    h = div_true(h, sqrtf(dot(h, h)));
    float ndotl = dot(n, l);
    float ldoth = dot(l, h);
    rgb base_color = get_color(m, sc);
    float FL = schlick(ndotl);
    float FV = schlick(ndotv);
    float RR = 2 * m.roughness * ldoth * ldoth;
    rgb f_d = div_rcp(base_color, PI_F)
              * ((1.0f - 0.5f * FL) * (1.0f - 0.5f * FV) + RR * (FL + FV + FL * FV * (RR - 1.0f)));
    return f_d * (1.0f - m.metal);
    }
static rgb brdf_specular(const Material& m, v3 l, v3 v, v3 n, rgb sc, float light_half_angle)
    {
    // Material.h:155-218
    float ndotv = dot(n, v);
    if (ndotv <= 0)
        return mk(0, 0, 0);
    v3 h = l + v;
    h = div_true(h, sqrtf(dot(h, h)));
    float ndotl = dot(n, l);
    float ldoth = dot(l, h);
    float ndoth = dot(n, h);
    rgb base_color = get_color(m, sc);
    float alpha = m.roughness * m.roughness;
    float alpha_sq = alpha * alpha;
    float denom_rt = (1.0f + (alpha_sq - 1.0f) * ndoth * ndoth);
    float D = alpha_sq / (PI_F * denom_rt * denom_rt);
    float alpha_prime;
    if (light_half_angle > 0.99f * PI_F / 2.0f)
        light_half_angle = 0.99f * PI_F / 2.0f;
    if (light_half_angle > 0.0f)
        {
        alpha_prime = alpha + 0.5f * tanf(light_half_angle);
        if (alpha_prime > 1.0f)
            alpha_prime = 1.0f;
        D = D * alpha_sq / (alpha_prime * alpha_prime);
        }
    rgb F0_dielectric = mk(0.08f * m.specular, 0.08f * m.specular, 0.08f * m.specular);
    rgb F0 = lerp(m.metal, F0_dielectric, base_color);
    rgb F = F0 + (mk(1.0f, 1.0f, 1.0f) - F0) * schlick(ldoth);
    float ndotv_sq = ndotv * ndotv;
    float ndotl_sq = ndotl * ndotl;
    float V1 = 1.0f / (ndotv + sqrtf(alpha_sq + ndotv_sq - alpha_sq * ndotv_sq));
    float V2 = 1.0f / (ndotl + sqrtf(alpha_sq + ndotl_sq - alpha_sq * ndotl_sq));
    float V = V1 * V2;
    return D * F * V;
    }
// cosf / sinf of the lobe and aperture samples (Material.h:240,287-288,
// Camera.h:201).  The C library's cosf/sinf are not specified to the last bit
// (glibc's, CUDA's and other libraries' differ in rare last-ulp cases); the
// checker takes the correctly rounded value (evaluated in double, rounded
// once), which those libraries return for all but near-midpoint arguments and
// which the device computes the same way.
static inline float cr_cosf(float x)
    {
    return (float)std::cos((double)x);
    }
static inline float cr_sinf(float x)
    {
    return (float)std::sin((double)x);
    }
static inline void tangent_frame(v3 n, v3& t_x, v3& t_y) // Material.h:243-249, 293-299
    {
    v3 up = mk(0, 0, 1.0f);
    if (fabs(n.z) > 0.999)
        up = mk(1.0f, 0, 0);
    t_x = cross(up, n);
    t_x = div_rcp(t_x, sqrtf(dot(t_x, t_x)));
    t_y = cross(n, t_x);
    }
static v3 sample_ggx(const Material& m, v2 xi, v3 v, v3 n) // Material.h:230-256
    {
    float alpha = m.roughness * m.roughness;
    float phi = 2.0f * PI_F * xi.x;
    float cos_theta = sqrtf((1.0f - xi.y) / (1.0f + (alpha * alpha - 1.0f) * xi.y));
    float sin_theta = sqrtf(1.0f - cos_theta * cos_theta);
    v3 h_t = mk(sin_theta * cr_cosf(phi), sin_theta * cr_sinf(phi), cos_theta);
    v3 t_x, t_y;
    tangent_frame(n, t_x, t_y);
    v3 h = t_x * h_t.x + t_y * h_t.y + n * h_t.z;
    return 2.0f * dot(v, h) * h - v;
    }
static float pdf_ggx(const Material& m, v3 l, v3 v, v3 n) // Material.h:258-280
    {
    v3 h = l + v;
    h = div_true(h, sqrtf(dot(h, h)));
    float ldoth = dot(l, h);
    float ndoth = dot(n, h);
    float alpha = m.roughness * m.roughness;
    float alpha_sq = alpha * alpha;
    float denom_rt = (1.0f + (alpha_sq - 1.0f) * ndoth * ndoth);
    float D = alpha_sq / (PI_F * denom_rt * denom_rt);
    float pdf_l = D * ndoth / (4.0f * ldoth);
    if (pdf_l > 0.0f)
        return pdf_l;
    return 0.0f;
    }
static v3 sample_diffuse(v2 xi, v3 v, v3 n) // Material.h:282-304
    {
    (void)v;
    const float r = sqrtf(xi.x);
    const float theta = 2 * PI_F * xi.y;
    const float x = r * cr_cosf(theta);
    const float y = r * cr_sinf(theta);
    v3 v_t = mk(x, y, sqrtf(1.0f - xi.x));
    v3 t_x, t_y;
    tangent_frame(n, t_x, t_y);
    return t_x * v_t.x + t_y * v_t.y + n * v_t.z;
    }
static float pdf_diffuse(v3 l, v3 n) // Material.h:306-313
    {
    float ndotl = dot(n, l);
    if (ndotl > 0.0f)
        return ndotl / PI_F;
    return 0.0f;
    }

// ---------------------------------------------------------------- camera
// Camera.h:55-309
struct UserCamera
    {
    v3 position, look_at, up;
    float h, f, f_stop, focus_distance;
    int model; // 0 orthographic, 1 perspective
    };
struct Camera
    {
    v3 p, u, v, w;
    float focal_h, focal_d, a;
    int model;
    unsigned W, H, seed;
    bool sample_aa;
    };
static Camera make_camera(const UserCamera& user, unsigned W, unsigned H, unsigned seed, bool aa)
    {
    Camera c;
    // CameraBasis, Camera.h:92-110
    v3 d = user.look_at - user.position;
    v3 vv = user.up;
    d = d * (1.0f / sqrtf(dot(d, d)));
    vv = vv * (1.0f / sqrtf(dot(vv, vv)));
    v3 uu = cross(d, vv);
    uu = uu * (1.0f / sqrtf(dot(uu, uu)));
    vv = cross(uu, d);
    vv = vv * (1.0f / sqrtf(dot(vv, vv)));
    c.u = uu;
    c.v = vv;
    c.w = -d;
    // Camera ctor, Camera.h:141-159
    c.p = user.position;
    c.focal_d = user.focus_distance;
    c.a = user.f / user.f_stop;
    c.model = user.model;
    c.W = W;
    c.H = H;
    c.seed = seed;
    c.sample_aa = aa;
    if (c.model == 0)
        c.focal_h = user.h;
    else
        c.focal_h = user.h / user.f * c.focal_d;
    return c;
    }
static v2 sample_aa(const Camera& c, unsigned i, unsigned j, unsigned sample) // Camera.h:226-269
    {
    float i_f, j_f;
    if (c.sample_aa)
        {
        u4 r = philox(u4 {{0, 0, sample, RNG_AA}}, j * c.W + i, c.seed);
        float r1 = u01(r.v[0]) * 2.0f;
        float r2 = u01(r.v[1]) * 2.0f;
        float dx, dy;
        if (r1 < 1.0f)
            dx = sqrtf(r1) - 1.0f;
        else
            dx = 1.0f - sqrtf(2.0f - r1);
        if (r2 < 1.0f)
            dy = sqrtf(r2) - 1.0f;
        else
            dy = 1.0f - sqrtf(2.0f - r2);
        const float aa_w = 0.707106781f;
        i_f = float(i) + 0.5f + dx * aa_w;
        j_f = float(j) + 0.5f + dy * aa_w;
        }
    else
        {
        i_f = float(i) + 0.5f;
        j_f = float(j) + 0.5f;
        }
    float ys = -1.0f * (j_f / float(c.H) - 0.5f);
    float xs = i_f / float(c.H) - 0.5f * float(c.W) / float(c.H);
    return v2 {xs, ys};
    }
static void generate_ray(const Camera& c, unsigned i, unsigned j, unsigned sample, v3& o, v3& d)
    {
    // Camera.h:169-209
    v2 s = sample_aa(c, i, j, sample);
    if (c.model == 0)
        {
        o = c.p + (s.y * c.v + s.x * c.u) * c.focal_h;
        d = -c.w;
        return;
        }
    v3 F = c.p - c.w * c.focal_d;
    v3 C = F + (s.y * c.v + s.x * c.u) * c.focal_h;
    u4 r = philox(u4 {{0, 0, sample, RNG_APERTURE}}, j * c.W + i, c.seed);
    float theta = u01(r.v[0]) * 2.0f * PI_F;
    float rr = u01(r.v[1]) * c.a * 0.5f;
    v3 offset = rr * cr_cosf(theta) * c.u + rr * cr_sinf(theta) * c.v;
    o = c.p + offset;
    d = C - o;
    d = d * (1.0f / sqrtf(dot(d, d)));
    }

// ---------------------------------------------------------------- lights
// Light.h:36-66 plus the per-light constants of TracerPathMethods.h:55-73
struct Lights
    {
    v3 dir[4];
    rgb color[4];
    float theta[4];
    unsigned N;
    };
static Lights lights_to_world(const Lights& in, const Camera& cam)
    {
    Lights out;
    out.N = in.N;
    for (unsigned i = 0; i < in.N; i++)
        {
        out.color[i] = in.color[i];
        v3 v = in.dir[i];
        v = v * (1.0f / sqrtf(dot(v, v)));
        out.dir[i] = v.x * cam.u + v.y * cam.v + v.z * cam.w;
        out.theta[i] = in.theta[i];
        }
    return out;
    }

// ---------------------------------------------------------------- geometry
enum GeomType
    {
    G_SPHERE = 0,
    G_CYLINDER = 1,
    G_POLYGON = 2,
    G_POLYHEDRON = 3,
    G_MESH = 4
    };

struct Geom
    {
    int type = 0;
    unsigned id = 0; // geomID: tie order and material table index
    bool enabled = true;
    Material mat {}, outline {};
    float outline_width = 0;
    unsigned N = 0; // number of primitives (mesh: instances * faces)
    // sphere / cylinder / polyhedron / mesh / polygon attribute arrays
    std::vector<v3> pos; // sphere, polyhedron, mesh instances
    std::vector<float> radius; // sphere, cylinder
    std::vector<v3> color; // per prim (cyl: 2 per prim, mesh: per vertex)
    std::vector<v3> points; // cylinder endpoints A,B interleaved
    std::vector<quat> orient; // polyhedron, mesh
    std::vector<v3> plane_origin, plane_normal; // polyhedron
    std::vector<rgb> plane_color;
    float poly_radius = 0;
    float color_by_face = 0;
    std::vector<v3> verts; // mesh (3F)
    unsigned n_faces = 0;
    std::vector<v2> pos2; // polygon
    std::vector<float> angle;
    std::vector<v2> verts2;
    float rounding_radius = 0;
    float xy_radius = 0;
    };

struct Box
    {
    v3 lo, hi;
    };

// per-primitive bounds, matching each reference bounds callback
static Box prim_bounds(const Geom& g, unsigned prim)
    {
    Box b;
    switch (g.type)
        {
    case G_SPHERE: // cpu/GeometrySphere.cc:60-73
        {
        v3 p = g.pos[prim];
        float r = g.radius[prim];
        b.lo = mk(p.x - r, p.y - r, p.z - r);
        b.hi = mk(p.x + r, p.y + r, p.z + r);
        break;
        }
    case G_CYLINDER: // cpu/GeometryCylinder.cc:55-70
        {
        v3 A = g.points[prim * 2 + 0], B = g.points[prim * 2 + 1];
        float r = g.radius[prim];
        b.lo = mk(std::min(A.x - r, B.x - r), std::min(A.y - r, B.y - r), std::min(A.z - r, B.z - r));
        b.hi = mk(std::max(A.x + r, B.x + r), std::max(A.y + r, B.y + r), std::max(A.z + r, B.z + r));
        break;
        }
    case G_POLYHEDRON: // cpu/GeometryConvexPolyhedron.cc:119-132
        {
        v3 p = g.pos[prim];
        float r = g.poly_radius;
        b.lo = mk(p.x - r, p.y - r, p.z - r);
        b.hi = mk(p.x + r, p.y + r, p.z + r);
        break;
        }
    case G_MESH: // cpu/GeometryMesh.cc:81-110
        {
        unsigned inst = prim / g.n_faces, face = prim % g.n_faces;
        v3 p3 = g.pos[inst];
        quat q = g.orient[inst];
        v3 a = rotate(q, g.verts[face * 3 + 0]) + p3;
        v3 bb = rotate(q, g.verts[face * 3 + 1]) + p3;
        v3 c = rotate(q, g.verts[face * 3 + 2]) + p3;
        b.lo = mk(std::min(a.x, std::min(bb.x, c.x)),
                  std::min(a.y, std::min(bb.y, c.y)),
                  std::min(a.z, std::min(bb.z, c.z)));
        b.hi = mk(std::max(a.x, std::max(bb.x, c.x)),
                  std::max(a.y, std::max(bb.y, c.y)),
                  std::max(a.z, std::max(bb.z, c.z)));
        break;
        }
    case G_POLYGON: // cpu/GeometryPolygon.cc:89-103 (z pad computed in double)
        {
        v2 p2 = g.pos2[prim];
        b.lo = mk(p2.x - g.xy_radius, p2.y - g.xy_radius, (float)(0.0f - 1e-5));
        b.hi = mk(p2.x + g.xy_radius, p2.y + g.xy_radius, (float)(0.0f + 1e-5));
        break;
        }
        }
    return b;
    }

// Shared float slab test.  Every primitive's intersector is gated on this test
// over the primitive's own bounds (see the file header).
static inline void slab(v3 o, v3 inv, const Box& b, float& tmin, float& tmax)
    {
    float tx0 = (b.lo.x - o.x) * inv.x, tx1 = (b.hi.x - o.x) * inv.x;
    float ty0 = (b.lo.y - o.y) * inv.y, ty1 = (b.hi.y - o.y) * inv.y;
    float tz0 = (b.lo.z - o.z) * inv.z, tz1 = (b.hi.z - o.z) * inv.z;
    tmin = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fminf(tz0, tz1));
    tmax = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1));
    }
// Cull rule by the current closest hit.  The reference's sphere solve
// (IntersectSphere.h: det = b*b - v.v + r*r) cancels for distant grazing hits,
// so the computed t can lie below the true entry distance by up to about
// sqrt(14 u)|v| ~ 9.2e-4 |v| (u = 2^-24, |v| = distance from the ray origin to
// the centre).  For a primitive inside a box whose slab interval is
// [tmin, tmax] with tmin >= 0, a computed hit needs the ray to pass within
// about its radius of the centre, so |v| <= tmax + 3 R, R the largest
// primitive half-extent of the scene.  No primitive of the box then has a
// computed distance below cull_bound(); boxes are culled only when that bound
// exceeds the closest hit, which makes the result independent of the
// hierarchy and of visiting order.  CULL_K = 2^-9 keeps a factor two.
static const float CULL_K = 0x1p-9f;
static inline float cull_bound(float tmin, float tmax, float margin)
    {
    return tmin - (CULL_K * fmaxf(tmax, 0.0f) + margin);
    }

// ---------------------------------------------------------------- intersectors
struct Hit
    {
    float t;
    v3 Ng;
    float d;
    rgb color;
    };

// IntersectSphere.h:36-85; returns the candidate root (first root > 1e-4 else
// the second); the caller applies tnear < t.
static inline bool isect_sphere(float& t, float& d_edge, v3& N, v3 o, v3 d, v3 p, float r)
    {
    const float eps = 1e-4f;
    v3 v = p - o;
    float b = dot(v, d);
    float det = b * b - dot(v, v) + r * r;
    if (det < 0)
        return false;
    const v3 w = cross(v, d);
    const float Dsq = dot(w, w);
    d_edge = r - sqrtf(Dsq);
    det = sqrtf(det);
    t = b - det;
    if (t > eps)
        {
        N = o + t * d - p;
        return true;
        }
    t = b + det;
    if (t > eps)
        {
        N = -(o + t * d - p);
        return true;
        }
    return false;
    }

// IntersectCylinder.h:126-204
static bool isect_cylinder(float& t, v3& N, v3 o, v3 d, v3 A, v3 B, float r)
    {
    const float eps = 1e-4f;
    const v3 Oa = o - A;
    const v3 C = B - A;
    float cdotc_inv = 1.0f / dot(C, C);
    float cdotd = dot(C, d);
    float cdotOa = dot(C, Oa);
    float a = (1.0f - cdotd * cdotd * cdotc_inv);
    float minus_b = -2.0f * (dot(Oa, d) - cdotOa * cdotd * cdotc_inv);
    float c = dot(Oa, Oa) - cdotOa * cdotOa * cdotc_inv - r * r;
    float det = minus_b * minus_b - 4 * a * c;
    if (det < 0)
        return false;
    det = sqrtf(det);
    bool hit = false;
    v3 P = mk(0, 0, 0);
    float pdotc = 0;
    t = (minus_b - det) / (2.0f * a);
    if (t > eps)
        {
        P = Oa + t * d;
        pdotc = dot(P, C);
        if (pdotc >= 0 && pdotc <= dot(C, C))
            hit = true;
        }
    if (!hit)
        {
        t = (minus_b + det) / (2.0f * a);
        if (t > eps)
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

// IntersectCylinder.h:222-318
static bool isect_spherocylinder(float& t, float& d_edge, v3& N, unsigned& color_index, v3 o, v3 d,
                                 v3 A, v3 B, float r)
    {
    t = HUGE_VALF;
    float t_sc = 0;
    float d_edge_sc = HUGE_VALF;
    v3 N_sc = mk(0, 0, 0);
    const v3 Oa = o - A;
    const v3 C = B - A;
    const float cdotc = dot(C, C);
    bool hit_A = false, hit_B = false, hit_cyl = false;
    hit_cyl = isect_cylinder(t_sc, N_sc, o, d, A, B, r);
    if (hit_cyl && t_sc < t)
        {
        t = t_sc;
        N = N_sc;
        }
    hit_A = isect_sphere(t_sc, d_edge_sc, N_sc, o, d, A, r);
    if (hit_A && (t_sc < t))
        {
        const v3 P_A = Oa + t_sc * d;
        if (dot(P_A, C) <= 0)
            {
            t = t_sc;
            N = N_sc;
            }
        else
            hit_A = false;
        }
    hit_B = isect_sphere(t_sc, d_edge_sc, N_sc, o, d, B, r);
    if (hit_B && (t_sc < t))
        {
        const v3 P_A = Oa + t_sc * d;
        if (dot(P_A, C) >= cdotc)
            {
            t = t_sc;
            N = N_sc;
            }
        else
            hit_B = false;
        }
    const v3 P = Oa + t * d;
    const float pdotc = dot(P, C);
    if (pdotc < cdotc / 2.0f)
        color_index = 0;
    else
        color_index = 1;
    const v3 C_vp = C - dot(C, d) * d;
    const v3 Oa_vp = Oa - dot(Oa, d) * d;
    const float oavpdotcvp = dot(Oa_vp, C_vp);
    const float cvpdotcvp = dot(C_vp, C_vp);
    if (oavpdotcvp < 0)
        d_edge = r - sqrtf(dot(Oa_vp, Oa_vp));
    else if (oavpdotcvp > cvpdotcvp)
        d_edge = r - sqrtf(dot(Oa_vp - C_vp, Oa_vp - C_vp));
    else
        {
        const v3 n_vp = Oa_vp - oavpdotcvp / cvpdotcvp * C_vp;
        d_edge = r - sqrtf(dot(n_vp, n_vp));
        }
    return hit_A || hit_B || hit_cyl;
    }

// IntersectTriangle.h:43-132
static bool isect_triangle(float& u, float& v, float& w, float& t, float& d_edge, v3& n, v3 p, v3 q,
                           v3 a, v3 b, v3 c)
    {
    v3 ab = b - a;
    v3 ac = c - a;
    v3 qp = p - q;
    n = cross(ab, ac);
    float d = dot(qp, n);
    if (d <= float(0.0))
        return false;
    v3 ap = p - a;
    t = dot(ap, n);
    if (t < float(0.0))
        return false;
    v3 e = cross(qp, ap);
    v = dot(ac, e);
    if (v < float(0.0) || v > d)
        return false;
    w = -dot(ab, e);
    if (w < float(0.0) || v + w > d)
        return false;
    float ood = float(1.0) / d;
    t *= ood;
    v *= ood;
    w *= ood;
    u = float(1.0) - v - w;
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
    v3 r_hit = p + t * (q - p);
    v3 vec = cross(edge, r_hit - pt);
    float dsq = dot(vec, vec) / dot(edge, edge);
    d_edge = sqrtf(dsq);
    return true;
    }

// GeometryMath.h:28-39
static inline float point_line_segment_distance(v2 p, v2 a, v2 b)
    {
    const v2 ab = b - a;
    const float L = dot(ab, ab);
    const float t = std::max(0.0f, std::min(1.0f, dot(p - a, ab) / L));
    const v2 projection = a + t * ab;
    const v2 v = p - projection;
    return sqrtf(dot(v, v));
    }
// cpu/GeometryPolygon.cc:116-142
static bool is_inside(float& min_d, v2 p, const std::vector<v2>& verts)
    {
    unsigned nvert = (unsigned)verts.size();
    min_d = FLT_MAX;
    unsigned i, j = nvert - 1;
    bool odd = false;
    for (i = 0; i < nvert; i++)
        {
        min_d = std::min(min_d, point_line_segment_distance(p, verts[i], verts[j]));
        if ((verts[i].y < p.y && verts[j].y >= p.y) || (verts[j].y < p.y && verts[i].y >= p.y))
            {
            if (verts[i].x + (p.y - verts[i].y) / (verts[j].y - verts[i].y) * (verts[j].x - verts[i].x)
                < p.x)
                odd = !odd;
            }
        j = i;
        }
    return odd;
    }

// Evaluate primitive `prim` of geometry `g`.  Returns true with a candidate
// hit (t > tnear) and its shading attributes.  This folds each reference
// intersect callback's acceptance test `tnear < t < tfar` into `tnear < t`;
// the `< tfar` part is the caller's total-order comparison.
static bool intersect_prim(const Geom& g, unsigned prim, v3 o, v3 d, float tnear, Hit& h)
    {
    switch (g.type)
        {
    case G_SPHERE: // cpu/GeometrySphere.cc:81-117
        {
        const v3 position = g.pos[prim];
        const float radius = g.radius[prim];
        float t = 0, de = 0;
        v3 N;
        bool hit = isect_sphere(t, de, N, o, d, position, radius);
        if (!(hit && tnear < t))
            return false;
        h.t = t;
        h.Ng = mk(o.x + t * d.x - position.x, o.y + t * d.y - position.y, o.z + t * d.z - position.z);
        h.d = de;
        h.color = g.color[prim];
        return true;
        }
    case G_CYLINDER: // cpu/GeometryCylinder.cc:76-118
        {
        const v3 A = g.points[prim * 2 + 0], B = g.points[prim * 2 + 1];
        float t = HUGE_VALF, de = HUGE_VALF;
        v3 N = mk(0, 0, 0);
        unsigned ci = 0;
        bool hit = isect_spherocylinder(t, de, N, ci, o, d, A, B, g.radius[prim]);
        if (!(hit && tnear < t))
            return false;
        h.t = t;
        h.Ng = N;
        h.d = de;
        h.color = g.color[prim * 2 + ci];
        return true;
        }
    case G_POLYHEDRON: // cpu/GeometryConvexPolyhedron.cc:140-335
        {
        int n_planes = (int)g.plane_normal.size();
        float t0 = -std::numeric_limits<float>::max();
        float t1 = std::numeric_limits<float>::max();
        const v3 pos_world = g.pos[prim];
        const quat q_world = g.orient[prim];
        v3 ray_dir_local = rotate(conj(q_world), d);
        v3 ray_org_local = rotate(conj(q_world), o - pos_world);
        v3 t0_n_local = mk(0, 0, 0), t0_p_local = mk(0, 0, 0);
        v3 t1_n_local = mk(0, 0, 0), t1_p_local = mk(0, 0, 0);
        int t0_plane_hit = 0, t1_plane_hit = 0;
        for (int i = 0; i < n_planes && t0 < t1; ++i)
            {
            v3 n = g.plane_normal[i];
            v3 p = g.plane_origin[i];
            float dd = -dot(n, p);
            float denom = dot(n, ray_dir_local);
            float t = -(dd + dot(n, ray_org_local)) / denom;
            if (fabs(denom) < 1e-5)
                {
                if (dot(ray_org_local - p, n) > 0)
                    return false;
                }
            else if (denom < 0)
                {
                if (t > t0)
                    {
                    t0 = t;
                    t0_n_local = n;
                    t0_p_local = p;
                    t0_plane_hit = i;
                    }
                }
            else
                {
                if (t < t1)
                    {
                    t1 = t;
                    t1_n_local = n;
                    t1_p_local = p;
                    t1_plane_hit = i;
                    }
                }
            }
        if (t0 > t1)
            return false;
        float t_hit;
        v3 n_hit, p_hit;
        int plane_hit;
        // entry plane when t0 is past tnear, otherwise the exit plane
        // (:217-249; a t0 beyond the current closest hit implies t1 is too)
        if (tnear < t0)
            {
            t_hit = t0;
            n_hit = t0_n_local;
            p_hit = t0_p_local;
            plane_hit = t0_plane_hit;
            }
        else if (tnear < t1)
            {
            t_hit = t1;
            n_hit = t1_n_local;
            p_hit = t1_p_local;
            plane_hit = t1_plane_hit;
            }
        else
            return false;
        h.t = t_hit;
        h.Ng = rotate(q_world, n_hit);
        h.color = lerp(g.color_by_face, g.color[prim], g.plane_color[plane_hit]);
        float min_d = std::numeric_limits<float>::max();
        v3 r_hit = ray_org_local + t_hit * ray_dir_local;
        for (int i = 0; i < n_planes; ++i)
            {
            v3 n = g.plane_normal[i];
            v3 p = g.plane_origin[i];
            v3 u = cross(n, n_hit);
            if (fabs(dot(u, u)) >= 1e-5)
                {
                int maxc;
                if (fabs(u.x) > fabs(u.y))
                    maxc = (fabs(u.x) > fabs(u.z)) ? 1 : 3;
                else
                    maxc = (fabs(u.y) > fabs(u.z)) ? 2 : 3;
                v3 x0;
                float d1 = -dot(n, p);
                float d2 = -dot(n_hit, p_hit);
                switch (maxc)
                    {
                case 1:
                    x0.x = 0;
                    x0.y = (d2 * n.z - d1 * n_hit.z) / u.x;
                    x0.z = (d1 * n_hit.y - d2 * n.y) / u.x;
                    break;
                case 2:
                    x0.x = (d1 * n_hit.z - d2 * n.z) / u.y;
                    x0.y = 0;
                    x0.z = (d2 * n.x - d1 * n_hit.x) / u.y;
                    break;
                default:
                    x0.x = (d2 * n.y - d1 * n_hit.y) / u.z;
                    x0.y = (d1 * n_hit.x - d2 * n.x) / u.z;
                    x0.z = 0;
                    }
                v3 view = div_rcp(-ray_dir_local, sqrtf(dot(ray_dir_local, ray_dir_local)));
                u = u - dot(u, view) * view;
                v3 w = x0 - r_hit;
                v3 w_perp = w - dot(w, view) * view;
                x0 = r_hit + w_perp;
                v3 vv = cross(u, x0 - r_hit);
                float dsq = dot(vv, vv) / dot(u, u);
                float dist = sqrtf(dsq);
                if (dist < min_d)
                    min_d = dist;
                }
            }
        h.d = min_d;
        return true;
        }
    case G_MESH: // cpu/GeometryMesh.cc:115-192
        {
        unsigned inst = prim / g.n_faces, face = prim % g.n_faces;
        const v3 p3 = g.pos[inst];
        const quat q_world = g.orient[inst];
        v3 ray_dir_local = rotate(conj(q_world), d);
        v3 ray_org_local = rotate(conj(q_world), o - p3);
        v3 v0 = g.verts[face * 3], v1 = g.verts[face * 3 + 1], v2_ = g.verts[face * 3 + 2];
        float u = 0, v = 0, w = 0, t = 0, de = 0;
        v3 n = mk(0, 0, 0);
        v3 qend = ray_org_local + ray_dir_local;
        if (!isect_triangle(u, v, w, t, de, n, ray_org_local, qend, v0, v1, v2_)
            && !isect_triangle(v, u, w, t, de, n, ray_org_local, qend, v1, v0, v2_))
            return false;
        if (!(tnear < t))
            return false;
        h.t = t;
        h.Ng = rotate(q_world, n);
        h.color = g.color[face * 3 + 0] * u + g.color[face * 3 + 1] * v + g.color[face * 3 + 2] * w;
        h.d = de;
        return true;
        }
    case G_POLYGON: // cpu/GeometryPolygon.cc:150-232
        {
        const v2 p2 = g.pos2[prim];
        const v3 pos_world = mk(p2.x, p2.y, 0.0f);
        const quat q_world = from_axis_angle(mk(0, 0, 1), g.angle[prim]);
        v3 ray_dir_local = rotate(conj(q_world), d);
        v3 ray_org_local = rotate(conj(q_world), o - pos_world);
        const v3 n = mk(0, 0, 1);
        const v3 p = mk(0, 0, 0);
        float dd = -dot(n, p);
        float denom = dot(n, ray_dir_local);
        float t_hit = -(dd + dot(n, ray_org_local)) / denom;
        if (fabs(denom) < 1e-5)
            return false;
        v3 r_hit = ray_org_local + t_hit * ray_dir_local;
        v2 r2 = v2 {r_hit.x, r_hit.y};
        float d_edge;
        bool inside = is_inside(d_edge, r2, g.verts2);
        if (inside)
            d_edge = -d_edge;
        if (d_edge > g.rounding_radius)
            return false;
        float min_d = g.rounding_radius - d_edge;
        if (!(tnear < t_hit))
            return false;
        v3 n_flip = (dot(n, ray_dir_local) < 0.0f) ? n : -n;
        h.t = t_hit;
        h.Ng = rotate(q_world, n_flip);
        h.color = g.color[prim];
        h.d = min_d;
        return true;
        }
        }
    return false;
    }

// ---------------------------------------------------------------- scene + BVH
struct PrimRef
    {
    unsigned geom; // index into Scene::geoms
    unsigned prim;
    };
struct Node
    {
    Box box;
    int left = -1, right = -1; // internal children
    unsigned first = 0, count = 0; // leaf range into refs
    };

struct Scene
    {
    UserCamera camera {};
    rgb bg {0, 0, 0};
    float bg_alpha = 0;
    Lights lights {};
    std::vector<Geom> geoms;
    std::vector<PrimRef> refs;
    std::vector<Box> ref_box;
    std::vector<Node> nodes;
    float cull_margin = 0.0f; // CULL_K * 3 * largest primitive half-extent
    bool brute = false;
    };

static Box box_union(const Box& a, const Box& b)
    {
    return Box {mk(std::min(a.lo.x, b.lo.x), std::min(a.lo.y, b.lo.y), std::min(a.lo.z, b.lo.z)),
                mk(std::max(a.hi.x, b.hi.x), std::max(a.hi.y, b.hi.y), std::max(a.hi.z, b.hi.z))};
    }

static int build_node(Scene& s, std::vector<unsigned>& idx, unsigned first, unsigned count,
                      std::vector<Box>& boxes)
    {
    int ni = (int)s.nodes.size();
    s.nodes.push_back(Node());
    Box b = boxes[idx[first]];
    for (unsigned k = 1; k < count; k++)
        b = box_union(b, boxes[idx[first + k]]);
    s.nodes[ni].box = b;
    if (count <= 4)
        {
        s.nodes[ni].first = first;
        s.nodes[ni].count = count;
        return ni;
        }
    // median split of centroids along the widest centroid axis
    v3 clo = mk(FLT_MAX, FLT_MAX, FLT_MAX), chi = mk(-FLT_MAX, -FLT_MAX, -FLT_MAX);
    for (unsigned k = 0; k < count; k++)
        {
        const Box& bb = boxes[idx[first + k]];
        v3 c = mk(0.5f * (bb.lo.x + bb.hi.x), 0.5f * (bb.lo.y + bb.hi.y), 0.5f * (bb.lo.z + bb.hi.z));
        clo = mk(std::min(clo.x, c.x), std::min(clo.y, c.y), std::min(clo.z, c.z));
        chi = mk(std::max(chi.x, c.x), std::max(chi.y, c.y), std::max(chi.z, c.z));
        }
    v3 ext = chi - clo;
    int axis = (ext.x >= ext.y && ext.x >= ext.z) ? 0 : (ext.y >= ext.z ? 1 : 2);
    auto key = [&](unsigned r)
    {
        const Box& bb = boxes[r];
        return axis == 0 ? bb.lo.x + bb.hi.x : (axis == 1 ? bb.lo.y + bb.hi.y : bb.lo.z + bb.hi.z);
    };
    unsigned mid = first + count / 2;
    std::nth_element(idx.begin() + first,
                     idx.begin() + mid,
                     idx.begin() + first + count,
                     [&](unsigned a, unsigned b) { return key(a) < key(b); });
    int l = build_node(s, idx, first, mid - first, boxes);
    int r = build_node(s, idx, mid, first + count - mid, boxes);
    s.nodes[ni].left = l;
    s.nodes[ni].right = r;
    return ni;
    }

static void commit(Scene& s)
    {
    s.refs.clear();
    s.ref_box.clear();
    s.nodes.clear();
    std::vector<Box> boxes;
    for (unsigned gi = 0; gi < s.geoms.size(); gi++)
        {
        const Geom& g = s.geoms[gi];
        if (!g.enabled)
            continue;
        for (unsigned p = 0; p < g.N; p++)
            {
            s.refs.push_back(PrimRef {gi, p});
            boxes.push_back(prim_bounds(g, p));
            }
        }
    s.cull_margin = 0.0f;
    if (s.refs.empty())
        return;
    float r_max = 0.0f;
    for (const Box& b : boxes)
        r_max = fmaxf(r_max, fmaxf(fmaxf(b.hi.x - b.lo.x, b.hi.y - b.lo.y), b.hi.z - b.lo.z) * 0.5f);
    s.cull_margin = CULL_K * (3.0f * r_max);
    std::vector<unsigned> idx(s.refs.size());
    for (unsigned k = 0; k < idx.size(); k++)
        idx[k] = k;
    s.nodes.reserve(2 * s.refs.size());
    build_node(s, idx, 0, (unsigned)idx.size(), boxes);
    // reorder refs into leaf order
    std::vector<PrimRef> refs2(s.refs.size());
    s.ref_box.resize(s.refs.size());
    for (unsigned k = 0; k < idx.size(); k++)
        {
        refs2[k] = s.refs[idx[k]];
        s.ref_box[k] = boxes[idx[k]];
        }
    s.refs.swap(refs2);
    }

struct HitRec
    {
    bool hit = false;
    float t = HUGE_VALF;
    unsigned geom_id = 0xffffffffu; // geomID (tie order)
    unsigned geom = 0; // index into Scene::geoms
    unsigned prim = 0xffffffffu;
    Hit attr;
    };

static inline bool better(float t, unsigned gid, unsigned prim, const HitRec& best)
    {
    if (!best.hit)
        return true;
    if (t != best.t)
        return t < best.t;
    if (gid != best.geom_id)
        return gid < best.geom_id;
    return prim < best.prim;
    }

static inline void test_ref(const Scene& s, unsigned r, v3 o, v3 d, v3 inv, float tnear, HitRec& best,
                            bool cull)
    {
    float tmin, tmax;
    slab(o, inv, s.ref_box[r], tmin, tmax);
    if (!(tmin <= tmax && tmax >= tnear))
        return;
    if (cull && best.hit && cull_bound(tmin, tmax, s.cull_margin) > best.t)
        return;
    const PrimRef& pr = s.refs[r];
    const Geom& g = s.geoms[pr.geom];
    Hit h;
    if (!intersect_prim(g, pr.prim, o, d, tnear, h))
        return;
    if (better(h.t, g.id, pr.prim, best))
        {
        best.hit = true;
        best.t = h.t;
        best.geom_id = g.id;
        best.geom = pr.geom;
        best.prim = pr.prim;
        best.attr = h;
        }
    }

static HitRec closest_hit(const Scene& s, v3 o, v3 d, float tnear)
    {
    HitRec best;
    if (s.refs.empty())
        return best;
    v3 inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    if (s.brute)
        {
        for (unsigned r = 0; r < s.refs.size(); r++)
            test_ref(s, r, o, d, inv, tnear, best, false);
        return best;
        }
    int stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp)
        {
        const Node& n = s.nodes[stack[--sp]];
        float tmin, tmax;
        slab(o, inv, n.box, tmin, tmax);
        if (!(tmin <= tmax && tmax >= tnear))
            continue;
        if (best.hit && cull_bound(tmin, tmax, s.cull_margin) > best.t)
            continue;
        if (n.left < 0)
            {
            for (unsigned k = 0; k < n.count; k++)
                test_ref(s, n.first + k, o, d, inv, tnear, best, true);
            continue;
            }
        // visit the nearer child first
        float l0, l1, r0, r1;
        slab(o, inv, s.nodes[n.left].box, l0, l1);
        slab(o, inv, s.nodes[n.right].box, r0, r1);
        if (l0 <= r0)
            {
            stack[sp++] = n.right;
            stack[sp++] = n.left;
            }
        else
            {
            stack[sp++] = n.left;
            stack[sp++] = n.right;
            }
        }
    return best;
    }

// ---------------------------------------------------------------- path tracer
// Per-pixel body of cpu/TracerPath.cc:96-249 with path_tracer_hit/miss
// (TracerPathMethods.h:36-168) and RayGen (RayGen.h:110-186).
struct PassParams
    {
    unsigned W, H, seed, n_samples, L;
    int highlight;
    rgb hw_color;
    };

static void path_pixel(const Scene& s, const Camera& cam, const Lights& lights, const PassParams& pp,
                       unsigned i, unsigned j, float* linear, uint8_t* srgb_out)
    {
    const unsigned key0 = j * pp.W + i;
    rgb result = mk(0, 0, 0);
    float a = 1.0f;
    v3 org, dir;
    generate_ray(cam, i, j, pp.n_samples, org, dir);
    const HitRec primary = closest_hit(s, org, dir, 1e-3f);
    for (unsigned ls = 0; ls < pp.L; ls++)
        {
        rgb att = mk(1.0f, 1.0f, 1.0f);
        bool done = false;
        v3 p_org = org, p_dir = dir;
        for (unsigned depth = 0;; depth++)
            {
            HitRec hr;
            v3 ro, rd;
            if (depth == 0)
                {
                hr = primary;
                ro = org;
                rd = dir;
                }
            else
                {
                ro = p_org;
                rd = p_dir;
                hr = closest_hit(s, ro, rd, 1e-3f);
                }
            if (hr.hit)
                {
                // path_tracer_hit, TracerPathMethods.h:79-168
                const Geom& g = s.geoms[hr.geom];
                const Material& m = (hr.attr.d >= g.outline_width) ? g.mat : g.outline;
                v3 Ng = hr.attr.Ng;
                v3 n = Ng * (1.0f / sqrtf(dot(Ng, Ng)));
                v3 v = -rd;
                if (is_solid(m))
                    {
                    result = result + att * get_color(m, hr.attr.color);
                    done = true;
                    }
                else
                    {
                    unsigned sidx = (pp.n_samples - 1) * pp.L + ls;
                    // RayGen::MISReflectionTransmission, RayGen.h:110-157
                    u4 ru = philox(u4 {{0, depth, sidx, RNG_MIS}}, key0, pp.seed);
                    v2 xi = v2 {u01(ru.v[0]), u01(ru.v[1])};
                    float choice_mis = u01(ru.v[2]);
                    float choice_trans = u01(ru.v[3]);
                    float factor = 1.0f;
                    v3 l;
                    bool transmit = (choice_trans <= m.spec_trans);
                    if (transmit)
                        l = -v;
                    else if (choice_mis <= 0.5f)
                        {
                        l = sample_diffuse(xi, v, n);
                        float pd = pdf_diffuse(l, n);
                        float pg = pdf_ggx(m, l, v, n);
                        float w_d = pd / (pd + pg);
                        factor = w_d / (0.5f * pd);
                        }
                    else
                        {
                        l = sample_ggx(m, xi, v, n);
                        float pd = pdf_diffuse(l, n);
                        float pg = pdf_ggx(m, l, v, n);
                        float w_g = pg / (pd + pg);
                        factor = w_g / (0.5f * pg);
                        }
                    if (transmit)
                        {
                        rgb tc = get_color(m, hr.attr.color);
                        tc = mk(sqrtf(tc.x), sqrtf(tc.y), sqrtf(tc.z));
                        att = att * tc;
                        }
                    else
                        {
                        float ndotl = dot(n, l);
                        if (dot(n, v) > 0.0f && ndotl > 0.0f)
                            att = att * (brdf(m, l, v, n, hr.attr.color) * ndotl * factor);
                        else
                            {
                            att = mk(0, 0, 0);
                            done = true;
                            }
                        }
                    p_org = ro + rd * hr.t;
                    p_dir = l;
                    // RayGen::shouldTerminatePath, RayGen.h:168-186
                    u4 rr = philox(u4 {{0, depth, sidx, RNG_RR}}, key0, pp.seed);
                    float p_continue = fmaxf(att.x, fmaxf(att.y, att.z));
                    p_continue = fminf(p_continue, 1.0f);
                    if (u01(rr.v[0]) > p_continue)
                        done = true;
                    else
                        att = div_true(att, p_continue);
                    }
                }
            else
                {
                // path_tracer_miss, TracerPathMethods.h:36-77
                if (depth == 0)
                    {
                    result = s.bg * float(pp.L);
                    a = s.bg_alpha;
                    }
                else
                    {
                    for (unsigned li = 0; li < lights.N; li++)
                        {
                        v3 l = lights.dir[li];
                        float half_angle = lights.theta[li];
                        float cos_half_angle = cosf(half_angle);
                        if (half_angle > PI_F / 2.0f)
                            half_angle = PI_F / 2.0f;
                        float ldotd = dot(l, rd);
                        if (ldotd >= cos_half_angle)
                            result = result + div_rcp(att * lights.color[li], sinf(half_angle));
                        }
                    }
                done = true;
                }
            if (done)
                break;
            }
        }
    // cpu/TracerPath.cc:231-249
    rgb rs = div_rcp(result, float(pp.L));
    float* px = linear + 4 * (size_t)key0;
    rgba old = rgba {px[0], px[1], px[2], px[3]};
    float q = 1.0f / float(pp.n_samples);
    rgba nw;
    nw.r = old.r + (rs.x - old.r) * q;
    nw.g = old.g + (rs.y - old.g) * q;
    nw.b = old.b + (rs.z - old.b) * q;
    nw.a = old.a + (a - old.a) * q;
    px[0] = nw.r;
    px[1] = nw.g;
    px[2] = nw.b;
    px[3] = nw.a;
    if (!pp.highlight || (nw.r <= 1.0f && nw.g <= 1.0f && nw.b <= 1.0f))
        srgb(nw, srgb_out + 4 * (size_t)key0);
    else
        srgb(rgba {pp.hw_color.x, pp.hw_color.y, pp.hw_color.z, nw.a}, srgb_out + 4 * (size_t)key0);
    }

// ---------------------------------------------------------------- preview
// Per-pixel body of cpu/TracerDirect.cc:77-218.
static void preview_pixel(const Scene& s, const Camera& cam, const Lights& lights, unsigned aa_n,
                          int highlight, rgb hw_color, unsigned i, unsigned j, float* linear,
                          uint8_t* srgb_out)
    {
    rgba avg {0, 0, 0, 0};
    for (unsigned sample = 0; sample < aa_n * aa_n; sample++)
        {
        v3 org, dir;
        generate_ray(cam, i, j, sample, org, dir);
        HitRec hr = closest_hit(s, org, dir, 0.0f);
        rgb c = s.bg;
        float a = s.bg_alpha;
        if (hr.hit)
            {
            v3 n = hr.attr.Ng;
            n = div_true(n, sqrtf(dot(n, n)));
            v3 v = div_rcp(-dir, sqrtf(dot(dir, dir)));
            const Geom& g = s.geoms[hr.geom];
            const Material& m = (hr.attr.d >= g.outline_width) ? g.mat : g.outline;
            if (is_solid(m))
                c = get_color(m, hr.attr.color);
            else
                {
                c = mk(0, 0, 0);
                for (unsigned li = 0; li < lights.N; li++)
                    {
                    v3 l = lights.dir[li];
                    v3 r = -v + (2.0f * n * dot(n, v));
                    float half_angle = lights.theta[li];
                    float cos_half_angle = cosf(half_angle);
                    float ldotr = dot(l, r);
                    if (ldotr < cos_half_angle)
                        {
                        v3 ax = cross(l, r);
                        ax = div_rcp(ax, sqrtf(dot(ax, ax)));
                        quat q = from_axis_angle(ax, -acosf(ldotr) + half_angle);
                        r = rotate(q, r);
                        }
                    rgb f_d;
                    float ndotl = dot(n, l);
                    if (ndotl >= 0.0f)
                        f_d = brdf_diffuse(m, l, v, n, hr.attr.color) * ndotl;
                    else
                        f_d = mk(0, 0, 0);
                    rgb f_s;
                    if (dot(n, r) >= 0.0f)
                        f_s = brdf_specular(m, r, v, n, hr.attr.color, half_angle) * dot(n, r);
                    else
                        f_s = mk(0, 0, 0);
                    c = c + (f_d + f_s) * PI_F * lights.color[li];
                    }
                }
            a = 1.0;
            }
        avg.r += c.x;
        avg.g += c.y;
        avg.b += c.z;
        avg.a += a;
        }
    float q = 1.0f / float(aa_n * aa_n);
    rgba outp {avg.r * q, avg.g * q, avg.b * q, avg.a * q};
    size_t px = (size_t)j * cam.W + i;
    linear[4 * px + 0] = outp.r;
    linear[4 * px + 1] = outp.g;
    linear[4 * px + 2] = outp.b;
    linear[4 * px + 3] = outp.a;
    if (!highlight || (outp.r <= 1.0f && outp.g <= 1.0f && outp.b <= 1.0f))
        srgb(outp, srgb_out + 4 * px);
    else
        srgb(rgba {hw_color.x, hw_color.y, hw_color.z, outp.a}, srgb_out + 4 * px);
    }

// 4x4 pixel tiles handed out dynamically to a pool of threads
// (cpu/TracerPath.cc:78-90 and cpu/Device.cc:17-40 use TBB the same way).
template<class F> static void parallel_tiles(unsigned W, unsigned H, int nthreads, F f)
    {
    const unsigned TS = 4;
    unsigned ntx = (W + TS - 1) / TS, nty = (H + TS - 1) / TS;
    unsigned ntiles = ntx * nty;
    std::atomic<unsigned> next {0};
    auto worker = [&]()
    {
        for (;;)
            {
            unsigned t = next.fetch_add(1);
            if (t >= ntiles)
                return;
            unsigned ty = t / ntx, tx = t - ty * ntx;
            unsigned x0 = tx * TS, y0 = ty * TS;
            unsigned x1 = std::min(x0 + TS, W), y1 = std::min(y0 + TS, H);
            for (unsigned j = y0; j < y1; j++)
                for (unsigned i = x0; i < x1; i++)
                    f(i, j);
            }
    };
    if (nthreads <= 0)
        nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
    std::vector<std::thread> pool;
    for (int k = 1; k < nthreads; k++)
        pool.emplace_back(worker);
    worker();
    for (auto& th : pool)
        th.join();
    }

    } // namespace orc

// ====================================================================== C ABI
using namespace orc;

static v3 ld3(const float* p)
    {
    return mk(p[0], p[1], p[2]);
    }
static Material ldmat(const float* m)
    {
    Material r;
    r.solid = m[0];
    r.color = mk(m[1], m[2], m[3]);
    r.primitive_color_mix = m[4];
    r.roughness = m[5];
    r.specular = m[6];
    r.metal = m[7];
    r.spec_trans = m[8];
    return r;
    }

extern "C"
    {
    // ---- RNG
    void oracle_philox(const uint32_t* in, size_t n, uint32_t* out)
        {
        for (size_t k = 0; k < n; k++)
            {
            const uint32_t* c = in + 6 * k;
            u4 r = philox(u4 {{c[0], c[1], c[2], c[3]}}, c[4], c[5]);
            for (int q = 0; q < 4; q++)
                out[4 * k + q] = r.v[q];
            }
        }
    void oracle_u01(const uint32_t* in, size_t n, float* out)
        {
        for (size_t k = 0; k < n; k++)
            out[k] = u01(in[k]);
        }

    // ---- scene construction
    void* oracle_scene_new()
        {
        return new Scene();
        }
    void oracle_scene_free(void* s)
        {
        delete (Scene*)s;
        }
    void oracle_scene_set_camera(void* sp,
                                 const float* pos,
                                 const float* look,
                                 const float* up,
                                 float h,
                                 float f,
                                 float f_stop,
                                 float focus,
                                 int model)
        {
        Scene& s = *(Scene*)sp;
        s.camera.position = ld3(pos);
        s.camera.look_at = ld3(look);
        s.camera.up = ld3(up);
        s.camera.h = h;
        s.camera.f = f;
        s.camera.f_stop = f_stop;
        s.camera.focus_distance = focus;
        s.camera.model = model;
        }
    void oracle_scene_set_background(void* sp, const float* c, float alpha)
        {
        Scene& s = *(Scene*)sp;
        s.bg = ld3(c);
        s.bg_alpha = alpha;
        }
    void oracle_scene_set_lights(void* sp,
                                 unsigned N,
                                 const float* dir,
                                 const float* color,
                                 const float* theta)
        {
        Scene& s = *(Scene*)sp;
        s.lights.N = N;
        for (unsigned i = 0; i < N && i < 4; i++)
            {
            s.lights.dir[i] = ld3(dir + 3 * i);
            s.lights.color[i] = ld3(color + 3 * i);
            s.lights.theta[i] = theta[i];
            }
        }
    void oracle_scene_set_brute(void* sp, int brute)
        {
        ((Scene*)sp)->brute = brute != 0;
        }
    static Geom& new_geom(Scene& s,
                          int type,
                          unsigned id,
                          int enabled,
                          const float* mat,
                          const float* outline,
                          float ow)
        {
        s.geoms.emplace_back();
        Geom& g = s.geoms.back();
        g.type = type;
        g.id = id;
        g.enabled = enabled != 0;
        g.mat = ldmat(mat);
        g.outline = ldmat(outline);
        g.outline_width = ow;
        return g;
        }
    void oracle_add_sphere(void* sp,
                           unsigned id,
                           int en,
                           const float* mat,
                           const float* outl,
                           float ow,
                           unsigned N,
                           const float* pos,
                           const float* radius,
                           const float* color)
        {
        Geom& g = new_geom(*(Scene*)sp, G_SPHERE, id, en, mat, outl, ow);
        g.N = N;
        for (unsigned k = 0; k < N; k++)
            {
            g.pos.push_back(ld3(pos + 3 * k));
            g.radius.push_back(radius[k]);
            g.color.push_back(ld3(color + 3 * k));
            }
        }
    void oracle_add_cylinder(void* sp,
                             unsigned id,
                             int en,
                             const float* mat,
                             const float* outl,
                             float ow,
                             unsigned N,
                             const float* points,
                             const float* radius,
                             const float* color)
        {
        Geom& g = new_geom(*(Scene*)sp, G_CYLINDER, id, en, mat, outl, ow);
        g.N = N;
        for (unsigned k = 0; k < 2 * N; k++)
            {
            g.points.push_back(ld3(points + 3 * k));
            g.color.push_back(ld3(color + 3 * k));
            }
        for (unsigned k = 0; k < N; k++)
            g.radius.push_back(radius[k]);
        }
    // normals are normalized as in cpu/GeometryConvexPolyhedron.cc:83-92
    void oracle_add_polyhedron(void* sp,
                               unsigned id,
                               int en,
                               const float* mat,
                               const float* outl,
                               float ow,
                               unsigned N,
                               const float* pos,
                               const float* orient,
                               const float* color,
                               unsigned P,
                               const float* origins,
                               const float* normals,
                               const float* pcolors,
                               float r,
                               float color_by_face)
        {
        Geom& g = new_geom(*(Scene*)sp, G_POLYHEDRON, id, en, mat, outl, ow);
        g.N = N;
        for (unsigned k = 0; k < N; k++)
            {
            g.pos.push_back(ld3(pos + 3 * k));
            g.orient.push_back(quat {orient[4 * k], ld3(orient + 4 * k + 1)});
            g.color.push_back(ld3(color + 3 * k));
            }
        for (unsigned k = 0; k < P; k++)
            {
            v3 n = ld3(normals + 3 * k);
            n = div_rcp(n, sqrtf(dot(n, n)));
            g.plane_origin.push_back(ld3(origins + 3 * k));
            g.plane_normal.push_back(n);
            g.plane_color.push_back(ld3(pcolors + 3 * k));
            }
        g.poly_radius = r;
        g.color_by_face = color_by_face;
        }
    void oracle_add_mesh(void* sp,
                         unsigned id,
                         int en,
                         const float* mat,
                         const float* outl,
                         float ow,
                         unsigned Ninst,
                         const float* pos,
                         const float* orient,
                         unsigned nverts,
                         const float* verts,
                         const float* vcolors)
        {
        Geom& g = new_geom(*(Scene*)sp, G_MESH, id, en, mat, outl, ow);
        g.n_faces = nverts / 3;
        g.N = Ninst * g.n_faces;
        for (unsigned k = 0; k < Ninst; k++)
            {
            g.pos.push_back(ld3(pos + 3 * k));
            g.orient.push_back(quat {orient[4 * k], ld3(orient + 4 * k + 1)});
            }
        for (unsigned k = 0; k < nverts; k++)
            {
            g.verts.push_back(ld3(verts + 3 * k));
            g.color.push_back(ld3(vcolors + 3 * k));
            }
        }
    // xy radius as in cpu/GeometryPolygon.cc:55-65
    void oracle_add_polygon(void* sp,
                            unsigned id,
                            int en,
                            const float* mat,
                            const float* outl,
                            float ow,
                            unsigned N,
                            const float* pos2,
                            const float* angle,
                            const float* color,
                            unsigned nv,
                            const float* verts2,
                            float rounding)
        {
        Geom& g = new_geom(*(Scene*)sp, G_POLYGON, id, en, mat, outl, ow);
        g.N = N;
        for (unsigned k = 0; k < N; k++)
            {
            g.pos2.push_back(v2 {pos2[2 * k], pos2[2 * k + 1]});
            g.angle.push_back(angle[k]);
            g.color.push_back(ld3(color + 3 * k));
            }
        g.xy_radius = 0;
        for (unsigned k = 0; k < nv; k++)
            {
            v2 p0 {verts2[2 * k], verts2[2 * k + 1]};
            g.verts2.push_back(p0);
            g.xy_radius = std::max(g.xy_radius, sqrtf(dot(p0, p0)));
            }
        g.rounding_radius = rounding;
        g.xy_radius += rounding;
        }
    void oracle_scene_commit(void* sp)
        {
        commit(*(Scene*)sp);
        }
    unsigned oracle_scene_num_prims(void* sp)
        {
        return (unsigned)((Scene*)sp)->refs.size();
        }

    // ---- rendering
    // One Path pass (one camera sample per pixel), cpu/TracerPath.cc:56-256.
    // linear is updated in place with the Welford mean; srgb is overwritten.
    void oracle_path_pass(void* sp,
                          unsigned W,
                          unsigned H,
                          unsigned seed,
                          unsigned n_samples,
                          unsigned L,
                          float* linear,
                          uint8_t* srgb_out,
                          int highlight,
                          const float* hw_color,
                          int nthreads)
        {
        const Scene& s = *(Scene*)sp;
        Camera cam = make_camera(s.camera, W, H, seed, true);
        Lights lights = lights_to_world(s.lights, cam);
        PassParams pp {W, H, seed, n_samples, L, highlight, ld3(hw_color)};
        parallel_tiles(W,
                       H,
                       nthreads,
                       [&](unsigned i, unsigned j)
                       { path_pixel(s, cam, lights, pp, i, j, linear, srgb_out); });
        }
    // Preview, cpu/TracerDirect.cc:33-225 (f_stop forced to infinity)
    void oracle_preview(void* sp,
                        unsigned W,
                        unsigned H,
                        unsigned seed,
                        unsigned aa_n,
                        float* linear,
                        uint8_t* srgb_out,
                        int highlight,
                        const float* hw_color,
                        int nthreads)
        {
        const Scene& s = *(Scene*)sp;
        UserCamera uc = s.camera;
        uc.f_stop = std::numeric_limits<float>::infinity();
        Camera cam = make_camera(uc, W, H, seed, aa_n != 1);
        Lights lights = lights_to_world(s.lights, cam);
        rgb hw = ld3(hw_color);
        parallel_tiles(W,
                       H,
                       nthreads,
                       [&](unsigned i, unsigned j)
                       { preview_pixel(s, cam, lights, aa_n, highlight, hw, i, j, linear, srgb_out); });
        }

    // ---- diagnostics used by the parity tests
    // Camera rays for every pixel: out[6*(j*W+i)] = org.xyz, dir.xyz
    void oracle_camera_rays(void* sp,
                            unsigned W,
                            unsigned H,
                            unsigned seed,
                            unsigned sample,
                            int aa,
                            float* out)
        {
        const Scene& s = *(Scene*)sp;
        Camera cam = make_camera(s.camera, W, H, seed, aa != 0);
        for (unsigned j = 0; j < H; j++)
            for (unsigned i = 0; i < W; i++)
                {
                v3 o, d;
                generate_ray(cam, i, j, sample, o, d);
                float* p = out + 6 * ((size_t)j * W + i);
                p[0] = o.x;
                p[1] = o.y;
                p[2] = o.z;
                p[3] = d.x;
                p[4] = d.y;
                p[5] = d.z;
                }
        }
    // Closest hit for arbitrary rays (6 floats each).  geom_id = 0xffffffff on
    // a miss.  attr[7*k] = Ng.xyz, d_edge, color.rgb
    void oracle_trace(void* sp,
                      const float* rays,
                      size_t n,
                      float tnear,
                      unsigned* geom_id,
                      unsigned* prim,
                      float* t,
                      float* attr,
                      int nthreads)
        {
        const Scene& s = *(Scene*)sp;
        if (nthreads <= 0)
            nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
        std::atomic<size_t> next {0};
        auto worker = [&]()
        {
            for (;;)
                {
                size_t k0 = next.fetch_add(256);
                if (k0 >= n)
                    return;
                for (size_t k = k0; k < std::min(n, k0 + 256); k++)
                    {
                    const float* r = rays + 6 * k;
                    HitRec h = closest_hit(s, ld3(r), ld3(r + 3), tnear);
                    geom_id[k] = h.hit ? h.geom_id : 0xffffffffu;
                    prim[k] = h.hit ? h.prim : 0xffffffffu;
                    t[k] = h.hit ? h.t : HUGE_VALF;
                    if (attr)
                        {
                        float* a = attr + 7 * k;
                        a[0] = h.attr.Ng.x;
                        a[1] = h.attr.Ng.y;
                        a[2] = h.attr.Ng.z;
                        a[3] = h.attr.d;
                        a[4] = h.attr.color.x;
                        a[5] = h.attr.color.y;
                        a[6] = h.attr.color.z;
                        }
                    }
                }
        };
        std::vector<std::thread> pool;
        for (int k = 1; k < nthreads; k++)
            pool.emplace_back(worker);
        worker();
        for (auto& th : pool)
            th.join();
        }
    // Count of candidate primitives whose reported t equals the winner's t
    // exactly (exact ties), for the hit of each ray.  A primitive whose
    // computed distance equals the winner's lies in no box whose cull bound
    // (a lower bound of the computed distances of its primitives) exceeds that
    // distance, so the hierarchy is searched with that bound; a brute-force
    // scene tests every primitive.
    void oracle_count_ties(void* sp, const float* rays, size_t n, float tnear, unsigned* ties)
        {
        const Scene& s = *(Scene*)sp;
        auto count_one = [&](size_t k)
        {
            const float* r = rays + 6 * k;
            v3 o = ld3(r), d = ld3(r + 3);
            HitRec h = closest_hit(s, o, d, tnear);
            unsigned c = 0;
            if (!h.hit)
                return c;
            v3 inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
            auto test = [&](unsigned q)
            {
                float tmin, tmax;
                slab(o, inv, s.ref_box[q], tmin, tmax);
                if (!(tmin <= tmax && tmax >= tnear))
                    return;
                Hit hh;
                const PrimRef& pr = s.refs[q];
                if (intersect_prim(s.geoms[pr.geom], pr.prim, o, d, tnear, hh) && hh.t == h.t)
                    c++;
            };
            if (s.brute)
                {
                for (unsigned q = 0; q < s.refs.size(); q++)
                    test(q);
                return c;
                }
            std::vector<int> stack;
            stack.push_back(0);
            while (!stack.empty())
                {
                const Node& nd = s.nodes[stack.back()];
                stack.pop_back();
                float tmin, tmax;
                slab(o, inv, nd.box, tmin, tmax);
                if (!(tmin <= tmax && tmax >= tnear))
                    continue;
                if (cull_bound(tmin, tmax, s.cull_margin) > h.t)
                    continue;
                if (nd.left < 0)
                    {
                    for (unsigned q = 0; q < nd.count; q++)
                        test(nd.first + q);
                    continue;
                    }
                stack.push_back(nd.left);
                stack.push_back(nd.right);
                }
            return c;
        };
        const int nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
        std::atomic<size_t> next {0};
        auto worker = [&]()
        {
            for (;;)
                {
                const size_t k0 = next.fetch_add(64);
                if (k0 >= n)
                    return;
                for (size_t k = k0; k < std::min(n, k0 + 64); k++)
                    ties[k] = count_one(k);
                }
        };
        std::vector<std::thread> pool;
        for (int k = 1; k < nthreads; k++)
            pool.emplace_back(worker);
        worker();
        for (auto& th : pool)
            th.join();
        }
    // FP64 sphere distance: the root of |o + t d - p|^2 = r^2 (d taken as
    // given, a = d.d) nearest past tnear, with the discriminant in the stable
    // Lagrange form a r^2 - |v x d|^2, evaluated in double from the float
    // inputs.  NaN on a miss.
    double oracle_sphere_t64(const float* o, const float* d, const float* p, float r, double tnear)
        {
        double dx = d[0], dy = d[1], dz = d[2];
        double a = dx * dx + dy * dy + dz * dz;
        double vx = (double)p[0] - o[0], vy = (double)p[1] - o[1], vz = (double)p[2] - o[2];
        double b = vx * dx + vy * dy + vz * dz;
        double cx = vy * dz - vz * dy, cy = vz * dx - vx * dz, cz = vx * dy - vy * dx;
        double disc = a * (double)r * (double)r - (cx * cx + cy * cy + cz * cz);
        if (disc < 0)
            return NAN;
        double s = std::sqrt(disc);
        double t0 = (b - s) / a, t1 = (b + s) / a;
        if (t0 > tnear)
            return t0;
        if (t1 > tnear)
            return t1;
        return NAN;
        }
    // The checker's cosf / sinf of lobe and aperture angles, for n arguments.
    void oracle_cr_sincosf(const float* x, size_t n, float* s, float* c)
        {
        for (size_t k = 0; k < n; k++)
            {
            s[k] = cr_sinf(x[k]);
            c[k] = cr_cosf(x[k]);
            }
        }
    // oracle_sphere_t64 for n rays (6 floats each) against the sphere
    // (xyz[3k..3k+2], r[k]) each one hit.
    void oracle_sphere_t64_batch(const float* rays, const float* xyz, const float* r, size_t n, double tnear,
                                 double* out)
        {
        for (size_t k = 0; k < n; k++)
            out[k] = oracle_sphere_t64(rays + 6 * k, rays + 6 * k + 3, xyz + 3 * k, r[k], tnear);
        }
    }
