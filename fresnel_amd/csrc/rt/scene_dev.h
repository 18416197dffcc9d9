// scene_dev.h -- device-resident scene, BVH and job records.
#pragma once

#include "dmath.h"
#include "fresnel_hip.h"

namespace fh
{
enum GeomType : uint32_t
    {
    G_SPHERE = FH_GEOM_SPHERE,
    G_CYLINDER = FH_GEOM_CYLINDER,
    G_POLYGON = FH_GEOM_POLYGON,
    G_POLYHEDRON = FH_GEOM_CONVEX_POLYHEDRON,
    G_MESH = FH_GEOM_MESH
    };

// Material record (fresnel/common/Material.h:47-56 field order)
struct Mat
    {
    float solid;
    float cr, cg, cb;
    float mix;
    float roughness;
    float specular;
    float metal;
    float spec_trans;
    };

// One geometry slot of a scene.  The slot index is the geometry id: it
// orders exact-t ties and selects the materials, as Embree's geomID does in
// the reference (fresnel/cpu/Scene.h:55-107).
struct GeomDev
    {
    uint32_t type;
    uint32_t n_prims;     // sphere/cyl/polyhedron/polygon: N; mesh: instances * faces
    uint32_t n_aux;       // polyhedron: planes; mesh: faces; polygon: vertices
    uint32_t pad0;
    float outline_width;
    float color_by_face;  // polyhedron
    float rounding;       // polygon rounding radius
    float xy_radius;      // polygon bounds radius (vertex radius + rounding)
    Mat mat;
    Mat outline;
    float pad1[2];
    // primitive attributes (layout per type, packed on upload)
    //   sphere:     a[i] = (x, y, z, r)
    //   cylinder:   a[i] = (Ax, Ay, Az, r), b[i] = (Bx, By, Bz, 0)
    //   polyhedron: a[i] = (x, y, z, 0),    b[i] = quaternion (s, x, y, z)
    //   mesh:       a[inst] = (x, y, z, 0), b[inst] = quaternion
    //   polygon:    a[i] = (x, y, 0, 0),    b[i] = quaternion about +z from the angle
    const float4* a;
    const float4* b;
    const float* color;    // the reference's color array layout
    const float4* aux0;    // polyhedron plane (nx, ny, nz, -dot(n, p)); mesh vertices; polygon vertices
    const float4* aux1;    // polyhedron plane origins
    const float* aux_color; // polyhedron plane colors
    };

// BVH2 node with both children's boxes (64 bytes).  A child is an inner node
// index, a leaf (LEAF bit | local primitive index, geometry slot in g), or
// EMPTY with a NaN box that every slab test rejects.
// Device work-counter slots (64-bit each): rays traced, BVH nodes visited
// and primitive tests of the primary, secondary and tail traversals; paths
// stopped at the depth cap; tile-list entries examined by primary rays (per
// ray, each entry a wavefront reads counts once per lane), then those of
// spherocylinders, convex polyhedra and mesh triangles; ray-queue segment
// overflow flag; traversal-stack overflow flag.
constexpr int CTR_SECONDARY = 0, CTR_PRIMARY = 3, CTR_TAIL = 6, CTR_CAPPED = 9, CTR_TILE_ENTRIES = 10,
              CTR_QUEUE_OVERFLOW = 14, CTR_OVERFLOW = 15, N_CTR = 16;

constexpr uint32_t LEAF = 0x80000000u;
constexpr uint32_t EMPTY = 0xffffffffu;
// Tile-list entry reference of a hierarchy subtree (NODE_REF | its root node);
// other references are indices into the primitive bounds arrays.
constexpr uint32_t NODE_REF = 0x80000000u;
struct __attribute__((aligned(64))) Node
    {
    float4 a_lo;  // child 0: lo.xyz, hi.x
    float4 a_hi;  // child 0: hi.y, hi.z | child 1: lo.x, lo.y
    float4 b_box; // child 1: lo.z, hi.x, hi.y, hi.z
    uint4 link;   // c0, c1, g0, g1
    };

// 4-wide node (128 bytes): two child pairs, each laid out as a binary node
// (children 2p and 2p + 1 of pair p: boxes, references, geometry slots), so a
// search step loads and tests one pair at a time.  A child reference is a
// 4-wide node index, LEAF | local primitive index (geometry slot alongside)
// or EMPTY with a NaN box.  In a spheres-only scene (SceneDev::sphere_leaves)
// every child holds a centre c in lo.xyz and half extents h in hi.xyz, and its
// box is c -+ h: a leaf child its sphere (h = radius on every axis, so the box
// is exactly the builder's bounds), an inner child a box that contains the
// exact one after rounding.
struct __attribute__((aligned(128))) Node4
    {
    Node pair[2];
    };

struct SceneDev
    {
    const Node* nodes;
    const Node4* nodes4; // root: node 0 (as for nodes) unless root is EMPTY
    const GeomDev* geoms;
    uint32_t root;      // EMPTY when the scene has no primitive
    uint32_t sphere_leaves; // 4-wide leaf children hold spheres (spheres-only scene)
    uint32_t n_geoms;
    uint32_t n_prims;
    float cull_margin;  // CULL_K * 3 * largest primitive half-extent
    };

// Camera constants derived on the host from the UserCamera
// (fresnel/common/Camera.h:92-159).
struct CamDev
    {
    v3 p, u, v, w;
    float focal_h, focal_d, a;
    int32_t model;
    // perspective with aperture 0 and no component of p equal to -0: the lens
    // offset is a signed zero and every ray origin is exactly p
    int32_t pinhole;
    };

// Lights in scene space plus per-light constants evaluated once on the host:
// cos(theta) and 1/sin(min(theta, pi/2)) (fresnel/common/TracerPathMethods.h:57-71),
// tan(min(theta, 0.99 pi/2)) for the preview (fresnel/common/Material.h:187-192).
struct LightsDev
    {
    v3 dir[4];
    v3 color[4];
    float theta[4];
    float cos_theta[4];
    float inv_sin[4];
    float tan_clamped[4];
    uint32_t N;
    };

// One render job: one (scene, tracer) pair for one batch of camera samples.
struct JobDev
    {
    SceneDev scene;
    CamDev cam;
    LightsDev lights;
    v3 bg;
    float bg_alpha;
    uint32_t W, H, npix;
    float xs_off;         // 0.5f * float(W) / float(H)
    uint32_t seed;
    uint32_t n_base;      // passes accumulated before this batch
    uint32_t K;           // camera samples in this batch
    uint32_t L;           // light samples per camera sample
    uint32_t highlight;
    v3 hw_color;
    uint32_t ray_base;    // first primary ray of this job in the batch
    FastDiv div_chunk;    // division by 64 K (primary rays per 64-pixel chunk)
    FastDiv div_tiles;    // division by (W / 8) (full 8x8 tiles per row)
    FastDiv div_w;        // division by W
    uint64_t rec_base;    // first termination record of this job
    double inv_h;         // 1 / H
    float4* linear;       // running mean (h, w) RGBA
    uchar4* srgb;         // sRGB output (h, w) RGBA
    // per-tile candidate lists of the primary rays (kernels/tiles.hip), used
    // when use_tiles is set: entries [tile_start[t], tile_end[t]) of tile
    // t = (y / 8) * tiles_x + x / 8, 8 words each (k_tile_pack), ascending in
    // their hit-distance lower bound; prim_lo/hi .w carry the geometry slot
    // and primitive of a box entry's index
    const uint32_t* tile_start;
    const uint32_t* tile_end;
    const uint4* tile_entries;
    const float4* prim_lo;
    const float4* prim_hi;
    uint32_t tiles_x, tiles_y;
    uint32_t use_tiles;
    uint32_t pad_tiles;
    };

// Termination record of one light-sample path.
//   kind 0: no contribution
//   kind 1: add (x, y, z)             solid emitter, TracerPathMethods.h:128-134
//   kind 2: lights in bits 2..5 with attenuation (x, y, z), TracerPathMethods.h:55-73
//   kind 3: primary miss (the whole camera sample is background), :44-49
constexpr uint32_t REC_NONE = 0, REC_ADD = 1, REC_LIGHTS = 2, REC_BG = 3;
    } // namespace fh
