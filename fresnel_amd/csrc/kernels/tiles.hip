// tiles.hip -- per-tile candidate lists for primary rays.
//
// For a pinhole perspective camera (aperture 0) or an orthographic camera,
// every primary ray of an 8x8 pixel tile passes through the tile's pixel
// footprint (widened by the anti-aliasing jitter of at most 0.7072 pixel).  A
// primitive whose bounding box projects outside that footprint cannot pass the
// slab test of any of those rays, so a list of the primitives whose projected
// box overlaps the tile, sorted by a lower bound of their computed hit
// distances, replaces the BVH search for these rays: each lane tests the list
// in order (slab test on the primitive's bounds, then its intersector) and the
// wavefront stops once every lane's closest hit lies below the next entry's
// bound.  The closest hit is the same (t, geometry, primitive) minimum over
// the primitives whose bounds pass the slab test as in the BVH search.
#include "rt/isect.h"

namespace fh
{
constexpr uint32_t TILE = 8;

// Projection constants of a camera (host-derived from CamDev).
struct TileCam
    {
    v3 p, u, v, w;     // camera position and basis (w points backwards)
    float scale;       // perspective: focal_d / focal_h; orthographic: 1 / focal_h
    float W, H;        // image size
    int32_t model;     // 0 orthographic, 1 perspective
    uint32_t tiles_x, tiles_y;
    float margin;      // CULL_K * 3 * largest primitive half-extent
    float plane_r;     // orthographic: largest distance of a ray origin from p; perspective: 0
    };

// Tile rectangle [tx0, tx1] x [ty0, ty1] of a primitive box and the lower
// bound q of its computed hit distance for rays of those tiles.  Returns false
// when no ray of a full tile can reach the box; *straddle is set when a
// perspective camera's plane cuts the box (such scenes use the BVH search).
__device__ __forceinline__ bool tile_rect(const TileCam& c, float4 lo, float4 hi, int& tx0, int& tx1, int& ty0,
                                          int& ty1, float& q, bool& straddle)
    {
    float imin = __builtin_huge_valf(), imax = -__builtin_huge_valf();
    float jmin = __builtin_huge_valf(), jmax = -__builtin_huge_valf();
    float zmin = __builtin_huge_valf(), zmax = -__builtin_huge_valf();
    float dmax2 = 0.0f;
    for (int k = 0; k < 8; k++)
        {
        const v3 X = mk3((k & 1) ? hi.x : lo.x, (k & 2) ? hi.y : lo.y, (k & 4) ? hi.z : lo.z);
        const v3 d = X - c.p;
        const float z = -dot(d, c.w);
        zmin = fminf(zmin, z);
        zmax = fmaxf(zmax, z);
        dmax2 = fmaxf(dmax2, dot(d, d));
        const float s = (c.model == 1) ? c.scale / z : c.scale;
        const float xs = dot(d, c.u) * s;
        const float ys = dot(d, c.v) * s;
        const float i_f = xs * c.H + 0.5f * c.W;
        const float j_f = (0.5f - ys) * c.H;
        imin = fminf(imin, i_f);
        imax = fmaxf(imax, i_f);
        jmin = fminf(jmin, j_f);
        jmax = fmaxf(jmax, j_f);
        }
    if (zmax <= 0.0f)
        return false; // behind the camera (rays advance along -w)
    if (c.model == 1 && zmin <= 0.0f)
        {
        straddle = true;
        return false;
        }
    // pixel i sees jittered positions in (i + 0.5 - 0.7072, i + 0.5 + 0.7072];
    // one more pixel covers rounding of the projection and of the camera rays
    const float lim = 1.0e6f;
    const int x0 = (int)floorf(fmaxf(imin - 2.25f, -lim)), x1 = (int)floorf(fminf(imax + 1.25f, lim));
    const int y0 = (int)floorf(fmaxf(jmin - 2.25f, -lim)), y1 = (int)floorf(fminf(jmax + 1.25f, lim));
    const int nx = (int)(c.tiles_x * TILE), ny = (int)(c.tiles_y * TILE);
    if (x1 < 0 || y1 < 0 || x0 >= nx || y0 >= ny)
        return false;
    tx0 = max(x0, 0) / (int)TILE;
    tx1 = min(x1, nx - 1) / (int)TILE;
    ty0 = max(y0, 0) / (int)TILE;
    ty1 = min(y1, ny - 1) / (int)TILE;
    // hit distances: perspective rays start at p with unit directions, so a
    // hit inside the box lies at least the box's distance from p; orthographic
    // rays start on the image plane and t is the depth along -w.  The computed
    // distance can be lower by CULL_K (|v| + 3R) with |v| at most the box's
    // farthest corner distance (plus its size for orthographic rays).
    float base, far;
    if (c.model == 1)
        {
        const float dx = fmaxf(fmaxf(lo.x - c.p.x, c.p.x - hi.x), 0.0f);
        const float dy = fmaxf(fmaxf(lo.y - c.p.y, c.p.y - hi.y), 0.0f);
        const float dz = fmaxf(fmaxf(lo.z - c.p.z, c.p.z - hi.z), 0.0f);
        base = sqrtf(dx * dx + dy * dy + dz * dz);
        far = sqrtf(dmax2);
        }
    else
        {
        base = zmin;
        far = zmax + ((hi.x - lo.x) + (hi.y - lo.y) + (hi.z - lo.z));
        }
    q = fmaxf(base * (1.0f - 0x1p-20f) - (CULL_K * far + c.margin), 0.0f);
    return true;
    }

// Pixel interval [a0, a1] (before the jitter margins) of the directions or
// origins whose rays pass within rho of a centre at camera coordinates
// (x, z): tangent slopes x/z of the circle for perspective rays, which needs
// z > 2 rho (false otherwise), the shifted interval for orthographic rays.
// The interval is widened by 2^-16 (1 + |slope|), far above the float error of
// the formula at z > 2 rho.
__device__ __forceinline__ bool ball_interval(const TileCam& c, float x, float z, float rho, float& a0, float& a1)
    {
    float m0, m1;
    if (c.model == 1)
        {
        if (!(z > 2.0f * rho))
            return false;
        const float den = z * z - rho * rho;
        const float root = rho * sqrtf(x * x + den);
        m0 = (x * z - root) / den;
        m1 = (x * z + root) / den;
        }
    else
        {
        m0 = x - rho;
        m1 = x + rho;
        }
    m0 -= 0x1p-16f * (1.0f + fabsf(m0));
    m1 += 0x1p-16f * (1.0f + fabsf(m1));
    a0 = m0 * c.scale;
    a1 = m1 * c.scale;
    return true;
    }

// Tighter rectangle and bound for a sphere (centre, radius s), intersected
// with the box rule's.  A ray whose computed sphere solve reports a hit passes
// the true centre within rho, rho^2 = r^2 (1 + 2^-20) + 2^-18 far^2: the float
// discriminant b*b - v.v + r*r errs by about 19 u |v|^2 (u = 2^-24, |v| <= far
// the distance from the ray origin to the centre).  Its computed distance is
// at least (distance of the centre, or its depth for orthographic rays) - r
// - CULL_K |v|, the same sqrt(14 u) |v| cancellation bound as the BVH cull.
__device__ __forceinline__ bool sphere_tile_rect(const TileCam& c, float4 s, float4 lo, float4 hi, int& tx0, int& tx1,
                                                 int& ty0, int& ty1, float& q, bool& straddle)
    {
    if (!tile_rect(c, lo, hi, tx0, tx1, ty0, ty1, q, straddle))
        return false;
    const v3 e = mk3(s.x, s.y, s.z) - c.p;
    const float dist = sqrtf(dot(e, e));
    const float far = dist * (1.0f + 0x1p-20f) + c.plane_r;
    const float rho = sqrtf(s.w * s.w * (1.0f + 0x1p-20f) + 0x1p-18f * far * far) * (1.0f + 0x1p-20f);
    const float cx = dot(e, c.u), cy = dot(e, c.v), cz = -dot(e, c.w);
    float i0, i1, j0, j1;
    if (ball_interval(c, cx, cz, rho, i0, i1) && ball_interval(c, cy, cz, rho, j0, j1))
        {
        // screen position -> pixel: i = xs H + W / 2, j = (1/2 - ys) H
        const float imin = i0 * c.H + 0.5f * c.W, imax = i1 * c.H + 0.5f * c.W;
        const float jmin = (0.5f - j1) * c.H, jmax = (0.5f - j0) * c.H;
        const float lim = 1.0e6f;
        const int x0 = (int)floorf(fmaxf(imin - 2.25f, -lim)), x1 = (int)floorf(fminf(imax + 1.25f, lim));
        const int y0 = (int)floorf(fmaxf(jmin - 2.25f, -lim)), y1 = (int)floorf(fminf(jmax + 1.25f, lim));
        const int nx = (int)(c.tiles_x * TILE), ny = (int)(c.tiles_y * TILE);
        if (x1 < 0 || y1 < 0 || x0 >= nx || y0 >= ny)
            return false;
        tx0 = max(tx0, max(x0, 0) / (int)TILE);
        tx1 = min(tx1, min(x1, nx - 1) / (int)TILE);
        ty0 = max(ty0, max(y0, 0) / (int)TILE);
        ty1 = min(ty1, min(y1, ny - 1) / (int)TILE);
        if (tx0 > tx1 || ty0 > ty1)
            return false;
        }
    const float base = ((c.model == 1) ? dist : cz) - s.w;
    const float qs = base - 0x1p-20f * (far + s.w) - (CULL_K * far + c.margin);
    q = fmaxf(q, fmaxf(qs, 0.0f));
    return true;
    }

// Tile rectangle and bound of primitive i: the box rule, tightened by the
// sphere rule in spheres-only scenes (the sphere through the bounds' slot and
// primitive bits).
template<bool SPHERES>
__device__ __forceinline__ bool prim_tile_rect(const TileCam& c, const float4* __restrict__ lo,
                                               const float4* __restrict__ hi, const GeomDev* __restrict__ geoms,
                                               uint32_t i, int& tx0, int& tx1, int& ty0, int& ty1, float& q,
                                               bool& straddle)
    {
    const float4 a = lo[i], b = hi[i];
    if constexpr (SPHERES)
        {
        const float4 s = geoms[__float_as_uint(a.w)].a[__float_as_uint(b.w)];
        return sphere_tile_rect(c, s, a, b, tx0, tx1, ty0, ty1, q, straddle);
        }
    else
        return tile_rect(c, a, b, tx0, tx1, ty0, ty1, q, straddle);
    }

template<bool SPHERES>
__global__ void k_tile_count(const float4* __restrict__ lo, const float4* __restrict__ hi,
                             const GeomDev* __restrict__ geoms, uint32_t N, TileCam c, uint32_t* counts,
                             uint32_t* flags)
    {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > N)
        return;
    if (i == N)
        {
        counts[N] = 0; // the scan's last element gives the total
        return;
        }
    int tx0, tx1, ty0, ty1;
    float q;
    bool straddle = false;
    uint32_t n = 0;
    if (prim_tile_rect<SPHERES>(c, lo, hi, geoms, i, tx0, tx1, ty0, ty1, q, straddle))
        n = (uint32_t)((tx1 - tx0 + 1) * (ty1 - ty0 + 1));
    counts[i] = n;
    if (straddle)
        atomicOr(flags, 1u);
    }

// Entries of primitive i at offsets[i]: tile index, bound bits, primitive.
template<bool SPHERES>
__global__ void k_tile_emit(const float4* __restrict__ lo, const float4* __restrict__ hi,
                            const GeomDev* __restrict__ geoms, uint32_t N, TileCam c,
                            const uint32_t* __restrict__ offsets, uint32_t* tile, uint32_t* qbits, uint32_t* ref,
                            uint32_t* order)
    {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N)
        return;
    int tx0, tx1, ty0, ty1;
    float q;
    bool straddle = false;
    if (!prim_tile_rect<SPHERES>(c, lo, hi, geoms, i, tx0, tx1, ty0, ty1, q, straddle))
        return;
    uint32_t e = offsets[i];
    const uint32_t e_end = offsets[i + 1]; // the count pass saw the same rectangle
    for (int ty = ty0; ty <= ty1; ty++)
        for (int tx = tx0; tx <= tx1 && e < e_end; tx++, e++)
            {
            tile[e] = (uint32_t)ty * c.tiles_x + (uint32_t)tx;
            qbits[e] = __float_as_uint(q); // q >= 0: the bits order like the values
            ref[e] = i;
            order[e] = e;
            }
    }

// ---------------------------------------------------------------- hierarchy cut
// In scenes other than spheres-only a tile list may hold subtrees of the
// binary hierarchy in place of their primitives.  The cut is the set of nodes
// whose tile rectangle covers at most `group` tiles while their parent's
// covers more (or the parent's box cuts the camera plane), plus the leaf
// primitives of such parents; the root is an element when it is small itself.
// A wavefront that reaches a subtree entry tests its box once per lane and
// searches the subtree only in the lanes whose ray passes it, instead of
// walking every primitive of a compact cluster (a mesh instance's triangles)
// as entries of their own.  The subtree search applies the slab and cull
// tests of the BVH search, and every box contains its primitives' boxes, so
// each ray tests the same primitives that could win and its closest hit is
// unchanged.
// Box of binary node n: the union of its children's boxes (an EMPTY child's
// NaN box drops out of fminf/fmaxf).  Every thread derives a node's box this
// way, so a parent and the node itself classify it alike.
__device__ __forceinline__ void node_box(const Node* __restrict__ nodes, uint32_t n, float4& lo, float4& hi)
    {
    const Node& N = nodes[n];
    lo = make_float4(fminf(N.a_lo.x, N.a_hi.z), fminf(N.a_lo.y, N.a_hi.w), fminf(N.a_lo.z, N.b_box.x), 0.0f);
    hi = make_float4(fmaxf(N.a_lo.w, N.b_box.y), fmaxf(N.a_hi.x, N.b_box.z), fmaxf(N.a_hi.y, N.b_box.w), 0.0f);
    }

// Whether a subtree box is split further: its rectangle covers more than
// `group` tiles or the camera plane cuts it.  A box no full tile sees is not
// split (nothing below it is seen either).
__device__ __forceinline__ bool box_large(const TileCam& c, float4 lo, float4 hi, uint32_t group)
    {
    int tx0, tx1, ty0, ty1;
    float q;
    bool straddle = false;
    if (!tile_rect(c, lo, hi, tx0, tx1, ty0, ty1, q, straddle))
        return straddle;
    return (uint32_t)((tx1 - tx0 + 1) * (ty1 - ty0 + 1)) > group;
    }

// Bounds index of leaf primitive `prim` of geometry `slot` (prefix[k] is the
// first bounds index of active slot slots[k]).
__device__ __forceinline__ uint32_t leaf_bounds_index(const uint32_t* __restrict__ prefix,
                                                      const uint32_t* __restrict__ slots, uint32_t n_slots,
                                                      uint32_t slot, uint32_t prim)
    {
    uint32_t base = 0;
    for (uint32_t k = 0; k < n_slots; k++)
        base = (slots[k] == slot) ? prefix[k] : base;
    return base + prim;
    }

// Entries of one cut element: the tiles of its rectangle, its bound, its
// reference.  Without EMIT, cnt[idx] = their number; with EMIT, they are
// written at cnt[idx] (the exclusive scan of the counts).
template<bool EMIT>
__device__ __forceinline__ void cut_entries(const TileCam& c, bool ok, int tx0, int tx1, int ty0, int ty1, float q,
                                            uint32_t refv, uint32_t* cnt, uint32_t idx, uint32_t* tile,
                                            uint32_t* qbits, uint32_t* ref, uint32_t* order)
    {
    if (!EMIT)
        {
        cnt[idx] = ok ? (uint32_t)((tx1 - tx0 + 1) * (ty1 - ty0 + 1)) : 0u;
        return;
        }
    if (!ok)
        return;
    uint32_t e = cnt[idx];
    const uint32_t e_end = cnt[idx + 1]; // the count pass saw the same rectangle
    for (int ty = ty0; ty <= ty1; ty++)
        for (int tx = tx0; tx <= tx1 && e < e_end; tx++, e++)
            {
            tile[e] = (uint32_t)ty * c.tiles_x + (uint32_t)tx;
            qbits[e] = __float_as_uint(q);
            ref[e] = refv;
            order[e] = e;
            }
    }

// Cut elements of the binary hierarchy (n_inner inner nodes, root 0): thread
// n < n_inner handles the two children of node n when n is split, thread
// n_inner the root; element slots 2 n + c and 2 n_inner, plus a zero slot
// 2 n_inner + 1 for the scan's total.  A leaf primitive whose box the camera
// plane cuts raises *flags, as in k_tile_count.
template<bool EMIT>
__global__ void k_tile_cut(const Node* __restrict__ nodes, uint32_t n_inner, const float4* __restrict__ lo,
                           const float4* __restrict__ hi, const uint32_t* __restrict__ prefix,
                           const uint32_t* __restrict__ slots, uint32_t n_slots, TileCam c, uint32_t group,
                           uint32_t* cnt, uint32_t* flags, uint32_t* tile, uint32_t* qbits, uint32_t* ref,
                           uint32_t* order)
    {
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n > n_inner)
        return;
    int tx0 = 0, tx1 = -1, ty0 = 0, ty1 = -1;
    float q = 0.0f;
    if (n == n_inner)
        {
        float4 a, b;
        node_box(nodes, 0, a, b);
        bool straddle = false;
        const bool ok = tile_rect(c, a, b, tx0, tx1, ty0, ty1, q, straddle)
                        && (uint32_t)((tx1 - tx0 + 1) * (ty1 - ty0 + 1)) <= group;
        cut_entries<EMIT>(c, ok, tx0, tx1, ty0, ty1, q, NODE_REF | 0u, cnt, 2 * n_inner, tile, qbits, ref, order);
        if (!EMIT)
            cnt[2 * n_inner + 1] = 0;
        return;
        }
    float4 a, b;
    node_box(nodes, n, a, b);
    const bool split = box_large(c, a, b, group);
    const uint4 lk = nodes[n].link;
    for (uint32_t k = 0; k < 2; k++)
        {
        const uint32_t child = k ? lk.y : lk.x, slot = k ? lk.w : lk.z;
        bool ok = false;
        uint32_t refv = 0;
        if (split && child != EMPTY)
            {
            bool straddle = false;
            if (child & LEAF)
                {
                refv = leaf_bounds_index(prefix, slots, n_slots, slot, child & ~LEAF);
                const float4 pa = lo[refv], pb = hi[refv];
                ok = tile_rect(c, pa, pb, tx0, tx1, ty0, ty1, q, straddle);
                if (straddle && !EMIT)
                    atomicOr(flags, 1u);
                }
            else
                {
                float4 ca, cb;
                node_box(nodes, child, ca, cb);
                refv = NODE_REF | child;
                ok = tile_rect(c, ca, cb, tx0, tx1, ty0, ty1, q, straddle)
                     && (uint32_t)((tx1 - tx0 + 1) * (ty1 - ty0 + 1)) <= group;
                }
            }
        cut_entries<EMIT>(c, ok, tx0, tx1, ty0, ty1, q, refv, cnt, 2 * n + k, tile, qbits, ref, order);
        }
    }

__global__ void k_gather_u32(const uint32_t* __restrict__ src, const uint32_t* __restrict__ idx, uint32_t n,
                             uint32_t* dst)
    {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        dst[i] = src[idx[i]];
    }

__global__ void k_fill_u32(uint32_t* p, uint32_t n, uint32_t v)
    {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = v;
    }

// Sorted entries packed as 8 words each, so a wavefront reads an entry with
// one scalar load: [0] bound bits, then for spheres-only scenes the sphere
// (x, y, z, r), geometry slot and primitive, otherwise the box (lo.xyz,
// hi.xyz) and the reference: a primitive's index into the bounds arrays, or
// NODE_REF | the root of a subtree (`nodes`: the binary hierarchy).
template<bool SPHERES>
__global__ void k_tile_pack(const uint32_t* __restrict__ order, const uint32_t* __restrict__ ref0,
                            const uint32_t* __restrict__ q0, const float4* __restrict__ lo,
                            const float4* __restrict__ hi, const GeomDev* __restrict__ geoms,
                            const Node* __restrict__ nodes, uint32_t n, uint4* __restrict__ out)
    {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n)
        return;
    const uint32_t e = order[j];
    const uint32_t r = ref0[e];
    if (SPHERES)
        {
        const float4 a = lo[r], b = hi[r];
        const uint32_t slot = __float_as_uint(a.w), prim = __float_as_uint(b.w);
        const float4 sp = geoms[slot].a[prim];
        out[2 * j] = make_uint4(q0[e], __float_as_uint(sp.x), __float_as_uint(sp.y), __float_as_uint(sp.z));
        out[2 * j + 1] = make_uint4(__float_as_uint(sp.w), slot, prim, 0u);
        }
    else
        {
        float4 a, b;
        if (r & NODE_REF)
            node_box(nodes, r & ~NODE_REF, a, b);
        else
            {
            a = lo[r];
            b = hi[r];
            }
        out[2 * j] = make_uint4(q0[e], __float_as_uint(a.x), __float_as_uint(a.y), __float_as_uint(a.z));
        out[2 * j + 1] = make_uint4(__float_as_uint(b.x), __float_as_uint(b.y), __float_as_uint(b.z), r);
        }
    }

// [start, end) of each tile's run in the tile-sorted entries.
__global__ void k_tile_bounds(const uint32_t* __restrict__ tile, uint32_t n, uint32_t* start, uint32_t* end)
    {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t t = tile[i];
    if (i == 0 || tile[i - 1] != t)
        start[t] = i;
    if (i == n - 1 || tile[i + 1] != t)
        end[t] = i + 1;
    }
    } // namespace fh
