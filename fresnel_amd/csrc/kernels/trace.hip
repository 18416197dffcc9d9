// trace.hip -- wavefront path tracer kernels for gfx950.
//
// One camera-sample batch of a Path render runs as
//   traverse(primary) -> shade_primary -> { traverse -> shade }* -> finish -> accumulate
// (primary rays are generated from their index where they are used)
// over ping-pong SoA ray queues (float4 origin|tnear, float4 dir|job,
// float4 attenuation, uint4 job|pixel|sample|depth).  shade appends the rays
// that continue with one atomic per wavefront (ballot + popcount) and writes a
// termination record per light-sample path; accumulate replays the records of
// each pixel in (sample, light sample) order, which is the order the
// reference adds them (fresnel/cpu/TracerPath.cc:141-239), so the image does
// not depend on how paths were scheduled.
#include "rt/isect.h"
#include "rt/shade.h"
#include "kernels/tiles.hip"

namespace fh
{
// >= 3 pushes per level of a 4-wide tree of depth 32 (half the depth of a
// Karras tree over 64-bit keys; the binary packet search pushes one per level)
constexpr uint32_t STACK_MAX = 96;

struct HitT
    {
    float t;
    uint32_t geom;
    uint32_t prim;
    };

__device__ __forceinline__ void slab(v3 o, v3 inv, float lx, float ly, float lz, float hx, float hy, float hz,
                                     float& tmin, float& tmax)
    {
    const float tx0 = (lx - o.x) * inv.x, tx1 = (hx - o.x) * inv.x;
    const float ty0 = (ly - o.y) * inv.y, ty1 = (hy - o.y) * inv.y;
    const float tz0 = (lz - o.z) * inv.z, tz1 = (hz - o.z) * inv.z;
    tmin = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fminf(tz0, tz1));
    tmax = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1));
    }

// Lower bound on the computed distance of any primitive inside a box whose
// slab interval is [tmin, tmax] (see CULL_K).  A box is culled when this
// bound exceeds the closest hit, so every primitive that could still win is
// tested and the result does not depend on the hierarchy or the visiting
// order.  margin = CULL_K * 3 * the scene's largest primitive half-extent.
__device__ __forceinline__ float cull_bound(float tmin, float tmax, float margin)
    {
    return tmin - (CULL_K * fmaxf(tmax, 0.0f) + margin);
    }

__device__ __forceinline__ bool better(float t, uint32_t g, uint32_t p, const HitT& b)
    {
    // every comparison evaluated (no short circuit): lane masks, no branches
    const bool lt = t < b.t, eq = t == b.t, g_lt = g < b.geom, g_eq = g == b.geom, p_lt = p < b.prim;
    return lt | (eq & (g_lt | (g_eq & p_lt)));
    }

// Stack with the top SS entries in LDS (interleaved by thread) and the rest
// in a per-thread global region.
template<int SS> struct LdsStack
    {
    uint2* lds;
    uint2* glob;
    uint32_t lstride, gstride;
    uint32_t sp;
    __device__ __forceinline__ void push(uint32_t n, float t, uint32_t& err)
        {
        const uint2 e = make_uint2(n, __float_as_uint(t));
        if (sp < SS)
            lds[sp * lstride] = e;
        else if (sp < STACK_MAX)
            glob[(sp - SS) * gstride] = e;
        else
            {
            err = 1;
            return;
            }
        sp++;
        }
    __device__ __forceinline__ void pop(uint32_t& n, float& t)
        {
        sp--;
        const uint2 e = (sp < SS) ? lds[sp * lstride] : glob[(sp - SS) * gstride];
        n = e.x;
        t = __uint_as_float(e.y);
        }
    };

// Stack in private memory (for the preview and diagnostic kernels).
struct LocalStack
    {
    uint2 s[STACK_MAX];
    uint32_t sp;
    __device__ __forceinline__ void push(uint32_t n, float t, uint32_t& err)
        {
        if (sp < STACK_MAX)
            s[sp++] = make_uint2(n, __float_as_uint(t));
        else
            err = 1;
        }
    __device__ __forceinline__ void pop(uint32_t& n, float& t)
        {
        sp--;
        n = s[sp].x;
        t = __uint_as_float(s[sp].y);
        }
    };


// Candidate distance of one primitive; TYPES lists the geometry types the
// scene can hold, so a spheres-only traversal compiles without the other
// intersectors (fewer registers, more wavefronts in flight).
template<uint32_t TYPES>
__device__ __forceinline__ bool leaf_test(const GeomDev& g, uint32_t prim, v3 o, v3 d, float tnear, float& t)
    {
    if constexpr (TYPES == T_SPHERES)
        {
        const float4 s = g.a[prim];
        float de;
        bool first;
        return sphere_t(o, d, mk3(s.x, s.y, s.z), s.w, t, de, first) && (tnear < t);
        }
    else
        return intersect_prim<false>(g, prim, o, d, tnear, t, nullptr);
    }

typedef const Node __attribute__((address_space(1))) GlobalNode;

// Closest-hit traversal state of one ray.
struct TraceState
    {
    v3 o, d, inv;
    float tnear;
    float margin;
    HitT best;
    uint32_t node;
    bool sph; // 4-wide leaf children hold spheres
    };

template<class Stack>
__device__ __forceinline__ void trace_begin(const SceneDev& sc, v3 o, v3 d, float tnear, TraceState& r, Stack& st)
    {
    r.o = o;
    r.d = d;
    r.inv = mk3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    r.tnear = tnear;
    r.margin = sc.cull_margin;
    r.best.t = __builtin_huge_valf();
    r.best.geom = EMPTY;
    r.best.prim = EMPTY;
    r.node = sc.root;
    r.sph = sc.sphere_leaves != 0;
    st.sp = 0;
    }

// Primitive test of a leaf child that passed the slab and cull tests.
template<bool COUNT, uint32_t TYPES>
__device__ __forceinline__ void visit_leaf(const GeomDev* __restrict__ geoms, uint32_t ref, uint32_t slot,
                                           TraceState& r, uint32_t& n_prims)
    {
    const uint32_t prim = ref & ~LEAF;
    float t;
    if (COUNT)
        n_prims++;
    if (leaf_test<TYPES>(geoms[slot], prim, r.o, r.d, r.tnear, t) && better(t, slot, prim, r.best))
        {
        r.best.t = t;
        r.best.geom = slot;
        r.best.prim = prim;
        }
    }

// Sort two (key, node, bound) entries by key.
__device__ __forceinline__ void order2(float& ka, uint32_t& na, float& ca, float& kb, uint32_t& nb, float& cb)
    {
    const bool sw = kb < ka;
    const float k = sw ? kb : ka, c = sw ? cb : ca;
    const uint32_t n = sw ? nb : na;
    kb = sw ? ka : kb;
    nb = sw ? na : nb;
    cb = sw ? ca : cb;
    ka = k;
    na = n;
    ca = c;
    }

// One node visit of the closest-hit search over the 4-wide hierarchy: the
// primitives of leaf children are tested, the nearest hit inner child is
// visited next and the others are pushed, farthest first.  Returns false when
// the search is complete.  The closest hit is the smallest (t, geometry,
// primitive) over the primitives whose bounds pass the slab test with
// tmax >= tnear; a subtree is culled when the lower bound of its primitives'
// distances (cull_bound) exceeds the closest hit, so neither the hierarchy
// nor the visiting order changes the result.
template<class Stack, bool COUNT, uint32_t TYPES>
__device__ __forceinline__ bool trace_visit4(const Node4* __restrict__ nodes, const GeomDev* __restrict__ geoms,
                                            TraceState& r, Stack& st, uint32_t& n_nodes, uint32_t& n_prims,
                                            uint32_t& err)
    {
    if (r.node == EMPTY)
        return false;
    if (COUNT)
        n_nodes++;
    // spheres-only kernels always see sphere leaves (checked on the host)
    const bool sph = (TYPES == T_SPHERES) || r.sph;
    float tmin[4], cb[4];
    bool h[4];
    uint32_t ch[4];
#pragma unroll
    for (int q = 0; q < 2; q++)
        {
        // one child pair: global address space, direct global loads
        const GlobalNode* np = (const GlobalNode*)(&nodes[r.node].pair[q]);
        const float4 A = make_float4(np->a_lo.x, np->a_lo.y, np->a_lo.z, np->a_lo.w);
        const float4 B = make_float4(np->a_hi.x, np->a_hi.y, np->a_hi.z, np->a_hi.w);
        const float4 C = make_float4(np->b_box.x, np->b_box.y, np->b_box.z, np->b_box.w);
        const uint4 lk = make_uint4(np->link.x, np->link.y, np->link.z, np->link.w);
        const float bx[2][6] = {{A.x, A.y, A.z, A.w, B.x, B.y}, {B.z, B.w, C.x, C.y, C.z, C.w}};
        const uint32_t ref[2] = {lk.x, lk.y}, slot[2] = {lk.z, lk.w};
#pragma unroll
        for (int c = 0; c < 2; c++)
            {
            const int k = 2 * q + c;
            // with sphere leaves every child is a centre and half extents
            // (k_bvh4_fill): a sphere leaf's box is centre -+ radius, as the
            // builder bounds it (GeometrySphere.cc:60-73)
            float x0 = bx[c][0], y0 = bx[c][1], z0 = bx[c][2], x1 = bx[c][3], y1 = bx[c][4], z1 = bx[c][5];
            if (sph)
                {
                x0 = bx[c][0] - bx[c][3];
                x1 = bx[c][0] + bx[c][3];
                y0 = bx[c][1] - bx[c][4];
                y1 = bx[c][1] + bx[c][4];
                z0 = bx[c][2] - bx[c][5];
                z1 = bx[c][2] + bx[c][5];
                }
            float tmax;
            slab(r.o, r.inv, x0, y0, z0, x1, y1, z1, tmin[k], tmax);
            cb[k] = cull_bound(tmin[k], tmax, r.margin);
            h[k] = (tmin[k] <= tmax) && (tmax >= r.tnear) && !(cb[k] > r.best.t);
            ch[k] = ref[c];
            }
        // leaf children of this pair
#pragma unroll
        for (int c = 0; c < 2; c++)
            {
            const int k = 2 * q + c;
            if (h[k] && (ref[c] & LEAF))
                {
                if (sph)
                    {
                    // the sphere is in the node: no memory access
                    const uint32_t prim = ref[c] & ~LEAF;
                    float t, de;
                    bool first;
                    if (COUNT)
                        n_prims++;
                    if (sphere_t(r.o, r.d, mk3(bx[c][0], bx[c][1], bx[c][2]), bx[c][3], t, de, first)
                        && (r.tnear < t) && better(t, slot[c], prim, r.best))
                        {
                        r.best.t = t;
                        r.best.geom = slot[c];
                        r.best.prim = prim;
                        }
                    }
                else
                    visit_leaf<COUNT, TYPES>(geoms, ref[c], slot[c], r, n_prims);
                h[k] = false;
                }
            }
        }
    // inner children still open, nearest entry first
    float key[4];
    uint32_t nd[4];
#pragma unroll
    for (int k = 0; k < 4; k++)
        {
        const bool open = h[k] && !(cb[k] > r.best.t);
        key[k] = open ? tmin[k] : __builtin_huge_valf();
        nd[k] = open ? ch[k] : EMPTY;
        }
    order2(key[0], nd[0], cb[0], key[1], nd[1], cb[1]);
    order2(key[2], nd[2], cb[2], key[3], nd[3], cb[3]);
    order2(key[0], nd[0], cb[0], key[2], nd[2], cb[2]);
    order2(key[1], nd[1], cb[1], key[3], nd[3], cb[3]);
    order2(key[1], nd[1], cb[1], key[2], nd[2], cb[2]);
    if (nd[0] != EMPTY)
        {
#pragma unroll
        for (int k = 3; k >= 1; k--)
            if (nd[k] != EMPTY)
                st.push(nd[k], cb[k], err);
        r.node = nd[0];
        return true;
        }
    while (st.sp > 0)
        {
        uint32_t n;
        float tm;
        st.pop(n, tm);
        if (!(tm > r.best.t))
            {
            r.node = n;
            return true;
            }
        }
    r.node = EMPTY;
    return false;
    }

// One node visit of the closest-hit search over the binary hierarchy (the
// kernels for scenes with other primitive types, whose larger intersectors
// leave no registers for 4-wide nodes): the primitives of leaf children
// are tested, the nearer hit inner child is visited next and the farther one
// pushed.  Returns false when the search is complete.  The closest hit is
// the smallest (t, geometry, primitive) over the primitives whose bounds pass
// the slab test with tmax >= tnear; a subtree is culled when the lower bound
// of its primitives' distances (cull_bound) exceeds the closest hit.
template<class Stack, bool COUNT, uint32_t TYPES>
__device__ __forceinline__ bool trace_visit2(const Node* __restrict__ nodes, const GeomDev* __restrict__ geoms,
                                            TraceState& r, Stack& st, uint32_t& n_nodes, uint32_t& n_prims,
                                            uint32_t& err)
    {
    if (r.node == EMPTY)
        return false;
    // global address space: direct global loads instead of flat ones
    const GlobalNode* np = (const GlobalNode*)(nodes + r.node);
    const float4 A = make_float4(np->a_lo.x, np->a_lo.y, np->a_lo.z, np->a_lo.w);
    const float4 B = make_float4(np->a_hi.x, np->a_hi.y, np->a_hi.z, np->a_hi.w);
    const float4 C = make_float4(np->b_box.x, np->b_box.y, np->b_box.z, np->b_box.w);
    const uint4 lk = make_uint4(np->link.x, np->link.y, np->link.z, np->link.w);
    if (COUNT)
        n_nodes++;
    float t0min, t0max, t1min, t1max;
    slab(r.o, r.inv, A.x, A.y, A.z, A.w, B.x, B.y, t0min, t0max);
    slab(r.o, r.inv, B.z, B.w, C.x, C.y, C.z, C.w, t1min, t1max);
    const float c0 = cull_bound(t0min, t0max, r.margin);
    const float c1 = cull_bound(t1min, t1max, r.margin);
    bool h0 = (t0min <= t0max) && (t0max >= r.tnear) && !(c0 > r.best.t);
    bool h1 = (t1min <= t1max) && (t1max >= r.tnear) && !(c1 > r.best.t);
    if (h0 && (lk.x & LEAF))
        {
        const uint32_t prim = lk.x & ~LEAF;
        float t;
        if (COUNT)
            n_prims++;
        if (leaf_test<TYPES>(geoms[lk.z], prim, r.o, r.d, r.tnear, t) && better(t, lk.z, prim, r.best))
            {
            r.best.t = t;
            r.best.geom = lk.z;
            r.best.prim = prim;
            }
        h0 = false;
        }
    if (h1 && (lk.y & LEAF))
        {
        const uint32_t prim = lk.y & ~LEAF;
        float t;
        if (COUNT)
            n_prims++;
        if (leaf_test<TYPES>(geoms[lk.w], prim, r.o, r.d, r.tnear, t) && better(t, lk.w, prim, r.best))
            {
            r.best.t = t;
            r.best.geom = lk.w;
            r.best.prim = prim;
            }
        h1 = false;
        }
    h0 = h0 && !(c0 > r.best.t);
    h1 = h1 && !(c1 > r.best.t);
    if (h0 && h1)
        {
        const bool swap = t1min < t0min;
        st.push(swap ? lk.x : lk.y, swap ? c0 : c1, err);
        r.node = swap ? lk.y : lk.x;
        return true;
        }
    if (h0)
        {
        r.node = lk.x;
        return true;
        }
    if (h1)
        {
        r.node = lk.y;
        return true;
        }
    while (st.sp > 0)
        {
        uint32_t n;
        float tm;
        st.pop(n, tm);
        if (!(tm > r.best.t))
            {
            r.node = n;
            return true;
            }
        }
    r.node = EMPTY;
    return false;
    }

// One node visit of the closest-hit search: spheres-only kernels search the
// 4-wide hierarchy, the others the binary one (same result either way).
template<class Stack, bool COUNT, uint32_t TYPES>
__device__ __forceinline__ bool trace_visit(const SceneDev& sc, TraceState& r, Stack& st, uint32_t& n_nodes,
                                            uint32_t& n_prims, uint32_t& err)
    {
    if constexpr (TYPES == T_SPHERES)
        return trace_visit4<Stack, COUNT, TYPES>(sc.nodes4, sc.geoms, r, st, n_nodes, n_prims, err);
    else
        return trace_visit2<Stack, COUNT, TYPES>(sc.nodes, sc.geoms, r, st, n_nodes, n_prims, err);
    }

template<class Stack, bool COUNT, uint32_t TYPES = T_ALL>
__device__ inline void trace_closest(const SceneDev& sc, v3 o, v3 d, float tnear, HitT& best, Stack& st,
                                     uint32_t& n_nodes, uint32_t& n_prims, uint32_t& err)
    {
    TraceState r;
    trace_begin(sc, o, d, tnear, r, st);
    while (trace_visit<Stack, COUNT, TYPES>(sc, r, st, n_nodes, n_prims, err))
        {
        }
    best = r.best;
    }

// One node visit of the binary-hierarchy search with the primitive tests
// postponed: leaf children that pass the slab and cull tests are appended to
// the lane's pending list (ps, pp; np entries, at most two per visit) instead
// of being tested; inner children are visited or pushed as in trace_visit2.
template<class Stack, bool COUNT>
__device__ __forceinline__ void visit2_collect(const Node* __restrict__ nodes, TraceState& r, Stack& st,
                                               uint32_t (&ps)[2], uint32_t (&pp)[2], uint32_t& np,
                                               uint32_t& n_nodes, uint32_t& err)
    {
    const GlobalNode* n = (const GlobalNode*)(nodes + r.node);
    const float4 A = make_float4(n->a_lo.x, n->a_lo.y, n->a_lo.z, n->a_lo.w);
    const float4 B = make_float4(n->a_hi.x, n->a_hi.y, n->a_hi.z, n->a_hi.w);
    const float4 C = make_float4(n->b_box.x, n->b_box.y, n->b_box.z, n->b_box.w);
    const uint4 lk = make_uint4(n->link.x, n->link.y, n->link.z, n->link.w);
    if (COUNT)
        n_nodes++;
    float t0min, t0max, t1min, t1max;
    slab(r.o, r.inv, A.x, A.y, A.z, A.w, B.x, B.y, t0min, t0max);
    slab(r.o, r.inv, B.z, B.w, C.x, C.y, C.z, C.w, t1min, t1max);
    const float c0 = cull_bound(t0min, t0max, r.margin);
    const float c1 = cull_bound(t1min, t1max, r.margin);
    bool h0 = (t0min <= t0max) && (t0max >= r.tnear) && !(c0 > r.best.t);
    bool h1 = (t1min <= t1max) && (t1max >= r.tnear) && !(c1 > r.best.t);
    if (h0 && (lk.x & LEAF))
        {
        ps[0] = lk.z;
        pp[0] = lk.x & ~LEAF;
        np = 1;
        h0 = false;
        }
    if (h1 && (lk.y & LEAF))
        {
        // constant indices keep the list in registers
        if (np == 0)
            {
            ps[0] = lk.w;
            pp[0] = lk.y & ~LEAF;
            }
        else
            {
            ps[1] = lk.w;
            pp[1] = lk.y & ~LEAF;
            }
        np++;
        h1 = false;
        }
    if (h0 && h1)
        {
        const bool swap = t1min < t0min;
        st.push(swap ? lk.x : lk.y, swap ? c0 : c1, err);
        r.node = swap ? lk.y : lk.x;
        return;
        }
    if (h0 || h1)
        {
        r.node = h0 ? lk.x : lk.y;
        return;
        }
    r.node = EMPTY;
    while (st.sp > 0)
        {
        uint32_t nn;
        float tm;
        st.pop(nn, tm);
        if (!(tm > r.best.t))
            {
            r.node = nn;
            return;
            }
        }
    }

// Postponed primitive tests of type G: the lanes whose pending entry k is a
// primitive of that type run its intersector together.
template<uint32_t G, uint32_t TYPES, bool COUNT>
__device__ __forceinline__ void test_pending(const GeomDev* __restrict__ geoms, TraceState& r, bool has, uint32_t ty,
                                             uint32_t slot, uint32_t prim, uint32_t& n_prims)
    {
    if constexpr ((TYPES & (1u << G)) != 0)
        {
        const bool mine = has && ty == G;
        if (__ballot(mine) == 0)
            return;
        if (mine)
            {
            float t;
            if (COUNT)
                n_prims++;
            if (intersect_prim<false, 1u << G>(geoms[slot], prim, r.o, r.d, r.tnear, t, nullptr)
                && better(t, slot, prim, r.best))
                {
                r.best.t = t;
                r.best.geom = slot;
                r.best.prim = prim;
                }
            }
        }
    }

// Closest hit over the binary hierarchy for scenes of several primitive
// types, one ray per lane.  A lane that reaches leaves keeps them pending and
// waits; once the waiting lanes are at least twice as many as those still
// searching (or none can go on), the pending primitives are tested one
// geometry type at a time,
// so each intersector runs once for all the lanes that need it instead of a
// few lanes at a time per node visit.  Culling meanwhile uses the closest hit
// found so far, which only prunes less; every primitive whose bounds pass the
// slab test and that could still win is tested, so the result is the same
// (t, geometry, primitive) minimum as trace_closest's.
template<class Stack, bool COUNT, uint32_t TYPES>
__device__ inline void trace_closest_grouped(const SceneDev& sc, v3 o, v3 d, float tnear, HitT& best, Stack& st,
                                             uint32_t& n_nodes, uint32_t& n_prims, uint32_t& err)
    {
    TraceState r;
    trace_begin(sc, o, d, tnear, r, st);
    uint32_t ps[2] = {0, 0}, pp[2] = {0, 0};
    uint32_t np = 0;
    for (;;)
        {
        for (;;)
            {
            const bool go = r.node != EMPTY && np == 0;
            const uint32_t n_go = (uint32_t)__popcll(__ballot(go));
            const uint32_t n_wait = (uint32_t)__popcll(__ballot(np != 0));
            if (n_go == 0 || n_wait >= 2 * n_go)
                break;
            if (go)
                visit2_collect<Stack, COUNT>(sc.nodes, r, st, ps, pp, np, n_nodes, err);
            }
        if (__ballot(np != 0) == 0)
            break;
#pragma unroll
        for (uint32_t k = 0; k < 2; k++)
            {
            const bool has = np > k;
            const uint32_t slot = ps[k], prim = pp[k];
            const uint32_t ty = has ? sc.geoms[slot].type : 0xffu;
            test_pending<G_SPHERE, TYPES, COUNT>(sc.geoms, r, has, ty, slot, prim, n_prims);
            test_pending<G_CYLINDER, TYPES, COUNT>(sc.geoms, r, has, ty, slot, prim, n_prims);
            test_pending<G_POLYHEDRON, TYPES, COUNT>(sc.geoms, r, has, ty, slot, prim, n_prims);
            test_pending<G_MESH, TYPES, COUNT>(sc.geoms, r, has, ty, slot, prim, n_prims);
            test_pending<G_POLYGON, TYPES, COUNT>(sc.geoms, r, has, ty, slot, prim, n_prims);
            }
        np = 0;
        }
    best = r.best;
    }

// A ray queue is cut into QSEG segments with one append counter each, the
// counters QSEG_STRIDE words (256 bytes) apart: a producer unit (a claimed
// chunk of a persistent kernel, or a workgroup) appends to segment
// unit % QSEG, so the append atomics spread over QSEG addresses instead of
// queueing on one (one counter's atomic rate bounded the shading kernels).
// Segment s keeps its entries at [s seg_cap, s seg_cap + count_s); the host
// sizes seg_cap from the producers' per-unit output bounds, so no segment can
// overflow (a full segment raises a flag instead of writing past it).
// Consumers see the entries as one dense index range: entry i of the queue
// lives at slot i + adj[s] for the last segment s with pre[s] <= i.
constexpr uint32_t QSEG = 16, QSEG_STRIDE = 64;
constexpr uint32_t QSEG_FLAG = QSEG * QSEG_STRIDE; // overflow flag word after the counters

struct Queue
    {
    float4* o;   // origin, tnear
    float4* d;   // direction, job (bits)
    float4* att; // attenuation
    uint4* meta; // job, pixel, sample-in-batch | light sample << 16, depth
    uint32_t* count;      // producer: append counters (count[s * QSEG_STRIDE]) and flag
    uint32_t seg_cap;     // slots per segment
    uint32_t pre[QSEG];   // consumer: first dense index of each segment
    uint32_t adj[QSEG];   // consumer: slot - index within segment s (s seg_cap - pre[s])
    };

// Consumer view of a queue, written on the device by k_queue_view from its
// producer's segment counters: the entry count, the segment prefix (as
// Queue::pre/adj) and which of the next kernels act on it.  A bounce step's
// kernels all read it, so no count goes back to the host between bounces.
constexpr uint32_t VIEW_IDLE = 0, VIEW_BOUNCE = 1, VIEW_FINISH = 2;
struct QueueView
    {
    uint32_t n;
    uint32_t mode;
    uint32_t pad[2];
    uint32_t pre[QSEG];
    uint32_t adj[QSEG];
    };

// A kernel of `mode` reading a queue: false when it has nothing to do;
// otherwise n and q's segment prefix come from the view (no view: the values
// passed by the host stand).  The view is the same for every thread.
__device__ __forceinline__ bool take_view(const QueueView* __restrict__ view, uint32_t mode, Queue& q, uint32_t& n)
    {
    if (!view)
        return true;
    if (view->mode != mode)
        return false;
    n = view->n;
#pragma unroll
    for (uint32_t s = 0; s < QSEG; s++)
        {
        q.pre[s] = view->pre[s];
        q.adj[s] = view->adj[s];
        }
    return true;
    }

// One wavefront: the consumer view of the queue whose segment counters are
// `count` (filled by the previous producer), then those counters and the
// traversal's work counters cleared for the next producer and consumer.  The
// view's mode is FINISH when at most `tail` entries are left or `force` is
// set, BOUNCE for more, IDLE when the queue is empty or a segment overflowed
// (then CTR_QUEUE_OVERFLOW is raised; the host reports it).  steps[2 step],
// steps[2 step + 1] record the count and mode.
__global__ __launch_bounds__(64) void k_queue_view(uint32_t* count, uint32_t seg_cap, uint32_t tail, uint32_t force,
                                                   QueueView* view, uint32_t* work, uint32_t* steps, uint32_t step,
                                                   unsigned long long* counters)
    {
    const uint32_t lane = threadIdx.x;
    const uint32_t c = lane < QSEG ? count[lane * QSEG_STRIDE] : 0u;
    const uint32_t flag = count[QSEG_FLAG];
    uint32_t incl = c;
    for (uint32_t off = 1; off < QSEG; off <<= 1)
        {
        const uint32_t t = __shfl_up(incl, off, 64);
        if (lane >= off)
            incl += t;
        }
    const uint32_t total = __shfl(incl, QSEG - 1, 64);
    const bool bad = flag != 0 || __ballot(lane < QSEG && c > seg_cap) != 0;
    const uint32_t mode = (bad || total == 0) ? VIEW_IDLE : ((total <= tail || force) ? VIEW_FINISH : VIEW_BOUNCE);
    if (lane < QSEG)
        {
        const uint32_t pre = incl - c;
        view->pre[lane] = pre;
        view->adj[lane] = lane * seg_cap - pre;
        count[lane * QSEG_STRIDE] = 0u;
        }
    if (lane < 8)
        work[lane * QSEG_STRIDE] = 0u;
    if (lane == 0)
        {
        view->n = total;
        view->mode = mode;
        steps[2 * step] = total;
        steps[2 * step + 1] = mode;
        count[QSEG_FLAG] = 0u;
        if (bad)
            atomicOr((unsigned int*)(counters + CTR_QUEUE_OVERFLOW), 1u);
        }
    }

// slot - index for dense index i (pre is non-decreasing, pre[0] = 0)
__device__ __forceinline__ uint32_t seg_adj(const Queue& q, uint32_t i)
    {
    uint32_t a = 0;
#pragma unroll
    for (uint32_t s = 1; s < QSEG; s++)
        a = (i >= q.pre[s]) ? q.adj[s] : a;
    return a;
    }

// Slot of dense index i, for lanes whose indices lie in [lo, hi] (both
// wavefront-uniform): one scalar offset when the range is in one segment.
__device__ __forceinline__ uint32_t q_slot(const Queue& q, uint32_t i, uint32_t lo, uint32_t hi)
    {
    lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)lo);
    hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)hi);
    const uint32_t a = seg_adj(q, lo);
    if (a == seg_adj(q, hi))
        return i + a;
    return i + seg_adj(q, i);
    }

// First slot of a reservation of `total` entries in segment seg, or EMPTY
// (and the flag set) when the segment is full.
__device__ __forceinline__ uint32_t seg_reserve(const Queue& q, uint32_t seg, uint32_t total)
    {
    const uint32_t p = atomicAdd(q.count + seg * QSEG_STRIDE, total);
    if ((uint64_t)p + total <= q.seg_cap)
        return seg * q.seg_cap + p;
    atomicOr(q.count + QSEG_FLAG, 1u);
    return EMPTY;
    }

__device__ __forceinline__ uint32_t lane_id()
    {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
    }

// Append slot for lanes with pred set: one atomic per wavefront.  Every lane
// of the wavefront must call it.
__device__ __forceinline__ uint32_t wave_append(bool pred, const Queue& q, uint32_t seg)
    {
    const uint64_t m = __ballot(pred);
    const uint32_t lane = lane_id();
    uint32_t base = EMPTY;
    if (lane == 0 && m)
        base = seg_reserve(q, seg, (uint32_t)__popcll(m));
    base = __shfl(base, 0, 64);
    return (base == EMPTY) ? EMPTY : base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    }

// Append slot in segment seg of q for lanes with pred set (EMPTY if the
// segment is full): one atomic per workgroup of BS threads (wavefront counts
// combined in LDS).  Every thread of the workgroup must call it; sh holds
// BS / 64 + 1 words.
template<int BS>
__device__ __forceinline__ uint32_t block_append(bool pred, const Queue& q, uint32_t seg, uint32_t* sh)
    {
    const uint64_t m = __ballot(pred);
    const uint32_t lane = lane_id();
    const uint32_t wave = threadIdx.x >> 6;
    if (lane == 0)
        sh[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0)
        {
        uint32_t total = 0;
        for (int w = 0; w < BS / 64; w++)
            {
            const uint32_t c = sh[w];
            sh[w] = total;
            total += c;
            }
        sh[BS / 64] = total ? seg_reserve(q, seg, total) : EMPTY;
        }
    __syncthreads();
    const uint32_t b = sh[BS / 64];
    const uint32_t slot = (b == EMPTY) ? EMPTY : b + sh[wave] + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    __syncthreads(); // sh is reused by the next call
    return slot;
    }

// Rank of this thread among the workgroup's threads with `pred` (in thread
// order) and, in `total`, their number.  Every thread must call it.
template<int BS>
__device__ __forceinline__ uint32_t block_rank(bool pred, uint32_t* sh, uint32_t& total)
    {
    const uint64_t m = __ballot(pred);
    const uint32_t lane = lane_id();
    const uint32_t wave = threadIdx.x >> 6;
    if (lane == 0)
        sh[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (uint32_t w = 0; w < BS / 64; w++)
        {
        const uint32_t c = sh[w];
        before += (w < wave) ? c : 0u;
        all += c;
        }
    __syncthreads(); // sh is reused by the next call
    total = all;
    return before + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    }

template<int BS>
__device__ __forceinline__ uint32_t block_max(uint32_t v, uint32_t* sh)
    {
    for (int off = 32; off > 0; off >>= 1)
        v = max(v, (uint32_t)__shfl_xor(v, off, 64));
    if (lane_id() == 0)
        sh[threadIdx.x >> 6] = v;
    __syncthreads();
    uint32_t r = 0;
    for (int w = 0; w < BS / 64; w++)
        r = max(r, sh[w]);
    __syncthreads();
    return r;
    }

__device__ __forceinline__ uint32_t wave_max(uint32_t v)
    {
    for (int off = 32; off > 0; off >>= 1)
        v = max(v, (uint32_t)__shfl_xor(v, off, 64));
    return v;
    }

__device__ __forceinline__ uint32_t find_job(const JobDev* jobs, uint32_t n_jobs, uint32_t i)
    {
    // last job with ray_base <= i
    uint32_t lo = 0, hi = n_jobs;
    while (hi - lo > 1)
        {
        const uint32_t mid = (lo + hi) >> 1;
        if (jobs[mid].ray_base <= i)
            lo = mid;
        else
            hi = mid;
        }
    return lo;
    }

__device__ __forceinline__ uint4 make_rec(v3 c, uint32_t code)
    {
    return make_uint4(__float_as_uint(c.x), __float_as_uint(c.y), __float_as_uint(c.z), code);
    }

// Termination record of light sample ls of camera sample k at pixel ordinal
// j (the primary order below): chunk-major, [j / 64][k][j % 64][ls], so the
// 64 paths of a wavefront's tile are stored and read back contiguously.
__device__ __forceinline__ uint64_t rec_index(const JobDev& J, uint32_t k, uint32_t j, uint32_t ls)
    {
    return J.rec_base + (((uint64_t)(j >> 6) * J.K + k) * 64u + (j & 63u)) * J.L + ls;
    }


// ---------------------------------------------------------------- traverse
// Persistent threads: each wavefront takes 64 rays at a time from a counter
// until the queue is exhausted.
// Primary ray i of a batch: job, camera sample and pixel from the index
// (fresnel/cpu/TracerPath.cc:100-136, tnear 1e-3).
// Order of a job's primary rays: the image is cut into chunks of 64 pixels
// (8x8 tiles over the largest multiple-of-8 region, then the right and bottom
// strips row by row), and the K samples of a chunk are consecutive.  A
// wavefront therefore traces one sample of one compact tile, and the rays in
// flight at any time cover a K times smaller part of the image than in
// sample-major order.  local in [0, npix K) -> (sample k, pixel px, py).
// Pixel of ordinal j of the primary order (below).
__device__ __forceinline__ void ordinal_pixel(const JobDev& J, uint32_t j, uint32_t& px, uint32_t& py)
    {
    const uint32_t W8 = J.W & ~7u, H8 = J.H & ~7u;
    if (j < W8 * H8)
        {
        const uint32_t tile = j >> 6, within = j & 63u, tiles_x = W8 >> 3;
        const uint32_t ty = fdiv(tile, J.div_tiles); // tile / tiles_x
        px = (tile - ty * tiles_x) * 8u + (within & 7u);
        py = ty * 8u + (within >> 3);
        return;
        }
    uint32_t r = j - W8 * H8;
    const uint32_t rw = J.W - W8;
    if (r < rw * H8)
        {
        py = r / rw;
        px = W8 + (r - py * rw);
        return;
        }
    r -= rw * H8;
    const uint32_t by = r / J.W;
    py = H8 + by;
    px = r - by * J.W;
    }

// Ordinal of pixel (px, py): the inverse of ordinal_pixel.
__device__ __forceinline__ uint32_t pixel_ordinal(const JobDev& J, uint32_t px, uint32_t py)
    {
    const uint32_t W8 = J.W & ~7u, H8 = J.H & ~7u;
    if (px < W8 && py < H8)
        return (((py >> 3) * (W8 >> 3) + (px >> 3)) << 6) | ((py & 7u) << 3) | (px & 7u);
    const uint32_t rw = J.W - W8;
    if (py < H8)
        return W8 * H8 + py * rw + (px - W8);
    return W8 * H8 + rw * H8 + (py - H8) * J.W + px;
    }

// Ordinal of the pixel with raster index `pixel`.
__device__ __forceinline__ uint32_t raster_ordinal(const JobDev& J, uint32_t pixel)
    {
    const uint32_t py = fdiv(pixel, J.div_w);
    return pixel_ordinal(J, pixel - py * J.W, py);
    }

__device__ __forceinline__ void primary_order(const JobDev& J, uint32_t local, uint32_t& k, uint32_t& px,
                                              uint32_t& py, uint32_t& j)
    {
    const uint32_t chunk = fdiv(local, J.div_chunk); // local / (64 K)
    const uint32_t rem = local - chunk * 64u * J.K;
    const uint32_t size = min(64u, J.npix - chunk * 64u);
    k = (size == 64u) ? (rem >> 6) : rem / size;
    j = chunk * 64u + (rem - k * size);
    ordinal_pixel(J, j, px, py);
    }

// Primary ray i of job J (the job whose ray range holds i).
__device__ __forceinline__ void primary_ray_of(const JobDev& J, uint32_t i, uint32_t& k, uint32_t& px, uint32_t& py,
                                               uint32_t& jord, v3& o, v3& d)
    {
    primary_order(J, i - J.ray_base, k, px, py, jord);
    camera_ray(J.cam, J.W, J.H, J.xs_off, J.inv_h, J.seed, px, py, J.n_base + k + 1, true, o, d);
    }

__device__ __forceinline__ void primary_ray(const JobDev* __restrict__ jobs, uint32_t n_jobs, uint32_t i, uint32_t& job,
                                            uint32_t& k, uint32_t& px, uint32_t& py, uint32_t& jord, v3& o, v3& d)
    {
    job = find_job(jobs, n_jobs, i);
    primary_ray_of(jobs[job], i, k, px, py, jord, o, d);
    }

__device__ __forceinline__ void primary_ray(const JobDev* __restrict__ jobs, uint32_t n_jobs, uint32_t i, uint32_t& job,
                                            uint32_t& k, uint32_t& pixel, uint32_t& jord, v3& o, v3& d)
    {
    uint32_t px, py;
    primary_ray(jobs, n_jobs, i, job, k, px, py, jord, o, d);
    pixel = py * jobs[job].W + px;
    }

// Persistent traversal: each wavefront claims `chunk` rays (a multiple of 64)
// per atomic on the work counter and claims its next chunk before tracing the
// current one, so the atomic's round trip overlaps the traversal.
// Work claims of the persistent traversal.  Without bands the wavefronts
// claim `chunk` rays at a time from [0, n).  With bands the queue's 16
// segments hold the rays of 16 image bands (k_primary) and the dense range is
// cut into 8 ranges of two segments, one per XCD: a workgroup starts on the
// range of the XCD it runs on (workgroups are dispatched to the 8 XCDs in
// turn) and moves to the next range once its own is exhausted, so each XCD's
// L2 mostly serves the hierarchy around two bands.  Each range's counter only
// grows, so every wavefront stops after at most 8 exhausted ranges.
struct Claims
    {
    uint32_t* work; // one counter per range, QSEG_STRIDE words apart
    uint32_t chunk, n, r, ranges, tried;
    // q.pre[2 k] (constant indices only, so the prefix stays in registers)
    __device__ __forceinline__ static uint32_t even_pre(const Queue& q, uint32_t k)
        {
        uint32_t v = 0;
#pragma unroll
        for (uint32_t j = 1; j < QSEG / 2; j++)
            v = (k == j) ? q.pre[2 * j] : v;
        return v;
        }
    __device__ __forceinline__ uint32_t lo(const Queue& q, uint32_t k) const
        {
        return ranges == 1 ? 0u : even_pre(q, k);
        }
    __device__ __forceinline__ uint32_t hi(const Queue& q, uint32_t k) const
        {
        return (ranges == 1 || k == ranges - 1) ? n : even_pre(q, k + 1);
        }
    // next claimed [base, end) of the calling wavefront, or base = EMPTY
    __device__ __forceinline__ void next(const Queue& q, uint32_t lane, uint32_t& base, uint32_t& end)
        {
        while (tried < ranges)
            {
            uint32_t c = 0;
            if (lane == 0)
                c = atomicAdd(work + r * QSEG_STRIDE, chunk);
            c = __shfl(c, 0, 64);
            const uint32_t a = lo(q, r), b = hi(q, r);
            if (c < b - a)
                {
                base = a + c;
                end = min(b, base + chunk);
                return;
                }
            r = (r + 1 == ranges) ? 0u : r + 1;
            tried++;
            }
        base = EMPTY;
        end = EMPTY;
        }
    };

// The search is latency-bound (dependent node loads), so wavefronts in flight
// matter: the spheres-only variant is held to 71 registers (7 wavefronts per
// SIMD, no spills; 8 spill and run 7 % slower), the others to 64 (8 per SIMD,
// a few values spill; on C3 the traversal takes 51 ms per frame at 8, 53 at
// 7, 57 at 6 and 65 at the 5 its 87 registers allow).
template<int BLOCK, int SS, bool COUNT, bool PRIMARY, uint32_t TYPES>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(TYPES == T_SPHERES ? 7 : 8))) void k_traverse(
    const JobDev* __restrict__ jobs, uint32_t n_jobs, Queue q, uint32_t n, const QueueView* __restrict__ view,
    uint4* hits, uint32_t* work, uint32_t chunk, uint32_t bands, uint2* gstack, uint32_t gstride,
    unsigned long long* counters)
    {
    if (!take_view(view, VIEW_BOUNCE, q, n))
        return;
    __shared__ uint2 lds[SS * BLOCK];
    const uint32_t gtid = blockIdx.x * BLOCK + threadIdx.x;
    LdsStack<SS> st;
    st.lds = lds + threadIdx.x;
    st.lstride = BLOCK;
    st.glob = gstack + gtid;
    st.gstride = gstride;
    st.sp = 0;
    const uint32_t lane = lane_id();
    uint32_t n_nodes = 0, n_prims = 0, err = 0, n_rays = 0;
    Claims cl;
    cl.work = work;
    cl.chunk = chunk;
    cl.n = n;
    cl.ranges = (bands && !PRIMARY) ? 8u : 1u;
    cl.r = (cl.ranges == 8u) ? (blockIdx.x & 7u) : 0u;
    cl.tried = 0;
    uint32_t base, end;
    cl.next(q, lane, base, end);
    while (base != EMPTY)
        {
        uint32_t nbase, nend; // the next chunk, claimed before tracing this one
        cl.next(q, lane, nbase, nend);
        for (uint32_t sub = base; sub < end; sub += 64)
        {
        const uint32_t i = sub + lane;
        const uint32_t sl = PRIMARY ? 0u : q_slot(q, i, sub, min(end, sub + 64u) - 1u);
        if (i < end)
            {
            v3 o, d;
            float tnear;
            uint32_t job;
            if (PRIMARY)
                {
                uint32_t k, pixel, jord;
                primary_ray(jobs, n_jobs, i, job, k, pixel, jord, o, d);
                tnear = 1e-3f;
                }
            else
                {
                const float4 a = q.o[sl];
                const float4 b = q.d[sl];
                o = mk3(a.x, a.y, a.z);
                d = mk3(b.x, b.y, b.z);
                tnear = a.w;
                job = __float_as_uint(b.w);
                }
            const SceneDev& sc = jobs[job].scene;
            HitT h;
            if constexpr (TYPES == T_SPHERES)
                trace_closest<LdsStack<SS>, COUNT, TYPES>(sc, o, d, tnear, h, st, n_nodes, n_prims, err);
            else
                trace_closest_grouped<LdsStack<SS>, COUNT, TYPES>(sc, o, d, tnear, h, st, n_nodes, n_prims, err);
            hits[i] = make_uint4(__float_as_uint(h.t), h.geom, h.prim, 0);
            if (COUNT)
                n_rays++;
            }
        }
        base = nbase;
        end = nend;
        }
    if (COUNT)
        {
        const int c = PRIMARY ? CTR_PRIMARY : CTR_SECONDARY;
        atomicAdd(counters + c + 0, (unsigned long long)n_rays);
        atomicAdd(counters + c + 1, (unsigned long long)n_nodes);
        atomicAdd(counters + c + 2, (unsigned long long)n_prims);
        }
    if (err)
        atomicOr((unsigned int*)(counters + CTR_OVERFLOW), 1u);
    }

// ---------------------------------------------------------------- shading
// Shading attributes of the closest hit: re-evaluate the winning primitive.
template<uint32_t TYPES = T_ALL>
__device__ __forceinline__ void hit_attributes(const SceneDev& sc, const HitT& h, v3 o, v3 d, float tnear, Attr& at,
                                               const GeomDev*& g)
    {
    g = sc.geoms + h.geom;
    if (TYPES == T_SPHERES || g->type == G_SPHERE)
        {
        // the search's t is the sphere solve's: no second solve
        // (GeometrySphere.cc:81-117 attributes, unflipped Ng)
        const float4 s = g->a[h.prim];
        const v3 p = xyz(s);
        const v3 w = cross(p - o, d);
        at.Ng = mk3(o.x + h.t * d.x - p.x, o.y + h.t * d.y - p.y, o.z + h.t * d.z - p.z);
        at.d = s.w - sqrtf(dot(w, w));
        at.color = ld3(g->color, h.prim);
        return;
        }
    float t2;
    intersect_prim<true, TYPES>(*g, h.prim, o, d, tnear, t2, &at);
    }

// Depth-0 shading: one primary hit spawns the L light-sample paths
// (fresnel/cpu/TracerPath.cc:141-155 reuse the cached primary hit).
// Depth-0 shading of a workgroup's primary hits: each hit spawns the job's L
// light-sample paths (fresnel/cpu/TracerPath.cc:141-155 reuse the cached
// primary hit); continuing paths are appended to `out` with one atomic per
// workgroup and light sample, terminated ones write their record.  Every
// thread of the workgroup must call it.
template<int BS, uint32_t TYPES>
__device__ __forceinline__ void shade_primary_block(const JobDev* __restrict__ jobs, bool valid, uint32_t job,
                                                    uint32_t k, uint32_t pixel, uint32_t jord, v3 o, v3 d,
                                                    const HitT& h,
                                                    const Queue& out, uint32_t seg, uint4* records, uint32_t* sh)
    {
    const float tnear = 1e-3f;
    uint32_t L = 0;
    const JobDev& J = jobs[job];
    if (valid)
        {
        L = J.L;
        if (h.geom == EMPTY)
            {
            // primary miss: background for the whole camera sample (TracerPathMethods.h:44-49)
            records[rec_index(J, k, jord, 0)] = make_rec(mk3(0, 0, 0), REC_BG);
            L = 0;
            }
        }
    Attr at;
    const GeomDev* g = nullptr;
    const Mat* m = nullptr;
    if (L)
        {
        hit_attributes<TYPES>(J.scene, h, o, d, tnear, at, g);
        m = (at.d >= g->outline_width) ? &g->mat : &g->outline;
        }
    const uint32_t Lmax = block_max<BS>(L, sh);
    const uint32_t n_sample = valid ? J.n_base + k + 1 : 0;
    for (uint32_t ls = 0; ls < Lmax; ls++)
        {
        bool cont = false;
        v3 att = mk3(1.0f, 1.0f, 1.0f), no, nd, add;
        if (ls < L)
            {
            bool addv;
            cont = path_hit(*m, at, o, d, h.t, pixel, J.seed, 0, (n_sample - 1) * J.L + ls, att, no, nd, add, addv);
            if (!cont)
                records[rec_index(J, k, jord, ls)] = make_rec(addv ? add : mk3(0, 0, 0), addv ? REC_ADD : REC_NONE);
            }
        const uint32_t slot = block_append<BS>(cont, out, seg, sh);
        if (cont && slot != EMPTY)
            {
            out.o[slot] = make_float4(no.x, no.y, no.z, 1e-3f);
            out.d[slot] = make_float4(nd.x, nd.y, nd.z, __uint_as_float(job));
            out.att[slot] = make_float4(att.x, att.y, att.z, 0.0f);
            out.meta[slot] = make_uint4(job, pixel, k | (ls << 16), 1);
            }
        }
    }

template<int BS, uint32_t TYPES>
__global__ __launch_bounds__(BS) void k_shade_primary(const JobDev* __restrict__ jobs, uint32_t n_jobs,
                                                      const uint4* __restrict__ hits, uint32_t n, Queue out,
                                                      uint4* records)
    {
    __shared__ uint32_t sh[BS / 64 + 1];
    const uint32_t i = blockIdx.x * BS + threadIdx.x;
    const bool valid = i < n;
    uint32_t job = 0, pixel = 0, k = 0, jord = 0;
    HitT h;
    h.geom = EMPTY;
    v3 o = mk3(0, 0, 0), d = mk3(0, 0, 1);
    if (valid)
        {
        // the primary ray is regenerated rather than stored (bit-identical)
        primary_ray(jobs, n_jobs, i, job, k, pixel, jord, o, d);
        const uint4 hh = hits[i];
        h.t = __uint_as_float(hh.x);
        h.geom = hh.y;
        h.prim = hh.z;
        }
    shade_primary_block<BS, TYPES>(jobs, valid, job, k, pixel, jord, o, d, h, out, blockIdx.x % QSEG, records, sh);
    }

// Shade one hit or miss of a path at depth >= 1.  Returns true when the path
// continues with (no, nd, att); otherwise rec holds its termination record.
template<uint32_t TYPES>
__device__ inline bool shade_bounce(const JobDev& J, const HitT& h, v3 o, v3 d, float tnear, uint32_t pixel,
                                    uint32_t k, uint32_t ls, uint32_t depth, v3& att, v3& no, v3& nd, uint4& rec)
    {
    if (h.geom == EMPTY)
        {
        // miss: lights whose cone contains the direction (TracerPathMethods.h:50-73)
        rec = make_rec(att, REC_LIGHTS | (light_mask(J.lights, d) << 2));
        return false;
        }
    Attr at;
    const GeomDev* g;
    hit_attributes<TYPES>(J.scene, h, o, d, tnear, at, g);
    const Mat& m = (at.d >= g->outline_width) ? g->mat : g->outline;
    const uint32_t n_sample = J.n_base + k + 1;
    v3 add;
    bool addv;
    const bool cont = path_hit(m, at, o, d, h.t, pixel, J.seed, depth, (n_sample - 1) * J.L + ls, att, no, nd, add,
                               addv);
    if (!cont)
        rec = make_rec(addv ? add : mk3(0, 0, 0), addv ? REC_ADD : REC_NONE);
    return cont;
    }

// Shade the queued hit h of job J (ray o, d from tnear, entry meta mt, path
// attenuation att): returns whether its path continues with (no, nd, att); a
// terminated path's record is written.
template<uint32_t TYPES>
__device__ __forceinline__ bool shade_queued(const JobDev& J, const HitT& h, v3 o, v3 d, float tnear, const uint4& mt,
                                             uint4* records, uint32_t max_depth, unsigned long long* counters,
                                             v3& att, v3& no, v3& nd)
    {
    const uint32_t k = mt.z & 0xffffu, ls = mt.z >> 16;
    uint4 rec;
    bool cont = shade_bounce<TYPES>(J, h, o, d, tnear, mt.y, k, ls, mt.w, att, no, nd, rec);
    if (cont && mt.w + 1 >= max_depth)
        {
        // depth cap (the reference has none): stop without contribution
        cont = false;
        rec = make_rec(mk3(0, 0, 0), REC_NONE);
        atomicAdd(counters + CTR_CAPPED, 1ull);
        }
    if (!cont)
        records[rec_index(J, k, raster_ordinal(J, mt.y), ls)] = rec;
    return cont;
    }

// Hit list of one shading unit: queue index, slot, geometry, primitive and
// distance of each ray that hit, in queue order.
template<int BS>
struct ShadeLds
    {
    uint4 ent[BS];
    float t[BS];
    uint32_t sh[BS / 64 + 1];
    };

// Shade the BS queued rays of `unit`.  A ray that missed ends on its own
// thread with its lights record; the rays that hit are listed in LDS in queue
// order and shaded by the workgroup's first threads, so a wavefront's lanes
// all run the surface interaction (on a lattice of spheres most bounce rays
// escape, and a wavefront of mixed rays would run it with few lanes).  The
// continuing paths are appended to segment unit % QSEG in queue order.  Each
// path's arithmetic depends only on its own entry, so the result is the same
// as shading every ray on its own thread.  A wavefront whose entries all
// belong to one job (every one of a single-job batch, most of a multi-job
// one, whose queues are filled chunk by chunk) reads the job record through
// scalar loads; otherwise each ray reads its own job's record.  Every thread
// of the workgroup must call it.
template<int BS, uint32_t TYPES>
__device__ __forceinline__ void shade_unit(const JobDev* __restrict__ jobs, uint32_t n_jobs, const Queue& in,
                                           uint32_t n, uint32_t unit, const uint4* __restrict__ hits,
                                           const Queue& out, uint4* records, uint32_t max_depth,
                                           unsigned long long* counters, ShadeLds<BS>& L)
    {
    const uint32_t first = unit * BS;
    const uint32_t i = first + threadIdx.x;
    const uint32_t sl = q_slot(in, i, first, min(n, first + BS) - 1u);
    bool hit = false;
    uint4 hh = make_uint4(0, 0, 0, 0);
    if (i < n)
        {
        hh = hits[i];
        hit = hh.y != EMPTY;
        if (!hit)
            {
            // miss: lights whose cone contains the direction (TracerPathMethods.h:50-73)
            const uint4 mt = in.meta[sl];
            const float4 b = in.d[sl];
            const float4 a4 = in.att[sl];
            const uint32_t j0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)mt.x);
            const JobDev& J = (n_jobs == 1 || __ballot(mt.x != j0) == 0) ? jobs[j0] : jobs[mt.x];
            const uint32_t k = mt.z & 0xffffu, ls = mt.z >> 16;
            records[rec_index(J, k, raster_ordinal(J, mt.y), ls)] =
                make_rec(mk3(a4.x, a4.y, a4.z), REC_LIGHTS | (light_mask(J.lights, mk3(b.x, b.y, b.z)) << 2));
            }
        }
    uint32_t n_hit;
    const uint32_t pos = block_rank<BS>(hit, L.sh, n_hit);
    if (hit)
        {
        L.ent[pos] = make_uint4(i, sl, hh.y, hh.z);
        L.t[pos] = __uint_as_float(hh.x);
        }
    __syncthreads();
    bool cont = false;
    v3 att, no, nd;
    uint4 mt = make_uint4(0, 0, 0, 0);
    if (threadIdx.x < n_hit)
        {
        const uint4 e = L.ent[threadIdx.x];
        HitT h;
        h.t = L.t[threadIdx.x];
        h.geom = e.z;
        h.prim = e.w;
        const uint32_t hs = e.y;
        mt = in.meta[hs];
        const float4 a = in.o[hs];
        const float4 b = in.d[hs];
        const float4 at4 = in.att[hs];
        att = mk3(at4.x, at4.y, at4.z);
        const v3 o = mk3(a.x, a.y, a.z), d = mk3(b.x, b.y, b.z);
        const uint32_t j0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)mt.x);
        if (n_jobs == 1 || __ballot(mt.x != j0) == 0)
            cont = shade_queued<TYPES>(jobs[j0], h, o, d, a.w, mt, records, max_depth, counters, att, no, nd);
        else
            cont = shade_queued<TYPES>(jobs[mt.x], h, o, d, a.w, mt, records, max_depth, counters, att, no, nd);
        }
    const uint32_t slot = block_append<BS>(cont, out, unit % QSEG, L.sh);
    if (cont && slot != EMPTY)
        {
        out.o[slot] = make_float4(no.x, no.y, no.z, 1e-3f);
        out.d[slot] = make_float4(nd.x, nd.y, nd.z, __uint_as_float(mt.x));
        out.att[slot] = make_float4(att.x, att.y, att.z, 0.0f);
        out.meta[slot] = make_uint4(mt.x, mt.y, mt.z, mt.w + 1);
        }
    }

// Units of BS rays from unit0 + blockIdx.x; LOOP: grid-stride over every
// further unit, else one unit per workgroup (the host sizes the grid from the
// live-path fraction recent renders had at this bounce and launches the LOOP
// form from the first unit beyond it, so no ray is left out).  Unit u appends
// to segment u % QSEG.  The LOOP form is held to 7 wavefronts per SIMD: its
// loop would otherwise keep the loop's constants (Philox, the sine/cosine
// polynomial) in registers across iterations instead of rematerializing them.
template<int BS, uint32_t TYPES, bool LOOP>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(LOOP ? 7 : 1))) void k_shade(
    const JobDev* __restrict__ jobs, uint32_t n_jobs, Queue in, uint32_t n, const QueueView* __restrict__ view,
    uint32_t unit0, const uint4* __restrict__ hits, Queue out, uint4* records, uint32_t max_depth,
    unsigned long long* counters)
    {
    if (!take_view(view, VIEW_BOUNCE, in, n))
        return;
    __shared__ ShadeLds<BS> lds;
    const uint32_t units = (n + BS - 1) / BS;
    for (uint32_t unit = unit0 + blockIdx.x; unit < units; unit += gridDim.x)
        {
        shade_unit<BS, TYPES>(jobs, n_jobs, in, n, unit, hits, out, records, max_depth, counters, lds);
        if (!LOOP)
            break;
        __syncthreads(); // the next unit rewrites the hit list
        }
    }

// Primary rays as wavefront packets.  The 64 rays of a wavefront come from
// one 8x8 pixel tile of one camera sample, so they need nearly the same
// nodes: the wavefront walks one node sequence (uniform node index, one
// fetch per visit) and every lane tests the visited node's children against
// its own ray, with its own closest hit, cull and leaf tests.  A lane acts
// only on nodes it wants (its own slab and cull tests passed at the parent),
// so each ray's result is the one the per-ray search returns; the wavefront
// descends into a child when any lane wants it, nearer child first by lane
// majority, and pushes the other with each lane's own cull bound.  Groups of
// 64 rays that straddle two jobs fall back to the per-ray search.
constexpr uint32_t UNWANTED = 0x80000000u; // stack flag: this lane does not need the node
typedef const float __attribute__((address_space(4))) ConstF;
typedef const uint32_t __attribute__((address_space(4))) ConstU;
typedef const GeomDev __attribute__((address_space(4))) ConstGeom;

// Leaf test of a primitive every lane of the wavefront shares: the geometry
// record and the primitive's data arrive through scalar loads.
template<uint32_t TYPES>
__device__ __forceinline__ bool leaf_test_uniform(const GeomDev* geoms, uint32_t slot, uint32_t prim, v3 o, v3 d,
                                                  float tnear, float& t)
    {
    if constexpr (TYPES == T_SPHERES)
        {
        const ConstGeom* g = (const ConstGeom*)geoms + slot;
        const ConstF* sp = (const ConstF*)g->a + 4 * (size_t)prim;
        float de;
        bool first;
        return sphere_t(o, d, mk3(sp[0], sp[1], sp[2]), sp[3], t, de, first) && (tnear < t);
        }
    else
        {
        const GeomDev g = uniform_geom(geoms, slot);
        return intersect_prim<false, TYPES, true>(g, prim, o, d, tnear, t, nullptr);
        }
    }

// Tile-list entries examined (counting kernels only): all, and those of
// spherocylinders, convex polyhedra and mesh triangles.
struct EntryCount
    {
    uint32_t all = 0, cyl = 0, poly = 0, mesh = 0;
    };

// Closest hits of one wavefront's 64 primary rays (lane ray o, d of job `job`
// at pixel px, py; `valid` lanes only): tile lists when all 64 rays form one
// full tile of a job that has them, wavefront-packet BVH search when they
// belong to one job, per-ray search otherwise.
template<int SS, bool COUNT, uint32_t TYPES>
__device__ __forceinline__ void primary_group_hit(const JobDev* __restrict__ jobs, bool valid, uint32_t job,
                                                  uint32_t px, uint32_t py, v3 o, v3 d, LdsStack<SS>& st, HitT& best,
                                                  uint32_t& n_nodes, uint32_t& n_prims, EntryCount& n_entries,
                                                  uint32_t& err)
    {
    const uint32_t job0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)job);
    const bool one_job = __ballot(valid && job != job0) == 0;
    // the tile of this group, if all 64 rays belong to one full 8x8 tile
    uint32_t tile = EMPTY;
    if (one_job && jobs[job0].use_tiles)
        {
        const JobDev& J = jobs[job0];
        if (valid && px < J.tiles_x * TILE && py < J.tiles_y * TILE)
            tile = (py / TILE) * J.tiles_x + px / TILE;
        const uint32_t t0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)tile);
        if (__ballot(tile != t0) != 0)
            tile = EMPTY;
        tile = (uint32_t)__builtin_amdgcn_readfirstlane((int)tile);
        }
    if (!one_job)
        {
        if (valid)
            trace_closest<LdsStack<SS>, COUNT, TYPES>(jobs[job].scene, o, d, 1e-3f, best, st, n_nodes,
                                                      n_prims, err);
        }
    else if (tile != EMPTY)
        {
        // candidate list of the tile (kernels/tiles.hip), nearest bound
        // first; the next entry is loaded while the current one is tested
        const JobDev& J = jobs[job0];
        const GeomDev* __restrict__ geoms = J.scene.geoms;
        const v3 inv = mk3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
        const float tnear = 1e-3f;
        best.t = __builtin_huge_valf();
        best.geom = EMPTY;
        best.prim = EMPTY;
        const uint32_t e0 = ((const ConstU*)J.tile_start)[tile];
        const uint32_t e1 = ((const ConstU*)J.tile_end)[tile];
        const ConstU* ent = (const ConstU*)J.tile_entries;
        uint32_t cur[8], nxt[8];
        if (e0 < e1)
            for (int k = 0; k < 8; k++)
                cur[k] = ent[8 * (size_t)e0 + k];
        for (uint32_t e = e0; e < e1; e++)
            {
            const size_t en = (e + 1 < e1) ? (size_t)e + 1 : (size_t)e;
            for (int k = 0; k < 8; k++)
                nxt[k] = ent[8 * en + k];
            const float q = __uint_as_float(cur[0]);
            const bool want = !(q > best.t);
            if (__ballot(want) == 0)
                break;
            if (COUNT)
                {
                n_entries.all++;
                if constexpr (TYPES != T_SPHERES)
                    if (!(cur[7] & NODE_REF))
                        {
                        const uint32_t type = J.scene.geoms[__float_as_uint(J.prim_lo[cur[7]].w)].type;
                        n_entries.cyl += type == G_CYLINDER;
                        n_entries.poly += type == G_POLYHEDRON;
                        n_entries.mesh += type == G_MESH;
                        }
                }
            if (want)
                {
                float lx, ly, lz, hx, hy, hz;
                float4 sph;
                uint32_t slot, prim;
                if constexpr (TYPES == T_SPHERES)
                    {
                    sph = make_float4(__uint_as_float(cur[1]), __uint_as_float(cur[2]),
                                      __uint_as_float(cur[3]), __uint_as_float(cur[4]));
                    // the sphere's bounds, as the builder computes them (GeometrySphere.cc:60-73)
                    lx = sph.x - sph.w;
                    ly = sph.y - sph.w;
                    lz = sph.z - sph.w;
                    hx = sph.x + sph.w;
                    hy = sph.y + sph.w;
                    hz = sph.z + sph.w;
                    slot = cur[5];
                    prim = cur[6];
                    }
                else
                    {
                    lx = __uint_as_float(cur[1]);
                    ly = __uint_as_float(cur[2]);
                    lz = __uint_as_float(cur[3]);
                    hx = __uint_as_float(cur[4]);
                    hy = __uint_as_float(cur[5]);
                    hz = __uint_as_float(cur[6]);
                    slot = 0;
                    prim = 0;
                    }
                float tmin, tmax;
                slab(o, inv, lx, ly, lz, hx, hy, hz, tmin, tmax);
                if ((tmin <= tmax) && (tmax >= tnear))
                    {
                    float t;
                    bool hit;
                    if constexpr (TYPES == T_SPHERES)
                        {
                        float de;
                        bool first;
                        if (COUNT)
                            n_prims++;
                        hit = sphere_t(o, d, mk3(sph.x, sph.y, sph.z), sph.w, t, de, first) && (tnear < t);
                        }
                    else
                        {
                        const uint32_t r = cur[7];
                        hit = false;
                        if (r & NODE_REF)
                            {
                            // a subtree: the BVH search below its root, from this
                            // lane's closest hit so far (same slab and cull tests)
                            if (!(cull_bound(tmin, tmax, J.scene.cull_margin) > best.t))
                                {
                                TraceState ts;
                                ts.o = o;
                                ts.d = d;
                                ts.inv = inv;
                                ts.tnear = tnear;
                                ts.margin = J.scene.cull_margin;
                                ts.best = best;
                                ts.node = r & ~NODE_REF;
                                ts.sph = false;
                                st.sp = 0;
                                while (trace_visit2<LdsStack<SS>, COUNT, TYPES>(J.scene.nodes, geoms, ts, st, n_nodes,
                                                                                n_prims, err))
                                    {
                                    }
                                best = ts.best;
                                }
                            }
                        else
                            {
                            if (COUNT)
                                n_prims++;
                            // the entry's primitive is the same in every lane: its
                            // record and data arrive through scalar loads
                            slot = __float_as_uint(((ConstF*)J.prim_lo)[4 * (size_t)r + 3]);
                            prim = __float_as_uint(((ConstF*)J.prim_hi)[4 * (size_t)r + 3]);
                            const GeomDev g = uniform_geom(geoms, slot);
                            hit = intersect_prim<false, TYPES, true>(g, prim, o, d, tnear, t, nullptr);
                            }
                        }
                    if (hit && better(t, slot, prim, best))
                        {
                        best.t = t;
                        best.geom = slot;
                        best.prim = prim;
                        }
                    }
                }
            for (int k = 0; k < 8; k++)
                cur[k] = nxt[k];
            }
        }
    else
        {
        const SceneDev& sc = jobs[job0].scene;
        const Node* __restrict__ nodes = sc.nodes;
        const GeomDev* __restrict__ geoms = sc.geoms;
        const v3 inv = mk3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
        const float tnear = 1e-3f;
        const float margin = sc.cull_margin;
        best.t = __builtin_huge_valf();
        best.geom = EMPTY;
        best.prim = EMPTY;
        st.sp = 0;
        uint32_t node = sc.root;
        bool want = valid;
        while (node != EMPTY)
            {
            // the node index is uniform: fetch through the constant address
            // space so the node arrives in scalar registers
            node = (uint32_t)__builtin_amdgcn_readfirstlane((int)node);
            const ConstF* nf = (const ConstF*)(nodes + node);
            const ConstU* nu = (const ConstU*)(nodes + node);
            const float4 A = make_float4(nf[0], nf[1], nf[2], nf[3]);
            const float4 B = make_float4(nf[4], nf[5], nf[6], nf[7]);
            const float4 C = make_float4(nf[8], nf[9], nf[10], nf[11]);
            const uint4 lk = make_uint4(nu[12], nu[13], nu[14], nu[15]);
            if (COUNT && want)
                n_nodes++;
            float t0min, t0max, t1min, t1max;
            slab(o, inv, A.x, A.y, A.z, A.w, B.x, B.y, t0min, t0max);
            slab(o, inv, B.z, B.w, C.x, C.y, C.z, C.w, t1min, t1max);
            const float c0 = cull_bound(t0min, t0max, margin);
            const float c1 = cull_bound(t1min, t1max, margin);
            bool w0 = want && (t0min <= t0max) && (t0max >= tnear) && !(c0 > best.t);
            bool w1 = want && (t1min <= t1max) && (t1max >= tnear) && !(c1 > best.t);
            if (lk.x & LEAF)
                {
                if (w0)
                    {
                    const uint32_t prim = lk.x & ~LEAF;
                    float t;
                    if (COUNT)
                        n_prims++;
                    if (leaf_test_uniform<TYPES>(geoms, lk.z, prim, o, d, tnear, t) && better(t, lk.z, prim, best))
                        {
                        best.t = t;
                        best.geom = lk.z;
                        best.prim = prim;
                        }
                    }
                w0 = false;
                }
            if (lk.y & LEAF)
                {
                if (w1)
                    {
                    const uint32_t prim = lk.y & ~LEAF;
                    float t;
                    if (COUNT)
                        n_prims++;
                    if (leaf_test_uniform<TYPES>(geoms, lk.w, prim, o, d, tnear, t) && better(t, lk.w, prim, best))
                        {
                        best.t = t;
                        best.geom = lk.w;
                        best.prim = prim;
                        }
                    }
                w1 = false;
                }
            w0 = w0 && !(c0 > best.t);
            w1 = w1 && !(c1 > best.t);
            const uint64_t b0 = __ballot(w0), b1 = __ballot(w1);
            if (b0 && b1)
                {
                const uint64_t both = b0 & b1;
                const uint64_t pref1 = __ballot(w0 && w1 && t1min < t0min);
                const bool near1 = 2 * __popcll(pref1) > __popcll(both);
                const uint32_t far_node = near1 ? lk.x : lk.y;
                const bool far_want = near1 ? w0 : w1;
                st.push(far_want ? far_node : (far_node | UNWANTED), near1 ? c0 : c1, err);
                node = near1 ? lk.y : lk.x;
                want = near1 ? w1 : w0;
                continue;
                }
            if (b0)
                {
                node = lk.x;
                want = w0;
                continue;
                }
            if (b1)
                {
                node = lk.y;
                want = w1;
                continue;
                }
            node = EMPTY;
            while (st.sp > 0)
                {
                uint32_t e;
                float cv;
                st.pop(e, cv);
                const bool w = !(e & UNWANTED) && !(cv > best.t);
                if (__ballot(w) != 0)
                    {
                    node = (uint32_t)__builtin_amdgcn_readfirstlane((int)e) & ~UNWANTED;
                    want = w;
                    break;
                    }
                }
            }
        }
    }

template<int BLOCK, int SS, bool COUNT, uint32_t TYPES>
__global__ __launch_bounds__(BLOCK) void k_traverse_packet(const JobDev* __restrict__ jobs, uint32_t n_jobs,
                                                           uint32_t n, uint4* hits, uint32_t* work, uint32_t chunk,
                                                           uint2* gstack, uint32_t gstride,
                                                           unsigned long long* counters)
    {
    __shared__ uint2 lds[SS * BLOCK];
    const uint32_t gtid = blockIdx.x * BLOCK + threadIdx.x;
    LdsStack<SS> st;
    st.lds = lds + threadIdx.x;
    st.lstride = BLOCK;
    st.glob = gstack + gtid;
    st.gstride = gstride;
    st.sp = 0;
    const uint32_t lane = lane_id();
    uint32_t n_nodes = 0, n_prims = 0, err = 0, n_rays = 0;
    EntryCount n_entries;
    uint32_t claim = 0;
    if (lane == 0)
        claim = atomicAdd(work, chunk);
    uint32_t base = __shfl(claim, 0, 64);
    while (base < n)
        {
        uint32_t next = 0;
        if (lane == 0)
            next = atomicAdd(work, chunk);
        const uint32_t end = min(n, base + chunk);
        for (uint32_t sub = base; sub < end; sub += 64)
            {
            const uint32_t i = sub + lane;
            const bool valid = i < end;
            v3 o = mk3(0.0f, 0.0f, 0.0f), d = mk3(0.0f, 0.0f, 1.0f);
            uint32_t job = 0, px = 0, py = 0;
            if (valid)
                {
                uint32_t k, jord;
                primary_ray(jobs, n_jobs, i, job, k, px, py, jord, o, d);
                }
            HitT best;
            primary_group_hit<SS, COUNT, TYPES>(jobs, valid, job, px, py, o, d, st, best, n_nodes, n_prims, n_entries,
                                                err);
            if (valid)
                {
                hits[i] = make_uint4(__float_as_uint(best.t), best.geom, best.prim, 0);
                if (COUNT)
                    n_rays++;
                }
            }
        base = __shfl(next, 0, 64);
        }
    if (COUNT)
        {
        atomicAdd(counters + CTR_PRIMARY + 0, (unsigned long long)n_rays);
        atomicAdd(counters + CTR_PRIMARY + 1, (unsigned long long)n_nodes);
        atomicAdd(counters + CTR_PRIMARY + 2, (unsigned long long)n_prims);
        atomicAdd(counters + CTR_TILE_ENTRIES, (unsigned long long)n_entries.all);
        atomicAdd(counters + CTR_TILE_ENTRIES + 1, (unsigned long long)n_entries.cyl);
        atomicAdd(counters + CTR_TILE_ENTRIES + 2, (unsigned long long)n_entries.poly);
        atomicAdd(counters + CTR_TILE_ENTRIES + 3, (unsigned long long)n_entries.mesh);
        }
    if (err)
        atomicOr((unsigned int*)(counters + CTR_OVERFLOW), 1u);
    }

// Primary rays traced and shaded in one persistent kernel: the workgroup
// claims `groups` x its size rays per atomic, each wavefront traces `groups`
// groups of 64 rays (primary_group_hit) and the workgroup shades them
// together (shade_primary_block), so no hit is stored and no primary ray is
// generated twice.  Same arithmetic as k_traverse_packet + k_shade_primary.
// The spheres-only variant is held to 128 registers (4 wavefronts per SIMD; a
// few values spill): with the double-precision sine and cosine of the depth-0
// lobe sample it would take 147 (3 wavefronts) and run 1.7x longer, and at 5
// wavefronts it spills more and runs 17 % longer.  The variant for other
// scenes, whose list search waits on primitive loads, runs best at 6 (C3's
// primary kernel: 31.2 ms per frame at 4, 29.8 at 5, 27.7 at 6, 28.2 at 8).
template<int BLOCK, int SS, bool COUNT, uint32_t TYPES>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(TYPES == T_SPHERES ? 4 : 6))) void k_primary(
    const JobDev* __restrict__ jobs, uint32_t n_jobs, uint32_t n, uint32_t* work, uint32_t groups, uint32_t bands,
    uint2* gstack, uint32_t gstride, Queue out, uint4* records, unsigned long long* counters)
    {
    __shared__ uint2 lds[SS * BLOCK];
    __shared__ uint32_t sh[BLOCK / 64 + 1];
    __shared__ uint32_t s_claim;
    const uint32_t gtid = blockIdx.x * BLOCK + threadIdx.x;
    LdsStack<SS> st;
    st.lds = lds + threadIdx.x;
    st.lstride = BLOCK;
    st.glob = gstack + gtid;
    st.gstride = gstride;
    st.sp = 0;
    const uint32_t lane = lane_id();
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t chunk = BLOCK * groups;
    uint32_t n_nodes = 0, n_prims = 0, err = 0, n_rays = 0;
    EntryCount n_entries;
    // Chunk u of n_units goes to queue segment u % QSEG, or with bands to
    // segment u / G (16 contiguous bands of G chunks, i.e. of the image); the
    // claims then visit the bands in turn, claim c -> chunk (c % QSEG) G + c /
    // QSEG, so the chunks in flight still append to all 16 segments.
    const uint32_t n_units = (n + chunk - 1) / chunk;
    const uint32_t G = (n_units + QSEG - 1) / QSEG;
    const uint32_t n_claims = bands ? QSEG * G : n_units;
    if (threadIdx.x == 0)
        s_claim = atomicAdd(work, 1u);
    __syncthreads();
    uint32_t c = s_claim;
    __syncthreads();
    while (c < n_claims)
        {
        if (threadIdx.x == 0)
            s_claim = atomicAdd(work, 1u); // the next chunk, claimed ahead
        const uint32_t u = bands ? (c % QSEG) * G + c / QSEG : c;
        const uint32_t seg = bands ? u / G : u % QSEG; // the chunk's output segment
        const uint32_t base = u * chunk;
        // one job for the whole chunk (every chunk of a single-job batch, all
        // but the chunks that straddle two jobs otherwise): its index is
        // workgroup-uniform, so the job record arrives by scalar loads; the
        // branch is taken by the whole workgroup, barriers included
        const uint32_t last = min(n, base + chunk) - 1u;
        const uint32_t job_first = (uint32_t)__builtin_amdgcn_readfirstlane((int)find_job(jobs, n_jobs, base));
        const uint32_t job_last = (uint32_t)__builtin_amdgcn_readfirstlane((int)find_job(jobs, n_jobs, last));
        if (u < n_units && job_first == job_last)
            {
            const JobDev& J = jobs[job_first];
            for (uint32_t g = 0; g < groups; g++)
                {
                const uint32_t i = base + (wave * groups + g) * 64 + lane;
                const bool valid = i < n;
                v3 o = mk3(0.0f, 0.0f, 0.0f), d = mk3(0.0f, 0.0f, 1.0f);
                uint32_t k = 0, px = 0, py = 0, jord = 0;
                if (valid)
                    primary_ray_of(J, i, k, px, py, jord, o, d);
                HitT best;
                best.t = __builtin_huge_valf();
                best.geom = EMPTY;
                best.prim = EMPTY;
                primary_group_hit<SS, COUNT, TYPES>(jobs, valid, job_first, px, py, o, d, st, best, n_nodes, n_prims,
                                                    n_entries, err);
                if (COUNT && valid)
                    n_rays++;
                shade_primary_block<BLOCK, TYPES>(jobs, valid, job_first, k, py * J.W + px, jord, o, d, best, out,
                                                  seg, records, sh);
                }
            }
        else
            for (uint32_t g = 0; g < groups && u < n_units; g++)
                {
                const uint32_t i = base + (wave * groups + g) * 64 + lane;
                const bool valid = i < n;
                v3 o = mk3(0.0f, 0.0f, 0.0f), d = mk3(0.0f, 0.0f, 1.0f);
                uint32_t job = 0, k = 0, px = 0, py = 0, jord = 0;
                if (valid)
                    primary_ray(jobs, n_jobs, i, job, k, px, py, jord, o, d);
                HitT best;
                best.t = __builtin_huge_valf();
                best.geom = EMPTY;
                best.prim = EMPTY;
                primary_group_hit<SS, COUNT, TYPES>(jobs, valid, job, px, py, o, d, st, best, n_nodes, n_prims,
                                                    n_entries, err);
                if (COUNT && valid)
                    n_rays++;
                shade_primary_block<BLOCK, TYPES>(jobs, valid, job, k, py * jobs[job].W + px, jord, o, d, best, out,
                                                  seg, records, sh);
                }
        __syncthreads();
        c = s_claim;
        __syncthreads();
        }
    if (COUNT)
        {
        atomicAdd(counters + CTR_PRIMARY + 0, (unsigned long long)n_rays);
        atomicAdd(counters + CTR_PRIMARY + 1, (unsigned long long)n_nodes);
        atomicAdd(counters + CTR_PRIMARY + 2, (unsigned long long)n_prims);
        atomicAdd(counters + CTR_TILE_ENTRIES, (unsigned long long)n_entries.all);
        atomicAdd(counters + CTR_TILE_ENTRIES + 1, (unsigned long long)n_entries.cyl);
        atomicAdd(counters + CTR_TILE_ENTRIES + 2, (unsigned long long)n_entries.poly);
        atomicAdd(counters + CTR_TILE_ENTRIES + 3, (unsigned long long)n_entries.mesh);
        }
    if (err)
        atomicOr((unsigned int*)(counters + CTR_OVERFLOW), 1u);
    }

// One bounce of the secondary rays in one persistent kernel: each wavefront
// claims `chunk` rays of the queue per atomic (the next chunk ahead), traces
// each ray and shades its hit or miss (shade_bounce); continuing paths are
// appended to `out` with one atomic per 64 rays, terminated ones write their
// record.  Same arithmetic as k_traverse + k_shade.
template<int BLOCK, int SS, bool COUNT, uint32_t TYPES>
__global__ __launch_bounds__(BLOCK) void k_bounce(const JobDev* __restrict__ jobs, Queue in, uint32_t n,
                                                  const QueueView* __restrict__ view, uint32_t* work, uint32_t chunk,
                                                  uint2* gstack, uint32_t gstride, Queue out, uint4* records,
                                                  uint32_t max_depth, unsigned long long* counters)
    {
    if (!take_view(view, VIEW_BOUNCE, in, n))
        return;
    __shared__ uint2 lds[SS * BLOCK];
    const uint32_t gtid = blockIdx.x * BLOCK + threadIdx.x;
    LdsStack<SS> st;
    st.lds = lds + threadIdx.x;
    st.lstride = BLOCK;
    st.glob = gstack + gtid;
    st.gstride = gstride;
    st.sp = 0;
    const uint32_t lane = lane_id();
    uint32_t n_nodes = 0, n_prims = 0, err = 0, n_rays = 0;
    uint32_t claim = 0;
    if (lane == 0)
        claim = atomicAdd(work, chunk);
    uint32_t base = __shfl(claim, 0, 64);
    while (base < n)
        {
        uint32_t next = 0;
        if (lane == 0)
            next = atomicAdd(work, chunk);
        const uint32_t end = min(n, base + chunk);
        const uint32_t seg = (base / chunk) % QSEG; // the chunk's output segment
        for (uint32_t sub = base; sub < end; sub += 64)
            {
            const uint32_t i = sub + lane;
            const uint32_t sl = q_slot(in, i, sub, min(end, sub + 64u) - 1u);
            bool cont = false;
            v3 att, no, nd;
            uint4 mt = make_uint4(0, 0, 0, 0);
            if (i < end)
                {
                mt = in.meta[sl];
                const JobDev& J = jobs[mt.x];
                const float4 a = in.o[sl];
                const float4 b = in.d[sl];
                const float4 at4 = in.att[sl];
                const v3 o = mk3(a.x, a.y, a.z), d = mk3(b.x, b.y, b.z);
                HitT h;
                trace_closest<LdsStack<SS>, COUNT, TYPES>(J.scene, o, d, a.w, h, st, n_nodes, n_prims, err);
                if (COUNT)
                    n_rays++;
                att = mk3(at4.x, at4.y, at4.z);
                const uint32_t k = mt.z & 0xffffu, ls = mt.z >> 16;
                uint4 rec;
                cont = shade_bounce<TYPES>(J, h, o, d, a.w, mt.y, k, ls, mt.w, att, no, nd, rec);
                if (cont && mt.w + 1 >= max_depth)
                    {
                    // depth cap (the reference has none): stop without contribution
                    cont = false;
                    rec = make_rec(mk3(0, 0, 0), REC_NONE);
                    atomicAdd(counters + CTR_CAPPED, 1ull);
                    }
                if (!cont)
                    records[rec_index(J, k, raster_ordinal(J, mt.y), ls)] = rec;
                }
            const uint32_t slot = wave_append(cont, out, seg);
            if (cont && slot != EMPTY)
                {
                out.o[slot] = make_float4(no.x, no.y, no.z, 1e-3f);
                out.d[slot] = make_float4(nd.x, nd.y, nd.z, __uint_as_float(mt.x));
                out.att[slot] = make_float4(att.x, att.y, att.z, 0.0f);
                out.meta[slot] = make_uint4(mt.x, mt.y, mt.z, mt.w + 1);
                }
            }
        base = __shfl(next, 0, 64);
        }
    if (COUNT)
        {
        atomicAdd(counters + CTR_SECONDARY + 0, (unsigned long long)n_rays);
        atomicAdd(counters + CTR_SECONDARY + 1, (unsigned long long)n_nodes);
        atomicAdd(counters + CTR_SECONDARY + 2, (unsigned long long)n_prims);
        }
    if (err)
        atomicOr((unsigned int*)(counters + CTR_OVERFLOW), 1u);
    }

// Persistent traversal with per-lane refill: a lane whose ray is finished
// starts a new one as soon as at least `refill` lanes of the wavefront are
// idle (or all are), so lanes do not wait for the slowest ray of a 64-ray
// batch.  Every lane holds its next ray already loaded: a refill starts the
// prefetched ray from registers and issues the loads of the one after it,
// which arrive while the wavefront keeps searching, so a refill never waits
// on memory.  Rays are claimed per wavefront from the chunks of Claims (XCD
// bands as in k_traverse).  Same per-ray arithmetic as k_traverse.
template<int BLOCK, int SS, bool COUNT, bool PRIMARY, uint32_t TYPES>
__global__ __launch_bounds__(BLOCK) void k_traverse_dyn(const JobDev* __restrict__ jobs, uint32_t n_jobs, Queue q,
                                                        uint32_t n, const QueueView* __restrict__ view, uint4* hits,
                                                        uint32_t* work, uint32_t chunk, uint32_t bands,
                                                        uint32_t refill, uint2* gstack, uint32_t gstride,
                                                        unsigned long long* counters)
    {
    if (!take_view(view, VIEW_BOUNCE, q, n))
        return;
    __shared__ uint2 lds[SS * BLOCK];
    const uint32_t gtid = blockIdx.x * BLOCK + threadIdx.x;
    LdsStack<SS> st;
    st.lds = lds + threadIdx.x;
    st.lstride = BLOCK;
    st.glob = gstack + gtid;
    st.gstride = gstride;
    st.sp = 0;
    const uint32_t lane = lane_id();
    const uint64_t below = (1ull << lane) - 1ull;
    uint32_t n_nodes = 0, n_prims = 0, err = 0, n_rays = 0;
    Claims cl;
    cl.work = work;
    cl.chunk = chunk;
    cl.n = n;
    cl.ranges = (bands && !PRIMARY) ? 8u : 1u;
    cl.r = (cl.ranges == 8u) ? (blockIdx.x & 7u) : 0u;
    cl.tried = 0;
    uint32_t pool, pool_end;
    cl.next(q, lane, pool, pool_end);
    // the lane's prefetched ray: index, origin + tnear, direction + job
    uint32_t pf = EMPTY;
    float4 pa = make_float4(0, 0, 0, 0), pb = make_float4(0, 0, 1, 0);
    // Give every lane without a prefetched ray the next unclaimed ray, in
    // lane order, and issue its loads (not waited for here).
    auto prefetch = [&]()
    {
        for (;;)
            {
            const uint64_t need = __ballot(pf == EMPTY);
            if (need == 0 || pool == EMPTY)
                return;
            const uint32_t avail = pool_end - pool;
            const uint32_t rank = (uint32_t)__popcll(need & below);
            if (pf == EMPTY && rank < avail)
                {
                pf = pool + rank;
                if (PRIMARY)
                    {
                    pa = make_float4(0, 0, 0, 1e-3f);
                    pb = make_float4(0, 0, 1, 0);
                    }
                else
                    {
                    const uint32_t sl = q_slot(q, pf, pool, pool_end - 1u);
                    pa = q.o[sl];
                    pb = q.d[sl];
                    }
                }
            const uint32_t took = min((uint32_t)__popcll(need), avail);
            pool += took;
            if (pool < pool_end)
                return;
            cl.next(q, lane, pool, pool_end);
            }
    };
    prefetch();
    bool active = false;
    uint32_t i = 0;
    TraceState r;
    const Node* nodes2 = nullptr;
    const Node4* nodes4 = nullptr;
    const GeomDev* geoms = nullptr;
    for (;;)
        {
        const uint32_t n_idle = (uint32_t)__popcll(__ballot(!active));
        const uint64_t ready = __ballot(!active && pf != EMPTY);
        if (ready != 0 && (n_idle >= refill || n_idle == 64))
            {
            if (!active && pf != EMPTY)
                {
                i = pf;
                pf = EMPTY;
                v3 o, d;
                float tnear;
                uint32_t job;
                if (PRIMARY)
                    {
                    uint32_t k, pixel, jord;
                    primary_ray(jobs, n_jobs, i, job, k, pixel, jord, o, d);
                    tnear = 1e-3f;
                    }
                else
                    {
                    o = mk3(pa.x, pa.y, pa.z);
                    d = mk3(pb.x, pb.y, pb.z);
                    tnear = pa.w;
                    job = __float_as_uint(pb.w);
                    }
                const SceneDev& sc = jobs[job].scene;
                nodes2 = sc.nodes;
                nodes4 = sc.nodes4;
                geoms = sc.geoms;
                trace_begin(sc, o, d, tnear, r, st);
                active = true;
                }
            prefetch();
            }
        if (__ballot(active) == 0)
            break;
        bool more = false;
        if (active)
            {
            if constexpr (TYPES == T_SPHERES)
                more = trace_visit4<LdsStack<SS>, COUNT, TYPES>(nodes4, geoms, r, st, n_nodes, n_prims, err);
            else
                more = trace_visit2<LdsStack<SS>, COUNT, TYPES>(nodes2, geoms, r, st, n_nodes, n_prims, err);
            }
        if (active && !more)
            {
            hits[i] = make_uint4(__float_as_uint(r.best.t), r.best.geom, r.best.prim, 0);
            active = false;
            if (COUNT)
                n_rays++;
            }
        }
    if (COUNT)
        {
        const int c = PRIMARY ? CTR_PRIMARY : CTR_SECONDARY;
        atomicAdd(counters + c + 0, (unsigned long long)n_rays);
        atomicAdd(counters + c + 1, (unsigned long long)n_nodes);
        atomicAdd(counters + c + 2, (unsigned long long)n_prims);
        }
    if (err)
        atomicOr((unsigned int*)(counters + CTR_OVERFLOW), 1u);
    }

// Tail: each thread finishes its path (trace + shade until termination).
template<int SS, bool COUNT, uint32_t TYPES>
__device__ __forceinline__ void finish_path(const JobDev& J, const uint4& mt, v3 o, v3 d, v3 att, LdsStack<SS>& st,
                                            uint4* records, uint32_t max_depth, unsigned long long* counters,
                                            uint32_t& n_nodes, uint32_t& n_prims, uint32_t& n_rays, uint32_t& err)
    {
    const uint32_t k = mt.z & 0xffffu, ls = mt.z >> 16;
    uint32_t depth = mt.w;
    uint4 rec;
    for (;;)
        {
        HitT h;
        if constexpr (TYPES == T_SPHERES)
            trace_closest<LdsStack<SS>, COUNT, TYPES>(J.scene, o, d, 1e-3f, h, st, n_nodes, n_prims, err);
        else
            trace_closest_grouped<LdsStack<SS>, COUNT, TYPES>(J.scene, o, d, 1e-3f, h, st, n_nodes, n_prims, err);
        if (COUNT)
            n_rays++;
        v3 no, nd;
        if (!shade_bounce<TYPES>(J, h, o, d, 1e-3f, mt.y, k, ls, depth, att, no, nd, rec))
            break;
        if (depth + 1 >= max_depth)
            {
            rec = make_rec(mk3(0, 0, 0), REC_NONE);
            atomicAdd(counters + CTR_CAPPED, 1ull);
            break;
            }
        o = no;
        d = nd;
        depth++;
        }
    records[rec_index(J, k, raster_ordinal(J, mt.y), ls)] = rec;
    }

// A wavefront whose paths all belong to one job reads the job record through
// scalar loads, as in k_shade.  Grid-stride: each thread finishes paths
// i, i + G, ... (G = grid threads; gstride must be G).
template<int BLOCK, int SS, bool COUNT, uint32_t TYPES>
__global__ __launch_bounds__(BLOCK) void k_finish(const JobDev* __restrict__ jobs, uint32_t n_jobs, Queue in,
                                                  uint32_t n, const QueueView* __restrict__ view, uint4* records,
                                                  uint32_t max_depth, uint2* gstack, uint32_t gstride,
                                                  unsigned long long* counters)
    {
    if (!take_view(view, VIEW_FINISH, in, n))
        return;
    __shared__ uint2 lds[SS * BLOCK];
    LdsStack<SS> st;
    st.lds = lds + threadIdx.x;
    st.lstride = BLOCK;
    st.glob = gstack + blockIdx.x * BLOCK + threadIdx.x;
    st.gstride = gstride;
    st.sp = 0;
    uint32_t n_nodes = 0, n_prims = 0, n_rays = 0, err = 0;
    for (uint32_t first = blockIdx.x * BLOCK; first < n; first += gridDim.x * BLOCK)
        {
        const uint32_t i = first + threadIdx.x;
        if (i >= n)
            break;
        const uint32_t sl = q_slot(in, i, first, min(n, first + BLOCK) - 1u);
        const uint4 mt = in.meta[sl];
        const float4 a = in.o[sl];
        const float4 b = in.d[sl];
        const float4 at4 = in.att[sl];
        const v3 o = mk3(a.x, a.y, a.z), d = mk3(b.x, b.y, b.z), att = mk3(at4.x, at4.y, at4.z);
        const uint32_t j0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)mt.x);
        if (n_jobs == 1 || __ballot(mt.x != j0) == 0)
            finish_path<SS, COUNT, TYPES>(jobs[j0], mt, o, d, att, st, records, max_depth, counters, n_nodes,
                                          n_prims, n_rays, err);
        else
            finish_path<SS, COUNT, TYPES>(jobs[mt.x], mt, o, d, att, st, records, max_depth, counters, n_nodes,
                                          n_prims, n_rays, err);
        }
    if (COUNT)
        {
        atomicAdd(counters + CTR_TAIL + 0, (unsigned long long)n_rays);
        atomicAdd(counters + CTR_TAIL + 1, (unsigned long long)n_nodes);
        atomicAdd(counters + CTR_TAIL + 2, (unsigned long long)n_prims);
        }
    if (err)
        atomicOr((unsigned int*)(counters + CTR_OVERFLOW), 1u);
    }

// ---------------------------------------------------------------- accumulate
// Replay each camera sample's records in light-sample order, then the Welford
// mean and sRGB of fresnel/cpu/TracerPath.cc:229-249.
__global__ void k_accumulate(const JobDev* __restrict__ jobs, const uint32_t* __restrict__ pix_base, uint32_t n_jobs,
                             uint32_t total, const uint4* __restrict__ records)
    {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total)
        return;
    uint32_t lo = 0, hi = n_jobs;
    while (hi - lo > 1)
        {
        const uint32_t mid = (lo + hi) >> 1;
        if (pix_base[mid] <= i)
            lo = mid;
        else
            hi = mid;
        }
    const JobDev& J = jobs[lo];
    // thread i handles pixel ordinal j (records are chunk-major, so a
    // wavefront reads 64 contiguous records per sample)
    const uint32_t jord = i - pix_base[lo];
    uint32_t px, py;
    ordinal_pixel(J, jord, px, py);
    const uint32_t pixel = py * J.W + px;
    float4 mean = J.linear[pixel];
    const uint32_t K = J.K;
    // the first records of four camera samples are loaded together, then
    // folded into the mean in sample order
    constexpr uint32_t U = 4;
    for (uint32_t k0 = 0; k0 < K; k0 += U)
    {
    uint4 rk[U];
#pragma unroll
    for (uint32_t u = 0; u < U; u++)
        rk[u] = (k0 + u < K) ? records[rec_index(J, k0 + u, jord, 0)] : make_uint4(0, 0, 0, 0);
#pragma unroll
    for (uint32_t u = 0; u < U; u++)
        {
        const uint32_t k = k0 + u;
        if (k < K)
        {
        v3 result = mk3(0, 0, 0);
        float a = 1.0f;
        const uint4 r0 = rk[u];
        if ((r0.w & 3u) == REC_BG)
            {
            result = J.bg * float(J.L);
            a = J.bg_alpha;
            }
        else
            {
            auto fold = [&](const uint4 r)
            {
                const v3 val = mk3(__uint_as_float(r.x), __uint_as_float(r.y), __uint_as_float(r.z));
                const uint32_t kind = r.w & 3u;
                if (kind == REC_ADD)
                    result = result + val;
                else if (kind == REC_LIGHTS)
                    {
                    const uint32_t mask = r.w >> 2;
                    for (uint32_t li = 0; li < J.lights.N; li++)
                        if (mask & (1u << li))
                            result = result + (val * J.lights.color[li]) * J.lights.inv_sin[li];
                    }
            };
            fold(r0);
            for (uint32_t ls = 1; ls < J.L; ls++)
                fold(records[rec_index(J, k, jord, ls)]);
            }
        const v3 s = div_rcp(result, float(J.L));
        const float q = 1.0f / float(J.n_base + k + 1);
        mean.x = mean.x + (s.x - mean.x) * q;
        mean.y = mean.y + (s.y - mean.y) * q;
        mean.z = mean.z + (s.z - mean.z) * q;
        mean.w = mean.w + (a - mean.w) * q;
        }
        }
    }
    J.linear[pixel] = mean;
    if (!J.highlight || (mean.x <= 1.0f && mean.y <= 1.0f && mean.z <= 1.0f))
        J.srgb[pixel] = to_srgb(mean);
    else
        J.srgb[pixel] = to_srgb(make_float4(J.hw_color.x, J.hw_color.y, J.hw_color.z, mean.w));
    }

// ---------------------------------------------------------------- preview
// fresnel/cpu/TracerDirect.cc:77-218: aa_n^2 primary rays with tnear 0 and
// representative-point direct lighting, plain mean.
__global__ void k_preview(const JobDev* __restrict__ jobp, uint32_t aa_n)
    {
    const JobDev& J = *jobp;
    const uint32_t pixel = blockIdx.x * blockDim.x + threadIdx.x;
    if (pixel >= J.npix)
        return;
    const uint32_t py = pixel / J.W, px = pixel - py * J.W;
    LocalStack st;
    uint32_t dummy = 0, err = 0;
    float4 avg = make_float4(0, 0, 0, 0);
    for (uint32_t sample = 0; sample < aa_n * aa_n; sample++)
        {
        v3 o, d;
        camera_ray(J.cam, J.W, J.H, J.xs_off, J.inv_h, J.seed, px, py, sample, aa_n != 1, o, d);
        HitT h;
        trace_closest<LocalStack, false>(J.scene, o, d, 0.0f, h, st, dummy, dummy, err);
        v3 c = J.bg;
        float a = J.bg_alpha;
        if (h.geom != EMPTY)
            {
            Attr at;
            const GeomDev* g;
            hit_attributes(J.scene, h, o, d, 0.0f, at, g);
            const v3 n = div_true(at.Ng, sqrtf(dot(at.Ng, at.Ng)));
            const v3 v = div_rcp(-d, sqrtf(dot(d, d)));
            const Mat& m = (at.d >= g->outline_width) ? g->mat : g->outline;
            if (m.solid > 0.5f)
                c = get_color(m, at.color);
            else
                {
                c = mk3(0, 0, 0);
                for (uint32_t li = 0; li < J.lights.N; li++)
                    {
                    const v3 l = J.lights.dir[li];
                    v3 r = -v + (2.0f * n * dot(n, v));
                    const float half = J.lights.theta[li];
                    const float ldotr = dot(l, r);
                    if (ldotr < J.lights.cos_theta[li])
                        {
                        v3 ax = cross(l, r);
                        ax = div_rcp(ax, sqrtf(dot(ax, ax)));
                        const float ang = -acosf(ldotr) + half;
                        quat qr;
                        qr.s = cosf(ang / 2.0f);
                        qr.v = sinf(ang / 2.0f) * ax;
                        r = rotate(qr, r);
                        }
                    const float ndotl = dot(n, l);
                    const v3 f_d = (ndotl >= 0.0f) ? brdf_diffuse(m, l, v, n, at.color) * ndotl : mk3(0, 0, 0);
                    const v3 f_s = (dot(n, r) >= 0.0f)
                                       ? brdf_specular(m, r, v, n, at.color, half, J.lights.tan_clamped[li]) * dot(n, r)
                                       : mk3(0, 0, 0);
                    c = c + (f_d + f_s) * PI_F * J.lights.color[li];
                    }
                }
            a = 1.0f;
            }
        avg.x += c.x;
        avg.y += c.y;
        avg.z += c.z;
        avg.w += a;
        }
    const float q = 1.0f / float(aa_n * aa_n);
    const float4 outp = make_float4(avg.x * q, avg.y * q, avg.z * q, avg.w * q);
    J.linear[pixel] = outp;
    if (!J.highlight || (outp.x <= 1.0f && outp.y <= 1.0f && outp.z <= 1.0f))
        J.srgb[pixel] = to_srgb(outp);
    else
        J.srgb[pixel] = to_srgb(make_float4(J.hw_color.x, J.hw_color.y, J.hw_color.z, outp.w));
    }

// ---------------------------------------------------------------- diagnostics
// Plain streaming copy, the reference point that scales the HBM byte counters.
__global__ void k_debug_copy(const float4* __restrict__ src, float4* __restrict__ dst, uint64_t n)
    {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
    }

__global__ void k_debug_philox(const uint32_t* in, uint32_t n, uint32_t* out)
    {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t* c = in + 6 * i;
    const u4 r = philox(c[0], c[1], c[2], c[3], c[4], c[5]);
    for (int q = 0; q < 4; q++)
        out[4 * i + q] = r.v[q];
    }

// The shading code's sinf/cosf (cr_sincosf) of n arguments.
__global__ void k_debug_sincos(const float* x, uint32_t n, float* s, float* c)
    {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    float a, b;
    cr_sincosf(x[i], a, b);
    s[i] = a;
    c[i] = b;
    }

// Primary ray and closest hit per pixel for camera sample `sample`.
__global__ void k_debug_primary(const JobDev* __restrict__ jobp, uint32_t sample, float* rays, uint4* hits)
    {
    const JobDev& J = *jobp;
    const uint32_t pixel = blockIdx.x * blockDim.x + threadIdx.x;
    if (pixel >= J.npix)
        return;
    const uint32_t py = pixel / J.W, px = pixel - py * J.W;
    v3 o, d;
    camera_ray(J.cam, J.W, J.H, J.xs_off, J.inv_h, J.seed, px, py, sample, true, o, d);
    LocalStack st;
    uint32_t dummy = 0, err = 0;
    HitT h;
    trace_closest<LocalStack, false>(J.scene, o, d, 1e-3f, h, st, dummy, dummy, err);
    float* r = rays + 6 * (size_t)pixel;
    r[0] = o.x;
    r[1] = o.y;
    r[2] = o.z;
    r[3] = d.x;
    r[4] = d.y;
    r[5] = d.z;
    hits[pixel] = make_uint4(__float_as_uint(h.t), h.geom, h.prim, err);
    }

__global__ void k_debug_trace(const SceneDev sc, const float* rays, uint32_t n, float tnear, uint4* hits)
    {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const float* r = rays + 6 * (size_t)i;
    LocalStack st;
    uint32_t dummy = 0, err = 0;
    HitT h;
    trace_closest<LocalStack, false>(sc, mk3(r[0], r[1], r[2]), mk3(r[3], r[4], r[5]), tnear, h, st, dummy, dummy,
                                     err);
    hits[i] = make_uint4(__float_as_uint(h.t), h.geom, h.prim, err);
    }
    } // namespace fh
