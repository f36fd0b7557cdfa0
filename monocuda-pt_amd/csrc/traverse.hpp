// traverse.hpp — stack traversal of the scene's two trees.
//
// Two walks return the same closest hit (t and primitive, bit for bit):
//   step_closest()  one iteration of BVH::traverseRay's loop
//                   (src/path_tracer.cu:1102-1147, split as stepTraversal,
//                   :1276-1321) on the binary tree: both child boxes against
//                   the running closest distance, leaf children tested at
//                   once (left, then right), left child first, the right one
//                   pushed when both are internal and hit. Its step and
//                   leaf-test counts are the reference's.
//   step_wide()     the four-wide quantized tree, nearest inner child first,
//                   with the reference's tie rule (a per-leaf rank in its
//                   leaf-test order) so exact-t ties resolve as the
//                   reference's order would resolve them. The default.
// step_wide_any() is the any-hit walk (shadow rays) on the four-wide tree.
//
// The traversal stack lives in LDS, one column per lane
// (stack[level * stride + lane_in_block]), so lanes at any depth hit
// distinct banks. Its height (stack_entries) is the binary tree's measured
// depth (<= 64: Karras prefixes grow strictly along a root-to-leaf path and
// stay below 64), rounded up to even for the wide walk's two-int groups.
#pragma once

#include "device.hpp"

namespace mcpt {

struct Trav {
    vec3 o, d, inv;
    float t;   // running closest distance (or the any-hit limit)
    float time;  // the ray's time (where moving spheres are)
    int leaf;  // closest leaf so far, -1 = none
    int node;
    int sp;
    uint32_t pend;  // deferred walk: slots of pend_ch still to test as leaves
    int4 pend_ch;   // the child references of the node they belong to
};

// EXACT keeps the IEEE reciprocals that the reference's slab test uses
// (step_closest). The fast orders clamp them to +-1e30: with an infinite
// reciprocal (a zero direction component) the fused box test would form
// inf - inf = NaN on that axis for every box, drop the axis, and walk most
// of the tree. Clamping only moves the slab ends of axes along which the
// ray advances less than 1e-26 across any scene, and every triangle sits
// 1e-3 inside its boxes, so no box that holds a reachable hit is culled.
template <bool EXACT = false>
MCPT_DEV void trav_begin(Trav& s, const vec3& o, const vec3& d, float tmax) {
    s.o = o;
    s.d = d;
    s.inv = vec3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    if (!EXACT)
        s.inv = vec3(fminf(fmaxf(s.inv.x, -1e30f), 1e30f), fminf(fmaxf(s.inv.y, -1e30f), 1e30f),
                     fminf(fmaxf(s.inv.z, -1e30f), 1e30f));
    s.t = tmax;
    s.time = 0.0f;
    s.leaf = -1;
    s.node = 0;
    s.sp = 0;
    s.pend = 0;
}

// original-order hot triangle (9 floats), used when shading
MCPT_DEV void load_tri(const float* base, int i, vec3& v0, vec3& v1, vec3& v2) {
    const float* t = base + (size_t)i * 9;
    v0 = vec3(t[0], t[1], t[2]);
    v1 = vec3(t[3], t[4], t[5]);
    v2 = vec3(t[6], t[7], t[8]);
}

// The ray against leaf i's primitive: hit_triangle with the stored normal,
// or hit_sphere at the ray's time (the primitive-type switch of
// testPrimitiveClosest / doesPrimitiveBlockRay, :1537-1563).
MCPT_DEV bool leaf_hit(const LeafTri* base, int i, const Trav& s, float& t) {
    const LeafTri* p = base + i;
    const float4 a = p->a, b = p->b, c = p->c;
    if (leaf_is_sphere(c)) {
        const vec3 c0(a.x, a.y, a.z);
        return hit_sphere(sphere_center(c0, vec3(b.x, b.y, b.z), b.w != 0.0f, s.time), a.w, s.o, s.d, t);
    }
    return hit_triangle_n(vec3(a.x, a.y, a.z), vec3(a.w, b.x, b.y), vec3(b.z, b.w, c.x), vec3(c.y, c.z, c.w), s.o,
                          s.d, t);
}

// testPrimitiveClosest (src/path_tracer.cu:1537-1551)
MCPT_DEV void test_leaf_closest(const BvhView& bv, Trav& s, int leaf) {
    float t;
    if (leaf_hit(bv.leaf_tri, leaf, s, t) && t < s.t && t > 0.0001f) {
        s.t = t;
        s.leaf = leaf;
    }
}

// doesPrimitiveBlockRay (src/path_tracer.cu:1553-1563)
MCPT_DEV bool test_leaf_any(const BvhView& bv, const Trav& s, int leaf) {
    float t;
    return leaf_hit(bv.leaf_tri, leaf, s, t) && t > 0.0001f && t < s.t;
}

// One loop iteration; returns true when the traversal is finished. The up
// to two leaf children are tested by one loop (left first, as in the
// reference), so a wave runs the triangle test code once per step rather
// than once per side.
template <bool COUNT>
MCPT_DEV bool step_closest(const BvhView& bv, Trav& s, int* stk, int stride, uint32_t& ns, uint32_t& nl) {
    const BvhNode* np = bv.nodes + s.node;
    const float4 a = np->a, b = np->b, c = np->c;
    const int4 r = np->r;
    const bool oL = slab(a.x, a.y, a.z, a.w, b.x, b.y, s.o, s.inv, s.t);
    const bool oR = slab(b.z, b.w, c.x, c.y, c.z, c.w, s.o, s.inv, s.t);
    const bool lL = r.x < 0, rL = r.y < 0;
    if (COUNT) ns++;
    int first = -1, second = -1;
    if (oL && lL) first = ~r.x;
    if (oR && rL) {
        if (first < 0) first = ~r.y;
        else second = ~r.y;
    }
    while (first >= 0) {
        if (COUNT) nl++;
        test_leaf_closest(bv, s, first);
        first = second;
        second = -1;
    }
    const bool tL = oL && !lL, tR = oR && !rL;
    if (!tL && !tR) {
        if (s.sp == 0) return true;
        s.sp--;
        s.node = stk[s.sp * stride];
    } else {
        if (tL && tR) {
            stk[s.sp * stride] = r.y;
            s.sp++;
        }
        s.node = tL ? r.x : r.y;
    }
    return false;
}

// ---------------------------------------------------- four-wide walk
// The fast orders walk the four-wide quantized tree (WideNode). Child box k
// of node w spans w.o + q * 2^e per axis; its slab distances are formed as
// fma(q, 2^e / d, (w.o - o) / d), i.e. (b - o) / d with four roundings, each
// relative to |w.o - o| or to the node's extent, as in the reference's
// (b - o) * (1/d). A slab end therefore moves by at most about
// 4 * 2^-24 * (|w.o - o| + extent) along its axis, below the 1e-3 every
// triangle sits inside its boxes while the node lies within ~4000 units of
// the ray's origin (the same bound under which the reference's own slab test
// cannot cull a reachable hit), whatever the absolute coordinates; and the
// quantized boxes contain the exact ones. So no box that holds a reachable
// hit is culled and the exact triangle test alone decides t and the
// primitive.
//
// Ties: the reference tests a leaf when it visits the leaf's parent, parents
// in left-first pre-order, left side first, and keeps the first of equal t.
// Every leaf that reaches the minimum t is tested in any order (its boxes are
// entered before that t), so keeping the leaf of smallest rank in that order
// (tie[], computed by the build) returns the reference's (t, primitive).

// Bit k set when child k's box is entered before s.t; tn[k] its entry.
// The slab that a ray enters first on an axis is the lower bound when the
// direction component is positive and the upper one when it is negative
// (fma(q, a, b) is monotone in q with the sign of a = 2^e / d), so each
// axis's near and far byte words are chosen once per node instead of a
// min and a max per child and axis; the distances are the same values.
MCPT_DEV uint32_t wide_boxes(const float4& o, const uint4& q0, const uint4& q1, const Trav& s, float tn[4]) {
    const uint32_t meta = __float_as_uint(o.w);
    const int cnt = (int)(meta >> 24);
    const float ax = __uint_as_float((meta & 255u) << 23) * s.inv.x;
    const float ay = __uint_as_float(((meta >> 8) & 255u) << 23) * s.inv.y;
    const float az = __uint_as_float(((meta >> 16) & 255u) << 23) * s.inv.z;
    const float bx = (o.x - s.o.x) * s.inv.x;
    const float by = (o.y - s.o.y) * s.inv.y;
    const float bz = (o.z - s.o.z) * s.inv.z;
    const bool rx = s.inv.x < 0.0f, ry = s.inv.y < 0.0f, rz = s.inv.z < 0.0f;
    const uint32_t nx = rx ? q0.w : q0.x, fx = rx ? q0.x : q0.w;
    const uint32_t ny = ry ? q1.x : q0.y, fy = ry ? q0.y : q1.x;
    const uint32_t nz = rz ? q1.y : q0.z, fz = rz ? q0.z : q1.y;
    uint32_t hit = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int sh = 8 * k;
        const float tnx = __builtin_fmaf((float)((nx >> sh) & 255u), ax, bx);
        const float tfx = __builtin_fmaf((float)((fx >> sh) & 255u), ax, bx);
        const float tny = __builtin_fmaf((float)((ny >> sh) & 255u), ay, by);
        const float tfy = __builtin_fmaf((float)((fy >> sh) & 255u), ay, by);
        const float tnz = __builtin_fmaf((float)((nz >> sh) & 255u), az, bz);
        const float tfz = __builtin_fmaf((float)((fz >> sh) & 255u), az, bz);
        const float tnear = fmaxf(fmaxf(0.0001f, tnx), fmaxf(tny, tnz));
        const float tfar = fminf(fminf(s.t, tfx), fminf(tfy, tfz));
        tn[k] = tnear;
        hit |= (uint32_t)(k < cnt && !(tfar <= tnear)) << k;
    }
    return hit;
}

// the reference in child slot j (0..3), as selects rather than branches
MCPT_DEV int pick4(const int c[4], int j) {
    const int lo = (j & 1) ? c[1] : c[0];
    const int hi = (j & 1) ? c[3] : c[2];
    return (j & 2) ? hi : lo;
}

// bit j set when child slot j holds a leaf (the sign bits of the references)
MCPT_DEV uint32_t leaf_slots(const int4& ch) {
    return ((uint32_t)ch.x >> 31) | (((uint32_t)ch.y >> 31) << 1) | (((uint32_t)ch.z >> 31) << 2) |
           (((uint32_t)ch.w >> 31) << 3);
}

MCPT_DEV void test_leaf_ranked(const BvhView& bv, Trav& s, int leaf) {
    float t;
    if (leaf_hit(bv.leaf_tri, leaf, s, t) && t > 0.0001f) {
        if (t < s.t || (t == s.t && s.leaf >= 0 && bv.tie[leaf] < bv.tie[s.leaf])) {
            s.t = t;
            s.leaf = leaf;
        }
    }
}


// The wide walk's stack holds one group per wide level: two ints, the wide
// index of the node's first inner child and the ranks (among its inner
// children, in slot order) of the children still to visit, in visit order
// (bits 0-5, two bits each) with their count (bits 8-9). A node's inner
// children have consecutive wide indices (the build numbers wide nodes in
// level, then binary index order), so a pop reads no memory. Height: two
// ints per wide level, i.e. the binary depth rounded up to even.
MCPT_DEV void push_group(int* stk, int stride, Trav& s, int base, uint32_t seq) {
    stk[(2 * s.sp) * stride] = base;
    stk[(2 * s.sp + 1) * stride] = (int)seq;
    s.sp++;
}

// rank of slot k among the inner slots of `inner` (bit j set: slot j inner)
MCPT_DEV uint32_t inner_rank(uint32_t inner, int k) { return (uint32_t)__builtin_popcount(inner & ((1u << k) - 1u)); }

// every inner slot of a node, hit or not (the ranks count them all): the
// used slots (count in the top byte of the origin word) that hold no leaf
MCPT_DEV uint32_t node_inner(const float4& o, uint32_t leaf_mask) {
    const uint32_t cnt = __float_as_uint(o.w) >> 24;
    return ~leaf_mask & ((1u << cnt) - 1u);
}

// Moves to the next child of the top group; false when the stack is empty.
MCPT_DEV bool pop_group(int* stk, int stride, Trav& s) {
    if (s.sp == 0) return false;
    const int top = s.sp - 1;
    const int base = stk[(2 * top) * stride];
    const uint32_t seq = (uint32_t)stk[(2 * top + 1) * stride];
    const uint32_t cnt = seq >> 8;
    s.node = base + (int)(seq & 3u);
    if (cnt > 1) stk[(2 * top + 1) * stride] = (int)(((seq & 0x3Fu) >> 2) | ((cnt - 1) << 8));
    else s.sp = top;
    return true;
}

MCPT_DEV void cswap(float& da, int& ka, float& db, int& kb) {
    if (db < da) {
        const float d = da;
        da = db;
        db = d;
        const int k = ka;
        ka = kb;
        kb = k;
    }
}

// One closest-hit step on a wide node: the hit leaf children are tested,
// then the nearest hit inner child is visited next and the others wait on
// the stack as one group, nearest first.
template <bool COUNT>
MCPT_DEV bool step_wide(const BvhView& bv, Trav& s, int* stk, int stride, uint32_t& ns, uint32_t& nl) {
    const WideNode* wp = bv.wide + s.node;
    const float4 o = wp->o;
    const uint4 q0 = wp->q0, q1 = wp->q1;
    const int4 ch = wp->child;
    if (COUNT) ns++;
    float tn[4];
    const uint32_t hit = wide_boxes(o, q0, q1, s, tn);
    const int c[4] = {ch.x, ch.y, ch.z, ch.w};
    const uint32_t lm = leaf_slots(ch);
    uint32_t leaves = hit & lm;
    const uint32_t inner = hit & ~lm;
    float d[4];
    int k[4] = {0, 1, 2, 3};
#pragma unroll
    for (int j = 0; j < 4; j++) d[j] = ((inner >> j) & 1u) ? tn[j] : __builtin_inff();
    while (leaves) {
        const int j = __builtin_ctz(leaves);
        leaves &= leaves - 1u;
        if (COUNT) nl++;
        test_leaf_ranked(bv, s, ~pick4(c, j));
    }
    if (!inner) return !pop_group(stk, stride, s);
    // order the inner children by entry distance (non-inner slots sort last)
    cswap(d[0], k[0], d[1], k[1]);
    cswap(d[2], k[2], d[3], k[3]);
    cswap(d[0], k[0], d[2], k[2]);
    cswap(d[1], k[1], d[3], k[3]);
    cswap(d[1], k[1], d[2], k[2]);
    const uint32_t n = __builtin_popcount(inner);
    if (n > 1) {
        const uint32_t all = node_inner(o, lm);
        push_group(stk, stride, s, pick4(c, __builtin_ctz(all)),
                   inner_rank(all, k[1]) | inner_rank(all, k[2]) << 2 | inner_rank(all, k[3]) << 4 | (n - 1) << 8);
    }
    s.node = pick4(c, k[0]);
    return false;
}

// The same step with the hit leaf children deferred: they are recorded in
// s.pend / s.pend_ch for test_pending, which the caller runs for many lanes
// at once (a wave otherwise runs the triangle test whenever any one lane
// reaches a leaf). The walk continues meanwhile with the running distance it
// has; the closest hit does not depend on the order of the tests. Returns
// true when no node is left to visit (pending leaves may remain).
template <bool COUNT>
MCPT_DEV bool step_wide_defer(const BvhView& bv, Trav& s, int* stk, int stride, uint32_t& ns) {
    const WideNode* wp = bv.wide + s.node;
    const float4 o = wp->o;
    const uint4 q0 = wp->q0, q1 = wp->q1;
    const int4 ch = wp->child;
    if (COUNT) ns++;
    float tn[4];
    const uint32_t hit = wide_boxes(o, q0, q1, s, tn);
    const int c[4] = {ch.x, ch.y, ch.z, ch.w};
    const uint32_t lm = leaf_slots(ch);
    uint32_t leaves = hit & lm;
    const uint32_t inner = hit & ~lm;
    float d[4];
    int k[4] = {0, 1, 2, 3};
#pragma unroll
    for (int j = 0; j < 4; j++) d[j] = ((inner >> j) & 1u) ? tn[j] : __builtin_inff();
    s.pend = leaves;
    s.pend_ch = ch;
    if (!inner) return !pop_group(stk, stride, s);
    cswap(d[0], k[0], d[1], k[1]);
    cswap(d[2], k[2], d[3], k[3]);
    cswap(d[0], k[0], d[2], k[2]);
    cswap(d[1], k[1], d[3], k[3]);
    cswap(d[1], k[1], d[2], k[2]);
    const uint32_t n = __builtin_popcount(inner);
    if (n > 1) {
        const uint32_t all = node_inner(o, lm);
        push_group(stk, stride, s, pick4(c, __builtin_ctz(all)),
                   inner_rank(all, k[1]) | inner_rank(all, k[2]) << 2 | inner_rank(all, k[3]) << 4 | (n - 1) << 8);
    }
    s.node = pick4(c, k[0]);
    return false;
}

// the leaves step_wide_defer left pending
template <bool COUNT>
MCPT_DEV void test_pending(const BvhView& bv, Trav& s, uint32_t& nl) {
    const int c[4] = {s.pend_ch.x, s.pend_ch.y, s.pend_ch.z, s.pend_ch.w};
    while (s.pend) {
        const int j = __builtin_ctz(s.pend);
        s.pend &= s.pend - 1u;
        if (COUNT) nl++;
        test_leaf_ranked(bv, s, ~pick4(c, j));
    }
}

// Any-hit step on a wide node, children in left-to-right order; returns 0
// while walking, 1 when the ray is unblocked, 2 when blocked.
MCPT_DEV int step_wide_any(const BvhView& bv, Trav& s, int* stk, int stride) {
    const WideNode* wp = bv.wide + s.node;
    const float4 o = wp->o;
    const uint4 q0 = wp->q0, q1 = wp->q1;
    const int4 ch = wp->child;
    float tn[4];
    const uint32_t hit = wide_boxes(o, q0, q1, s, tn);
    const int c[4] = {ch.x, ch.y, ch.z, ch.w};
    const uint32_t lm = leaf_slots(ch);
    uint32_t leaves = hit & lm;
    const uint32_t inner = hit & ~lm;
    while (leaves) {
        const int j = __builtin_ctz(leaves);
        leaves &= leaves - 1u;
        if (test_leaf_any(bv, s, ~pick4(c, j))) return 2;
    }
    if (!inner) return pop_group(stk, stride, s) ? 0 : 1;
    // the leftmost inner child next, the others as one group, left to right
    const int first = __builtin_ctz(inner);
    uint32_t rest = inner & (inner - 1u);
    const uint32_t n = __builtin_popcount(rest);
    if (n) {
        const uint32_t all = node_inner(o, lm);
        uint32_t seq = 0;
        for (int i = 0; i < 3 && rest; i++) {
            seq |= inner_rank(all, __builtin_ctz(rest)) << (2 * i);
            rest &= rest - 1u;
        }
        push_group(stk, stride, s, pick4(c, __builtin_ctz(all)), seq | n << 8);
    }
    s.node = pick4(c, first);
    return 0;
}

// Full closest-hit traversal. REF selects the reference's visit order on the
// binary tree (for step and leaf-test counts that match the CPU oracle);
// otherwise the four-wide tree, near child first. Both return the same t and
// leaf.
template <bool COUNT, bool REF = false>
MCPT_DEV void trace_closest(const BvhView& bv, Trav& s, int* stk, int stride, uint32_t& ns, uint32_t& nl) {
    if (REF) {
        while (!step_closest<COUNT>(bv, s, stk, stride, ns, nl)) {
        }
    } else {
        while (!step_wide<COUNT>(bv, s, stk, stride, ns, nl)) {
        }
    }
}

// extendKernelPersistent (src/path_tracer.cu:3416-3468): lanes keep their
// traversal across iterations; a wave with at least `refill` idle lanes takes
// that many ray indices from *work with one atomic and hands them out by
// ballot rank. begin(q, s, bv) starts ray q's walk (and may point the lane's
// view `bv` at the ray's own scene); finish(q, s, bv) takes the finished
// walk. In the fast order leaf tests are deferred and run for all lanes that
// have some pending once kLeafBatch lanes do (or no lane has a node left to
// visit); REF steps the binary tree in the reference's order.
constexpr int kLeafBatch = 16;
template <bool COUNT, bool REF, class View, class Begin, class Finish>
MCPT_DEV void persistent_closest(View& bv, int* work, int total, int refill, int* stk, int stride, Begin begin,
                                 Finish finish, uint32_t& ns, uint32_t& nl) {
    const int lane = lane_id();
    Trav s;
    s.pend = 0;
    bool has = false, walked = false;  // walked: no node left, only pending leaves
    int slot = -1;
    bool exhausted = false;  // uniform across the wave
    while (true) {
        if (!exhausted) {
            const uint64_t idle = ballot64(!has);
            const int nidle = __popcll(idle);
            if (nidle >= refill) {
                const int leader = __builtin_ctzll(idle);
                int first = 0;
                if (lane == leader) first = atomicAdd(work, nidle);
                first = __shfl(first, leader);
                if (!has) {
                    const int q = first + rank_below(idle);
                    if (q < total) {
                        begin(q, s, bv);
                        slot = q;
                        has = true;
                    }
                }
                if (first + nidle >= total) exhausted = true;
            }
        }
        if (exhausted && ballot64(has) == 0) break;
        if (REF) {
            if (has && step_closest<COUNT>(bv, s, stk, stride, ns, nl)) {
                finish(slot, s, bv);
                has = false;
            }
            continue;
        }
        // node steps for lanes with nothing pending; leaf tests once enough
        // lanes have leaves pending, or when no lane has a node to visit
        if (has && !s.pend && !walked) walked = step_wide_defer<COUNT>(bv, s, stk, stride, ns);
        const bool want = has && s.pend != 0;
        const uint64_t wm = ballot64(want);
        if (wm && (__popcll(wm) >= kLeafBatch || ballot64(has && !s.pend && !walked) == 0)) {
            if (want) test_pending<COUNT>(bv, s, nl);
        }
        if (has && walked && !s.pend) {
            finish(slot, s, bv);
            has = false;
            walked = false;
        }
    }
}

// BVH::traverseAnyHit (src/path_tracer.cu:1152-1207) on the four-wide tree:
// boxes are culled with the fixed limit and the walk stops at the first
// blocker, children in left-to-right order (measured faster than near-first
// for shadow rays, which mostly start on a surface inside overlapping
// boxes). The answer does not depend on the order.
MCPT_DEV bool trace_any(const BvhView& bv, Trav& s, int* stk, int stride) {
    int st;
    while ((st = step_wide_any(bv, s, stk, stride)) == 0) {
    }
    return st == 2;
}

// ------------------------------------------- group walks for short queues
// A launch lasts as long as its slowest ray, and one lane walking one ray
// visits that ray's nodes one dependent fetch after another: the late,
// thin bounces of a render take ~0.6 ms each whatever their ray count. For
// a short queue kGroup lanes walk each ray together. The group shares one
// stack in LDS of (wide node, entry distance) pairs and pops up to kGroup of
// them per iteration, one per lane; it tests the leaves that the popped
// nodes hit with the tests spread over its lanes, takes the minimum
// (t, tie rank) over its lanes, and pushes the inner children still entered
// before that distance, every node's nearest child last, so that the next
// pops take the nearest child of each node just expanded.
//
// Same hits: every box that holds a hit at t is entered before t (the 1e-3
// padding), so an entry is culled only when its box holds nothing at or
// before the group's best distance; every leaf that reaches the final t is
// tested, and the tie rank picks among equal t as in step_wide. The order of
// the tests changes nothing.
//
// Stack bound: an iteration pops at most max(1, (cap - sp) / 3) entries and
// each pushes at most four, so an iteration that pops two or more leaves at
// most cap entries; a run of single pops is a depth-first walk on top of
// those, which keeps at most three pending siblings per wide level. The
// stack never holds more than cap + 3 * levels + 1 pairs
// (group_stack_pairs on the host).
constexpr int kGroup = 16;

// LDS ints per group: its stack of cap_phys (node, entry distance) pairs and
// the leaf list of one iteration (at most four per popped node)
__host__ __device__ constexpr int group_lds_ints(int cap_phys) { return 2 * cap_phys + 4 * kGroup; }

// The soft cap of a group's stack (every lane of a group pops while the
// stack holds up to 48 pairs), and the stack's bound for a tree whose
// single-lane walk takes stack_ints ints (two per wide level): the cap,
// three pending siblings per wide level, one more
constexpr int kGroupCap = 96;
__host__ __device__ constexpr int group_stack_pairs(int stack_ints) { return kGroupCap + 3 * (stack_ints / 2 + 1) + 4; }

MCPT_DEV int group_pops(int sp, int cap) {
    int n = (cap - sp) / 3;
    n = n < 1 ? 1 : (n > kGroup ? kGroup : n);
    return n < sp ? n : sp;
}

// LDS written by one lane of a wave and read by another: the wave's LDS
// operations complete in order, and the fences keep the compiler from
// moving them across this point
MCPT_DEV void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct GroupLanes {
    int gl;         // lane within the group
    int g0;         // the group's first lane in the wave
    uint64_t gm;    // the group's lanes
    uint64_t below; // the wave's lanes below this one
};

MCPT_DEV GroupLanes group_lanes() {
    GroupLanes g;
    const int lane = lane_id();
    g.gl = lane & (kGroup - 1);
    g.g0 = lane & ~(kGroup - 1);
    g.gm = ((1ull << kGroup) - 1ull) << g.g0;
    g.below = (1ull << lane) - 1ull;
    return g;
}

// The group's (t, leaf) minimum in every lane: smaller t, and among equal t
// the leaf of smaller tie rank (a total order on the hits, so the butterfly
// leaves the same pair in every lane).
MCPT_DEV void group_min_hit(const BvhView& bv, Trav& s) {
#pragma unroll
    for (int m = 1; m < kGroup; m <<= 1) {
        const float ot = __shfl_xor(s.t, m);
        const int ol = __shfl_xor(s.leaf, m);
        bool take = ot < s.t;
        if (!take && ot == s.t && ol >= 0 && s.leaf >= 0 && ol != s.leaf) take = bv.tie[ol] < bv.tie[s.leaf];
        if (take) {
            s.t = ot;
            s.leaf = ol;
        }
    }
}

// The leaves in `leaves` (child slots of c) of every lane of the group into
// lf[], in lane order; returns how many the group holds.
MCPT_DEV int group_leaf_list(const GroupLanes& g, uint32_t leaves, const int c[4], int* lf) {
    const int nl = __builtin_popcount(leaves);
    int at = 0, total = 0;
#pragma unroll
    for (int b = 0; b < 4; b++) {
        const uint64_t m = ballot64(nl > b) & g.gm;
        at += __popcll(m & g.below);
        total += __popcll(m);
    }
    while (leaves) {
        const int j = __builtin_ctz(leaves);
        leaves &= leaves - 1u;
        lf[at++] = ~pick4(c, j);
    }
    return total;
}

// Pops up to group_pops entries, one per lane, and expands the popped node
// of this lane (its child boxes against s.t): the hit leaf slots in *leaves,
// the hit inner children in k[]/d[] by entry distance (n of them; the other
// d[] are infinite).
MCPT_DEV void group_expand(const BvhView& bv, const GroupLanes& g, const Trav& s, const int2* stk, int& sp, int cap,
                           bool cull, int c[4], uint32_t& leaves, float d[4], int k[4]) {
    const int npop = group_pops(sp, cap);
    const int top = sp;
    sp -= npop;
    bool valid = false;
    int node = 0;
    if (g.gl < npop) {
        const int2 e = stk[top - 1 - g.gl];
        node = e.x;
        valid = !cull || __int_as_float(e.y) < s.t;
    }
    leaves = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        c[j] = 0;
        d[j] = __builtin_inff();
        k[j] = j;
    }
    if (valid) {
        const WideNode* wp = bv.wide + node;
        const float4 o = wp->o;
        const uint4 q0 = wp->q0, q1 = wp->q1;
        const int4 ch = wp->child;
        float tn[4];
        const uint32_t hit = wide_boxes(o, q0, q1, s, tn);
        c[0] = ch.x;
        c[1] = ch.y;
        c[2] = ch.z;
        c[3] = ch.w;
        const uint32_t lm = leaf_slots(ch);
        leaves = hit & lm;
        const uint32_t inner = hit & ~lm;
#pragma unroll
        for (int j = 0; j < 4; j++) d[j] = ((inner >> j) & 1u) ? tn[j] : __builtin_inff();
        cswap(d[0], k[0], d[1], k[1]);
        cswap(d[2], k[2], d[3], k[3]);
        cswap(d[0], k[0], d[2], k[2]);
        cswap(d[1], k[1], d[3], k[3]);
        cswap(d[1], k[1], d[2], k[2]);
    }
}

// Pushes each lane's first m children of k[]/d[], the group's nearest ones
// last (rank-major: every lane's fourth, then third, second, nearest).
MCPT_DEV void group_push(const GroupLanes& g, int2* stk, int& sp, int cap_phys, int m, const int c[4], const float d[4],
                         const int k[4]) {
#pragma unroll
    for (int r = 3; r >= 0; r--) {
        const uint64_t b = ballot64(m > r) & g.gm;
        if (m > r) {
            const int pos = sp + __popcll(b & g.below);
            if (pos < cap_phys) stk[pos] = make_int2(pick4(c, k[r]), __float_as_int(d[r]));
        }
        sp += __popcll(b);
    }
    sp = sp < cap_phys ? sp : cap_phys;  // never reached (see the bound above)
}

// Closest hits of rays [0, total) by groups of kGroup lanes, each group taking
// its next ray from *work. begin(q, s, bv) starts ray q in every lane of the
// group; finish(q, s, bv) runs in the group's first lane with the hit.
//
// The leaves an iteration's nodes hit are tested in the next iteration, their
// records loaded beside that iteration's nodes, so an iteration waits for one
// round of fetches rather than a node fetch and then a leaf fetch; the boxes
// of that iteration are culled with the best distance of the one before,
// which only ever lets more through. lf holds two leaf lists, used in turn.
template <class View, class Begin, class Finish>
MCPT_DEV void group_closest(View& bv, int* work, int total, int2* stk, int* lf, int cap, int cap_phys, Begin begin,
                            Finish finish) {
    const GroupLanes g = group_lanes();
    Trav s;
    int ray = -1, sp = 0;
    while (true) {
        if (ray == -1) {
            int q = 0;
            if (g.gl == 0) q = atomicAdd(work, 1);
            q = __shfl(q, g.g0);
            if (q < total) {
                begin(q, s, bv);
                ray = q;
                sp = 1;
                if (g.gl == 0) stk[0] = make_int2(0, __float_as_int(0.0f));
                wave_lds_sync();
            } else {
                ray = -2;
            }
        }
        if (ballot64(ray >= 0) == 0) break;
        if (ray < 0) continue;
        int c[4], k[4];
        float d[4];
        uint32_t leaves;
        group_expand(bv, g, s, stk, sp, cap, true, c, leaves, d, k);
        wave_lds_sync();  // pops read before pushes overwrite them
        const int nleaf = group_leaf_list(g, leaves, c, lf);
        wave_lds_sync();
        for (int i = g.gl; i < nleaf; i += kGroup) test_leaf_ranked(bv, s, lf[i]);
        group_min_hit(bv, s);
        int m = 0;
#pragma unroll
        for (int j = 0; j < 4; j++) m += d[j] < s.t ? 1 : 0;
        wave_lds_sync();  // leaf list read before the next iteration writes it
        group_push(g, stk, sp, cap_phys, m, c, d, k);
        wave_lds_sync();
        if (sp == 0) {
            if (g.gl == 0) finish(ray, s, bv);
            ray = -1;
        }
    }
}

// Any-hit test of rays [0, total) by groups of kGroup lanes (s.t is the
// ray's fixed limit); finish(q, blocked) runs in the group's first lane.
// Leaves are tested one iteration late as in group_closest.
template <class View, class Begin, class Finish>
MCPT_DEV void group_any(View& bv, int* work, int total, int2* stk, int* lf, int cap, int cap_phys, Begin begin,
                        Finish finish) {
    const GroupLanes g = group_lanes();
    Trav s;
    int ray = -1, sp = 0;
    while (true) {
        if (ray == -1) {
            int q = 0;
            if (g.gl == 0) q = atomicAdd(work, 1);
            q = __shfl(q, g.g0);
            if (q < total) {
                begin(q, s, bv);
                ray = q;
                sp = 1;
                if (g.gl == 0) stk[0] = make_int2(0, __float_as_int(0.0f));
                wave_lds_sync();
            } else {
                ray = -2;
            }
        }
        if (ballot64(ray >= 0) == 0) break;
        if (ray < 0) continue;
        int c[4], k[4];
        float d[4];
        uint32_t leaves;
        group_expand(bv, g, s, stk, sp, cap, false, c, leaves, d, k);
        wave_lds_sync();
        const int nleaf = group_leaf_list(g, leaves, c, lf);
        wave_lds_sync();
        bool blk = false;
        for (int i = g.gl; i < nleaf && !blk; i += kGroup) blk = test_leaf_any(bv, s, lf[i]);
        const bool blocked = (ballot64(blk) & g.gm) != 0;
        int m = 0;
#pragma unroll
        for (int j = 0; j < 4; j++) m += d[j] < __builtin_inff() ? 1 : 0;
        wave_lds_sync();
        if (!blocked) group_push(g, stk, sp, cap_phys, m, c, d, k);
        wave_lds_sync();
        if (blocked || sp == 0) {
            if (g.gl == 0) finish(ray, blocked);
            ray = -1;
        }
    }
}

}  // namespace mcpt
