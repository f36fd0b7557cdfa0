// traverse.hpp — stack traversal of the two-box BVH layout.
//
// One call of step_closest() is one iteration of BVH::traverseRay's loop
// (src/path_tracer.cu:1102-1147, split the same way as stepTraversal,
// :1276-1321): test both child boxes against the running closest distance,
// test leaf children immediately (left, then right), then descend left
// first and push the right child when both children are internal and hit.
// The visit order is therefore the reference's, and so are the closest-hit
// distance, the winning primitive on exact ties, and the step and leaf-test
// counts.
//
// The traversal stack lives in LDS, one column per lane
// (stack[level * stride + lane_in_block]), so lanes at any depth hit
// distinct banks. Its height is the tree's measured depth (<= 64: Karras
// prefixes grow strictly along a root-to-leaf path and stay below 64).
#pragma once

#include "device.hpp"

namespace mcpt {

struct Trav {
    vec3 o, d, inv;
    float t;   // running closest distance (or the any-hit limit)
    int leaf;  // closest leaf so far, -1 = none
    int node;
    int sp;
};

MCPT_DEV void trav_begin(Trav& s, const vec3& o, const vec3& d, float tmax) {
    s.o = o;
    s.d = d;
    s.inv = vec3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    s.t = tmax;
    s.leaf = -1;
    s.node = 0;
    s.sp = 0;
}

// original-order hot triangle (9 floats), used when shading
MCPT_DEV void load_tri(const float* base, int i, vec3& v0, vec3& v1, vec3& v2) {
    const float* t = base + (size_t)i * 9;
    v0 = vec3(t[0], t[1], t[2]);
    v1 = vec3(t[3], t[4], t[5]);
    v2 = vec3(t[6], t[7], t[8]);
}

MCPT_DEV void load_leaf(const LeafTri* base, int i, vec3& v0, vec3& v1, vec3& v2, vec3& n) {
    const LeafTri* p = base + i;
    const float4 a = p->a, b = p->b, c = p->c;
    v0 = vec3(a.x, a.y, a.z);
    v1 = vec3(a.w, b.x, b.y);
    v2 = vec3(b.z, b.w, c.x);
    n = vec3(c.y, c.z, c.w);
}

// testPrimitiveClosest (src/path_tracer.cu:1537-1551)
MCPT_DEV void test_leaf_closest(const SceneView& sv, Trav& s, int leaf) {
    vec3 v0, v1, v2, n;
    load_leaf(sv.leaf_tri, leaf, v0, v1, v2, n);
    float t;
    if (hit_triangle_n(v0, v1, v2, n, s.o, s.d, t) && t < s.t && t > 0.0001f) {
        s.t = t;
        s.leaf = leaf;
    }
}

// doesPrimitiveBlockRay (src/path_tracer.cu:1553-1563)
MCPT_DEV bool test_leaf_any(const SceneView& sv, const Trav& s, int leaf) {
    vec3 v0, v1, v2, n;
    load_leaf(sv.leaf_tri, leaf, v0, v1, v2, n);
    float t;
    return hit_triangle_n(v0, v1, v2, n, s.o, s.d, t) && t > 0.0001f && t < s.t;
}

// One loop iteration; returns true when the traversal is finished. The up
// to two leaf children are tested by one loop (left first, as in the
// reference), so a wave runs the triangle test code once per step rather
// than once per side.
template <bool COUNT>
MCPT_DEV bool step_closest(const SceneView& sv, Trav& s, int* stk, int stride, uint32_t& ns, uint32_t& nl) {
    const BvhNode* np = sv.nodes + s.node;
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
        test_leaf_closest(sv, s, first);
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

template <bool COUNT>
MCPT_DEV void trace_closest(const SceneView& sv, Trav& s, int* stk, int stride, uint32_t& ns, uint32_t& nl) {
    while (!step_closest<COUNT>(sv, s, stk, stride, ns, nl)) {
    }
}

// BVH::traverseAnyHit (src/path_tracer.cu:1152-1207): the boxes are culled
// with the fixed limit and the walk stops at the first blocker.
MCPT_DEV bool trace_any(const SceneView& sv, Trav& s, int* stk, int stride) {
    while (true) {
        const BvhNode* np = sv.nodes + s.node;
        const float4 a = np->a, b = np->b, c = np->c;
        const int4 r = np->r;
        const bool oL = slab(a.x, a.y, a.z, a.w, b.x, b.y, s.o, s.inv, s.t);
        const bool oR = slab(b.z, b.w, c.x, c.y, c.z, c.w, s.o, s.inv, s.t);
        const bool lL = r.x < 0, rL = r.y < 0;
        int first = -1, second = -1;
        if (oL && lL) first = ~r.x;
        if (oR && rL) {
            if (first < 0) first = ~r.y;
            else second = ~r.y;
        }
        bool blocked = false;
        while (first >= 0 && !blocked) {
            blocked = test_leaf_any(sv, s, first);
            first = second;
            second = -1;
        }
        if (blocked) return true;
        const bool tL = oL && !lL, tR = oR && !rL;
        if (!tL && !tR) {
            if (s.sp == 0) return false;
            s.sp--;
            s.node = stk[s.sp * stride];
        } else {
            if (tL && tR) {
                stk[s.sp * stride] = r.y;
                s.sp++;
            }
            s.node = tL ? r.x : r.y;
        }
    }
}

}  // namespace mcpt
