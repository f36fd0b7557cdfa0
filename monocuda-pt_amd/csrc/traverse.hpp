// traverse.hpp — stack traversal of the two-box BVH layout.
//
// Two visit orders return the same closest hit (t and primitive, bit for
// bit):
//   step_closest()  one iteration of BVH::traverseRay's loop
//                   (src/path_tracer.cu:1102-1147, split as stepTraversal,
//                   :1276-1321): both child boxes against the running
//                   closest distance, leaf children tested at once (left,
//                   then right), left child first, the right one pushed when
//                   both are internal and hit. Its step and leaf-test counts
//                   are the reference's.
//   step_ordered()  the nearer child first, with the reference's tie rule
//                   (ref_earlier) so exact-t ties resolve as the reference's
//                   order would resolve them. The default.
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
    vec3 oi;   // o * inv, for the fused box test of the fast orders
    float t;   // running closest distance (or the any-hit limit)
    int leaf;  // closest leaf so far, -1 = none
    int best;  // near-first order: the closest leaf's parent << 1 | side
    int node;
    int sp;
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
    s.oi = vec3(o.x * s.inv.x, o.y * s.inv.y, o.z * s.inv.z);
    s.t = tmax;
    s.leaf = -1;
    s.best = -1;
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
MCPT_DEV void test_leaf_closest(const BvhView& bv, Trav& s, int leaf) {
    vec3 v0, v1, v2, n;
    load_leaf(bv.leaf_tri, leaf, v0, v1, v2, n);
    float t;
    if (hit_triangle_n(v0, v1, v2, n, s.o, s.d, t) && t < s.t && t > 0.0001f) {
        s.t = t;
        s.leaf = leaf;
    }
}

// doesPrimitiveBlockRay (src/path_tracer.cu:1553-1563)
MCPT_DEV bool test_leaf_any(const BvhView& bv, const Trav& s, int leaf) {
    vec3 v0, v1, v2, n;
    load_leaf(bv.leaf_tri, leaf, v0, v1, v2, n);
    float t;
    return hit_triangle_n(v0, v1, v2, n, s.o, s.d, t) && t > 0.0001f && t < s.t;
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

// ------------------------------------------------------ near-first order
// The reference's closest hit does not depend on the visit order except on
// exact ties of t between different triangles: every triangle lies at least
// 1e-3 inside its box (the padding of calculate_bounding_box, :651-655), so a
// box culled against the running distance never holds a triangle at or
// before it, and the minimum t is found in any order. On a tie the
// reference keeps the triangle it tests first. It tests the leaf children of
// an internal node when it visits that node (left, then right) and visits
// internal nodes in left-first pre-order, so of two candidate leaves under
// parents P and Q the reference tests P's first when P == Q and P's side is
// left, when P is an ancestor of Q (P's range contains Q's), or when P's
// range lies left of Q's. ref_earlier() applies that rule, which lets the
// traversal below descend into the nearer child first and still return the
// reference's (t, primitive) bit for bit.
MCPT_DEV bool ref_earlier(const BvhView& bv, int p, int p_other_end, int p_side, int q_code) {
    const int q = q_code >> 1, q_side = q_code & 1;
    if (p == q) return p_side < q_side;
    const int fp = min(p, p_other_end), lp = max(p, p_other_end);
    const int q_other_end = bv.nodes[q].r.w;
    const int fq = min(q, q_other_end), lq = max(q, q_other_end);
    if (fp <= fq && lq <= lp) return true;
    if (fq <= fp && lp <= lq) return false;
    return lp < fq;
}

MCPT_DEV void test_leaf_ordered(const BvhView& bv, Trav& s, int leaf, int parent, int parent_other_end, int side) {
    vec3 v0, v1, v2, n;
    load_leaf(bv.leaf_tri, leaf, v0, v1, v2, n);
    float t;
    if (hit_triangle_n(v0, v1, v2, n, s.o, s.d, t) && t > 0.0001f) {
        if (t < s.t || (t == s.t && s.leaf >= 0 && ref_earlier(bv, parent, parent_other_end, side, s.best))) {
            s.t = t;
            s.leaf = leaf;
            s.best = (parent << 1) | side;
        }
    }
}

// One iteration: both boxes, leaf children, then the nearer internal child
// (the left one on equal entry distances) with the other one pushed.
template <bool COUNT>
MCPT_DEV bool step_ordered(const BvhView& bv, Trav& s, int* stk, int stride, uint32_t& ns, uint32_t& nl) {
    const BvhNode* np = bv.nodes + s.node;
    const float4 a = np->a, b = np->b, c = np->c;
    const int4 r = np->r;
    float nL, nR;
    const bool oL = slab_fma(a.x, a.y, a.z, a.w, b.x, b.y, s.oi, s.inv, s.t, nL);
    const bool oR = slab_fma(b.z, b.w, c.x, c.y, c.z, c.w, s.oi, s.inv, s.t, nR);
    const bool lL = r.x < 0, rL = r.y < 0;
    if (COUNT) ns++;
    int first = -1, second = -1;
    if (oL && lL) first = 0;
    if (oR && rL) {
        if (first < 0) first = 1;
        else second = 1;
    }
    while (first >= 0) {
        if (COUNT) nl++;
        test_leaf_ordered(bv, s, first ? ~r.y : ~r.x, s.node, r.w, first);
        first = second;
        second = -1;
    }
    const bool tL = oL && !lL, tR = oR && !rL;
    if (!tL && !tR) {
        if (s.sp == 0) return true;
        s.sp--;
        s.node = stk[s.sp * stride];
    } else if (tL && tR) {
        const bool left_first = !(nR < nL);
        stk[s.sp * stride] = left_first ? r.y : r.x;
        s.sp++;
        s.node = left_first ? r.x : r.y;
    } else {
        s.node = tL ? r.x : r.y;
    }
    return false;
}

// Full closest-hit traversal. REF selects the reference's visit order (for
// step and leaf-test counts that match the CPU checker); otherwise near
// child first. Both return the same t and leaf.
template <bool COUNT, bool REF = false>
MCPT_DEV void trace_closest(const BvhView& bv, Trav& s, int* stk, int stride, uint32_t& ns, uint32_t& nl) {
    if (REF) {
        while (!step_closest<COUNT>(bv, s, stk, stride, ns, nl)) {
        }
    } else {
        while (!step_ordered<COUNT>(bv, s, stk, stride, ns, nl)) {
        }
    }
}

// BVH::traverseAnyHit (src/path_tracer.cu:1152-1207): boxes are culled with
// the fixed limit and the walk stops at the first blocker, in the
// reference's left-first order (measured faster than near-first for shadow
// rays, which mostly start on a surface inside overlapping boxes). The
// answer does not depend on the order. One iteration per call: returns 0
// while walking, 1 when the ray is unblocked, 2 when blocked.
MCPT_DEV int step_any(const BvhView& bv, Trav& s, int* stk, int stride) {
    const BvhNode* np = bv.nodes + s.node;
    const float4 a = np->a, b = np->b, c = np->c;
    const int4 r = np->r;
    float nL, nR;
    const bool oL = slab_fma(a.x, a.y, a.z, a.w, b.x, b.y, s.oi, s.inv, s.t, nL);
    const bool oR = slab_fma(b.z, b.w, c.x, c.y, c.z, c.w, s.oi, s.inv, s.t, nR);
    const bool lL = r.x < 0, rL = r.y < 0;
    int first = -1, second = -1;
    if (oL && lL) first = ~r.x;
    if (oR && rL) {
        if (first < 0) first = ~r.y;
        else second = ~r.y;
    }
    bool blocked = false;
    while (first >= 0 && !blocked) {
        blocked = test_leaf_any(bv, s, first);
        first = second;
        second = -1;
    }
    if (blocked) return 2;
    const bool tL = oL && !lL, tR = oR && !rL;
    if (!tL && !tR) {
        if (s.sp == 0) return 1;
        s.sp--;
        s.node = stk[s.sp * stride];
    } else {
        if (tL && tR) {
            stk[s.sp * stride] = r.y;
            s.sp++;
        }
        s.node = tL ? r.x : r.y;
    }
    return 0;
}

MCPT_DEV bool trace_any(const BvhView& bv, Trav& s, int* stk, int stride) {
    int st;
    while ((st = step_any(bv, s, stk, stride)) == 0) {
    }
    return st == 2;
}

}  // namespace mcpt
