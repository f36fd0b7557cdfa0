// batch.hip — many scenes at once: their LBVH builds as one launch per stage,
// their renders as one wavefront whose queues hold the rays of all of them,
// and an autotuner for that wavefront's shape. The reference builds and
// renders one scene per process run; nothing here has a counterpart in it
// except the per-scene stages, which are lbvh.hip's and render.hip's steps
// applied per segment.
//
// Build: the triangles of scene s are the global indices [off_s, off_s+n_s).
// Bounds are reduced per scene (chunks, then one block per scene), Morton codes use their own
// scene's bounds, and the stable LSD radix sort takes four 8-bit passes over
// the codes and a fifth over the scene index: the result is ordered by
// (scene, code, index), i.e. each scene's segment is exactly the (code,
// index) order of its own build. Leaves, the Karras tree and the node levels
// then run over all segments in one launch each, and the refit in one launch
// per level over the nodes of that level in every scene, writing each
// scene's own node, leaf and box buffers, so each BVH is bit-identical to
// mcpt_lbvh_build of that scene alone.
//
// Render: path = k * P + q, where q runs over the pixels of all scenes
// (scene s owns q in [foff_s, foff_s + W_s H_s)) and k over the samples of
// the pass. Rays carry their scene index; the persistent extend loads that
// scene's BVH view per ray, shade and shadow its scene view. Finalize adds a
// pixel's samples in sample order, so every film is bit-identical to
// mcpt_render of its scene.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "lbvh_common.hpp"
#include "scene.hpp"
#include "shade.hpp"

namespace mcpt {
namespace {

struct Seg {  // one scene of a batched build
    int off, n, ioff, pad;  // first triangle, count, first internal node
    const float* hot;
    BvhNode* nodes;
    LeafTri* leaf_tri;
    int* leaf_prim;
    float* root_box;
    int* tie;
    WideNode* wide;
};

// scene owning global index g (segments ordered by off)
__device__ __forceinline__ int seg_of(const Seg* seg, int nseg, int g) {
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (seg[mid].off <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}
// scene owning internal node g (offsets ioff)
__device__ __forceinline__ int seg_of_node(const Seg* seg, int nseg, int g) {
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (seg[mid].ioff <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// scene bounds in two levels (min/max are exact in any order): chunks of
// kChunk triangles of one scene per block, then one block per scene over its
// chunks. chunk_off[s] is the first chunk of scene s.
constexpr int kChunk = 8192;

__device__ void reduce_box(float acc[6], float* out) {
    __shared__ float sh[6][kBlock];
    for (int k = 0; k < 6; k++) sh[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            for (int k = 0; k < 3; k++) sh[k][threadIdx.x] = fminf(sh[k][threadIdx.x], sh[k][threadIdx.x + s]);
            for (int k = 3; k < 6; k++) sh[k][threadIdx.x] = fmaxf(sh[k][threadIdx.x], sh[k][threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) out[threadIdx.x] = sh[threadIdx.x][0];
}

__global__ __launch_bounds__(kBlock) void k_b_bounds_chunks(const Seg* seg, int nseg, const int* chunk_off,
                                                           float* partial) {
    const int c = blockIdx.x;
    int lo = 0, hi = nseg - 1;  // scene of chunk c
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (chunk_off[mid] <= c) lo = mid;
        else hi = mid - 1;
    }
    const Seg S = seg[lo];
    const int b0 = (c - chunk_off[lo]) * kChunk, b1 = min(S.n, b0 + kChunk);
    float acc[6] = {__builtin_inff(), __builtin_inff(), __builtin_inff(), -__builtin_inff(), -__builtin_inff(),
                    -__builtin_inff()};
    for (int i = b0 + threadIdx.x; i < b1; i += kBlock) {
        float t[9], b[6];
        for (int k = 0; k < 9; k++) t[k] = S.hot[(size_t)i * 9 + k];
        tri_box(t, b);
        for (int k = 0; k < 3; k++) acc[k] = fminf(acc[k], b[k]);
        for (int k = 3; k < 6; k++) acc[k] = fmaxf(acc[k], b[k]);
    }
    reduce_box(acc, partial + 6 * c);
}

__global__ __launch_bounds__(kBlock) void k_b_bounds_final(const int* chunk_off, const float* partial, float* sb) {
    const int s = blockIdx.x;
    float acc[6] = {__builtin_inff(), __builtin_inff(), __builtin_inff(), -__builtin_inff(), -__builtin_inff(),
                    -__builtin_inff()};
    for (int c = chunk_off[s] + threadIdx.x; c < chunk_off[s + 1]; c += kBlock) {
        for (int k = 0; k < 3; k++) acc[k] = fminf(acc[k], partial[6 * c + k]);
        for (int k = 3; k < 6; k++) acc[k] = fmaxf(acc[k], partial[6 * c + k]);
    }
    reduce_box(acc, sb + 6 * s);
}

// morton3D/quantize (:913-935) against the triangle's own scene bounds
__global__ __launch_bounds__(kBlock) void k_b_morton(const Seg* seg, int nseg, int total, const float* sbs,
                                                    uint32_t* morton, uint32_t* keys, int* vals) {
    const int g = blockIdx.x * kBlock + threadIdx.x;
    if (g >= total) return;
    const int s = seg_of(seg, nseg, g);
    const float* t = seg[s].hot + (size_t)(g - seg[s].off) * 9;
    const float* sb = sbs + 6 * s;
    const vec3 c = (vec3(t[0], t[1], t[2]) + vec3(t[3], t[4], t[5]) + vec3(t[6], t[7], t[8])) / 3.0f;
    const uint32_t qx = (uint32_t)((c.x - sb[0]) / (sb[3] - sb[0]) * 1023.0f);
    const uint32_t qy = (uint32_t)((c.y - sb[1]) / (sb[4] - sb[1]) * 1023.0f);
    const uint32_t qz = (uint32_t)((c.z - sb[2]) / (sb[5] - sb[2]) * 1023.0f);
    const uint32_t m = expand_bits(qx) | expand_bits(qy) << 1 | expand_bits(qz) << 2;
    morton[g] = m;
    keys[g] = m;
    vals[g] = g;
}

// scene index of each element in its current order, the key of the last pass
__global__ __launch_bounds__(kBlock) void k_b_scene_keys(const Seg* seg, int nseg, int total, const int* vals,
                                                        uint32_t* keys) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < total) keys[i] = (uint32_t)seg_of(seg, nseg, vals[i]);
}

// sorted codes by gathering through the sorted ids
__global__ __launch_bounds__(kBlock) void k_b_gather_keys(const uint32_t* morton, const int* vals, int total,
                                                         uint32_t* keys) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < total) keys[i] = morton[vals[i]];
}

// leaf records per scene, in its sorted order (k_leaves of lbvh.hip)
__global__ __launch_bounds__(kBlock) void k_b_leaves(const Seg* seg, int nseg, int total, const int* ids) {
    const int j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= total) return;
    const int s = seg_of(seg, nseg, j);
    const Seg S = seg[s];
    const int jl = j - S.off, p = ids[j] - S.off;
    S.leaf_prim[jl] = p;
    const float* t = S.hot + (size_t)p * 9;
    const vec3 v0(t[0], t[1], t[2]), v1(t[3], t[4], t[5]), v2(t[6], t[7], t[8]);
    const vec3 nn = triangle_normal(v0, v1, v2);
    LeafTri L;
    L.a = make_float4(v0.x, v0.y, v0.z, v1.x);
    L.b = make_float4(v1.y, v1.z, v2.x, v2.y);
    L.c = make_float4(v2.z, nn.x, nn.y, nn.z);
    S.leaf_tri[jl] = L;
}

// Karras tree of each scene with >= 2 triangles (buildInternalNodesKernel,
// :1417-1447, determineRange/findSplit :1349-1386); leaf_parent is indexed
// by global leaf position
__global__ __launch_bounds__(kBlock) void k_b_karras(const Seg* seg, int nseg, int total_internal,
                                                    const uint32_t* skeys, int* leaf_parent) {
    const int g = blockIdx.x * kBlock + threadIdx.x;
    if (g >= total_internal) return;
    const Seg S = seg[seg_of_node(seg, nseg, g)];
    const int idx = g - S.ioff, n = S.n;
    const uint32_t* keys = skeys + S.off;
    const int d = (delta(keys, n, idx, idx + 1) - delta(keys, n, idx, idx - 1)) >= 0 ? 1 : -1;
    const int dmin = delta(keys, n, idx, idx - d);
    int lmax = 2;
    while (delta(keys, n, idx, idx + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (delta(keys, n, idx, idx + (l + t) * d) > dmin) l += t;
    const int j = idx + l * d;
    const int first = min(idx, j), last = max(idx, j);
    const int cp = delta(keys, n, first, last);
    int split = first, step = last - first;
    do {
        step = (step + 1) >> 1;
        const int ns = split + step;
        if (ns < last && delta(keys, n, first, ns) > cp) split = ns;
    } while (step > 1);
    int* rf = reinterpret_cast<int*>(S.nodes + idx) + 12;
    const int kids[2] = {split, split + 1};
    const bool leafs[2] = {split == first, split + 1 == last};
    for (int side = 0; side < 2; side++) {
        const int c = kids[side];
        if (leafs[side]) {
            rf[side] = ~c;
            leaf_parent[S.off + c] = (idx << 1) | side;
            float b[6];
            const LeafTri L = S.leaf_tri[c];
            const float tt[9] = {L.a.x, L.a.y, L.a.z, L.a.w, L.b.x, L.b.y, L.b.z, L.b.w, L.c.x};
            tri_box(tt, b);
            store_slot(S.nodes, idx, side, b);
        } else {
            rf[side] = c;
            reinterpret_cast<int*>(S.nodes + c)[14] = (idx << 1) | side;
        }
    }
    if (idx == 0) rf[2] = -1;
    rf[3] = j;
}

// Levels of every scene's internal nodes (keys and ids by global node index),
// each leaf's rank in the reference's leaf-test order, each scene's deepest
// leaf and its number of even-level (wide) nodes: a wave reduces its lanes'
// values per scene (nearly always one) before a single atomic each.
__global__ __launch_bounds__(kBlock) void k_b_levels(const Seg* seg, int nseg, int total, const int* leaf_parent,
                                                    uint32_t* level, int* ids, int* depth, int* wcount) {
    const int g = blockIdx.x * kBlock + threadIdx.x;
    const bool valid = g < total;
    int s = -1, d = 0, ne = 0;
    if (valid) {
        s = seg_of(seg, nseg, g);
        const Seg& S = seg[s];
        const int j = g - S.off;
        if (j < S.n - 1) ids[S.ioff + j] = S.ioff + j;
        const LeafClimb c = leaf_levels(S.nodes, leaf_parent[g], level + S.ioff);
        S.tie[j] = c.rank;
        d = c.depth;
        ne = c.n_even;
    }
    uint64_t todo = ballot64(valid);
    while (todo) {
        const int leader = __builtin_ctzll(todo);
        const int s0 = __shfl(s, leader);
        const bool mine = valid && s == s0;
        int v = mine ? d : 0, w = mine ? ne : 0;
        for (int off = 32; off > 0; off >>= 1) {
            v = max(v, __shfl_xor(v, off));
            w += __shfl_xor(w, off);
        }
        if (lane_id() == leader) {
            atomicMax(&depth[s0], v);
            if (w) atomicAdd(&wcount[s0], w);
        }
        todo &= ~ballot64(mine);
    }
}

// scene (segment) of each node of ids[]
__global__ __launch_bounds__(kBlock) void k_b_node_scene(const Seg* seg, int nseg, const int* ids, int n,
                                                        uint32_t* key) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) key[i] = (uint32_t)seg_of_node(seg, nseg, ids[i]);
}

// one wide node per even-level binary node of every scene (global index q)
__global__ __launch_bounds__(kBlock) void k_b_collapse(const Seg* seg, int nseg, const uint32_t* level,
                                                      const int* widx, int n) {
    const int q = blockIdx.x * kBlock + threadIdx.x;
    if (q >= n || (level[q] & 1u)) return;
    const Seg& S = seg[seg_of_node(seg, nseg, q)];
    collapse_node(S.nodes, q - S.ioff, widx, S.ioff, S.wide);
}

// refitBoundsKernel (:1398-1415) for the nodes of one level, all scenes
__global__ __launch_bounds__(kBlock) void k_b_refit_level(const Seg* seg, int nseg, const int* ids, int cnt) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= cnt) return;
    const int q = ids[i];
    const Seg& S = seg[seg_of_node(seg, nseg, q)];
    refit_node(S.nodes, q - S.ioff, S.root_box);
}

__global__ void k_b_fill(int* p, int n, int v) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = v;
}

int blocks_for(int64_t n, int b = kBlock) { return (int)((n + b - 1) / b); }

// --------------------------------------------------------- batched render
struct BScene {  // one scene of a batched render
    SceneView sv;
    mcpt_camera cam;
    float* accum;
    int foff;  // first pixel index of the scene across all scenes
    int npix;
};

enum { C_NS0 = 0, C_NS1 = 32, C_WORK = 64, C_EXT = 96, C_SHD = 128, C_WORK_SHADOW = 160, C_NCOUNTERS = 192 };

__device__ __forceinline__ unsigned long long* ns_slot(int* c, int q) {
    return reinterpret_cast<unsigned long long*>(c + (q ? C_NS1 : C_NS0));
}
__device__ __forceinline__ int queue_len(const int* c, int q) { return c[q ? C_NS1 : C_NS0]; }
__device__ __forceinline__ int shadow_len(const int* c, int q) { return c[(q ? C_NS1 : C_NS0) + 1]; }

__device__ __forceinline__ int scene_of_pixel(const BScene* tab, int nsc, int q) {
    int lo = 0, hi = nsc - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tab[mid].foff <= q) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

struct BWs {
    RayItem* q[2];
    mcpt_packed_hit* hits;
    ShadowItem* sq;
    float4* thr;
    float4* smp;
    uint64_t* rng;
    int* counters;
    int tb;  // 24: ray records carry scene and time (see pack_scene_time)
};

__global__ void k_bb_init(int* c, int count) {
    if (threadIdx.x == 0) {
        *ns_slot(c, 0) = (unsigned long long)(uint32_t)count;
        *ns_slot(c, 1) = 0;
        c[C_WORK] = 0;
    }
}
__global__ void k_bb_bounce(int* c, int cur) {
    if (threadIdx.x == 0) {
        *ns_slot(c, 1 - cur) = 0;
        c[C_WORK] = 0;
        *reinterpret_cast<unsigned long long*>(c + C_EXT) += (unsigned long long)queue_len(c, cur);
    }
}
__global__ void k_bb_shadow_total(int* c, int cur) {
    if (threadIdx.x == 0) {
        *reinterpret_cast<unsigned long long*>(c + C_SHD) += (unsigned long long)shadow_len(c, 1 - cur);
        c[C_WORK_SHADOW] = 0;
    }
}

// generateCameraRaysKernel (:3342-3374) for every scene's pixels and the
// pass's samples
__global__ __launch_bounds__(kBlock) void k_bb_camera(const BScene* tab, int nsc, int P, uint32_t sample0, int nb,
                                                     uint64_t seed, BWs ws) {
    const int path = blockIdx.x * kBlock + threadIdx.x;
    if (path >= P * nb) return;
    const int k = path / P, q = path - k * P;
    const int s = scene_of_pixel(tab, nsc, q);
    const mcpt_camera& cam = tab[s].cam;
    const int pixel = q - tab[s].foff;
    Rng rng;
    mcpt_pcg32_for_path(&rng.g, (uint32_t)pixel, sample0 + (uint32_t)k, seed);
    vec3 o, d;
    float time;
    camera_ray(cam, pixel % cam.width, pixel / cam.width, rng, o, d, time);
    RayItem it;
    it.ot = make_float4(o.x, o.y, o.z, pack_scene_time(s, time, ws.tb));
    it.dp = make_float4(d.x, d.y, d.z, __int_as_float(path));
    ws.q[0][path] = it;
    ws.thr[path] = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
    ws.smp[path] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    ws.rng[path] = rng.g.state;
}

// persistent extend over the rays of every scene (extendKernelPersistent,
// :3416-3468) with deferred leaf tests; a lane reloads the BVH view only
// when its scene changes
// the 16-byte hit record; prim_index keeps the leaf's code (triangle, or
// ~sphere)
__device__ __forceinline__ mcpt_packed_hit bb_hit(const Trav& s, const BvhView& v) {
    mcpt_packed_hit h;
    h.t = s.t;
    h.prim_type = s.leaf >= 0 ? 1 : -1;
    h.prim_index = s.leaf >= 0 ? v.leaf_prim[s.leaf] : -1;
    h.did_hit = s.leaf >= 0 ? 1 : 0;
    return h;
}

// A queue shorter than group_below is k_bb_extend_group's.
__global__ __launch_bounds__(kBlock) void k_bb_extend(const BvhView* bvhs, BWs ws, int cur, int refill,
                                                     int group_below) {
    extern __shared__ int lds_stack[];
    const int total = queue_len(ws.counters, cur);
    if (total < group_below) return;
    if ((int)blockIdx.x * kBlock >= total) return;  // blocks the queue does not need
    BvhView bv;
    int bv_scene = -1;
    uint32_t ns = 0, nl = 0;
    const RayItem* q = ws.q[cur];
    persistent_closest<false, false>(
        bv, &ws.counters[C_WORK], total, refill, lds_stack + threadIdx.x, kBlock,
        [&](int i, Trav& s, BvhView& v) {
            const RayItem it = q[i];
            const int sc = unpack_scene(it.ot.w, ws.tb);
            if (sc != bv_scene) {
                v = bvhs[sc];
                bv_scene = sc;
            }
            trav_begin(s, vec3(it.ot.x, it.ot.y, it.ot.z), vec3(it.dp.x, it.dp.y, it.dp.z), 1e30f);
            s.time = unpack_time(it.ot.w, ws.tb);
        },
        [&](int i, const Trav& s, const BvhView& v) { ws.hits[i] = bb_hit(s, v); }, ns, nl);
}

// a short queue: 16 lanes per ray (group_closest), each group's view that
// of its ray's scene
__global__ __launch_bounds__(kBlock) void k_bb_extend_group(const BvhView* bvhs, BWs ws, int cur, int group_below,
                                                           int cap, int cap_phys) {
    extern __shared__ int lds[];
    const int total = queue_len(ws.counters, cur);
    if (total >= group_below || (int)blockIdx.x * (kBlock / kGroup) >= total) return;
    int* base = lds + (threadIdx.x / kGroup) * group_lds_ints(cap_phys);
    const RayItem* q = ws.q[cur];
    BvhView bv;
    group_closest(
        bv, &ws.counters[C_WORK], total, reinterpret_cast<int2*>(base), base + 2 * cap_phys, cap, cap_phys,
        [&](int i, Trav& s, BvhView& v) {
            const RayItem it = q[i];
            v = bvhs[unpack_scene(it.ot.w, ws.tb)];
            trav_begin(s, vec3(it.ot.x, it.ot.y, it.ot.z), vec3(it.dp.x, it.dp.y, it.dp.z), 1e30f);
            s.time = unpack_time(it.ot.w, ws.tb);
        },
        [&](int i, const Trav& s, const BvhView& v) { ws.hits[i] = bb_hit(s, v); });
}

// shadeKernel (:3470-3696) with the ray's own scene and camera
// Each block's slots are shaded in shading-bin order (block_bin_sort).
constexpr int kShadeBlock = 1024;
__global__ __launch_bounds__(kShadeBlock) void k_bb_shade(const BScene* tab, BWs ws, int cur, int depth, int P) {
    __shared__ int lds[bin_sort_lds_ints<kShadeBlock>()];
    const int count = queue_len(ws.counters, cur);
    if ((int)blockIdx.x * kShadeBlock >= count) return;  // the whole block: past the queue
    int slot = blockIdx.x * kShadeBlock + threadIdx.x, mat = -1, bin = kShadeBins;
    if (slot < count) {
        const mcpt_packed_hit hit = ws.hits[slot];
        const BScene& B = tab[unpack_scene(ws.q[cur][slot].ot.w, ws.tb)];
        bin = shade_bin(B.sv, hit.did_hit != 0, hit.prim_index, mat);
    }
    block_bin_sort<kShadeBlock>(bin, slot, mat, lds);
    const bool active = slot >= 0;
    Bounce out;
    out.next = false;
    out.shadow = false;
    int path = 0, scene = 0;
    float time = 0.0f;
    if (active) {
        const RayItem it = ws.q[cur][slot];
        const mcpt_packed_hit hit = ws.hits[slot];
        path = __float_as_int(it.dp.w);
        scene = unpack_scene(it.ot.w, ws.tb);
        time = unpack_time(it.ot.w, ws.tb);
        const BScene& B = tab[scene];
        const float4 thr = ws.thr[path], smp = ws.smp[path];
        PathState ps;
        ps.thr = vec3(thr.x, thr.y, thr.z);
        ps.prev_pdf = thr.w;
        ps.smp = vec3(smp.x, smp.y, smp.z);
        ps.prev_spec = smp.w != 0.0f;
        Rng rng;
        rng.g.state = ws.rng[path];
        const int pixel = path % P - B.foff;
        rng.g.inc = ((uint64_t)(uint32_t)pixel << 1u) | 1u;
        shade_bounce(B.sv, load3(B.cam.background), B.cam.rr_cap, depth, vec3(it.ot.x, it.ot.y, it.ot.z),
                     vec3(it.dp.x, it.dp.y, it.dp.z), hit.did_hit != 0, hit.t, hit.prim_index, ps, rng, out, time,
                     mat);
        ws.thr[path] = make_float4(ps.thr.x, ps.thr.y, ps.thr.z, ps.prev_pdf);
        ws.smp[path] = make_float4(ps.smp.x, ps.smp.y, ps.smp.z, ps.prev_spec ? 1.0f : 0.0f);
        ws.rng[path] = rng.g.state;
    }
    int nslot, sslot;
    wave_push2(ns_slot(ws.counters, 1 - cur), out.next, out.shadow, nslot, sslot);
    if (out.next) {
        RayItem it;
        it.ot = make_float4(out.no.x, out.no.y, out.no.z, pack_scene_time(scene, time, ws.tb));
        it.dp = make_float4(out.nd.x, out.nd.y, out.nd.z, __int_as_float(path));
        ws.q[1 - cur][nslot] = it;
    }
    if (out.shadow) {
        ShadowItem si;
        si.om = make_float4(out.so.x, out.so.y, out.so.z, out.smax);
        si.dp = make_float4(out.sd.x, out.sd.y, out.sd.z, __int_as_float(path));
        si.c = make_float4(out.contrib.x, out.contrib.y, out.contrib.z, pack_scene_time(scene, time, ws.tb));
        ws.sq[sslot] = si;
    }
}

// the light a shadow ray carries, into its path's sample
__device__ __forceinline__ void add_shadow(float4* smp, const ShadowItem& si, float tr) {
    const int path = __float_as_int(si.dp.w);
    float4 v = smp[path];
    v.x = v.x + si.c.x * tr;
    v.y = v.y + si.c.y * tr;
    v.z = v.z + si.c.z * tr;
    smp[path] = v;
}

// shadowKernel (:3698-3728) for the shadow rays of every scene, on the
// persistent scheme of k_bb_extend: a lane walks the any-hit steps of its
// ray and refills when done. A ray of a scene with a medium (closest hits
// through the medium) is handled whole in one iteration.
// A queue shorter than group_below is k_bb_shadow_group's.
__global__ __launch_bounds__(kBlock) void k_bb_shadow_persistent(const BScene* tab, BWs ws, int cur, int refill,
                                                                int group_below) {
    extern __shared__ int lds_stack[];
    int* stk = lds_stack + threadIdx.x;
    const int total = shadow_len(ws.counters, 1 - cur);
    if (total < group_below) return;
    if ((int)blockIdx.x * kBlock >= total) return;  // blocks the queue does not need
    const int lane = lane_id();
    Trav s;
    BvhView bv;
    int bv_scene = -1;
    bool has = false;
    int slot = -1;
    bool exhausted = false;
    while (true) {
        if (!exhausted) {
            const uint64_t idle = ballot64(!has);
            const int nidle = __popcll(idle);
            if (nidle >= refill) {
                const int leader = __builtin_ctzll(idle);
                int first = 0;
                if (lane == leader) first = atomicAdd(&ws.counters[C_WORK_SHADOW], nidle);
                first = __shfl(first, leader);
                if (!has) {
                    const int q = first + rank_below(idle);
                    if (q < total) {
                        const ShadowItem si = ws.sq[q];
                        const int sc = unpack_scene(si.c.w, ws.tb);
                        const float tm = unpack_time(si.c.w, ws.tb);
                        const vec3 o(si.om.x, si.om.y, si.om.z), d(si.dp.x, si.dp.y, si.dp.z);
                        if (tab[sc].sv.has_medium) {
                            const float tr = shadow_transmittance(tab[sc].sv, o, d, si.om.w, stk, kBlock, tm);
                            if (tr > 0.0f) add_shadow(ws.smp, si, tr);
                        } else {
                            if (sc != bv_scene) {
                                bv = tab[sc].sv.bvh;
                                bv_scene = sc;
                            }
                            trav_begin(s, o, d, si.om.w);
                            s.time = tm;
                            slot = q;
                            has = true;
                        }
                    }
                }
                if (first + nidle >= total) exhausted = true;
            }
        }
        if (exhausted && ballot64(has) == 0) break;
        if (has) {
            const int st = step_wide_any(bv, s, stk, kBlock);
            if (st != 0) {
                if (st == 1) add_shadow(ws.smp, ws.sq[slot], 1.0f);
                has = false;
            }
        }
    }
}

// a short shadow queue: 16 lanes per ray (group_any); only for batches
// without a medium (group_below is 0 otherwise)
__global__ __launch_bounds__(kBlock) void k_bb_shadow_group(const BScene* tab, BWs ws, int cur, int group_below,
                                                           int cap, int cap_phys) {
    extern __shared__ int lds[];
    const int total = shadow_len(ws.counters, 1 - cur);
    if (total >= group_below || (int)blockIdx.x * (kBlock / kGroup) >= total) return;
    int* base = lds + (threadIdx.x / kGroup) * group_lds_ints(cap_phys);
    BvhView bv;
    group_any(
        bv, &ws.counters[C_WORK_SHADOW], total, reinterpret_cast<int2*>(base), base + 2 * cap_phys, cap, cap_phys,
        [&](int i, Trav& s, BvhView& v) {
            const ShadowItem si = ws.sq[i];
            v = tab[unpack_scene(si.c.w, ws.tb)].sv.bvh;
            trav_begin(s, vec3(si.om.x, si.om.y, si.om.z), vec3(si.dp.x, si.dp.y, si.dp.z), si.om.w);
            s.time = unpack_time(si.c.w, ws.tb);
        },
        [&](int i, bool blocked) {
            if (!blocked) add_shadow(ws.smp, ws.sq[i], 1.0f);
        });
}

// finalizeSampleKernel (:3740-3755): a pixel's samples in sample order
__global__ __launch_bounds__(kBlock) void k_bb_finalize(const BScene* tab, int nsc, int P, int nb, const float4* smp) {
    const int q = blockIdx.x * kBlock + threadIdx.x;
    if (q >= P) return;
    const int s = scene_of_pixel(tab, nsc, q);
    float* a = tab[s].accum + 3 * (q - tab[s].foff);
    float r = a[0], g = a[1], b = a[2];
    for (int k = 0; k < nb; k++) {
        const float4 v = smp[(size_t)k * P + q];
        const vec3 c = clamp_sample(vec3(v.x, v.y, v.z));
        r = r + c.x;
        g = g + c.y;
        b = b + c.z;
    }
    a[0] = r;
    a[1] = g;
    a[2] = b;
}

// Rays of the second bounce an autotune launch traces: enough that a launch
// lasts well over the ~0.5 ms its slowest rays take by themselves.
constexpr int kCalRays = 2 << 20;

constexpr size_t kPathBytes = 2 * sizeof(RayItem) + sizeof(mcpt_packed_hit) + sizeof(ShadowItem) + 2 * sizeof(float4) +
                              sizeof(uint64_t);

// Device tables and one wavefront workspace of a batched render; the
// memory goes back to the context when the object does.
struct BatchRun {
    Ctx* ctx = nullptr;
    std::vector<BScene> tab;
    std::vector<BvhView> bvh;
    int n = 0;
    int64_t P = 0;  // pixels of all scenes
    int max_stack = 1;
    bool moving = false;  // ray records then carry their time too
    bool medium = false;  // a scene with a medium: shadow rays one lane each
    void* mem = nullptr;
    size_t bytes = 0;
    BScene* dtab = nullptr;
    BvhView* dbvh = nullptr;
    BWs ws{};
    size_t lds = 0;
    int per_cu = 1;  // occupancy limit of the persistent extend
    BatchRun() = default;
    BatchRun(const BatchRun&) = delete;
    BatchRun& operator=(const BatchRun&) = delete;
    ~BatchRun() {
        if (mem) dev_free(ctx, mem, bytes);
    }
};

// the scene table (accum[i] overrides scene i's film: the autotuner's
// scratch films; nullptr -> films)
int batch_tables(const mcpt_scene* scenes, const mcpt_camera* cams, int n, float* const* accum, BatchRun& br) {
    br.ctx = scenes[0]->ctx;
    br.n = n;
    br.tab.resize(n);
    br.bvh.resize(n);
    for (int i = 0; i < n; i++) {
        br.moving = br.moving || scenes[i]->moving_spheres;
        br.medium = br.medium || scenes[i]->has_medium;
    }
    for (int i = 0; i < n; i++) {
        br.tab[i].sv = scenes[i]->view();
        br.tab[i].cam = cams[i];
        br.tab[i].accum = accum[i];
        br.tab[i].foff = (int)br.P;
        br.tab[i].npix = cams[i].width * cams[i].height;
        br.bvh[i] = br.tab[i].sv.bvh;
        br.P += br.tab[i].npix;
        br.max_stack = std::max(br.max_stack, scenes[i]->max_depth);
    }
    if (br.P > (1 << 30)) return set_error(MCPT_ERR_INVALID, "too many pixels in one batch");
    return MCPT_OK;
}

// device tables and a workspace of `capacity` paths in one allocation; the
// tables are copied on the context's stream
int batch_alloc(BatchRun& br, int64_t capacity) {
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const int n = br.n;
    const size_t parts[] = {al(n * sizeof(BScene)),        al(n * sizeof(BvhView)),
                            al(capacity * sizeof(RayItem)), al(capacity * sizeof(RayItem)),
                            al(capacity * sizeof(mcpt_packed_hit)), al(capacity * sizeof(ShadowItem)),
                            al(capacity * sizeof(float4)),  al(capacity * sizeof(float4)),
                            al(capacity * sizeof(uint64_t)), al(C_NCOUNTERS * sizeof(int))};
    size_t total = 0;
    for (size_t x : parts) total += x;
    MCPT_TRY(dev_alloc(br.ctx, &br.mem, total));
    br.bytes = total;
    char* b = (char*)br.mem;
    int k = 0;
    br.dtab = (BScene*)b;
    b += parts[k++];
    br.dbvh = (BvhView*)b;
    b += parts[k++];
    BWs& ws = br.ws;
    ws.q[0] = (RayItem*)b, b += parts[k++];
    ws.q[1] = (RayItem*)b, b += parts[k++];
    ws.hits = (mcpt_packed_hit*)b, b += parts[k++];
    ws.sq = (ShadowItem*)b, b += parts[k++];
    ws.thr = (float4*)b, b += parts[k++];
    ws.smp = (float4*)b, b += parts[k++];
    ws.rng = (uint64_t*)b, b += parts[k++];
    ws.counters = (int*)b;
    ws.tb = br.moving ? 24 : 0;
    br.lds = (size_t)kBlock * stack_entries(br.max_stack) * sizeof(int);
    int per_cu = 0;
    MCPT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_bb_extend, kBlock, br.lds));
    br.per_cu = std::max(1, per_cu);
    hipStream_t st = br.ctx->stream;
    MCPT_HIP(hipMemcpyAsync(br.dtab, br.tab.data(), n * sizeof(BScene), hipMemcpyHostToDevice, st));
    MCPT_HIP(hipMemcpyAsync(br.dbvh, br.bvh.data(), n * sizeof(BvhView), hipMemcpyHostToDevice, st));
    return MCPT_OK;
}

// persistent blocks of a tune's grid
int tuned_blocks(const BatchRun& br, const mcpt_batch_tune* tune) {
    int per_cu = br.per_cu;
    if (tune && tune->blocks_per_cu > 0) per_cu = std::min(per_cu, tune->blocks_per_cu);
    return per_cu * br.ctx->cu_count;
}

// samples of every scene per pass: the tuned value, or as many as fit a
// quarter of the budget (at most 64 Mi paths)
int default_samples_per_pass(const Ctx* ctx, int64_t P, int spp) {
    const int64_t cap = std::min<int64_t>(int64_t(64) << 20, (int64_t)(ctx->budget / 4) / (int64_t)kPathBytes);
    int nb = (int)std::max<int64_t>(1, std::min<int64_t>(spp, cap / std::max<int64_t>(1, P)));
    if (P * nb > (int64_t(1) << 30)) nb = std::max(1, (int)((int64_t(1) << 30) / P));
    return nb;
}

// One batched render.
int render_batch(const mcpt_scene* scenes, const mcpt_camera* cams, int n, const mcpt_render_opts* o,
                 float* const* accum, const mcpt_batch_tune* tune, mcpt_render_stats* stats) {
    BatchRun br;
    MCPT_TRY(batch_tables(scenes, cams, n, accum, br));
    Ctx* ctx = br.ctx;
    hipStream_t st = ctx->stream;
    const int64_t P = br.P;
    int nb = tune && tune->samples_per_pass > 0 ? tune->samples_per_pass : default_samples_per_pass(ctx, P, o->spp);
    nb = std::max(1, std::min(nb, o->spp));
    if (P * nb > (int64_t(1) << 30)) nb = std::max(1, (int)((int64_t(1) << 30) / P));
    const int refill = tune && tune->refill_lanes > 0 ? std::min(64, tune->refill_lanes) : 32;
    MCPT_TRY(batch_alloc(br, P * nb));
    const BWs& ws = br.ws;
    const int persistent_blocks = tuned_blocks(br, tune);
    const size_t lds = br.lds;
    // short queues: 16 lanes per ray (mcpt_render's default threshold)
    const int gb = o->group_below < 0 ? 0 : o->group_below > 0 ? o->group_below : 32768;
    const int gbs = br.medium ? 0 : gb;
    const int gcap_phys = group_stack_pairs(stack_entries(br.max_stack));
    const size_t glds = (size_t)(kBlock / kGroup) * group_lds_ints(gcap_phys) * sizeof(int);
    int group_blocks = 0;
    if (gb > 0) {
        int pc = 0;
        MCPT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&pc, k_bb_extend_group, kBlock, glds));
        group_blocks = std::min(std::max(1, pc) * ctx->cu_count, (gb + kBlock / kGroup - 1) / (kBlock / kGroup));
    }
    // Each bounce's shadow rays run on the context's second stream beside
    // the next bounce's closest-hit rays: they add light to the samples of
    // their paths (smp), which only the next shade and finalize read, and
    // the extend reads neither the shadow queue nor the samples. The next
    // shade waits for them (and so its shadow items never overwrite a
    // queue still being read).
    hipStream_t sd = ctx->side;
    hipEvent_t e0 = nullptr, e1 = nullptr, ev_shade = nullptr, ev_shadow = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ev_shade, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ev_shadow, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventRecord(e0, st);
    uint64_t launches = 0;
    if (e == hipSuccess) {
        k_b_fill<<<1, C_NCOUNTERS, 0, st>>>(ws.counters, C_NCOUNTERS, 0);
        for (int k0 = 0; k0 < o->spp && e == hipSuccess; k0 += nb) {
            const int cnt = std::min(nb, o->spp - k0);
            const int64_t paths = P * cnt;
            const int grid = blocks_for(paths);
            int cur = 0;
            k_bb_init<<<1, 64, 0, st>>>(ws.counters, (int)paths);
            k_bb_camera<<<grid, kBlock, 0, st>>>(br.dtab, n, (int)P, (uint32_t)(o->sample_begin + k0), cnt, o->seed,
                                                 ws);
            for (int d = 0; d < o->max_depth && e == hipSuccess; d++) {
                k_bb_bounce<<<1, 64, 0, st>>>(ws.counters, cur);
                k_bb_extend<<<std::min(grid, persistent_blocks), kBlock, lds, st>>>(br.dbvh, ws, cur, refill, gb);
                if (gb > 0)
                    k_bb_extend_group<<<std::min(grid, group_blocks), kBlock, glds, st>>>(br.dbvh, ws, cur, gb,
                                                                                        kGroupCap, gcap_phys);
                if (d > 0) e = hipStreamWaitEvent(st, ev_shadow, 0);  // the previous bounce's shadow rays
                k_bb_shade<<<(int)((paths + kShadeBlock - 1) / kShadeBlock), kShadeBlock, 0, st>>>(br.dtab, ws, cur, d,
                                                                                                  (int)P);
                if (e == hipSuccess) e = hipEventRecord(ev_shade, st);
                if (e == hipSuccess) e = hipStreamWaitEvent(sd, ev_shade, 0);
                k_bb_shadow_total<<<1, 64, 0, sd>>>(ws.counters, cur);
                k_bb_shadow_persistent<<<std::min(grid, persistent_blocks), kBlock, lds, sd>>>(br.dtab, ws, cur,
                                                                                                refill, gbs);
                if (gbs > 0)
                    k_bb_shadow_group<<<std::min(grid, group_blocks), kBlock, glds, sd>>>(br.dtab, ws, cur, gbs,
                                                                                        kGroupCap, gcap_phys);
                if (e == hipSuccess) e = hipEventRecord(ev_shadow, sd);
                launches++;
                cur = 1 - cur;
            }
            if (e == hipSuccess) e = hipStreamWaitEvent(st, ev_shadow, 0);
            k_bb_finalize<<<blocks_for(P), kBlock, 0, st>>>(br.dtab, n, (int)P, cnt, ws.smp);
        }
        if (e == hipSuccess) e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    uint64_t tot[2] = {0, 0};
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e == hipSuccess) e = hipMemcpy(&tot[0], ws.counters + C_EXT, sizeof(uint64_t), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(&tot[1], ws.counters + C_SHD, sizeof(uint64_t), hipMemcpyDeviceToHost);
    if (e != hipSuccess) (void)hipStreamSynchronize(sd);
    for (hipEvent_t x : {e0, e1, ev_shade, ev_shadow})
        if (x) (void)hipEventDestroy(x);
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(st);
        return set_error(MCPT_ERR_HIP, "batched render: %s", hipGetErrorString(e));
    }
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->total_ms = ms;
        stats->extend_launches = launches;
        stats->extend_rays = tot[0];
        stats->shadow_rays = tot[1];
        stats->paths = (uint64_t)P * o->spp;
    }
    return MCPT_OK;
}

int check_batch(const mcpt_scene* scenes, int32_t n) {
    if (!scenes || n < 1) return set_error(MCPT_ERR_INVALID, "no scenes");
    for (int i = 0; i < n; i++)
        if (!scenes[i] || scenes[i]->ctx != scenes[0]->ctx)
            return set_error(MCPT_ERR_INVALID, "scene %d is missing or belongs to another context", i);
    return MCPT_OK;
}

// ray records carry scene and time in one slot for at most 256 scenes
int check_moving_spheres(const mcpt_scene* scenes, int32_t n) {
    bool moving = false;
    for (int i = 0; i < n; i++) moving = moving || scenes[i]->moving_spheres;
    if (moving && n > 256) return set_error(MCPT_ERR_INVALID, "a batch with moving spheres takes at most 256 scenes");
    return MCPT_OK;
}

int check_render_batch(const mcpt_scene* scenes, const mcpt_camera* cams, int32_t n, const mcpt_render_opts* o) {
    MCPT_TRY(check_batch(scenes, n));
    if (!cams || !o) return set_error(MCPT_ERR_INVALID, "bad arguments");
    if (o->mode != MCPT_MODE_PERSISTENT) return set_error(MCPT_ERR_INVALID, "batched renders use the persistent mode");
    if (o->spp < 1 || o->max_depth < 1 || o->sample_begin < 0)
        return set_error(MCPT_ERR_INVALID, "spp, max_depth must be >= 1 and sample_begin >= 0");
    for (int i = 0; i < n; i++) {
        if (!scenes[i]->built) return set_error(MCPT_ERR_INVALID, "scene %d has no BVH", i);
        if (cams[i].width < 1 || cams[i].height < 1) return set_error(MCPT_ERR_INVALID, "camera %d is empty", i);
    }
    return MCPT_OK;
}

}  // namespace
}  // namespace mcpt

using namespace mcpt;

extern "C" {

int mcpt_lbvh_build_batch(const mcpt_scene* scenes, int32_t n, float* build_ms) {
    MCPT_TRY(check_batch(scenes, n));
    Ctx* ctx = scenes[0]->ctx;
    MCPT_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    // scenes with one primitive (their special root) or with spheres take
    // the single-scene path
    std::vector<int> multi;
    float single_ms = 0.f;
    for (int i = 0; i < n; i++) {
        if (scenes[i]->n_prims() < 1) return set_error(MCPT_ERR_INVALID, "scene %d is empty", i);
        if (scenes[i]->n_prims() == 1 || scenes[i]->n_spheres > 0) {
            MCPT_TRY(lbvh_build(scenes[i], nullptr));
            single_ms += scenes[i]->last_build_ms;
        } else {
            multi.push_back(i);
        }
    }
    if (multi.empty()) {
        if (build_ms) *build_ms = single_ms;
        return MCPT_OK;
    }
    const int ns = (int)multi.size();
    std::vector<Seg> seg(ns);
    std::vector<int> chunk_off(ns + 1, 0);
    int64_t total = 0, total_int = 0;
    for (int k = 0; k < ns; k++) {
        mcpt_scene_s* s = scenes[multi[k]];
        MCPT_TRY(s->nodes.alloc(ctx, s->n - 1));
        MCPT_TRY(s->leaf_tri.alloc(ctx, s->n));
        MCPT_TRY(s->leaf_prim.alloc(ctx, s->n));
        MCPT_TRY(s->root_box.alloc(ctx, 6));
        MCPT_TRY(s->tie.alloc(ctx, s->n));
        MCPT_TRY(s->wide.alloc(ctx, wide_capacity(s->n)));
        seg[k] = Seg{(int)total, s->n, (int)total_int, 0, s->hot.ptr, s->nodes.ptr, s->leaf_tri.ptr, s->leaf_prim.ptr,
                     s->root_box.ptr, s->tie.ptr, s->wide.ptr};
        total += s->n;
        total_int += s->n - 1;
        chunk_off[k + 1] = chunk_off[k] + (s->n + kChunk - 1) / kChunk;
    }
    const int n_chunks = chunk_off[ns];
    if (total > (int64_t(1) << 30)) return set_error(MCPT_ERR_INVALID, "too many triangles in one batch");
    const int tiles = blocks_for(total, kTile);
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t parts[] = {al(ns * sizeof(Seg)),         al(ns * 6 * sizeof(float)),  al(total * sizeof(uint32_t)),
                            al(total * sizeof(uint32_t)), al(total * sizeof(uint32_t)), al(total * sizeof(int)),
                            al(total * sizeof(int)),      al((size_t)256 * tiles * sizeof(int)),
                            al(total * sizeof(int)),      al(total_int * sizeof(int)),  al((2 * ns + 65) * sizeof(int)),
                            al((ns + 1) * sizeof(int)),   al((size_t)n_chunks * 6 * sizeof(float)),
                            al(total_int * sizeof(uint32_t)), al(scan_chunks(256 * (int64_t)tiles) * sizeof(int))};
    size_t bytes = 0;
    for (size_t x : parts) bytes += x;
    void* mem = nullptr;
    MCPT_TRY(dev_alloc(ctx, &mem, bytes));
    char* b = (char*)mem;
    int k = 0;
    Seg* dseg = (Seg*)b;
    b += parts[k++];
    float* sb = (float*)b;
    b += parts[k++];
    uint32_t* morton = (uint32_t*)b;
    b += parts[k++];
    uint32_t* keys_a = (uint32_t*)b;
    b += parts[k++];
    uint32_t* keys_b = (uint32_t*)b;
    b += parts[k++];
    int* vals_a = (int*)b;
    b += parts[k++];
    int* vals_b = (int*)b;
    b += parts[k++];
    int* counts = (int*)b;
    b += parts[k++];
    int* leaf_parent = (int*)b;
    b += parts[k++];
    int* ids_sorted = (int*)b;
    b += parts[k++];
    int* depth = (int*)b;  // per-scene depth, the 65 level offsets, per-scene wide node counts
    b += parts[k++];
    int* dchunk = (int*)b;
    b += parts[k++];
    float* partial = (float*)b;
    b += parts[k++];
    uint32_t* lvl_sorted = (uint32_t*)b;
    b += parts[k++];
    int* ssum = (int*)b;  // chunk sums of the histogram scans
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipMemcpyAsync(dseg, seg.data(), ns * sizeof(Seg), hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dchunk, chunk_off.data(), (ns + 1) * sizeof(int), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipEventRecord(e0, st);
    uint32_t* lvl_keys = nullptr;  // level of each global internal node
    int* widx = nullptr;           // its wide index (after the level sort)
    uint32_t* free_keys = nullptr;  // Morton sort buffers, free after the Karras tree
    int* free_vals = nullptr;
    if (e == hipSuccess) {
        const int n_all = (int)total, n_int = (int)total_int;
        k_b_bounds_chunks<<<n_chunks, kBlock, 0, st>>>(dseg, ns, dchunk, partial);
        k_b_bounds_final<<<ns, kBlock, 0, st>>>(dchunk, partial, sb);
        k_b_morton<<<blocks_for(n_all), kBlock, 0, st>>>(dseg, ns, n_all, sb, morton, keys_a, vals_a);
        uint32_t *kin = keys_a, *kout = keys_b;
        int *vin = vals_a, *vout = vals_b;
        for (int shift = 0; shift < 30; shift += 8) {
            k_radix_hist<<<tiles, kBlock, 0, st>>>(kin, n_all, shift, tiles, counts);
            scan_exclusive(counts, 256 * tiles, ssum, st);
            k_radix_scatter<<<tiles, kBlock, 0, st>>>(kin, vin, kout, vout, n_all, shift, tiles, counts);
            std::swap(kin, kout);
            std::swap(vin, vout);
        }
        // then the scene index, 8 bits per pass (segments stay in scene order)
        for (int shift = 0; (ns - 1) >> shift; shift += 8) {
            k_b_scene_keys<<<blocks_for(n_all), kBlock, 0, st>>>(dseg, ns, n_all, vin, kin);
            k_radix_hist<<<tiles, kBlock, 0, st>>>(kin, n_all, shift, tiles, counts);
            scan_exclusive(counts, 256 * tiles, ssum, st);
            k_radix_scatter<<<tiles, kBlock, 0, st>>>(kin, vin, kout, vout, n_all, shift, tiles, counts);
            std::swap(kin, kout);
            std::swap(vin, vout);
        }
        k_b_gather_keys<<<blocks_for(n_all), kBlock, 0, st>>>(morton, vin, n_all, kin);
        k_b_leaves<<<blocks_for(n_all), kBlock, 0, st>>>(dseg, ns, n_all, vin);
        k_b_fill<<<blocks_for(2 * ns + 65), kBlock, 0, st>>>(depth, 2 * ns + 65, 0);
        k_b_karras<<<blocks_for(n_int), kBlock, 0, st>>>(dseg, ns, n_int, kin, leaf_parent);
        // refit level by level over all scenes: levels into kout/vout (free
        // after the sort), one radix pass by level, offsets to the host
        const int itiles = blocks_for(n_int, kTile);
        k_b_levels<<<blocks_for(n_all), kBlock, 0, st>>>(dseg, ns, n_all, leaf_parent, kout, vout, depth,
                                                         depth + ns + 65);
        k_radix_hist<<<itiles, kBlock, 0, st>>>(kout, n_int, 0, itiles, counts);
        scan_exclusive(counts, 256 * itiles, ssum, st);
        k_radix_scatter<<<itiles, kBlock, 0, st>>>(kout, vout, lvl_sorted, ids_sorted, n_int, 0, itiles, counts);
        k_level_offsets<<<1, 128, 0, st>>>(counts, itiles, n_int, depth + ns);
        lvl_keys = kout;
        widx = vout;
        free_keys = kin;
        free_vals = vin;
        e = hipGetLastError();
    }
    std::vector<int> depths(2 * ns + 65, 0);
    if (e == hipSuccess)
        e = hipMemcpyAsync(depths.data(), depth, (2 * ns + 65) * sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    int bad = -1, max_d = 0;
    for (int k2 = 0; k2 < ns && e == hipSuccess; k2++) {
        if (depths[k2] < 1 || depths[k2] > 64) bad = k2;
        max_d = std::max(max_d, depths[k2]);
    }
    if (e == hipSuccess && bad < 0) {
        const int n_int = (int)total_int;
        const int* off = depths.data() + ns;
        for (int L = max_d - 1; L >= 0 && e == hipSuccess; L--)
            if (off[L + 1] > off[L])
                k_b_refit_level<<<blocks_for(off[L + 1] - off[L]), kBlock, 0, st>>>(dseg, ns, ids_sorted + off[L],
                                                                                    off[L + 1] - off[L]);
        // wide indices: each scene's even-level nodes in (level, index)
        // order, as the single-scene build numbers them (a stable sort of
        // the level-sorted nodes by scene, then ranks), into the level
        // sort's value buffer; then the collapse
        if (e == hipSuccess) {
            const int itiles = blocks_for(n_int, kTile);
            k_b_node_scene<<<blocks_for(n_int), kBlock, 0, st>>>(dseg, ns, ids_sorted, n_int, free_keys);
            uint32_t *ka = free_keys, *kb = lvl_sorted;
            int *va = ids_sorted, *vb = free_vals;
            for (int shift = 0; (ns - 1) >> shift; shift += 8) {
                k_radix_hist<<<itiles, kBlock, 0, st>>>(ka, n_int, shift, itiles, counts);
                scan_exclusive(counts, 256 * itiles, ssum, st);
                k_radix_scatter<<<itiles, kBlock, 0, st>>>(ka, va, kb, vb, n_int, shift, itiles, counts);
                std::swap(ka, kb);
                std::swap(va, vb);
            }
            int* rank = reinterpret_cast<int*>(ka);  // the scene keys are done with
            k_wide_flags<<<blocks_for(n_int), kBlock, 0, st>>>(va, lvl_keys, n_int, rank);
            scan_exclusive(rank, n_int, ssum, st);
            k_wide_scatter<<<blocks_for(n_int), kBlock, 0, st>>>(va, lvl_keys, rank, n_int, widx);
            k_b_collapse<<<blocks_for(n_int), kBlock, 0, st>>>(dseg, ns, lvl_keys, widx, n_int);
            e = hipGetLastError();
        }
    }
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    float ms = 0.f;
    if (e == hipSuccess && bad < 0) e = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    dev_free(ctx, mem, bytes);
    if (e != hipSuccess) return set_error(MCPT_ERR_HIP, "batched LBVH build: %s", hipGetErrorString(e));
    if (bad >= 0)
        return set_error(MCPT_ERR_INVALID, "scene %d: BVH depth %d outside [1, 64]", multi[bad], depths[bad]);
    for (int k2 = 0; k2 < ns; k2++) {
        mcpt_scene_s* s = scenes[multi[k2]];
        s->max_depth = depths[k2];
        s->n_wide = depths[ns + 65 + k2];
        s->built = true;
        s->last_build_ms = ms;
    }
    if (build_ms) *build_ms = ms + single_ms;
    return MCPT_OK;
}

int mcpt_render_batch(const mcpt_scene* scenes, const mcpt_camera* cams, int32_t n, const mcpt_render_opts* o,
                      mcpt_film* films, const mcpt_batch_tune* tune, mcpt_render_stats* stats) {
    MCPT_TRY(check_render_batch(scenes, cams, n, o));
    for (int i = 0; i < n; i++) MCPT_TRY(scenes[i]->check_textures());
    MCPT_TRY(check_moving_spheres(scenes, n));
    if (!films) return set_error(MCPT_ERR_INVALID, "films is NULL");
    std::vector<float*> acc(n);
    for (int i = 0; i < n; i++) {
        if (!films[i] || films[i]->w != cams[i].width || films[i]->h != cams[i].height)
            return set_error(MCPT_ERR_INVALID, "film %d missing or not %dx%d", i, cams[i].width, cams[i].height);
        acc[i] = films[i]->accum.ptr;
    }
    MCPT_HIP(hipSetDevice(scenes[0]->ctx->device));
    return render_batch(scenes, cams, n, o, acc.data(), tune, stats);
}

int mcpt_batch_autotune(const mcpt_scene* scenes, const mcpt_camera* cams, int32_t n, const mcpt_render_opts* o,
                        mcpt_batch_tune* best, double* search_ms) {
    MCPT_TRY(check_render_batch(scenes, cams, n, o));
    for (int i = 0; i < n; i++) MCPT_TRY(scenes[i]->check_textures());
    MCPT_TRY(check_moving_spheres(scenes, n));
    if (!best) return set_error(MCPT_ERR_INVALID, "best is NULL");
    MCPT_HIP(hipSetDevice(scenes[0]->ctx->device));
    // Calibration: the camera rays of one sample after the requested range
    // (the same kind of paths) and their first bounce, with the default
    // shape; then the extend launch of the second bounce, timed on the
    // first kCalRays of its queue for each candidate shape. The extend
    // launch is most of a batched render's time and reads nothing it
    // writes, so it can run again on the same queue; one launch per shape
    // costs ~1.5 ms where a whole calibration pass costs ~9 ms.
    BatchRun br;
    std::vector<float*> no_films(n, nullptr);  // the calibration writes no film
    MCPT_TRY(batch_tables(scenes, cams, n, no_films.data(), br));
    Ctx* ctx = br.ctx;
    hipStream_t st = ctx->stream;
    const int64_t P = br.P;
    MCPT_TRY(batch_alloc(br, P));
    const BWs& ws = br.ws;
    const int grid = blocks_for(P);
    std::vector<hipEvent_t> ev;
    auto event = [&](hipEvent_t* out) {
        hipEvent_t e = nullptr;
        MCPT_HIP(hipEventCreate(&e));
        ev.push_back(e);
        MCPT_HIP(hipEventRecord(e, st));
        if (out) *out = e;
        return MCPT_OK;
    };
    auto cleanup = [&](int rc) {
        (void)hipStreamSynchronize(st);
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        return rc;
    };
    hipEvent_t e_start = nullptr;
    int rc = event(&e_start);
    if (rc < 0) return cleanup(rc);
    k_b_fill<<<1, C_NCOUNTERS, 0, st>>>(ws.counters, C_NCOUNTERS, 0);
    k_bb_init<<<1, 64, 0, st>>>(ws.counters, (int)P);
    k_bb_camera<<<grid, kBlock, 0, st>>>(br.dtab, n, (int)P, (uint32_t)(o->sample_begin + o->spp), 1, o->seed, ws);
    k_bb_bounce<<<1, 64, 0, st>>>(ws.counters, 0);
    k_bb_extend<<<std::min(grid, tuned_blocks(br, nullptr)), kBlock, br.lds, st>>>(br.dbvh, ws, 0, 32, 0);
    if (o->max_depth > 1) k_bb_shade<<<(int)((P + kShadeBlock - 1) / kShadeBlock), kShadeBlock, 0, st>>>(
        br.dtab, ws, 0, 0, (int)P);
    if (hipGetLastError() != hipSuccess) return cleanup(set_error(MCPT_ERR_HIP, "autotune: launch failed"));
    int len[2] = {0, 0};
    if (hipMemcpyAsync(len, ws.counters + (o->max_depth > 1 ? C_NS1 : C_NS0), sizeof(int), hipMemcpyDeviceToHost,
                       st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return cleanup(set_error(MCPT_ERR_HIP, "autotune: reading the queue length failed"));
    const int q = o->max_depth > 1 ? 1 : 0;  // the queue the timed launches trace
    const int m = std::min(len[0], kCalRays);
    mcpt_batch_tune cur = {default_samples_per_pass(ctx, P, o->spp), 32, 0, 0};
    hipEvent_t e_end = e_start;
    if (m > 0) {
        k_b_fill<<<1, 64, 0, st>>>(ws.counters + (q ? C_NS1 : C_NS0), 1, m);
        auto launch = [&](const mcpt_batch_tune& t) {
            k_b_fill<<<1, 64, 0, st>>>(ws.counters + C_WORK, 1, 0);
            k_bb_extend<<<std::min(blocks_for(m), tuned_blocks(br, &t)), kBlock, br.lds, st>>>(br.dbvh, ws, q,
                                                                                                t.refill_lanes, 0);
        };
        auto timed = [&](const mcpt_batch_tune& t, float& ms) {
            hipEvent_t a = nullptr, b = nullptr;
            MCPT_TRY(event(&a));
            launch(t);
            MCPT_TRY(event(&b));
            MCPT_HIP(hipEventSynchronize(b));
            MCPT_HIP(hipEventElapsedTime(&ms, a, b));
            e_end = b;
            return MCPT_OK;
        };
        launch(cur);  // warm-up: the queue's rays and nodes into the caches
        float cur_ms = 0.f;
        rc = timed(cur, cur_ms);
        // the refill threshold, then the grid, each kept when it is faster
        const int refill_c[] = {16, 48};
        const int blocks_c[] = {4, 2};
        for (int v : refill_c) {
            if (rc < 0) break;
            mcpt_batch_tune t = cur;
            t.refill_lanes = v;
            float ms = 0.f;
            rc = timed(t, ms);
            if (rc >= 0 && ms < cur_ms) cur = t, cur_ms = ms;
        }
        for (int v : blocks_c) {
            if (rc < 0 || v >= br.per_cu) continue;
            mcpt_batch_tune t = cur;
            t.blocks_per_cu = v;
            float ms = 0.f;
            rc = timed(t, ms);
            if (rc >= 0 && ms < cur_ms) cur = t, cur_ms = ms;
        }
        if (rc < 0) return cleanup(rc);
    }
    float spent = 0.f;
    if (e_end != e_start) {
        if (hipEventElapsedTime(&spent, e_start, e_end) != hipSuccess)
            return cleanup(set_error(MCPT_ERR_HIP, "autotune: event timing failed"));
    }
    *best = cur;
    if (search_ms) *search_ms = spent;
    return cleanup(MCPT_OK);
}

}  // extern "C"
