// lbvh.hip — linear BVH build on gfx950.
//
// Stages (the reference does stages 1-3 on the host, src/path_tracer.cu
// :1930-1972 and :4224-4245, and stages 4-5 on the GPU, :1063-1090):
//   1. triangle boxes padded by 1e-3 and the scene bounds (two-level
//      min/max reduction; exact, so order does not matter)
//   2. 30-bit Morton codes of the centroids (expandBits/quantize/morton3D,
//      :913-935)
//   3. stable LSD radix sort of (code, index): 4 passes of 8 bits, each a
//      tile histogram, one exclusive scan and an LDS-ranked scatter whose
//      ranks come from wave64 ballots (8 ballots match the lanes holding
//      the same digit, mbcnt ranks them)
//   4. Karras radix tree (delta/determineRange/findSplit, :1337-1386,
//      :1417-1447); each internal node records its children's references,
//      and a leaf child's box is written straight into its parent's slot
//   5. refit (refitBoundsKernel, :1398-1415) level by level: node levels
//      from the leaves' climbs, one radix pass orders the nodes by level,
//      then one launch per level, deepest first, merges each node's child
//      boxes into its own slot in the parent. The climbs also rank every
//      leaf in the reference's leaf-test order (the tie rule of the fast
//      traversal).
//   6. the four-wide quantized tree the fast traversal walks: one wide node
//      per even-level binary node, children quantized outward to 8 bits.
#include <algorithm>
#include <cstring>

#include "lbvh_common.hpp"
#include "scene.hpp"

namespace mcpt {

namespace {

__device__ __forceinline__ void merge6(float a[6], const float b[6]) {
    for (int k = 0; k < 3; k++) a[k] = fminf(a[k], b[k]);
    for (int k = 3; k < 6; k++) a[k] = fmaxf(a[k], b[k]);
}

// block-wide min/max of a 6-float box through LDS
__device__ void block_reduce_box(float b[6], float* out) {
    __shared__ float sh[6][kBlock];
    for (int k = 0; k < 6; k++) sh[k][threadIdx.x] = b[k];
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            for (int k = 0; k < 3; k++) sh[k][threadIdx.x] = fminf(sh[k][threadIdx.x], sh[k][threadIdx.x + s]);
            for (int k = 3; k < 6; k++) sh[k][threadIdx.x] = fmaxf(sh[k][threadIdx.x], sh[k][threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 6; k++) out[k] = sh[k][0];
}

// Primitive i of the build: the spheres first, then the triangles (the
// order of finalizeScene's primitive list, :1955-1972).
__device__ __forceinline__ void prim_box(const float* hot, const mcpt_sphere* sph, int S, int i, float b[6]) {
    if (i < S) {
        const mcpt_sphere& sp = sph[i];
        sphere_box(load3(sp.center), load3(sp.center1), sp.moving != 0, sp.radius, b);
        return;
    }
    float t[9];
    for (int k = 0; k < 9; k++) t[k] = hot[(size_t)(i - S) * 9 + k];
    tri_box(t, b);
}

__global__ __launch_bounds__(kBlock) void k_bounds_partial(const float* hot, const mcpt_sphere* sph, int S, int n,
                                                           float* partial) {
    float acc[6] = {__builtin_inff(), __builtin_inff(), __builtin_inff(), -__builtin_inff(), -__builtin_inff(),
                    -__builtin_inff()};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        float b[6];
        prim_box(hot, sph, S, i, b);
        merge6(acc, b);
    }
    block_reduce_box(acc, partial + 6 * blockIdx.x);
}

__global__ __launch_bounds__(kBlock) void k_bounds_final(const float* partial, int np, float* out) {
    float acc[6] = {__builtin_inff(), __builtin_inff(), __builtin_inff(), -__builtin_inff(), -__builtin_inff(),
                    -__builtin_inff()};
    for (int i = threadIdx.x; i < np; i += kBlock) merge6(acc, partial + 6 * i);
    block_reduce_box(acc, out);
}

__global__ __launch_bounds__(kBlock) void k_morton(const float* hot, const mcpt_sphere* sph, int S, int n,
                                                   const float* sb, uint32_t* morton, uint32_t* keys, int* vals) {
    int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    vec3 c;
    if (i < S) {  // sphere::calculateCentroid (:858-860)
        const mcpt_sphere& sp = sph[i];
        c = sp.moving ? (load3(sp.center) + load3(sp.center1)) * 0.5f : load3(sp.center);
    } else {
        const float* t = hot + (size_t)(i - S) * 9;
        c = (vec3(t[0], t[1], t[2]) + vec3(t[3], t[4], t[5]) + vec3(t[6], t[7], t[8])) / 3.0f;
    }
    // quantize (:922-925): (v - min) / (max - min) * 1023, truncated
    uint32_t qx = (uint32_t)((c.x - sb[0]) / (sb[3] - sb[0]) * 1023.0f);
    uint32_t qy = (uint32_t)((c.y - sb[1]) / (sb[4] - sb[1]) * 1023.0f);
    uint32_t qz = (uint32_t)((c.z - sb[2]) / (sb[5] - sb[2]) * 1023.0f);
    uint32_t m = expand_bits(qx) | expand_bits(qy) << 1 | expand_bits(qz) << 2;
    morton[i] = m;
    keys[i] = m;
    vals[i] = i;
}

// ------------------------------------------------------------- leaves
// leaf records in Morton order: vertices plus the unit normal hit_triangle
// forms from them (same operations, so the same bits), or the sphere
// record; leaf_prim holds the triangle index, or ~ the sphere index
__global__ __launch_bounds__(kBlock) void k_leaves(const float* hot, const mcpt_sphere* sph, int S, const int* ids,
                                                   int n, LeafTri* leaf_tri, int* leaf_prim) {
    int j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= n) return;
    int p = ids[j];
    if (p < S) {
        leaf_prim[j] = ~p;
        leaf_tri[j] = sphere_leaf(sph[p]);
        return;
    }
    p -= S;
    leaf_prim[j] = p;
    const float* t = hot + (size_t)p * 9;
    const vec3 v0(t[0], t[1], t[2]), v1(t[3], t[4], t[5]), v2(t[6], t[7], t[8]);
    const vec3 nn = triangle_normal(v0, v1, v2);
    LeafTri L;
    L.a = make_float4(v0.x, v0.y, v0.z, v1.x);
    L.b = make_float4(v1.y, v1.z, v2.x, v2.y);
    L.c = make_float4(v2.z, nn.x, nn.y, nn.z);
    leaf_tri[j] = L;
}

// --------------------------------------------------------- Karras tree
__global__ __launch_bounds__(kBlock) void k_karras(const uint32_t* keys, int n, const LeafTri* leaf_tri, BvhNode* nodes,
                                                   int* leaf_parent) {
    int idx = blockIdx.x * kBlock + threadIdx.x;
    if (idx >= n - 1) return;
    // determineRange (:1349-1369)
    int d = (delta(keys, n, idx, idx + 1) - delta(keys, n, idx, idx - 1)) >= 0 ? 1 : -1;
    int dmin = delta(keys, n, idx, idx - d);
    int lmax = 2;
    while (delta(keys, n, idx, idx + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (delta(keys, n, idx, idx + (l + t) * d) > dmin) l += t;
    int j = idx + l * d;
    int first = min(idx, j), last = max(idx, j);
    // findSplit (:1371-1386)
    int cp = delta(keys, n, first, last);
    int split = first, step = last - first;
    do {
        step = (step + 1) >> 1;
        int ns = split + step;
        if (ns < last && delta(keys, n, first, ns) > cp) split = ns;
    } while (step > 1);

    int* rf = reinterpret_cast<int*>(nodes + idx) + 12;
    const int kids[2] = {split, split + 1};
    const bool leafs[2] = {split == first, split + 1 == last};
    for (int side = 0; side < 2; side++) {
        int c = kids[side];
        if (leafs[side]) {
            rf[side] = ~c;
            leaf_parent[c] = (idx << 1) | side;
            float b[6];
            leaf_box(leaf_tri[c], b);
            store_slot(nodes, idx, side, b);
        } else {
            rf[side] = c;
            reinterpret_cast<int*>(nodes + c)[14] = (idx << 1) | side;
        }
    }
    if (idx == 0) rf[2] = -1;
    rf[3] = j;  // the node covers leaves [min(idx, j), max(idx, j)]
}

// 5a. level of every internal node, each leaf's rank in the reference's
// leaf-test order and the deepest leaf (one atomic per wave); node ids
// beside the levels for the sort by level
__global__ __launch_bounds__(kBlock) void k_levels(const BvhNode* nodes, const int* leaf_parent, int n,
                                                   uint32_t* level, int* ids, int* tie, int* max_depth) {
    const int j = blockIdx.x * kBlock + threadIdx.x;
    int d = 0;
    if (j < n) {
        if (j < n - 1) ids[j] = j;
        const LeafClimb c = leaf_levels(nodes, leaf_parent[j], level);
        d = c.depth;
        tie[j] = c.rank;
    }
    for (int off = 32; off > 0; off >>= 1) d = max(d, __shfl_xor(d, off));
    if (lane_id() == 0 && d > 0) atomicMax(max_depth, d);
}

// 5b. refitBoundsKernel (:1398-1415) for the nodes of one level
__global__ __launch_bounds__(kBlock) void k_refit_level(BvhNode* nodes, const int* ids, int cnt, float* root_box) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < cnt) refit_node(nodes, ids[i], root_box);
}

// 6. the four-wide tree: one wide node per even-level binary node
__global__ __launch_bounds__(kBlock) void k_collapse(const BvhNode* nodes, const uint32_t* level, const int* widx,
                                                     int n, WideNode* wide) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n && !(level[i] & 1u)) collapse_node(nodes, i, widx, 0, wide);
}

__global__ void k_fill_int(int* p, int n, int v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

int blocks_for(int64_t n, int b = kBlock) { return (int)((n + b - 1) / b); }

void to_aabb(const float* b, mcpt_aabb* o) {
    for (int k = 0; k < 3; k++) {
        o->min[k] = b[k];
        o->max[k] = b[3 + k];
    }
    o->pad0 = o->pad1 = 0.f;
}

}  // namespace

int lbvh_build(mcpt_scene_s* s, mcpt_lbvh_dump* dump) {
    Ctx* ctx = s->ctx;
    hipStream_t st = ctx->stream;
    const int S = s->n_spheres;
    const int n = s->n + S;  // primitives: spheres, then triangles
    if (n < 1) return set_error(MCPT_ERR_INVALID, "empty scene");
    hipEvent_t e0, e1;
    MCPT_HIP(hipEventCreate(&e0));
    MCPT_HIP(hipEventCreate(&e1));
    MCPT_HIP(hipEventRecord(e0, st));

    const int n_int = std::max(1, n - 1);
    MCPT_TRY(s->nodes.alloc(ctx, n_int));
    MCPT_TRY(s->leaf_tri.alloc(ctx, n));
    MCPT_TRY(s->leaf_prim.alloc(ctx, n));
    MCPT_TRY(s->root_box.alloc(ctx, 6));
    MCPT_TRY(s->tie.alloc(ctx, n));
    MCPT_TRY(s->wide.alloc(ctx, wide_capacity(n)));

    const int tiles = blocks_for(n, kTile);
    const int n_partial = std::min(1024, blocks_for(n));
    DevBuf<float> partial, sb;
    DevBuf<uint32_t> morton, keys_a, keys_b;
    DevBuf<uint32_t> lvl_sorted;
    DevBuf<int> vals_a, vals_b, counts, ssum, leaf_parent, ids_sorted, depth;
    int status = MCPT_OK;
    auto fail = [&](int code) {
        partial.release(ctx);
        sb.release(ctx);
        morton.release(ctx);
        keys_a.release(ctx);
        keys_b.release(ctx);
        vals_a.release(ctx);
        vals_b.release(ctx);
        counts.release(ctx);
        ssum.release(ctx);
        leaf_parent.release(ctx);
        lvl_sorted.release(ctx);
        ids_sorted.release(ctx);
        depth.release(ctx);
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        return code;
    };
#define STEP(x)                       \
    do {                              \
        status = (x);                 \
        if (status < 0) return fail(status); \
    } while (0)
#define HSTEP(x)                                                                                  \
    do {                                                                                          \
        hipError_t he_ = (x);                                                                     \
        if (he_ != hipSuccess)                                                                    \
            return fail(set_error(MCPT_ERR_HIP, "lbvh %s: %s", #x, hipGetErrorString(he_)));      \
    } while (0)
    STEP(partial.alloc(ctx, (size_t)n_partial * 6));
    STEP(sb.alloc(ctx, 6));
    STEP(morton.alloc(ctx, n));
    STEP(keys_a.alloc(ctx, n));
    STEP(keys_b.alloc(ctx, n));
    STEP(vals_a.alloc(ctx, n));
    STEP(vals_b.alloc(ctx, n));
    STEP(counts.alloc(ctx, (size_t)256 * tiles));
    STEP(ssum.alloc(ctx, scan_chunks(256 * (int64_t)tiles)));
    STEP(leaf_parent.alloc(ctx, n));
    STEP(lvl_sorted.alloc(ctx, n_int));
    STEP(ids_sorted.alloc(ctx, n_int));
    STEP(depth.alloc(ctx, 66));  // max depth, then the 65 level offsets

    // 1-2. bounds and Morton codes
    k_bounds_partial<<<n_partial, kBlock, 0, st>>>(s->hot.ptr, s->spheres.ptr, S, n, partial.ptr);
    k_bounds_final<<<1, kBlock, 0, st>>>(partial.ptr, n_partial, sb.ptr);
    k_morton<<<blocks_for(n), kBlock, 0, st>>>(s->hot.ptr, s->spheres.ptr, S, n, sb.ptr, morton.ptr, keys_a.ptr,
                                               vals_a.ptr);
    HSTEP(hipGetLastError());
    // 3. stable radix sort, 8 bits per pass over the 30 code bits
    uint32_t* kin = keys_a.ptr;
    uint32_t* kout = keys_b.ptr;
    int* vin = vals_a.ptr;
    int* vout = vals_b.ptr;
    for (int shift = 0; shift < 30; shift += 8) {
        k_radix_hist<<<tiles, kBlock, 0, st>>>(kin, n, shift, tiles, counts.ptr);
        scan_exclusive(counts.ptr, 256 * tiles, ssum.ptr, st);
        k_radix_scatter<<<tiles, kBlock, 0, st>>>(kin, vin, kout, vout, n, shift, tiles, counts.ptr);
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    HSTEP(hipGetLastError());
    // 4. leaves and the Karras tree
    k_leaves<<<blocks_for(n), kBlock, 0, st>>>(s->hot.ptr, s->spheres.ptr, S, vin, n, s->leaf_tri.ptr,
                                               s->leaf_prim.ptr);
    k_fill_int<<<1, 64, 0, st>>>(depth.ptr, 1, 0);
    if (n == 1) {
        // One primitive: no internal node exists in the reference. The root
        // here holds leaf 0 on both sides, so traversal tests it once per
        // side and keeps the first (strict '<').
        HSTEP(hipStreamSynchronize(st));
        LeafTri L;
        float b[6];
        HSTEP(hipMemcpy(&L, s->leaf_tri.ptr, sizeof(L), hipMemcpyDeviceToHost));
        leaf_box(L, b);
        BvhNode root;
        std::memset(&root, 0, sizeof(root));
        float* f = reinterpret_cast<float*>(&root);
        for (int k = 0; k < 6; k++) f[k] = f[6 + k] = b[k];
        root.r = make_int4(~0, ~0, -1, 0);
        int lp = 0;
        HSTEP(hipMemcpy(s->nodes.ptr, &root, sizeof(root), hipMemcpyHostToDevice));
        HSTEP(hipMemcpy(s->root_box.ptr, b, sizeof(b), hipMemcpyHostToDevice));
        HSTEP(hipMemcpy(leaf_parent.ptr, &lp, sizeof(int), hipMemcpyHostToDevice));
        HSTEP(hipMemcpy(s->tie.ptr, &lp, sizeof(int), hipMemcpyHostToDevice));
        // the wide root likewise holds leaf 0 twice
        const float cb[2][6] = {{b[0], b[1], b[2], b[3], b[4], b[5]}, {b[0], b[1], b[2], b[3], b[4], b[5]}};
        const int refs[2] = {~0, ~0};
        WideNode wroot;
        wide_encode(cb, refs, 2, &wroot);
        HSTEP(hipMemcpy(s->wide.ptr, &wroot, sizeof(wroot), hipMemcpyHostToDevice));
        s->n_wide = 1;
        s->max_depth = 1;
    } else {
        k_karras<<<blocks_for(n - 1), kBlock, 0, st>>>(kin, n, s->leaf_tri.ptr, s->nodes.ptr, leaf_parent.ptr);
        // 5. refit: levels (keys into kout, ids into vout, both free after
        // the sort), one radix pass by level, then one launch per level
        const int ni = n - 1, itiles = blocks_for(ni, kTile);
        k_levels<<<blocks_for(n), kBlock, 0, st>>>(s->nodes.ptr, leaf_parent.ptr, n, kout, vout, s->tie.ptr,
                                                   depth.ptr);
        k_radix_hist<<<itiles, kBlock, 0, st>>>(kout, ni, 0, itiles, counts.ptr);
        scan_exclusive(counts.ptr, 256 * itiles, ssum.ptr, st);
        k_radix_scatter<<<itiles, kBlock, 0, st>>>(kout, vout, lvl_sorted.ptr, ids_sorted.ptr, ni, 0, itiles,
                                                   counts.ptr);
        k_level_offsets<<<1, 128, 0, st>>>(counts.ptr, itiles, ni, depth.ptr + 1);
        HSTEP(hipGetLastError());
        int hd[66];
        HSTEP(hipMemcpyAsync(hd, depth.ptr, sizeof(hd), hipMemcpyDeviceToHost, st));
        HSTEP(hipStreamSynchronize(st));
        s->max_depth = hd[0];
        if (s->max_depth < 1 || s->max_depth > 64)
            return fail(set_error(MCPT_ERR_INVALID, "BVH depth %d outside [1, 64]", s->max_depth));
        const int* off = hd + 1;
        for (int L = s->max_depth - 1; L >= 0; L--)
            if (off[L + 1] > off[L])
                k_refit_level<<<blocks_for(off[L + 1] - off[L]), kBlock, 0, st>>>(s->nodes.ptr,
                                                                                  ids_sorted.ptr + off[L],
                                                                                  off[L + 1] - off[L],
                                                                                  s->root_box.ptr);
        // 6. four-wide tree: wide indices of the even-level nodes in the
        // level sort's (level, index) order (ranks in lvl_sorted, free after
        // the offsets, indices into vout, free after the sort), then one wide
        // node per such node
        int n_wide = 0;
        for (int L = 0; L < 64; L += 2) n_wide += off[L + 1] - off[L];
        s->n_wide = n_wide;
        int* rank = reinterpret_cast<int*>(lvl_sorted.ptr);
        k_wide_flags<<<blocks_for(ni), kBlock, 0, st>>>(ids_sorted.ptr, kout, ni, rank);
        scan_exclusive(rank, ni, ssum.ptr, st);
        k_wide_scatter<<<blocks_for(ni), kBlock, 0, st>>>(ids_sorted.ptr, kout, rank, ni, vout);
        k_collapse<<<blocks_for(ni), kBlock, 0, st>>>(s->nodes.ptr, kout, vout, ni, s->wide.ptr);
        HSTEP(hipGetLastError());
    }
    HSTEP(hipEventRecord(e1, st));
    HSTEP(hipEventSynchronize(e1));
    float ms = 0.f;
    HSTEP(hipEventElapsedTime(&ms, e0, e1));
    s->last_build_ms = ms;
    if (s->max_depth < 1 || s->max_depth > 64)
        return fail(set_error(MCPT_ERR_INVALID, "BVH depth %d outside [1, 64]", s->max_depth));

    if (dump) {
        std::vector<BvhNode> nodes(n_int);
        std::vector<int> lpar(n);
        float root[6];
        HSTEP(hipMemcpy(nodes.data(), s->nodes.ptr, n_int * sizeof(BvhNode), hipMemcpyDeviceToHost));
        HSTEP(hipMemcpy(lpar.data(), leaf_parent.ptr, n * sizeof(int), hipMemcpyDeviceToHost));
        HSTEP(hipMemcpy(root, s->root_box.ptr, sizeof(root), hipMemcpyDeviceToHost));
        if (dump->morton) HSTEP(hipMemcpy(dump->morton, morton.ptr, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
        if (dump->sorted_keys) HSTEP(hipMemcpy(dump->sorted_keys, kin, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
        if (dump->sorted_ids) HSTEP(hipMemcpy(dump->sorted_ids, vin, n * sizeof(int), hipMemcpyDeviceToHost));
        if (dump->scene_bounds) {
            float b[6];
            HSTEP(hipMemcpy(b, sb.ptr, sizeof(b), hipMemcpyDeviceToHost));
            to_aabb(b, dump->scene_bounds);
        }
        auto slot = [&](int parent_code, float out[6]) {
            const float* f = reinterpret_cast<const float*>(&nodes[parent_code >> 1]);
            const int o = (parent_code & 1) ? 6 : 0;
            for (int k = 0; k < 6; k++) out[k] = f[o + k];
        };
        if (n > 1) {
            for (int i = 0; i < n - 1; i++) {
                const int4 r = nodes[i].r;
                if (dump->left) dump->left[i] = r.x >= 0 ? r.x : ~r.x;
                if (dump->right) dump->right[i] = r.y >= 0 ? r.y : ~r.y;
                if (dump->left_is_leaf) dump->left_is_leaf[i] = r.x < 0;
                if (dump->right_is_leaf) dump->right_is_leaf[i] = r.y < 0;
                if (dump->parent) dump->parent[i] = r.z < 0 ? -1 : (r.z >> 1);
                if (dump->internal_bounds) {
                    float b[6];
                    if (r.z < 0) std::memcpy(b, root, sizeof(b));
                    else slot(r.z, b);
                    to_aabb(b, &dump->internal_bounds[i]);
                }
            }
        }
        for (int j = 0; j < n; j++) {
            if (dump->leaf_parent) dump->leaf_parent[j] = lpar[j] >> 1;
            if (dump->leaf_bounds) {
                float b[6];
                slot(lpar[j], b);
                to_aabb(b, &dump->leaf_bounds[j]);
            }
        }
        dump->max_depth = s->max_depth;
        dump->build_ms = ms;
    }
    s->built = true;
    fail(MCPT_OK);
#undef STEP
#undef HSTEP
    return MCPT_OK;
}

}  // namespace mcpt
