// lbvh_common.hpp — pieces of the LBVH build shared by the single-scene
// build (lbvh.hip) and the batched one (batch.hip): the padded triangle box,
// the Morton bit spread, the Karras prefix length, the node-slot store, the
// level-synchronous refit and the stable LSD radix sort (tile histogram,
// one-block exclusive scan, LDS-ranked scatter whose ranks come from wave64
// ballots).
#pragma once

#include "device.hpp"

namespace mcpt {
namespace {

constexpr int kBlock = 256;
constexpr int kItems = 16;  // keys per thread per radix tile
constexpr int kTile = kBlock * kItems;

__device__ __forceinline__ void tri_box(const float* t, float b[6]) {
    const float eps = 1e-3f;  // triangle::calculate_bounding_box (:651-655)
    b[0] = fminf(t[0], fminf(t[3], t[6])) - eps;
    b[1] = fminf(t[1], fminf(t[4], t[7])) - eps;
    b[2] = fminf(t[2], fminf(t[5], t[8])) - eps;
    b[3] = fmaxf(t[0], fmaxf(t[3], t[6])) + eps;
    b[4] = fmaxf(t[1], fmaxf(t[4], t[7])) + eps;
    b[5] = fmaxf(t[2], fmaxf(t[5], t[8])) + eps;
}

__device__ __forceinline__ uint32_t expand_bits(uint32_t p) {  // expandBits (:913-920)
    uint32_t e = 0;
    for (int i = 0; i < 10; i++) e |= ((p >> i) & 1u) << (3 * i);
    return e;
}

__device__ __forceinline__ int delta(const uint32_t* k, int n, int i, int j) {  // :1337-1347
    if (j < 0 || j >= n) return -1;
    uint32_t a = k[i], b = k[j];
    if (a == b) return 32 + __clz((uint32_t)(i ^ j));
    return __clz(a ^ b);
}

__device__ __forceinline__ void store_slot(BvhNode* nodes, int node, int side, const float b[6]) {
    float* f = reinterpret_cast<float*>(nodes + node);
    const int o = side ? 6 : 0;
    for (int k = 0; k < 6; k++) f[o + k] = b[k];
}

// parent code of internal node p: (parent << 1) | side, -1 at the root
__device__ __forceinline__ int parent_word(const BvhNode* nodes, int p) {
    return reinterpret_cast<const int*>(nodes + p)[14];
}

// ---------------------------------------------------------------- refit
// The refit (refitBoundsKernel, src/path_tracer.cu:1398-1415) runs level by
// level instead of with per-node arrival counters: every internal node gets
// its level (root 0), the nodes are sorted by level, and one launch per level,
// deepest first, merges each node's two child boxes and stores the result in
// its parent's slot. Kernel boundaries order the levels, so no atomics and no
// agent-scope cache write-back/invalidate (which on a multi-XCD part costs a
// whole-L2 operation per climbing step) are needed, and the result is the
// same whatever order the threads run in.

// Levels along the climb from one leaf, given its parent code. The leaf
// first counts its internal ancestors (d), then climbs again while it is a
// left child, writing the level of each node whose leaf range it starts:
// every internal node has exactly one such leaf, so it is written once.
//
// The same climb gives the leaf's rank in the reference's leaf-test order
// (leaves are tested when their parent is visited, parents in left-first
// pre-order, left side first): rank = 2 * preorder(P) + side, where the
// pre-order number of internal node P is depth(P) + first(P) - m(P), with
// first(P) the first leaf of P's range and m(P) the number of right-child
// edges on the path from the root to P (each such edge skips a left subtree
// of k leaves and k - 1 internal nodes).
struct LeafClimb {
    int depth;   // internal ancestors of the leaf
    int rank;    // position in the reference's leaf-test order
    int n_even;  // even-level nodes this leaf wrote the level of
};

__device__ __forceinline__ LeafClimb leaf_levels(const BvhNode* nodes, int code, uint32_t* level) {
    LeafClimb r;
    int d = 1, m = 0;
    for (int p = code >> 1, pp; (pp = parent_word(nodes, p)) >= 0; p = pp >> 1) {
        d++;
        m += pp & 1;
    }
    const int P = code >> 1;
    const int first = min(P, reinterpret_cast<const int*>(nodes + P)[15]);
    r.depth = d;
    r.rank = 2 * (d - 1 + first - m) + (code & 1);
    r.n_even = 0;
    for (int c = code, lvl = d - 1; c >= 0 && !(c & 1); lvl--) {
        const int q = c >> 1;
        level[q] = (uint32_t)lvl;
        r.n_even += !(lvl & 1);
        c = parent_word(nodes, q);
    }
    return r;
}

// merge the two child boxes held in node p into the node's own box: its slot
// in the parent, or the root box
__device__ __forceinline__ void refit_node(BvhNode* nodes, int p, float* root_box) {
    const float* f = reinterpret_cast<const float*>(nodes + p);
    float b[6];
    for (int k = 0; k < 3; k++) b[k] = fminf(f[k], f[6 + k]);
    for (int k = 3; k < 6; k++) b[k] = fmaxf(f[k], f[6 + k]);
    const int pp = parent_word(nodes, p);
    if (pp < 0) {
        for (int k = 0; k < 6; k++) root_box[k] = b[k];
        return;
    }
    store_slot(nodes, pp >> 1, pp & 1, b);
}

// ------------------------------------------------------- four-wide tree
// Every binary node at an even level becomes a wide node whose children are
// its children that are leaves and the children of its inner children, in
// left-to-right order; wide index = exclusive prefix count of even-level
// nodes in node order, so the root stays 0. A tree of n leaves has at most
// (2n - 1) / 3 even-level nodes: with E even- and O odd-level internal nodes,
// every odd node has an even parent and every even node but the root an odd
// one, so E <= 2 O + 1 with E + O = n - 1.
inline int wide_capacity(int n) { return n < 2 ? 1 : (int)((2 * (int64_t)n + 1) / 3); }

MCPT_HDI float exp2i(int e) {  // 2^e for e in [-126, 127]
    union {
        uint32_t u;
        float f;
    } v;
    v.u = (uint32_t)(e + 127) << 23;
    return v.f;
}

// Quantizes up to four child boxes (min xyz, max xyz) outward against
// their union: per axis the scale is the smallest power of two (>= 2^-126)
// that spans the union in 255 steps, lo = floor and hi = ceil, each then
// checked by the decode the traversal bounds (origin + q * 2^e, one
// rounding of an exact product) and stepped outward until it contains the
// exact bound.
MCPT_HDI void wide_encode(const float (*b)[6], const int* ref, int cnt, WideNode* out) {
    float org[3];
    uint32_t ebits = 0;
    uint32_t qlo[3] = {0, 0, 0}, qhi[3] = {0, 0, 0};
    for (int a = 0; a < 3; a++) {
        float lo = b[0][a], hi = b[0][3 + a];
        for (int k = 1; k < cnt; k++) {
            lo = fminf(lo, b[k][a]);
            hi = fmaxf(hi, b[k][3 + a]);
        }
        org[a] = lo;
        const double ext = (double)hi - (double)lo;
        int e = -126;
        if (ext > 0.0) {
            int ex = 0;
            frexp(ext / 255.0, &ex);
            e = ex < -126 ? -126 : (ex > 127 ? 127 : ex);
        }
        while (true) {
            const float sc = exp2i(e);
            bool ok = true;
            uint32_t ql = 0, qh = 0;
            for (int k = 0; k < cnt && ok; k++) {
                double fl = floor(((double)b[k][a] - (double)lo) / (double)sc);
                double fh = ceil(((double)b[k][3 + a] - (double)lo) / (double)sc);
                int l = fl < 0.0 ? 0 : (fl > 255.0 ? 255 : (int)fl);
                int h = fh < 0.0 ? 0 : (fh > 255.0 ? 255 : (int)fh);
                while (l > 0 && lo + (float)l * sc > b[k][a]) l--;
                while (h < 255 && lo + (float)h * sc < b[k][3 + a]) h++;
                if (lo + (float)h * sc < b[k][3 + a]) ok = false;
                ql |= (uint32_t)l << (8 * k);
                qh |= (uint32_t)h << (8 * k);
            }
            if (ok || e >= 127) {
                qlo[a] = ql;
                qhi[a] = qh;
                break;
            }
            e++;
        }
        ebits |= (uint32_t)(e + 127) << (8 * a);
    }
    ebits |= (uint32_t)cnt << 24;
    union {
        uint32_t u;
        float f;
    } w;
    w.u = ebits;
    out->o = make_float4(org[0], org[1], org[2], w.f);
    out->q0 = make_uint4(qlo[0], qlo[1], qlo[2], qhi[0]);
    out->q1 = make_uint4(qhi[1], qhi[2], 0u, 0u);
    out->child = make_int4(ref[0], cnt > 1 ? ref[1] : 0, cnt > 2 ? ref[2] : 0, cnt > 3 ? ref[3] : 0);
}

// Wide node numbering. Wide node w is the w-th even-level binary node in
// (level, binary index) order. In a Karras tree every internal node inside
// a node's key range is its descendant or ancestor (ranges nest), so the
// binary nodes two levels below node p whose indices fall in p's range are
// exactly p's grandchildren: in (level, index) order they are consecutive,
// in slot order. Each wide node's inner children therefore have consecutive
// wide indices, and the walk finds a pending child as base + rank instead
// of re-reading the parent's references.
//
// flag[i] = 1 when node ids[i] (of an array in that order) is at an even
// level; scanned in place, it becomes the wide index
__global__ __launch_bounds__(kBlock) void k_wide_flags(const int* ids, const uint32_t* level, int n, int* flag) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) flag[i] = (level[ids[i]] & 1u) ? 0 : 1;
}
// widx[ids[i]] = rank[i] for the even-level nodes
__global__ __launch_bounds__(kBlock) void k_wide_scatter(const int* ids, const uint32_t* level, const int* rank,
                                                         int n, int* widx) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n && !(level[ids[i]] & 1u)) widx[ids[i]] = rank[i];
}

// The wide node of binary node p (at an even level): child boxes come from
// p's slots and its inner children's slots; wide references are
// widx[gbase + node] - widx[gbase] (the scan restarted at the tree's root)
// for inner grandchildren, ~leaf for leaves.
__device__ __forceinline__ void collapse_node(const BvhNode* nodes, int p, const int* widx, int gbase,
                                              WideNode* wide) {
    const int wbase = widx[gbase];
    float b[4][6];
    int ref[4];
    int cnt = 0;
    const float* fp = reinterpret_cast<const float*>(nodes + p);
    const int* rp = reinterpret_cast<const int*>(nodes + p) + 12;
    for (int side = 0; side < 2; side++) {
        const int c = rp[side];
        if (c < 0) {
            for (int k = 0; k < 6; k++) b[cnt][k] = fp[6 * side + k];
            ref[cnt++] = c;
        } else {
            const float* fc = reinterpret_cast<const float*>(nodes + c);
            const int* rc = reinterpret_cast<const int*>(nodes + c) + 12;
            for (int s2 = 0; s2 < 2; s2++) {
                const int g = rc[s2];
                for (int k = 0; k < 6; k++) b[cnt][k] = fc[6 * s2 + k];
                ref[cnt++] = g < 0 ? g : widx[gbase + g] - wbase;
            }
        }
    }
    wide_encode(b, ref, cnt, wide + (widx[gbase + p] - wbase));
}

// First sorted position of each level 0..63 after one radix pass over the
// level keys (the exclusive scan of the digit-major tile histograms), and n
// at index 64. Levels stay below 64: a Karras tree over 30-bit codes with
// the index tie-break is at most 62 deep, and the host checks the depth.
__global__ void k_level_offsets(const int* counts, int tiles, int n, int* off) {
    const int L = threadIdx.x;
    if (L < 64) off[L] = counts[L * tiles];
    else if (L == 64) off[L] = n;
}

// ------------------------------------------------------------ radix sort
__global__ __launch_bounds__(kBlock) void k_radix_hist(const uint32_t* keys, int n, int shift, int tiles,
                                                       int* counts) {
    __shared__ int hist[256];
    hist[threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * kTile;
    for (int r = 0; r < kItems; r++) {
        int i = base + r * kBlock + threadIdx.x;
        if (i < n) atomicAdd(&hist[(keys[i] >> shift) & 255u], 1);
    }
    __syncthreads();
    counts[threadIdx.x * tiles + blockIdx.x] = hist[threadIdx.x];
}

// exclusive scan of `total` ints in place, one block of 1024 threads
__global__ __launch_bounds__(1024) void k_scan_exclusive(int* data, int total) {
    __shared__ int sums[1024];
    const int per = (total + 1023) / 1024;
    const int b = threadIdx.x * per;
    const int e = min(total, b + per);
    int s = 0;
    for (int i = b; i < e; i++) s += data[i];
    sums[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {  // Hillis-Steele inclusive scan
        int v = (int)threadIdx.x >= off ? sums[threadIdx.x - off] : 0;
        __syncthreads();
        sums[threadIdx.x] += v;
        __syncthreads();
    }
    int run = sums[threadIdx.x] - s;
    for (int i = b; i < e; i++) {
        int v = data[i];
        data[i] = run;
        run += v;
    }
}

// Multi-block exclusive scan of the tile histograms: each block of kBlock
// threads owns kTile consecutive ints. Pass 1 sums every chunk (coalesced
// loads), k_scan_exclusive scans the chunk sums in one block, pass 2 stages
// its chunk in LDS (padded one int per 16 so the per-thread runs of 16 are
// conflict-free), scans it and adds the chunk's offset.
__device__ __forceinline__ int scan_pad(int i) { return i + (i >> 4); }

__global__ __launch_bounds__(kBlock) void k_scan_chunk_sums(const int* data, int total, int* sums) {
    __shared__ int wsum[kBlock / 64];
    const int base = blockIdx.x * kTile;
    int s = 0;
    for (int r = 0; r < kItems; r++) {
        const int i = base + r * kBlock + (int)threadIdx.x;
        if (i < total) s += data[i];
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (lane_id() == 0) wsum[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < kBlock / 64; w++) t += wsum[w];
        sums[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(kBlock) void k_scan_chunk_apply(int* data, int total, const int* sums) {
    __shared__ int buf[kTile + kTile / 16];
    __shared__ int wsum[kBlock / 64];
    const int base = blockIdx.x * kTile, tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
    for (int r = 0; r < kItems; r++) {
        const int i = r * kBlock + tid;
        buf[scan_pad(i)] = base + i < total ? data[base + i] : 0;
    }
    __syncthreads();
    int v[kItems], s = 0;
    for (int k = 0; k < kItems; k++) {
        v[k] = buf[scan_pad(tid * kItems + k)];
        s += v[k];
    }
    // exclusive scan of the per-thread sums: within the wave, then across waves
    int inc = s;
    for (int off = 1; off < 64; off <<= 1) {
        const int u = __shfl_up(inc, off);
        if (lane >= off) inc += u;
    }
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    int run = sums[blockIdx.x] + inc - s;
    for (int w = 0; w < wave; w++) run += wsum[w];
    for (int k = 0; k < kItems; k++) {
        buf[scan_pad(tid * kItems + k)] = run;
        run += v[k];
    }
    __syncthreads();
    for (int r = 0; r < kItems; r++) {
        const int i = r * kBlock + tid;
        if (base + i < total) data[base + i] = buf[scan_pad(i)];
    }
}

// chunk-sum slots scan_exclusive needs for `total` ints
inline int scan_chunks(int64_t total) { return (int)((total + kTile - 1) / kTile); }

// exclusive scan of data[0, total) in place; sums holds scan_chunks(total) ints
inline void scan_exclusive(int* data, int total, int* sums, hipStream_t st) {
    const int nb = scan_chunks(total);
    k_scan_chunk_sums<<<nb, kBlock, 0, st>>>(data, total, sums);
    k_scan_exclusive<<<1, 1024, 0, st>>>(sums, nb);
    k_scan_chunk_apply<<<nb, kBlock, 0, st>>>(data, total, sums);
}

__global__ __launch_bounds__(kBlock) void k_radix_scatter(const uint32_t* kin, const int* vin, uint32_t* kout,
                                                          int* vout, int n, int shift, int tiles, const int* offsets) {
    __shared__ int running[256];
    __shared__ int wave_cnt[kBlock / 64][256];
    const int tid = threadIdx.x, wave = tid >> 6, lane = lane_id();
    running[tid] = offsets[tid * tiles + blockIdx.x];
    for (int w = 0; w < kBlock / 64; w++) wave_cnt[w][tid] = 0;
    __syncthreads();
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int base = blockIdx.x * kTile;
    for (int r = 0; r < kItems; r++) {
        int i = base + r * kBlock + tid;
        bool valid = i < n;
        uint32_t key = valid ? kin[i] : 0u;
        int val = valid ? vin[i] : 0;
        uint32_t digit = (key >> shift) & 255u;
        // lanes of this wave holding the same digit
        uint64_t peers = ballot64(valid);
        for (int bit = 0; bit < 8; bit++) {
            uint64_t b = ballot64((digit >> bit) & 1u);
            peers &= ((digit >> bit) & 1u) ? b : ~b;
        }
        int wave_rank = __popcll(peers & below);
        if (valid && wave_rank == 0) wave_cnt[wave][digit] = __popcll(peers);
        __syncthreads();
        if (valid) {
            int pre = running[digit] + wave_rank;
            for (int w = 0; w < wave; w++) pre += wave_cnt[w][digit];
            kout[pre] = key;
            vout[pre] = val;
        }
        __syncthreads();
        int add = 0;
        for (int w = 0; w < kBlock / 64; w++) {
            add += wave_cnt[w][tid];
            wave_cnt[w][tid] = 0;
        }
        running[tid] += add;
        __syncthreads();
    }
}

}  // namespace
}  // namespace mcpt
