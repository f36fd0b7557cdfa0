// lbvh_common.hpp — pieces of the LBVH build shared by the single-scene
// build (lbvh.hip) and the batched one (batch.hip): the padded triangle box,
// the Morton bit spread, the Karras prefix length, the node-slot store and
// the stable LSD radix sort (tile histogram, one-block exclusive scan,
// LDS-ranked scatter whose ranks come from wave64 ballots).
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

// Refit arrival at node counter c: returns the count before this thread's
// increment (0 = first child done, the caller stops; 1 = both children done,
// the caller merges). The explicit s_waitcnt makes the box this thread stored
// on its previous climb step complete before the release's L2 write-back:
// inside the climb loop the compiler places the write-back straight after the
// loop-carried stores without waiting for them, and a sibling on another XCD
// could then read the slot before it reaches memory.
__device__ __forceinline__ int refit_arrive(int* c) {
    __builtin_amdgcn_s_waitcnt(0);
    return __hip_atomic_fetch_add(c, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
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
