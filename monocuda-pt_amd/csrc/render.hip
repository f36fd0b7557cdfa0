// render.hip — the integrators on gfx950.
//
//   megakernel   one lane per pixel path (trace_sample_kernel,
//                src/path_tracer.cu:2900-3059)
//   wavefront    camera -> [extend -> shade -> shadow] x depth -> finalize
//                (renderWithWavefront, :3757-3840), with
//                - ray queues of 32-byte records (origin+time, direction+path)
//                  so extend reads each ray with two coalesced 16-byte loads,
//                - the 16-byte PackedHit written per queue slot,
//                - queue pushes compacted by wave64 ballots: one atomic per
//                  wave and per queue (:3573, :3625, :3691 take one per lane),
//                - queue lengths kept on the device: every kernel reads its
//                  count from HBM, so a whole sample is enqueued without the
//                  per-bounce host round trip of :3816.
//   persistent   the extend kernel runs a fixed grid; a wave refills its idle
//                lanes from a global work counter once `refill` lanes are
//                idle (extendKernelPersistent, :3416-3468, on 64-lane waves).
#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "scene.hpp"
#include "shade.hpp"

namespace mcpt {

namespace {

constexpr int kBlock = 256;

// counters[] slots, one 128-byte line each: every wave of a launch takes
// slots with atomics on them, and atomics on one line serialize in one L2
// channel. C_NS0 / C_NS1 are u64 per ray queue: the queue's length (low 32
// bits) and the length of the shadow queue filled in the same bounce (high
// 32 bits), so shade takes both slots with one atomic. The *_TOTAL, STEPS
// and LEAVES slots hold u64.
enum {
    C_NS0 = 0,
    C_NS1 = 32,
    C_WORK = 64,
    C_WORK_SHADOW = 96,
    C_EXT_TOTAL = 128,
    C_SHD_TOTAL = 160,
    C_STEPS = 192,
    C_LEAVES = 224,
    C_NCOUNTERS = 256
};

// Path state travels with its ray: thr/smp/rng[q] hold the state of the
// path in slot i of ray queue q, and shade writes a surviving path's state
// beside its next ray, so every state access of a wave stays inside the
// slots the wave (or its block) works on. A path that ends leaves its sample
// in done[path]; a shadow ray adds its light to the sample where the path
// is now (next-queue slot, or done[path]; ShadowItem.dp.w).
struct Workspace {
    RayItem* q[2];
    mcpt_packed_hit* hits;
    ShadowItem* sq;
    float4* thr[2];  // throughput rgb, previous bsdf pdf
    float4* smp[2];  // sample rgb, previous-bounce-specular flag
    uint64_t* rng[2];
    float4* done;    // by path: the finished sample
    int* counters;   // int slots; the *_TOTAL and count slots at 4.. are u64
};

// where a shadow ray's light goes: the path's next-queue slot (>= 0) or,
// for a path that ended, ~path into done[]
__device__ __forceinline__ float4* shadow_target(const Workspace& ws, int cur, int target) {
    return target >= 0 ? ws.smp[1 - cur] + target : ws.done + ~target;
}

__device__ __forceinline__ uint64_t* u64slot(int* c, int i) { return reinterpret_cast<uint64_t*>(c + i); }
__device__ __forceinline__ unsigned long long* ns_slot(int* c, int q) {
    return reinterpret_cast<unsigned long long*>(c + (q ? C_NS1 : C_NS0));
}
__device__ __forceinline__ int queue_len(const int* c, int q) { return c[q ? C_NS1 : C_NS0]; }
__device__ __forceinline__ int shadow_len(const int* c, int q) { return c[(q ? C_NS1 : C_NS0) + 1]; }

// ------------------------------------------------------------ counters
__global__ void k_counters_init(int* c, int cur, int count) {
    if (threadIdx.x == 0) {
        *ns_slot(c, cur) = (unsigned long long)(uint32_t)count;
        *ns_slot(c, 1 - cur) = 0;
        c[C_WORK] = 0;
        c[C_WORK_SHADOW] = 0;
    }
}
// start of a bounce: record the previous bounce's queue lengths, clear the
// queues this bounce fills (the shadow counters belong to the shadow launch,
// which may still run from the previous bounce)
__global__ void k_counters_bounce(int* c, int cur) {
    if (threadIdx.x == 0) {
        *ns_slot(c, 1 - cur) = 0;
        c[C_WORK] = 0;
        *u64slot(c, C_EXT_TOTAL) += (uint64_t)queue_len(c, cur);
    }
}
// before a shadow launch: its work counter, and the shadow-ray total
__global__ void k_counters_shadow_total(int* c, int cur) {
    if (threadIdx.x == 0) {
        c[C_WORK_SHADOW] = 0;
        *u64slot(c, C_SHD_TOTAL) += (uint64_t)shadow_len(c, 1 - cur);
    }
}
__global__ void k_counters_zero_totals(int* c) {
    if (threadIdx.x < C_NCOUNTERS) c[threadIdx.x] = 0;
}

// ---------------------------------------------------------------- camera
// generateCameraRaysKernel (:3342-3374). A pass carries `batch` samples:
// path = k * W*H + pixel renders sample `sample0 + k` of `pixel`.
__global__ __launch_bounds__(kBlock) void k_camera(mcpt_camera cam, int sample0, int batch, uint64_t seed,
                                                   Workspace ws, int cur) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    const int W = cam.width, npix = cam.width * cam.height;
    if (p >= npix * batch) return;
    const int pixel = p % npix, k = p / npix;
    Rng rng;
    mcpt_pcg32_for_path(&rng.g, (uint32_t)pixel, (uint32_t)(sample0 + k), seed);
    vec3 o, d;
    float time;
    camera_ray(cam, pixel % W, pixel / W, rng, o, d, time);
    RayItem it;
    it.ot = make_float4(o.x, o.y, o.z, time);
    it.dp = make_float4(d.x, d.y, d.z, __int_as_float(p));
    ws.q[cur][p] = it;
    ws.thr[cur][p] = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
    ws.smp[cur][p] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    ws.rng[cur][p] = rng.g.state;
}

// ---------------------------------------------------------------- extend
template <bool COUNT>
__device__ __forceinline__ void count_flush(int* counters, uint32_t ns, uint32_t nl) {
    if (!COUNT) return;
    // wave-reduce, one atomic per wave
    uint64_t s = ns, l = nl;
    for (int off = 32; off > 0; off >>= 1) {
        s += __shfl_xor(s, off);
        l += __shfl_xor(l, off);
    }
    if (lane_id() == 0) {
        atomicAdd(reinterpret_cast<unsigned long long*>(u64slot(counters, C_STEPS)), (unsigned long long)s);
        atomicAdd(reinterpret_cast<unsigned long long*>(u64slot(counters, C_LEAVES)), (unsigned long long)l);
    }
}

__device__ __forceinline__ mcpt_packed_hit pack_hit(const BvhView& bv, const Trav& s) {
    mcpt_packed_hit h;
    h.t = s.t;
    if (s.leaf >= 0) {
        const int p = bv.leaf_prim[s.leaf];  // triangle, or ~sphere
        h.prim_type = p >= 0 ? 1 : 0;
        h.prim_index = p >= 0 ? p : ~p;
        h.did_hit = 1;
    } else {
        h.prim_type = -1;
        h.prim_index = -1;
        h.did_hit = 0;
    }
    return h;
}

// extendKernel (:3376-3397): one lane per queued ray
template <bool COUNT, bool REF>
__global__ __launch_bounds__(kBlock) void k_extend(SceneView sv, Workspace ws, int cur) {
    extern __shared__ int lds_stack[];
    const int slot = blockIdx.x * kBlock + threadIdx.x;
    const int count = queue_len(ws.counters, cur);
    uint32_t ns = 0, nl = 0;
    if (slot < count) {
        const RayItem it = ws.q[cur][slot];
        Trav s;
        trav_begin<REF>(s, vec3(it.ot.x, it.ot.y, it.ot.z), vec3(it.dp.x, it.dp.y, it.dp.z), 1e30f);
        s.time = it.ot.w;
        trace_closest<COUNT, REF>(sv.bvh, s, lds_stack + threadIdx.x, kBlock, ns, nl);
        ws.hits[slot] = pack_hit(sv.bvh, s);
    }
    count_flush<COUNT>(ws.counters, ns, nl);
}

// A queue shorter than group_below: kGroup lanes per ray (group_closest),
// each block's groups with their stacks side by side in LDS.
__device__ __forceinline__ void extend_group(const SceneView& sv, const Workspace& ws, int cur, int total, int* lds,
                                             int cap, int cap_phys) {
    if ((int)blockIdx.x * (kBlock / kGroup) >= total) return;  // blocks the queue does not need
    int* base = lds + (threadIdx.x / kGroup) * group_lds_ints(cap_phys);
    const RayItem* q = ws.q[cur];
    BvhView bv = sv.bvh;
    group_closest(
        bv, &ws.counters[C_WORK], total, reinterpret_cast<int2*>(base), base + 2 * cap_phys, cap, cap_phys,
        [&](int i, Trav& s, const BvhView&) {
            const RayItem it = q[i];
            trav_begin(s, vec3(it.ot.x, it.ot.y, it.ot.z), vec3(it.dp.x, it.dp.y, it.dp.z), 1e30f);
            s.time = it.ot.w;
        },
        [&](int i, const Trav& s, const BvhView&) { ws.hits[i] = pack_hit(sv.bvh, s); });
}

// A short queue's launch of the group walk (the persistent kernel below
// returns at once for it, and this one for a long queue). Kept a launch of
// its own: in the same kernel the group walk's registers (75 VGPRs against
// 59) cost the other lane's kernels room beside every persistent launch
// (64 spp on the 1M-triangle mesh: 1,437 against 1,465 Mrays/s).
__global__ __launch_bounds__(kBlock) void k_extend_group(SceneView sv, Workspace ws, int cur, int group_below,
                                                          int cap, int cap_phys) {
    extern __shared__ int lds[];
    const int total = queue_len(ws.counters, cur);
    if (total >= group_below) return;
    extend_group(sv, ws, cur, total, lds, cap, cap_phys);
}

// extendKernelPersistent (:3416-3468) on the shared persistent loop
template <bool COUNT, bool REF>
__global__ __launch_bounds__(kBlock) void k_extend_persistent(SceneView sv, Workspace ws, int cur, int refill,
                                                               int group_below) {
    extern __shared__ int lds_stack[];
    const int total = queue_len(ws.counters, cur);
    if (total < group_below) return;                 // k_extend_group walks a short queue
    if ((int)blockIdx.x * kBlock >= total) return;  // blocks the queue does not need
    uint32_t ns = 0, nl = 0;
    const RayItem* q = ws.q[cur];
    persistent_closest<COUNT, REF>(
        sv.bvh, &ws.counters[C_WORK], total, refill, lds_stack + threadIdx.x, kBlock,
        [&](int i, Trav& s, const BvhView&) {
            const RayItem it = q[i];
            trav_begin<REF>(s, vec3(it.ot.x, it.ot.y, it.ot.z), vec3(it.dp.x, it.dp.y, it.dp.z), 1e30f);
            s.time = it.ot.w;
        },
        [&](int i, const Trav& s, const BvhView&) { ws.hits[i] = pack_hit(sv.bvh, s); }, ns, nl);
    count_flush<COUNT>(ws.counters, ns, nl);
}

// ----------------------------------------------------------------- shade
// shadeKernel (:3470-3696)
// Grid-stride over chunks of kShadeBlock queue slots (a grid sized to the
// occupancy limit, so the thin late bounces launch no empty waves), each
// chunk ordered by shading bin before it is shaded. 256-thread blocks: a
// 1024-thread block waits for a whole CU to drain beside the other lane's
// persistent launches, where small blocks fill the gaps (1024 -> 256 cost
// 8% in a serial launch and gained 2-6% in the overlapped render).
constexpr int kShadeBlock = 256;
// Lanes of an untimed wavefront render (enqueue_render), and whether each
// lane's shadow launches run on a stream of their own. Two lanes with their
// shadow streams use the four hardware queues a process gets; measured on
// the 1 M-triangle mesh at 8 / 20 / 32 / 64 spp: 1,001 / 1,273 / 1,370 /
// 1,472 Mrays/s, against 768 / 1,077 / 1,188 / 1,334 without the shadow
// streams and at most 824 / 1,150 / 1,286 / 1,417 with three or four lanes
// (with or without them), whose streams share hardware queues.
constexpr int kRenderLanes = 2;
constexpr bool kLaneSideStreams = true;
__global__ __launch_bounds__(kShadeBlock) void k_shade(SceneView sv, Workspace ws, int cur, int depth, int max_depth,
                                                       vec3 bg, float rr_cap, int npix) {
    __shared__ int lds[bin_sort_lds_ints<kShadeBlock>()];
    const int count = queue_len(ws.counters, cur);
    for (int base = blockIdx.x * kShadeBlock; base < count; base += gridDim.x * kShadeBlock) {
        int slot = base + (int)threadIdx.x, mat = -1, bin = kShadeBins;
        mcpt_packed_hit hit;
        if (slot < count) {
            hit = ws.hits[slot];
            bin = shade_bin(sv, hit.did_hit != 0, hit.prim_type == 0 ? ~hit.prim_index : hit.prim_index, mat);
        }
        block_bin_sort<kShadeBlock>(bin, slot, mat, lds);
        const bool active = slot >= 0;
        Bounce out;
        out.next = false;
        out.shadow = false;
        int path = 0;
        float time = 0.f;
        PathState ps;
        Rng rng;
        if (active) {
            const RayItem it = ws.q[cur][slot];
            hit = ws.hits[slot];
            path = __float_as_int(it.dp.w);
            time = it.ot.w;
            const vec3 ro(it.ot.x, it.ot.y, it.ot.z), rd(it.dp.x, it.dp.y, it.dp.z);
            const float4 thr = ws.thr[cur][slot], smp = ws.smp[cur][slot];
            ps.thr = vec3(thr.x, thr.y, thr.z);
            ps.prev_pdf = thr.w;
            ps.smp = vec3(smp.x, smp.y, smp.z);
            ps.prev_spec = smp.w != 0.0f;
            rng.g.state = ws.rng[cur][slot];
            rng.g.inc = ((uint64_t)(uint32_t)(path % npix) << 1u) | 1u;  // the stream of the path's pixel
            const int prim = hit.prim_type == 0 ? ~hit.prim_index : hit.prim_index;
            shade_bounce(sv, bg, rr_cap, depth, ro, rd, hit.did_hit != 0, hit.t, prim, ps, rng, out, time, mat);
            // the pass ends after bounce max_depth - 1: nothing continues past it
            if (depth + 1 >= max_depth) out.next = false;
        }
        // compacted pushes: every lane of the wave takes part
        int nslot, sslot;
        wave_push2(ns_slot(ws.counters, 1 - cur), out.next, out.shadow, nslot, sslot);
        if (active) {
            const float4 smp_out = make_float4(ps.smp.x, ps.smp.y, ps.smp.z, ps.prev_spec ? 1.0f : 0.0f);
            if (out.next) {  // the state goes beside the path's next ray
                RayItem it;
                it.ot = make_float4(out.no.x, out.no.y, out.no.z, time);
                it.dp = make_float4(out.nd.x, out.nd.y, out.nd.z, __int_as_float(path));
                ws.q[1 - cur][nslot] = it;
                ws.thr[1 - cur][nslot] = make_float4(ps.thr.x, ps.thr.y, ps.thr.z, ps.prev_pdf);
                ws.smp[1 - cur][nslot] = smp_out;
                ws.rng[1 - cur][nslot] = rng.g.state;
            } else {
                ws.done[path] = smp_out;
            }
        }
        if (out.shadow) {
            ShadowItem si;
            si.om = make_float4(out.so.x, out.so.y, out.so.z, out.smax);
            si.dp = make_float4(out.sd.x, out.sd.y, out.sd.z, __int_as_float(out.next ? nslot : ~path));
            si.c = make_float4(out.contrib.x, out.contrib.y, out.contrib.z, time);
            ws.sq[sslot] = si;
        }
    }
}

// ---------------------------------------------------------------- shadow
// the light of an unblocked shadow ray into its path's sample
__device__ __forceinline__ void shadow_deliver(const Workspace& ws, int cur, const ShadowItem& si) {
    float4* dst = shadow_target(ws, cur, __float_as_int(si.dp.w));
    float4 v = *dst;
    v.x = v.x + si.c.x;
    v.y = v.y + si.c.y;
    v.z = v.z + si.c.z;
    *dst = v;
}

// shadowKernel (:3698-3728)
__global__ __launch_bounds__(kBlock) void k_shadow(SceneView sv, Workspace ws, int cur) {
    extern __shared__ int lds_stack[];
    const int slot = blockIdx.x * kBlock + threadIdx.x;
    if (slot >= shadow_len(ws.counters, 1 - cur)) return;
    const ShadowItem si = ws.sq[slot];
    const float tr = shadow_transmittance(sv, vec3(si.om.x, si.om.y, si.om.z), vec3(si.dp.x, si.dp.y, si.dp.z),
                                          si.om.w, lds_stack + threadIdx.x, kBlock, si.c.w);
    if (tr > 0.0f) {
        float4* dst = shadow_target(ws, cur, __float_as_int(si.dp.w));
        float4 s = *dst;
        s.x = s.x + si.c.x * tr;
        s.y = s.y + si.c.y * tr;
        s.z = s.z + si.c.z * tr;
        *dst = s;
    }
}

// A shadow queue shorter than group_below: kGroup lanes per ray (group_any).
__device__ __forceinline__ void shadow_group(const SceneView& sv, const Workspace& ws, int cur, int total, int* lds,
                                             int cap, int cap_phys) {
    if ((int)blockIdx.x * (kBlock / kGroup) >= total) return;  // blocks the queue does not need
    int* base = lds + (threadIdx.x / kGroup) * group_lds_ints(cap_phys);
    BvhView bv = sv.bvh;
    group_any(
        bv, &ws.counters[C_WORK_SHADOW], total, reinterpret_cast<int2*>(base), base + 2 * cap_phys, cap, cap_phys,
        [&](int i, Trav& s, const BvhView&) {
            const ShadowItem si = ws.sq[i];
            trav_begin(s, vec3(si.om.x, si.om.y, si.om.z), vec3(si.dp.x, si.dp.y, si.dp.z), si.om.w);
            s.time = si.c.w;
        },
        [&](int i, bool blocked) {
            if (!blocked) shadow_deliver(ws, cur, ws.sq[i]);
        });
}

__global__ __launch_bounds__(kBlock) void k_shadow_group(SceneView sv, Workspace ws, int cur, int group_below,
                                                          int cap, int cap_phys) {
    extern __shared__ int lds[];
    const int total = shadow_len(ws.counters, 1 - cur);
    if (total >= group_below) return;
    shadow_group(sv, ws, cur, total, lds, cap, cap_phys);
}

// Shadow rays on the persistent scheme of the extend kernel (scenes without
// a medium: a plain visibility test), the persistent mode's default. The
// reference tried this on 32-lane warps and reverted it (README.md:170); on
// 64-lane waves over the four-wide tree it took 28 ms where one lane per ray
// (k_shadow) took 44 ms on the 1M-triangle mesh at 1024x1024, 64 spp.
__global__ __launch_bounds__(kBlock) void k_shadow_persistent(SceneView sv, Workspace ws, int cur, int refill,
                                                               int group_below) {
    extern __shared__ int lds_stack[];
    int* stk = lds_stack + threadIdx.x;
    const int total = shadow_len(ws.counters, 1 - cur);
    if (total < group_below) return;                 // k_shadow_group tests a short queue
    if ((int)blockIdx.x * kBlock >= total) return;  // blocks the queue does not need
    const int lane = lane_id();
    Trav s;
    bool has = false;
    int slot = -1;
    bool exhausted = false;  // uniform across the wave
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
                        trav_begin(s, vec3(si.om.x, si.om.y, si.om.z), vec3(si.dp.x, si.dp.y, si.dp.z), si.om.w);
                        s.time = si.c.w;
                        slot = q;
                        has = true;
                    }
                }
                if (first + nidle >= total) exhausted = true;
            }
        }
        if (exhausted && ballot64(has) == 0) break;
        if (has) {
            const int st = step_wide_any(sv.bvh, s, stk, kBlock);
            if (st != 0) {
                if (st == 1) shadow_deliver(ws, cur, ws.sq[slot]);
                has = false;
            }
        }
    }
}

// -------------------------------------------------------------- finalize
// finalizeSampleKernel (:3740-3755): the pass's samples of a pixel are added
// in sample order, so the film holds the same FP32 sums however many
// samples a pass carries.
__global__ __launch_bounds__(kBlock) void k_finalize(Workspace ws, float* accum, int npix, int batch) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= npix) return;
    float r = accum[3 * p], g = accum[3 * p + 1], b = accum[3 * p + 2];
    for (int k = 0; k < batch; k++) {
        const float4 s = ws.done[(size_t)k * npix + p];
        const vec3 c = clamp_sample(vec3(s.x, s.y, s.z));
        r = r + c.x;
        g = g + c.y;
        b = b + c.z;
    }
    accum[3 * p] = r;
    accum[3 * p + 1] = g;
    accum[3 * p + 2] = b;
}

// ------------------------------------------------------------ megakernel
// trace_sample_kernel (:2900-3059); 16x16 pixel tiles, one lane per path
__global__ __launch_bounds__(kBlock) void k_megakernel(SceneView sv, mcpt_camera cam, int sample, uint64_t seed,
                                                       int max_depth, float* accum, int* counters) {
    extern __shared__ int lds_stack[];
    int* stk = lds_stack + threadIdx.x;
    const int px = blockIdx.x * 16 + (threadIdx.x & 15);
    const int py = blockIdx.y * 16 + (threadIdx.x >> 4);
    uint32_t n_ext = 0, n_shd = 0;
    if (px < cam.width && py < cam.height) {
        const int p = py * cam.width + px;
        Rng rng;
        mcpt_pcg32_for_path(&rng.g, (uint32_t)p, (uint32_t)sample, seed);
        vec3 ro, rd;
        float time;
        camera_ray(cam, px, py, rng, ro, rd, time);
        PathState ps;
        ps.thr = vec3(1.0f, 1.0f, 1.0f);
        ps.smp = vec3(0.0f, 0.0f, 0.0f);
        ps.prev_pdf = 1.0f;
        ps.prev_spec = true;
        const vec3 bg = load3(cam.background);
        for (int depth = 0; depth < max_depth; depth++) {
            Trav s;
            trav_begin(s, ro, rd, 1e30f);
            s.time = time;
            uint32_t ns = 0, nl = 0;
            trace_closest<false>(sv.bvh, s, stk, kBlock, ns, nl);
            n_ext++;
            const int prim = s.leaf >= 0 ? sv.bvh.leaf_prim[s.leaf] : -1;
            Bounce out;
            shade_bounce(sv, bg, cam.rr_cap, depth, ro, rd, s.leaf >= 0, s.t, prim, ps, rng, out, time);
            if (out.shadow) {
                n_shd++;
                const float tr = shadow_transmittance(sv, out.so, out.sd, out.smax, stk, kBlock, time);
                if (tr > 0.0f) ps.smp = ps.smp + out.contrib * tr;
            }
            if (!out.next) break;
            ro = out.no;
            rd = out.nd;
        }
        const vec3 c = clamp_sample(ps.smp);
        accum[3 * p] = accum[3 * p] + c.x;
        accum[3 * p + 1] = accum[3 * p + 1] + c.y;
        accum[3 * p + 2] = accum[3 * p + 2] + c.z;
    }
    uint64_t e = n_ext, h = n_shd;
    for (int off = 32; off > 0; off >>= 1) {
        e += __shfl_xor(e, off);
        h += __shfl_xor(h, off);
    }
    if (lane_id() == 0) {
        atomicAdd(reinterpret_cast<unsigned long long*>(u64slot(counters, C_EXT_TOTAL)), (unsigned long long)e);
        atomicAdd(reinterpret_cast<unsigned long long*>(u64slot(counters, C_SHD_TOTAL)), (unsigned long long)h);
    }
}

// --------------------------------------------------------------- resolve
// linearToSrgb / toneMapChannel / resolve_kernel (:3062-3100)
__device__ float linear_to_srgb(float x) {
    x = fmaxf(x, 0.0f);
    if (x <= 0.0031308f) return 12.92f * x;
    return 1.055f * mf_pow(x, 1.0f / 2.4f) - 0.055f;
}
__device__ float tonemap_channel(float x, int tm) {
    x = fmaxf(x, 0.f);
    if (tm == MCPT_TONEMAP_REINHARD) return x / (1.f + x);
    if (tm == MCPT_TONEMAP_ACES) {
        float m = (x * (2.51f * x + 0.03f)) / (x * (2.43f * x + 0.59f) + 0.14f);
        return fminf(fmaxf(m, 0.0f), 1.0f);
    }
    return x;
}
__global__ __launch_bounds__(kBlock) void k_resolve(const float* accum, float* out, int n, int spp, int tm,
                                                    float exposure) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= 3 * n) return;
    const float a = accum[i] / float(spp) * exposure;
    out[i] = linear_to_srgb(tonemap_channel(a, tm));
}

// resolve_kernel plus the output conversion of main (:4366-4380): the NaN
// shield, clamp to [0, 1] and (unsigned char)(255.99 x), three bytes per
// pixel written on the device
__global__ __launch_bounds__(kBlock) void k_resolve_rgb8(const float* accum, uint8_t* out, int n, int spp, int tm,
                                                         float exposure) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= 3 * n) return;
    const float a = accum[i] / float(spp) * exposure;
    const float c = linear_to_srgb(tonemap_channel(a, tm));
    const float v = __builtin_isfinite(c) ? fminf(fmaxf(c, 0.0f), 1.0f) : 0.0f;
    out[i] = (uint8_t)(255.99f * v);
}

// logLuminanceSumKernel (:3101-3111): ln(1e-4 + luminance) per pixel (a
// non-finite luminance counts as 0), summed per block in FP64 in a fixed
// order; the host adds the block sums in block order (the reference adds
// warp sums with float atomics, in no fixed order)
__global__ __launch_bounds__(kBlock) void k_log_luminance(const float* accum, int n, int spp, double* partial) {
    __shared__ double sh[kBlock];
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double v = 0.0;
    if (i < n) {
        const float r = accum[3 * i] / float(spp), g = accum[3 * i + 1] / float(spp), b = accum[3 * i + 2] / float(spp);
        const float lum = 0.2126f * r + 0.7152f * g + 0.0722f * b;
        v = (double)mf_log(1e-4f + (__builtin_isfinite(lum) ? fmaxf(lum, 0.0f) : 0.0f));
    }
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// ------------------------------------------------------------ trace hooks
template <bool COUNT, bool REF>
__global__ __launch_bounds__(kBlock) void k_trace_closest(SceneView sv, const mcpt_ray* rays, int n,
                                                          mcpt_packed_hit* hits, uint32_t* steps, uint32_t* leaves) {
    extern __shared__ int lds_stack[];
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const mcpt_ray r = rays[i];
    Trav s;
    trav_begin<REF>(s, load3(r.origin), load3(r.dir), r.tmax);
    s.time = r.time;
    uint32_t ns = 0, nl = 0;
    trace_closest<COUNT, REF>(sv.bvh, s, lds_stack + threadIdx.x, kBlock, ns, nl);
    hits[i] = pack_hit(sv.bvh, s);
    if (COUNT) {
        steps[i] = ns;
        leaves[i] = nl;
    }
}
// the hook without counts: the persistent near-first walk of the extend
// kernel over the caller's rays (their tmax and time included)
__global__ __launch_bounds__(kBlock) void k_trace_closest_persistent(SceneView sv, const mcpt_ray* rays, int n,
                                                                     mcpt_packed_hit* hits, int* work, int refill) {
    extern __shared__ int lds_stack[];
    if ((int)blockIdx.x * kBlock >= n) return;
    uint32_t ns = 0, nl = 0;
    persistent_closest<false, false>(
        sv.bvh, work, n, refill, lds_stack + threadIdx.x, kBlock,
        [&](int i, Trav& s, const BvhView&) {
            const mcpt_ray r = rays[i];
            trav_begin(s, load3(r.origin), load3(r.dir), r.tmax);
            s.time = r.time;
        },
        [&](int i, const Trav& s, const BvhView&) { hits[i] = pack_hit(sv.bvh, s); }, ns, nl);
}
// the any-hit hook on the persistent scheme of k_shadow_persistent
__global__ __launch_bounds__(kBlock) void k_trace_any_persistent(SceneView sv, const mcpt_ray* rays, int n,
                                                                 uint8_t* occ, int* work, int refill) {
    extern __shared__ int lds_stack[];
    int* stk = lds_stack + threadIdx.x;
    if ((int)blockIdx.x * kBlock >= n) return;
    const int lane = lane_id();
    Trav s;
    bool has = false, exhausted = false;
    int slot = -1;
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
                    if (q < n) {
                        const mcpt_ray r = rays[q];
                        trav_begin(s, load3(r.origin), load3(r.dir), r.tmax);
                        s.time = r.time;
                        slot = q;
                        has = true;
                    }
                }
                if (first + nidle >= n) exhausted = true;
            }
        }
        if (exhausted && ballot64(has) == 0) break;
        if (has) {
            const int st = step_wide_any(sv.bvh, s, stk, kBlock);
            if (st != 0) {
                occ[slot] = st == 2 ? 1 : 0;
                has = false;
            }
        }
    }
}
__global__ void k_zero_int(int* p) {
    if (threadIdx.x == 0) *p = 0;
}

int blocks_for(int64_t n) { return (int)((n + kBlock - 1) / kBlock); }
size_t stack_bytes(const mcpt_scene_s* s) { return (size_t)kBlock * stack_entries(s->max_depth) * sizeof(int); }

// Group walks (k_extend_group, k_shadow_group): a queue shorter than
// kGroupBelow rays is walked by kGroup lanes per ray.
constexpr int kGroupBelow = 32768;
// Default refill threshold of the persistent extend and shadow waves (idle
// lanes of 64 before a refill). Each refill is one atomic on the launch's
// work counter whose round trip the wave waits for, so lower thresholds
// trade idle lanes for stalls, and the stalls grow with the waves that share
// the counter. Measured on the 1M-triangle mesh at 1024^2, two to seven runs
// each: with two lanes (each launch on half of the grid) 16 / 24 / 32 / 48
// idle lanes gave 1,329 / 1,355 / 1,340 / 1,193 Mrays/s at 20 spp and
// 1,416 / 1,516 / 1,499 / 1,322 at 64 spp; with one lane on the full grid
// (the kernel-timed render) 24 took extend from 43.1 to 45.3 ms at 20 spp.
constexpr int kRefillTwoLanes = 24;
constexpr int kRefillOneLane = 32;
int scene_group_pairs(const mcpt_scene_s* s) { return group_stack_pairs(stack_entries(s->max_depth)); }
size_t group_lds_bytes(const mcpt_scene_s* s) {
    return (size_t)(kBlock / kGroup) * group_lds_ints(scene_group_pairs(s)) * sizeof(int);
}

// Layout of a wavefront workspace of `capacity` paths from `base`: two ray
// queues, hit records, the shadow queue, path state, then the counters.
constexpr size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
constexpr size_t workspace_bytes(int64_t capacity) {
    return 2 * align256(sizeof(RayItem) * capacity) + align256(sizeof(mcpt_packed_hit) * capacity) +
           align256(sizeof(ShadowItem) * capacity) + 5 * align256(sizeof(float4) * capacity) +
           2 * align256(sizeof(uint64_t) * capacity) + C_NCOUNTERS * sizeof(int);
}
void layout_workspace(char* b, int64_t capacity, Workspace& ws) {
    const size_t cap_q = align256(sizeof(RayItem) * capacity), cap_h = align256(sizeof(mcpt_packed_hit) * capacity),
                 cap_s = align256(sizeof(ShadowItem) * capacity), cap_4 = align256(sizeof(float4) * capacity),
                 cap_r = align256(sizeof(uint64_t) * capacity);
    ws.q[0] = (RayItem*)b;
    b += cap_q;
    ws.q[1] = (RayItem*)b;
    b += cap_q;
    ws.hits = (mcpt_packed_hit*)b;
    b += cap_h;
    ws.sq = (ShadowItem*)b;
    b += cap_s;
    for (int q = 0; q < 2; q++) {
        ws.thr[q] = (float4*)b;
        b += cap_4;
        ws.smp[q] = (float4*)b;
        b += cap_4;
        ws.rng[q] = (uint64_t*)b;
        b += cap_r;
    }
    ws.done = (float4*)b;
    b += cap_4;
    ws.counters = (int*)b;
}

// The wavefront workspace lives in the context and grows to the largest
// size requested: `nlanes` workspaces of `capacity` paths each, one after
// the other.
int ensure_workspace(Ctx* ctx, int64_t capacity, int nlanes, Workspace* ws) {
    const size_t one = workspace_bytes(capacity), total = one * nlanes;
    if (ctx->ws_bytes < total) {
        if (ctx->ws) dev_free(ctx, ctx->ws, ctx->ws_bytes);
        ctx->ws = nullptr;
        ctx->ws_bytes = 0;
        void* p = nullptr;
        MCPT_TRY(dev_alloc(ctx, &p, total));
        ctx->ws = p;
        ctx->ws_bytes = total;
    }
    for (int l = 0; l < nlanes; l++) layout_workspace((char*)ctx->ws + l * one, capacity, ws[l]);
    return MCPT_OK;
}

// The fixed workspace of mcpt_render_noalloc: one copy per device, held
// until the process ends. Host code reaches it only through the address
// hipGetSymbolAddress returns for the device (kept in g_scratch_ptr); the
// array's name in host code is not a device address.
constexpr size_t kScratchBytes = size_t(64) << 20;
__device__ __align__(256) char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_init_mu;
char* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu[kMaxDevices];  // one no-allocation render per device at a time
// the largest path count whose workspace fits the fixed array
constexpr int64_t scratch_capacity() {
    int64_t c = (int64_t)((kScratchBytes - C_NCOUNTERS * sizeof(int)) / 224);
    while (c > 0 && workspace_bytes(c) > kScratchBytes) c--;
    return c;
}
static_assert(scratch_capacity() == MCPT_NOALLOC_MAX_PATHS, "the header's path limit is the workspace's");

struct Timer {
    bool on;
    hipStream_t st;
    std::vector<hipEvent_t> ev;
    std::vector<int> kind;
    explicit Timer(bool enable, hipStream_t s) : on(enable), st(s) {}
    int mark(int k) {  // records an event that closes a span of kind k
        if (!on) return MCPT_OK;
        hipEvent_t e;
        MCPT_HIP(hipEventCreate(&e));
        MCPT_HIP(hipEventRecord(e, st));
        ev.push_back(e);
        kind.push_back(k);
        return MCPT_OK;
    }
    ~Timer() {
        for (auto e : ev) hipEventDestroy(e);
    }
};

}  // namespace

// ------------------------------------------------------------ host driver
namespace {

int check_render_args(const mcpt_scene_s* s, const mcpt_camera* cam, const mcpt_render_opts* o,
                      const mcpt_film_s* film) {
    if (!s->built) return set_error(MCPT_ERR_INVALID, "scene has no BVH (call mcpt_lbvh_build)");
    if (film->w != cam->width || film->h != cam->height)
        return set_error(MCPT_ERR_INVALID, "film %dx%d does not match camera %dx%d", film->w, film->h, cam->width,
                         cam->height);
    if (o->spp < 1 || o->max_depth < 1 || o->sample_begin < 0)
        return set_error(MCPT_ERR_INVALID, "spp, max_depth must be >= 1 and sample_begin >= 0");
    return MCPT_OK;
}

// Enqueues a whole render over the workspaces ws[0..nlanes-1], `batch`
// samples per wavefront pass; `tm` (may be NULL) marks the kernel spans.
// Returns the number of extend launches in *extend_launches.
//
// Without `tm`:
//   * each bounce's shadow rays run on a second stream beside the next
//     bounce's closest-hit rays: neither reads what the other writes (the
//     shadow rays add to the states of the next queue, which only the next
//     shade reads), so one launch's last slow rays overlap the other's start;
//     the next shade waits for them;
//   * with kRenderLanes lanes, consecutive passes go round the lanes'
//     workspaces and stream pairs, so one pass's launches fill the tails of
//     the others', each lane's persistent launches on its share of the grid.
//     Passes add to the film in pass order (each finalize waits for the
//     previous one), so the film is the same bit for bit.
// With `tm` every launch runs in order on one stream, so each kernel's span
// is its own.
int enqueue_render(mcpt_scene_s* s, const mcpt_camera* cam, const mcpt_render_opts* o, mcpt_film_s* film,
                   const Workspace* ws, int nlanes, int batch, Timer* tm, uint64_t* extend_launches_out) {
    Ctx* ctx = s->ctx;
    const bool overlap = tm == nullptr;
    if (!overlap) nlanes = 1;
    nlanes = std::min(nlanes, (int)Ctx::kMaxLanes);
    hipStream_t mains[Ctx::kMaxLanes], sides[Ctx::kMaxLanes];
    for (int l = 0; l < Ctx::kMaxLanes; l++) {
        mains[l] = l == 0 ? ctx->stream : ctx->lane_stream[l];
        sides[l] = !overlap ? mains[l] : !kLaneSideStreams ? mains[l] : l == 0 ? ctx->side : ctx->lane_side[l];
    }
    const int W = cam->width, H = cam->height, npix = W * H;
    const SceneView sv = s->view();
    const vec3 bg = load3(cam->background);
    const size_t lds = stack_bytes(s);
    const bool count = (o->flags & MCPT_RENDER_COUNT_TRAVERSAL) != 0;
    const bool ref = (o->flags & MCPT_RENDER_REFERENCE_ORDER) != 0;
    int refill = o->refill_lanes > 0 ? std::min(64, o->refill_lanes) : nlanes > 1 ? kRefillTwoLanes : kRefillOneLane;
    enum { K_CAMERA, K_EXTEND, K_SHADE, K_SHADOW, K_FINALIZE, K_OTHER };
    auto mark = [&](int k) { return tm ? tm->mark(k) : MCPT_OK; };

    for (int l = 0; l < nlanes; l++) k_counters_zero_totals<<<1, C_NCOUNTERS, 0, ctx->stream>>>(ws[l].counters);
    if (nlanes > 1) {  // the other lanes start after the clear above
        MCPT_HIP(hipEventRecord(ctx->ev_join, ctx->stream));
        for (int l = 1; l < nlanes; l++) MCPT_HIP(hipStreamWaitEvent(mains[l], ctx->ev_join, 0));
    }
    MCPT_TRY(mark(K_OTHER));

    int persistent_blocks = 0, shadow_blocks = 0, group_blocks = 0;
    const bool persistent_shadow =
        o->mode == MCPT_MODE_PERSISTENT && !s->has_medium && !(o->flags & MCPT_RENDER_SHADOW_PER_RAY);
    // Short queues go to the group walks; the counting and reference-order
    // passes keep one lane per ray (their step counts are the single walk's).
    const int group_below = o->mode != MCPT_MODE_PERSISTENT || count || ref || o->group_below < 0 ? 0
                            : o->group_below > 0                                                ? o->group_below
                                                                                                 : kGroupBelow;
    const int gcap_phys = scene_group_pairs(s);
    const size_t glds = group_lds_bytes(s);
    if (o->mode == MCPT_MODE_PERSISTENT) {
        int per_cu = 0;
        // With two lanes each lane's persistent launches take half of a CU's
        // blocks, so both lanes' launches are resident together instead of
        // one waiting for the other's blocks to retire (config 2: 1,185 ->
        // 1,277 Mrays/s at 20 spp, 1,420 -> 1,471 at 64 spp; a third or a
        // full grid per lane measured slower).
        MCPT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_extend_persistent<false, false>, kBlock, lds));
        persistent_blocks = std::max(1, (per_cu + nlanes - 1) / nlanes) * ctx->cu_count;
        MCPT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_shadow_persistent, kBlock, lds));
        shadow_blocks = std::max(1, (per_cu + nlanes - 1) / nlanes) * ctx->cu_count;
        if (group_below > 0) {  // a group per ray at most
            MCPT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_extend_group, kBlock, glds));
            const int need = (group_below + kBlock / kGroup - 1) / (kBlock / kGroup);
            group_blocks = std::min(need, std::max(1, (per_cu + nlanes - 1) / nlanes) * ctx->cu_count);
        }
    }
    int shade_blocks = 0;
    {
        int per_cu = 0;
        MCPT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_shade, kShadeBlock, 0));
        shade_blocks = std::max(1, per_cu) * ctx->cu_count;
    }
    uint64_t extend_launches = 0;
    if (o->mode == MCPT_MODE_MEGAKERNEL) {
        for (int k = 0; k < o->spp; k++) {
            dim3 grid((W + 15) / 16, (H + 15) / 16);
            k_megakernel<<<grid, kBlock, lds, ctx->stream>>>(sv, *cam, o->sample_begin + k, o->seed, o->max_depth,
                                                            film->accum.ptr, ws[0].counters);
            MCPT_TRY(mark(K_EXTEND));
        }
    }
    int pass = 0;
    for (int k0 = 0; o->mode != MCPT_MODE_MEGAKERNEL && k0 < o->spp; k0 += batch, pass++) {
        const int lane = pass % nlanes;
        const Workspace& w = ws[lane];
        hipStream_t st = mains[lane], shs = sides[lane];
        hipEvent_t ev_shade = ctx->ev_shade[lane], ev_shadow = ctx->ev_shadow[lane];
        const int nb = std::min(batch, o->spp - k0);
        const int npath = npix * nb;
        const int grid = blocks_for(npath);
        int cur = 0;
        k_counters_init<<<1, 64, 0, st>>>(w.counters, cur, npath);
        k_camera<<<grid, kBlock, 0, st>>>(*cam, o->sample_begin + k0, nb, o->seed, w, cur);
        MCPT_TRY(mark(K_CAMERA));
        for (int depth = 0; depth < o->max_depth; depth++) {
            k_counters_bounce<<<1, 64, 0, st>>>(w.counters, cur);
            MCPT_TRY(mark(K_OTHER));
            if (o->mode == MCPT_MODE_PERSISTENT) {
                const int blocks = std::min(grid, persistent_blocks);
                auto kern = count ? (ref ? k_extend_persistent<true, true> : k_extend_persistent<true, false>)
                                  : (ref ? k_extend_persistent<false, true> : k_extend_persistent<false, false>);
                kern<<<blocks, kBlock, lds, st>>>(sv, w, cur, refill, group_below);
                if (group_below > 0)
                    k_extend_group<<<std::min(grid, group_blocks), kBlock, glds, st>>>(sv, w, cur, group_below,
                                                                                     kGroupCap, gcap_phys);
            } else {
                auto kern = count ? (ref ? k_extend<true, true> : k_extend<true, false>)
                                  : (ref ? k_extend<false, true> : k_extend<false, false>);
                kern<<<grid, kBlock, lds, st>>>(sv, w, cur);
            }
            extend_launches++;
            MCPT_TRY(mark(K_EXTEND));
            if (overlap && depth > 0) MCPT_HIP(hipStreamWaitEvent(st, ev_shadow, 0));  // the last shadow rays
            k_shade<<<std::min((npath + kShadeBlock - 1) / kShadeBlock, shade_blocks), kShadeBlock, 0, st>>>(
                sv, w, cur, depth, o->max_depth, bg, cam->rr_cap, npix);
            MCPT_TRY(mark(K_SHADE));
            if (overlap) {
                MCPT_HIP(hipEventRecord(ev_shade, st));
                MCPT_HIP(hipStreamWaitEvent(shs, ev_shade, 0));
            }
            k_counters_shadow_total<<<1, 64, 0, shs>>>(w.counters, cur);
            if (persistent_shadow) {
                k_shadow_persistent<<<std::min(grid, shadow_blocks), kBlock, lds, shs>>>(sv, w, cur, refill,
                                                                                        group_below);
                if (group_below > 0)
                    k_shadow_group<<<std::min(grid, group_blocks), kBlock, glds, shs>>>(sv, w, cur, group_below,
                                                                                      kGroupCap, gcap_phys);
            } else
                k_shadow<<<grid, kBlock, lds, shs>>>(sv, w, cur);
            if (overlap) MCPT_HIP(hipEventRecord(ev_shadow, shs));
            MCPT_TRY(mark(K_SHADOW));
            cur = 1 - cur;
        }
        if (overlap) MCPT_HIP(hipStreamWaitEvent(st, ev_shadow, 0));
        if (nlanes > 1 && pass > 0)  // film order: after the previous pass's update
            MCPT_HIP(hipStreamWaitEvent(st, ctx->ev_fin[(lane + nlanes - 1) % nlanes], 0));
        k_finalize<<<blocks_for(npix), kBlock, 0, st>>>(w, film->accum.ptr, npix, nb);
        if (nlanes > 1) MCPT_HIP(hipEventRecord(ctx->ev_fin[lane], st));
        MCPT_TRY(mark(K_FINALIZE));
    }
    if (nlanes > 1 && pass > 0) {  // the caller's stream ends after the last film update
        const int last = (pass - 1) % nlanes;
        if (last != 0) MCPT_HIP(hipStreamWaitEvent(ctx->stream, ctx->ev_fin[last], 0));
    }
    MCPT_HIP(hipGetLastError());
    *extend_launches_out = o->mode == MCPT_MODE_MEGAKERNEL ? (uint64_t)o->spp : extend_launches;
    return MCPT_OK;
}

// sums of the lanes' u64 totals: extend rays, shadow rays, steps, leaves
int read_totals(const Workspace* ws, int nlanes, uint64_t totals[4]) {
    const int at[4] = {C_EXT_TOTAL, C_SHD_TOTAL, C_STEPS, C_LEAVES};
    for (int i = 0; i < 4; i++) totals[i] = 0;
    for (int l = 0; l < nlanes; l++) {
        int ctr[C_NCOUNTERS];
        MCPT_HIP(hipMemcpy(ctr, ws[l].counters, sizeof(ctr), hipMemcpyDeviceToHost));
        for (int i = 0; i < 4; i++) {
            uint64_t v;
            std::memcpy(&v, ctr + at[i], sizeof(uint64_t));
            totals[i] += v;
        }
    }
    return MCPT_OK;
}

}  // namespace

int render(mcpt_scene_s* s, const mcpt_camera* cam, const mcpt_render_opts* o, mcpt_film_s* film,
           mcpt_render_stats* stats) {
    MCPT_TRY(check_render_args(s, cam, o, film));
    Ctx* ctx = s->ctx;
    hipStream_t st = ctx->stream;
    const int npix = cam->width * cam->height;
    const bool timed = (o->flags & MCPT_RENDER_TIME_KERNELS) != 0;

    // Samples per wavefront pass. Every bounce ends with the slowest rays of
    // its launch, so the late, thin bounces of one sample leave most of the
    // GPU idle; carrying many samples per pass keeps them full. Default: up
    // to 64 Mi paths (about 14 GB of queues and path state at 224 B per
    // path), at most a quarter of the context's HBM budget. An untimed
    // wavefront render of two or more samples splits them over two lanes
    // (enqueue_render), each with half the paths.
    const int64_t per_path = 2 * 32 + 16 + 48 + 2 * (2 * 16 + 8) + 16;
    int64_t cap_paths = std::min<int64_t>(int64_t(64) << 20, (int64_t)(ctx->budget / 4) / per_path);
    int batch = o->batch > 0 ? o->batch : (int)std::max<int64_t>(1, cap_paths / npix);
    batch = std::max(1, std::min(batch, o->spp));
    if ((int64_t)npix * batch > (int64_t(1) << 30)) batch = std::max(1, (1 << 30) / npix);
    const int nlanes =
        (!timed && o->mode != MCPT_MODE_MEGAKERNEL) ? std::max(1, std::min(kRenderLanes, o->spp)) : 1;
    if (nlanes > 1 && o->batch <= 0) batch = std::max(1, (std::min(batch, o->spp) + nlanes - 1) / nlanes);

    Workspace ws[Ctx::kMaxLanes];
    MCPT_TRY(ensure_workspace(ctx, o->mode == MCPT_MODE_MEGAKERNEL ? 1 : (int64_t)npix * batch, nlanes, ws));
    Timer tm(timed, st);
    hipEvent_t e0, e1;
    MCPT_HIP(hipEventCreate(&e0));
    MCPT_HIP(hipEventCreate(&e1));
    MCPT_HIP(hipEventRecord(e0, st));
    uint64_t extend_launches = 0;
    int rc = enqueue_render(s, cam, o, film, ws, nlanes, batch, timed ? &tm : nullptr, &extend_launches);
    if (rc < 0) {
        for (int l = 1; l < Ctx::kMaxLanes; l++) {
            (void)hipStreamSynchronize(ctx->lane_stream[l]);
            (void)hipStreamSynchronize(ctx->lane_side[l]);
        }
        for (hipStream_t x : {ctx->side, st}) (void)hipStreamSynchronize(x);
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        return rc;
    }
    MCPT_HIP(hipEventRecord(e1, st));
    MCPT_HIP(hipEventSynchronize(e1));
    float total_ms = 0.f;
    MCPT_HIP(hipEventElapsedTime(&total_ms, e0, e1));
    uint64_t totals[4];  // extend rays, shadow rays, steps, leaves
    MCPT_TRY(read_totals(ws, nlanes, totals));
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->total_ms = total_ms;
        stats->extend_launches = extend_launches;
        stats->extend_rays = totals[0];
        stats->shadow_rays = totals[1];
        stats->extend_steps = totals[2];
        stats->extend_leaves = totals[3];
        stats->paths = (uint64_t)npix * o->spp;
        if (timed) {
            double* acc[6] = {&stats->camera_ms, &stats->extend_ms, &stats->shade_ms, &stats->shadow_ms,
                              &stats->finalize_ms, nullptr};
            hipEvent_t prev = e0;
            for (size_t i = 0; i < tm.ev.size(); i++) {
                float ms = 0.f;
                MCPT_HIP(hipEventElapsedTime(&ms, prev, tm.ev[i]));
                if (acc[tm.kind[i]]) *acc[tm.kind[i]] += ms;
                prev = tm.ev[i];
            }
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return MCPT_OK;
}

int scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices)
        return set_error(MCPT_ERR_INVALID, "device %d is beyond the %d devices the fixed workspace table holds", device,
                         kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_init_mu);
    if (g_scratch_ptr[device]) return MCPT_OK;
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return set_error(MCPT_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    g_scratch_ptr[device] = (char*)p;
    return MCPT_OK;
}

int render_noalloc(mcpt_scene_s* s, const mcpt_camera* cam, const mcpt_render_opts* o, mcpt_film_s* film) {
    MCPT_TRY(check_render_args(s, cam, o, film));
    if (o->flags & MCPT_RENDER_TIME_KERNELS)
        return set_error(MCPT_ERR_INVALID, "kernel timing creates events; use mcpt_render for timed renders");
    Ctx* ctx = s->ctx;
    const int dev = ctx->device;
    if (dev < 0 || dev >= kMaxDevices || !g_scratch_ptr[dev])
        return set_error(MCPT_ERR_INVALID, "the fixed workspace of device %d was not set up", dev);
    const int64_t npix = (int64_t)cam->width * cam->height;
    const int64_t cap = scratch_capacity();
    if (o->mode != MCPT_MODE_MEGAKERNEL && npix > cap)
        return set_error(MCPT_ERR_INVALID,
                         "%lld pixels exceed the %lld paths the fixed %zu-byte workspace holds in one pass; "
                         "use mcpt_render", (long long)npix, (long long)cap, kScratchBytes);
    int batch = (int)std::max<int64_t>(1, std::min<int64_t>(o->batch > 0 ? o->batch : cap / npix, cap / npix));
    batch = std::max(1, std::min(batch, o->spp));
    Workspace ws;
    layout_workspace(g_scratch_ptr[dev], o->mode == MCPT_MODE_MEGAKERNEL ? 1 : (int64_t)npix * batch, ws);
    std::lock_guard<std::mutex> lock(g_scratch_mu[dev]);
    uint64_t launches = 0;
    int rc = enqueue_render(s, cam, o, film, &ws, 1, batch, nullptr, &launches);
    if (rc < 0) (void)hipStreamSynchronize(ctx->side);
    hipError_t e = hipStreamSynchronize(ctx->stream);  // the array is free for the next call once this returns
    if (rc < 0) return rc;
    if (e != hipSuccess) return set_error(MCPT_ERR_HIP, "render_noalloc: %s", hipGetErrorString(e));
    return MCPT_OK;
}

int resolve(mcpt_film_s* film, int spp, int tonemap, float exposure, float* out) {
    Ctx* ctx = film->ctx;
    const int n = film->w * film->h;
    DevBuf<float> o;
    MCPT_TRY(o.alloc(ctx, (size_t)n * 3));
    k_resolve<<<blocks_for(3 * n), kBlock, 0, ctx->stream>>>(film->accum.ptr, o.ptr, n, spp, tonemap, exposure);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out, o.ptr, (size_t)n * 3 * sizeof(float), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    o.release(ctx);
    if (e != hipSuccess) return set_error(MCPT_ERR_HIP, "resolve: %s", hipGetErrorString(e));
    return MCPT_OK;
}

int resolve_rgb8(mcpt_film_s* film, int spp, int tonemap, float exposure, uint8_t* out) {
    Ctx* ctx = film->ctx;
    const int n = film->w * film->h;
    DevBuf<uint8_t> o;
    MCPT_TRY(o.alloc(ctx, (size_t)n * 3));
    k_resolve_rgb8<<<blocks_for(3 * n), kBlock, 0, ctx->stream>>>(film->accum.ptr, o.ptr, n, spp, tonemap, exposure);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out, o.ptr, (size_t)n * 3, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    o.release(ctx);
    if (e != hipSuccess) return set_error(MCPT_ERR_HIP, "resolve_rgb8: %s", hipGetErrorString(e));
    return MCPT_OK;
}

// computeAutoExposure (:3113-3124): target / exp(mean log luminance)
int auto_exposure(mcpt_film_s* film, int spp, float target, float* exposure) {
    Ctx* ctx = film->ctx;
    const int n = film->w * film->h;
    const int nb = blocks_for(n);
    DevBuf<double> part;
    MCPT_TRY(part.alloc(ctx, (size_t)nb));
    std::vector<double> h(nb);
    k_log_luminance<<<nb, kBlock, 0, ctx->stream>>>(film->accum.ptr, n, spp, part.ptr);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), part.ptr, nb * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    part.release(ctx);
    if (e != hipSuccess) return set_error(MCPT_ERR_HIP, "auto_exposure: %s", hipGetErrorString(e));
    double sum = 0.0;
    for (double v : h) sum += v;
    *exposure = target / mf_exp((float)(sum / (double)n));
    return MCPT_OK;
}

int trace_closest_host(mcpt_scene_s* s, const mcpt_ray* rays, int n, mcpt_packed_hit* hits, uint32_t* steps,
                       uint32_t* leaves) {
    if (!s->built) return set_error(MCPT_ERR_INVALID, "scene has no BVH");
    if (n < 0 || (n > 0 && (!rays || !hits))) return set_error(MCPT_ERR_INVALID, "bad trace arguments");
    if (n == 0) return MCPT_OK;
    Ctx* ctx = s->ctx;
    hipStream_t st = ctx->stream;
    const bool count = steps && leaves;
    DevBuf<mcpt_ray> dr;
    DevBuf<mcpt_packed_hit> dh;
    DevBuf<uint32_t> dcount;
    DevBuf<int> dwork;
    int status = dr.alloc(ctx, n);
    if (status >= 0) status = dh.alloc(ctx, n);
    if (status >= 0 && count) status = dcount.alloc(ctx, (size_t)2 * n);
    if (status >= 0 && !count) status = dwork.alloc(ctx, 1);
    hipError_t e = hipSuccess;
    if (status >= 0) {
        e = hipMemcpyAsync(dr.ptr, rays, sizeof(mcpt_ray) * n, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) {
            // with counts: the reference's visit order, so the counts are the
            // reference's; without: near-first order (same hits)
            if (count)
                k_trace_closest<true, true><<<blocks_for(n), kBlock, stack_bytes(s), st>>>(
                    s->view(), dr.ptr, n, dh.ptr, dcount.ptr, dcount.ptr + n);
            else {
                int per_cu = 0;
                e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_trace_closest_persistent, kBlock,
                                                                 stack_bytes(s));
                k_zero_int<<<1, 64, 0, st>>>(dwork.ptr);
                k_trace_closest_persistent<<<std::min(blocks_for(n), std::max(1, per_cu) * ctx->cu_count), kBlock,
                                             stack_bytes(s), st>>>(s->view(), dr.ptr, n, dh.ptr, dwork.ptr, 32);
            }
            if (e == hipSuccess) e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipMemcpyAsync(hits, dh.ptr, sizeof(mcpt_packed_hit) * n, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess && count) e = hipMemcpyAsync(steps, dcount.ptr, 4 * (size_t)n, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess && count)
            e = hipMemcpyAsync(leaves, dcount.ptr + n, 4 * (size_t)n, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
    }
    dr.release(ctx);
    dh.release(ctx);
    dcount.release(ctx);
    dwork.release(ctx);
    if (status < 0) return status;
    if (e != hipSuccess) return set_error(MCPT_ERR_HIP, "trace_closest: %s", hipGetErrorString(e));
    return MCPT_OK;
}

int trace_any_host(mcpt_scene_s* s, const mcpt_ray* rays, int n, uint8_t* occluded) {
    if (!s->built) return set_error(MCPT_ERR_INVALID, "scene has no BVH");
    if (n < 0 || (n > 0 && (!rays || !occluded))) return set_error(MCPT_ERR_INVALID, "bad trace arguments");
    if (n == 0) return MCPT_OK;
    Ctx* ctx = s->ctx;
    hipStream_t st = ctx->stream;
    DevBuf<mcpt_ray> dr;
    DevBuf<uint8_t> dout;
    DevBuf<int> dwork;
    int status = dr.alloc(ctx, n);
    if (status >= 0) status = dout.alloc(ctx, n);
    if (status >= 0) status = dwork.alloc(ctx, 1);
    hipError_t e = hipSuccess;
    if (status >= 0) {
        e = hipMemcpyAsync(dr.ptr, rays, sizeof(mcpt_ray) * n, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) {
            int per_cu = 0;
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_trace_any_persistent, kBlock, stack_bytes(s));
            k_zero_int<<<1, 64, 0, st>>>(dwork.ptr);
            k_trace_any_persistent<<<std::min(blocks_for(n), std::max(1, per_cu) * ctx->cu_count), kBlock,
                                     stack_bytes(s), st>>>(s->view(), dr.ptr, n, dout.ptr, dwork.ptr, 32);
            if (e == hipSuccess) e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipMemcpyAsync(occluded, dout.ptr, n, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
    }
    dr.release(ctx);
    dout.release(ctx);
    dwork.release(ctx);
    if (status < 0) return status;
    if (e != hipSuccess) return set_error(MCPT_ERR_HIP, "trace_any: %s", hipGetErrorString(e));
    return MCPT_OK;
}

}  // namespace mcpt
