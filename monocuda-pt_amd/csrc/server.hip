// server.hip — continuous-batching render server on one gfx950 device.
//
// The reference renders one image per run (main, src/path_tracer.cu:4038);
// serving is new here. A server keeps a pool of paths in flight that every
// admitted job shares. One step of the pool is one bounce for every path in
// it: generate camera paths of admitted jobs into the slots freed by the
// previous step, extend (persistent, the ray's scene picked per lane),
// shade, shadow, then finish the paths that terminated. Paths of jobs on
// different scenes, resolutions and sample counts sit in the same queues.
//
// The device schedules: the job table lives in HBM, and each step's plan
// kernel orders the running jobs by the server's policy and hands the free
// pool slots out as per-job quotas, with no host round trip. The host thread
// keeps one step queued ahead of the one it reads back: it enqueues step
// n + 1, then waits for step n's status copy (each job's progress) to retire
// finished jobs and admit waiting ones. A step's shadow rays and finished
// paths run on a second stream beside the next step's camera paths and
// closest-hit launch; that step's commits and shade wait for them.
//
// Determinism: a job's paths are keyed by (pixel, sample, seed) like
// mcpt_render's. A terminated path stores its clamped sample into the
// job's frame buffer for that sample; when all W*H paths of the oldest open
// sample are finished, the commit kernels add the frame into the film. Films
// therefore receive the samples of each pixel in sample order, with the
// same FP32 additions as mcpt_render, whatever else shares the pool.
//
// Admission: a job needs a job slot and W*H*12*(1 + frames) bytes of HBM
// within the server's film budget; otherwise it waits in FIFO order and is
// admitted when a finished job is released.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <thread>

#include "scene.hpp"
#include "shade.hpp"

namespace mcpt {

namespace {

constexpr int kBlock = 256;
constexpr int kMaxJobs = 1024;      // job slots of one server at most
constexpr int kStatusRing = 4;      // status copies in flight (two are enough)
// queues shorter than this are walked by 16 lanes per ray (group_closest,
// group_any), as in mcpt_render
constexpr int kSrvGroupBelow = 32768;

// A job slot's record. The host writes it whole at admission; from then on
// the device owns the progress fields.
struct JobDev {
    mcpt_camera cam;
    float* film;    // W*H*3 linear sums
    float* frames;  // nframes * W*H*3 clamped samples awaiting commit
    int scene;
    int max_depth;
    int npix;
    int nframes;
    uint64_t seed;
    uint64_t id;   // the host's job id
    int spp;
    uint32_t seq;  // admission order
    // progress
    int active;     // from admission until every sample is in the film
    int committed;  // samples added to the film
    int opened;     // samples whose frame slot is in use
    int ncommit;    // samples the last commit step added (committed - ncommit is the first)
    int64_t gen;    // next camera path: sample * npix + pixel
};

struct Quota {  // this step's new camera paths of one job slot
    int64_t start;  // linear index (sample * npix + pixel) of the first
    int count;
    int prefix;     // paths of the lower job slots this step
};

// Step counters, each in its own 128-byte line: every wave of a step takes
// slots with atomics on them, and atomics on one line serialize in one L2
// channel. S_ND holds the next-queue count (low 32 bits) and the done count
// (high 32 bits) for one combined push; S_GEN the first slot and the count
// of the step's camera paths.
enum {
    S_QCOUNT = 0,
    S_ND = 32,
    S_SHADOW = 64,
    S_WORK = 96,
    S_WORK_SHADOW = 128,
    S_GEN = 160,
    S_NCOUNTERS = 192
};

__device__ __forceinline__ unsigned long long* nd_counter(int* c) {
    return reinterpret_cast<unsigned long long*>(c + S_ND);
}

// Path state lives beside its ray: slot i of queue c owns thr[c][i],
// smp[c][i], rng[c][i] and meta[c][i], and shade moves the state with the
// ray when it compacts the survivors into the other queue. Every state
// access of a wave is then one coalesced run, where ids handed out by a
// free list would turn into a random permutation within a few steps.
struct Pool {
    RayItem* q[2];
    float4* thr[2];   // throughput rgb, previous bsdf pdf
    float4* smp[2];   // sample rgb, previous-bounce-specular flag
    uint64_t* rng[2];
    int4* meta[2];    // job slot, pixel, sample, depth
    mcpt_packed_hit* hits;
    ShadowItem* sq;   // dp.w: target state slot (>= 0 next queue, < 0 ~done slot)
    float4* done_smp; // terminated paths of the step
    int4* done_meta;
    int* counters;
    int* frame_done;  // max_jobs * max_frames
    JobDev* jobs;
    Quota* quotas;    // max_jobs
    const SceneView* scenes;
    const BvhView* bvhs;
    int max_frames;
    int max_jobs;
    int capacity;
    int tb;  // 24: ray records carry scene and time (see pack_scene_time)
};

// ------------------------------------------------------------------ kernels
// wave-wide inclusive prefix sum
__device__ __forceinline__ long long wave_incl_scan(long long v) {
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const long long u = __shfl_up(v, off);
        if (lane_id() >= off) v += u;
    }
    return v;
}

// in-place inclusive prefix sum of a[0, n) by one wave
__device__ void wave_scan_array(long long* a, int n) {
    long long carry = 0;
    for (int base = 0; base < n; base += 64) {
        const int i = base + lane_id();
        const long long incl = wave_incl_scan(i < n ? a[i] : 0);
        if (i < n) a[i] = carry + incl;
        carry += __shfl(incl, 63);
    }
}

// One wave (small, so that it finds room beside the other stream's
// persistent launch), job slots j = lane + 64k. The running jobs are ranked
// by policy (FIFO: admission order; shortest first: fewest paths left to
// start, ties by admission; aged: shortest first, but the oldest job first
// takes up to half of the free slots), and the pool's free slots (capacity -
// paths in flight after the last shade) go to them in that order, each job
// taking the paths of its open samples not yet started (a sample is open
// while fewer than nframes samples of the job await their commit). LDS:
// plan_lds_bytes(max_jobs).
__global__ __launch_bounds__(64) void k_srv_plan(Pool p, int policy) {
    extern __shared__ long long lds_plan[];
    const int nj = p.max_jobs;
    long long* s_key = lds_plan;           // [nj] policy key
    long long* s_want = lds_plan + nj;     // [nj] paths wanted, by rank; then their prefix sums
    long long* s_take = lds_plan + 2 * nj; // [nj] paths taken, by slot; then their prefix sums
    unsigned* s_seq = reinterpret_cast<unsigned*>(lds_plan + 3 * nj);  // [nj] admission order, 0 = not running
    const int lane = lane_id();
    const int cur_count = (int)(uint32_t)*nd_counter(p.counters);  // paths the last shade kept
    const long long free0 = p.capacity > cur_count ? p.capacity - cur_count : 0;
    for (int j = lane; j < nj; j += 64) {
        const JobDev& J = p.jobs[j];
        long long key = 0;
        unsigned seq = 0;
        if (J.active) {
            seq = J.seq;
            key = policy == MCPT_SCHED_FIFO ? (long long)seq : (long long)J.spp * J.npix - J.gen;
        }
        s_key[j] = key;
        s_seq[j] = seq;
        s_want[j] = 0;
    }
    __syncthreads();
    // the oldest running job, for the aged policy
    int oldest = -1;
    if (policy == MCPT_SCHED_SHORTEST_FIRST_AGED) {
        unsigned best = 0xFFFFFFFFu;
        for (int i = 0; i < nj; i++)
            if (s_seq[i] && s_seq[i] < best) {
                best = s_seq[i];
                oldest = i;
            }
    }
    // wants, in rank order; (key, admission order) is a total order
    long long reserve = 0;
    for (int j = lane; j < nj; j += 64) {
        if (!s_seq[j]) continue;
        const JobDev& J = p.jobs[j];
        const int opened = max(J.opened, min(J.spp, J.committed + J.nframes));
        const long long want = (long long)opened * J.npix - J.gen;
        int rank = 0;
        for (int i = 0; i < nj; i++)
            rank += s_seq[i] && (s_key[i] < s_key[j] || (s_key[i] == s_key[j] && s_seq[i] < s_seq[j])) ? 1 : 0;
        const long long r = j == oldest ? min(want, free0 / 2) : 0;
        if (j == oldest) reserve = r;
        s_want[rank] = want - r;
    }
    // the oldest job's first share (aged), then the rest in rank order
    for (int off = 32; off > 0; off >>= 1) reserve += __shfl_xor(reserve, off);
    const long long free1 = free0 - reserve;
    __syncthreads();
    wave_scan_array(s_want, nj);
    __syncthreads();
    for (int j = lane; j < nj; j += 64) {
        long long take = 0;
        if (s_seq[j]) {
            const JobDev& J = p.jobs[j];
            const int opened = max(J.opened, min(J.spp, J.committed + J.nframes));
            const long long want = (long long)opened * J.npix - J.gen;
            int rank = 0;
            for (int i = 0; i < nj; i++)
                rank += s_seq[i] && (s_key[i] < s_key[j] || (s_key[i] == s_key[j] && s_seq[i] < s_seq[j])) ? 1 : 0;
            const long long r = j == oldest ? min(want, free0 / 2) : 0;
            const long long want1 = want - r;
            const long long before = s_want[rank] - want1;
            take = r + max(0LL, min(want1, free1 - before));
        }
        s_take[j] = take;
    }
    __syncthreads();
    wave_scan_array(s_take, nj);
    __syncthreads();
    for (int j = lane; j < nj; j += 64) {
        JobDev& J = p.jobs[j];
        const long long incl = s_take[j], take = incl - (j ? s_take[j - 1] : 0);
        Quota q;
        q.start = s_seq[j] ? J.gen : 0;
        q.count = (int)take;
        q.prefix = (int)(incl - take);
        p.quotas[j] = q;
        if (s_seq[j]) {
            J.gen += take;
            J.opened = max(J.opened, min(J.spp, J.committed + J.nframes));
        }
    }
    if (lane == 0) {
        const int total = (int)s_take[nj - 1];
        p.counters[S_GEN] = cur_count;
        p.counters[S_GEN + 1] = total;
        p.counters[S_QCOUNT] = cur_count + total;
        p.counters[S_WORK] = 0;
    }
}

int plan_lds_bytes(int max_jobs) { return max_jobs * (3 * (int)sizeof(long long) + (int)sizeof(unsigned)); }

// camera paths of the step's quotas, appended to the current queue after
// the paths in flight (mcpt_render's camera kernel per path)
__global__ __launch_bounds__(kBlock) void k_srv_generate(Pool p, int cur) {
    const int base = p.counters[S_GEN], total = p.counters[S_GEN + 1];
    for (int g = blockIdx.x * kBlock + threadIdx.x; g < total; g += gridDim.x * kBlock) {
        int lo = 0, hi = p.max_jobs - 1;  // last job slot with prefix <= g
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (p.quotas[mid].prefix <= g) lo = mid;
            else hi = mid - 1;
        }
        const Quota q = p.quotas[lo];
        const JobDev& J = p.jobs[lo];
        const int64_t L = q.start + (g - q.prefix);
        const int sample = (int)(L / J.npix);
        const int pixel = (int)(L - (int64_t)sample * J.npix);
        Rng rng;
        mcpt_pcg32_for_path(&rng.g, (uint32_t)pixel, (uint32_t)sample, J.seed);
        vec3 o, d;
        float time;
        camera_ray(J.cam, pixel % J.cam.width, pixel / J.cam.width, rng, o, d, time);
        const int slot = base + g;
        RayItem it;
        it.ot = make_float4(o.x, o.y, o.z, pack_scene_time(J.scene, time, p.tb));
        it.dp = make_float4(d.x, d.y, d.z, 0.0f);
        p.q[cur][slot] = it;
        p.thr[cur][slot] = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
        p.smp[cur][slot] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
        p.rng[cur][slot] = rng.g.state;
        p.meta[cur][slot] = make_int4(lo, pixel, sample, 0);
    }
}

// the 16-byte hit record; prim_index keeps the leaf's code (triangle, or
// ~sphere)
__device__ __forceinline__ mcpt_packed_hit srv_hit(const Trav& s, const BvhView& v) {
    mcpt_packed_hit h;
    h.t = s.t;
    h.prim_type = s.leaf >= 0 ? 1 : -1;
    h.prim_index = s.leaf >= 0 ? v.leaf_prim[s.leaf] : -1;
    h.did_hit = s.leaf >= 0 ? 1 : 0;
    return h;
}

// persistent extend over rays of mixed scenes; the scene index rides in the
// ray record (ot.w) and a lane reloads the BVH view only when it changes.
// A queue shorter than group_below is k_srv_extend_group's.
__global__ __launch_bounds__(kBlock) void k_srv_extend(Pool p, int cur, int refill, int group_below) {
    extern __shared__ int lds_stack[];
    const int total = p.counters[S_QCOUNT];
    if (total < group_below) return;
    if ((int)blockIdx.x * kBlock >= total) return;  // blocks the queue does not need
    BvhView bv;
    int bv_scene = -1;  // scene whose view `bv` holds: refills rarely change it
    uint32_t ns = 0, nl = 0;
    const RayItem* q = p.q[cur];
    persistent_closest<false, false>(
        bv, &p.counters[S_WORK], total, refill, lds_stack + threadIdx.x, kBlock,
        [&](int i, Trav& s, BvhView& v) {
            const RayItem it = q[i];
            const int sc = unpack_scene(it.ot.w, p.tb);
            if (sc != bv_scene) {
                v = p.bvhs[sc];
                bv_scene = sc;
            }
            trav_begin(s, vec3(it.ot.x, it.ot.y, it.ot.z), vec3(it.dp.x, it.dp.y, it.dp.z), 1e30f);
            s.time = unpack_time(it.ot.w, p.tb);
        },
        [&](int i, const Trav& s, const BvhView& v) { p.hits[i] = srv_hit(s, v); }, ns, nl);
}

// a short queue: 16 lanes per ray (group_closest), each group's view that
// of its ray's scene
__global__ __launch_bounds__(kBlock) void k_srv_extend_group(Pool p, int cur, int group_below, int cap,
                                                              int cap_phys) {
    extern __shared__ int lds[];
    const int total = p.counters[S_QCOUNT];
    if (total >= group_below || (int)blockIdx.x * (kBlock / kGroup) >= total) return;
    int* base = lds + (threadIdx.x / kGroup) * group_lds_ints(cap_phys);
    const RayItem* q = p.q[cur];
    BvhView bv;
    group_closest(
        bv, &p.counters[S_WORK], total, reinterpret_cast<int2*>(base), base + 2 * cap_phys, cap, cap_phys,
        [&](int i, Trav& s, BvhView& v) {
            const RayItem it = q[i];
            v = p.bvhs[unpack_scene(it.ot.w, p.tb)];
            trav_begin(s, vec3(it.ot.x, it.ot.y, it.ot.z), vec3(it.dp.x, it.dp.y, it.dp.z), 1e30f);
            s.time = unpack_time(it.ot.w, p.tb);
        },
        [&](int i, const Trav& s, const BvhView& v) { p.hits[i] = srv_hit(s, v); });
}

// shadeKernel (src/path_tracer.cu:3470-3696) per queue slot; survivors
// move to the other queue with their state, terminated paths to the done
// list. Grid-stride over chunks of kShadeBlock slots (the grid is sized to
// occupancy: the host does not know the queue's length), each chunk shaded
// in shading-bin order (block_bin_sort).
constexpr int kShadeBlock = 1024;
__global__ __launch_bounds__(kShadeBlock) void k_srv_shade(Pool p, int cur) {
    __shared__ int lds[bin_sort_lds_ints<kShadeBlock>()];
    const int count = p.counters[S_QCOUNT];
    for (int chunk = blockIdx.x * kShadeBlock; chunk < count; chunk += gridDim.x * kShadeBlock) {
        int slot = chunk + (int)threadIdx.x, mat = -1, bin = kShadeBins;
        if (slot < count) {
            const mcpt_packed_hit hit = p.hits[slot];
            const SceneView& sv = p.scenes[unpack_scene(p.q[cur][slot].ot.w, p.tb)];
            bin = shade_bin(sv, hit.did_hit != 0, hit.prim_index, mat);
        }
        block_bin_sort<kShadeBlock>(bin, slot, mat, lds);
        const bool active = slot >= 0;
        Bounce out;
        out.next = false;
        out.shadow = false;
        bool finished = false;
        int scene = 0;
        float time = 0.0f;
        PathState ps;
        Rng rng;
        int4 m = make_int4(0, 0, 0, 0);
        if (active) {
            const RayItem it = p.q[cur][slot];
            const mcpt_packed_hit hit = p.hits[slot];
            scene = unpack_scene(it.ot.w, p.tb);
            time = unpack_time(it.ot.w, p.tb);
            m = p.meta[cur][slot];
            const JobDev& J = p.jobs[m.x];
            const SceneView& sv = p.scenes[scene];
            const float4 thr = p.thr[cur][slot], smp = p.smp[cur][slot];
            ps.thr = vec3(thr.x, thr.y, thr.z);
            ps.prev_pdf = thr.w;
            ps.smp = vec3(smp.x, smp.y, smp.z);
            ps.prev_spec = smp.w != 0.0f;
            rng.g.state = p.rng[cur][slot];
            rng.g.inc = ((uint64_t)(uint32_t)m.y << 1u) | 1u;
            shade_bounce(sv, load3(J.cam.background), J.cam.rr_cap, m.w, vec3(it.ot.x, it.ot.y, it.ot.z),
                         vec3(it.dp.x, it.dp.y, it.dp.z), hit.did_hit != 0, hit.t, hit.prim_index, ps, rng, out,
                         time, mat);
            // the reference's depth loop ends after shading bounce max_depth-1
            if (out.next && m.w + 1 >= J.max_depth) out.next = false;
            finished = !out.next;
        }
        // compacted pushes: every lane of the wave takes part
        int nslot, dslot;
        wave_push2(nd_counter(p.counters), out.next, finished, nslot, dslot);
        const int sslot = wave_push(&p.counters[S_SHADOW], out.shadow);
        const float4 smp_out = make_float4(ps.smp.x, ps.smp.y, ps.smp.z, ps.prev_spec ? 1.0f : 0.0f);
        if (out.next) {
            const int n = 1 - cur;
            RayItem it;
            it.ot = make_float4(out.no.x, out.no.y, out.no.z, pack_scene_time(scene, time, p.tb));
            it.dp = make_float4(out.nd.x, out.nd.y, out.nd.z, 0.0f);
            p.q[n][nslot] = it;
            p.thr[n][nslot] = make_float4(ps.thr.x, ps.thr.y, ps.thr.z, ps.prev_pdf);
            p.smp[n][nslot] = smp_out;
            p.rng[n][nslot] = rng.g.state;
            p.meta[n][nslot] = make_int4(m.x, m.y, m.z, m.w + 1);
        }
        if (finished) {
            p.done_smp[dslot] = smp_out;
            p.done_meta[dslot] = m;
        }
        if (out.shadow) {
            ShadowItem si;
            si.om = make_float4(out.so.x, out.so.y, out.so.z, out.smax);
            si.dp = make_float4(out.sd.x, out.sd.y, out.sd.z, __int_as_float(out.next ? nslot : ~dslot));
            si.c = make_float4(out.contrib.x, out.contrib.y, out.contrib.z, pack_scene_time(scene, time, p.tb));
            p.sq[sslot] = si;
        }
    }
}

// the light a shadow ray carries, into the sample of its path: the path's
// next slot, or (target < 0) the done-list entry of a path that ended
__device__ __forceinline__ void srv_add_shadow(const Pool& p, int cur, const ShadowItem& si, float tr) {
    const int target = __float_as_int(si.dp.w);
    float4* dst = target >= 0 ? p.smp[1 - cur] + target : p.done_smp + ~target;
    float4 v = *dst;
    v.x = v.x + si.c.x * tr;
    v.y = v.y + si.c.y * tr;
    v.z = v.z + si.c.z * tr;
    *dst = v;
}

// shadowKernel (src/path_tracer.cu:3698-3728) on the persistent scheme of
// k_srv_extend: a lane walks the any-hit steps of its ray and refills when
// done; a ray of a scene with a medium is handled whole in one iteration.
// The light sample is added to the path's state wherever shade moved it.
// A queue shorter than group_below is k_srv_shadow_group's.
__global__ __launch_bounds__(kBlock) void k_srv_shadow_persistent(Pool p, int cur, int refill, int group_below) {
    extern __shared__ int lds_stack[];
    int* stk = lds_stack + threadIdx.x;
    const int total = p.counters[S_SHADOW];
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
                if (lane == leader) first = atomicAdd(&p.counters[S_WORK_SHADOW], nidle);
                first = __shfl(first, leader);
                if (!has) {
                    const int q = first + rank_below(idle);
                    if (q < total) {
                        const ShadowItem si = p.sq[q];
                        const int sc = unpack_scene(si.c.w, p.tb);
                        const float tm = unpack_time(si.c.w, p.tb);
                        const vec3 o(si.om.x, si.om.y, si.om.z), d(si.dp.x, si.dp.y, si.dp.z);
                        if (p.scenes[sc].has_medium) {
                            const float tr = shadow_transmittance(p.scenes[sc], o, d, si.om.w, stk, kBlock, tm);
                            if (tr > 0.0f) srv_add_shadow(p, cur, si, tr);
                        } else {
                            if (sc != bv_scene) {
                                bv = p.bvhs[sc];
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
                if (st == 1) srv_add_shadow(p, cur, p.sq[slot], 1.0f);
                has = false;
            }
        }
    }
}

// a short shadow queue: 16 lanes per ray (group_any); only for servers
// whose scenes have no medium (group_below is 0 otherwise)
__global__ __launch_bounds__(kBlock) void k_srv_shadow_group(Pool p, int cur, int group_below, int cap,
                                                              int cap_phys) {
    extern __shared__ int lds[];
    const int total = p.counters[S_SHADOW];
    if (total >= group_below || (int)blockIdx.x * (kBlock / kGroup) >= total) return;
    int* base = lds + (threadIdx.x / kGroup) * group_lds_ints(cap_phys);
    BvhView bv;
    group_any(
        bv, &p.counters[S_WORK_SHADOW], total, reinterpret_cast<int2*>(base), base + 2 * cap_phys, cap, cap_phys,
        [&](int i, Trav& s, BvhView& v) {
            const ShadowItem si = p.sq[i];
            v = p.bvhs[unpack_scene(si.c.w, p.tb)];
            trav_begin(s, vec3(si.om.x, si.om.y, si.om.z), vec3(si.dp.x, si.dp.y, si.dp.z), si.om.w);
            s.time = unpack_time(si.c.w, p.tb);
        },
        [&](int i, bool blocked) {
            if (!blocked) srv_add_shadow(p, cur, p.sq[i], 1.0f);
        });
}

// terminated paths: clamped sample into the frame buffer of the path's
// sample (finalizeSampleKernel's clamp, :3740-3755), and the count of
// finished paths per frame. The terminated paths of a chunk nearly always
// belong to one or two frames: wave leaders add their group counts into a
// small LDS table and the block adds each table entry with one atomic.
// Grid-stride over chunks of kFinishBlock paths.
constexpr int kFinishBlock = 1024;
constexpr int kFinishKeys = 32;
__global__ __launch_bounds__(kFinishBlock) void k_srv_finish(Pool p) {
    __shared__ int lkey[kFinishKeys], lcnt[kFinishKeys];
    const int ndone = (int)(*nd_counter(p.counters) >> 32);
    for (int chunk = blockIdx.x * kFinishBlock; chunk < ndone; chunk += gridDim.x * kFinishBlock) {
        if (threadIdx.x < kFinishKeys) {
            lkey[threadIdx.x] = -1;
            lcnt[threadIdx.x] = 0;
        }
        __syncthreads();
        const int d = chunk + (int)threadIdx.x;
        const bool valid = d < ndone;
        int key = -1;
        if (valid) {
            const int4 m = p.done_meta[d];
            const JobDev& J = p.jobs[m.x];
            const int fslot = m.z % J.nframes;
            const float4 s = p.done_smp[d];
            const vec3 c = clamp_sample(vec3(s.x, s.y, s.z));
            float* f = J.frames + ((size_t)fslot * J.npix + m.y) * 3;
            f[0] = c.x;
            f[1] = c.y;
            f[2] = c.z;
            key = m.x * p.max_frames + fslot;
        }
        uint64_t todo = ballot64(valid);
        while (todo) {  // one group per distinct key in the wave
            const int leader = __builtin_ctzll(todo);
            const int k = __shfl(key, leader);
            const uint64_t same = ballot64(valid && key == k);
            if (lane_id() == leader) {
                const int n = __popcll(same);
                bool placed = false;
                for (int i = 0; i < kFinishKeys && !placed; i++) {
                    const int e = (k + i) & (kFinishKeys - 1);
                    const int prev = atomicCAS(&lkey[e], -1, k);
                    if (prev == -1 || prev == k) {
                        atomicAdd(&lcnt[e], n);
                        placed = true;
                    }
                }
                if (!placed) atomicAdd(&p.frame_done[k], n);  // table full
            }
            todo &= ~same;
        }
        __syncthreads();
        if (threadIdx.x < kFinishKeys && lcnt[threadIdx.x] > 0)
            atomicAdd(&p.frame_done[lkey[threadIdx.x]], lcnt[threadIdx.x]);
        __syncthreads();
    }
}

// After the previous step's finished paths are counted: each running job's
// oldest open samples whose every path has finished are committed (their
// frame counters cleared, `committed` advanced, `ncommit` set for
// k_srv_commit_apply); a job with every sample committed stops running. The
// step's shade counters are cleared here, after the previous step's shadow
// and finish launches have read them.
__global__ __launch_bounds__(kMaxJobs) void k_srv_commit_plan(Pool p) {
    const int j = threadIdx.x;
    if (j < p.max_jobs) {
        JobDev& J = p.jobs[j];
        int n = 0;
        if (J.active) {
            int* fd = p.frame_done + j * p.max_frames;
            while (J.committed + n < J.opened && fd[(J.committed + n) % J.nframes] == J.npix) {
                fd[(J.committed + n) % J.nframes] = 0;
                n++;
            }
            J.committed += n;
            if (J.committed == J.spp) J.active = 0;
        }
        J.ncommit = n;
    }
    if (j == 0) {
        *nd_counter(p.counters) = 0;
        p.counters[S_SHADOW] = 0;
        p.counters[S_WORK_SHADOW] = 0;
    }
}

// film += the committed frames, in sample order (the same FP32 additions as
// mcpt_render's finalize); grid-stride over each committing job's film
__global__ __launch_bounds__(kBlock) void k_srv_commit_apply(Pool p) {
    for (int j = 0; j < p.max_jobs; j++) {
        const JobDev& J = p.jobs[j];
        const int n = J.ncommit;
        if (n == 0) continue;
        const int first = J.committed - n;
        const size_t n3 = (size_t)J.npix * 3;
        for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n3; i += (size_t)gridDim.x * kBlock) {
            float v = J.film[i];
            for (int c = 0; c < n; c++) v = v + J.frames[(size_t)((first + c) % J.nframes) * n3 + i];
            J.film[i] = v;
        }
    }
}

__global__ void k_srv_fill_f(float* p, size_t n, float v) {
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_srv_fill_i(int* p, int n, int v) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = v;
}

int blocks_for(int64_t n) { return (int)((n + kBlock - 1) / kBlock); }

double now_ms(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace
}  // namespace mcpt

using namespace mcpt;

struct JobHost {
    uint64_t id = 0;
    mcpt_job job{};
    int npix = 0, nframes = 1;
    int slot = -1;  // device job slot while running
    int status = MCPT_DEFERRED;
    double submit_ms = 0, start_ms = 0, finish_ms = 0;
    float* film = nullptr;
    float* frames = nullptr;
    size_t film_bytes = 0, frame_bytes = 0;
    uint64_t paths = 0;  // camera paths started (from the status copies)
};

// A step's status as the host reads it back: the counters after its shade,
// then the job table.
struct StatusSlot {
    int* counters = nullptr;  // pinned, S_NCOUNTERS ints
    JobDev* jobs = nullptr;   // pinned, max_jobs records
    hipEvent_t begin = nullptr, end = nullptr;
};

struct mcpt_server_s {
    Ctx* ctx = nullptr;
    int device = 0;
    hipStream_t stream = nullptr;  // plan, camera paths, closest hits, commits, shade
    hipStream_t side = nullptr;    // shadow rays and finished paths
    hipEvent_t ev_shade = nullptr, ev_fin = nullptr;
    std::chrono::steady_clock::time_point t0;
    // configuration
    int capacity = 0, max_jobs = 0, frames_opt = 0, refill = 32, max_frames = 64, policy = 0;
    int group_below = 0, shadow_group_below = 0, gcap_phys = 0;
    int extend_blocks = 1, shadow_blocks = 1, shade_blocks = 1, finish_blocks = 1, gen_blocks = 1;
    int group_blocks = 1, commit_blocks = 1;
    uint64_t film_budget = 0, film_in_use = 0;
    size_t lds = 0, glds = 0;
    // device memory
    void* pool_mem = nullptr;
    size_t pool_bytes = 0;
    Pool pool{};
    SceneView* scene_tab = nullptr;
    BvhView* bvh_tab = nullptr;
    int n_scenes = 0;
    int max_depth_scene = 1;
    std::vector<mcpt_scene> scenes;
    // pinned staging
    JobDev* h_job = nullptr;  // admission records, one per slot
    StatusSlot ring[kStatusRing];
    // host state
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<uint64_t> waiting;
    std::map<uint64_t, JobHost> jobs;
    std::vector<uint64_t> running;  // admission order
    std::vector<int> free_slots;
    uint64_t next_id = 1;
    uint32_t next_seq = 1;
    bool stop = false;
    int error = 0;
    std::string error_msg;
    std::thread worker;
    int cur = 0;       // the queue the next step's rays are in
    int64_t steps = 0; // steps enqueued
    // the closest-hit and shadow queue lengths of the last step read back
    // (the worker's own): the group kernels are launched only when these
    // were short, since a launch that finds a long queue still waits for
    // room beside the other stream's persistent kernel
    int last_rays = 0, last_shadow = 0;
    mcpt_server_stats st{};
};

namespace {

// Samples of one job in flight. A sample's frame is committed only when its
// slowest path has ended, while most paths end within a bounce or two, so
// the ring has to run far ahead of the oldest open sample to keep the pool
// full: all of the job's samples up to 64, as many as the film budget
// holds next to the film.
int frames_for(const mcpt_server_s* s, int npix, int spp) {
    int f = s->frames_opt > 0 ? s->frames_opt : s->max_frames;
    f = std::max(1, std::min(std::min(f, spp), s->max_frames));
    const uint64_t film = (uint64_t)npix * 3 * sizeof(float);
    const uint64_t fit = s->film_budget / film;  // film + frames
    if (fit >= 2 && (uint64_t)f + 1 > fit) f = (int)(fit - 1);
    return f;
}

void srv_fail(mcpt_server_s* s, const char* what, hipError_t e) {
    std::lock_guard<std::mutex> lock(s->mu);
    s->error = MCPT_ERR_HIP;
    s->error_msg = std::string(what) + ": " + hipGetErrorString(e);
    s->stop = true;
    s->cv_done.notify_all();
}

#define SRV_HIP(x)                              \
    do {                                        \
        hipError_t e_ = (x);                    \
        if (e_ != hipSuccess) {                 \
            srv_fail(s, #x, e_);                \
            return false;                       \
        }                                       \
    } while (0)

// Admits waiting jobs in FIFO order while slots and memory allow (mu held):
// the job's record, a cleared film and cleared frame counters, in stream
// order before any later plan kernel.
int srv_admit_locked(mcpt_server_s* s) {
    int admitted = 0;
    while (!s->waiting.empty() && !s->free_slots.empty()) {
        JobHost& j = s->jobs[s->waiting.front()];
        const size_t need = j.film_bytes + j.frame_bytes;
        if (s->film_in_use + need > s->film_budget) break;
        void* a = nullptr;
        void* b = nullptr;
        if (dev_alloc(s->ctx, &a, j.film_bytes) < 0) break;
        if (dev_alloc(s->ctx, &b, j.frame_bytes) < 0) {
            dev_free(s->ctx, a, j.film_bytes);
            break;
        }
        j.film = (float*)a;
        j.frames = (float*)b;
        s->film_in_use += need;
        j.slot = s->free_slots.back();
        s->free_slots.pop_back();
        j.status = MCPT_PENDING;
        j.start_ms = now_ms(s->t0);
        s->waiting.pop_front();
        s->running.push_back(j.id);
        admitted++;
        JobDev jd;
        std::memset(&jd, 0, sizeof(jd));
        jd.cam = j.job.camera;
        jd.film = j.film;
        jd.frames = j.frames;
        jd.scene = j.job.scene;
        jd.max_depth = j.job.max_depth;
        jd.npix = j.npix;
        jd.nframes = j.nframes;
        jd.seed = j.job.seed;
        jd.id = j.id;
        jd.spp = j.job.spp;
        jd.seq = s->next_seq++;
        jd.active = 1;
        // the pinned record is not rewritten before this copy has run: the
        // slot is reused only after a later status copy showed its job done
        s->h_job[j.slot] = jd;
        (void)hipMemcpyAsync(s->pool.jobs + j.slot, s->h_job + j.slot, sizeof(JobDev), hipMemcpyHostToDevice,
                             s->stream);
        k_srv_fill_f<<<blocks_for((int64_t)j.npix * 3), kBlock, 0, s->stream>>>(j.film, (size_t)j.npix * 3, 0.0f);
        k_srv_fill_i<<<1, kBlock, 0, s->stream>>>(s->pool.frame_done + j.slot * s->max_frames, s->max_frames, 0);
    }
    return admitted;
}

// Enqueues one bounce step of the pool (no host synchronisation).
bool srv_enqueue_step(mcpt_server_s* s) {
    Pool& p = s->pool;
    hipStream_t st = s->stream, sd = s->side;
    const int cur = s->cur;
    StatusSlot& rs = s->ring[s->steps % kStatusRing];
    SRV_HIP(hipEventRecord(rs.begin, st));
    k_srv_plan<<<1, 64, plan_lds_bytes(s->max_jobs), st>>>(p, s->policy);
    k_srv_generate<<<s->gen_blocks, kBlock, 0, st>>>(p, cur);
    // a queue shorter than gb goes to the group walk; the persistent
    // kernel takes every queue when no group launch follows
    const int gb = s->last_rays < 2 * s->group_below ? s->group_below : 0;
    const int gbs = s->last_shadow < 2 * s->shadow_group_below ? s->shadow_group_below : 0;
    k_srv_extend<<<s->extend_blocks, kBlock, s->lds, st>>>(p, cur, s->refill, gb);
    if (gb > 0)
        k_srv_extend_group<<<s->group_blocks, kBlock, s->glds, st>>>(p, cur, gb, kGroupCap, s->gcap_phys);
    // the previous step's shadow rays add to this queue's states and its
    // finished paths fill the frames: both before the commits and the shade
    if (s->steps > 0) SRV_HIP(hipStreamWaitEvent(st, s->ev_fin, 0));
    k_srv_commit_plan<<<1, (s->max_jobs + 63) / 64 * 64, 0, st>>>(p);
    k_srv_commit_apply<<<s->commit_blocks, kBlock, 0, st>>>(p);
    k_srv_shade<<<s->shade_blocks, kShadeBlock, 0, st>>>(p, cur);
    SRV_HIP(hipEventRecord(s->ev_shade, st));
    SRV_HIP(hipMemcpyAsync(rs.counters, p.counters, S_NCOUNTERS * sizeof(int), hipMemcpyDeviceToHost, st));
    SRV_HIP(hipMemcpyAsync(rs.jobs, p.jobs, s->max_jobs * sizeof(JobDev), hipMemcpyDeviceToHost, st));
    SRV_HIP(hipEventRecord(rs.end, st));
    SRV_HIP(hipStreamWaitEvent(sd, s->ev_shade, 0));
    k_srv_shadow_persistent<<<s->shadow_blocks, kBlock, s->lds, sd>>>(p, cur, s->refill, gbs);
    if (gbs > 0)
        k_srv_shadow_group<<<s->group_blocks, kBlock, s->glds, sd>>>(p, cur, gbs, kGroupCap, s->gcap_phys);
    k_srv_finish<<<s->finish_blocks, kFinishBlock, 0, sd>>>(p);
    SRV_HIP(hipEventRecord(s->ev_fin, sd));
    SRV_HIP(hipGetLastError());
    s->cur = 1 - cur;
    s->steps++;
    return true;
}

// Reads step `n`'s status back (waiting for it): statistics, the jobs that
// finished (released to their callers, their slots to waiting jobs).
bool srv_consume(mcpt_server_s* s, int64_t n) {
    StatusSlot& rs = s->ring[n % kStatusRing];
    SRV_HIP(hipEventSynchronize(rs.end));
    float busy = 0.f;
    SRV_HIP(hipEventElapsedTime(&busy, rs.begin, rs.end));
    std::lock_guard<std::mutex> lock(s->mu);
    s->st.iterations++;
    s->st.busy_ms += busy;
    s->st.paths_started += (uint64_t)rs.counters[S_GEN + 1];
    s->st.extend_rays += (uint64_t)rs.counters[S_QCOUNT];
    s->st.shadow_rays += (uint64_t)rs.counters[S_SHADOW];
    s->last_rays = rs.counters[S_QCOUNT];
    s->last_shadow = rs.counters[S_SHADOW];
    bool any = false;
    for (size_t r = 0; r < s->running.size();) {
        JobHost& j = s->jobs[s->running[r]];
        const JobDev& d = rs.jobs[j.slot];
        if (d.id == j.id) j.paths = (uint64_t)d.gen;  // the slot's record is this job's
        if (d.id == j.id && !d.active) {
            j.status = MCPT_OK;
            j.finish_ms = now_ms(s->t0);
            s->free_slots.push_back(j.slot);
            j.slot = -1;
            s->running.erase(s->running.begin() + r);
            s->st.jobs_done++;
            any = true;
        } else {
            r++;
        }
    }
    if (any) {
        srv_admit_locked(s);
        s->cv_done.notify_all();
    }
    return true;
}

// The worker: while jobs run, one step is enqueued ahead of the one read
// back, so the device never waits for the host's bookkeeping.
void srv_loop(mcpt_server_s* s) {
    if (hipSetDevice(s->device) != hipSuccess) {
        srv_fail(s, "hipSetDevice", hipErrorInvalidDevice);
        return;
    }
    int64_t outstanding = -1;  // a step enqueued and not yet read back
    while (true) {
        bool run = false;
        {
            std::unique_lock<std::mutex> lock(s->mu);
            srv_admit_locked(s);
            s->cv_work.wait(lock, [&] { return s->stop || !s->running.empty() || outstanding >= 0; });
            if (s->stop) break;
            run = !s->running.empty();
        }
        const int64_t n = s->steps;
        if (run && !srv_enqueue_step(s)) return;
        if (outstanding >= 0 && !srv_consume(s, outstanding)) return;
        outstanding = run ? n : -1;
    }
    if (outstanding >= 0) (void)hipEventSynchronize(s->ring[outstanding % kStatusRing].end);
}

}  // namespace

extern "C" {

int mcpt_server_create(mcpt_ctx ctx, const mcpt_scene* scenes, int32_t n_scenes, const mcpt_server_opts* opts,
                       mcpt_server* out) {
    if (!ctx || !scenes || n_scenes < 1 || !out) return set_error(MCPT_ERR_INVALID, "bad server arguments");
    *out = nullptr;
    for (int i = 0; i < n_scenes; i++) {
        if (!scenes[i] || scenes[i]->ctx != &ctx->c) return set_error(MCPT_ERR_INVALID, "scene %d is invalid", i);
        if (!scenes[i]->built) return set_error(MCPT_ERR_INVALID, "scene %d has no BVH", i);
        MCPT_TRY(scenes[i]->check_textures());
    }
    bool moving = false, medium = false;
    for (int i = 0; i < n_scenes; i++) {
        moving = moving || scenes[i]->moving_spheres;
        medium = medium || scenes[i]->has_medium;
    }
    if (moving && n_scenes > 256)
        return set_error(MCPT_ERR_INVALID, "a server with moving spheres takes at most 256 scenes");
    const int max_jobs = (opts && opts->max_jobs > 0) ? opts->max_jobs : 64;
    if (max_jobs > kMaxJobs) return set_error(MCPT_ERR_INVALID, "max_jobs %d above %d", max_jobs, kMaxJobs);
    const int policy = opts ? opts->policy : MCPT_SCHED_SHORTEST_FIRST;
    if (policy != MCPT_SCHED_SHORTEST_FIRST && policy != MCPT_SCHED_FIFO && policy != MCPT_SCHED_SHORTEST_FIRST_AGED)
        return set_error(MCPT_ERR_INVALID, "unknown scheduling policy %d", policy);
    Ctx* c = &ctx->c;
    MCPT_HIP(hipSetDevice(c->device));
    auto* s = new mcpt_server_s();
    s->ctx = c;
    s->device = c->device;
    s->t0 = std::chrono::steady_clock::now();
    s->capacity = (opts && opts->path_capacity > 0) ? opts->path_capacity : (8 << 20);
    s->max_jobs = max_jobs;
    s->frames_opt = opts ? opts->frames_per_job : 0;
    s->refill = (opts && opts->refill_lanes > 0) ? std::min(64, opts->refill_lanes) : 32;
    s->policy = policy;
    s->group_below = !opts || opts->group_below == 0 ? kSrvGroupBelow : std::max(0, opts->group_below);
    s->shadow_group_below = medium ? 0 : s->group_below;  // ratio tracking walks a medium ray whole
    s->n_scenes = n_scenes;
    s->scenes.assign(scenes, scenes + n_scenes);
    for (int i = 0; i < n_scenes; i++) s->max_depth_scene = std::max(s->max_depth_scene, scenes[i]->max_depth);
    s->lds = (size_t)kBlock * stack_entries(s->max_depth_scene) * sizeof(int);
    s->gcap_phys = group_stack_pairs(stack_entries(s->max_depth_scene));
    s->glds = (size_t)(kBlock / kGroup) * group_lds_ints(s->gcap_phys) * sizeof(int);
    // device memory: pool + tables
    const size_t C = s->capacity;
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t parts[] = {al(C * sizeof(RayItem)),    al(C * sizeof(RayItem)),  al(C * sizeof(float4)),
                            al(C * sizeof(float4)),     al(C * sizeof(float4)),   al(C * sizeof(float4)),
                            al(C * sizeof(uint64_t)),   al(C * sizeof(uint64_t)), al(C * sizeof(int4)),
                            al(C * sizeof(int4)),       al(C * sizeof(mcpt_packed_hit)), al(C * sizeof(ShadowItem)),
                            al(C * sizeof(float4)),     al(C * sizeof(int4)),     al(S_NCOUNTERS * sizeof(int)),
                            al((size_t)s->max_jobs * s->max_frames * sizeof(int)), al(s->max_jobs * sizeof(JobDev)),
                            al(s->max_jobs * sizeof(Quota)), al(n_scenes * sizeof(SceneView)),
                            al(n_scenes * sizeof(BvhView))};
    size_t total = 0;
    for (size_t x : parts) total += x;
    int stt = dev_alloc(c, &s->pool_mem, total);
    if (stt < 0) {
        delete s;
        return stt;
    }
    s->pool_bytes = total;
    char* b = (char*)s->pool_mem;
    int k = 0;
    Pool& p = s->pool;
    p.q[0] = (RayItem*)b, b += parts[k++];
    p.q[1] = (RayItem*)b, b += parts[k++];
    p.thr[0] = (float4*)b, b += parts[k++];
    p.thr[1] = (float4*)b, b += parts[k++];
    p.smp[0] = (float4*)b, b += parts[k++];
    p.smp[1] = (float4*)b, b += parts[k++];
    p.rng[0] = (uint64_t*)b, b += parts[k++];
    p.rng[1] = (uint64_t*)b, b += parts[k++];
    p.meta[0] = (int4*)b, b += parts[k++];
    p.meta[1] = (int4*)b, b += parts[k++];
    p.hits = (mcpt_packed_hit*)b, b += parts[k++];
    p.sq = (ShadowItem*)b, b += parts[k++];
    p.done_smp = (float4*)b, b += parts[k++];
    p.done_meta = (int4*)b, b += parts[k++];
    p.counters = (int*)b, b += parts[k++];
    p.frame_done = (int*)b, b += parts[k++];
    p.jobs = (JobDev*)b, b += parts[k++];
    p.quotas = (Quota*)b, b += parts[k++];
    s->scene_tab = (SceneView*)b, b += parts[k++];
    s->bvh_tab = (BvhView*)b, b += parts[k++];
    if (k != (int)(sizeof(parts) / sizeof(parts[0]))) std::abort();  // every part placed once
    p.scenes = s->scene_tab;
    p.bvhs = s->bvh_tab;
    p.max_frames = s->max_frames;
    p.max_jobs = s->max_jobs;
    p.capacity = s->capacity;
    p.tb = moving ? 24 : 0;
    hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&s->ev_shade, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&s->ev_fin, hipEventDisableTiming);
    for (int r = 0; r < kStatusRing && e == hipSuccess; r++) {
        e = hipEventCreate(&s->ring[r].begin);
        if (e == hipSuccess) e = hipEventCreate(&s->ring[r].end);
        if (e == hipSuccess)
            e = hipHostMalloc((void**)&s->ring[r].counters, S_NCOUNTERS * sizeof(int), hipHostMallocDefault);
        if (e == hipSuccess)
            e = hipHostMalloc((void**)&s->ring[r].jobs, s->max_jobs * sizeof(JobDev), hipHostMallocDefault);
    }
    // grids: the persistent and grid-stride kernels run on the occupancy
    // limit (the host does not know a step's queue lengths)
    auto per_cu = [&](auto f, int threads, size_t lds) {
        int n = 0;
        if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, f, threads, lds);
        return std::max(1, n) * c->cu_count;
    };
    s->extend_blocks = per_cu(k_srv_extend, kBlock, s->lds);
    s->shadow_blocks = per_cu(k_srv_shadow_persistent, kBlock, s->lds);
    s->shade_blocks = per_cu(k_srv_shade, kShadeBlock, 0);
    s->finish_blocks = per_cu(k_srv_finish, kFinishBlock, 0);
    s->gen_blocks = per_cu(k_srv_generate, kBlock, 0);
    s->group_blocks = std::min(per_cu(k_srv_extend_group, kBlock, s->glds),
                               (s->group_below + kBlock / kGroup - 1) / (kBlock / kGroup) + 1);
    s->commit_blocks = 2 * c->cu_count;
    std::vector<SceneView> sv(n_scenes);
    std::vector<BvhView> bv(n_scenes);
    for (int i = 0; i < n_scenes; i++) {
        sv[i] = scenes[i]->view();
        bv[i] = sv[i].bvh;
    }
    if (e == hipSuccess) e = hipMemcpy(s->scene_tab, sv.data(), n_scenes * sizeof(SceneView), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->bvh_tab, bv.data(), n_scenes * sizeof(BvhView), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipHostMalloc((void**)&s->h_job, s->max_jobs * sizeof(JobDev), hipHostMallocDefault);
    if (e == hipSuccess) {
        // counters, frame counters and the job table (every slot inactive)
        k_srv_fill_i<<<1, kBlock, 0, s->stream>>>(p.counters, S_NCOUNTERS, 0);
        k_srv_fill_i<<<blocks_for(s->max_jobs * s->max_frames), kBlock, 0, s->stream>>>(p.frame_done,
                                                                                         s->max_jobs * s->max_frames, 0);
        const int job_ints = (int)(s->max_jobs * sizeof(JobDev) / sizeof(int));
        k_srv_fill_i<<<blocks_for(job_ints), kBlock, 0, s->stream>>>((int*)p.jobs, job_ints, 0);
        e = hipStreamSynchronize(s->stream);
    }
    if (e != hipSuccess) {
        for (StatusSlot& r : s->ring) {
            if (r.counters) (void)hipHostFree(r.counters);
            if (r.jobs) (void)hipHostFree(r.jobs);
            if (r.begin) (void)hipEventDestroy(r.begin);
            if (r.end) (void)hipEventDestroy(r.end);
        }
        if (s->h_job) (void)hipHostFree(s->h_job);
        if (s->ev_shade) (void)hipEventDestroy(s->ev_shade);
        if (s->ev_fin) (void)hipEventDestroy(s->ev_fin);
        if (s->side) (void)hipStreamDestroy(s->side);
        if (s->stream) (void)hipStreamDestroy(s->stream);
        dev_free(c, s->pool_mem, s->pool_bytes);
        delete s;
        return set_error(MCPT_ERR_HIP, "server set-up: %s", hipGetErrorString(e));
    }
    for (int i = s->max_jobs - 1; i >= 0; i--) s->free_slots.push_back(i);
    {
        std::lock_guard<std::mutex> lock(c->alloc_mu);
        const uint64_t avail = c->budget > c->used ? c->budget - c->used : 0;
        s->film_budget = (opts && opts->film_budget_bytes) ? opts->film_budget_bytes : avail / 2;
    }
    s->st.film_budget = s->film_budget;
    s->worker = std::thread(srv_loop, s);
    for (mcpt_scene sc : s->scenes) sc->servers++;  // held until mcpt_server_destroy
    c->children++;
    *out = s;
    return MCPT_OK;
}

int mcpt_server_submit(mcpt_server s, const mcpt_job* job, uint64_t* job_id) {
    if (!s || !job) return set_error(MCPT_ERR_INVALID, "bad arguments");
    if (job->scene < 0 || job->scene >= s->n_scenes) return set_error(MCPT_ERR_INVALID, "scene %d unknown", job->scene);
    if (job->spp < 1 || job->max_depth < 1) return set_error(MCPT_ERR_INVALID, "spp and max_depth must be >= 1");
    const int64_t npix = (int64_t)job->camera.width * job->camera.height;
    if (npix < 1 || npix > s->capacity) return set_error(MCPT_ERR_INVALID, "image size outside the path pool");
    std::lock_guard<std::mutex> lock(s->mu);
    if (s->error) return set_error(s->error, "%s", s->error_msg.c_str());
    JobHost j;
    j.id = s->next_id++;
    j.job = *job;
    j.npix = (int)npix;
    j.nframes = frames_for(s, j.npix, job->spp);
    j.film_bytes = (size_t)npix * 3 * sizeof(float);
    j.frame_bytes = j.film_bytes * j.nframes;
    if (j.film_bytes + j.frame_bytes > s->film_budget)
        return set_error(MCPT_ERR_BUDGET, "job needs %zu bytes, more than the server's film budget %llu",
                         j.film_bytes + j.frame_bytes, (unsigned long long)s->film_budget);
    j.submit_ms = now_ms(s->t0);
    const uint64_t id = j.id;
    s->jobs[id] = j;
    s->waiting.push_back(id);
    s->st.jobs_submitted++;
    if (job_id) *job_id = id;
    // admission enqueues the job's set-up on the server stream before the
    // worker can see it among the running jobs
    MCPT_HIP(hipSetDevice(s->device));
    srv_admit_locked(s);
    s->cv_work.notify_all();
    return s->jobs[id].status == MCPT_PENDING ? MCPT_OK : MCPT_DEFERRED;
}

int mcpt_server_poll(mcpt_server s, uint64_t id, mcpt_job_result* r) {
    if (!s) return set_error(MCPT_ERR_INVALID, "server is NULL");
    std::lock_guard<std::mutex> lock(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return set_error(MCPT_ERR_INVALID, "unknown job %llu", (unsigned long long)id);
    const JobHost& j = it->second;
    if (r) {
        r->status = j.status;
        r->width = j.job.camera.width;
        r->height = j.job.camera.height;
        r->spp = j.job.spp;
        r->submit_ms = j.submit_ms;
        r->start_ms = j.start_ms;
        r->finish_ms = j.finish_ms;
        r->paths = j.paths;
    }
    if (s->error) return set_error(s->error, "%s", s->error_msg.c_str());
    return j.status;
}

int mcpt_server_wait(mcpt_server s, uint64_t id, mcpt_job_result* r, float* rgb) {
    if (!s) return set_error(MCPT_ERR_INVALID, "server is NULL");
    float* film = nullptr;
    size_t bytes = 0;
    {
        std::unique_lock<std::mutex> lock(s->mu);
        auto it = s->jobs.find(id);
        if (it == s->jobs.end()) return set_error(MCPT_ERR_INVALID, "unknown job %llu", (unsigned long long)id);
        s->cv_done.wait(lock, [&] { return s->error || s->jobs[id].status == MCPT_OK; });
        if (s->error) return set_error(s->error, "%s", s->error_msg.c_str());
        film = s->jobs[id].film;
        bytes = s->jobs[id].film_bytes;
    }
    if (rgb) {
        if (!film) return set_error(MCPT_ERR_INVALID, "job %llu was released", (unsigned long long)id);
        MCPT_HIP(hipMemcpy(rgb, film, bytes, hipMemcpyDeviceToHost));
    }
    return mcpt_server_poll(s, id, r);
}

int mcpt_server_release(mcpt_server s, uint64_t id) {
    if (!s) return set_error(MCPT_ERR_INVALID, "server is NULL");
    std::lock_guard<std::mutex> lock(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return set_error(MCPT_ERR_INVALID, "unknown job %llu", (unsigned long long)id);
    JobHost& j = it->second;
    if (j.status != MCPT_OK) return set_error(MCPT_ERR_INVALID, "job %llu is not finished", (unsigned long long)id);
    if (j.film) {
        dev_free(s->ctx, j.film, j.film_bytes);
        dev_free(s->ctx, j.frames, j.frame_bytes);
        s->film_in_use -= j.film_bytes + j.frame_bytes;
        j.film = j.frames = nullptr;
    }
    srv_admit_locked(s);  // waiting jobs may fit now
    s->cv_work.notify_all();
    return MCPT_OK;
}

int mcpt_server_stats_get(mcpt_server s, mcpt_server_stats* out) {
    if (!s || !out) return set_error(MCPT_ERR_INVALID, "bad arguments");
    std::lock_guard<std::mutex> lock(s->mu);
    *out = s->st;
    out->jobs_waiting = s->waiting.size();
    out->jobs_running = s->running.size();
    out->film_bytes_in_use = s->film_in_use;
    out->film_budget = s->film_budget;
    return MCPT_OK;
}

int mcpt_server_destroy(mcpt_server s) {
    if (!s) return MCPT_OK;
    {
        std::lock_guard<std::mutex> lock(s->mu);
        s->stop = true;
        s->cv_work.notify_all();
        s->cv_done.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    (void)hipSetDevice(s->device);
    (void)hipStreamSynchronize(s->stream);
    (void)hipStreamSynchronize(s->side);
    for (auto& kv : s->jobs) {
        if (kv.second.film) {
            dev_free(s->ctx, kv.second.film, kv.second.film_bytes);
            dev_free(s->ctx, kv.second.frames, kv.second.frame_bytes);
        }
    }
    for (StatusSlot& r : s->ring) {
        (void)hipHostFree(r.counters);
        (void)hipHostFree(r.jobs);
        (void)hipEventDestroy(r.begin);
        (void)hipEventDestroy(r.end);
    }
    (void)hipHostFree(s->h_job);
    (void)hipEventDestroy(s->ev_shade);
    (void)hipEventDestroy(s->ev_fin);
    (void)hipStreamDestroy(s->side);
    (void)hipStreamDestroy(s->stream);
    dev_free(s->ctx, s->pool_mem, s->pool_bytes);
    for (mcpt_scene sc : s->scenes) sc->servers--;
    s->ctx->children--;
    delete s;
    return MCPT_OK;
}

}  // extern "C"
