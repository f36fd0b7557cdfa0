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
// Determinism: a job's paths are keyed by (pixel, sample, seed) like
// mcpt_render's. A terminated path stores its clamped sample into the
// job's frame buffer for that sample; when all W*H paths of the oldest open
// sample are finished, a commit kernel adds the frame into the film. Films
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
constexpr int kMaxQuotas = 4096;

struct JobDev {
    mcpt_camera cam;
    float* film;    // W*H*3 linear sums
    float* frames;  // frames_per_job * W*H*3 clamped samples awaiting commit
    int scene;
    int max_depth;
    int npix;
    int nframes;
    uint64_t seed;
};

struct Quota {  // one run of new camera paths: pixels [pixel0, pixel0+count) of `sample`
    int job, sample, pixel0, count, prefix;
};

// Step counters, each in its own 128-byte line: every wave of a step takes
// slots with atomics on them, and atomics on one line serialize in one L2
// channel. S_ND holds the next-queue count (low 32 bits) and the done count
// (high 32 bits) for one combined push.
enum { S_QCOUNT = 0, S_ND = 32, S_SHADOW = 64, S_WORK = 96, S_WORK_SHADOW = 128, S_NCOUNTERS = 160 };

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
    Quota* quotas;
    const SceneView* scenes;
    const BvhView* bvhs;
    int max_frames;
    int tb;  // 24: ray records carry scene and time (see pack_scene_time)
};

// ------------------------------------------------------------------ kernels
__global__ void k_srv_begin(int* c) {
    if (threadIdx.x == 0) {
        *nd_counter(c) = 0;
        c[S_SHADOW] = 0;
        c[S_WORK] = 0;
        c[S_WORK_SHADOW] = 0;
    }
}

// camera paths for the quota runs, appended to the current queue at
// base + g (mcpt_render's camera kernel per path)
__global__ __launch_bounds__(kBlock) void k_srv_generate(Pool p, int cur, int total, int nq, int base) {
    const int g = blockIdx.x * kBlock + threadIdx.x;
    if (g == 0) p.counters[S_QCOUNT] = base + total;
    if (g >= total) return;
    int lo = 0, hi = nq - 1;  // last quota with prefix <= g
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (p.quotas[mid].prefix <= g) lo = mid;
        else hi = mid - 1;
    }
    const Quota q = p.quotas[lo];
    const JobDev& J = p.jobs[q.job];
    const int pixel = q.pixel0 + (g - q.prefix);
    Rng rng;
    mcpt_pcg32_for_path(&rng.g, (uint32_t)pixel, (uint32_t)q.sample, J.seed);
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
    p.meta[cur][slot] = make_int4(q.job, pixel, q.sample, 0);
}

// persistent extend over rays of mixed scenes with deferred leaf tests; the
// scene index rides in the ray record (ot.w) and a lane reloads the BVH
// view only when it changes
__global__ __launch_bounds__(kBlock) void k_srv_extend(Pool p, int cur, int refill) {
    extern __shared__ int lds_stack[];
    const int total = p.counters[S_QCOUNT];
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
        [&](int i, const Trav& s, const BvhView& v) {
            mcpt_packed_hit h;  // prim_index keeps the leaf's code: triangle, or ~sphere
            h.t = s.t;
            h.prim_type = s.leaf >= 0 ? 1 : -1;
            h.prim_index = s.leaf >= 0 ? v.leaf_prim[s.leaf] : -1;
            h.did_hit = s.leaf >= 0 ? 1 : 0;
            p.hits[i] = h;
        },
        ns, nl);
}

// shadeKernel (src/path_tracer.cu:3470-3696) per queue slot; survivors
// move to the other queue with their state, terminated paths to the done
// list
// Each block's slots are shaded in shading-bin order (block_bin_sort).
constexpr int kShadeBlock = 1024;
__global__ __launch_bounds__(kShadeBlock) void k_srv_shade(Pool p, int cur) {
    __shared__ int lds[bin_sort_lds_ints<kShadeBlock>()];
    const int count = p.counters[S_QCOUNT];
    int slot = blockIdx.x * kShadeBlock + threadIdx.x, mat = -1, bin = kShadeBins;
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
                     vec3(it.dp.x, it.dp.y, it.dp.z), hit.did_hit != 0, hit.t, hit.prim_index, ps, rng, out, time,
                     mat);
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
__global__ __launch_bounds__(kBlock) void k_srv_shadow_persistent(Pool p, int cur, int refill) {
    extern __shared__ int lds_stack[];
    int* stk = lds_stack + threadIdx.x;
    const int total = p.counters[S_SHADOW];
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

// terminated paths: clamped sample into the frame buffer of the path's
// sample (finalizeSampleKernel's clamp, :3740-3755), and the count of
// finished paths per frame. The terminated paths of a block nearly always
// belong to one or two frames: wave leaders add their group counts into a
// small LDS table and the block adds each table entry with one atomic.
constexpr int kFinishBlock = 1024;
constexpr int kFinishKeys = 32;
__global__ __launch_bounds__(kFinishBlock) void k_srv_finish(Pool p) {
    __shared__ int lkey[kFinishKeys], lcnt[kFinishKeys];
    if (threadIdx.x < kFinishKeys) {
        lkey[threadIdx.x] = -1;
        lcnt[threadIdx.x] = 0;
    }
    __syncthreads();
    const int d = blockIdx.x * kFinishBlock + threadIdx.x;
    const bool valid = d < (int)(*nd_counter(p.counters) >> 32);
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
    if (threadIdx.x < kFinishKeys && lcnt[threadIdx.x] > 0) atomicAdd(&p.frame_done[lkey[threadIdx.x]], lcnt[threadIdx.x]);
}

// film += frame (the sample order is the host's commit order)
__global__ __launch_bounds__(kBlock) void k_srv_commit(float* film, const float* frame, int n3) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n3) film[i] = film[i] + frame[i];
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
    // progress
    int committed = 0;  // samples added to the film
    int opened = 0;     // samples whose frame slot is in use
    int gen_sample = 0, gen_pixel = 0;
    uint64_t paths = 0;
    int64_t unstarted() const { return (int64_t)(job.spp - gen_sample) * npix - gen_pixel; }
};

struct mcpt_server_s {
    Ctx* ctx = nullptr;
    int device = 0;
    hipStream_t stream = nullptr;
    std::chrono::steady_clock::time_point t0;
    // configuration
    int capacity = 0, max_jobs = 0, frames_opt = 0, refill = 32, max_frames = 64, policy = 0;
    hipEvent_t ev[2] = {nullptr, nullptr};
    int extend_blocks = 1;  // resident blocks of the persistent extend
    std::vector<JobHost*> order;
    uint64_t film_budget = 0, film_in_use = 0;
    size_t lds = 0;
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
    Quota* h_quotas = nullptr;
    int* h_status = nullptr;  // counters + frame_done
    JobDev* h_job = nullptr;
    // host state
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<uint64_t> waiting;
    std::map<uint64_t, JobHost> jobs;
    std::vector<uint64_t> running;  // admission order
    std::vector<int> free_slots;
    uint64_t next_id = 1;
    bool stop = false;
    int error = 0;
    std::string error_msg;
    std::thread worker;
    // stream state
    int cur = 0, cur_count = 0;  // current queue and its length (paths in flight)
    mcpt_server_stats st{};
};

namespace {

int srv_admit_locked(mcpt_server_s* s);

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
            return;                             \
        }                                       \
    } while (0)

// admit waiting jobs in FIFO order while slots and memory allow (mu held)
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
        // device record and cleared film / frame counters (stream order)
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
        s->h_job[j.slot] = jd;
        (void)hipMemcpyAsync(s->pool.jobs + j.slot, s->h_job + j.slot, sizeof(JobDev), hipMemcpyHostToDevice,
                             s->stream);
        k_srv_fill_f<<<blocks_for((int64_t)j.npix * 3), kBlock, 0, s->stream>>>(j.film, (size_t)j.npix * 3, 0.0f);
        k_srv_fill_i<<<1, kBlock, 0, s->stream>>>(s->pool.frame_done + j.slot * s->max_frames, s->max_frames, 0);
    }
    return admitted;
}

// one bounce step of the pool
void srv_step(mcpt_server_s* s) {
    Pool& p = s->pool;
    hipStream_t st = s->stream;
    // ---- plan generation into the free part of the current queue
    int free = s->capacity - s->cur_count;
    int nq = 0, total = 0;
    {
        std::lock_guard<std::mutex> lock(s->mu);
        std::vector<JobHost*>& order = s->order;
        order.clear();
        for (uint64_t id : s->running) order.push_back(&s->jobs[id]);
        // the oldest running job (running holds admission order) and its
        // reserve: half of this step's free slots under the aged policy
        JobHost* oldest = order.empty() ? nullptr : order.front();
        int reserve = s->policy == MCPT_SCHED_SHORTEST_FIRST_AGED ? free / 2 : 0;
        if (s->policy != MCPT_SCHED_FIFO)  // least paths left to start first, ties by admission
            std::stable_sort(order.begin(), order.end(), [](const JobHost* a, const JobHost* b) {
                return a->unstarted() < b->unstarted();
            });
        if (reserve > 0 && oldest) {
            order.erase(std::find(order.begin(), order.end(), oldest));
            order.insert(order.begin(), oldest);
        }
        // new paths of job j into at most `cap` free slots, in its sample order
        auto take = [&](JobHost& j, int cap) {
            while (j.opened < j.job.spp && j.opened - j.committed < j.nframes) j.opened++;  // open frames
            while (free > 0 && cap > 0 && nq < kMaxQuotas && j.gen_sample < j.opened) {
                const int n = std::min(j.npix - j.gen_pixel, std::min(free, cap));
                s->h_quotas[nq] = Quota{j.slot, j.gen_sample, j.gen_pixel, n, total};
                nq++;
                total += n;
                free -= n;
                cap -= n;
                j.paths += n;
                j.gen_pixel += n;
                if (j.gen_pixel == j.npix) {
                    j.gen_sample++;
                    j.gen_pixel = 0;
                }
            }
        };
        // the oldest job first takes at most its reserve; the other slots go
        // to the jobs in policy order, the oldest included again
        for (JobHost* jp : order) take(*jp, (reserve > 0 && jp == oldest) ? reserve : free);
        if (reserve > 0 && free > 0) take(*oldest, free);
    }
    if (s->cur_count == 0 && total == 0) return;
    const int cur = s->cur;
    hipEvent_t e0 = s->ev[0], e1 = s->ev[1];
    SRV_HIP(hipEventRecord(e0, st));
    k_srv_begin<<<1, 64, 0, st>>>(p.counters);
    if (nq) SRV_HIP(hipMemcpyAsync(p.quotas, s->h_quotas, nq * sizeof(Quota), hipMemcpyHostToDevice, st));
    k_srv_generate<<<std::max(1, blocks_for(total)), kBlock, 0, st>>>(p, cur, total, nq, s->cur_count);
    const int inflight = s->cur_count + total;
    const int grid = blocks_for(inflight);
    k_srv_extend<<<std::min(grid, s->extend_blocks), kBlock, s->lds, st>>>(p, cur, s->refill);
    k_srv_shade<<<(inflight + kShadeBlock - 1) / kShadeBlock, kShadeBlock, 0, st>>>(p, cur);
    k_srv_shadow_persistent<<<std::min(grid, s->extend_blocks), kBlock, s->lds, st>>>(p, cur, s->refill);
    k_srv_finish<<<(inflight + kFinishBlock - 1) / kFinishBlock, kFinishBlock, 0, st>>>(p);
    SRV_HIP(hipGetLastError());
    const size_t status_bytes = (S_NCOUNTERS + (size_t)s->max_jobs * s->max_frames) * sizeof(int);
    SRV_HIP(hipMemcpyAsync(s->h_status, p.counters, S_NCOUNTERS * sizeof(int), hipMemcpyDeviceToHost, st));
    SRV_HIP(hipMemcpyAsync(s->h_status + S_NCOUNTERS, p.frame_done, status_bytes - S_NCOUNTERS * sizeof(int),
                           hipMemcpyDeviceToHost, st));
    SRV_HIP(hipEventRecord(e1, st));
    SRV_HIP(hipEventSynchronize(e1));
    float busy = 0.f;
    SRV_HIP(hipEventElapsedTime(&busy, e0, e1));
    const int next_count = s->h_status[S_ND];  // low half of the combined next/done counter
    // ---- bookkeeping
    std::vector<uint64_t> finished;
    {
        std::lock_guard<std::mutex> lock(s->mu);
        s->st.iterations++;
        s->st.busy_ms += busy;
        s->st.paths_started += total;
        s->st.extend_rays += inflight;
        s->st.shadow_rays += s->h_status[S_SHADOW];
        s->cur_count = next_count;
        s->cur = 1 - cur;
        const int* fd = s->h_status + S_NCOUNTERS;
        for (uint64_t id : s->running) {
            JobHost& j = s->jobs[id];
            // commit the oldest open samples whose every path has finished
            while (j.committed < j.opened) {
                const int fslot = j.committed % j.nframes;
                if (fd[j.slot * s->max_frames + fslot] < j.npix) break;
                k_srv_commit<<<blocks_for((int64_t)j.npix * 3), kBlock, 0, st>>>(
                    j.film, j.frames + (size_t)fslot * j.npix * 3, j.npix * 3);
                k_srv_fill_i<<<1, 64, 0, st>>>(p.frame_done + j.slot * s->max_frames + fslot, 1, 0);
                j.committed++;
            }
            if (j.committed == j.job.spp) finished.push_back(id);
        }
    }
    if (!finished.empty()) {
        SRV_HIP(hipStreamSynchronize(st));
        std::lock_guard<std::mutex> lock(s->mu);
        for (uint64_t id : finished) {
            JobHost& j = s->jobs[id];
            j.status = MCPT_OK;
            j.finish_ms = now_ms(s->t0);
            s->free_slots.push_back(j.slot);
            j.slot = -1;
            s->running.erase(std::find(s->running.begin(), s->running.end(), id));
            s->st.jobs_done++;
        }
        srv_admit_locked(s);
        s->cv_done.notify_all();
    }
}

void srv_loop(mcpt_server_s* s) {
    if (hipSetDevice(s->device) != hipSuccess) {
        srv_fail(s, "hipSetDevice", hipErrorInvalidDevice);
        return;
    }
    while (true) {
        {
            std::unique_lock<std::mutex> lock(s->mu);
            srv_admit_locked(s);
            s->cv_work.wait(lock, [&] { return s->stop || !s->running.empty(); });
            if (s->stop) break;
        }
        srv_step(s);
    }
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
    bool moving = false;
    for (int i = 0; i < n_scenes; i++) moving = moving || scenes[i]->moving_spheres;
    if (moving && n_scenes > 256)
        return set_error(MCPT_ERR_INVALID, "a server with moving spheres takes at most 256 scenes");
    Ctx* c = &ctx->c;
    MCPT_HIP(hipSetDevice(c->device));
    auto* s = new mcpt_server_s();
    s->ctx = c;
    s->device = c->device;
    s->t0 = std::chrono::steady_clock::now();
    s->capacity = (opts && opts->path_capacity > 0) ? opts->path_capacity : (8 << 20);
    s->max_jobs = (opts && opts->max_jobs > 0) ? opts->max_jobs : 64;
    s->frames_opt = opts ? opts->frames_per_job : 0;
    s->refill = (opts && opts->refill_lanes > 0) ? std::min(64, opts->refill_lanes) : 32;
    s->policy = opts ? opts->policy : MCPT_SCHED_SHORTEST_FIRST;
    if (s->policy != MCPT_SCHED_SHORTEST_FIRST && s->policy != MCPT_SCHED_FIFO &&
        s->policy != MCPT_SCHED_SHORTEST_FIRST_AGED) {
        delete s;
        return set_error(MCPT_ERR_INVALID, "unknown scheduling policy %d", opts->policy);
    }
    s->n_scenes = n_scenes;
    s->scenes.assign(scenes, scenes + n_scenes);
    for (int i = 0; i < n_scenes; i++) s->max_depth_scene = std::max(s->max_depth_scene, scenes[i]->max_depth);
    s->lds = (size_t)kBlock * stack_entries(s->max_depth_scene) * sizeof(int);
    // device memory: pool + tables
    const size_t C = s->capacity;
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t parts[] = {al(C * sizeof(RayItem)),    al(C * sizeof(RayItem)),  al(C * sizeof(float4)),
                            al(C * sizeof(float4)),     al(C * sizeof(float4)),   al(C * sizeof(float4)),
                            al(C * sizeof(uint64_t)),   al(C * sizeof(uint64_t)), al(C * sizeof(int4)),
                            al(C * sizeof(int4)),       al(C * sizeof(mcpt_packed_hit)), al(C * sizeof(ShadowItem)),
                            al(C * sizeof(float4)),     al(C * sizeof(int4)),     al(S_NCOUNTERS * sizeof(int)),
                            al((size_t)s->max_jobs * s->max_frames * sizeof(int)), al(s->max_jobs * sizeof(JobDev)),
                            al(kMaxQuotas * sizeof(Quota)), al(n_scenes * sizeof(SceneView)),
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
    p.tb = moving ? 24 : 0;
    hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&s->ev[0]);
    if (e == hipSuccess) e = hipEventCreate(&s->ev[1]);
    int per_cu = 0;
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_srv_extend, kBlock, s->lds);
    s->extend_blocks = std::max(1, per_cu) * c->cu_count;
    std::vector<SceneView> sv(n_scenes);
    std::vector<BvhView> bv(n_scenes);
    for (int i = 0; i < n_scenes; i++) {
        sv[i] = scenes[i]->view();
        bv[i] = sv[i].bvh;
    }
    if (e == hipSuccess) e = hipMemcpy(s->scene_tab, sv.data(), n_scenes * sizeof(SceneView), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->bvh_tab, bv.data(), n_scenes * sizeof(BvhView), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipHostMalloc((void**)&s->h_quotas, kMaxQuotas * sizeof(Quota), hipHostMallocDefault);
    if (e == hipSuccess)
        e = hipHostMalloc((void**)&s->h_status, (S_NCOUNTERS + (size_t)s->max_jobs * s->max_frames) * sizeof(int),
                          hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void**)&s->h_job, s->max_jobs * sizeof(JobDev), hipHostMallocDefault);
    if (e == hipSuccess) {
        k_srv_fill_i<<<1, 64, 0, s->stream>>>(p.counters, S_NCOUNTERS, 0);
        k_srv_fill_i<<<blocks_for(s->max_jobs * s->max_frames), kBlock, 0, s->stream>>>(p.frame_done,
                                                                                         s->max_jobs * s->max_frames, 0);
        e = hipStreamSynchronize(s->stream);
    }
    if (e != hipSuccess) {
        if (s->h_quotas) (void)hipHostFree(s->h_quotas);
        if (s->h_status) (void)hipHostFree(s->h_status);
        if (s->h_job) (void)hipHostFree(s->h_job);
        for (hipEvent_t ev : s->ev)
            if (ev) (void)hipEventDestroy(ev);
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
    for (auto& kv : s->jobs) {
        if (kv.second.film) {
            dev_free(s->ctx, kv.second.film, kv.second.film_bytes);
            dev_free(s->ctx, kv.second.frames, kv.second.frame_bytes);
        }
    }
    (void)hipHostFree(s->h_quotas);
    (void)hipHostFree(s->h_status);
    (void)hipHostFree(s->h_job);
    (void)hipEventDestroy(s->ev[0]);
    (void)hipEventDestroy(s->ev[1]);
    (void)hipStreamDestroy(s->stream);
    dev_free(s->ctx, s->pool_mem, s->pool_bytes);
    for (mcpt_scene sc : s->scenes) sc->servers--;
    s->ctx->children--;
    delete s;
    return MCPT_OK;
}

}  // extern "C"
