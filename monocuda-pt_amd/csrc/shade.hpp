// shade.hpp — per-path light transport shared by the wavefront shade kernel
// and the megakernel. Same formulas and the same random-draw order as the
// reference's trace_sample_kernel (src/path_tracer.cu:2900-3059) and
// shadeKernel (:3470-3696), so both execution models produce the same bits.
#pragma once

#include "traverse.hpp"

namespace mcpt {

struct Rng {
    mcpt_pcg32 g;
    MCPT_DEV float u() { return mcpt_pcg32_uniform(&g); }
};

// ------------------------------------------------------------- sampling
MCPT_DEV vec3 random_unit_vector(Rng& r) {  // :498-509
    while (true) {
        float x = r.u() * 2.f - 1.f;
        float y = r.u() * 2.f - 1.f;
        float z = r.u() * 2.f - 1.f;
        vec3 v(x, y, z);
        if (v.length_squared() < 1.0f) return v.normalize();
    }
}
MCPT_DEV vec3 lambertian_dir(const vec3& n, Rng& r) {  // :511-518
    vec3 d = n + random_unit_vector(r);
    if (d.length_squared() < 1e-8f) return n;
    return d.normalize();
}
MCPT_DEV float schlick(float c, float eta) {  // :476-480
    float r0 = (1.f - eta) / (1.f + eta);
    r0 = r0 * r0;
    return r0 + (1.f - r0) * mf_pow((1 - c), 5.f);
}
MCPT_DEV vec3 dielectric_dir(const vec3& in, const vec3& n, float eta, Rng& r) {  // :524-540
    vec3 ud = in.normalize();
    float c = fminf((-ud).dot(n), 1.0f);
    float sn = __builtin_sqrtf(1.0f - c * c);
    bool cannot = (eta * sn) > 1.0f;
    float pr = cannot ? 1.0f : schlick(c, eta);
    if (r.u() < pr) return ud.reflect(n);
    return ud.refract(n, eta);
}
MCPT_DEV vec3 isotropic_dir(Rng& r) {  // :542-551
    float u1 = r.u();
    float u2 = r.u();
    float z = 1.0f - 2.0f * u1;
    float rr = __builtin_sqrtf(fmaxf(0.0f, 1.0f - z * z));
    float phi = 2.0f * MCPT_PI_F * u2;
    return vec3(rr * mf_cos(phi), rr * mf_sin(phi), z);
}
MCPT_DEV vec3 hg_dir(const vec3& in, float g, Rng& r) {  // :556-578
    if (fabsf(g) < 1e-3f) return isotropic_dir(r);
    float u1 = r.u();
    float u2 = r.u();
    float sq = (1.0f - g * g) / (1.0f + g - 2.0f * g * u1);
    float ct = (1.0f + g * g - sq * sq) / (2.0f * g);
    ct = fminf(fmaxf(ct, -1.0f), 1.0f);
    float st = __builtin_sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
    float phi = 2.0f * MCPT_PI_F * u2;
    vec3 fw = in.normalize();
    vec3 helper = (fabsf(fw.x) > 0.9f) ? vec3(0.f, 1.f, 0.f) : vec3(1.f, 0.f, 0.f);
    vec3 tg = helper.cross(fw).normalize();
    vec3 bt = fw.cross(tg);
    return tg * (st * mf_cos(phi)) + bt * (st * mf_sin(phi)) + fw * ct;
}
MCPT_DEV vec3 sample_triangle(const mcpt_tri_hot& t, Rng& r) {  // :744-755
    float r1 = r.u();
    float r2 = r.u();
    float s1 = __builtin_sqrtf(r1);
    float w0 = 1.0f - s1;
    float w1 = r2 * s1;
    float w2 = 1.0f - w0 - w1;
    return load3(t.v0) * w0 + load3(t.v1) * w1 + load3(t.v2) * w2;
}
MCPT_DEV float power_heuristic(float a, float b) {  // :757-761
    float a2 = a * a, b2 = b * b;
    return a2 / (a2 + b2);
}
MCPT_DEV float lambert_pdf(float c) { return c / MCPT_PI_F; }  // :763-765
MCPT_DEV float hg_pdf(float ct, float g) {                      // :775-780
    float g2 = g * g;
    float den = fmaxf(1.0f + g2 - 2.0f * g * ct, 1e-6f);
    return (1.0f - g2) / (4.0f * MCPT_PI_F * den * __builtin_sqrtf(den));
}

// ---------------------------------------------------------- environment
MCPT_DEV void env_uv(const EnvView& e, const vec3& ud, float& u, float& v) {  // :408-412
    float uu = mf_atan2(ud.z, ud.x) / (2.0f * MCPT_PI_F) + 0.5f + e.rot / (2.0f * MCPT_PI_F);
    v = mf_acos(fminf(fmaxf(ud.y, -1.0f), 1.0f)) / MCPT_PI_F;
    u = uu - floorf(uu);
}
MCPT_DEV vec3 env_direction(const EnvView& e, float u, float v) {  // :414-420
    float phi = (u - 0.5f) * 2.0f * MCPT_PI_F - e.rot;
    float theta = v * MCPT_PI_F;
    float st = mf_sin(theta);
    return vec3(mf_cos(phi) * st, mf_cos(theta), mf_sin(phi) * st);
}
// bilinear fetch with wrap in u and clamp in v, full FP32 weights (the
// reference uses the texture unit's filtering, :387-406, :424)
MCPT_DEV vec3 env_fetch(const EnvView& e, float u, float v) {
    float x = u * (float)e.w - 0.5f;
    float y = v * (float)e.h - 0.5f;
    float x0 = floorf(x), y0 = floorf(y);
    float fx = x - x0, fy = y - y0;
    int ix0 = (int)x0, iy0 = (int)y0;
    int ix1 = ix0 + 1, iy1 = iy0 + 1;
    ix0 = ((ix0 % e.w) + e.w) % e.w;
    ix1 = ((ix1 % e.w) + e.w) % e.w;
    iy0 = min(max(iy0, 0), e.h - 1);
    iy1 = min(max(iy1, 0), e.h - 1);
    float4 t00 = e.texels[(size_t)iy0 * e.w + ix0];
    float4 t10 = e.texels[(size_t)iy0 * e.w + ix1];
    float4 t01 = e.texels[(size_t)iy1 * e.w + ix0];
    float4 t11 = e.texels[(size_t)iy1 * e.w + ix1];
    float gx = 1.0f - fx, gy = 1.0f - fy;
    return vec3((t00.x * gx + t10.x * fx) * gy + (t01.x * gx + t11.x * fx) * fy,
                (t00.y * gx + t10.y * fx) * gy + (t01.y * gx + t11.y * fx) * fy,
                (t00.z * gx + t10.z * fx) * gy + (t01.z * gx + t11.z * fx) * fy);
}
MCPT_DEV vec3 env_radiance(const EnvView& e, const vec3& d, const vec3& fallback) {  // :422-427
    if (!e.enabled) return fallback;
    float u, v;
    env_uv(e, d.normalize(), u, v);
    return env_fetch(e, u, v) * e.intensity;
}
MCPT_DEV float env_select_prob(const EnvView& e, int nl) {  // :429-432
    if (!e.enabled || e.total <= 0.0f) return 0.0f;
    return (nl > 0) ? 0.5f : 1.0f;
}
MCPT_DEV int find_cdf_interval(const float* cdf, int count, float value) {  // :434-443
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (cdf[mid + 1] <= value) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}
MCPT_DEV float env_texel_pdf(const EnvView& e, int row, int col, float st) {  // :445-450
    if (st < 1e-6f) return 0.0f;
    const float* cdf = e.col_cdf + (size_t)row * (e.w + 1);
    float tw = cdf[col + 1] - cdf[col];
    return tw * e.w * e.h / (e.total * 2.0f * MCPT_PI_F * MCPT_PI_F * st);
}
MCPT_DEV float env_dir_pdf(const EnvView& e, const vec3& d) {  // :452-458
    vec3 ud = d.normalize();
    float u, v;
    env_uv(e, ud, u, v);
    int col = min((int)(u * e.w), e.w - 1);
    int row = min((int)(v * e.h), e.h - 1);
    return env_texel_pdf(e, row, col, __builtin_sqrtf(fmaxf(1.0f - ud.y * ud.y, 0.0f)));
}
MCPT_DEV vec3 env_sample(const EnvView& e, Rng& r, float& pdf) {  // :460-467
    int row = find_cdf_interval(e.row_cdf, e.h, r.u() * e.total);
    const float* cdf = e.col_cdf + (size_t)row * (e.w + 1);
    int col = find_cdf_interval(cdf, e.w, r.u() * cdf[e.w]);
    float u = (col + r.u()) / e.w;
    float v = (row + r.u()) / e.h;
    pdf = env_texel_pdf(e, row, col, mf_sin(v * MCPT_PI_F));
    return env_direction(e, u, v);
}

// ------------------------------------------------------ hit reconstruction
struct HitInfo {
    vec3 p, normal;
    bool front;
    float u, v;
};
// the fields testPrimitiveHit fills (:1490-1533), rebuilt from the 16-byte
// hit record as shadeKernel does (:3508-3535). prim >= 0 is a triangle,
// prim < 0 the sphere ~prim, whose centre is taken at the ray's time.
// need_uv: the caller reads h.u, h.v (a textured material). A triangle
// without vertex normals whose uv is not needed is shaded from its vertices
// alone: its shading normal is the face normal, and the cold record and the
// barycentric weights are skipped (h.u = h.v = 0).
MCPT_DEV void rebuild_hit(const SceneView& sv, int prim, const vec3& o, const vec3& d, float t, float time,
                          HitInfo& h, bool need_uv = true) {
    if (prim < 0) {
        const mcpt_sphere& sp = sv.spheres[~prim];
        const vec3 c = sphere_center(load3(sp.center), load3(sp.center1), sp.moving != 0, time);
        h.p = o + d * t;
        const vec3 on = (h.p - c).normalize();
        h.front = d.dot(on) < 0.0f;
        h.normal = h.front ? on : -on;
        sphere_uv(on, h.u, h.v);
        return;
    }
    vec3 v0, v1, v2;
    load_tri(sv.hot, prim, v0, v1, v2);
    h.p = o + d * t;
    vec3 n = (v1 - v0).cross(v2 - v0).normalize();
    h.front = d.dot(n) < 0.0f;
    if (!need_uv && !(sv.tri_info[prim] & TRI_VERTEX_NORMALS)) {
        h.normal = h.front ? n : -n;
        h.u = h.v = 0.0f;
        return;
    }
    const mcpt_tri_cold& c = sv.cold[prim];
    // computeTriangleShadingNormal (:698-713)
    vec3 sn = n;
    float a2 = (v1 - v0).cross(v2 - v0).length();
    float w0 = (v1 - h.p).cross(v2 - h.p).length() / a2;
    float w1 = (v0 - h.p).cross(v2 - h.p).length() / a2;
    float w2 = (v1 - h.p).cross(v0 - h.p).length() / a2;
    vec3 n0 = load3(c.n0);
    if (!(n0.length_squared() == 0.0f)) {
        float ws = w0 + w1 + w2;
        if ((ws > 0.0f) && !(ws > 1.01f)) {
            vec3 s = n0 * w0 + load3(c.n1) * w1 + load3(c.n2) * w2;
            if (!(s.length_squared() < 1e-12f)) {
                s = s.normalize();
                if (s.dot(n) < 0.0f) s = -s;
                sn = s;
            }
        }
    }
    h.normal = h.front ? sn : -sn;
    // computeTriangleUvFromHitPoint (:687-696)
    h.u = ((c.uv0[0] * w0) + (c.uv1[0] * w1)) + (c.uv2[0] * w2);
    h.v = ((c.uv0[1] * w0) + (c.uv1[1] * w1)) + (c.uv2[1] * w2);
}

// the material of a primitive in the hit encoding above
MCPT_DEV int prim_material(const SceneView& sv, int prim) {
    return prim >= 0 ? (int)(sv.tri_info[prim] & TRI_MATERIAL_MASK) : sv.spheres[~prim].material;
}

MCPT_DEV float tri_area(const float* hot, int prim) {  // :837-839
    vec3 v0, v1, v2;
    load_tri(hot, prim, v0, v1, v2);
    return 0.5f * (v1 - v0).cross(v2 - v0).length();
}

// heterogeneousFreeFlightDistance (:796-812): delta tracking against the
// majorant sigma * max_density; a tentative collision is real with
// probability density / max_density. Returns maxd when the ray leaves
// without a real collision (or after 8192 steps).
MCPT_DEV float grid_free_flight(const GridView& g, float sigma, const vec3& o, const vec3& d, float maxd, Rng& rng) {
    const float maj = sigma * g.max_density;
    if (!(maj > 0.0f)) return maxd;
    float t = 0.0f;
    for (int step = 0; step < 8192; step++) {
        t = t - mf_log(1.0f - rng.u()) / maj;
        if (t >= maxd) return maxd;
        const float rho = grid_density(g, o + d * t);
        if (rng.u() < rho / g.max_density) return t;
    }
    return maxd;
}

// heterogeneousTransmittance (:817-835): ratio tracking, each tentative
// collision scales the estimate by 1 - density / max_density; below 1e-4
// the segment counts as opaque.
MCPT_DEV float grid_transmittance(const GridView& g, float sigma, const vec3& o, const vec3& d, float len, Rng& rng) {
    const float maj = sigma * g.max_density;
    if (!(maj > 0.0f)) return 1.0f;
    float tr = 1.0f, t = 0.0f;
    for (int step = 0; step < 8192; step++) {
        t = t - mf_log(1.0f - rng.u()) / maj;
        if (t >= len) break;
        const float rho = grid_density(g, o + d * t);
        tr = tr * (1.0f - rho / g.max_density);
        if (tr < 1e-4f) return 0.0f;
    }
    return tr;
}

// mediumTransmittance (:2874-2898): closest hits along the segment; each
// stretch that ends on a back face is inside a medium and contributes
// exp(-sigma d), or with a density grid a ratio-tracking estimate whose
// numbers come from the shadow ray's own stream (mcpt_pcg32_for_shadow).
MCPT_DEV float medium_transmittance(const SceneView& sv, vec3 o, const vec3& d, float maxd, int* stk, int stride,
                                    float time) {
    float tr = 1.0f;
    float travelled = 0.0f;
    Rng srng;
    if (sv.grid.data) {
        const uint32_t bits[7] = {f2u(o.x), f2u(o.y), f2u(o.z), f2u(d.x), f2u(d.y), f2u(d.z), f2u(maxd)};
        mcpt_pcg32_for_shadow(&srng.g, bits);
    }
    for (int k = 0; k < 4; k++) {
        float remaining = maxd - travelled;
        if (remaining <= 0.0f) break;
        Trav s;
        trav_begin(s, o, d, remaining);
        s.time = time;
        uint32_t ns = 0, nl = 0;
        trace_closest<false>(sv.bvh, s, stk, stride, ns, nl);
        if (s.leaf < 0) break;
        int prim = sv.bvh.leaf_prim[s.leaf];
        const mcpt_material& m = sv.mats[prim_material(sv, prim)];
        if (m.type != MCPT_MAT_ISOTROPIC) return 0.0f;
        bool front;
        if (prim < 0) {  // the sphere's outward normal at the hit
            const mcpt_sphere& sp = sv.spheres[~prim];
            const vec3 c = sphere_center(load3(sp.center), load3(sp.center1), sp.moving != 0, time);
            front = d.dot(((o + d * s.t) - c).normalize()) < 0.0f;
        } else {
            vec3 v0, v1, v2;
            load_tri(sv.hot, prim, v0, v1, v2);
            vec3 n = (v1 - v0).cross(v2 - v0).normalize();
            front = d.dot(n) < 0.0f;
        }
        if (!front)
            tr = tr * (sv.grid.data ? grid_transmittance(sv.grid, m.density, o, d, s.t, srng) : mf_exp(-m.density * s.t));
        travelled = travelled + s.t;
        o = o + d * s.t;
    }
    return tr;
}

// any-hit visibility, or transmittance when the scene holds a medium
// (shadowKernel, :3698-3728)
MCPT_DEV float shadow_transmittance(const SceneView& sv, const vec3& o, const vec3& d, float maxd, int* stk,
                                    int stride, float time = 0.0f) {
    if (sv.has_medium) return medium_transmittance(sv, o, d, maxd, stk, stride, time);
    Trav s;
    trav_begin(s, o, d, maxd);
    s.time = time;
    return trace_any(sv.bvh, s, stk, stride) ? 0.0f : 1.0f;
}

// ------------------------------------------------------------ one bounce
struct PathState {
    vec3 thr;        // throughput ("attenuation")
    vec3 smp;        // radiance gathered so far
    float prev_pdf;  // previousBsdfPdf
    bool prev_spec;  // previousBounceWasSpecular
};

struct Bounce {
    bool next;  // the path continues with (no, nd)
    vec3 no, nd;
    bool shadow;  // a shadow ray (so, sd, smax) carries `contrib`
    vec3 so, sd;
    float smax;
    vec3 contrib;
};

// ------------------------------------------------------- shading bins
// Lanes of one wave run every branch any of them takes, and the branches of
// a bounce differ by material: only diffuse surfaces and media sample a
// light, mirrors and glass each scatter their own way. Shade kernels
// therefore order each block's queue slots by bin before shading them, so
// most waves take one branch. The order of shading does not change any
// path's result (every path draws from its own stream and owns its state).
constexpr int kShadeBins = 5;  // miss, emitter, light-sampling surface or medium, conductor, dielectric
MCPT_DEV int shade_bin(const SceneView& sv, bool did_hit, int prim, int& mat) {
    if (!did_hit) {
        mat = -1;
        return 0;
    }
    mat = prim_material(sv, prim);
    const int type = sv.mats[mat].type;
    return type == MCPT_MAT_EMISSIVE ? 1 : type == MCPT_MAT_CONDUCTOR ? 3 : type == MCPT_MAT_DIELECTRIC ? 4 : 2;
}

// LDS ints block_bin_sort<B> needs
template <int B>
constexpr int bin_sort_lds_ints() {
    return 2 * kShadeBins * (B / 64) + 2 * B;
}

// Stable block-wide counting sort of (item, aux) pairs by bin (bin =
// kShadeBins marks a thread without an item). Afterwards thread i holds the
// i-th pair in (bin, thread) order; threads past the number of items get
// item = -1. Every thread of the block must call it.
template <int B>
MCPT_DEV void block_bin_sort(int bin, int& item, int& aux, int* lds) {
    constexpr int W = B / 64;
    int* cnt = lds;                      // [kShadeBins][W]
    int* off = lds + kShadeBins * W;     // [kShadeBins][W]
    int* s_item = lds + 2 * kShadeBins * W;
    int* s_aux = s_item + B;
    const int wave = (int)threadIdx.x >> 6;
    int rank = 0;
#pragma unroll
    for (int b = 0; b < kShadeBins; b++) {
        const uint64_t m = ballot64(bin == b);
        if (bin == b) rank = rank_below(m);
        if (lane_id() == 0) cnt[b * W + wave] = __popcll(m);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int i = 0; i < kShadeBins * W; i++) {
            off[i] = acc;
            acc += cnt[i];
        }
        cnt[0] = acc;  // the number of items, read below
    }
    __syncthreads();
    if (bin < kShadeBins) {
        const int at = off[bin * W + wave] + rank;
        s_item[at] = item;
        s_aux[at] = aux;
    }
    __syncthreads();
    const int n = cnt[0];
    item = (int)threadIdx.x < n ? s_item[threadIdx.x] : -1;
    aux = (int)threadIdx.x < n ? s_aux[threadIdx.x] : -1;
    __syncthreads();  // the arrays are free for the next call
}

// Everything one bounce does after the closest hit is known: the miss and
// emission terms, free flight in a medium, albedo, next-event estimation,
// Russian roulette and the scatter (:2917-3048, :3486-3693).
// mat is the primitive's material index when the caller has it (-1: look
// it up).
MCPT_DEV void shade_bounce(const SceneView& sv, const vec3& bg, float rr_cap, int depth, const vec3& ro,
                           const vec3& rd, bool did_hit, float t, int prim, PathState& ps, Rng& rng, Bounce& out,
                           float time = 0.0f, int mat = -1) {
    out.next = false;
    out.shadow = false;
    const int nl = sv.n_lights;
    if (!did_hit) {
        float w = 1.0f;
        float pe = env_select_prob(sv.env, nl);
        if (!(depth == 0 || ps.prev_spec) && pe > 0.0f) w = power_heuristic(ps.prev_pdf, pe * env_dir_pdf(sv.env, rd));
        ps.smp = ps.smp + ps.thr * env_radiance(sv.env, rd, bg) * w;
        return;
    }
    const mcpt_material m = sv.mats[mat >= 0 ? mat : prim_material(sv, prim)];
    HitInfo h;
    rebuild_hit(sv, prim, ro, rd, t, time, h, m.texture > 0);
    if (m.type == MCPT_MAT_EMISSIVE) {
        float ew = 1.0f;
        if (!(depth == 0 || ps.prev_spec) && prim >= 0) {  // spheres are not on the light list
            float cl = fmaxf(fabsf(rd.dot(h.normal)), 1e-4f);
            float area = tri_area(sv.hot, prim);
            float lp = ((1.0f - env_select_prob(sv.env, nl)) / float(nl)) * (t * t) / (cl * area);
            ew = power_heuristic(ps.prev_pdf, lp);
        }
        ps.smp = ps.smp + ps.thr * load3(m.emission) * ew;
        return;
    }
    vec3 so = h.p;
    if (m.type == MCPT_MAT_ISOTROPIC) {
        float sd;
        if (h.front) sd = t;  // this hit only enters the medium
        else if (sv.grid.data) sd = grid_free_flight(sv.grid, m.density, ro, rd, t, rng);
        else sd = -mf_log(1.0f - rng.u()) / m.density;
        if (sd >= t) {  // pass through the boundary, same direction
            out.next = true;
            out.no = h.p;
            out.nd = rd;
            return;
        }
        so = ro + rd * sd;
    }
    ps.thr = ps.thr * sample_albedo(sv, m, h.u, h.v);

    const float pe = env_select_prob(sv.env, nl);
    if ((m.type == MCPT_MAT_LAMBERTIAN || m.type == MCPT_MAT_ISOTROPIC) && (nl > 0 || pe > 0.0f)) {
        vec3 dl, lc;
        float dist, sapdf;
        if (pe > 0.0f && (nl == 0 || rng.u() < pe)) {
            float dp;
            dl = env_sample(sv.env, rng, dp);
            lc = env_radiance(sv.env, dl, vec3(0.f, 0.f, 0.f));
            dist = 1e30f;
            sapdf = pe * dp;
        } else {
            int li = min((int)(rng.u() * nl), nl - 1);
            const mcpt_light L = sv.lights[li];
            float lsp = (1.0f - pe) / float(nl);
            vec3 pl = sample_triangle(L.tri, rng);
            vec3 tl = pl - so;
            dist = tl.length();
            dl = tl / dist;
            lc = load3(L.emission);
            float cl = fabsf(load3(L.normal).dot(dl));
            sapdf = (cl > 1e-4f) ? lsp * (dist * dist) / (cl * L.area) : 0.0f;
        }
        const bool iso = (m.type == MCPT_MAT_ISOTROPIC);
        float cs = iso ? 1.0f : h.normal.dot(dl);
        if (cs > 0.0f && sapdf > 0.0f) {
            float php = hg_pdf(rd.dot(dl), m.phase_g);
            float spdf = iso ? php : lambert_pdf(cs);
            float sval = iso ? php : cs / MCPT_PI_F;
            float nw = power_heuristic(sapdf, spdf);
            out.shadow = true;
            out.so = so;
            out.sd = dl;
            out.smax = dist * 0.999f;
            out.contrib = ps.thr * lc * (sval / sapdf) * nw;
        }
    }
    if (depth > 3) {  // Russian roulette (:3011-3017)
        float mc = fmaxf(ps.thr.x, fmaxf(ps.thr.y, ps.thr.z));
        float cp = fminf(mc, rr_cap);
        if (rng.u() > cp) return;
        ps.thr = ps.thr / cp;
    }
    vec3 nd;
    switch (m.type) {
        case MCPT_MAT_CONDUCTOR: {
            vec3 refl = rd.normalize().reflect(h.normal);
            nd = (refl + random_unit_vector(rng) * m.fuzz).normalize();
            break;
        }
        case MCPT_MAT_DIELECTRIC: {
            float eta = h.front ? 1.0f / m.ior : m.ior;
            nd = dielectric_dir(rd, h.normal, eta, rng);
            break;
        }
        case MCPT_MAT_ISOTROPIC: nd = hg_dir(rd, m.phase_g, rng); break;
        default: nd = lambertian_dir(h.normal, rng); break;
    }
    ps.prev_spec = (m.type != MCPT_MAT_LAMBERTIAN && m.type != MCPT_MAT_ISOTROPIC);
    ps.prev_pdf = (m.type == MCPT_MAT_ISOTROPIC) ? hg_pdf(rd.dot(nd), m.phase_g)
                                                 : lambert_pdf(fmaxf(nd.dot(h.normal), 0.0f));
    out.next = true;
    out.no = so;
    out.nd = nd;
}

// Camera::getRay (:1648-1658): lens disk by rejection, then the time draw;
// no subpixel jitter.
MCPT_DEV void camera_ray(const mcpt_camera& cam, int px, int py, Rng& rng, vec3& o, vec3& d, float& time) {
    vec3 pc = load3(cam.pixel00) + (load3(cam.delta_u) * float(px)) + (load3(cam.delta_v) * float(py));
    vec3 lens;
    while (true) {
        float rx = rng.u() * 2.f - 1.f;
        float ry = rng.u() * 2.f - 1.f;
        vec3 cand(rx, ry, 0.0f);
        if (cand.length_squared() < 1.0f) {
            lens = cand;
            break;
        }
    }
    lens = lens * cam.lens_radius;
    vec3 off = load3(cam.u) * lens.x + load3(cam.v) * lens.y;
    o = load3(cam.center) + off;
    time = rng.u();
    d = (pc - o).normalize();
}

// firefly clamp (:3050-3055)
MCPT_DEV vec3 clamp_sample(vec3 c) {
    float b = fmaxf(c.x, fmaxf(c.y, c.z));
    if (b > 10.0f) c = c * (10.0f / b);
    return c;
}

}  // namespace mcpt
