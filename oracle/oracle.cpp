// oracle.cpp — CPU oracle for the path-tracing core. TEST INFRASTRUCTURE:
// only the tests, the smoke check and the benchmark's CPU-baseline leg use
// it. It restates MonoCUDA-PT's device functions (src/path_tracer.cu, cited
// per function) with the same operation order, in host C++17 built with
// -ffp-contract=off. cuRAND is replaced by PCG32 keyed by (pixel, sample)
// and CUDA's hardware bilinear texture fetch by a software bilinear filter.
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "mcpt_fmath.h"
#include "mcpt_pcg32.h"

namespace {

// ---------------------------------------------------------------- vec3
// src/path_tracer.cu:105-163
struct V3 {
    float x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(float a, float b, float c) : x(a), y(b), z(c) {}
    V3 operator+(const V3& o) const { return V3(x + o.x, y + o.y, z + o.z); }
    V3 operator-(const V3& o) const { return V3(x - o.x, y - o.y, z - o.z); }
    V3 operator*(const V3& o) const { return V3(x * o.x, y * o.y, z * o.z); }
    V3 operator-() const { return V3(-x, -y, -z); }
    V3 operator*(float n) const { return V3(x * n, y * n, z * n); }
    V3 operator/(float n) const { return V3(x / n, y / n, z / n); }
    float length_squared() const { return x * x + y * y + z * z; }
    float length() const { return sqrtf(length_squared()); }
    V3 cross(const V3& o) const { return V3(y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x); }
    float dot(const V3& o) const { return x * o.x + y * o.y + z * o.z; }
    V3 normalize() const {
        float l = sqrtf(length_squared());
        return V3(x / l, y / l, z / l);
    }
    V3 reflect(const V3& n) const { return *this - n * (2.0f * dot(n)); }
    // src/path_tracer.cu:146-163
    V3 refract(const V3& n, float eta) const {
        V3 d = normalize();
        float cos1 = fminf((-d).dot(n), 1.0f);
        float theta1 = mf_acos(cos1);
        float sin2 = mf_sin(theta1) * eta;
        if (sin2 > 1.f) return d.reflect(n);
        float theta2 = mf_asin(sin2);
        V3 tangent = d + n * cos1;
        float tl = tangent.length();
        if (tl > 1e-8f) tangent = tangent / tl;
        return tangent * mf_sin(theta2) - n * mf_cos(theta2);
    }
};
V3 v3(const float* p) { return V3(p[0], p[1], p[2]); }

struct AABB {
    V3 mn, mx;
};

// merge (src/path_tracer.cu:1456-1465)
AABB merge(const AABB& a, const AABB& b) {
    AABB m;
    m.mx.x = fmaxf(a.mx.x, b.mx.x);
    m.mx.y = fmaxf(a.mx.y, b.mx.y);
    m.mx.z = fmaxf(a.mx.z, b.mx.z);
    m.mn.x = fminf(a.mn.x, b.mn.x);
    m.mn.y = fminf(a.mn.y, b.mn.y);
    m.mn.z = fminf(a.mn.z, b.mn.z);
    return m;
}

// triangle::calculate_bounding_box (src/path_tracer.cu:651-655)
AABB tri_box(const mcpt_tri_hot& t) {
    const float eps = 1e-3f;
    AABB b;
    b.mn = V3(fminf(t.v0[0], fminf(t.v1[0], t.v2[0])) - eps, fminf(t.v0[1], fminf(t.v1[1], t.v2[1])) - eps,
              fminf(t.v0[2], fminf(t.v1[2], t.v2[2])) - eps);
    b.mx = V3(fmaxf(t.v0[0], fmaxf(t.v1[0], t.v2[0])) + eps, fmaxf(t.v0[1], fmaxf(t.v1[1], t.v2[1])) + eps,
              fmaxf(t.v0[2], fmaxf(t.v1[2], t.v2[2])) + eps);
    return b;
}

// hit_aabb (src/path_tracer.cu:1467-1488)
bool hit_aabb(const AABB& box, const V3& o, const V3& d, float& tnear, float& tfar) {
    const float os[3] = {o.x, o.y, o.z}, ds[3] = {d.x, d.y, d.z};
    const float mn[3] = {box.mn.x, box.mn.y, box.mn.z}, mx[3] = {box.mx.x, box.mx.y, box.mx.z};
    for (int a = 0; a < 3; a++) {
        float inv = 1.0f / ds[a];
        float t0 = (mn[a] - os[a]) * inv;
        float t1 = (mx[a] - os[a]) * inv;
        if (inv < 0.0f) std::swap(t0, t1);
        tnear = fmaxf(tnear, t0);
        tfar = fminf(tfar, t1);
        if (tfar <= tnear) return false;
    }
    return true;
}

// hit_triangle (src/path_tracer.cu:715-741)
bool hit_triangle(const mcpt_tri_hot& tri, const V3& o, const V3& d, float& tout) {
    V3 v0 = v3(tri.v0), v1 = v3(tri.v1), v2 = v3(tri.v2);
    V3 e1 = v1 - v0;
    V3 e2 = v2 - v0;
    V3 n = e1.cross(e2).normalize();
    float den = d.dot(n);
    if (fabsf(den) < 1e-6f) return false;
    float t = (v0 - o).dot(n) / den;
    if (t < 0.0001f) return false;
    V3 p = o + d * t;
    if (n.dot((v1 - v0).cross(p - v0)) < 0) return false;
    if (n.dot((v2 - v1).cross(p - v1)) < 0) return false;
    if (n.dot((v0 - v2).cross(p - v2)) < 0) return false;
    tout = t;
    return true;
}

// sphere::calculate_bounding_box (src/path_tracer.cu:862-869): centre +-
// radius, merged over both ends of the motion for a moving sphere
AABB sphere_box(const mcpt_sphere& sp) {
    const float r = sp.radius;
    AABB b;
    b.mn = V3(sp.center[0] - r, sp.center[1] - r, sp.center[2] - r);
    b.mx = V3(sp.center[0] + r, sp.center[1] + r, sp.center[2] + r);
    if (sp.moving) {
        AABB b1;
        b1.mn = V3(sp.center1[0] - r, sp.center1[1] - r, sp.center1[2] - r);
        b1.mx = V3(sp.center1[0] + r, sp.center1[1] + r, sp.center1[2] + r);
        b = merge(b, b1);
    }
    return b;
}

// sphere::center_at (src/path_tracer.cu:870-872)
V3 sphere_center_at(const mcpt_sphere& sp, float time) {
    return sp.moving ? v3(sp.center) + (v3(sp.center1) - v3(sp.center)) * time : v3(sp.center);
}

// hit_sphere (src/path_tracer.cu:886-911)
bool hit_sphere(const mcpt_sphere& sp, const V3& o, const V3& d, float time, float& tout) {
    V3 c = sphere_center_at(sp, time);
    V3 to_center = c - o;
    float along = to_center.dot(d) / d.dot(d);
    V3 closest = o + d * along;
    float dist2 = (c - closest).length_squared();
    float disc = sp.radius * sp.radius - dist2;
    if (disc < 0.0f) return false;
    float half = sqrtf(disc);
    float t = along - half;
    if (t < 0.0001f) t = along + half;
    if (t < 0.0001f) return false;
    tout = t;
    return true;
}

// computeSphereUvFromOutwardNormal (src/path_tracer.cu:877-884)
void sphere_uv(const V3& n, float& u, float& v) {
    float ny = fminf(1.0f, fmaxf(-1.0f, -n.y));
    float theta = mf_acos(ny);
    float phi = mf_atan2(-n.z, n.x) + MCPT_PI_F;
    u = phi / (2.0f * MCPT_PI_F);
    v = theta / MCPT_PI_F;
}

// Morton (src/path_tracer.cu:913-935)
uint32_t expand_bits(uint32_t p) {
    uint32_t e = 0;
    for (int i = 0; i < 10; i++) e |= ((p >> i) & 1u) << (3 * i);
    return e;
}
uint32_t quantize(float v, float lo, float hi) {
    float nrm = (v - lo) / (hi - lo);
    return (uint32_t)(nrm * 1023.0f);
}

// delta (src/path_tracer.cu:1337-1347)
int delta(const uint32_t* k, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    if (k[i] == k[j]) return 32 + __builtin_clz((uint32_t)(i ^ j));
    return __builtin_clz(k[i] ^ k[j]);
}

struct Node {  // internal node of the Karras tree (BVH_Node, :1006-1015)
    int left, right;
    bool left_leaf, right_leaf;
    int parent;
    AABB box;
};

struct Env {
    bool enabled = false;
    int w = 0, h = 0;
    float intensity = 1.f, rot = 0.f, total = 0.f;
    std::vector<float> rgba, row_cdf, col_cdf;
};

}  // namespace

namespace {
struct Tex {  // a checker (linear colours) or an RGBA8 image
    int type = 0, w = 0, h = 0;
    float scale = 1.f;
    float ca[3] = {0, 0, 0}, cb[3] = {0, 0, 0};
    std::vector<uint32_t> px;
};
}  // namespace

struct oracle_scene_s {
    std::vector<mcpt_sphere> spheres;
    std::vector<mcpt_tri_hot> hot;
    std::vector<mcpt_tri_cold> cold;
    std::vector<mcpt_material> mats;
    std::vector<mcpt_light> lights;
    Env env;
    std::vector<Tex> tex;
    float srgb8[256];  // sRGB byte -> linear
    // BVH over the primitive list of finalizeScene (:1955-1972): spheres
    // 0..S-1, then triangles; leaf_prim holds a triangle index, or ~ a
    // sphere index
    int n = 0;
    std::vector<uint32_t> morton, keys;
    std::vector<int32_t> ids;
    std::vector<AABB> tri_boxes;  // original order
    AABB bounds;
    std::vector<Node> nodes;       // n-1 (or 1 for n == 1)
    std::vector<int> leaf_parent;  // n
    std::vector<int> leaf_prim;    // n
    std::vector<AABB> leaf_box;    // n
    int max_depth = 0;
    // density grid of every isotropic material (empty: homogeneous media)
    std::vector<float> grid;
    int gdim[3] = {0, 0, 0};
    float gorg[3] = {0, 0, 0}, ginv[3] = {0, 0, 0};
    float gmax = 0.0f;
};

namespace {

// ---------------------------------------------------------------- traversal
struct Closest {
    float t;
    int prim;  // triangle index, or ~ sphere index
    bool hit;
};

// the primitive-type switch of testPrimitiveClosest (src/path_tracer.cu:1537-1551)
bool prim_hit(const oracle_scene_s& s, int p, const V3& o, const V3& d, float time, float& t) {
    if (p < 0) return hit_sphere(s.spheres[~p], o, d, time, t);
    return hit_triangle(s.hot[p], o, d, t);
}

// BVH::traverseRay / stepTraversal / testPrimitiveClosest
// (src/path_tracer.cu:1092-1150, :1262-1321, :1537-1551). Returns the closest
// candidate; counts internal-node iterations and leaf tests.
Closest traverse_closest(const oracle_scene_s& s, const V3& o, const V3& d, float tmax, float time,
                         uint32_t* steps, uint32_t* leaves) {
    Closest c{tmax, 0, false};
    int stack[64];
    int sp = 0;
    int node = 0;
    uint32_t ns = 0, nl = 0;
    while (true) {
        const Node& nd = s.nodes[node];
        ns++;
        const AABB& bl = nd.left_leaf ? s.leaf_box[nd.left] : s.nodes[nd.left].box;
        const AABB& br = nd.right_leaf ? s.leaf_box[nd.right] : s.nodes[nd.right].box;
        float nL = 0.0001f, fL = c.t;
        bool oL = hit_aabb(bl, o, d, nL, fL);
        float nR = 0.0001f, fR = c.t;
        bool oR = hit_aabb(br, o, d, nR, fR);
        if (oL && nd.left_leaf) {
            nl++;
            int p = s.leaf_prim[nd.left];
            float t;
            if (prim_hit(s, p, o, d, time, t) && t < c.t && t > 0.0001f) c = Closest{t, p, true};
        }
        if (oR && nd.right_leaf) {
            nl++;
            int p = s.leaf_prim[nd.right];
            float t;
            if (prim_hit(s, p, o, d, time, t) && t < c.t && t > 0.0001f) c = Closest{t, p, true};
        }
        bool tL = oL && !nd.left_leaf;
        bool tR = oR && !nd.right_leaf;
        if (!tL && !tR) {
            if (sp == 0) break;
            node = stack[--sp];
        } else {
            if (tL && tR) stack[sp++] = nd.right;
            node = tL ? nd.left : nd.right;
        }
    }
    if (steps) *steps = ns;
    if (leaves) *leaves = nl;
    return c;
}

// BVH::traverseAnyHit + doesPrimitiveBlockRay (src/path_tracer.cu:1152-1207, :1553-1563)
bool traverse_any(const oracle_scene_s& s, const V3& o, const V3& d, float tmax, float time) {
    int stack[64];
    int sp = 0;
    int node = 0;
    while (true) {
        const Node& nd = s.nodes[node];
        const AABB& bl = nd.left_leaf ? s.leaf_box[nd.left] : s.nodes[nd.left].box;
        const AABB& br = nd.right_leaf ? s.leaf_box[nd.right] : s.nodes[nd.right].box;
        float nL = 0.0001f, fL = tmax;
        bool oL = hit_aabb(bl, o, d, nL, fL);
        float nR = 0.0001f, fR = tmax;
        bool oR = hit_aabb(br, o, d, nR, fR);
        float t;
        if (oL && nd.left_leaf && prim_hit(s, s.leaf_prim[nd.left], o, d, time, t) && t > 0.0001f && t < tmax)
            return true;
        if (oR && nd.right_leaf && prim_hit(s, s.leaf_prim[nd.right], o, d, time, t) && t > 0.0001f && t < tmax)
            return true;
        bool tL = oL && !nd.left_leaf;
        bool tR = oR && !nd.right_leaf;
        if (!tL && !tR) {
            if (sp == 0) break;
            node = stack[--sp];
        } else {
            if (tL && tR) stack[sp++] = nd.right;
            node = tL ? nd.left : nd.right;
        }
    }
    return false;
}

// --------------------------------------------------------- hit reconstruction
struct Hit {
    bool did = false;
    float t = 1e30f;
    V3 p, normal;
    bool front = true;
    int prim = 0;  // triangle index, or ~ sphere index
    float u = 0, v = 0;
    const mcpt_material* mat = nullptr;
};

// computeTriangleUvFromHitPoint (src/path_tracer.cu:687-696)
void tri_uv(const mcpt_tri_hot& h, const mcpt_tri_cold& c, const V3& p, float& u, float& v) {
    V3 v0 = v3(h.v0), v1 = v3(h.v1), v2 = v3(h.v2);
    float a2 = (v1 - v0).cross(v2 - v0).length();
    float w0 = (v1 - p).cross(v2 - p).length() / a2;
    float w1 = (v0 - p).cross(v2 - p).length() / a2;
    float w2 = (v1 - p).cross(v0 - p).length() / a2;
    u = ((c.uv0[0] * w0) + (c.uv1[0] * w1)) + (c.uv2[0] * w2);
    v = ((c.uv0[1] * w0) + (c.uv1[1] * w1)) + (c.uv2[1] * w2);
}

// computeTriangleShadingNormal (src/path_tracer.cu:698-713)
V3 tri_shading_normal(const mcpt_tri_hot& h, const mcpt_tri_cold& c, const V3& p, const V3& face) {
    V3 n0 = v3(c.n0), n1 = v3(c.n1), n2 = v3(c.n2);
    if (n0.length_squared() == 0.0f) return face;
    V3 v0 = v3(h.v0), v1 = v3(h.v1), v2 = v3(h.v2);
    float a2 = (v1 - v0).cross(v2 - v0).length();
    float w0 = (v1 - p).cross(v2 - p).length() / a2;
    float w1 = (v0 - p).cross(v2 - p).length() / a2;
    float w2 = (v1 - p).cross(v0 - p).length() / a2;
    float ws = w0 + w1 + w2;
    if (!(ws > 0.0f) || ws > 1.01f) return face;
    V3 sn = n0 * w0 + n1 * w1 + n2 * w2;
    if (sn.length_squared() < 1e-12f) return face;
    sn = sn.normalize();
    if (sn.dot(face) < 0.0f) sn = -sn;
    return sn;
}

// the hitRecord fields testPrimitiveHit fills (src/path_tracer.cu:1490-1533)
Hit make_hit(const oracle_scene_s& s, const V3& o, const V3& d, float time, const Closest& c) {
    Hit h;
    if (!c.hit) {
        h.t = c.t;
        return h;
    }
    if (c.prim < 0) {
        const mcpt_sphere& sp = s.spheres[~c.prim];
        h.did = true;
        h.t = c.t;
        h.prim = c.prim;
        h.p = o + d * c.t;
        h.mat = &s.mats[sp.material];
        V3 outward = (h.p - sphere_center_at(sp, time)).normalize();
        h.front = d.dot(outward) < 0.0f;
        h.normal = h.front ? outward : -outward;
        sphere_uv(outward, h.u, h.v);
        return h;
    }
    const mcpt_tri_hot& th = s.hot[c.prim];
    const mcpt_tri_cold& tc = s.cold[c.prim];
    h.did = true;
    h.t = c.t;
    h.prim = c.prim;
    h.p = o + d * c.t;
    h.mat = &s.mats[tc.material];
    V3 n = (v3(th.v1) - v3(th.v0)).cross(v3(th.v2) - v3(th.v0)).normalize();
    h.front = d.dot(n) < 0.0f;
    V3 sn = tri_shading_normal(th, tc, h.p, n);
    h.normal = h.front ? sn : -sn;
    tri_uv(th, tc, h.p, h.u, h.v);
    return h;
}

Hit trace_ray(const oracle_scene_s& s, const V3& o, const V3& d, float tmax, float time, uint32_t* steps,
              uint32_t* leaves) {
    Closest c = traverse_closest(s, o, d, tmax, time, steps, leaves);
    return make_hit(s, o, d, time, c);
}

// ---------------------------------------------------------------- sampling
struct Rng {
    mcpt_pcg32 g;
    float u() { return mcpt_pcg32_uniform(&g); }
};

// random_unit_vector (src/path_tracer.cu:498-509)
template <class R>
V3 random_unit_vector(R& r) {
    while (true) {
        float x = r.u() * 2.f - 1.f;
        float y = r.u() * 2.f - 1.f;
        float z = r.u() * 2.f - 1.f;
        V3 v(x, y, z);
        if (v.length_squared() < 1.0f) return v.normalize();
    }
}
// lambertian_scatter_direction (:511-518)
template <class R>
V3 lambertian_dir(const V3& n, R& r) {
    V3 d = n + random_unit_vector(r);
    if (d.length_squared() < 1e-8f) return n;
    return d.normalize();
}
// the Conductor case of the scatter switch (:3657-3661): the mirror of
// specular_scatter_direction (:520-522) plus fuzz times a unit vector
template <class R>
V3 conductor_dir(const V3& in, const V3& n, float fuzz, R& r) {
    V3 refl = in.normalize().reflect(n);
    return (refl + random_unit_vector(r) * fuzz).normalize();
}
// schlick_approx (:476-480)
float schlick(float c, float eta) {
    float r0 = (1.f - eta) / (1.f + eta);
    r0 = r0 * r0;
    return r0 + (1.f - r0) * mf_pow((1 - c), 5.f);
}
// dielectric_scatter_direction (:524-540); R is any source of uniforms
// (the PCG32 stream, or fixed values in the closed-form tests)
template <class R>
V3 dielectric_dir(const V3& in, const V3& n, float eta, R& r) {
    V3 ud = in.normalize();
    float c = fminf((-ud).dot(n), 1.0f);
    float sn = sqrtf(1.0f - c * c);
    bool cannot = (eta * sn) > 1.0f;
    float pr = cannot ? 1.0f : schlick(c, eta);
    if (r.u() < pr) return ud.reflect(n);
    return ud.refract(n, eta);
}
// isotropic_scatter_direction (:542-551)
template <class R>
V3 isotropic_dir(R& r) {
    float u1 = r.u();
    float u2 = r.u();
    float z = 1.0f - 2.0f * u1;
    float rr = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    float phi = 2.0f * MCPT_PI_F * u2;
    return V3(rr * mf_cos(phi), rr * mf_sin(phi), z);
}
// henyeyGreenstein_scatter_direction (:556-578)
template <class R>
V3 hg_dir(const V3& in, float g, R& r) {
    if (fabsf(g) < 1e-3f) return isotropic_dir(r);
    float u1 = r.u();
    float u2 = r.u();
    float sq = (1.0f - g * g) / (1.0f + g - 2.0f * g * u1);
    float ct = (1.0f + g * g - sq * sq) / (2.0f * g);
    ct = fminf(fmaxf(ct, -1.0f), 1.0f);
    float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
    float phi = 2.0f * MCPT_PI_F * u2;
    V3 fw = in.normalize();
    V3 helper = (fabsf(fw.x) > 0.9f) ? V3(0.f, 1.f, 0.f) : V3(1.f, 0.f, 0.f);
    V3 tg = helper.cross(fw).normalize();
    V3 bt = fw.cross(tg);
    return tg * (st * mf_cos(phi)) + bt * (st * mf_sin(phi)) + fw * ct;
}
// sample_triangle (:744-755)
V3 sample_triangle(const mcpt_tri_hot& t, Rng& r) {
    float r1 = r.u();
    float r2 = r.u();
    float s1 = sqrtf(r1);
    float w0 = 1.0f - s1;
    float w1 = r2 * s1;
    float w2 = 1.0f - w0 - w1;
    return v3(t.v0) * w0 + v3(t.v1) * w1 + v3(t.v2) * w2;
}
float power_heuristic(float a, float b) {  // :757-761
    float a2 = a * a, b2 = b * b;
    return a2 / (a2 + b2);
}
float lambert_pdf(float c) { return c / MCPT_PI_F; }  // :763-765
float hg_pdf(float ct, float g) {                     // :775-780
    float g2 = g * g;
    float den = fmaxf(1.0f + g2 - 2.0f * g * ct, 1e-6f);
    return (1.0f - g2) / (4.0f * MCPT_PI_F * den * sqrtf(den));
}
float tri_area(const mcpt_tri_hot& t) {  // :837-839
    return 0.5f * (v3(t.v1) - v3(t.v0)).cross(v3(t.v2) - v3(t.v0)).length();
}

// ------------------------------------------------------------ environment
// environmentUv / environmentDirection (src/path_tracer.cu:408-420)
void env_uv(const Env& e, const V3& ud, float& u, float& v) {
    float uu = mf_atan2(ud.z, ud.x) / (2.0f * MCPT_PI_F) + 0.5f + e.rot / (2.0f * MCPT_PI_F);
    v = mf_acos(fminf(fmaxf(ud.y, -1.0f), 1.0f)) / MCPT_PI_F;
    u = uu - floorf(uu);
}
V3 env_direction(const Env& e, float u, float v) {
    float phi = (u - 0.5f) * 2.0f * MCPT_PI_F - e.rot;
    float theta = v * MCPT_PI_F;
    float st = mf_sin(theta);
    return V3(mf_cos(phi) * st, mf_cos(theta), mf_sin(phi) * st);
}
// Software replacement of tex2D with linear filtering, wrap in u, clamp in
// v (the texture set up at src/path_tracer.cu:387-406): full FP32 weights.
V3 env_fetch(const Env& e, float u, float v) {
    float x = u * (float)e.w - 0.5f;
    float y = v * (float)e.h - 0.5f;
    float x0 = floorf(x), y0 = floorf(y);
    float fx = x - x0, fy = y - y0;
    int ix0 = (int)x0, iy0 = (int)y0;
    int ix1 = ix0 + 1, iy1 = iy0 + 1;
    ix0 = ((ix0 % e.w) + e.w) % e.w;
    ix1 = ((ix1 % e.w) + e.w) % e.w;
    iy0 = std::min(std::max(iy0, 0), e.h - 1);
    iy1 = std::min(std::max(iy1, 0), e.h - 1);
    const float* t00 = &e.rgba[((size_t)iy0 * e.w + ix0) * 4];
    const float* t10 = &e.rgba[((size_t)iy0 * e.w + ix1) * 4];
    const float* t01 = &e.rgba[((size_t)iy1 * e.w + ix0) * 4];
    const float* t11 = &e.rgba[((size_t)iy1 * e.w + ix1) * 4];
    float gx = 1.0f - fx, gy = 1.0f - fy;
    float c[3];
    for (int k = 0; k < 3; k++) c[k] = (t00[k] * gx + t10[k] * fx) * gy + (t01[k] * gx + t11[k] * fx) * fy;
    return V3(c[0], c[1], c[2]);
}
// environmentRadiance (:422-427)
V3 env_radiance(const Env& e, const V3& d, const V3& fallback) {
    if (!e.enabled) return fallback;
    float u, v;
    env_uv(e, d.normalize(), u, v);
    return env_fetch(e, u, v) * e.intensity;
}
float env_select_prob(const Env& e, int nl) {  // :429-432
    if (!e.enabled || e.total <= 0.0f) return 0.0f;
    return (nl > 0) ? 0.5f : 1.0f;
}
int find_cdf_interval(const float* cdf, int count, float value) {  // :434-443
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (cdf[mid + 1] <= value) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}
float env_texel_pdf(const Env& e, int row, int col, float st) {  // :445-450
    if (st < 1e-6f) return 0.0f;
    const float* cdf = e.col_cdf.data() + (size_t)row * (e.w + 1);
    float tw = cdf[col + 1] - cdf[col];
    return tw * e.w * e.h / (e.total * 2.0f * MCPT_PI_F * MCPT_PI_F * st);
}
float env_dir_pdf(const Env& e, const V3& d) {  // :452-458
    V3 ud = d.normalize();
    float u, v;
    env_uv(e, ud, u, v);
    int col = std::min((int)(u * e.w), e.w - 1);
    int row = std::min((int)(v * e.h), e.h - 1);
    return env_texel_pdf(e, row, col, sqrtf(fmaxf(1.0f - ud.y * ud.y, 0.0f)));
}
template <class R>
V3 env_sample(const Env& e, R& r, float& pdf) {  // :460-467
    int row = find_cdf_interval(e.row_cdf.data(), e.h, r.u() * e.total);
    const float* cdf = e.col_cdf.data() + (size_t)row * (e.w + 1);
    int col = find_cdf_interval(cdf, e.w, r.u() * cdf[e.w]);
    float u = (col + r.u()) / e.w;
    float v = (row + r.u()) / e.h;
    pdf = env_texel_pdf(e, row, col, mf_sin(v * MCPT_PI_F));
    return env_direction(e, u, v);
}

// --------------------------------------------------------------- integrator
struct PathCounts {
    uint64_t ext = 0, shd = 0, steps = 0, leaves = 0;
};

// srgbToLinear (src/path_tracer.cu:200-202)
float srgb_to_linear(float e) { return e <= 0.04045f ? e / 12.92f : mf_pow((e + 0.055f) / 1.055f, 2.4f); }

// sampleTexture (src/path_tracer.cu:216-234); images are read with the
// texture unit's normalised-coordinate bilinear filter and wrap addressing,
// restated in FP32: x = u w - 1/2, y = (1 - v) h - 1/2, the four texels
// decoded from sRGB, then blended by the fractional parts.
V3 albedo_at(const oracle_scene_s& s, const mcpt_material& m, float u, float v) {
    if (m.texture <= 0) return v3(m.albedo);
    const Tex& t = s.tex[m.texture - 1];
    if (t.type == MCPT_TEX_CHECKER) {
        const int cu = (int)floorf(u * t.scale), cv = (int)floorf(v * t.scale);
        return ((cu + cv) & 1) == 0 ? V3(t.ca[0], t.ca[1], t.ca[2]) : V3(t.cb[0], t.cb[1], t.cb[2]);
    }
    const float x = fminf(fmaxf(u * (float)t.w - 0.5f, -1e9f), 1e9f);
    const float y = fminf(fmaxf((1.0f - v) * (float)t.h - 0.5f, -1e9f), 1e9f);
    const float x0 = floorf(x), y0 = floorf(y);
    const float fx = x - x0, fy = y - y0;
    const int ix0 = (((int)((long long)x0 % t.w)) + t.w) % t.w, iy0 = (((int)((long long)y0 % t.h)) + t.h) % t.h;
    const int ix1 = (ix0 + 1) % t.w, iy1 = (iy0 + 1) % t.h;
    const uint32_t p00 = t.px[(size_t)iy0 * t.w + ix0], p10 = t.px[(size_t)iy0 * t.w + ix1];
    const uint32_t p01 = t.px[(size_t)iy1 * t.w + ix0], p11 = t.px[(size_t)iy1 * t.w + ix1];
    const float gx = 1.0f - fx, gy = 1.0f - fy;
    float c[3];
    for (int k = 0; k < 3; k++) {
        const int sh = 8 * k;
        const float a00 = s.srgb8[(p00 >> sh) & 255u], a10 = s.srgb8[(p10 >> sh) & 255u];
        const float a01 = s.srgb8[(p01 >> sh) & 255u], a11 = s.srgb8[(p11 >> sh) & 255u];
        c[k] = (a00 * gx + a10 * fx) * gy + (a01 * gx + a11 * fx) * fy;
    }
    return V3(c[0], c[1], c[2]);
}

// sampleVolumeDensity (src/path_tracer.cu:784-789) over a dense grid:
// index point q = (p - origin) * (1 / voxel), trilinear between the eight
// voxels around floor(q) (NanoVDB's TrilinearSampler: lerp along z, then
// y, then x), voxels outside the grid read 0.
float grid_at(const oracle_scene_s& s, long long i, long long j, long long k) {
    if (i < 0 || j < 0 || k < 0 || i >= s.gdim[0] || j >= s.gdim[1] || k >= s.gdim[2]) return 0.0f;
    return s.grid[((size_t)k * s.gdim[1] + (size_t)j) * s.gdim[0] + (size_t)i];
}
float grid_density(const oracle_scene_s& s, const V3& p) {
    const float q[3] = {(p.x - s.gorg[0]) * s.ginv[0], (p.y - s.gorg[1]) * s.ginv[1], (p.z - s.gorg[2]) * s.ginv[2]};
    long long c[3];
    float f[3];
    for (int a = 0; a < 3; a++) {
        if (!(q[a] > -1.0f && q[a] < (float)s.gdim[a])) return 0.0f;  // every neighbour outside, or NaN
        const float fl = floorf(q[a]);
        c[a] = (long long)fl;
        f[a] = q[a] - fl;
    }
    auto lerp = [](float a, float b, float w) { return a + w * (b - a); };
    float xs[2];
    for (int dx = 0; dx < 2; dx++) {
        float ys[2];
        for (int dy = 0; dy < 2; dy++)
            ys[dy] = lerp(grid_at(s, c[0] + dx, c[1] + dy, c[2]), grid_at(s, c[0] + dx, c[1] + dy, c[2] + 1), f[2]);
        xs[dx] = lerp(ys[0], ys[1], f[1]);
    }
    return lerp(xs[0], xs[1], f[0]);
}

// heterogeneousFreeFlightDistance (src/path_tracer.cu:796-812)
float grid_free_flight(const oracle_scene_s& s, float sigma, const V3& o, const V3& d, float maxd, Rng& rng) {
    const float majorant = sigma * s.gmax;
    if (!(majorant > 0.0f)) return maxd;
    float t = 0.0f;
    for (int step = 0; step < 8192; step++) {
        t = t - mf_log(1.0f - rng.u()) / majorant;
        if (t >= maxd) return maxd;
        if (rng.u() < grid_density(s, o + d * t) / s.gmax) return t;
    }
    return maxd;
}

// heterogeneousTransmittance (src/path_tracer.cu:817-835)
float grid_transmittance(const oracle_scene_s& s, float sigma, const V3& o, const V3& d, float len, Rng& rng) {
    const float majorant = sigma * s.gmax;
    if (!(majorant > 0.0f)) return 1.0f;
    float tr = 1.0f, t = 0.0f;
    for (int step = 0; step < 8192; step++) {
        t = t - mf_log(1.0f - rng.u()) / majorant;
        if (t >= len) break;
        tr = tr * (1.0f - grid_density(s, o + d * t) / s.gmax);
        if (tr < 1e-4f) return 0.0f;
    }
    return tr;
}

uint32_t bits_of(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

// mediumTransmittance (src/path_tracer.cu:2874-2898). With a density grid,
// ratio tracking draws from the shadow ray's own stream, keyed by the ray's
// bits (mcpt_pcg32_for_shadow) rather than by the path's state.
float medium_transmittance(const oracle_scene_s& s, V3 o, const V3& d, float maxd, float time, PathCounts& pc) {
    float tr = 1.0f;
    float travelled = 0.0f;
    Rng srng;
    if (!s.grid.empty()) {
        const uint32_t key[7] = {bits_of(o.x), bits_of(o.y), bits_of(o.z), bits_of(d.x),
                                 bits_of(d.y), bits_of(d.z), bits_of(maxd)};
        mcpt_pcg32_for_shadow(&srng.g, key);
    }
    for (int k = 0; k < 4; k++) {
        float remaining = maxd - travelled;
        if (remaining <= 0.0f) break;
        pc.shd++;
        Hit b = trace_ray(s, o, d, remaining, time, nullptr, nullptr);
        if (!b.did) break;
        if (b.mat->type != MCPT_MAT_ISOTROPIC) return 0.0f;
        if (!b.front)
            tr = tr * (s.grid.empty() ? mf_exp(-b.mat->density * b.t)
                                      : grid_transmittance(s, b.mat->density, o, d, b.t, srng));
        travelled = travelled + b.t;
        o = b.p;
    }
    return tr;
}

// Camera::getRay (src/path_tracer.cu:1633-1658): the pixel centre (no
// jitter), a lens point drawn by rejection from the unit disk, then the
// path's time
template <class R>
void camera_ray(const mcpt_camera& cam, int px, int py, R& rng, V3& ro, V3& rd, float& time) {
    V3 pc_ = v3(cam.pixel00) + (v3(cam.delta_u) * float(px)) + (v3(cam.delta_v) * float(py));
    V3 lens;
    while (true) {
        float rx = rng.u() * 2.f - 1.f;
        float ry = rng.u() * 2.f - 1.f;
        V3 cand(rx, ry, 0.0f);
        if (cand.length_squared() < 1.0f) {
            lens = cand;
            break;
        }
    }
    lens = lens * cam.lens_radius;
    V3 off = v3(cam.u) * lens.x + v3(cam.v) * lens.y;
    ro = v3(cam.center) + off;
    time = rng.u();  // the path's time (moving spheres)
    rd = (pc_ - ro).normalize();
}

// trace_sample_kernel body (src/path_tracer.cu:2900-3059) for one path.
V3 trace_path(const oracle_scene_s& s, const mcpt_camera& cam, int px, int py, Rng& rng, int max_depth,
              PathCounts& pc) {
    V3 ro, rd;
    float time;
    camera_ray(cam, px, py, rng, ro, rd, time);

    const int nl = (int)s.lights.size();
    const V3 bg = v3(cam.background);
    V3 att(1.0f, 1.0f, 1.0f);
    V3 col(0.0f, 0.0f, 0.0f);
    bool prev_spec = true;
    float prev_pdf = 1.0f;
    for (int depth = 0; depth < max_depth; depth++) {
        pc.ext++;
        uint32_t st = 0, lf = 0;
        Hit hit = trace_ray(s, ro, rd, 1e30f, time, &st, &lf);
        pc.steps += st;
        pc.leaves += lf;
        if (!hit.did) {
            float w = 1.0f;
            float pe = env_select_prob(s.env, nl);
            if (!(depth == 0 || prev_spec) && pe > 0.0f) w = power_heuristic(prev_pdf, pe * env_dir_pdf(s.env, rd));
            col = col + att * env_radiance(s.env, rd, bg) * w;
            break;
        }
        const mcpt_material& m = *hit.mat;
        float ew = 1.0f;
        bool from_bsdf = !(depth == 0 || prev_spec);
        if (from_bsdf && m.type == MCPT_MAT_EMISSIVE && hit.prim >= 0) {  // triangles only (:2930)
            float cl = fmaxf(fabsf(rd.dot(hit.normal)), 1e-4f);
            float area = tri_area(s.hot[hit.prim]);
            float lp = ((1.0f - env_select_prob(s.env, nl)) / float(nl)) * (hit.t * hit.t) / (cl * area);
            ew = power_heuristic(prev_pdf, lp);
        }
        col = col + att * v3(m.emission) * ew;
        if (m.type == MCPT_MAT_EMISSIVE) break;

        V3 so = hit.p;
        if (m.type == MCPT_MAT_ISOTROPIC) {
            float sd;
            if (hit.front) sd = hit.t;
            else if (!s.grid.empty()) sd = grid_free_flight(s, m.density, ro, rd, hit.t, rng);
            else sd = -mf_log(1.0f - rng.u()) / m.density;
            if (sd >= hit.t) {
                ro = hit.p;
                continue;
            }
            so = ro + rd * sd;
        }
        att = att * albedo_at(s, m, hit.u, hit.v);

        // next-event estimation (src/path_tracer.cu:2962-3007)
        float pe = env_select_prob(s.env, nl);
        if ((m.type == MCPT_MAT_LAMBERTIAN || m.type == MCPT_MAT_ISOTROPIC) && (nl > 0 || pe > 0.0f)) {
            V3 dl, lc;
            float dist, sapdf;
            if (pe > 0.0f && (nl == 0 || rng.u() < pe)) {
                float dp;
                dl = env_sample(s.env, rng, dp);
                lc = env_radiance(s.env, dl, V3(0.f, 0.f, 0.f));
                dist = 1e30f;
                sapdf = pe * dp;
            } else {
                int li = std::min((int)(rng.u() * nl), nl - 1);
                const mcpt_light& L = s.lights[li];
                float lsp = (1.0f - pe) / float(nl);
                V3 pl = sample_triangle(L.tri, rng);
                V3 tl = pl - so;
                dist = tl.length();
                dl = tl / dist;
                lc = v3(L.emission);
                float cl = fabsf(v3(L.normal).dot(dl));
                sapdf = (cl > 1e-4f) ? lsp * (dist * dist) / (cl * L.area) : 0.0f;
            }
            bool iso = (m.type == MCPT_MAT_ISOTROPIC);
            float cs = iso ? 1.0f : hit.normal.dot(dl);
            if (cs > 0.0f && sapdf > 0.0f) {
                float tr = medium_transmittance(s, so, dl, dist * 0.999f, time, pc);
                if (tr > 0.0f) {
                    float php = hg_pdf(rd.dot(dl), m.phase_g);
                    float spdf = iso ? php : lambert_pdf(cs);
                    float sval = iso ? php : cs / MCPT_PI_F;
                    float nw = power_heuristic(sapdf, spdf);
                    col = col + att * lc * (sval / sapdf) * nw * tr;
                }
            }
        }
        // Russian roulette (src/path_tracer.cu:3011-3017)
        if (depth > 3) {
            float mc = fmaxf(att.x, fmaxf(att.y, att.z));
            float cp = fminf(mc, cam.rr_cap);
            if (rng.u() > cp) break;
            att = att / cp;
        }
        V3 nd;
        switch (m.type) {
            case MCPT_MAT_CONDUCTOR:
                nd = conductor_dir(rd, hit.normal, m.fuzz, rng);
                break;
            case MCPT_MAT_DIELECTRIC: {
                float eta = hit.front ? 1.0f / m.ior : m.ior;
                nd = dielectric_dir(rd, hit.normal, eta, rng);
                break;
            }
            case MCPT_MAT_ISOTROPIC:
                nd = hg_dir(rd, m.phase_g, rng);
                break;
            default:
                nd = lambertian_dir(hit.normal, rng);
                break;
        }
        prev_spec = (m.type != MCPT_MAT_LAMBERTIAN && m.type != MCPT_MAT_ISOTROPIC);
        prev_pdf = (m.type == MCPT_MAT_ISOTROPIC) ? hg_pdf(rd.dot(nd), m.phase_g)
                                                  : lambert_pdf(fmaxf(nd.dot(hit.normal), 0.0f));
        ro = so;
        rd = nd;
    }
    // firefly clamp (src/path_tracer.cu:3050-3055)
    float b = fmaxf(col.x, fmaxf(col.y, col.z));
    if (b > 10.0f) col = col * (10.0f / b);
    return col;
}

int hw_threads(int t) {
    if (t > 0) return t;
    unsigned h = std::thread::hardware_concurrency();
    return h ? (int)h : 1;
}

template <class F>
void parallel_for(int n, int chunk, int threads, F f) {
    threads = hw_threads(threads);
    std::atomic<int> next{0};
    auto work = [&]() {
        while (true) {
            int b = next.fetch_add(chunk);
            if (b >= n) break;
            int e = std::min(n, b + chunk);
            for (int i = b; i < e; i++) f(i);
        }
    };
    if (threads == 1) {
        work();
        return;
    }
    std::vector<std::thread> ts;
    for (int i = 0; i < threads; i++) ts.emplace_back(work);
    for (auto& t : ts) t.join();
}

}  // namespace

// ============================================================== C interface
extern "C" {

int oracle_camera_init(const mcpt_camera_desc* d, mcpt_camera* c) {
    // Camera::initialize (src/path_tracer.cu:1610-1631)
    float aspect = d->width / float(d->height);
    int width = d->width;
    int height = int(width / aspect);
    if (height != d->height) return -1;
    V3 from = v3(d->look_from), at = v3(d->look_at), up = v3(d->up);
    V3 w = (from - at).normalize();
    V3 u = up.cross(w).normalize();
    V3 v = w.cross(u);
    float lens_radius = d->aperture / 2.;
    V3 center = from;
    float theta = d->vfov_degrees * (3.14159265358979323846f / 180.0f);
    float h = tanf(theta / 2.0f);
    float vh = 2.0f * h * d->focus_distance;
    float vw = vh * (float(width) / height);
    V3 vu = u * vw;
    V3 vv = v * -vh;
    V3 du = vu / float(width);
    V3 dv = vv / float(height);
    V3 ul = center - (w * d->focus_distance) - vu / 2.0f - vv / 2.0f;
    V3 p00 = ul + (du + dv) * 0.5f;
    auto put = [](float* o, const V3& x) {
        o[0] = x.x;
        o[1] = x.y;
        o[2] = x.z;
    };
    std::memset(c, 0, sizeof(*c));
    put(c->center, center);
    put(c->pixel00, p00);
    put(c->delta_u, du);
    put(c->delta_v, dv);
    put(c->u, u);
    put(c->v, v);
    put(c->w, w);
    c->lens_radius = lens_radius;
    c->width = width;
    c->height = height;
    for (int k = 0; k < 3; k++) c->background[k] = d->background[k];
    c->rr_cap = d->rr_cap;
    return 0;
}

void oracle_env_tables(const float* rgba, int w, int h, float* row_cdf, float* col_cdf, float* total_weight) {
    // buildEnvironmentSamplingTables (src/path_tracer.cu:370-385)
    for (int i = 0; i <= h; i++) row_cdf[i] = 0.0f;
    for (size_t i = 0; i < (size_t)h * (w + 1); i++) col_cdf[i] = 0.0f;
    for (int row = 0; row < h; row++) {
        float st = mf_sin(MCPT_PI_F * (row + 0.5f) / h);
        float* cdf = &col_cdf[(size_t)row * (w + 1)];
        cdf[0] = 0.0f;
        for (int col = 0; col < w; col++) {
            const float* t = &rgba[((size_t)row * w + col) * 4];
            float lum = 0.2126f * t[0] + 0.7152f * t[1] + 0.0722f * t[2];
            cdf[col + 1] = cdf[col] + fmaxf(lum, 0.0f) * st;
        }
        row_cdf[row + 1] = row_cdf[row] + cdf[w];
    }
    *total_weight = row_cdf[h];
}

namespace {
struct FixedUniforms {  // replays given uniforms in draw order
    const float* v;
    int i = 0;
    float u() { return v[i++]; }
};
}  // namespace

int oracle_sampling(int fn, const float* in, int n, float* out) {
    static const int ins[] = {7, 2, 8, 6, 2}, outs[] = {3, 1, 3, 3, 1};
    if (fn < 0 || fn > 4) return -1;
    for (int i = 0; i < n; i++) {
        const float* a = in + (size_t)i * ins[fn];
        float* o = out + (size_t)i * outs[fn];
        V3 d;
        switch (fn) {
            case 0: d = v3(a).refract(v3(a + 3), a[6]); break;
            case 1: o[0] = schlick(a[0], a[1]); continue;
            case 2: {
                FixedUniforms r{a + 7};
                d = dielectric_dir(v3(a), v3(a + 3), a[6], r);
                break;
            }
            case 3: {
                FixedUniforms r{a + 4};
                d = hg_dir(v3(a), a[3], r);
                break;
            }
            default: o[0] = hg_pdf(a[0], a[1]); continue;
        }
        o[0] = d.x;
        o[1] = d.y;
        o[2] = d.z;
    }
    return 0;
}

void oracle_conductor(const float* in_dir, const float* normal, float fuzz, int n, uint64_t seed, float* out) {
    for (int i = 0; i < n; i++) {
        Rng rng;
        mcpt_pcg32_for_path(&rng.g, (uint32_t)i, 0u, seed);
        const V3 d = conductor_dir(v3(in_dir), v3(normal), fuzz, rng);
        out[3 * i] = d.x;
        out[3 * i + 1] = d.y;
        out[3 * i + 2] = d.z;
    }
}

void oracle_camera_rays(const mcpt_camera* cam, const int32_t* pixels, int n, uint32_t sample, uint64_t seed,
                        float* out) {
    for (int i = 0; i < n; i++) {
        const int px = pixels[2 * i], py = pixels[2 * i + 1];
        Rng rng;
        mcpt_pcg32_for_path(&rng.g, (uint32_t)(py * cam->width + px), sample, seed);
        V3 ro, rd;
        float time;
        camera_ray(*cam, px, py, rng, ro, rd, time);
        float* o = out + 7 * i;
        o[0] = ro.x, o[1] = ro.y, o[2] = ro.z;
        o[3] = rd.x, o[4] = rd.y, o[5] = rd.z;
        o[6] = time;
    }
}

void oracle_tri_interp(const mcpt_tri_hot* hot, const mcpt_tri_cold* cold, const float* points, int n, float* uv,
                       float* normal) {
    const V3 face = (v3(hot->v1) - v3(hot->v0)).cross(v3(hot->v2) - v3(hot->v0)).normalize();
    for (int i = 0; i < n; i++) {
        const V3 p = v3(points + 3 * i);
        tri_uv(*hot, *cold, p, uv[2 * i], uv[2 * i + 1]);
        const V3 sn = tri_shading_normal(*hot, *cold, p, face);
        normal[3 * i] = sn.x, normal[3 * i + 1] = sn.y, normal[3 * i + 2] = sn.z;
    }
}

int oracle_env_probe(const float* rgba, int w, int h, float rot, int fn, const float* in, int n, float* out) {
    if (fn < 0 || fn > 3 || w < 1 || h < 1) return -1;
    Env e;
    e.enabled = true;
    e.w = w;
    e.h = h;
    e.rot = rot;
    e.rgba.assign(rgba, rgba + (size_t)w * h * 4);
    e.row_cdf.resize(h + 1);
    e.col_cdf.resize((size_t)h * (w + 1));
    oracle_env_tables(e.rgba.data(), w, h, e.row_cdf.data(), e.col_cdf.data(), &e.total);
    static const int ins[] = {4, 3, 3, 2}, outs[] = {4, 1, 2, 3};
    for (int i = 0; i < n; i++) {
        const float* a = in + (size_t)i * ins[fn];
        float* o = out + (size_t)i * outs[fn];
        if (fn == 0) {
            FixedUniforms r{a};
            const V3 d = env_sample(e, r, o[3]);
            o[0] = d.x;
            o[1] = d.y;
            o[2] = d.z;
        } else if (fn == 1) {
            o[0] = env_dir_pdf(e, v3(a));
        } else if (fn == 2) {
            env_uv(e, v3(a).normalize(), o[0], o[1]);
        } else {
            const V3 d = env_direction(e, a[0], a[1]);
            o[0] = d.x;
            o[1] = d.y;
            o[2] = d.z;
        }
    }
    return 0;
}

uint32_t oracle_expand_bits(uint32_t v) { return expand_bits(v); }
uint32_t oracle_quantize(float v, float lo, float hi) { return quantize(v, lo, hi); }

// finalizeScene (src/path_tracer.cu:1930-1972) over the primitive list:
// spheres 0..S-1, then the triangles
static void finalize_prims(const mcpt_tri_hot* hot, int T, const mcpt_sphere* sph, int S, mcpt_aabb* tri_bounds,
                           mcpt_aabb* scene_bounds, float* centroids, uint32_t* morton) {
    const int n = S + T;
    std::vector<AABB> boxes(n);
    AABB sb;
    for (int i = 0; i < n; i++) {
        boxes[i] = i < S ? sphere_box(sph[i]) : tri_box(hot[i - S]);
        sb = (i == 0) ? boxes[i] : merge(sb, boxes[i]);
    }
    for (int i = 0; i < n; i++) {
        V3 c;
        if (i < S)  // sphere::calculateCentroid (:858-860)
            c = sph[i].moving ? (v3(sph[i].center) + v3(sph[i].center1)) * 0.5f : v3(sph[i].center);
        else
            c = (v3(hot[i - S].v0) + v3(hot[i - S].v1) + v3(hot[i - S].v2)) / 3.;
        uint32_t qx = quantize(c.x, sb.mn.x, sb.mx.x);
        uint32_t qy = quantize(c.y, sb.mn.y, sb.mx.y);
        uint32_t qz = quantize(c.z, sb.mn.z, sb.mx.z);
        if (morton) morton[i] = expand_bits(qx) | expand_bits(qy) << 1 | expand_bits(qz) << 2;
        if (centroids) {
            centroids[3 * i] = c.x;
            centroids[3 * i + 1] = c.y;
            centroids[3 * i + 2] = c.z;
        }
        if (tri_bounds) {
            mcpt_aabb& b = tri_bounds[i];
            b.min[0] = boxes[i].mn.x;
            b.min[1] = boxes[i].mn.y;
            b.min[2] = boxes[i].mn.z;
            b.max[0] = boxes[i].mx.x;
            b.max[1] = boxes[i].mx.y;
            b.max[2] = boxes[i].mx.z;
            b.pad0 = b.pad1 = 0.f;
        }
    }
    if (scene_bounds && n > 0) {
        scene_bounds->min[0] = sb.mn.x;
        scene_bounds->min[1] = sb.mn.y;
        scene_bounds->min[2] = sb.mn.z;
        scene_bounds->max[0] = sb.mx.x;
        scene_bounds->max[1] = sb.mx.y;
        scene_bounds->max[2] = sb.mx.z;
        scene_bounds->pad0 = scene_bounds->pad1 = 0.f;
    }
}

void oracle_finalize(const mcpt_tri_hot* hot, int n, mcpt_aabb* tri_bounds, mcpt_aabb* scene_bounds, float* centroids,
                     uint32_t* morton) {
    finalize_prims(hot, n, nullptr, 0, tri_bounds, scene_bounds, centroids, morton);
}

void oracle_sort_stable(const uint32_t* morton, int n, uint32_t* keys, int32_t* ids) {
    std::vector<int32_t> idx(n);
    for (int i = 0; i < n; i++) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return morton[a] < morton[b]; });
    for (int i = 0; i < n; i++) {
        ids[i] = idx[i];
        keys[i] = morton[idx[i]];
    }
}

void oracle_sort_reference(const uint32_t* morton, int n, uint32_t* keys, int32_t* ids) {
    struct Ref {
        uint32_t code;
        int32_t idx;
    };
    std::vector<Ref> v(n);
    for (int i = 0; i < n; i++) v[i] = Ref{morton[i], i};
    std::sort(v.begin(), v.end(), [](const Ref& a, const Ref& b) { return a.code < b.code; });
    for (int i = 0; i < n; i++) {
        keys[i] = v[i].code;
        ids[i] = v[i].idx;
    }
}

void oracle_karras(const uint32_t* k, int n, int32_t* left, int32_t* right, uint8_t* left_leaf, uint8_t* right_leaf,
                   int32_t* parent, int32_t* leaf_parent) {
    if (n > 1) parent[0] = -1;
    for (int idx = 0; idx < n - 1; idx++) {
        // determineRange (src/path_tracer.cu:1349-1369)
        int d = (delta(k, n, idx, idx + 1) - delta(k, n, idx, idx - 1)) >= 0 ? 1 : -1;
        int dmin = delta(k, n, idx, idx - d);
        int lmax = 2;
        while (delta(k, n, idx, idx + lmax * d) > dmin) lmax *= 2;
        int l = 0;
        for (int t = lmax / 2; t >= 1; t /= 2)
            if (delta(k, n, idx, idx + (l + t) * d) > dmin) l += t;
        int j = idx + l * d;
        int first = std::min(idx, j), last = std::max(idx, j);
        // findSplit (src/path_tracer.cu:1371-1386)
        int cp = delta(k, n, first, last);
        int split = first, step = last - first;
        do {
            step = (step + 1) >> 1;
            int ns = split + step;
            if (ns < last && delta(k, n, first, ns) > cp) split = ns;
        } while (step > 1);
        // buildInternalNodesKernel child wiring (src/path_tracer.cu:1425-1446)
        left[idx] = split;
        left_leaf[idx] = (split == first);
        if (split == first) leaf_parent[split] = idx;
        else parent[split] = idx;
        right[idx] = split + 1;
        right_leaf[idx] = (split + 1 == last);
        if (split + 1 == last) leaf_parent[split + 1] = idx;
        else parent[split + 1] = idx;
    }
}

int oracle_scene_create(const mcpt_tri_hot* hot, const mcpt_tri_cold* cold, int n, const mcpt_material* mats,
                        int n_mats, const mcpt_env_desc* env, oracle_scene* out) {
    if (n < 1) return -1;
    return oracle_scene_create_ex(hot, cold, n, nullptr, 0, mats, n_mats, env, out);
}

int oracle_scene_create_ex(const mcpt_tri_hot* hot, const mcpt_tri_cold* cold, int n_tris, const mcpt_sphere* spheres,
                           int n_spheres, const mcpt_material* mats, int n_mats, const mcpt_env_desc* env,
                           oracle_scene* out) {
    const int T = n_tris, S = n_spheres, n = S + T;
    if (T < 0 || S < 0 || n < 1 || (T && (!hot || !cold)) || (S && !spheres) || !mats || n_mats < 1) return -1;
    for (int i = 0; i < T; i++)
        if (cold[i].material < 0 || cold[i].material >= n_mats) return -1;
    for (int i = 0; i < S; i++)
        if (spheres[i].material < 0 || spheres[i].material >= n_mats || !(spheres[i].radius > 0.0f)) return -1;
    oracle_scene_s* s = new oracle_scene_s();
    s->hot.assign(hot, hot + T);
    s->cold.assign(cold, cold + T);
    s->spheres.assign(spheres, spheres + S);
    s->mats.assign(mats, mats + n_mats);
    s->n = n;
    // lights (finalizeScene, src/path_tracer.cu:1973-1985): emissive triangles
    for (int i = 0; i < T; i++) {
        if (mats[cold[i].material].type != MCPT_MAT_EMISSIVE) continue;
        V3 e1 = v3(hot[i].v1) - v3(hot[i].v0);
        V3 e2 = v3(hot[i].v2) - v3(hot[i].v0);
        V3 av = e1.cross(e2);
        mcpt_light L;
        L.tri = hot[i];
        for (int k = 0; k < 3; k++) L.emission[k] = mats[cold[i].material].emission[k];
        L.area = 0.5f * av.length();
        V3 nn = av.normalize();
        L.normal[0] = nn.x;
        L.normal[1] = nn.y;
        L.normal[2] = nn.z;
        s->lights.push_back(L);
    }
    if (env && env->rgba && env->width > 0 && env->height > 0) {
        Env& e = s->env;
        e.enabled = true;
        e.w = env->width;
        e.h = env->height;
        e.intensity = env->intensity;
        e.rot = env->rotation_radians;
        e.rgba.assign(env->rgba, env->rgba + (size_t)e.w * e.h * 4);
        e.row_cdf.resize(e.h + 1);
        e.col_cdf.resize((size_t)e.h * (e.w + 1));
        oracle_env_tables(e.rgba.data(), e.w, e.h, e.row_cdf.data(), e.col_cdf.data(), &e.total);
    }
    // BVH
    s->morton.resize(n);
    std::vector<mcpt_aabb> tb(n);
    mcpt_aabb sb;
    finalize_prims(hot, T, spheres, S, tb.data(), &sb, nullptr, s->morton.data());
    s->tri_boxes.resize(n);
    for (int i = 0; i < n; i++) {
        s->tri_boxes[i].mn = v3(tb[i].min);
        s->tri_boxes[i].mx = v3(tb[i].max);
    }
    s->bounds.mn = v3(sb.min);
    s->bounds.mx = v3(sb.max);
    s->keys.resize(n);
    s->ids.resize(n);
    oracle_sort_stable(s->morton.data(), n, s->keys.data(), s->ids.data());
    s->leaf_prim.resize(n);
    for (int i = 0; i < n; i++) s->leaf_prim[i] = s->ids[i] < S ? ~s->ids[i] : s->ids[i] - S;
    s->leaf_box.resize(n);
    for (int i = 0; i < n; i++) s->leaf_box[i] = s->tri_boxes[s->ids[i]];
    s->leaf_parent.assign(n, -1);
    if (n == 1) {
        // A single primitive has no internal node in the reference (its
        // traversal would read d_internalNodes[0] out of bounds). Here the
        // root holds the leaf on the left and an empty box on the right.
        Node r;
        r.left = 0;
        r.left_leaf = true;
        r.right = 0;
        r.right_leaf = true;
        r.parent = -1;
        r.box = s->leaf_box[0];
        s->nodes.push_back(r);
        s->leaf_parent[0] = 0;
        s->max_depth = 1;
        *out = s;
        return 0;
    }
    std::vector<int32_t> L(n - 1), R(n - 1), P(n - 1);
    std::vector<uint8_t> LL(n - 1), RL(n - 1);
    oracle_karras(s->keys.data(), n, L.data(), R.data(), LL.data(), RL.data(), P.data(), s->leaf_parent.data());
    s->nodes.resize(n - 1);
    for (int i = 0; i < n - 1; i++) {
        s->nodes[i].left = L[i];
        s->nodes[i].right = R[i];
        s->nodes[i].left_leaf = LL[i];
        s->nodes[i].right_leaf = RL[i];
        s->nodes[i].parent = P[i];
    }
    // refit (refitBoundsKernel, src/path_tracer.cu:1398-1415): a post-order
    // walk computes the same fmin/fmax merges the atomic climb does.
    std::vector<int> order;
    order.reserve(n - 1);
    std::vector<std::pair<int, int>> st;  // node, depth
    st.push_back({0, 1});
    int maxd = 0;
    while (!st.empty()) {
        auto [x, dep] = st.back();
        st.pop_back();
        order.push_back(x);
        const Node& nd = s->nodes[x];
        if (nd.left_leaf || nd.right_leaf) maxd = std::max(maxd, dep);
        if (!nd.left_leaf) st.push_back({nd.left, dep + 1});
        if (!nd.right_leaf) st.push_back({nd.right, dep + 1});
    }
    s->max_depth = maxd;
    for (int i = (int)order.size() - 1; i >= 0; i--) {
        Node& nd = s->nodes[order[i]];
        const AABB& a = nd.left_leaf ? s->leaf_box[nd.left] : s->nodes[nd.left].box;
        const AABB& b = nd.right_leaf ? s->leaf_box[nd.right] : s->nodes[nd.right].box;
        nd.box = merge(a, b);
    }
    *out = s;
    return 0;
}

void oracle_scene_destroy(oracle_scene s) { delete s; }

void oracle_sphere_uv(const float* normals, int n, float* uv) {
    for (int i = 0; i < n; i++) sphere_uv(v3(normals + 3 * i), uv[2 * i], uv[2 * i + 1]);
}

void oracle_sample_albedo(oracle_scene s, const mcpt_material* m, const float* uv, int n, float* out) {
    for (int i = 0; i < n; i++) {
        const V3 c = albedo_at(*s, *m, uv[2 * i], uv[2 * i + 1]);
        out[3 * i] = c.x;
        out[3 * i + 1] = c.y;
        out[3 * i + 2] = c.z;
    }
}

// A density grid for every isotropic material (the reference attaches a
// loaded volume to them all, src/path_tracer.cu:2509-2519); max_density is
// the grid's largest value. Returns -1 on a malformed grid.
int oracle_scene_set_density_grid(oracle_scene s, const mcpt_density_grid* g, float* max_density) {
    for (int a = 0; a < 3; a++)
        if (g->dims[a] < 1 || !(g->voxel_size[a] > 0.0f)) return -1;
    const size_t n = (size_t)g->dims[0] * g->dims[1] * g->dims[2];
    s->grid.assign(g->density, g->density + n);
    float mx = 0.0f;
    for (float v : s->grid) mx = std::max(mx, v);
    for (int a = 0; a < 3; a++) {
        s->gdim[a] = g->dims[a];
        s->gorg[a] = g->origin[a];
        s->ginv[a] = 1.0f / g->voxel_size[a];
    }
    s->gmax = mx;
    if (max_density) *max_density = mx;
    return 0;
}

void oracle_grid_density(oracle_scene s, const float* pts, int n, float* out) {
    for (int i = 0; i < n; i++) out[i] = grid_density(*s, V3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]));
}

// fn 0: delta-tracking free flight, 1: ratio-tracking transmittance, for n
// rays of 7 floats (origin, direction, length), ray i drawing from the path
// stream (pixel i, sample 0, seed); sigma is the extinction scale
void oracle_grid_track(oracle_scene s, int fn, float sigma, const float* rays, int n, uint64_t seed, float* out) {
    for (int i = 0; i < n; i++) {
        const float* r = rays + 7 * i;
        Rng rng;
        mcpt_pcg32_for_path(&rng.g, (uint32_t)i, 0u, seed);
        const V3 o(r[0], r[1], r[2]), d(r[3], r[4], r[5]);
        out[i] = fn == 0 ? grid_free_flight(*s, sigma, o, d, r[6], rng) : grid_transmittance(*s, sigma, o, d, r[6], rng);
    }
}

int oracle_scene_add_texture(oracle_scene s, const mcpt_texture_desc* d) {
    Tex t;
    t.type = d->type;
    if (d->type == MCPT_TEX_CHECKER) {
        t.scale = d->scale;
        for (int k = 0; k < 3; k++) {  // make_checker_texture (:211-213)
            t.ca[k] = srgb_to_linear(d->color_a[k]);
            t.cb[k] = srgb_to_linear(d->color_b[k]);
        }
    } else if (d->type == MCPT_TEX_IMAGE && d->width > 0 && d->height > 0 && d->rgba8) {
        t.w = d->width;
        t.h = d->height;
        t.px.resize((size_t)t.w * t.h);
        std::memcpy(t.px.data(), d->rgba8, t.px.size() * 4);
    } else {
        return -1;
    }
    for (int i = 0; i < 256; i++) s->srgb8[i] = srgb_to_linear((float)i / 255.0f);
    s->tex.push_back(std::move(t));
    return (int)s->tex.size() - 1;
}

int oracle_scene_lights(oracle_scene s, mcpt_light* out, int cap) {
    int nl = (int)s->lights.size();
    for (int i = 0; i < nl && i < cap; i++) out[i] = s->lights[i];
    return nl;
}

int oracle_scene_dump(oracle_scene s, mcpt_lbvh_dump* d) {
    int n = s->n;
    auto box = [](mcpt_aabb& o, const AABB& b) {
        o.min[0] = b.mn.x;
        o.min[1] = b.mn.y;
        o.min[2] = b.mn.z;
        o.max[0] = b.mx.x;
        o.max[1] = b.mx.y;
        o.max[2] = b.mx.z;
        o.pad0 = o.pad1 = 0.f;
    };
    if (d->morton) std::copy(s->morton.begin(), s->morton.end(), d->morton);
    if (d->sorted_keys) std::copy(s->keys.begin(), s->keys.end(), d->sorted_keys);
    if (d->sorted_ids) std::copy(s->ids.begin(), s->ids.end(), d->sorted_ids);
    int ni = n - 1;
    for (int i = 0; i < ni; i++) {
        const Node& nd = s->nodes[i];
        if (d->left) d->left[i] = nd.left;
        if (d->right) d->right[i] = nd.right;
        if (d->left_is_leaf) d->left_is_leaf[i] = nd.left_leaf;
        if (d->right_is_leaf) d->right_is_leaf[i] = nd.right_leaf;
        if (d->parent) d->parent[i] = nd.parent;
        if (d->internal_bounds) box(d->internal_bounds[i], nd.box);
    }
    for (int i = 0; i < n; i++) {
        if (d->leaf_parent) d->leaf_parent[i] = s->leaf_parent[i];
        if (d->leaf_bounds) box(d->leaf_bounds[i], s->leaf_box[i]);
    }
    if (d->scene_bounds) box(*d->scene_bounds, s->bounds);
    d->max_depth = s->max_depth;
    d->build_ms = 0.f;
    return 0;
}

void oracle_trace_closest(oracle_scene s, const mcpt_ray* rays, int n, mcpt_packed_hit* hits, uint32_t* steps,
                          uint32_t* leaves, int threads) {
    parallel_for(n, 256, threads, [&](int i) {
        const mcpt_ray& r = rays[i];
        Closest c = traverse_closest(*s, v3(r.origin), v3(r.dir), r.tmax, r.time, steps ? &steps[i] : nullptr,
                                     leaves ? &leaves[i] : nullptr);
        hits[i].t = c.t;
        hits[i].prim_index = !c.hit ? -1 : c.prim >= 0 ? c.prim : ~c.prim;
        hits[i].prim_type = !c.hit ? -1 : c.prim >= 0 ? 1 : 0;
        hits[i].did_hit = c.hit ? 1 : 0;
    });
}

void oracle_trace_any(oracle_scene s, const mcpt_ray* rays, int n, uint8_t* occluded, int threads) {
    parallel_for(n, 256, threads, [&](int i) {
        const mcpt_ray& r = rays[i];
        occluded[i] = traverse_any(*s, v3(r.origin), v3(r.dir), r.tmax, r.time) ? 1 : 0;
    });
}

void oracle_brute_closest(const mcpt_tri_hot* hot, int n_tris, const mcpt_ray* rays, int n, mcpt_packed_hit* hits) {
    parallel_for(n, 64, 0, [&](int i) {
        const mcpt_ray& r = rays[i];
        V3 o = v3(r.origin), d = v3(r.dir);
        Closest c{r.tmax, -1, false};
        for (int k = 0; k < n_tris; k++) {
            float t;
            if (hit_triangle(hot[k], o, d, t) && t < c.t && t > 0.0001f) c = Closest{t, k, true};
        }
        hits[i].t = c.t;
        hits[i].prim_index = c.hit ? c.prim : -1;
        hits[i].prim_type = c.hit ? 1 : -1;
        hits[i].did_hit = c.hit ? 1 : 0;
    });
}

int oracle_render(oracle_scene s, const mcpt_camera* cam, int spp, int sample_begin, int max_depth, uint64_t seed,
                  int row_begin, int row_end, float* accum, double* accum64, int threads,
                  oracle_render_stats* stats) {
    const int W = cam->width, H = cam->height;
    if (row_begin < 0) row_begin = 0;
    if (row_end <= 0 || row_end > H) row_end = H;
    if (row_end <= row_begin || spp < 1) return -1;
    const int tiles_x = (W + 15) / 16;
    const int tiles_y = (row_end - row_begin + 15) / 16;
    std::atomic<uint64_t> ext{0}, shd{0}, stp{0}, lvs{0};
    auto t0 = std::chrono::steady_clock::now();
    parallel_for(tiles_x * tiles_y, 1, threads, [&](int tile) {
        int tx = tile % tiles_x, ty = tile / tiles_x;
        PathCounts pc;
        for (int y = row_begin + ty * 16; y < std::min(row_end, row_begin + ty * 16 + 16); y++) {
            for (int x = tx * 16; x < std::min(W, tx * 16 + 16); x++) {
                int p = y * W + x;
                for (int k = 0; k < spp; k++) {
                    Rng rng;
                    mcpt_pcg32_for_path(&rng.g, (uint32_t)p, (uint32_t)(sample_begin + k), seed);
                    V3 c = trace_path(*s, *cam, x, y, rng, max_depth, pc);
                    accum[3 * p] = accum[3 * p] + c.x;
                    accum[3 * p + 1] = accum[3 * p + 1] + c.y;
                    accum[3 * p + 2] = accum[3 * p + 2] + c.z;
                    if (accum64) {
                        accum64[3 * p] += c.x;
                        accum64[3 * p + 1] += c.y;
                        accum64[3 * p + 2] += c.z;
                    }
                }
            }
        }
        ext += pc.ext;
        shd += pc.shd;
        stp += pc.steps;
        lvs += pc.leaves;
    });
    auto t1 = std::chrono::steady_clock::now();
    if (stats) {
        stats->extend_rays = ext;
        stats->shadow_rays = shd;
        stats->extend_steps = stp;
        stats->extend_leaves = lvs;
        stats->paths = (uint64_t)W * (row_end - row_begin) * spp;
        stats->seconds = std::chrono::duration<double>(t1 - t0).count();
        stats->threads = hw_threads(threads);
    }
    return 0;
}

// linearToSrgb / toneMapChannel / resolve_kernel (src/path_tracer.cu:3062-3100)
static float linear_to_srgb(float x) {
    x = fmaxf(x, 0.0f);
    if (x <= 0.0031308f) return 12.92f * x;
    return 1.055f * mf_pow(x, 1.0f / 2.4f) - 0.055f;
}
static float tonemap_channel(float x, int tm) {
    x = fmaxf(x, 0.f);
    if (tm == MCPT_TONEMAP_REINHARD) return x / (1.f + x);
    if (tm == MCPT_TONEMAP_ACES) {
        float m = (x * (2.51f * x + 0.03f)) / (x * (2.43f * x + 0.59f) + 0.14f);
        return fminf(fmaxf(m, 0.0f), 1.0f);
    }
    return x;
}
void oracle_resolve(const float* accum, int w, int h, int spp, int tonemap, float exposure, float* out) {
    for (int p = 0; p < w * h; p++) {
        for (int k = 0; k < 3; k++) {
            float a = accum[3 * p + k] / float(spp) * exposure;
            out[3 * p + k] = linear_to_srgb(tonemap_channel(a, tonemap));
        }
    }
}

// main's output conversion after resolve_kernel (src/path_tracer.cu:4366-4380)
void oracle_resolve_rgb8(const float* accum, int w, int h, int spp, int tonemap, float exposure, uint8_t* out) {
    std::vector<float> f((size_t)w * h * 3);
    oracle_resolve(accum, w, h, spp, tonemap, exposure, f.data());
    for (size_t i = 0; i < f.size(); i++) {
        const float c = f[i];
        const float v = std::isfinite(c) ? fminf(fmaxf(c, 0.0f), 1.0f) : 0.0f;
        out[i] = (uint8_t)(255.99f * v);
    }
}

// logLuminanceSumKernel + computeAutoExposure (src/path_tracer.cu:3101-3124),
// the per-pixel terms in FP32, their sum in FP64 in pixel order
float oracle_auto_exposure(const float* accum, int w, int h, int spp, float target) {
    double sum = 0.0;
    const int n = w * h;
    for (int p = 0; p < n; p++) {
        const float r = accum[3 * p] / float(spp), g = accum[3 * p + 1] / float(spp), b = accum[3 * p + 2] / float(spp);
        const float lum = 0.2126f * r + 0.7152f * g + 0.0722f * b;
        sum += (double)mf_log(1e-4f + (std::isfinite(lum) ? fmaxf(lum, 0.0f) : 0.0f));
    }
    return target / mf_exp((float)(sum / (double)n));
}

void oracle_fmath(int fn, const float* x, const float* y, float* out, int n) {
    for (int i = 0; i < n; i++) {
        switch (fn) {
            case 0: out[i] = mf_sin(x[i]); break;
            case 1: out[i] = mf_cos(x[i]); break;
            case 2: out[i] = mf_asin(x[i]); break;
            case 3: out[i] = mf_acos(x[i]); break;
            case 4: out[i] = mf_atan2(x[i], y[i]); break;
            case 5: out[i] = mf_log(x[i]); break;
            case 6: out[i] = mf_exp(x[i]); break;
            default: out[i] = mf_pow(x[i], y[i]); break;
        }
    }
}

void oracle_pcg32(uint32_t pixel, uint32_t sample, uint64_t seed, int n, uint32_t* out) {
    mcpt_pcg32 g;
    mcpt_pcg32_for_path(&g, pixel, sample, seed);
    for (int i = 0; i < n; i++) out[i] = mcpt_pcg32_next(&g);
}

}  // extern "C"
