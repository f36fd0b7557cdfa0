// device.hpp — types and per-ray arithmetic shared by the gfx950 kernels.
//
// Every floating-point expression keeps the operand order of the
// reference's device functions (MonoCUDA-PT src/path_tracer.cu, cited per
// function) and the library is compiled with -ffp-contract=off and IEEE
// division/sqrt, so a ray produces the same bits here as in a host build of
// the same formulas. Elementary functions come from mcpt_fmath.h, random
// numbers from mcpt_pcg32.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "mcpt.h"
#include "mcpt_fmath.h"
#include "mcpt_pcg32.h"

#define MCPT_DEV __device__ __forceinline__
#define MCPT_HDI __host__ __device__ __forceinline__

namespace mcpt {

// ------------------------------------------------------------------ vec3
// src/path_tracer.cu:105-163
struct vec3 {
    float x, y, z;
    MCPT_HDI vec3() : x(0), y(0), z(0) {}
    MCPT_HDI vec3(float a, float b, float c) : x(a), y(b), z(c) {}
    MCPT_HDI vec3 operator+(const vec3& o) const { return vec3(x + o.x, y + o.y, z + o.z); }
    MCPT_HDI vec3 operator-(const vec3& o) const { return vec3(x - o.x, y - o.y, z - o.z); }
    MCPT_HDI vec3 operator*(const vec3& o) const { return vec3(x * o.x, y * o.y, z * o.z); }
    MCPT_HDI vec3 operator-() const { return vec3(-x, -y, -z); }
    MCPT_HDI vec3 operator*(float n) const { return vec3(x * n, y * n, z * n); }
    MCPT_HDI vec3 operator/(float n) const { return vec3(x / n, y / n, z / n); }
    MCPT_HDI float length_squared() const { return x * x + y * y + z * z; }
    MCPT_HDI float length() const { return __builtin_sqrtf(length_squared()); }
    MCPT_HDI vec3 cross(const vec3& o) const { return vec3(y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x); }
    MCPT_HDI float dot(const vec3& o) const { return x * o.x + y * o.y + z * o.z; }
    MCPT_HDI vec3 normalize() const {
        float l = __builtin_sqrtf(length_squared());
        return vec3(x / l, y / l, z / l);
    }
    MCPT_HDI vec3 reflect(const vec3& n) const { return *this - n * (2.0f * dot(n)); }
    // src/path_tracer.cu:146-163
    MCPT_HDI vec3 refract(const vec3& n, float eta) const {
        vec3 d = normalize();
        float cos1 = fminf((-d).dot(n), 1.0f);
        float theta1 = mf_acos(cos1);
        float sin2 = mf_sin(theta1) * eta;
        if (sin2 > 1.f) return d.reflect(n);
        float theta2 = mf_asin(sin2);
        vec3 tangent = d + n * cos1;
        float tl = tangent.length();
        if (tl > 1e-8f) tangent = tangent / tl;
        return tangent * mf_sin(theta2) - n * mf_cos(theta2);
    }
};
MCPT_HDI vec3 load3(const float* p) { return vec3(p[0], p[1], p[2]); }

// ------------------------------------------------------------- BVH layout
// One 64-byte record per internal node holding both children's boxes and
// references, so a traversal step is four 16-byte loads from one 64-byte
// segment. (The reference stores one box per node and reads the two
// children's 64-byte BVH_Node records, src/path_tracer.cu:1006-1015.)
// Child reference r >= 0: internal node r; r < 0: leaf ~r.
struct alignas(16) BvhNode {
    float4 a;  // left.min.xyz, left.max.x
    float4 b;  // left.max.yz, right.min.xy
    float4 c;  // right.min.z, right.max.xyz
    int4 r;    // left ref, right ref, parent (node<<1 | side, root = -1),
               // the other end of the node's leaf range (Karras: one end is
               // the node index itself)
};
static_assert(sizeof(BvhNode) == 64, "BVH node must be one 64-byte segment");

// Four-wide quantized node, what the fast traversal orders walk: the binary
// node at an even level with its two to four children (its grandchildren,
// or its children where those are leaves), in left-to-right leaf order.
// Child boxes are 8-bit offsets from the node's origin in units of a power
// of two per axis, rounded outward when built: origin + q * 2^e is one
// monotone rounding of an exact product, so a decoded box never shrinks.
// Half the bytes of two binary levels, in one 64-byte segment.
struct alignas(16) WideNode {
    float4 o;     // origin xyz; w holds the bits of ex | ey << 8 | ez << 16 |
                  // count << 24 (exponents biased by 127 like a float's)
    uint4 q0;     // byte k of each word is child k's: lo.x, lo.y, lo.z, hi.x
    uint4 q1;     // hi.y, hi.z, unused, unused
    int4 child;   // >= 0 wide node, < 0 leaf ~r; slots >= count unused
};
static_assert(sizeof(WideNode) == 64, "wide node must be one 64-byte segment");

struct Box {
    vec3 mn, mx;
};

// hit_aabb (src/path_tracer.cu:1467-1488) with 1/d hoisted out of the loop
// (the same IEEE division per ray). Evaluating all three slabs before the
// single test gives the same answer as the per-axis early exit: tnear only
// grows, tfar only shrinks, and fmaxf/fminf drop NaN operands.
MCPT_DEV bool slab(float mnx, float mny, float mnz, float mxx, float mxy, float mxz, const vec3& o, const vec3& inv,
                   float tfar) {
    float tnear = 0.0001f;
    float t0 = (mnx - o.x) * inv.x, t1 = (mxx - o.x) * inv.x;
    if (inv.x < 0.0f) {
        float s = t0;
        t0 = t1;
        t1 = s;
    }
    tnear = fmaxf(tnear, t0);
    tfar = fminf(tfar, t1);
    t0 = (mny - o.y) * inv.y;
    t1 = (mxy - o.y) * inv.y;
    if (inv.y < 0.0f) {
        float s = t0;
        t0 = t1;
        t1 = s;
    }
    tnear = fmaxf(tnear, t0);
    tfar = fminf(tfar, t1);
    t0 = (mnz - o.z) * inv.z;
    t1 = (mxz - o.z) * inv.z;
    if (inv.z < 0.0f) {
        float s = t0;
        t0 = t1;
        t1 = s;
    }
    tnear = fmaxf(tnear, t0);
    tfar = fminf(tfar, t1);
    return !(tfar <= tnear);
}

// unit normal exactly as hit_triangle forms it (src/path_tracer.cu:716-718)
MCPT_HDI vec3 triangle_normal(const vec3& v0, const vec3& v1, const vec3& v2) {
    vec3 e1 = v1 - v0;
    vec3 e2 = v2 - v0;
    return e1.cross(e2).normalize();
}

// hit_triangle (src/path_tracer.cu:715-741) with the normal, a pure function
// of the vertices, precomputed once per leaf by the BVH build.
MCPT_DEV bool hit_triangle_n(const vec3& v0, const vec3& v1, const vec3& v2, const vec3& n, const vec3& o,
                             const vec3& d, float& tout) {
    float den = d.dot(n);
    if (fabsf(den) < 1e-6f) return false;
    float t = (v0 - o).dot(n) / den;
    if (t < 0.0001f) return false;
    vec3 p = o + d * t;
    if (n.dot((v1 - v0).cross(p - v0)) < 0) return false;
    if (n.dot((v2 - v1).cross(p - v1)) < 0) return false;
    if (n.dot((v0 - v2).cross(p - v2)) < 0) return false;
    tout = t;
    return true;
}

// ---------------------------------------------------------------- spheres
MCPT_HDI uint32_t f2u(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    return u;
}
MCPT_HDI float u2f(uint32_t u) {
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

// A leaf record (below) holds a triangle or a sphere. A sphere's record is
// (centre, radius), (centre at time 1, 1 if moving else 0), (0, 0, 0, tag)
// with the tag a signalling-NaN bit pattern in the slot of a triangle's
// normal z. Arithmetic never yields a signalling NaN (the normal of a
// degenerate triangle is a quiet NaN), so the slot tells the two apart.
constexpr uint32_t kSphereTag = 0x7f80abcdu;

// sphere::center_at (src/path_tracer.cu:870-872)
MCPT_HDI vec3 sphere_center(const vec3& c0, const vec3& c1, bool moving, float time) {
    return moving ? c0 + (c1 - c0) * time : c0;
}

// hit_sphere (src/path_tracer.cu:886-911): the chord around the ray's
// closest approach to the centre; the near root, or the far one when the
// near one is behind (the ray starts inside)
MCPT_HDI bool hit_sphere(const vec3& c, float r, const vec3& o, const vec3& d, float& tout) {
    const vec3 oc = c - o;
    const float tc = oc.dot(d) / d.dot(d);
    const vec3 cp = o + d * tc;
    const float d2 = (c - cp).length_squared();
    const float disc = r * r - d2;
    if (disc < 0.0f) return false;
    const float h = __builtin_sqrtf(disc);
    float t = tc - h;
    if (t < 0.0001f) t = tc + h;
    if (t < 0.0001f) return false;
    tout = t;
    return true;
}

// computeSphereUvFromOutwardNormal (src/path_tracer.cu:877-884)
MCPT_HDI void sphere_uv(const vec3& n, float& u, float& v) {
    const float cy = fminf(1.0f, fmaxf(-1.0f, -n.y));
    const float theta = mf_acos(cy);
    const float phi = mf_atan2(-n.z, n.x) + MCPT_PI_F;
    u = phi / (2.0f * MCPT_PI_F);
    v = theta / MCPT_PI_F;
}

// sphere::calculate_bounding_box (:862-869): centre +- radius, the union
// over both ends of the motion for a moving sphere (no padding)
MCPT_HDI void sphere_box(const vec3& c0, const vec3& c1, bool moving, float r, float b[6]) {
    b[0] = c0.x - r;
    b[1] = c0.y - r;
    b[2] = c0.z - r;
    b[3] = c0.x + r;
    b[4] = c0.y + r;
    b[5] = c0.z + r;
    if (moving) {
        b[0] = fminf(b[0], c1.x - r);
        b[1] = fminf(b[1], c1.y - r);
        b[2] = fminf(b[2], c1.z - r);
        b[3] = fmaxf(b[3], c1.x + r);
        b[4] = fmaxf(b[4], c1.y + r);
        b[5] = fmaxf(b[5], c1.z + r);
    }
}

// Rays of a queue shared by several scenes (server, batched render) carry
// their scene in the ray record's time slot. With tb = 24 the slot also
// holds the time, as scene << 24 | k: a camera time is k * 2^-24 with
// k < 2^24 (the uniform draw, mcpt_pcg32.h), so it is kept exactly; this
// form serves up to 256 scenes. With tb = 0 the slot is the scene alone and
// the time 0 (scenes without moving spheres do not read it).
MCPT_DEV float pack_scene_time(int scene, float time, int tb) {
    return __int_as_float(tb ? (int)(((uint32_t)scene << 24) | (uint32_t)(time * 16777216.0f)) : scene);
}
MCPT_DEV int unpack_scene(float w, int tb) {
    const uint32_t b = (uint32_t)__float_as_int(w);
    return tb ? (int)(b >> 24) : (int)b;
}
MCPT_DEV float unpack_time(float w, int tb) {
    return tb ? (float)((uint32_t)__float_as_int(w) & 0xFFFFFFu) * (1.0f / 16777216.0f) : 0.0f;
}

// ------------------------------------------------------------ scene view
struct EnvView {
    const float4* texels = nullptr;
    const float* row_cdf = nullptr;
    const float* col_cdf = nullptr;
    int w = 0, h = 0;
    float intensity = 1.f, rot = 0.f, total = 0.f;
    int enabled = 0;
};

// Everything a kernel needs to trace and shade, passed by value (no
// __constant__ globals, so several scenes can be in flight at once).
// Leaf record in Morton (leaf) order: the triangle's 36-byte hot vertices
// followed by the unit normal hit_triangle would compute, 48 bytes, read as
// three 16-byte loads.
struct alignas(16) LeafTri {
    float4 a;  // v0.xyz, v1.x
    float4 b;  // v1.yz, v2.xy
    float4 c;  // v2.z, n.xyz
};
static_assert(sizeof(LeafTri) == 48, "leaf record is 48 bytes");

MCPT_HDI bool leaf_is_sphere(const float4& c) { return f2u(c.w) == kSphereTag; }

MCPT_HDI LeafTri sphere_leaf(const mcpt_sphere& sp) {
    LeafTri L;
    L.a = make_float4(sp.center[0], sp.center[1], sp.center[2], sp.radius);
    L.b = make_float4(sp.center1[0], sp.center1[1], sp.center1[2], sp.moving ? 1.0f : 0.0f);
    L.c = make_float4(0.0f, 0.0f, 0.0f, u2f(kSphereTag));
    return L;
}

// the primitive's box from its leaf record: padded triangle box
// (triangle::calculate_bounding_box, :651-655) or sphere box
MCPT_HDI void leaf_box(const LeafTri& L, float b[6]) {
    if (leaf_is_sphere(L.c)) {
        sphere_box(vec3(L.a.x, L.a.y, L.a.z), vec3(L.b.x, L.b.y, L.b.z), L.b.w != 0.0f, L.a.w, b);
        return;
    }
    const float eps = 1e-3f;
    b[0] = fminf(L.a.x, fminf(L.a.w, L.b.z)) - eps;
    b[1] = fminf(L.a.y, fminf(L.b.x, L.b.w)) - eps;
    b[2] = fminf(L.a.z, fminf(L.b.y, L.c.x)) - eps;
    b[3] = fmaxf(L.a.x, fmaxf(L.a.w, L.b.z)) + eps;
    b[4] = fmaxf(L.a.y, fmaxf(L.b.x, L.b.w)) + eps;
    b[5] = fmaxf(L.a.z, fmaxf(L.b.y, L.c.x)) + eps;
}

// What traversal needs of a scene (40 bytes, loaded per ray when the rays of
// several scenes share one queue).
struct BvhView {
    const BvhNode* nodes;     // binary tree (the reference's visit order)
    const WideNode* wide;     // four-wide quantized tree (fast orders)
    const LeafTri* leaf_tri;  // leaf (Morton) order
    const int* leaf_prim;     // leaf -> original triangle index, ~sphere index
    const int* tie;           // leaf -> rank in the reference's leaf-test order
};

// A texture a material can take its albedo from (sampleTexture,
// src/path_tracer.cu:216-234; the solid albedo is the material's own).
struct TexView {
    int type;            // MCPT_TEX_CHECKER / MCPT_TEX_IMAGE
    int w, h;            // image size
    float scale;         // checker cells per unit of u and v
    float ca[3], cb[3];  // checker colours, linear
    const uint32_t* px;  // image RGBA8, top row first
};

// Heterogeneous medium: a dense density grid, x fastest, whose voxel
// (i, j, k) sits at world position origin + (i, j, k) * voxel size; the
// reference samples a NanoVDB float grid the same way (sampleVolumeDensity,
// src/path_tracer.cu:784-789). data == nullptr: every medium is homogeneous.
struct GridView {
    const float* data;
    int nx, ny, nz;
    float ox, oy, oz;  // world position of voxel (0, 0, 0)
    float ix, iy, iz;  // 1 / voxel size
    float max_density;
};

constexpr uint32_t TRI_VERTEX_NORMALS = 1u << 31;
constexpr uint32_t TRI_MATERIAL_MASK = (1u << 30) - 1;

struct SceneView {
    BvhView bvh;
    const float* hot;           // 9 floats per triangle, original order
    const mcpt_tri_cold* cold;  // original order
    // per triangle, original order: the material index (bits 0-29) and
    // TRI_VERTEX_NORMALS when the record has vertex normals (n0 != 0); 4 bytes
    // where the 64-byte cold record would be gathered for one field
    const uint32_t* tri_info;
    const mcpt_material* mats;
    const mcpt_light* lights;
    int n_lights;
    int n_tris;
    int has_medium;
    int stack_depth;  // deepest leaf in internal levels (<= 64)
    EnvView env;
    const mcpt_sphere* spheres;  // original order
    int n_spheres;
    const TexView* tex;   // the scene's textures
    const float* srgb8;   // sRGB byte -> linear, 256 entries
    GridView grid;        // density grid of every isotropic material
};

// Trilinear density at world point p (NanoVDB's TrilinearSampler over
// worldToIndexF): index coordinates q = (p - origin) / voxel, the eight
// voxels around floor(q), voxels outside the grid read 0 (the background of
// a fog volume), interpolated along z, then y, then x as a + w (b - a).
MCPT_DEV float grid_voxel(const GridView& g, int i, int j, int k) {
    if (i < 0 || j < 0 || k < 0 || i >= g.nx || j >= g.ny || k >= g.nz) return 0.0f;
    return g.data[((size_t)k * g.ny + j) * g.nx + i];
}
MCPT_DEV float grid_density(const GridView& g, const vec3& p) {
    const float x = (p.x - g.ox) * g.ix, y = (p.y - g.oy) * g.iy, z = (p.z - g.oz) * g.iz;
    // all eight voxels outside (or a non-finite point): background
    if (!(x > -1.0f && y > -1.0f && z > -1.0f && x < (float)g.nx && y < (float)g.ny && z < (float)g.nz)) return 0.0f;
    const float x0 = floorf(x), y0 = floorf(y), z0 = floorf(z);
    const float u = x - x0, v = y - y0, w = z - z0;
    const int i = (int)x0, j = (int)y0, k = (int)z0;
    float a[2][2];
    for (int di = 0; di < 2; di++)
        for (int dj = 0; dj < 2; dj++) {
            const float c0 = grid_voxel(g, i + di, j + dj, k), c1 = grid_voxel(g, i + di, j + dj, k + 1);
            a[di][dj] = c0 + w * (c1 - c0);
        }
    const float b0 = a[0][0] + v * (a[0][1] - a[0][0]);
    const float b1 = a[1][0] + v * (a[1][1] - a[1][0]);
    return b0 + u * (b1 - b0);
}

// sampleTexture (src/path_tracer.cu:216-234). The image filter restates the
// texture unit's normalised-coordinate bilinear read with wrap addressing in
// FP32: x = u w - 1/2, y = (1 - v) h - 1/2, the four texels around (x, y)
// decoded from sRGB and blended by the fractional parts.
MCPT_DEV vec3 sample_albedo(const SceneView& sv, const mcpt_material& m, float u, float v) {
    if (m.texture <= 0) return load3(m.albedo);
    const TexView& t = sv.tex[m.texture - 1];
    if (t.type == MCPT_TEX_CHECKER) {
        const int cu = (int)floorf(u * t.scale), cv = (int)floorf(v * t.scale);
        return ((cu + cv) & 1) == 0 ? vec3(t.ca[0], t.ca[1], t.ca[2]) : vec3(t.cb[0], t.cb[1], t.cb[2]);
    }
    // far-off or non-finite coordinates are brought into range first
    const float x = fminf(fmaxf(u * (float)t.w - 0.5f, -1e9f), 1e9f);
    const float y = fminf(fmaxf((1.0f - v) * (float)t.h - 0.5f, -1e9f), 1e9f);
    const float x0 = floorf(x), y0 = floorf(y);
    const float fx = x - x0, fy = y - y0;
    const int ix0 = (((int)((long long)x0 % t.w)) + t.w) % t.w, iy0 = (((int)((long long)y0 % t.h)) + t.h) % t.h;
    const int ix1 = (ix0 + 1) % t.w, iy1 = (iy0 + 1) % t.h;
    const float* lut = sv.srgb8;
    const uint32_t p00 = t.px[(size_t)iy0 * t.w + ix0], p10 = t.px[(size_t)iy0 * t.w + ix1];
    const uint32_t p01 = t.px[(size_t)iy1 * t.w + ix0], p11 = t.px[(size_t)iy1 * t.w + ix1];
    const float gx = 1.0f - fx, gy = 1.0f - fy;
    float c[3];
    for (int k = 0; k < 3; k++) {
        const int sh = 8 * k;
        const float a00 = lut[(p00 >> sh) & 255u], a10 = lut[(p10 >> sh) & 255u];
        const float a01 = lut[(p01 >> sh) & 255u], a11 = lut[(p11 >> sh) & 255u];
        c[k] = (a00 * gx + a10 * fx) * gy + (a01 * gx + a11 * fx) * fy;
    }
    return vec3(c[0], c[1], c[2]);
}

// ------------------------------------------------------- queue records
struct alignas(16) RayItem {  // 32 bytes: two 16-byte loads per ray
    float4 ot;                // origin xyz, time
    float4 dp;                // direction xyz, path index (int bits)
};
struct alignas(16) ShadowItem {
    float4 om;  // origin xyz, max distance
    float4 dp;  // direction xyz, path index (int bits)
    float4 c;   // contribution rgb, the path's time
};

// ------------------------------------------------------ wave64 utilities
MCPT_DEV int lane_id() { return __lane_id(); }
MCPT_DEV uint64_t ballot64(bool p) { return __ballot(p); }
// number of set bits of m below the calling lane
MCPT_DEV int rank_below(uint64_t m) {
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
// One atomicAdd per wave for all lanes with `take`; returns each taking
// lane's slot. Must be called by every active lane of the wave.
MCPT_DEV int wave_push(int* counter, bool take) {
    uint64_t m = ballot64(take);
    int base = 0;
    if (m) {
        int leader = __builtin_ctzll(m);
        if (lane_id() == leader) base = atomicAdd(counter, __popcll(m));
        base = __shfl(base, leader);
    }
    return base + rank_below(m);
}
// Two pushes with one atomic: the counter holds two 32-bit counts (a in the
// low half, b in the high half), so lanes taking a slot in either list cost
// the wave a single round trip.
MCPT_DEV void wave_push2(unsigned long long* counter, bool take_a, bool take_b, int& slot_a, int& slot_b) {
    const uint64_t ma = ballot64(take_a), mb = ballot64(take_b);
    unsigned long long base = 0;
    if (ma | mb) {
        const int leader = __builtin_ctzll(ma | mb);
        if (lane_id() == leader)
            base = atomicAdd(counter, (unsigned long long)__popcll(ma) + ((unsigned long long)__popcll(mb) << 32));
        base = __shfl(base, leader);
    }
    slot_a = (int)(uint32_t)base + rank_below(ma);
    slot_b = (int)(uint32_t)(base >> 32) + rank_below(mb);
}

}  // namespace mcpt
