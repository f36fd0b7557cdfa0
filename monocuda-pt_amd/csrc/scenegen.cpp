// scenegen.cpp — procedural scenes for the workloads, built on the host with
// the reference's scene-construction helpers restated: add_quad / add_box
// (src/path_tracer.cu:1677-1708), buildCornellRoom (:1747-1763), the
// sRGB-to-linear conversion of solid colours (:200-209) and the sky formula
// of scripts/make_sky.py:5-27. Generation uses FP64 libm and PCG32 streams
// seeded by the caller, so every host produces the same arrays.
#include <cmath>
#include <cstring>
#include <vector>

#include "host_common.hpp"
#include "mcpt_fmath.h"
#include "mcpt_pcg32.h"

namespace mcpt {

namespace {

struct HostScene {  // hostScene (src/path_tracer.cu:1667-1675), triangles only
    std::vector<mcpt_tri_hot> hot;
    std::vector<mcpt_tri_cold> cold;
    std::vector<mcpt_material> mats;
};


int add_material(HostScene& s, int type, const float albedo_srgb[3], const float emission[3], float fuzz, float ior,
                 float density, float g) {
    mcpt_material m;
    std::memset(&m, 0, sizeof(m));
    m.type = type;
    for (int k = 0; k < 3; k++) {
        m.albedo[k] = srgb_to_linear(albedo_srgb[k]);
        m.emission[k] = emission[k];
    }
    m.fuzz = fuzz < 1 ? fuzz : 1.f;  // make_conductor clamps (:589-591)
    m.ior = ior;
    m.density = density;
    m.phase_g = g;
    s.mats.push_back(m);
    return (int)s.mats.size() - 1;
}
int lambertian(HostScene& s, float r, float g, float b) {
    const float a[3] = {r, g, b}, e[3] = {0, 0, 0};
    return add_material(s, MCPT_MAT_LAMBERTIAN, a, e, 0.f, 0.f, 0.f, 0.f);
}
int conductor(HostScene& s, float r, float g, float b, float fuzz) {
    const float a[3] = {r, g, b}, e[3] = {0, 0, 0};
    return add_material(s, MCPT_MAT_CONDUCTOR, a, e, fuzz, 0.f, 0.f, 0.f);
}
int dielectric(HostScene& s, float ior) {
    const float a[3] = {1, 1, 1}, e[3] = {0, 0, 0};
    return add_material(s, MCPT_MAT_DIELECTRIC, a, e, 0.f, ior, 0.f, 0.f);
}
int emissive(HostScene& s, float r, float g, float b) {
    const float a[3] = {0, 0, 0}, e[3] = {r, g, b};
    return add_material(s, MCPT_MAT_EMISSIVE, a, e, 0.f, 0.f, 0.f, 0.f);
}
int isotropic(HostScene& s, float r, float g, float b, float density, float phase_g) {
    const float a[3] = {r, g, b}, e[3] = {0, 0, 0};
    return add_material(s, MCPT_MAT_ISOTROPIC, a, e, 0.f, 0.f, density, phase_g);
}

struct P3 {
    float x, y, z;
};

void push_tri(HostScene& s, P3 a, P3 b, P3 c, int mat, const float uv[6]) {
    mcpt_tri_hot h = {{a.x, a.y, a.z}, {b.x, b.y, b.z}, {c.x, c.y, c.z}};
    mcpt_tri_cold k;
    std::memset(&k, 0, sizeof(k));
    if (uv) std::memcpy(k.uv0, uv, 6 * sizeof(float));
    k.material = mat;
    s.hot.push_back(h);
    s.cold.push_back(k);
}

// add_quad (src/path_tracer.cu:1677-1691)
void add_quad(HostScene& s, P3 a, P3 b, P3 c, P3 d, int m) {
    const float uv1[6] = {0.f, 0.f, 1.f, 0.f, 1.f, 1.f};
    const float uv2[6] = {0.f, 0.f, 1.f, 1.f, 0.f, 1.f};
    push_tri(s, a, b, c, m, uv1);
    push_tri(s, a, c, d, m, uv2);
}

// add_box (src/path_tracer.cu:1693-1708)
void add_box(HostScene& s, P3 center, P3 half, float rot_y, int m) {
    float cs = mf_cos(rot_y), sn = mf_sin(rot_y);
    P3 p[8];
    for (int i = 0; i < 8; i++) {
        float x = (i & 1) ? half.x : -half.x;
        float y = (i & 2) ? half.y : -half.y;
        float z = (i & 4) ? half.z : -half.z;
        p[i] = P3{center.x + (x * cs + z * sn), center.y + y, center.z + (-x * sn + z * cs)};
    }
    add_quad(s, p[0], p[1], p[3], p[2], m);
    add_quad(s, p[4], p[5], p[7], p[6], m);
    add_quad(s, p[0], p[2], p[6], p[4], m);
    add_quad(s, p[1], p[3], p[7], p[5], m);
    add_quad(s, p[0], p[1], p[5], p[4], m);
    add_quad(s, p[2], p[3], p[7], p[6], m);
}

// Cornell room with a grey floor (the reference uses a checker floor,
// buildCornellRoom :1747-1763), a tall box as in buildCornellScene
// (:1772-1773) and a short box in place of the mirror sphere: 36 triangles.
void build_cornell(HostScene& s) {
    int white = lambertian(s, 0.73f, 0.73f, 0.73f);
    int red = lambertian(s, 0.65f, 0.05f, 0.05f);
    int green = lambertian(s, 0.12f, 0.45f, 0.15f);
    int light = emissive(s, 15.0f, 15.0f, 15.0f);
    const float L = -2.f, R = 2.f, B = 0.f, T = 4.f, F = -4.f, N = 0.f;
    add_quad(s, {L, B, N}, {R, B, N}, {R, B, F}, {L, B, F}, white);  // floor
    add_quad(s, {L, T, N}, {R, T, N}, {R, T, F}, {L, T, F}, white);  // ceiling
    add_quad(s, {L, B, F}, {R, B, F}, {R, T, F}, {L, T, F}, white);  // back
    add_quad(s, {L, B, N}, {L, B, F}, {L, T, F}, {L, T, N}, red);    // left
    add_quad(s, {R, B, N}, {R, B, F}, {R, T, F}, {R, T, N}, green);  // right
    add_quad(s, {-1.f, T - 0.01f, -2.6f}, {1.f, T - 0.01f, -2.6f}, {1.f, T - 0.01f, -1.4f}, {-1.f, T - 0.01f, -1.4f},
             light);
    add_box(s, {-0.8f, 1.1f, -2.8f}, {0.55f, 1.1f, 0.55f}, 0.3f, white);
    add_box(s, {0.85f, 0.6f, -1.7f}, {0.6f, 0.6f, 0.6f}, -0.3f, white);
}

struct Gen {
    mcpt_pcg32 g;
    explicit Gen(uint64_t seed, uint64_t stream) { mcpt_pcg32_seed(&g, seed, stream); }
    double u() {  // 53-bit uniform in [0, 1)
        uint64_t hi = mcpt_pcg32_next(&g), lo = mcpt_pcg32_next(&g);
        return (double)(((hi << 32) | lo) >> 11) * (1.0 / 9007199254740992.0);
    }
    double normal() {  // Box-Muller
        double u1 = u(), u2 = u();
        return std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(2.0 * M_PI * u2);
    }
    void unit(double v[3]) {
        double n2;
        do {
            v[0] = normal();
            v[1] = normal();
            v[2] = normal();
            n2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        } while (n2 < 1e-12);
        double n = std::sqrt(n2);
        v[0] /= n;
        v[1] /= n;
        v[2] /= n;
    }
};

// one triangle: vertex k = centre + unit_k * exp(mu + sigma N(0,1))
void random_triangle(HostScene& s, Gen& g, const double c[3], double mu, double sigma, int mat) {
    P3 v[3];
    for (int k = 0; k < 3; k++) {
        double d[3];
        g.unit(d);
        double len = std::exp(mu + sigma * g.normal());
        v[k] = P3{(float)(c[0] + d[0] * len), (float)(c[1] + d[1] * len), (float)(c[2] + d[2] * len)};
    }
    push_tri(s, v[0], v[1], v[2], mat, nullptr);
}

// Triangle soup of the incoherent-mesh workload: n triangles with centres
// uniform in the unit cube, vertex offsets lognormal(ln 0.004, 0.5) along
// random unit directions; materials uniform over diffuse / conductor (fuzz
// 0.1) / dielectric (IOR 1.5) with a palette of colours; then 16 emissive
// triangles (radiance 15, offsets lognormal(ln 0.05, 0.25)).
void build_soup(HostScene& s, uint64_t seed, int64_t n) {
    const float palette[8][3] = {{0.73f, 0.73f, 0.73f}, {0.65f, 0.05f, 0.05f}, {0.12f, 0.45f, 0.15f},
                                 {0.2f, 0.3f, 0.8f},    {0.9f, 0.75f, 0.2f},  {0.5f, 0.2f, 0.6f},
                                 {0.1f, 0.6f, 0.6f},    {0.85f, 0.45f, 0.3f}};
    int diffuse[8], metal[8];
    for (int i = 0; i < 8; i++) diffuse[i] = lambertian(s, palette[i][0], palette[i][1], palette[i][2]);
    for (int i = 0; i < 8; i++)
        metal[i] = conductor(s, 0.5f + 0.5f * palette[i][0], 0.5f + 0.5f * palette[i][1], 0.5f + 0.5f * palette[i][2],
                             0.1f);
    int glass = dielectric(s, 1.5f);
    int light = emissive(s, 15.0f, 15.0f, 15.0f);
    s.hot.reserve((size_t)n + 16);
    s.cold.reserve((size_t)n + 16);
    Gen g(seed, 0x50u);
    const double mu = std::log(0.004), sigma = 0.5;
    for (int64_t i = 0; i < n; i++) {
        double c[3] = {g.u(), g.u(), g.u()};
        int kind = (int)(g.u() * 3.0);
        int colour = (int)(g.u() * 8.0);
        int mat = kind == 0 ? diffuse[colour] : kind == 1 ? metal[colour] : glass;
        random_triangle(s, g, c, mu, sigma, mat);
    }
    for (int i = 0; i < 16; i++) {
        double c[3] = {0.1 + 0.8 * g.u(), 0.1 + 0.8 * g.u(), 0.1 + 0.8 * g.u()};
        random_triangle(s, g, c, std::log(0.05), 0.25, light);
    }
}

// Displaced icosphere (radius 0.35 about the cube centre, radial noise of
// amplitude 0.05 from a seeded sum of sines) plus a random-triangle fill up
// to `n` triangles, lit by 4 emissive triangles.
void build_icosphere(HostScene& s, uint64_t seed, int64_t n) {
    int mats[3] = {lambertian(s, 0.73f, 0.73f, 0.73f), conductor(s, 0.9f, 0.8f, 0.6f, 0.1f), dielectric(s, 1.5f)};
    int light = emissive(s, 15.0f, 15.0f, 15.0f);
    Gen g(seed, 0x1c05u);
    // icosahedron
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    std::vector<double> V = {-1, t, 0, 1, t, 0, -1, -t, 0, 1, -t, 0, 0, -1, t, 0, 1, t,
                             0, -1, -t, 0, 1, -t, t, 0, -1, t, 0, 1, -t, 0, -1, -t, 0, 1};
    std::vector<int> F = {0, 11, 5, 0, 5,  1,  0, 1, 7, 0, 7,  10, 0, 10, 11, 1, 5, 9, 5, 11,
                          4, 11, 10, 2, 10, 7, 6, 7, 1, 8, 3,  9,  4, 3,  4,  2, 3, 2, 6, 3,
                          6, 8,  3, 8, 9,  4, 9, 5, 2, 4, 11, 6,  2, 10, 8,  6, 7, 9, 8, 1};
    int64_t target_sphere = n / 2;
    while ((int64_t)F.size() / 3 * 4 <= target_sphere) {
        std::vector<int> nf;
        std::vector<std::pair<long long, int>> cache;
        auto mid = [&](int a, int b) {
            long long key = a < b ? ((long long)a << 32) | b : ((long long)b << 32) | a;
            for (auto& kv : cache)
                if (kv.first == key) return kv.second;
            int id = (int)V.size() / 3;
            V.push_back((V[3 * a] + V[3 * b]) / 2);
            V.push_back((V[3 * a + 1] + V[3 * b + 1]) / 2);
            V.push_back((V[3 * a + 2] + V[3 * b + 2]) / 2);
            cache.push_back({key, id});
            if (cache.size() > 64) cache.erase(cache.begin());
            return id;
        };
        for (size_t f = 0; f < F.size(); f += 3) {
            int a = F[f], b = F[f + 1], c = F[f + 2];
            int ab = mid(a, b), bc = mid(b, c), ca = mid(c, a);
            int add[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
            nf.insert(nf.end(), add, add + 12);
        }
        F.swap(nf);
    }
    double freq[6][3], phase[6];
    for (int k = 0; k < 6; k++) {
        for (int a = 0; a < 3; a++) freq[k][a] = (g.u() * 2 - 1) * (2.0 + 6.0 * k);
        phase[k] = g.u() * 2 * M_PI;
    }
    std::vector<P3> P(V.size() / 3);
    for (size_t i = 0; i < P.size(); i++) {
        double x = V[3 * i], y = V[3 * i + 1], z = V[3 * i + 2];
        double l = std::sqrt(x * x + y * y + z * z);
        x /= l;
        y /= l;
        z /= l;
        double noise = 0;
        for (int k = 0; k < 6; k++)
            noise += std::sin(freq[k][0] * x + freq[k][1] * y + freq[k][2] * z + phase[k]) / (1 << k);
        double r = 0.35 * (1.0 + 0.05 * noise / 2.0);
        P[i] = P3{(float)(0.5 + r * x), (float)(0.5 + r * y), (float)(0.5 + r * z)};
    }
    int sphere_mat = mats[(int)(g.u() * 3.0)];
    for (size_t f = 0; f < F.size(); f += 3) push_tri(s, P[F[f]], P[F[f + 1]], P[F[f + 2]], sphere_mat, nullptr);
    int64_t fill = n - (int64_t)F.size() / 3 - 4;
    const double mu = std::log(0.004), sigma = 0.5;
    for (int64_t i = 0; i < fill; i++) {
        double c[3] = {g.u(), g.u(), g.u()};
        random_triangle(s, g, c, mu, sigma, mats[(int)(g.u() * 3.0)]);
    }
    for (int i = 0; i < 4; i++) {
        double c[3] = {0.1 + 0.8 * g.u(), 0.9, 0.1 + 0.8 * g.u()};
        random_triangle(s, g, c, std::log(0.06), 0.2, light);
    }
}

// soup + an axis-aligned box (12 triangles) of homogeneous isotropic medium,
// sigma_t 0.5, albedo 0.8, Henyey-Greenstein g 0.3
void build_medium(HostScene& s, uint64_t seed, int64_t n) {
    build_soup(s, seed, n);
    int fog = isotropic(s, 0.8f, 0.8f, 0.8f, 0.5f, 0.3f);
    add_box(s, {0.5f, 0.5f, 0.5f}, {0.3f, 0.3f, 0.3f}, 0.0f, fog);
}

}  // namespace

float srgb_to_linear(float e) {  // src/path_tracer.cu:200-202
    return e <= 0.04045f ? e / 12.92f : mf_pow((e + 0.055f) / 1.055f, 2.4f);
}

int gen_scene(int kind, uint64_t seed, int64_t param, mcpt_tri_hot* hot, mcpt_tri_cold* cold,
              mcpt_material* materials, mcpt_gen_sizes* sizes) {
    HostScene s;
    switch (kind) {
        case MCPT_GEN_CORNELL: build_cornell(s); break;
        case MCPT_GEN_SOUP: build_soup(s, seed, param > 0 ? param : 1048576); break;
        case MCPT_GEN_ICOSPHERE: build_icosphere(s, seed, param > 0 ? param : 65536); break;
        case MCPT_GEN_MEDIUM: build_medium(s, seed, param > 0 ? param : 2097152); break;
        default: return set_error(MCPT_ERR_INVALID, "unknown scene kind %d", kind);
    }
    if (sizes) {
        sizes->n_tris = (int32_t)s.hot.size();
        sizes->n_materials = (int32_t)s.mats.size();
    }
    if (hot) std::memcpy(hot, s.hot.data(), s.hot.size() * sizeof(mcpt_tri_hot));
    if (cold) std::memcpy(cold, s.cold.data(), s.cold.size() * sizeof(mcpt_tri_cold));
    if (materials) std::memcpy(materials, s.mats.data(), s.mats.size() * sizeof(mcpt_material));
    return MCPT_OK;
}

// scripts/make_sky.py:5-27 evaluated per texel in FP64, top row first.
int gen_sky(int w, int h, float* rgba) {
    if (w < 1 || h < 1 || !rgba) return set_error(MCPT_ERR_INVALID, "bad sky size");
    const double elev = 50.0 * M_PI / 180.0, azim = 30.0 * M_PI / 180.0;
    const double sun[3] = {std::cos(elev) * std::cos(azim), std::sin(elev), std::cos(elev) * std::sin(azim)};
    const double zen[3] = {0.20, 0.45, 1.00}, hor[3] = {0.90, 0.95, 1.00}, gnd[3] = {0.25, 0.22, 0.18};
    const double sun_rgb[3] = {600.0, 560.0, 480.0}, glow_rgb[3] = {1.2, 1.0, 0.7};
    for (int j = 0; j < h; j++) {
        double v = (j + 0.5) / h;
        double theta = v * M_PI;
        for (int i = 0; i < w; i++) {
            double u = (i + 0.5) / w;
            double phi = (u - 0.5) * 2 * M_PI;
            double d[3] = {std::cos(phi) * std::sin(theta), std::cos(theta), std::sin(phi) * std::sin(theta)};
            double up = std::fmin(std::fmax(d[1], -1.0), 1.0);
            double t = std::sqrt(std::fmin(std::fmax(up, 0.0), 1.0));
            double ca = std::fmin(std::fmax(d[0] * sun[0] + d[1] * sun[1] + d[2] * sun[2], -1.0), 1.0);
            double ang = std::acos(ca);
            double glow = std::exp(-std::pow(ang / (12.0 * M_PI / 180.0), 2));
            float* o = &rgba[((size_t)j * w + i) * 4];
            for (int k = 0; k < 3; k++) {
                double c = (up >= 0 ? hor[k] * (1 - t) + zen[k] * t : gnd[k]) * 0.8;
                if (ang < 2.0 * M_PI / 180.0) c += sun_rgb[k];
                c += glow * glow_rgb[k];
                o[k] = (float)c;
            }
            o[3] = 1.0f;
        }
    }
    return MCPT_OK;
}

int gen_camera(int kind, int w, int h, mcpt_camera_desc* d) {
    if (!d || w < 1 || h < 1) return set_error(MCPT_ERR_INVALID, "bad camera request");
    std::memset(d, 0, sizeof(*d));
    d->up[1] = 1.f;
    d->width = w;
    d->height = h;
    d->rr_cap = 0.95f;
    d->aperture = 0.f;
    if (kind == MCPT_GEN_CORNELL) {  // main's defaults (src/path_tracer.cu:4114-4124)
        const float from[3] = {0, 2, 5.5f}, at[3] = {0, 2, -2};
        std::memcpy(d->look_from, from, sizeof(from));
        std::memcpy(d->look_at, at, sizeof(at));
        d->vfov_degrees = 40.f;
        d->focus_distance = 7.5f;
    } else {
        const float from[3] = {0.5f, 0.5f, 2.6f}, at[3] = {0.5f, 0.5f, 0.5f};
        std::memcpy(d->look_from, from, sizeof(from));
        std::memcpy(d->look_at, at, sizeof(at));
        d->vfov_degrees = 38.f;
        d->focus_distance = 2.1f;
    }
    return MCPT_OK;
}

}  // namespace mcpt
