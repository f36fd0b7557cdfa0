// capi.cpp — the extern "C" boundary declared in include/mcpt.h: contexts,
// scenes, films and argument checking around the kernel launchers.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>
#include <exception>
#include <string>

#include "scene.hpp"

namespace mcpt {

namespace {
thread_local std::string g_last_error;
}

int set_error(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}
void clear_error() { g_last_error.clear(); }

int dev_alloc(Ctx* ctx, void** p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) return MCPT_OK;
    std::lock_guard<std::mutex> lock(ctx->alloc_mu);
    if (ctx->used + bytes > ctx->budget)
        return set_error(MCPT_ERR_BUDGET, "allocation of %zu bytes exceeds the HBM budget (%llu of %llu used)", bytes,
                         (unsigned long long)ctx->used, (unsigned long long)ctx->budget);
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return set_error(MCPT_ERR_HIP, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    }
    ctx->used += bytes;
    return MCPT_OK;
}

// the context's streams and events but `stream`, as pointers
static std::vector<hipStream_t*> ctx_streams(Ctx* c) {
    std::vector<hipStream_t*> v = {&c->side};
    for (int l = 1; l < Ctx::kMaxLanes; l++) {
        v.push_back(&c->lane_stream[l]);
        v.push_back(&c->lane_side[l]);
    }
    return v;
}
static std::vector<hipEvent_t*> ctx_events(Ctx* c) {
    std::vector<hipEvent_t*> v = {&c->ev_join};
    for (int l = 0; l < Ctx::kMaxLanes; l++) {
        v.push_back(&c->ev_shade[l]);
        v.push_back(&c->ev_shadow[l]);
        v.push_back(&c->ev_fin[l]);
    }
    return v;
}

int ctx_sync_objects_create(Ctx* c) {
    for (hipStream_t* sp : ctx_streams(c))
        if (hipStreamCreateWithFlags(sp, hipStreamNonBlocking) != hipSuccess) return MCPT_ERR_HIP;
    for (hipEvent_t* ep : ctx_events(c))
        if (hipEventCreateWithFlags(ep, hipEventDisableTiming) != hipSuccess) return MCPT_ERR_HIP;
    return MCPT_OK;
}

void ctx_sync_objects_destroy(Ctx* c) {
    for (hipStream_t* sp : ctx_streams(c))
        if (*sp) {
            (void)hipStreamSynchronize(*sp);
            (void)hipStreamDestroy(*sp);
            *sp = nullptr;
        }
    for (hipEvent_t* ep : ctx_events(c))
        if (*ep) {
            (void)hipEventDestroy(*ep);
            *ep = nullptr;
        }
}

void dev_free(Ctx* ctx, void* p, size_t bytes) {
    if (!p) return;
    std::lock_guard<std::mutex> lock(ctx->alloc_mu);
    (void)hipFree(p);
    ctx->used -= bytes;
}

namespace {

// Camera::initialize (src/path_tracer.cu:1610-1631)
int camera_init(const mcpt_camera_desc* d, mcpt_camera* c) {
    if (d->width < 1 || d->height < 1) return set_error(MCPT_ERR_INVALID, "camera size must be positive");
    float aspect = d->width / float(d->height);
    int width = d->width;
    int height = int(width / aspect);
    if (height != d->height)
        return set_error(MCPT_ERR_INVALID, "image height %d is not reproduced by width/aspect (%d)", d->height, height);
    vec3 from = load3(d->look_from), at = load3(d->look_at), up = load3(d->up);
    vec3 w = (from - at).normalize();
    vec3 u = up.cross(w).normalize();
    vec3 v = w.cross(u);
    float lens_radius = d->aperture / 2.;
    vec3 center = from;
    float theta = d->vfov_degrees * (3.14159265358979323846f / 180.0f);
    float h = tanf(theta / 2.0f);
    float vh = 2.0f * h * d->focus_distance;
    float vw = vh * (float(width) / height);
    vec3 vu = u * vw;
    vec3 vv = v * -vh;
    vec3 du = vu / float(width);
    vec3 dv = vv / float(height);
    vec3 ul = center - (w * d->focus_distance) - vu / 2.0f - vv / 2.0f;
    vec3 p00 = ul + (du + dv) * 0.5f;
    auto put = [](float* o, const vec3& x) {
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
    return MCPT_OK;
}

// buildEnvironmentSamplingTables (src/path_tracer.cu:370-385)
void env_tables(const float* rgba, int w, int h, std::vector<float>& row_cdf, std::vector<float>& col_cdf,
                float& total) {
    row_cdf.assign(h + 1, 0.0f);
    col_cdf.assign((size_t)h * (w + 1), 0.0f);
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
    total = row_cdf[h];
}

template <class T>
int upload(Ctx* ctx, DevBuf<T>& b, const T* host, size_t n) {
    MCPT_TRY(b.alloc(ctx, n));
    if (n) MCPT_HIP(hipMemcpyAsync(b.ptr, host, n * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
    return MCPT_OK;
}

}  // namespace
}  // namespace mcpt

using namespace mcpt;

#define CHECK_ARG(cond, msg)                                          \
    do {                                                              \
        if (!(cond)) return set_error(MCPT_ERR_INVALID, "%s", msg);   \
    } while (0)

// The file readers run C++ library code that can throw (out of memory on
// a huge but consistent file); nothing may cross the C boundary.
template <class F>
static int guarded(const char* what, F f) {
    try {
        return f();
    } catch (const std::exception& e) {
        return set_error(MCPT_ERR_INVALID, "%s: %s", what, e.what());
    } catch (...) {
        return set_error(MCPT_ERR_INVALID, "%s failed", what);
    }
}

extern "C" {

const char* mcpt_last_error(void) { return g_last_error.c_str(); }
const char* mcpt_version(void) { return "mcpt 0.1 gfx950"; }

int mcpt_ctx_create(int device, uint64_t hbm_budget_bytes, mcpt_ctx* out) {
    CHECK_ARG(out, "out is NULL");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1)
        return set_error(MCPT_ERR_NODEVICE, "no HIP device available (the library runs on MI355X / gfx950)");
    CHECK_ARG(device >= 0 && device < n, "device index out of range");
    MCPT_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    MCPT_HIP(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return set_error(MCPT_ERR_NODEVICE, "device %d is %s; this library is built for gfx950 only", device,
                         prop.gcnArchName);
    auto* c = new mcpt_ctx_s();
    c->c.device = device;
    c->c.cu_count = prop.multiProcessorCount;
    c->c.clock_khz = prop.clockRate;
    std::snprintf(c->c.name, sizeof(c->c.name), "%s (%s)", prop.name, prop.gcnArchName);
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
        delete c;
        return set_error(MCPT_ERR_HIP, "hipMemGetInfo failed");
    }
    c->c.budget = hbm_budget_bytes ? hbm_budget_bytes : (uint64_t)free_b;
    if (hipStreamCreateWithFlags(&c->c.stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return set_error(MCPT_ERR_HIP, "hipStreamCreate failed");
    }
    if (ctx_sync_objects_create(&c->c) < 0) {
        ctx_sync_objects_destroy(&c->c);
        hipStreamDestroy(c->c.stream);
        delete c;
        return set_error(MCPT_ERR_HIP, "hipStreamCreate / hipEventCreate failed");
    }
    // the fixed workspace of mcpt_render_noalloc, once per device, with the
    // device current
    int st = scratch_init(device);
    if (st < 0) {
        ctx_sync_objects_destroy(&c->c);
        hipStreamDestroy(c->c.stream);
        delete c;
        return st;
    }
    *out = c;
    return MCPT_OK;
}

int mcpt_ctx_destroy(mcpt_ctx ctx) {
    if (!ctx) return MCPT_OK;
    const int live = ctx->c.children.load();
    if (live > 0)
        return set_error(MCPT_ERR_INVALID,
                         "the context still has %d live scene/film/server/network handle(s); destroy them first", live);
    hipSetDevice(ctx->c.device);
    hipStreamSynchronize(ctx->c.stream);
    if (ctx->c.ws) dev_free(&ctx->c, ctx->c.ws, ctx->c.ws_bytes);
    ctx_sync_objects_destroy(&ctx->c);
    hipStreamDestroy(ctx->c.stream);
    delete ctx;
    return MCPT_OK;
}

int mcpt_ctx_memory(mcpt_ctx ctx, uint64_t* used, uint64_t* budget) {
    CHECK_ARG(ctx, "ctx is NULL");
    if (used) *used = ctx->c.used;
    if (budget) *budget = ctx->c.budget;
    return MCPT_OK;
}

int mcpt_ctx_describe(mcpt_ctx ctx, char* buf, size_t buflen) {
    CHECK_ARG(ctx && buf && buflen > 0, "bad arguments");
    std::snprintf(buf, buflen, "{\"device\": %d, \"name\": \"%s\", \"cus\": %d, \"clock_khz\": %d, \"budget\": %llu}",
                  ctx->c.device, ctx->c.name, ctx->c.cu_count, ctx->c.clock_khz,
                  (unsigned long long)ctx->c.budget);
    return MCPT_OK;
}

int mcpt_scene_create(mcpt_ctx ctx, const mcpt_tri_hot* hot, const mcpt_tri_cold* cold, int32_t n_tris,
                      const mcpt_material* materials, int32_t n_materials, const mcpt_env_desc* env,
                      mcpt_scene* out) {
    CHECK_ARG(n_tris >= 1, "a scene needs at least one triangle");
    return mcpt_scene_create_ex(ctx, hot, cold, n_tris, nullptr, 0, materials, n_materials, env, out);
}

int mcpt_scene_create_ex(mcpt_ctx ctx, const mcpt_tri_hot* hot, const mcpt_tri_cold* cold, int32_t n_tris,
                         const mcpt_sphere* spheres, int32_t n_spheres, const mcpt_material* materials,
                         int32_t n_materials, const mcpt_env_desc* env, mcpt_scene* out) {
    CHECK_ARG(ctx && out, "ctx/out is NULL");
    *out = nullptr;
    CHECK_ARG(n_tris >= 0 && n_spheres >= 0 && n_tris + n_spheres >= 1, "a scene needs at least one primitive");
    CHECK_ARG((int64_t)n_tris + n_spheres < (1LL << 30), "too many primitives");
    CHECK_ARG(n_tris == 0 || (hot && cold), "triangle arrays are NULL");
    CHECK_ARG(n_spheres == 0 || spheres, "sphere array is NULL");
    CHECK_ARG(n_materials >= 1 && materials, "a scene needs at least one material");
    CHECK_ARG(n_materials <= (int32_t)TRI_MATERIAL_MASK, "too many materials");
    for (int i = 0; i < n_tris; i++)
        if (cold[i].material < 0 || cold[i].material >= n_materials)
            return set_error(MCPT_ERR_INVALID, "triangle %d has material %d outside [0, %d)", i, cold[i].material,
                             n_materials);
    for (int i = 0; i < n_spheres; i++) {
        const mcpt_sphere& sp = spheres[i];
        if (sp.material < 0 || sp.material >= n_materials)
            return set_error(MCPT_ERR_INVALID, "sphere %d has material %d outside [0, %d)", i, sp.material,
                             n_materials);
        if (!(sp.radius > 0.0f) || !std::isfinite(sp.radius))
            return set_error(MCPT_ERR_INVALID, "sphere %d: radius must be positive and finite", i);
    }
    int max_tex = 0;
    for (int m = 0; m < n_materials; m++) {
        if (materials[m].texture < 0) return set_error(MCPT_ERR_INVALID, "material %d: negative texture", m);
        max_tex = std::max(max_tex, (int)materials[m].texture);
    }
    Ctx* c = &ctx->c;
    MCPT_HIP(hipSetDevice(c->device));
    auto* s = new mcpt_scene_s();
    s->ctx = c;
    s->n = n_tris;
    s->n_spheres = n_spheres;
    s->n_mats = n_materials;
    s->max_material_texture = max_tex;
    for (int m = 0; m < n_materials; m++)
        if (materials[m].type == MCPT_MAT_ISOTROPIC)
            for (int i = 0; i < n_tris; i++)
                if (cold[i].material == m) s->has_medium = true;
    for (int i = 0; i < n_spheres; i++) {
        if (materials[spheres[i].material].type == MCPT_MAT_ISOTROPIC) s->has_medium = true;
        if (spheres[i].moving) s->moving_spheres = true;
    }
    // light list (finalizeScene, src/path_tracer.cu:1973-1985)
    for (int i = 0; i < n_tris; i++) {
        const mcpt_material& m = materials[cold[i].material];
        if (m.type != MCPT_MAT_EMISSIVE) continue;
        vec3 e1 = load3(hot[i].v1) - load3(hot[i].v0);
        vec3 e2 = load3(hot[i].v2) - load3(hot[i].v0);
        vec3 av = e1.cross(e2);
        mcpt_light L;
        L.tri = hot[i];
        for (int k = 0; k < 3; k++) L.emission[k] = m.emission[k];
        L.area = 0.5f * av.length();
        vec3 nn = av.normalize();
        L.normal[0] = nn.x;
        L.normal[1] = nn.y;
        L.normal[2] = nn.z;
        s->host_lights.push_back(L);
    }
    int st = upload(c, s->hot, reinterpret_cast<const float*>(hot), (size_t)n_tris * 9);
    if (st >= 0) st = upload(c, s->cold, cold, (size_t)n_tris);
    if (st >= 0 && n_tris) {
        std::vector<uint32_t> info(n_tris);
        for (int i = 0; i < n_tris; i++) {
            const bool vn = !(cold[i].n0[0] * cold[i].n0[0] + cold[i].n0[1] * cold[i].n0[1] +
                                  cold[i].n0[2] * cold[i].n0[2] == 0.0f);
            info[i] = (uint32_t)cold[i].material | (vn ? TRI_VERTEX_NORMALS : 0u);
        }
        st = upload(c, s->tri_info, info.data(), info.size());
        if (st >= 0 && hipStreamSynchronize(c->stream) != hipSuccess)  // `info` leaves scope
            st = set_error(MCPT_ERR_HIP, "scene upload failed");
    }
    if (st >= 0) st = upload(c, s->mats, materials, (size_t)n_materials);
    if (st >= 0 && n_spheres) st = upload(c, s->spheres, spheres, (size_t)n_spheres);
    if (st >= 0 && !s->host_lights.empty()) st = upload(c, s->lights, s->host_lights.data(), s->host_lights.size());
    if (st >= 0 && env && env->rgba) {
        if (env->width < 1 || env->height < 1) st = set_error(MCPT_ERR_INVALID, "bad environment size");
        else {
            s->env_enabled = true;
            s->env_w = env->width;
            s->env_h = env->height;
            s->env_intensity = env->intensity;
            s->env_rot = env->rotation_radians;
            env_tables(env->rgba, env->width, env->height, s->host_row_cdf, s->host_col_cdf, s->env_total);
            st = upload(c, s->env_texels, reinterpret_cast<const float4*>(env->rgba),
                        (size_t)env->width * env->height);
            if (st >= 0) st = upload(c, s->env_row_cdf, s->host_row_cdf.data(), s->host_row_cdf.size());
            if (st >= 0) st = upload(c, s->env_col_cdf, s->host_col_cdf.data(), s->host_col_cdf.size());
        }
    }
    if (st >= 0 && hipStreamSynchronize(c->stream) != hipSuccess) st = set_error(MCPT_ERR_HIP, "scene upload failed");
    if (st < 0) {
        s->release();
        delete s;
        return st;
    }
    c->children++;
    *out = s;
    return MCPT_OK;
}

int mcpt_scene_s::check_textures() const {
    if (max_material_texture > (int)tex_host.size())
        return set_error(MCPT_ERR_INVALID, "a material names texture %d but the scene has %d", max_material_texture,
                         (int)tex_host.size());
    return MCPT_OK;
}

int mcpt_scene_add_texture(mcpt_scene scene, const mcpt_texture_desc* d, int32_t* index) {
    CHECK_ARG(scene && d, "bad arguments");
    if (scene->servers.load() > 0) return set_error(MCPT_ERR_INVALID, "the scene is in use by a server");
    Ctx* c = scene->ctx;
    MCPT_HIP(hipSetDevice(c->device));
    TexView v;
    std::memset(&v, 0, sizeof(v));
    v.type = d->type;
    DevBuf<uint32_t> px;
    if (d->type == MCPT_TEX_CHECKER) {
        v.scale = d->scale;
        for (int k = 0; k < 3; k++) {  // make_checker_texture converts (:211-213)
            v.ca[k] = srgb_to_linear(d->color_a[k]);
            v.cb[k] = srgb_to_linear(d->color_b[k]);
        }
    } else if (d->type == MCPT_TEX_IMAGE) {
        CHECK_ARG(d->width >= 1 && d->height >= 1 && d->rgba8, "an image texture needs pixels");
        v.w = d->width;
        v.h = d->height;
        MCPT_TRY(upload(c, px, reinterpret_cast<const uint32_t*>(d->rgba8), (size_t)d->width * d->height));
        v.px = px.ptr;
    } else {
        return set_error(MCPT_ERR_INVALID, "unknown texture type %d", d->type);
    }
    if (!scene->srgb8.ptr) {  // the sRGB decode of a texel byte, once per scene
        float lut[256];
        for (int i = 0; i < 256; i++) lut[i] = srgb_to_linear((float)i / 255.0f);
        int st = upload(c, scene->srgb8, lut, 256);
        if (st < 0) {
            px.release(c);
            return st;
        }
    }
    scene->tex_host.push_back(v);
    int st = upload(c, scene->tex, scene->tex_host.data(), scene->tex_host.size());
    if (st >= 0 && hipStreamSynchronize(c->stream) != hipSuccess) st = set_error(MCPT_ERR_HIP, "texture upload failed");
    if (st < 0) {
        scene->tex_host.pop_back();
        px.release(c);
        return st;
    }
    if (px.ptr) scene->tex_px.push_back(px);  // the scene owns the pixels now
    if (index) *index = (int32_t)scene->tex_host.size() - 1;
    return MCPT_OK;
}

int mcpt_scene_set_density_grid(mcpt_scene scene, const mcpt_density_grid* g, float* max_density) {
    CHECK_ARG(scene && g, "bad arguments");
    if (scene->servers.load() > 0) return set_error(MCPT_ERR_INVALID, "the scene is in use by a server");
    const int nx = g->dims[0], ny = g->dims[1], nz = g->dims[2];
    CHECK_ARG(nx >= 1 && ny >= 1 && nz >= 1 && g->density, "a density grid needs voxels");
    const size_t n = (size_t)nx * (size_t)ny * (size_t)nz;
    CHECK_ARG(n <= ((size_t)1 << 34), "density grid too large");
    float inv[3];
    for (int k = 0; k < 3; k++) {
        CHECK_ARG(std::isfinite(g->origin[k]) && std::isfinite(g->voxel_size[k]) && g->voxel_size[k] > 0.0f,
                  "bad density grid origin or voxel size");
        inv[k] = 1.0f / g->voxel_size[k];
        CHECK_ARG(std::isfinite(inv[k]) && inv[k] > 0.0f, "bad density grid voxel size");
    }
    float mx = 0.0f;  // the grid's largest value: the majorant's grid part
    for (size_t i = 0; i < n; i++) {
        const float v = g->density[i];
        if (!(std::isfinite(v) && v >= 0.0f))
            return set_error(MCPT_ERR_INVALID, "density grid value %zu is negative or not finite", i);
        mx = std::max(mx, v);
    }
    Ctx* c = scene->ctx;
    MCPT_HIP(hipSetDevice(c->device));
    DevBuf<float> buf;
    MCPT_TRY(upload(c, buf, g->density, n));
    if (hipStreamSynchronize(c->stream) != hipSuccess) {
        buf.release(c);
        return set_error(MCPT_ERR_HIP, "density grid upload failed");
    }
    scene->grid.release(c);
    scene->grid = buf;
    GridView& v = scene->grid_view;
    v.nx = nx;
    v.ny = ny;
    v.nz = nz;
    v.ox = g->origin[0];
    v.oy = g->origin[1];
    v.oz = g->origin[2];
    v.ix = inv[0];
    v.iy = inv[1];
    v.iz = inv[2];
    v.max_density = mx;
    if (max_density) *max_density = mx;
    return MCPT_OK;
}

static int load_ppm_file(const char* path, int32_t* width, int32_t* height, uint8_t* rgba8);

int mcpt_load_ppm(const char* path, int32_t* width, int32_t* height, uint8_t* rgba8) {
    return guarded("mcpt_load_ppm", [&] { return load_ppm_file(path, width, height, rgba8); });
}

int mcpt_load_png(const char* path, int32_t* width, int32_t* height, uint8_t* rgba8) {
    return guarded("mcpt_load_png", [&]() -> int {
        CHECK_ARG(path && width && height, "bad arguments");
        int w = 0, h = 0;
        std::vector<uint8_t> px;
        MCPT_TRY(read_png(path, w, h, rgba8 ? &px : nullptr));
        *width = w;
        *height = h;
        if (rgba8) std::memcpy(rgba8, px.data(), px.size());
        return MCPT_OK;
    });
}

int mcpt_load_tga(const char* path, int32_t* width, int32_t* height, uint8_t* rgba8) {
    return guarded("mcpt_load_tga", [&]() -> int {
        CHECK_ARG(path && width && height, "bad arguments");
        int w = 0, h = 0;
        std::vector<uint8_t> px;
        MCPT_TRY(read_tga(path, w, h, rgba8 ? &px : nullptr));
        *width = w;
        *height = h;
        if (rgba8) std::memcpy(rgba8, px.data(), px.size());
        return MCPT_OK;
    });
}

int mcpt_load_jpeg(const char* path, int32_t* width, int32_t* height, uint8_t* rgba8) {
    return guarded("mcpt_load_jpeg", [&]() -> int {
        CHECK_ARG(path && width && height, "bad arguments");
        int w = 0, h = 0;
        std::vector<uint8_t> px;
        MCPT_TRY(read_jpeg(path, w, h, rgba8 ? &px : nullptr));
        *width = w;
        *height = h;
        if (rgba8) std::memcpy(rgba8, px.data(), px.size());
        return MCPT_OK;
    });
}

static int load_ppm_file(const char* path, int32_t* width, int32_t* height, uint8_t* rgba8) {
    CHECK_ARG(path && width && height, "bad arguments");
    int w = 0, h = 0;
    std::vector<uint8_t> px;
    MCPT_TRY(read_ppm(path, w, h, rgba8 ? &px : nullptr));
    *width = w;
    *height = h;
    if (rgba8) std::memcpy(rgba8, px.data(), px.size());
    return MCPT_OK;
}

// the environment readers share the two-call shape
static int load_env_file(int (*reader)(const char*, int&, int&, std::vector<float>*), const char* path,
                         int32_t* width, int32_t* height, float* rgba) {
    CHECK_ARG(path && width && height, "bad arguments");
    int w = 0, h = 0;
    std::vector<float> px;
    MCPT_TRY(reader(path, w, h, rgba ? &px : nullptr));
    *width = w;
    *height = h;
    if (rgba) std::memcpy(rgba, px.data(), px.size() * sizeof(float));
    return MCPT_OK;
}

int mcpt_load_pfm(const char* path, int32_t* width, int32_t* height, float* rgba) {
    return guarded("mcpt_load_pfm", [&] { return load_env_file(read_pfm, path, width, height, rgba); });
}

int mcpt_load_hdr(const char* path, int32_t* width, int32_t* height, float* rgba) {
    return guarded("mcpt_load_hdr", [&] { return load_env_file(read_hdr, path, width, height, rgba); });
}

int mcpt_load_exr(const char* path, int32_t* width, int32_t* height, float* rgba) {
    return guarded("mcpt_load_exr", [&] { return load_env_file(read_exr, path, width, height, rgba); });
}

int mcpt_load_nvdb(const char* path, mcpt_density_grid* grid, float* density, float* max_density) {
    if (!path || !grid) return set_error(MCPT_ERR_INVALID, "bad arguments");
    return guarded("mcpt_load_nvdb", [&] { return read_nvdb(path, grid, density, max_density); });
}

int mcpt_scene_destroy(mcpt_scene scene) {
    if (!scene) return MCPT_OK;
    const int srv = scene->servers.load();
    if (srv > 0)
        return set_error(MCPT_ERR_INVALID, "the scene is in use by %d server(s); destroy them first", srv);
    Ctx* c = scene->ctx;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    scene->release();
    delete scene;
    c->children--;
    return MCPT_OK;
}

int mcpt_scene_light_count(mcpt_scene scene, int32_t* n) {
    CHECK_ARG(scene && n, "bad arguments");
    *n = (int32_t)scene->host_lights.size();
    return MCPT_OK;
}

int mcpt_scene_lights(mcpt_scene scene, mcpt_light* out) {
    CHECK_ARG(scene && out, "bad arguments");
    std::memcpy(out, scene->host_lights.data(), scene->host_lights.size() * sizeof(mcpt_light));
    return MCPT_OK;
}

int mcpt_scene_env_tables(mcpt_scene scene, float* row_cdf, float* col_cdf, float* total_weight) {
    CHECK_ARG(scene, "scene is NULL");
    CHECK_ARG(scene->env_enabled, "scene has no environment map");
    if (row_cdf) std::memcpy(row_cdf, scene->host_row_cdf.data(), scene->host_row_cdf.size() * sizeof(float));
    if (col_cdf) std::memcpy(col_cdf, scene->host_col_cdf.data(), scene->host_col_cdf.size() * sizeof(float));
    if (total_weight) *total_weight = scene->env_total;
    return MCPT_OK;
}

int mcpt_lbvh_build(mcpt_scene scene, mcpt_lbvh_dump* dump) {
    CHECK_ARG(scene, "scene is NULL");
    if (scene->servers.load() > 0) return set_error(MCPT_ERR_INVALID, "the scene is in use by a server");
    MCPT_HIP(hipSetDevice(scene->ctx->device));
    return lbvh_build(scene, dump);
}

int mcpt_trace_closest(mcpt_scene scene, const mcpt_ray* rays, int32_t n, mcpt_packed_hit* hits, uint32_t* steps,
                       uint32_t* leaves) {
    CHECK_ARG(scene, "scene is NULL");
    MCPT_HIP(hipSetDevice(scene->ctx->device));
    return trace_closest_host(scene, rays, n, hits, steps, leaves);
}

int mcpt_trace_any(mcpt_scene scene, const mcpt_ray* rays, int32_t n, uint8_t* occluded) {
    CHECK_ARG(scene, "scene is NULL");
    MCPT_HIP(hipSetDevice(scene->ctx->device));
    return trace_any_host(scene, rays, n, occluded);
}

int mcpt_lbvh_dump_wide(mcpt_scene scene, void* nodes, int32_t cap, int32_t* n_wide, int32_t* tie) {
    CHECK_ARG(scene && n_wide && cap >= 0, "bad arguments");
    if (!scene->built) return set_error(MCPT_ERR_INVALID, "scene is not built");
    MCPT_HIP(hipSetDevice(scene->ctx->device));
    MCPT_HIP(hipStreamSynchronize(scene->ctx->stream));
    *n_wide = scene->n_wide;
    if (nodes && cap > 0)
        MCPT_HIP(hipMemcpy(nodes, scene->wide.ptr, (size_t)std::min(cap, scene->n_wide) * sizeof(WideNode),
                           hipMemcpyDeviceToHost));
    if (tie) MCPT_HIP(hipMemcpy(tie, scene->tie.ptr, (size_t)scene->n_prims() * sizeof(int), hipMemcpyDeviceToHost));
    return MCPT_OK;
}

int mcpt_camera_init(const mcpt_camera_desc* desc, mcpt_camera* out) {
    CHECK_ARG(desc && out, "bad arguments");
    return camera_init(desc, out);
}

int mcpt_film_create(mcpt_ctx ctx, int32_t width, int32_t height, mcpt_film* out) {
    CHECK_ARG(ctx && out, "bad arguments");
    CHECK_ARG(width >= 1 && height >= 1, "film size must be positive");
    *out = nullptr;
    MCPT_HIP(hipSetDevice(ctx->c.device));
    auto* f = new mcpt_film_s();
    f->ctx = &ctx->c;
    f->w = width;
    f->h = height;
    int st = f->accum.alloc(&ctx->c, (size_t)width * height * 3);
    if (st >= 0 && hipMemsetAsync(f->accum.ptr, 0, f->accum.bytes(), ctx->c.stream) != hipSuccess)
        st = set_error(MCPT_ERR_HIP, "film clear failed");
    if (st >= 0 && hipStreamSynchronize(ctx->c.stream) != hipSuccess) st = set_error(MCPT_ERR_HIP, "film clear failed");
    if (st < 0) {
        f->accum.release(&ctx->c);
        delete f;
        return st;
    }
    ctx->c.children++;
    *out = f;
    return MCPT_OK;
}

int mcpt_film_destroy(mcpt_film film) {
    if (!film) return MCPT_OK;
    Ctx* c = film->ctx;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    film->accum.release(c);
    delete film;
    c->children--;
    return MCPT_OK;
}

int mcpt_film_clear(mcpt_film film) {
    CHECK_ARG(film, "film is NULL");
    MCPT_HIP(hipSetDevice(film->ctx->device));
    MCPT_HIP(hipMemsetAsync(film->accum.ptr, 0, film->accum.bytes(), film->ctx->stream));
    MCPT_HIP(hipStreamSynchronize(film->ctx->stream));
    return MCPT_OK;
}

int mcpt_film_read(mcpt_film film, float* rgb) {
    CHECK_ARG(film && rgb, "bad arguments");
    MCPT_HIP(hipSetDevice(film->ctx->device));
    MCPT_HIP(hipMemcpyAsync(rgb, film->accum.ptr, film->accum.bytes(), hipMemcpyDeviceToHost, film->ctx->stream));
    MCPT_HIP(hipStreamSynchronize(film->ctx->stream));
    return MCPT_OK;
}

int mcpt_film_write(mcpt_film film, const float* rgb) {
    CHECK_ARG(film && rgb, "bad arguments");
    MCPT_HIP(hipSetDevice(film->ctx->device));
    MCPT_HIP(hipMemcpyAsync(film->accum.ptr, rgb, film->accum.bytes(), hipMemcpyHostToDevice, film->ctx->stream));
    MCPT_HIP(hipStreamSynchronize(film->ctx->stream));
    return MCPT_OK;
}

int mcpt_render(mcpt_scene scene, const mcpt_camera* camera, const mcpt_render_opts* opts, mcpt_film film,
                mcpt_render_stats* stats) {
    CHECK_ARG(scene && camera && opts && film, "bad arguments");
    CHECK_ARG(film->ctx == scene->ctx, "film and scene belong to different contexts");
    CHECK_ARG(opts->mode >= MCPT_MODE_MEGAKERNEL && opts->mode <= MCPT_MODE_PERSISTENT, "unknown render mode");
    MCPT_TRY(scene->check_textures());
    MCPT_HIP(hipSetDevice(scene->ctx->device));
    return render(scene, camera, opts, film, stats);
}

int mcpt_render_noalloc(mcpt_scene scene, const mcpt_camera* camera, const mcpt_render_opts* opts, mcpt_film film) {
    CHECK_ARG(scene && camera && opts && film, "bad arguments");
    CHECK_ARG(film->ctx == scene->ctx, "film and scene belong to different contexts");
    CHECK_ARG(opts->mode >= MCPT_MODE_MEGAKERNEL && opts->mode <= MCPT_MODE_PERSISTENT, "unknown render mode");
    MCPT_TRY(scene->check_textures());
    MCPT_HIP(hipSetDevice(scene->ctx->device));
    return render_noalloc(scene, camera, opts, film);
}

int mcpt_resolve(mcpt_film film, int32_t samples_so_far, int32_t tonemap, float exposure, float* out_rgb) {
    CHECK_ARG(film && out_rgb, "bad arguments");
    CHECK_ARG(samples_so_far >= 1, "samples_so_far must be >= 1");
    CHECK_ARG(tonemap >= MCPT_TONEMAP_NONE && tonemap <= MCPT_TONEMAP_ACES, "unknown tone map");
    MCPT_HIP(hipSetDevice(film->ctx->device));
    return resolve(film, samples_so_far, tonemap, exposure, out_rgb);
}

int mcpt_resolve_rgb8(mcpt_film film, int32_t samples_so_far, int32_t tonemap, float exposure, uint8_t* out_rgb8) {
    CHECK_ARG(film && out_rgb8 && samples_so_far > 0, "bad arguments");
    CHECK_ARG(tonemap >= MCPT_TONEMAP_NONE && tonemap <= MCPT_TONEMAP_ACES, "unknown tone map");
    MCPT_HIP(hipSetDevice(film->ctx->device));
    return resolve_rgb8(film, samples_so_far, tonemap, exposure, out_rgb8);
}

int mcpt_auto_exposure(mcpt_film film, int32_t samples_so_far, float target_gray, float* exposure) {
    CHECK_ARG(film && exposure && samples_so_far > 0, "bad arguments");
    MCPT_HIP(hipSetDevice(film->ctx->device));
    return auto_exposure(film, samples_so_far, target_gray, exposure);
}

int mcpt_write_png(const char* path, int32_t width, int32_t height, const uint8_t* rgb8) {
    return write_png(path, width, height, rgb8);
}

int mcpt_gen_scene(int32_t kind, uint64_t seed, int64_t param, mcpt_tri_hot* hot, mcpt_tri_cold* cold,
                   mcpt_material* materials, mcpt_gen_sizes* sizes) {
    return gen_scene(kind, seed, param, hot, cold, materials, sizes);
}

int mcpt_gen_sky(int32_t width, int32_t height, float* rgba) { return gen_sky(width, height, rgba); }

int mcpt_load_obj(const char* path, const mcpt_obj_opts* opts, mcpt_tri_hot* hot, mcpt_tri_cold* cold,
                  mcpt_material* materials, mcpt_obj_stats* stats) {
    return guarded("mcpt_load_obj", [&] { return load_obj(path, opts, hot, cold, materials, stats); });
}

int mcpt_obj_texture(int32_t index, int32_t* width, int32_t* height, uint8_t* rgba8) {
    return obj_texture(index, width, height, rgba8);
}

int mcpt_load_pbrt(const char* path, int64_t max_prims, mcpt_tri_hot* hot, mcpt_tri_cold* cold, mcpt_sphere* spheres,
                   mcpt_material* materials, mcpt_pbrt_info* info) {
    return guarded("mcpt_load_pbrt", [&] { return load_pbrt(path, max_prims, hot, cold, spheres, materials, info); });
}

int mcpt_pbrt_texture(int32_t index, int32_t* width, int32_t* height, uint8_t* rgba8) {
    return pbrt_texture(index, width, height, rgba8);
}

int mcpt_gen_camera(int32_t kind, int32_t width, int32_t height, mcpt_camera_desc* out) {
    return gen_camera(kind, width, height, out);
}

}  // extern "C"
