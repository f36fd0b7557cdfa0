// host_common.hpp — host-side plumbing shared by the library's translation
// units: error reporting, the budgeted device allocator, and the internal
// objects behind the opaque handles of mcpt.h.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <mutex>
#include <vector>

#include "mcpt.h"

namespace mcpt {

int set_error(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

#define MCPT_HIP(call)                                                                                       \
    do {                                                                                                     \
        hipError_t e_ = (call);                                                                              \
        if (e_ != hipSuccess)                                                                                \
            return ::mcpt::set_error(MCPT_ERR_HIP, "%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
    } while (0)

#define MCPT_TRY(call)          \
    do {                        \
        int s_ = (call);        \
        if (s_ < 0) return s_;  \
    } while (0)

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // Streams (default priority) and events of a render's lanes. Lane 0 runs
    // on `stream` with shadow launches on `side`, lane l > 0 on
    // lane_stream[l] and lane_side[l]; ev_shade / ev_shadow order a lane's
    // shade and shadow launches, ev_fin the lanes' film updates, ev_join the
    // lanes' start after `stream` and `stream`'s wait for the last lane.
    static constexpr int kMaxLanes = 2;
    hipStream_t side = nullptr;
    hipStream_t lane_stream[kMaxLanes] = {}, lane_side[kMaxLanes] = {};
    hipEvent_t ev_shade[kMaxLanes] = {}, ev_shadow[kMaxLanes] = {};
    hipEvent_t ev_fin[kMaxLanes] = {}, ev_join = nullptr;
    std::mutex alloc_mu;  // guards `used` (a server thread allocates too)
    uint64_t budget = 0;
    uint64_t used = 0;
    int cu_count = 0;
    int clock_khz = 0;
    char name[256] = {0};
    // live scenes, films, servers and networks made from this context; the
    // context may not be destroyed while any remains (they hold its stream
    // and allocator)
    std::atomic<int> children{0};
    // reusable wavefront workspace (see render.hip)
    void* ws = nullptr;
    size_t ws_bytes = 0;
};

// the render lanes' streams and events (capi.cpp)
int ctx_sync_objects_create(Ctx* c);
void ctx_sync_objects_destroy(Ctx* c);

// every device allocation of a context goes through these two calls, which
// keep `used` within `budget`
int dev_alloc(Ctx* ctx, void** p, size_t bytes);
void dev_free(Ctx* ctx, void* p, size_t bytes);

template <class T>
struct DevBuf {
    T* ptr = nullptr;
    size_t count = 0;
    int alloc(Ctx* ctx, size_t n) {
        release(ctx);
        if (n == 0) return MCPT_OK;
        void* p = nullptr;
        int s = dev_alloc(ctx, &p, n * sizeof(T));
        if (s < 0) return s;
        ptr = (T*)p;
        count = n;
        return MCPT_OK;
    }
    void release(Ctx* ctx) {
        if (ptr) dev_free(ctx, ptr, count * sizeof(T));
        ptr = nullptr;
        count = 0;
    }
    size_t bytes() const { return count * sizeof(T); }
};

// image files (imageio.cpp); a NULL vector reads the header only
int read_ppm(const char* path, int& w, int& h, std::vector<uint8_t>* rgba);
int read_pfm(const char* path, int& w, int& h, std::vector<float>* rgba);
int read_hdr(const char* path, int& w, int& h, std::vector<float>* rgba);
int read_exr(const char* path, int& w, int& h, std::vector<float>* rgba);
// a NanoVDB float grid as a dense density grid (nvdb.cpp); density NULL
// fills the description only
int read_nvdb(const char* path, mcpt_density_grid* out, float* density, float* max_density);
int read_png(const char* path, int& w, int& h, std::vector<uint8_t>* rgba);
int read_tga(const char* path, int& w, int& h, std::vector<uint8_t>* rgba);
int read_jpeg(const char* path, int& w, int& h, std::vector<uint8_t>* rgba);
// an 8-bit texture: TGA by its extension, PNG or JPEG by its signature, else binary PPM
int read_image8(const char* path, int& w, int& h, std::vector<uint8_t>* rgba);
// pbrt-v3 import (pbrtload.cpp)
int load_pbrt(const char* path, int64_t max_prims, mcpt_tri_hot* hot, mcpt_tri_cold* cold, mcpt_sphere* spheres,
              mcpt_material* materials, mcpt_pbrt_info* info);
int pbrt_texture(int index, int32_t* width, int32_t* height, uint8_t* rgba8);
// OBJ + MTL import (objload.cpp)
int obj_texture(int index, int32_t* width, int32_t* height, uint8_t* rgba8);
int load_obj(const char* path, const mcpt_obj_opts* opts, mcpt_tri_hot* hot, mcpt_tri_cold* cold,
             mcpt_material* materials, mcpt_obj_stats* stats);
// scene generators (scenegen.cpp); srgb_to_linear is srgbToLinear
// (src/path_tracer.cu:200-202), the albedo conversion of make_solid_texture
float srgb_to_linear(float e);
int gen_scene(int kind, uint64_t seed, int64_t param, mcpt_tri_hot* hot, mcpt_tri_cold* cold,
              mcpt_material* materials, mcpt_gen_sizes* sizes);
int gen_sky(int w, int h, float* rgba);
int gen_camera(int kind, int w, int h, mcpt_camera_desc* d);

}  // namespace mcpt
