// scene.hpp — device-resident scene, film and the launchers each kernel
// file provides.
#pragma once

#include <atomic>

#include "device.hpp"
#include "host_common.hpp"

struct mcpt_ctx_s {
    mcpt::Ctx c;
};

struct mcpt_scene_s {
    mcpt::Ctx* ctx = nullptr;
    int n = 0;          // triangles
    int n_spheres = 0;  // the BVH's primitives are the spheres, then the triangles
    bool moving_spheres = false;
    int n_mats = 0;
    bool has_medium = false;
    // inputs (original triangle order)
    mcpt::DevBuf<float> hot;  // 9 floats per triangle
    mcpt::DevBuf<mcpt_tri_cold> cold;
    mcpt::DevBuf<uint32_t> tri_info;  // material index | TRI_VERTEX_NORMALS
    mcpt::DevBuf<mcpt_material> mats;
    mcpt::DevBuf<mcpt_sphere> spheres;
    mcpt::DevBuf<mcpt_light> lights;
    std::vector<mcpt_light> host_lights;
    // environment
    bool env_enabled = false;
    int env_w = 0, env_h = 0;
    float env_intensity = 1.f, env_rot = 0.f, env_total = 0.f;
    mcpt::DevBuf<float4> env_texels;
    mcpt::DevBuf<float> env_row_cdf, env_col_cdf;
    std::vector<float> host_row_cdf, host_col_cdf;
    // BVH
    bool built = false;
    int max_depth = 0;
    mcpt::DevBuf<mcpt::BvhNode> nodes;
    mcpt::DevBuf<mcpt::LeafTri> leaf_tri;
    mcpt::DevBuf<int> leaf_prim;
    mcpt::DevBuf<float> root_box;  // 6 floats: min xyz, max xyz
    // textures (mcpt_scene_add_texture): the table the kernels read, each
    // image's pixels, the sRGB byte table, and the highest texture index a
    // material names (checked before a render)
    std::vector<mcpt::TexView> tex_host;
    mcpt::DevBuf<mcpt::TexView> tex;
    std::vector<mcpt::DevBuf<uint32_t>> tex_px;
    mcpt::DevBuf<float> srgb8;
    int max_material_texture = 0;
    // density grid of every isotropic material (mcpt_scene_set_density_grid)
    mcpt::DevBuf<float> grid;
    mcpt::GridView grid_view{};
    // servers holding views of this scene: it may not be rebuilt or get
    // textures (the views would point at freed buffers) until they end
    std::atomic<int> servers{0};
    mcpt::DevBuf<mcpt::WideNode> wide;  // four-wide quantized tree, root 0
    mcpt::DevBuf<int> tie;              // leaf -> reference leaf-test rank
    int n_wide = 0;
    float last_build_ms = 0.f;

    mcpt::SceneView view() const {
        mcpt::SceneView v;
        v.bvh.nodes = nodes.ptr;
        v.bvh.leaf_tri = leaf_tri.ptr;
        v.bvh.leaf_prim = leaf_prim.ptr;
        v.bvh.wide = wide.ptr;
        v.bvh.tie = tie.ptr;
        v.hot = hot.ptr;
        v.cold = cold.ptr;
        v.tri_info = tri_info.ptr;
        v.mats = mats.ptr;
        v.lights = lights.ptr;
        v.n_lights = (int)host_lights.size();
        v.n_tris = n;
        v.has_medium = has_medium ? 1 : 0;
        v.stack_depth = max_depth;
        v.env.texels = env_texels.ptr;
        v.env.row_cdf = env_row_cdf.ptr;
        v.env.col_cdf = env_col_cdf.ptr;
        v.env.w = env_w;
        v.env.h = env_h;
        v.env.intensity = env_intensity;
        v.env.rot = env_rot;
        v.env.total = env_total;
        v.env.enabled = env_enabled ? 1 : 0;
        v.spheres = spheres.ptr;
        v.n_spheres = n_spheres;
        v.tex = tex.ptr;
        v.srgb8 = srgb8.ptr;
        v.grid = grid_view;
        v.grid.data = grid.ptr;
        return v;
    }
    void release() {
        hot.release(ctx);
        cold.release(ctx);
        tri_info.release(ctx);
        mats.release(ctx);
        spheres.release(ctx);
        lights.release(ctx);
        env_texels.release(ctx);
        env_row_cdf.release(ctx);
        env_col_cdf.release(ctx);
        nodes.release(ctx);
        leaf_tri.release(ctx);
        leaf_prim.release(ctx);
        root_box.release(ctx);
        wide.release(ctx);
        tie.release(ctx);
        tex.release(ctx);
        for (auto& b : tex_px) b.release(ctx);
        tex_px.clear();
        tex_host.clear();
        srgb8.release(ctx);
        grid.release(ctx);
    }
    int n_prims() const { return n + n_spheres; }
    // every texture a material names has been added
    int check_textures() const;

};

namespace mcpt {
// Traversal stack ints per lane for a tree whose deepest leaf has max_depth
// internal ancestors: the binary walk keeps at most one pending node per
// level, the four-wide walk one two-int group per wide level (two binary
// levels).
inline int stack_entries(int max_depth) {
    const int d = max_depth < 1 ? 1 : max_depth;
    return 2 * ((d + 1) / 2);
}
}  // namespace mcpt

struct mcpt_film_s {
    mcpt::Ctx* ctx = nullptr;
    int w = 0, h = 0;
    mcpt::DevBuf<float> accum;  // W*H*3
};

namespace mcpt {

// lbvh.hip
int lbvh_build(mcpt_scene_s* s, mcpt_lbvh_dump* dump);

// trace.hip
int trace_closest_host(mcpt_scene_s* s, const mcpt_ray* rays, int n, mcpt_packed_hit* hits, uint32_t* steps,
                       uint32_t* leaves);
int trace_any_host(mcpt_scene_s* s, const mcpt_ray* rays, int n, uint8_t* occluded);

// render.hip
int render(mcpt_scene_s* s, const mcpt_camera* cam, const mcpt_render_opts* o, mcpt_film_s* film,
           mcpt_render_stats* stats);
// the same over the fixed per-device workspace (no device allocation);
// scratch_init finds that workspace's address on the current device
int render_noalloc(mcpt_scene_s* s, const mcpt_camera* cam, const mcpt_render_opts* o, mcpt_film_s* film);
int scratch_init(int device);
int resolve(mcpt_film_s* film, int spp, int tonemap, float exposure, float* out);
int resolve_rgb8(mcpt_film_s* film, int spp, int tonemap, float exposure, uint8_t* out);
int auto_exposure(mcpt_film_s* film, int spp, float target, float* exposure);
int write_png(const char* path, int w, int h, const uint8_t* rgb8);

}  // namespace mcpt
