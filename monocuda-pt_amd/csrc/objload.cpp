// objload.cpp — Wavefront OBJ + MTL import into the library's triangle
// records (loadObject / loadMtlFile / parseObjFaceToken, src/path_tracer.cu
// :2527-2859), host code.
//
// The reference's rules, kept:
//   * v / vt / vn / f / mtllib / usemtl; OBJ indices start at 1 and negative
//     ones count back from the end of the list read so far; a face token is
//     "p", "p/t", "p/t/n" or "p//n"; polygons are fan-triangulated
//     (0, k, k + 1).
//   * The model is centred on x and z, its lowest point put at y = 0 and its
//     largest side scaled to 1, then uniform_scale, a rotation about +y and
//     the translation (where the bottom centre ends up) are applied, in FP32
//     in the reference's operation order (libm cosf / sinf of the angle).
//   * Faces with an index outside the lists, or whose cross product is
//     shorter than 1e-9, are skipped.
//   * Vertex normals from vn are rotated and normalised; a model without vn
//     gets area-weighted averaged vertex normals (cross products summed in
//     face order), zero vectors staying zero (shading then uses the face
//     normal).
//   * MTL: newmtl / Kd / Ke / map_Kd; a material with a positive Ke becomes
//     emissive (emission = Ke), otherwise Lambertian with albedo Kd converted
//     from sRGB to linear (make_solid_texture, :207-209). usemtl of an
//     unknown name uses the library's only material when it has exactly one,
//     otherwise the fallback material.
//   * map_Kd (with opts.load_textures): the path with '\\' turned into '/'
//     and its extension replaced by ".ppm", relative to the OBJ file's
//     folder (convertTexturePathToPpm, :2580-2591), is read as a binary PPM
//     once per path (getOrLoadObjTexture, :2594-2611); a material whose
//     image loads is Lambertian with that image texture (material.texture =
//     1 + its position in the list mcpt_obj_texture hands out), otherwise
//     it keeps its Kd colour and counts as missing, as in the reference.
//     Other image formats, which the reference decodes with stb_image, are
//     not read: those materials keep Kd.
#include <sys/stat.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "device.hpp"
#include "host_common.hpp"

namespace mcpt {

namespace {

struct Corner {
    int p, t, n;  // resolved indices, -1 = missing
};

struct Face {
    Corner c[3];
    int material;  // index into the MTL library, -1 = fallback
};

struct Vec2 {
    float x = 0.f, y = 0.f;
};

int resolve_index(int raw, int size) {  // resolveObjIndex (:2546-2550)
    if (raw > 0) return raw - 1;
    if (raw < 0) return size + raw;
    return -1;
}

// parseObjFaceToken (:2553-2571): up to three '/'-separated integers
Corner parse_corner(const std::string& tok, int np, int nt, int nn) {
    int raw[3] = {0, 0, 0};
    size_t start = 0;
    for (int part = 0; part < 3; part++) {
        const size_t slash = tok.find('/', start);
        const std::string piece = tok.substr(start, slash == std::string::npos ? std::string::npos : slash - start);
        if (!piece.empty()) raw[part] = std::atoi(piece.c_str());
        if (slash == std::string::npos) break;
        start = slash + 1;
    }
    return Corner{resolve_index(raw[0], np), resolve_index(raw[1], nt), resolve_index(raw[2], nn)};
}

std::string directory_of(const std::string& path) {
    const size_t s = path.find_last_of("/\\");
    return s == std::string::npos ? std::string(".") : path.substr(0, s);
}

struct Image {
    int w = 0, h = 0;
    std::vector<uint8_t> rgba;
};

struct Library {
    std::vector<mcpt_material> mats;
    std::map<std::string, int> by_name;
    int texture_requests = 0;  // map_Kd materials
    // texture loading (opts.load_textures)
    bool load_textures = false;
    std::string obj_dir;
    std::map<std::string, int> tex_by_path;  // -1: the file could not be read
    std::shared_ptr<std::vector<Image>> images = std::make_shared<std::vector<Image>>();
    int missing_textures = 0;
};

// convertTexturePathToPpm (:2580-2591)
std::string ppm_path_of(std::string p) {
    for (char& ch : p)
        if (ch == '\\') ch = '/';
    const size_t dot = p.find_last_of('.'), slash = p.find_last_of('/');
    if (dot != std::string::npos && (slash == std::string::npos || dot > slash)) p = p.substr(0, dot);
    return p + ".ppm";
}

// getOrLoadObjTexture (:2594-2611): one read per .ppm path, failures
// remembered; the .ppm first, then the named file itself (PNG or PPM here,
// where the reference's fallback is stb_image, :2602)
int texture_for(Library& lib, const std::string& map_kd) {
    const std::string ppm = lib.obj_dir + "/" + ppm_path_of(map_kd);
    auto it = lib.tex_by_path.find(ppm);
    if (it != lib.tex_by_path.end()) return it->second;
    std::string named = map_kd;
    for (char& ch : named)
        if (ch == '\\') ch = '/';
    named = lib.obj_dir + "/" + named;
    Image im;
    int idx = -1;
    if (read_ppm(ppm.c_str(), im.w, im.h, &im.rgba) == MCPT_OK ||
        (named != ppm && read_image8(named.c_str(), im.w, im.h, &im.rgba) == MCPT_OK)) {
        idx = (int)lib.images->size();
        lib.images->push_back(std::move(im));
    }
    clear_error();  // a missing texture is not an error of the load
    lib.tex_by_path[ppm] = idx;
    return idx;
}

mcpt_material lambertian_srgb(const vec3& srgb) {
    mcpt_material m;
    std::memset(&m, 0, sizeof(m));
    m.type = MCPT_MAT_LAMBERTIAN;
    m.albedo[0] = srgb_to_linear(srgb.x);
    m.albedo[1] = srgb_to_linear(srgb.y);
    m.albedo[2] = srgb_to_linear(srgb.z);
    return m;
}

// loadMtlFile (:2617-2678)
bool load_mtl(const std::string& path, Library& lib) {
    std::ifstream f(path);
    if (!f.is_open()) return false;
    std::string name, texture;
    vec3 kd(0.8f, 0.8f, 0.8f), ke(0.f, 0.f, 0.f);
    bool open = false;
    auto finish = [&]() {
        if (!open) return;
        mcpt_material m = lambertian_srgb(kd);
        if (std::fmax(ke.x, std::fmax(ke.y, ke.z)) > 0.f) {
            std::memset(&m, 0, sizeof(m));
            m.type = MCPT_MAT_EMISSIVE;
            m.emission[0] = ke.x;
            m.emission[1] = ke.y;
            m.emission[2] = ke.z;
        } else if (!texture.empty()) {
            lib.texture_requests++;
            if (lib.load_textures) {
                const int idx = texture_for(lib, texture);
                if (idx >= 0)
                    m.texture = idx + 1;
                else
                    lib.missing_textures++;
            }
        }
        lib.by_name[name] = (int)lib.mats.size();
        lib.mats.push_back(m);
        open = false;
    };
    std::string line;
    while (std::getline(f, line)) {
        std::istringstream ls(line);
        std::string key;
        ls >> key;
        if (key == "newmtl") {
            finish();
            name.clear();
            ls >> name;
            kd = vec3(0.8f, 0.8f, 0.8f);
            ke = vec3(0.f, 0.f, 0.f);
            texture.clear();
            open = true;
        } else if (key == "Kd" && open) {
            float r = 0.f, g = 0.f, b = 0.f;
            ls >> r >> g >> b;
            kd = vec3(r, g, b);
        } else if (key == "Ke" && open) {
            float r = 0.f, g = 0.f, b = 0.f;
            ls >> r >> g >> b;
            ke = vec3(r, g, b);
        } else if (key == "map_Kd" && open) {
            ls >> texture;
        }
    }
    finish();
    return true;
}

struct Loaded {
    std::vector<mcpt_tri_hot> hot;
    std::vector<mcpt_tri_cold> cold;
    std::vector<mcpt_material> mats;
    std::shared_ptr<std::vector<Image>> images;
    mcpt_obj_stats stats;
};

// whitespace as the stream extraction the reference parses with skips it
bool is_space(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }

// next whitespace-separated token of a NUL-terminated line
bool next_token(char*& p, std::string& tok) {
    while (is_space(*p)) p++;
    if (!*p) return false;
    char* q = p;
    while (*q && !is_space(*q)) q++;
    tok.assign(p, q);
    p = q;
    return true;
}

// up to n floats (strtof, the conversion stream extraction uses); a missing
// or malformed one stays 0 and ends the line's reading
int read_floats(char*& p, float* v, int n) {
    for (int i = 0; i < n; i++) {
        while (is_space(*p)) p++;
        char* e = nullptr;
        const float x = std::strtof(p, &e);
        if (e == p) return i;
        v[i] = x;
        p = e;
    }
    return n;
}

int load(const char* path, const mcpt_obj_opts* o, Loaded& out) {
    // the whole file in memory, lines cut at their newline: the parse is
    // the reference's line-by-line stream parse without a stream per line
    std::FILE* fp = std::fopen(path, "rb");
    if (!fp) return set_error(MCPT_ERR_INVALID, "cannot open OBJ file '%s'", path);
    std::string buf;
    char chunk[1 << 16];
    size_t got;
    while ((got = std::fread(chunk, 1, sizeof(chunk), fp)) > 0) buf.append(chunk, got);
    std::fclose(fp);
    for (char& ch : buf)
        if (ch == '\n') ch = '\0';
    buf.push_back('\0');
    const std::string dir = directory_of(path);
    std::vector<vec3> pos, nrm;
    std::vector<Vec2> uvs;
    std::vector<Face> faces;
    Library lib;
    lib.load_textures = o->load_textures != 0;
    lib.obj_dir = dir;
    int current = -1, missing_libs = 0;
    std::string key, tok;
    std::vector<Corner> cs;
    for (size_t at = 0; at < buf.size();) {
        char* p = &buf[at];
        const size_t len = std::strlen(p);
        at += len + 1;
        if (!next_token(p, key)) continue;
        if (key == "v" || key == "vn") {
            float v[3] = {0.f, 0.f, 0.f};
            read_floats(p, v, 3);
            (key == "v" ? pos : nrm).push_back(vec3(v[0], v[1], v[2]));
        } else if (key == "vt") {
            float v[2] = {0.f, 0.f};
            read_floats(p, v, 2);
            Vec2 t;
            t.x = v[0];
            t.y = v[1];
            uvs.push_back(t);
        } else if (key == "mtllib") {
            while (next_token(p, tok))
                if (!load_mtl(dir + "/" + tok, lib)) missing_libs++;
        } else if (key == "usemtl") {
            std::string name;
            next_token(p, name);
            auto it = lib.by_name.find(name);
            current = it != lib.by_name.end() ? it->second : (lib.mats.size() == 1 ? 0 : -1);
        } else if (key == "f") {
            cs.clear();
            while (next_token(p, tok)) cs.push_back(parse_corner(tok, (int)pos.size(), (int)uvs.size(), (int)nrm.size()));
            for (size_t k = 1; k + 1 < cs.size(); k++) faces.push_back(Face{{cs[0], cs[k], cs[k + 1]}, current});
        }
    }
    if (pos.empty() || faces.empty()) return set_error(MCPT_ERR_INVALID, "OBJ file '%s' has no vertices or faces", path);

    // normalise: bottom centre to the origin, largest side to 1 (:2733-2742)
    vec3 lo = pos[0], hi = pos[0];
    for (size_t i = 1; i < pos.size(); i++) {
        lo = vec3(std::fmin(lo.x, pos[i].x), std::fmin(lo.y, pos[i].y), std::fmin(lo.z, pos[i].z));
        hi = vec3(std::fmax(hi.x, pos[i].x), std::fmax(hi.y, pos[i].y), std::fmax(hi.z, pos[i].z));
    }
    const vec3 size = hi - lo;
    const float largest = std::fmax(size.x, std::fmax(size.y, size.z));
    const vec3 bottom((lo.x + hi.x) * 0.5f, lo.y, (lo.z + hi.z) * 0.5f);
    const float c = cosf(o->rotation_y), s = sinf(o->rotation_y);
    const float scale = o->uniform_scale / largest;
    const vec3 t(o->translation[0], o->translation[1], o->translation[2]);
    std::vector<vec3> xp(pos.size());
    for (size_t i = 0; i < pos.size(); i++) {
        const vec3 q = (pos[i] - bottom) * scale;
        xp[i] = vec3(q.x * c + q.z * s, q.y, -q.x * s + q.z * c) + t;
    }
    // area-weighted vertex normals when the model has no vn (:2752-2766)
    const int np = (int)pos.size();
    std::vector<vec3> avg;
    if (nrm.empty()) {
        avg.assign(pos.size(), vec3(0.f, 0.f, 0.f));
        for (const Face& fc : faces) {
            const int i0 = fc.c[0].p, i1 = fc.c[1].p, i2 = fc.c[2].p;
            if (i0 < 0 || i1 < 0 || i2 < 0 || i0 >= np || i1 >= np || i2 >= np) continue;
            const vec3 a = (xp[i1] - xp[i0]).cross(xp[i2] - xp[i0]);
            avg[i0] = avg[i0] + a;
            avg[i1] = avg[i1] + a;
            avg[i2] = avg[i2] + a;
        }
    }
    // materials: 0 = fallback, 1 + i = library material i
    out.mats.clear();
    out.mats.push_back(o->fallback);
    for (const mcpt_material& m : lib.mats) out.mats.push_back(m);
    std::memset(&out.stats, 0, sizeof(out.stats));
    out.hot.clear();
    out.cold.clear();
    for (const Face& fc : faces) {
        bool ok = true;
        for (int k = 0; k < 3; k++) ok = ok && fc.c[k].p >= 0 && fc.c[k].p < np;
        if (!ok) {
            out.stats.skipped_faces++;
            continue;
        }
        const vec3 v0 = xp[fc.c[0].p], v1 = xp[fc.c[1].p], v2 = xp[fc.c[2].p];
        if ((v1 - v0).cross(v2 - v0).length() < 1e-9f) {
            out.stats.skipped_faces++;
            continue;
        }
        mcpt_tri_hot h = {{v0.x, v0.y, v0.z}, {v1.x, v1.y, v1.z}, {v2.x, v2.y, v2.z}};
        mcpt_tri_cold cd;
        std::memset(&cd, 0, sizeof(cd));
        cd.material = (fc.material >= 0 && fc.material < (int)lib.mats.size()) ? 1 + fc.material : 0;
        float* uv[3] = {cd.uv0, cd.uv1, cd.uv2};
        float* nn[3] = {cd.n0, cd.n1, cd.n2};
        for (int k = 0; k < 3; k++) {
            const int ti = fc.c[k].t, ni = fc.c[k].n;
            if (ti >= 0 && ti < (int)uvs.size()) {
                uv[k][0] = uvs[ti].x;
                uv[k][1] = uvs[ti].y;
            }
            vec3 n(0.f, 0.f, 0.f);
            if (ni >= 0 && ni < (int)nrm.size() && nrm[ni].length_squared() > 0.0f) {
                const vec3 r = nrm[ni];
                n = vec3(r.x * c + r.z * s, r.y, -r.x * s + r.z * c).normalize();
            } else if (!avg.empty()) {
                const vec3& a = avg[fc.c[k].p];
                if (a.length_squared() > 0.0f) n = a.normalize();
            }
            nn[k][0] = n.x;
            nn[k][1] = n.y;
            nn[k][2] = n.z;
        }
        out.hot.push_back(h);
        out.cold.push_back(cd);
    }
    out.stats.n_tris = (int32_t)out.hot.size();
    out.stats.n_materials = (int32_t)out.mats.size();
    out.stats.texture_requests = lib.texture_requests;
    out.stats.missing_libraries = missing_libs;
    out.stats.textures_loaded = (int32_t)lib.images->size();
    out.stats.missing_textures = lib.missing_textures;
    out.images = lib.images;
    if (out.hot.empty()) return set_error(MCPT_ERR_INVALID, "OBJ file '%s': every face was skipped", path);
    return MCPT_OK;
}

}  // namespace

// The two-call pattern would parse a large file twice: a sizing call (no
// arrays) keeps its result for this thread's next call on the same file
// (path, options, size and modification time unchanged).
namespace {
struct Cached {
    bool valid = false;
    std::string path;
    mcpt_obj_opts opts;
    long long size = 0, mtime = 0;
    Loaded data;
};
thread_local Cached g_cache;
// the images of this thread's last successful load, for obj_texture
thread_local std::shared_ptr<std::vector<Image>> g_images;

bool file_id(const char* path, long long& size, long long& mtime) {
    struct stat sb;
    if (stat(path, &sb) != 0) return false;
    size = (long long)sb.st_size;
    mtime = (long long)sb.st_mtim.tv_sec * 1000000000LL + sb.st_mtim.tv_nsec;
    return true;
}
}  // namespace

int load_obj(const char* path, const mcpt_obj_opts* opts, mcpt_tri_hot* hot, mcpt_tri_cold* cold,
             mcpt_material* materials, mcpt_obj_stats* stats) {
    if (!path || !opts) return set_error(MCPT_ERR_INVALID, "path and opts are required");
    if (!(opts->uniform_scale > 0.f)) return set_error(MCPT_ERR_INVALID, "uniform_scale must be positive");
    long long size = -1, mtime = -1;
    const bool have_id = file_id(path, size, mtime);
    Cached& c = g_cache;
    const bool hit = c.valid && have_id && c.path == path && c.size == size && c.mtime == mtime &&
                     std::memcmp(&c.opts, opts, sizeof(mcpt_obj_opts)) == 0;
    if (!hit) {
        c.valid = false;
        MCPT_TRY(load(path, opts, c.data));
        c.path = path;
        c.opts = *opts;
        c.size = size;
        c.mtime = mtime;
        c.valid = have_id;
    }
    const Loaded& l = c.data;
    g_images = l.images;
    if (stats) *stats = l.stats;
    if (hot) std::memcpy(hot, l.hot.data(), l.hot.size() * sizeof(mcpt_tri_hot));
    if (cold) std::memcpy(cold, l.cold.data(), l.cold.size() * sizeof(mcpt_tri_cold));
    if (materials) std::memcpy(materials, l.mats.data(), l.mats.size() * sizeof(mcpt_material));
    if (hot || cold || materials) {  // the arrays are filled: free the copy
        c = Cached();
    }
    return MCPT_OK;
}

int obj_texture(int index, int32_t* width, int32_t* height, uint8_t* rgba8) {
    if (!width || !height) return set_error(MCPT_ERR_INVALID, "bad arguments");
    const int n = g_images ? (int)g_images->size() : 0;
    if (index < 0 || index >= n)
        return set_error(MCPT_ERR_INVALID, "texture %d: the last OBJ load on this thread holds %d", index, n);
    const Image& im = (*g_images)[index];
    *width = im.w;
    *height = im.h;
    if (rgba8) std::memcpy(rgba8, im.rgba.data(), im.rgba.size());
    return MCPT_OK;
}

}  // namespace mcpt
