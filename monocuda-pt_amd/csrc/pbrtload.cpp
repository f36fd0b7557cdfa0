// pbrtload.cpp — pbrt-v3 scene import (loadPbrtScene, src/path_tracer.cu
// :1990-2483), host code.
//
// The reference parses pbrt files with minipbrt (github.com/vilya/minipbrt
// at a6f9dad, fetched by its CMakeLists.txt:33-36) and flattens the result
// into its scene. minipbrt is not part of this image, so the parse here is
// written from the pbrt-v3 file format (pbrt.org/fileformat-v3):
//   * tokens: quoted strings, numbers, bare words, '[' ']', '#' comments;
//     parameters are "type name" followed by one value or a bracketed list;
//   * transforms: Identity, Translate, Scale, Rotate, LookAt, Transform,
//     ConcatTransform (column-major 4x4), CoordinateSystem,
//     CoordSysTransform; AttributeBegin/End and TransformBegin/End stacks;
//   * Camera (cameraToWorld = inverse of the transform at the directive),
//     Film resolution, Material / MakeNamedMaterial / NamedMaterial,
//     Texture, AreaLightSource, LightSource, MakeNamedMedium /
//     MediumInterface, ObjectBegin / ObjectEnd / ObjectInstance,
//     Include / Import; other directives are read and ignored;
//   * shapes: trianglemesh, plymesh (ASCII and binary PLY) and sphere;
//     other shapes are counted as skipped, as the reference skips them.
// What becomes of the parse follows the reference's mapping:
//   * materials (pbrtMaterial, :2124-2171): matte, plastic, substrate, uber
//     and disney become Lambertian with their diffuse colour, glass
//     dielectric, mirror and metal conductors (metal's colour is the
//     normal-incidence reflectance of its complex index, :2111-2122), mix
//     its first material, anything else grey 0.5 Lambertian; pbrt colours
//     are linear;
//   * colour textures (pbrtColorTexture, :2089-2109): an image map read as a
//     PPM file (other formats are not decoded here), constant, and scale
//     (its first texture);
//   * shapes (pbrtAddShape, :2243-2288): an area light makes the shape
//     emissive, an inside medium makes it an isotropic medium
//     (pbrtMediumMaterial, :2173-2185), a material of type "none" skips it;
//     triangles closer than 1e-24 in squared cross product are skipped;
//     file normals are transformed and normalised; a sphere's radius is
//     scaled by the mean axis length of its transform, and an emissive
//     sphere is tessellated into 32 x 16 quads (pbrtAddTessellatedSphere,
//     :2225-2252) so that it joins the light list;
//   * instances: every shape of the object with instanceToWorld x its own
//     transform (pbrt-v3's ObjectInstance);
//   * lights (:2380-2437): an infinite light gives the background colour
//     or the environment map path with its rotation; distant, point and spot
//     lights become emissive quads and boxes at the reference's sizes;
//   * the camera (:2443-2461): look-from, look-at and up from
//     cameraToWorld, the vertical field of view from pbrt's shorter-side one.
// Defaults are pbrt-v3's where the file gives no value.
#include <sys/stat.h>

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "device.hpp"
#include "host_common.hpp"

namespace mcpt {

namespace {

constexpr float kPi = 3.14159265358979323846f;

// ------------------------------------------------------------- matrices
struct M4 {
    float m[4][4];
};
M4 identity() {
    M4 r;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) r.m[i][j] = i == j ? 1.0f : 0.0f;
    return r;
}
// pbrtMultiply (:2056-2062)
M4 mul(const M4& a, const M4& b) {
    M4 r;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            r.m[i][j] = a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j] + a.m[i][2] * b.m[2][j] + a.m[i][3] * b.m[3][j];
    return r;
}
// Gauss-Jordan with partial pivoting, in double; false when singular
bool inverse(const M4& a, M4& out) {
    double w[4][8];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 8; j++) w[i][j] = j < 4 ? a.m[i][j] : (j - 4 == i ? 1.0 : 0.0);
    for (int c = 0; c < 4; c++) {
        int p = c;
        for (int r = c + 1; r < 4; r++)
            if (std::fabs(w[r][c]) > std::fabs(w[p][c])) p = r;
        if (std::fabs(w[p][c]) < 1e-30) return false;
        for (int j = 0; j < 8; j++) std::swap(w[c][j], w[p][j]);
        const double inv = 1.0 / w[c][c];
        for (int j = 0; j < 8; j++) w[c][j] *= inv;
        for (int r = 0; r < 4; r++)
            if (r != c) {
                const double f = w[r][c];
                for (int j = 0; j < 8; j++) w[r][j] -= f * w[c][j];
            }
    }
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) out.m[i][j] = (float)w[i][j + 4];
    return true;
}
// pbrtTransformPoint / pbrtTransformVector (:2064-2074)
vec3 xpoint(const M4& t, const vec3& p) {
    return vec3(t.m[0][0] * p.x + t.m[0][1] * p.y + t.m[0][2] * p.z + t.m[0][3],
                t.m[1][0] * p.x + t.m[1][1] * p.y + t.m[1][2] * p.z + t.m[1][3],
                t.m[2][0] * p.x + t.m[2][1] * p.y + t.m[2][2] * p.z + t.m[2][3]);
}
vec3 xvector(const M4& t, const vec3& v) {
    return vec3(t.m[0][0] * v.x + t.m[0][1] * v.y + t.m[0][2] * v.z, t.m[1][0] * v.x + t.m[1][1] * v.y + t.m[1][2] * v.z,
                t.m[2][0] * v.x + t.m[2][1] * v.y + t.m[2][2] * v.z);
}
// pbrt-v3 Translate / Scale / Rotate / LookAt (core/transform.cpp)
M4 translate(float x, float y, float z) {
    M4 r = identity();
    r.m[0][3] = x;
    r.m[1][3] = y;
    r.m[2][3] = z;
    return r;
}
M4 scale(float x, float y, float z) {
    M4 r = identity();
    r.m[0][0] = x;
    r.m[1][1] = y;
    r.m[2][2] = z;
    return r;
}
M4 rotate(float deg, vec3 a) {
    a = a.normalize();
    const float th = deg * kPi / 180.0f, s = sinf(th), c = cosf(th);
    M4 r = identity();
    r.m[0][0] = a.x * a.x + (1 - a.x * a.x) * c;
    r.m[0][1] = a.x * a.y * (1 - c) - a.z * s;
    r.m[0][2] = a.x * a.z * (1 - c) + a.y * s;
    r.m[1][0] = a.x * a.y * (1 - c) + a.z * s;
    r.m[1][1] = a.y * a.y + (1 - a.y * a.y) * c;
    r.m[1][2] = a.y * a.z * (1 - c) - a.x * s;
    r.m[2][0] = a.x * a.z * (1 - c) - a.y * s;
    r.m[2][1] = a.y * a.z * (1 - c) + a.x * s;
    r.m[2][2] = a.z * a.z + (1 - a.z * a.z) * c;
    return r;
}
bool look_at(const vec3& pos, const vec3& look, const vec3& up, M4& world_to_camera) {
    M4 c2w = identity();
    const vec3 dir = (look - pos).normalize();
    const vec3 upn = up.normalize();
    const vec3 right0 = upn.cross(dir);
    if (right0.length() == 0.0f) return false;
    const vec3 right = right0.normalize();
    const vec3 nu = dir.cross(right);
    const float cols[4][3] = {{right.x, right.y, right.z}, {nu.x, nu.y, nu.z}, {dir.x, dir.y, dir.z},
                              {pos.x, pos.y, pos.z}};
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < 3; r++) c2w.m[r][c] = cols[c][r];
    return inverse(c2w, world_to_camera);
}

// --------------------------------------------------------------- tokens
struct Tok {
    enum Kind { Word, Str, Num, Open, Close } kind;
    std::string text;
    double num = 0.0;
};

bool read_file(const std::string& path, std::string& out) {
    std::FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    char buf[1 << 16];
    size_t got;
    out.clear();
    while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0) out.append(buf, got);
    std::fclose(f);
    return true;
}

bool tokenize(const std::string& s, std::vector<Tok>& out, std::string& err) {
    size_t i = 0;
    while (i < s.size()) {
        const char c = s[i];
        if (std::isspace((unsigned char)c)) {
            i++;
        } else if (c == '#') {
            while (i < s.size() && s[i] != '\n') i++;
        } else if (c == '[' || c == ']') {
            out.push_back(Tok{c == '[' ? Tok::Open : Tok::Close, std::string(1, c)});
            i++;
        } else if (c == '"') {
            const size_t e = s.find('"', i + 1);
            if (e == std::string::npos) {
                err = "unterminated string";
                return false;
            }
            out.push_back(Tok{Tok::Str, s.substr(i + 1, e - i - 1)});
            i = e + 1;
        } else {
            size_t e = i;
            while (e < s.size() && !std::isspace((unsigned char)s[e]) && s[e] != '[' && s[e] != ']' && s[e] != '"' &&
                   s[e] != '#')
                e++;
            Tok t{Tok::Word, s.substr(i, e - i)};
            char* end = nullptr;
            const double v = std::strtod(t.text.c_str(), &end);
            if (end && *end == '\0' && !t.text.empty()) {
                t.kind = Tok::Num;
                t.num = v;
            }
            out.push_back(t);
            i = e;
        }
    }
    return true;
}

// ----------------------------------------------------------- parameters
struct Param {
    std::string type, name;
    std::vector<double> nums;
    std::vector<std::string> strs;
    bool boolean = false;
};

struct Params {
    std::vector<Param> ps;
    const Param* find(const char* name) const {
        for (const Param& p : ps)
            if (p.name == name) return &p;
        return nullptr;
    }
    float f(const char* name, float def) const {
        const Param* p = find(name);
        return p && !p->nums.empty() ? (float)p->nums[0] : def;
    }
    int i(const char* name, int def) const { return (int)f(name, (float)def); }
    std::string s(const char* name) const {
        const Param* p = find(name);
        return p && !p->strs.empty() ? p->strs[0] : std::string();
    }
    // an RGB-like value: rgb / color / spectrum with 3 values, or one float
    // for all channels; `tex` receives the texture name of a texture param
    bool rgb(const char* name, float out[3], std::string* tex = nullptr) const {
        const Param* p = find(name);
        if (!p) return false;
        if (p->type == "texture") {
            if (tex && !p->strs.empty()) *tex = p->strs[0];
            return false;
        }
        if (p->nums.size() >= 3 && (p->type == "rgb" || p->type == "color" || p->type == "spectrum")) {
            for (int k = 0; k < 3; k++) out[k] = (float)p->nums[k];
            return true;
        }
        if (p->nums.size() == 1) {
            out[0] = out[1] = out[2] = (float)p->nums[0];
            return true;
        }
        return false;
    }
};

struct Texture {
    std::string cls;  // imagemap, constant, scale, ...
    Params params;
};

struct MatDesc {
    std::string type;
    Params params;
};

struct Medium {
    float sigma_a[3] = {0.0011f, 0.0024f, 0.014f};  // pbrt-v3 homogeneous defaults
    float sigma_s[3] = {2.55f, 3.21f, 3.77f};
    float scale = 1.0f;
    float g = 0.0f;
};

struct AreaLight {
    float L[3] = {1.0f, 1.0f, 1.0f};
    float scale[3] = {1.0f, 1.0f, 1.0f};
};

struct ShapeDesc {
    std::string type;
    Params params;
    M4 to_world;
    std::shared_ptr<MatDesc> mat;
    int area = -1;          // index into area lights
    std::string medium;     // inside medium name
    std::string dir;        // folder of the file that named it (plymesh paths)
};

struct ObjectDef {
    std::vector<ShapeDesc> shapes;
};

struct LightDesc {
    std::string type;
    Params params;
    M4 to_world;
    std::string dir;
};

struct GState {
    M4 ctm = identity();
    std::shared_ptr<MatDesc> mat;
    int area = -1;
    std::string medium;
};

struct ImageTex {
    int w = 0, h = 0;
    std::vector<uint8_t> rgba;
};

struct Parsed {
    std::vector<ShapeDesc> shapes;
    std::map<std::string, ObjectDef> objects;
    std::vector<std::pair<std::string, std::pair<M4, int>>> instances;  // object, (toWorld, area light)
    std::vector<LightDesc> lights;
    std::vector<AreaLight> areas;
    std::map<std::string, Medium> media;
    std::map<std::string, Texture> textures;
    bool has_camera = false;
    M4 camera_to_world = identity();
    float fov = 90.0f;
    int width = 0, height = 0;
};

struct Parser {
    Parsed& out;
    std::string err;
    GState gs;
    std::vector<GState> attr_stack;
    std::vector<M4> xform_stack;
    std::map<std::string, M4> coord_sys;
    std::map<std::string, std::shared_ptr<MatDesc>> named;
    std::string object;  // inside ObjectBegin
    int depth = 0;       // Include nesting

    explicit Parser(Parsed& o) : out(o) { gs.mat = std::make_shared<MatDesc>(MatDesc{"matte", Params()}); }

    bool fail(const std::string& m) {
        if (err.empty()) err = m;
        return false;
    }

    // "type name" value-or-list pairs after a directive's fixed arguments
    bool params(const std::vector<Tok>& t, size_t& i, Params& p) {
        while (i < t.size() && t[i].kind == Tok::Str) {
            const std::string decl = t[i].text;
            const size_t sp = decl.find_first_of(" \t");
            const size_t nm = sp == std::string::npos ? sp : decl.find_first_not_of(" \t", sp);
            if (nm == std::string::npos) return true;  // a plain string: the next directive's argument
            Param q;
            q.type = decl.substr(0, sp);
            q.name = decl.substr(nm);
            i++;
            auto take = [&](const Tok& v) {
                if (v.kind == Tok::Num) q.nums.push_back(v.num);
                else if (v.kind == Tok::Str) q.strs.push_back(v.text);
                else if (v.text == "true" || v.text == "false") q.boolean = v.text == "true";
            };
            if (i < t.size() && t[i].kind == Tok::Open) {
                i++;
                while (i < t.size() && t[i].kind != Tok::Close) take(t[i++]);
                if (i >= t.size()) return fail("unclosed '['");
                i++;
            } else if (i < t.size()) {
                take(t[i++]);
            }
            if (q.type == "bool" && !q.strs.empty()) q.boolean = q.strs[0] == "true";
            p.ps.push_back(q);
        }
        return true;
    }

    bool nums(const std::vector<Tok>& t, size_t& i, int n, float* v) {
        bool bracket = i < t.size() && t[i].kind == Tok::Open;
        if (bracket) i++;
        for (int k = 0; k < n; k++) {
            if (i >= t.size() || t[i].kind != Tok::Num) return fail("expected a number");
            v[k] = (float)t[i++].num;
        }
        if (bracket) {
            if (i >= t.size() || t[i].kind != Tok::Close) return fail("expected ']'");
            i++;
        }
        return true;
    }

    bool str(const std::vector<Tok>& t, size_t& i, std::string& s) {
        if (i >= t.size() || t[i].kind != Tok::Str) return fail("expected a string");
        s = t[i++].text;
        return true;
    }

    bool run_file(const std::string& path) {
        if (depth > 16) return fail("Include nesting too deep");
        std::string text;
        if (!read_file(path, text)) return fail("cannot open '" + path + "'");
        std::vector<Tok> t;
        if (!tokenize(text, t, err)) return false;
        const size_t s = path.find_last_of('/');
        const std::string dir = s == std::string::npos ? "." : path.substr(0, s);
        depth++;
        const bool ok = run(t, dir);
        depth--;
        return ok;
    }

    // pbrtResolvePath (:2076-2087)
    static std::string resolve(const std::string& dir, const std::string& name) {
        if (name.empty() || name[0] == '/') return name;
        if (std::FILE* f = std::fopen(name.c_str(), "rb")) {
            std::fclose(f);
            return name;
        }
        return dir + "/" + name;
    }

    bool run(const std::vector<Tok>& t, const std::string& dir) {
        size_t i = 0;
        while (i < t.size()) {
            if (t[i].kind != Tok::Word) return fail("expected a directive, got '" + t[i].text + "'");
            const std::string d = t[i++].text;
            float v[16];
            std::string a, b;
            Params p;
            if (d == "Identity") {
                gs.ctm = identity();
            } else if (d == "Translate") {
                if (!nums(t, i, 3, v)) return false;
                gs.ctm = mul(gs.ctm, translate(v[0], v[1], v[2]));
            } else if (d == "Scale") {
                if (!nums(t, i, 3, v)) return false;
                gs.ctm = mul(gs.ctm, scale(v[0], v[1], v[2]));
            } else if (d == "Rotate") {
                if (!nums(t, i, 4, v)) return false;
                gs.ctm = mul(gs.ctm, rotate(v[0], vec3(v[1], v[2], v[3])));
            } else if (d == "LookAt") {
                if (!nums(t, i, 9, v)) return false;
                M4 w2c;
                if (!look_at(vec3(v[0], v[1], v[2]), vec3(v[3], v[4], v[5]), vec3(v[6], v[7], v[8]), w2c))
                    return fail("LookAt: up is parallel to the view direction");
                gs.ctm = mul(gs.ctm, w2c);
            } else if (d == "Transform" || d == "ConcatTransform") {
                if (!nums(t, i, 16, v)) return false;
                M4 m;  // the file's 16 numbers are column-major
                for (int r = 0; r < 4; r++)
                    for (int c = 0; c < 4; c++) m.m[r][c] = v[c * 4 + r];
                gs.ctm = d == "Transform" ? m : mul(gs.ctm, m);
            } else if (d == "CoordinateSystem") {
                if (!str(t, i, a)) return false;
                coord_sys[a] = gs.ctm;
            } else if (d == "CoordSysTransform") {
                if (!str(t, i, a)) return false;
                auto it = coord_sys.find(a);
                if (it != coord_sys.end()) gs.ctm = it->second;
            } else if (d == "AttributeBegin") {
                attr_stack.push_back(gs);
            } else if (d == "AttributeEnd") {
                if (attr_stack.empty()) return fail("AttributeEnd without AttributeBegin");
                gs = attr_stack.back();
                attr_stack.pop_back();
            } else if (d == "TransformBegin") {
                xform_stack.push_back(gs.ctm);
            } else if (d == "TransformEnd") {
                if (xform_stack.empty()) return fail("TransformEnd without TransformBegin");
                gs.ctm = xform_stack.back();
                xform_stack.pop_back();
            } else if (d == "Camera") {
                if (!str(t, i, a) || !params(t, i, p)) return false;
                M4 c2w;
                if (!inverse(gs.ctm, c2w)) return fail("Camera: singular transform");
                out.has_camera = true;
                out.camera_to_world = c2w;
                out.fov = p.f("fov", 90.0f);
                coord_sys["camera"] = c2w;
            } else if (d == "Film") {
                if (!str(t, i, a) || !params(t, i, p)) return false;
                out.width = p.i("xresolution", 1280);
                out.height = p.i("yresolution", 720);
            } else if (d == "WorldBegin") {
                gs.ctm = identity();
                coord_sys["world"] = gs.ctm;
            } else if (d == "WorldEnd") {
                // nothing to flush
            } else if (d == "Material") {
                if (!str(t, i, a) || !params(t, i, p)) return false;
                gs.mat = std::make_shared<MatDesc>(MatDesc{a, p});
            } else if (d == "MakeNamedMaterial") {
                if (!str(t, i, a) || !params(t, i, p)) return false;
                named[a] = std::make_shared<MatDesc>(MatDesc{p.s("type"), p});
            } else if (d == "NamedMaterial") {
                if (!str(t, i, a)) return false;
                auto it = named.find(a);
                gs.mat = it != named.end() ? it->second : nullptr;
            } else if (d == "Texture") {
                std::string kind;
                if (!str(t, i, a) || !str(t, i, kind) || !str(t, i, b) || !params(t, i, p)) return false;
                Texture tx{b, p};
                if (b == "imagemap") {  // resolved now, relative to this file
                    for (Param& q : tx.params.ps)
                        if (q.name == "filename" && !q.strs.empty()) q.strs[0] = resolve(dir, q.strs[0]);
                }
                out.textures[a] = tx;
            } else if (d == "AreaLightSource") {
                if (!str(t, i, a) || !params(t, i, p)) return false;
                AreaLight al;
                p.rgb("L", al.L);
                p.rgb("scale", al.scale);
                out.areas.push_back(al);
                gs.area = (int)out.areas.size() - 1;
            } else if (d == "LightSource") {
                if (!str(t, i, a) || !params(t, i, p)) return false;
                out.lights.push_back(LightDesc{a, p, gs.ctm, dir});
            } else if (d == "MakeNamedMedium") {
                if (!str(t, i, a) || !params(t, i, p)) return false;
                Medium m;
                p.rgb("sigma_a", m.sigma_a);
                p.rgb("sigma_s", m.sigma_s);
                m.scale = p.f("scale", 1.0f);
                m.g = p.f("g", 0.0f);
                if (p.s("type") == "homogeneous" || p.s("type").empty()) out.media[a] = m;
            } else if (d == "MediumInterface") {
                if (!str(t, i, a)) return false;
                if (i < t.size() && t[i].kind == Tok::Str && t[i].text.find(' ') == std::string::npos) b = t[i++].text;
                gs.medium = a;
            } else if (d == "Shape") {
                if (!str(t, i, a) || !params(t, i, p)) return false;
                ShapeDesc sd{a, p, gs.ctm, gs.mat, gs.area, gs.medium, dir};
                if (!object.empty()) out.objects[object].shapes.push_back(sd);
                else out.shapes.push_back(sd);
            } else if (d == "ObjectBegin") {
                if (!str(t, i, a)) return false;
                attr_stack.push_back(gs);
                object = a;
                out.objects[a];
            } else if (d == "ObjectEnd") {
                object.clear();
                if (!attr_stack.empty()) {
                    gs = attr_stack.back();
                    attr_stack.pop_back();
                }
            } else if (d == "ObjectInstance") {
                if (!str(t, i, a)) return false;
                out.instances.push_back({a, {gs.ctm, gs.area}});
            } else if (d == "Include" || d == "Import") {
                if (!str(t, i, a)) return false;
                if (!run_file(resolve(dir, a))) return false;
            } else if (d == "ReverseOrientation") {
                // no state that the renderer uses
            } else if (d == "ActiveTransform") {
                if (i < t.size()) i++;
            } else if (d == "TransformTimes") {
                if (!nums(t, i, 2, v)) return false;
            } else {
                // Sampler, Integrator, PixelFilter, Accelerator, ...: a type
                // name and parameters that the renderer has no use for
                if (i < t.size() && t[i].kind == Tok::Str) i++;
                if (!params(t, i, p)) return false;
            }
        }
        return true;
    }
};

// ------------------------------------------------------------ PLY files
struct PlyMesh {
    std::vector<float> P, N, uv;
    std::vector<int> idx;
};

bool read_ply(const std::string& path, PlyMesh& m, std::string& err) {
    std::string data;
    if (!read_file(path, data)) {
        err = "cannot open PLY file '" + path + "'";
        return false;
    }
    size_t hdr_end = data.find("end_header");
    if (data.compare(0, 3, "ply") != 0 || hdr_end == std::string::npos) {
        err = "'" + path + "' is not a PLY file";
        return false;
    }
    size_t body = data.find('\n', hdr_end);
    body = body == std::string::npos ? data.size() : body + 1;
    struct Prop {
        std::string name, type, count_type;
        bool list = false;
    };
    struct Elem {
        std::string name;
        long long count = 0;
        std::vector<Prop> props;
    };
    std::vector<Elem> elems;
    int format = -1;  // 0 ascii, 1 little endian, 2 big endian
    {
        size_t at = 0;
        while (at < hdr_end) {
            size_t nl = data.find('\n', at);
            if (nl == std::string::npos || nl > hdr_end) nl = hdr_end;
            std::string line = data.substr(at, nl - at);
            at = nl + 1;
            char w0[64] = {0}, w1[64] = {0}, w2[64] = {0}, w3[64] = {0}, w4[64] = {0};
            const int n = std::sscanf(line.c_str(), "%63s %63s %63s %63s %63s", w0, w1, w2, w3, w4);
            if (n < 1) continue;
            const std::string k = w0;
            if (k == "format") {
                const std::string f = w1;
                format = f == "ascii" ? 0 : f == "binary_little_endian" ? 1 : f == "binary_big_endian" ? 2 : -1;
            } else if (k == "element" && n >= 3) {
                elems.push_back(Elem{w1, std::atoll(w2), {}});
            } else if (k == "property" && !elems.empty()) {
                if (std::string(w1) == "list" && n >= 5) elems.back().props.push_back(Prop{w4, w3, w2, true});
                else if (n >= 3) elems.back().props.push_back(Prop{w2, w1, "", false});
            }
        }
    }
    if (format < 0) {
        err = "'" + path + "': unknown PLY format";
        return false;
    }
    auto size_of = [](const std::string& t) -> int {
        if (t == "char" || t == "uchar" || t == "int8" || t == "uint8") return 1;
        if (t == "short" || t == "ushort" || t == "int16" || t == "uint16") return 2;
        if (t == "int" || t == "uint" || t == "float" || t == "int32" || t == "uint32" || t == "float32") return 4;
        if (t == "double" || t == "float64") return 8;
        return 0;
    };
    size_t at = body;
    std::vector<std::string> words;  // ASCII: the body split once
    size_t wpos = 0;
    if (format == 0) {
        size_t k = body;
        while (k < data.size()) {
            while (k < data.size() && std::isspace((unsigned char)data[k])) k++;
            size_t e = k;
            while (e < data.size() && !std::isspace((unsigned char)data[e])) e++;
            if (e > k) words.push_back(data.substr(k, e - k));
            k = e;
        }
    }
    auto read_val = [&](const std::string& type, double& out) -> bool {
        if (format == 0) {
            if (wpos >= words.size()) return false;
            out = std::strtod(words[wpos++].c_str(), nullptr);
            return true;
        }
        const int sz = size_of(type);
        if (sz == 0 || at + sz > data.size()) return false;
        unsigned char b[8];
        std::memcpy(b, &data[at], sz);
        at += sz;
        if (format == 2)
            for (int k = 0; k < sz / 2; k++) std::swap(b[k], b[sz - 1 - k]);
        if (type == "char" || type == "int8") out = (int8_t)b[0];
        else if (type == "uchar" || type == "uint8") out = b[0];
        else if (type == "short" || type == "int16") { int16_t x; std::memcpy(&x, b, 2); out = x; }
        else if (type == "ushort" || type == "uint16") { uint16_t x; std::memcpy(&x, b, 2); out = x; }
        else if (type == "int" || type == "int32") { int32_t x; std::memcpy(&x, b, 4); out = x; }
        else if (type == "uint" || type == "uint32") { uint32_t x; std::memcpy(&x, b, 4); out = x; }
        else if (type == "float" || type == "float32") { float x; std::memcpy(&x, b, 4); out = x; }
        else { double x; std::memcpy(&x, b, 8); out = x; }
        return true;
    };
    for (const Elem& e : elems) {
        if (e.count < 0 || e.count > (1LL << 30)) {
            err = "'" + path + "': bad element count";
            return false;
        }
        const bool vert = e.name == "vertex", face = e.name == "face";
        int ix = -1, iy = -1, iz = -1, inx = -1, iny = -1, inz = -1, iu = -1, iv = -1, ilist = -1;
        for (int k = 0; k < (int)e.props.size(); k++) {
            const std::string& n = e.props[k].name;
            if (n == "x") ix = k;
            else if (n == "y") iy = k;
            else if (n == "z") iz = k;
            else if (n == "nx") inx = k;
            else if (n == "ny") iny = k;
            else if (n == "nz") inz = k;
            else if (n == "u" || n == "s" || n == "texture_u" || n == "texture_s") iu = k;
            else if (n == "v" || n == "t" || n == "texture_v" || n == "texture_t") iv = k;
            else if (e.props[k].list && (n == "vertex_indices" || n == "vertex_index")) ilist = k;
        }
        const bool has_n = inx >= 0 && iny >= 0 && inz >= 0, has_uv = iu >= 0 && iv >= 0;
        std::vector<double> vals(e.props.size());
        std::vector<int> poly;
        for (long long r = 0; r < e.count; r++) {
            poly.clear();
            for (int k = 0; k < (int)e.props.size(); k++) {
                const Prop& pr = e.props[k];
                if (pr.list) {
                    double cnt;
                    if (!read_val(pr.count_type, cnt) || cnt < 0 || cnt > 1 << 20) {
                        err = "'" + path + "': truncated or bad list";
                        return false;
                    }
                    for (int j = 0; j < (int)cnt; j++) {
                        double x;
                        if (!read_val(pr.type, x)) {
                            err = "'" + path + "': truncated list";
                            return false;
                        }
                        if (k == ilist) poly.push_back((int)x);
                    }
                } else if (!read_val(pr.type, vals[k])) {
                    err = "'" + path + "': truncated data";
                    return false;
                }
            }
            if (vert) {
                if (ix < 0 || iy < 0 || iz < 0) {
                    err = "'" + path + "': vertices without x y z";
                    return false;
                }
                m.P.insert(m.P.end(), {(float)vals[ix], (float)vals[iy], (float)vals[iz]});
                if (has_n) m.N.insert(m.N.end(), {(float)vals[inx], (float)vals[iny], (float)vals[inz]});
                if (has_uv) m.uv.insert(m.uv.end(), {(float)vals[iu], (float)vals[iv]});
            } else if (face) {
                for (size_t j = 1; j + 1 < poly.size(); j++) m.idx.insert(m.idx.end(), {poly[0], poly[j], poly[j + 1]});
            }
        }
    }
    return true;
}

// --------------------------------------------------------------- output
struct Loaded {
    std::vector<mcpt_tri_hot> hot;
    std::vector<mcpt_tri_cold> cold;
    std::vector<mcpt_sphere> spheres;
    std::vector<mcpt_material> mats;
    std::shared_ptr<std::vector<ImageTex>> images = std::make_shared<std::vector<ImageTex>>();
    mcpt_pbrt_info info;
};

struct Builder {
    const Parsed& P;
    Loaded& L;
    int64_t budget;
    std::string err;
    std::map<std::string, int> tex_by_path;  // image path -> image index, -1 when unreadable
    std::map<std::string, int> mat_index;    // material record bytes -> table index
    std::map<const MatDesc*, mcpt_material> mat_cache;

    int material(const mcpt_material& m) {
        const std::string key((const char*)&m, sizeof(m));
        auto it = mat_index.find(key);
        if (it != mat_index.end()) return it->second;
        L.mats.push_back(m);
        return mat_index[key] = (int)L.mats.size() - 1;
    }

    static mcpt_material lambert(const float c[3]) {
        mcpt_material m;
        std::memset(&m, 0, sizeof(m));
        m.type = MCPT_MAT_LAMBERTIAN;
        for (int k = 0; k < 3; k++) m.albedo[k] = c[k];
        return m;
    }

    // pbrtColorTexture (:2089-2109): the colour, or an image texture
    int color(const Params& p, const char* name, const float def[3], float out[3]) {
        for (int k = 0; k < 3; k++) out[k] = def[k];
        std::string tex;
        p.rgb(name, out, &tex);
        for (int hops = 0; !tex.empty() && hops < 16; hops++) {
            auto it = P.textures.find(tex);
            if (it == P.textures.end()) break;
            const Texture& t = it->second;
            tex.clear();
            if (t.cls == "imagemap") {
                const std::string path = t.params.s("filename");
                auto c = tex_by_path.find(path);
                int idx;
                if (c != tex_by_path.end()) {
                    idx = c->second;
                } else {
                    ImageTex im;
                    idx = -1;
                    if (read_image8(path.c_str(), im.w, im.h, &im.rgba) == MCPT_OK) {
                        idx = (int)L.images->size();
                        L.images->push_back(std::move(im));
                    } else {
                        clear_error();
                    }
                    tex_by_path[path] = idx;
                }
                if (idx >= 0) return idx + 1;
                L.info.unsupported_textures++;
            } else if (t.cls == "constant") {
                t.params.rgb("value", out);
            } else if (t.cls == "scale") {  // its first texture
                t.params.rgb("tex1", out, &tex);
            } else {
                L.info.unsupported_textures++;
            }
        }
        return 0;
    }

    // pbrtMaterial (:2124-2171)
    mcpt_material map_material(const MatDesc* d, int hops = 0) {
        const float grey[3] = {0.5f, 0.5f, 0.5f};
        mcpt_material m = lambert(grey);
        if (!d) return m;
        auto it = mat_cache.find(d);
        if (it != mat_cache.end()) return it->second;
        const std::string& t = d->type;
        const Params& p = d->params;
        float c[3];
        if (t == "matte" || t == "plastic" || t == "substrate" || t == "uber" || t == "disney") {
            const float def[3] = {t == "plastic" || t == "uber" ? 0.25f : 0.5f, t == "plastic" || t == "uber" ? 0.25f : 0.5f,
                                  t == "plastic" || t == "uber" ? 0.25f : 0.5f};
            const int tex = color(p, t == "disney" ? "color" : "Kd", def, c);
            m = lambert(c);
            m.texture = tex;
            if (t != "matte") L.info.approximated_materials++;
        } else if (t == "glass") {
            std::memset(&m, 0, sizeof(m));
            m.type = MCPT_MAT_DIELECTRIC;
            m.ior = p.f("eta", p.f("index", 1.5f));
        } else if (t == "mirror") {
            std::memset(&m, 0, sizeof(m));
            m.type = MCPT_MAT_CONDUCTOR;
            float kr[3] = {0.9f, 0.9f, 0.9f};
            p.rgb("Kr", kr);
            for (int k = 0; k < 3; k++) m.albedo[k] = kr[k];
        } else if (t == "metal") {
            // copper, pbrt-v3's default metal, as RGB
            float eta[3] = {0.2004f, 0.9240f, 1.1022f}, kk[3] = {3.9129f, 2.4528f, 2.1421f};
            p.rgb("eta", eta);
            p.rgb("k", kk);
            std::memset(&m, 0, sizeof(m));
            m.type = MCPT_MAT_CONDUCTOR;
            for (int ch = 0; ch < 3; ch++) {  // pbrtConductorColor (:2111-2122)
                const float em = eta[ch] - 1.0f, ep = eta[ch] + 1.0f, k2 = kk[ch] * kk[ch];
                m.albedo[ch] = (em * em + k2) / (ep * ep + k2);
            }
            const float rough = p.f("uroughness", p.f("roughness", 0.01f));
            m.fuzz = fminf(fmaxf(rough, 0.0f), 1.0f);
        } else if (t == "mix" && hops < 8) {
            L.info.approximated_materials++;
            const std::string first = p.s("namedmaterial1");
            auto f = named_materials->find(first);
            m = f != named_materials->end() ? map_material(f->second.get(), hops + 1) : m;
        } else {
            L.info.approximated_materials++;
        }
        mat_cache[d] = m;
        return m;
    }

    const std::map<std::string, std::shared_ptr<MatDesc>>* named_materials = nullptr;

    bool budget_ok() {
        if ((int64_t)L.hot.size() + (int64_t)L.spheres.size() > budget) {
            err = "the scene expands to more primitives than the budget allows";
            return false;
        }
        return true;
    }

    void add_tri(const vec3& a, const vec3& b, const vec3& c, int mat, const float* uv = nullptr,
                 const vec3* n = nullptr) {
        mcpt_tri_hot h = {{a.x, a.y, a.z}, {b.x, b.y, b.z}, {c.x, c.y, c.z}};
        mcpt_tri_cold cd;
        std::memset(&cd, 0, sizeof(cd));
        cd.material = mat;
        if (uv)
            for (int k = 0; k < 2; k++) {
                cd.uv0[k] = uv[k];
                cd.uv1[k] = uv[2 + k];
                cd.uv2[k] = uv[4 + k];
            }
        if (n) {
            const vec3* nn[3] = {&n[0], &n[1], &n[2]};
            float* dst[3] = {cd.n0, cd.n1, cd.n2};
            for (int k = 0; k < 3; k++) {
                dst[k][0] = nn[k]->x;
                dst[k][1] = nn[k]->y;
                dst[k][2] = nn[k]->z;
            }
        }
        L.hot.push_back(h);
        L.cold.push_back(cd);
    }

    // add_quad (:1677-1691): two triangles with unit-square uvs
    void add_quad(const vec3& a, const vec3& b, const vec3& c, const vec3& d, int mat) {
        const float uv1[6] = {0, 0, 1, 0, 1, 1}, uv2[6] = {0, 0, 1, 1, 0, 1};
        add_tri(a, b, c, mat, uv1);
        add_tri(a, c, d, mat, uv2);
    }

    // add_box (:1693-1708) without rotation
    void add_box(const vec3& center, const vec3& half, int mat) {
        vec3 p[8];
        for (int i = 0; i < 8; i++) {
            const float x = (i & 1) ? half.x : -half.x, y = (i & 2) ? half.y : -half.y, z = (i & 4) ? half.z : -half.z;
            p[i] = center + vec3(x, y, z);
        }
        add_quad(p[0], p[1], p[3], p[2], mat);
        add_quad(p[4], p[5], p[7], p[6], mat);
        add_quad(p[0], p[2], p[6], p[4], mat);
        add_quad(p[1], p[3], p[7], p[5], mat);
        add_quad(p[0], p[1], p[5], p[4], mat);
        add_quad(p[2], p[3], p[7], p[6], mat);
    }

    // pbrtAddShape (:2243-2288) for one placement of a shape
    bool add_shape(const ShapeDesc& s, const M4& to_world, int area_override) {
        if (s.type != "trianglemesh" && s.type != "plymesh" && s.type != "sphere") {
            L.info.skipped_shapes++;
            return true;
        }
        mcpt_material m;
        const int area = area_override >= 0 ? area_override : s.area;
        auto med = s.medium.empty() ? P.media.end() : P.media.find(s.medium);
        if (area >= 0 && area < (int)P.areas.size()) {
            const AreaLight& al = P.areas[area];
            std::memset(&m, 0, sizeof(m));
            m.type = MCPT_MAT_EMISSIVE;
            for (int k = 0; k < 3; k++) m.emission[k] = al.L[k] * al.scale[k];
        } else if (med != P.media.end()) {  // pbrtMediumMaterial (:2173-2185)
            const Medium& md = med->second;
            std::memset(&m, 0, sizeof(m));
            m.type = MCPT_MAT_ISOTROPIC;
            float density = 0.0f;
            for (int ch = 0; ch < 3; ch++) {
                const float ext = (md.sigma_a[ch] + md.sigma_s[ch]) * md.scale;
                m.albedo[ch] = ext > 0.0f ? md.sigma_s[ch] * md.scale / ext : 0.0f;
                density += ext / 3.0f;
            }
            m.density = density;
            m.phase_g = fminf(fmaxf(md.g, -0.99f), 0.99f);
        } else {
            if (s.mat && s.mat->type == "none") {  // an invisible interface
                L.info.skipped_shapes++;
                return true;
            }
            m = map_material(s.mat.get());
        }
        const int mi = material(m);
        if (s.type == "sphere") {
            const vec3 center = xpoint(to_world, vec3(0.0f, 0.0f, 0.0f));
            float axes = 0.0f;
            for (int a = 0; a < 3; a++)
                axes += xvector(to_world, vec3(a == 0 ? 1.0f : 0.0f, a == 1 ? 1.0f : 0.0f, a == 2 ? 1.0f : 0.0f)).length();
            const float radius = s.params.f("radius", 1.0f) * axes / 3.0f;
            if (!(radius > 0.0f)) {
                L.info.skipped_shapes++;
                return true;
            }
            if (m.type == MCPT_MAT_EMISSIVE) {  // pbrtAddTessellatedSphere (:2225-2252)
                const int mer = 32, par = 16;
                for (int pi = 0; pi < par; pi++) {
                    const float t0 = kPi * pi / par, t1 = kPi * (pi + 1) / par;
                    for (int mj = 0; mj < mer; mj++) {
                        const float p0 = 2.f * kPi * mj / mer, p1 = 2.f * kPi * (mj + 1) / mer;
                        const float th[4] = {t0, t1, t1, t0}, ph[4] = {p0, p0, p1, p1};
                        vec3 c[4];
                        for (int k = 0; k < 4; k++)
                            c[k] = center + vec3(sinf(th[k]) * cosf(ph[k]), cosf(th[k]), sinf(th[k]) * sinf(ph[k])) * radius;
                        const int tri[2][3] = {{0, 1, 2}, {0, 2, 3}};
                        for (int q = 0; q < 2; q++) {
                            const vec3 v0 = c[tri[q][0]], v1 = c[tri[q][1]], v2 = c[tri[q][2]];
                            if ((v1 - v0).cross(v2 - v0).length_squared() < 1e-24f) continue;
                            add_tri(v0, v1, v2, mi);
                        }
                    }
                }
                return budget_ok();
            }
            mcpt_sphere sp;
            std::memset(&sp, 0, sizeof(sp));
            sp.center[0] = center.x;
            sp.center[1] = center.y;
            sp.center[2] = center.z;
            sp.radius = radius;
            sp.material = mi;
            L.spheres.push_back(sp);
            return budget_ok();
        }
        // triangle mesh (a plymesh read from its file)
        PlyMesh ply;
        const std::vector<float>* Pp;
        const std::vector<float>* Np;
        const std::vector<float>* UVp;
        std::vector<float> Pv, Nv, UVv;
        std::vector<int> Iv;
        const std::vector<int>* Ip;
        if (s.type == "plymesh") {
            const std::string path = Parser::resolve(s.dir, s.params.s("filename"));
            if (!read_ply(path, ply, err)) return false;
            Pp = &ply.P, Np = &ply.N, UVp = &ply.uv, Ip = &ply.idx;
        } else {
            auto grab = [&](const char* n, std::vector<float>& v) {
                if (const Param* q = s.params.find(n))
                    for (double x : q->nums) v.push_back((float)x);
            };
            grab("P", Pv);
            grab("N", Nv);
            grab("uv", UVv);
            if (UVv.empty()) grab("st", UVv);
            if (const Param* q = s.params.find("indices"))
                for (double x : q->nums) Iv.push_back((int)x);
            if (Iv.empty() && Pv.size() == 9) Iv = {0, 1, 2};
            Pp = &Pv, Np = &Nv, UVp = &UVv, Ip = &Iv;
        }
        const int nv = (int)(Pp->size() / 3);
        const bool has_n = Np->size() == Pp->size(), has_uv = (int)UVp->size() == 2 * nv;
        for (size_t t = 0; t + 2 < Ip->size(); t += 3) {  // pbrtAddMesh (:2187-2223)
            const int ci[3] = {(*Ip)[t], (*Ip)[t + 1], (*Ip)[t + 2]};
            if (ci[0] < 0 || ci[1] < 0 || ci[2] < 0 || ci[0] >= nv || ci[1] >= nv || ci[2] >= nv) continue;
            vec3 c[3], n[3];
            float uv[6];
            for (int k = 0; k < 3; k++) {
                const float* q = &(*Pp)[(size_t)ci[k] * 3];
                c[k] = xpoint(to_world, vec3(q[0], q[1], q[2]));
            }
            if ((c[1] - c[0]).cross(c[2] - c[0]).length_squared() < 1e-24f) continue;
            for (int k = 0; k < 3; k++) {
                if (has_uv) {
                    uv[2 * k] = (*UVp)[(size_t)ci[k] * 2];
                    uv[2 * k + 1] = (*UVp)[(size_t)ci[k] * 2 + 1];
                }
                n[k] = vec3(0.f, 0.f, 0.f);
                if (has_n) {
                    const float* q = &(*Np)[(size_t)ci[k] * 3];
                    const vec3 w = xvector(to_world, vec3(q[0], q[1], q[2]));
                    if (w.length_squared() > 0.f) n[k] = w.normalize();
                }
            }
            add_tri(c[0], c[1], c[2], mi, has_uv ? uv : nullptr, has_n ? n : nullptr);
            if ((int64_t)L.hot.size() > budget) return budget_ok();
        }
        return true;
    }

    bool build() {
        for (const ShapeDesc& s : P.shapes) {
            L.info.shapes++;
            if (!add_shape(s, s.to_world, -1)) return false;
        }
        for (const auto& inst : P.instances) {
            auto it = P.objects.find(inst.first);
            if (it == P.objects.end()) continue;
            L.info.instances++;
            for (const ShapeDesc& s : it->second.shapes)
                if (!add_shape(s, mul(inst.second.first, s.to_world), inst.second.second)) return false;
        }
        // the lights that are not geometry need the extent of what was loaded
        vec3 lo(1e30f, 1e30f, 1e30f), hi(-1e30f, -1e30f, -1e30f);
        for (const mcpt_tri_hot& t : L.hot) {
            const float* vs[3] = {t.v0, t.v1, t.v2};
            for (const float* v : vs) {
                lo = vec3(fminf(lo.x, v[0]), fminf(lo.y, v[1]), fminf(lo.z, v[2]));
                hi = vec3(fmaxf(hi.x, v[0]), fmaxf(hi.y, v[1]), fmaxf(hi.z, v[2]));
            }
        }
        const vec3 center = (lo + hi) * 0.5f;
        const float radius = fmaxf((hi - lo).length() * 0.5f, 1e-3f);
        for (const LightDesc& l : P.lights) {
            float scl[3] = {1.f, 1.f, 1.f};
            l.params.rgb("scale", scl);
            if (l.type == "infinite") {
                float Lr[3] = {1.f, 1.f, 1.f};
                l.params.rgb("L", Lr);
                const std::string map = l.params.s("mapname");
                if (!map.empty()) {
                    const std::string path = Parser::resolve(l.dir, map);
                    std::snprintf(L.info.env_path, sizeof(L.info.env_path), "%s", path.c_str());
                    L.info.has_env_map = 1;
                    L.info.env_intensity = scl[0];
                    const vec3 ay = xvector(l.to_world, vec3(0.f, 1.f, 0.f));
                    const float rot = -0.5f * kPi - atan2f(ay.z, ay.x);
                    L.info.env_rotation_degrees = rot * 180.f / kPi;
                } else {
                    for (int k = 0; k < 3; k++) L.info.background[k] = Lr[k] * scl[k];
                }
            } else if (l.type == "distant") {
                float from[3] = {0.f, 0.f, 0.f}, to[3] = {0.f, 0.f, 1.f}, Lr[3] = {1.f, 1.f, 1.f};
                if (const Param* q = l.params.find("from"))
                    for (int k = 0; k < 3 && k < (int)q->nums.size(); k++) from[k] = (float)q->nums[k];
                if (const Param* q = l.params.find("to"))
                    for (int k = 0; k < 3 && k < (int)q->nums.size(); k++) to[k] = (float)q->nums[k];
                l.params.rgb("L", Lr);
                const vec3 toward = xvector(l.to_world, vec3(from[0] - to[0], from[1] - to[1], from[2] - to[2])).normalize();
                const float tan_half = tanf(1.0f * kPi / 180.f);
                const float dist = radius * 10.f, half = dist * tan_half;
                mcpt_material m;
                std::memset(&m, 0, sizeof(m));
                m.type = MCPT_MAT_EMISSIVE;
                for (int k = 0; k < 3; k++) m.emission[k] = (Lr[k] * scl[k]) / (4.f * tan_half * tan_half);
                const vec3 helper = fabsf(toward.y) < 0.9f ? vec3(0.f, 1.f, 0.f) : vec3(1.f, 0.f, 0.f);
                const vec3 su = helper.cross(toward).normalize() * half;
                const vec3 sv = toward.cross(su).normalize() * half;
                const vec3 c = center + toward * dist;
                add_quad(c - su - sv, c + su - sv, c + su + sv, c - su + sv, material(m));
                L.info.approximated_lights++;
            } else if (l.type == "point" || l.type == "spot") {
                float from[3] = {0.f, 0.f, 0.f}, I[3] = {1.f, 1.f, 1.f};
                if (const Param* q = l.params.find("from"))
                    for (int k = 0; k < 3 && k < (int)q->nums.size(); k++) from[k] = (float)q->nums[k];
                l.params.rgb("I", I);
                const vec3 pos = xpoint(l.to_world, vec3(from[0], from[1], from[2]));
                const float half = radius * 0.005f;
                mcpt_material m;
                std::memset(&m, 0, sizeof(m));
                m.type = MCPT_MAT_EMISSIVE;
                for (int k = 0; k < 3; k++) m.emission[k] = (I[k] * scl[k]) / (6.f * half * half);
                add_box(pos, vec3(half, half, half), material(m));
                L.info.approximated_lights++;
            } else {
                L.info.approximated_lights++;
            }
        }
        L.info.width = P.width;
        L.info.height = P.height;
        if (P.has_camera) {  // :2443-2461
            const M4& c = P.camera_to_world;
            const vec3 from = xpoint(c, vec3(0.f, 0.f, 0.f));
            const vec3 fwd = xvector(c, vec3(0.f, 0.f, 1.f)).normalize();
            const vec3 at = from + fwd;
            const vec3 up = xvector(c, vec3(0.f, 1.f, 0.f)).normalize();
            L.info.has_camera = 1;
            const float f3[3][3] = {{from.x, from.y, from.z}, {at.x, at.y, at.z}, {up.x, up.y, up.z}};
            for (int k = 0; k < 3; k++) {
                L.info.look_from[k] = f3[0][k];
                L.info.look_at[k] = f3[1][k];
                L.info.up[k] = f3[2][k];
            }
            const float w = (float)(P.width > 0 ? P.width : 1), h = (float)(P.height > 0 ? P.height : 1);
            L.info.vfov_degrees = w >= h ? P.fov : 2.f * atanf(tanf(P.fov * 0.5f * kPi / 180.f) * h / w) * 180.f / kPi;
        }
        if (L.mats.empty()) {
            const float grey[3] = {0.5f, 0.5f, 0.5f};
            material(lambert(grey));
        }
        L.info.n_tris = (int32_t)L.hot.size();
        L.info.n_spheres = (int32_t)L.spheres.size();
        L.info.n_materials = (int32_t)L.mats.size();
        L.info.textures_loaded = (int32_t)L.images->size();
        return true;
    }
};

int load(const char* path, int64_t budget, Loaded& out) {
    std::memset(&out.info, 0, sizeof(out.info));
    Parsed parsed;
    Parser parser(parsed);
    if (!parser.run_file(path)) return set_error(MCPT_ERR_INVALID, "pbrt '%s': %s", path, parser.err.c_str());
    Builder b{parsed, out, budget > 0 ? budget : (int64_t(1) << 30)};
    b.named_materials = &parser.named;
    if (!b.build()) return set_error(MCPT_ERR_INVALID, "pbrt '%s': %s", path, b.err.c_str());
    if (out.hot.empty() && out.spheres.empty())
        return set_error(MCPT_ERR_INVALID, "pbrt '%s': no triangles or spheres were added", path);
    return MCPT_OK;
}

// the sizing call keeps its result for this thread's filling call
struct Cached {
    bool valid = false;
    std::string path;
    int64_t budget = 0;
    long long size = 0, mtime = 0;
    Loaded data;
};
thread_local Cached g_cache;
thread_local std::shared_ptr<std::vector<ImageTex>> g_images;

bool file_id(const char* path, long long& size, long long& mtime) {
    struct stat sb;
    if (stat(path, &sb) != 0) return false;
    size = (long long)sb.st_size;
    mtime = (long long)sb.st_mtim.tv_sec * 1000000000LL + sb.st_mtim.tv_nsec;
    return true;
}

}  // namespace

int load_pbrt(const char* path, int64_t max_prims, mcpt_tri_hot* hot, mcpt_tri_cold* cold, mcpt_sphere* spheres,
              mcpt_material* materials, mcpt_pbrt_info* info) {
    if (!path) return set_error(MCPT_ERR_INVALID, "path is NULL");
    long long size = -1, mtime = -1;
    const bool have_id = file_id(path, size, mtime);
    Cached& c = g_cache;
    const bool hit = c.valid && have_id && c.path == path && c.size == size && c.mtime == mtime && c.budget == max_prims;
    if (!hit) {
        c.valid = false;
        MCPT_TRY(load(path, max_prims, c.data));
        c.path = path;
        c.budget = max_prims;
        c.size = size;
        c.mtime = mtime;
        c.valid = have_id;
    }
    const Loaded& l = c.data;
    g_images = l.images;
    if (info) *info = l.info;
    if (hot) std::memcpy(hot, l.hot.data(), l.hot.size() * sizeof(mcpt_tri_hot));
    if (cold) std::memcpy(cold, l.cold.data(), l.cold.size() * sizeof(mcpt_tri_cold));
    if (spheres) std::memcpy(spheres, l.spheres.data(), l.spheres.size() * sizeof(mcpt_sphere));
    if (materials) std::memcpy(materials, l.mats.data(), l.mats.size() * sizeof(mcpt_material));
    if (hot || cold || spheres || materials) c = Cached();
    return MCPT_OK;
}

int pbrt_texture(int index, int32_t* width, int32_t* height, uint8_t* rgba8) {
    if (!width || !height) return set_error(MCPT_ERR_INVALID, "bad arguments");
    const int n = g_images ? (int)g_images->size() : 0;
    if (index < 0 || index >= n)
        return set_error(MCPT_ERR_INVALID, "texture %d: the last pbrt load on this thread holds %d", index, n);
    const ImageTex& im = (*g_images)[index];
    *width = im.w;
    *height = im.h;
    if (rgba8) std::memcpy(rgba8, im.rgba.data(), im.rgba.size());
    return MCPT_OK;
}

}  // namespace mcpt
