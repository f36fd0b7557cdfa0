// nvdb.cpp — NanoVDB float grids (.nvdb) as the dense density grid of
// mcpt_scene_set_density_grid.
//
// The reference reads a volume with nanovdb::io::readGrid and keeps the
// grid's bytes for the device (loadNanoVdbGrid, src/volume_io.cpp:45-53),
// samples it with a trilinear sampler in index space (sampleVolumeDensity,
// src/path_tracer.cu:784-789) and takes the root's maximum statistic as the
// majorant (takeFloatGrid, src/volume_io.cpp:20-26). NanoVDB (32.7, part of
// OpenVDB 12) is not in this image; this reader follows its published
// binary layout, little-endian, and densifies the sparse tree over the
// root's index bounding box, so the dense grid's trilinear sample at a
// world point is the sparse grid's (voxels outside the box read 0, a fog
// volume's background).
//
// File (nanovdb/io): a segment is a 16-byte header (magic "NanoVDB0" or
// "NanoVDB2", version, grid count, codec), one 176-byte record plus name per
// grid, then the grids' buffers in order (raw, or for ZIP a u64 size and a
// zlib stream). Grid buffer: GridData (672 bytes: magic, version, flags,
// index, count, size, name[256], the affine map, world box, voxel size,
// class, type, blind data), TreeData (64 bytes: byte offsets from the tree
// to its first leaf, lower, upper node and root; node and tile counts),
// RootData (index box, tile count, background, min, max, average, deviation;
// 64 bytes) and its 32-byte tiles (key, child offset from the root, state,
// value), internal nodes of 32^3 and 16^3 children (box, flags, value and
// child masks, statistics, then a 32-byte-aligned table of 8-byte entries:
// a value, or a child offset from the node) and 8^3 float leaves (min
// corner, box extent, flags, value mask, statistics, then 512 values at
// byte 96).
#include <zlib.h>

#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "host_common.hpp"

namespace mcpt {

namespace {

constexpr uint64_t kMagicOld = 0x304244566f6e614eULL;   // "NanoVDB0"
constexpr uint64_t kMagicGrid = 0x314244566f6e614eULL;  // "NanoVDB1"
constexpr uint64_t kMagicFile = 0x324244566f6e614eULL;  // "NanoVDB2"
constexpr uint32_t kGridTypeFloat = 1;
constexpr uint16_t kCodecNone = 0, kCodecZip = 1;
constexpr size_t kFileHeader = 16, kFileMeta = 176, kGridData = 672, kTreeData = 64, kRootData = 64, kRootTile = 32;
constexpr size_t kUpperTable = 8256, kLowerTable = 1088, kLeafValues = 96;
constexpr size_t kUpperMask = 4096, kLowerMask = 512;

template <class T>
T get(const std::vector<char>& b, size_t at) {
    T v;
    std::memcpy(&v, b.data() + at, sizeof(T));
    return v;
}

struct Grid {
    const std::vector<char>* buf = nullptr;
    size_t tree = 0, root = 0;
    int bmin[3] = {0, 0, 0}, bmax[3] = {-1, -1, -1};
    float background = 0.0f, maximum = 0.0f;
    double mat[9] = {}, vec[3] = {};
};

struct Dense {
    int lo[3], n[3];
    float* v;  // x fastest
    // fills [a, a + s) per axis (clipped to the dense box) with `value`
    void fill(const int a[3], int s, float value) const {
        int from[3], to[3];
        for (int c = 0; c < 3; c++) {
            from[c] = std::max(a[c], lo[c]) - lo[c];
            to[c] = std::min(a[c] + s, lo[c] + n[c]) - lo[c];
            if (from[c] >= to[c]) return;
        }
        for (int z = from[2]; z < to[2]; z++)
            for (int y = from[1]; y < to[1]; y++)
                for (int x = from[0]; x < to[0]; x++) v[((size_t)z * n[1] + y) * n[0] + x] = value;
    }
};

int bad(const char* path, const char* what) { return set_error(MCPT_ERR_INVALID, "%s: %s", path, what); }

bool in_range(const std::vector<char>& b, size_t at, size_t n) { return at <= b.size() && n <= b.size() - at; }

// the grid buffer of the first float grid in `path`
int read_buffer(const char* path, std::vector<char>& grid, std::string& name) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return set_error(MCPT_ERR_INVALID, "cannot open %s", path);
    std::vector<char> file((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    size_t at = 0;
    while (in_range(file, at, kFileHeader)) {  // segments
        const uint64_t magic = get<uint64_t>(file, at);
        if (magic != kMagicOld && magic != kMagicFile) return bad(path, "not a NanoVDB file");
        const uint32_t version = get<uint32_t>(file, at + 8);
        if ((version >> 21) != 32) return bad(path, "NanoVDB major version is not 32");
        const int count = get<uint16_t>(file, at + 12);
        at += kFileHeader;
        struct Meta {
            uint64_t grid_size, file_size;
            uint32_t type;
            uint16_t codec;
            std::string name;
        };
        std::vector<Meta> metas;
        for (int g = 0; g < count; g++) {
            if (!in_range(file, at, kFileMeta)) return bad(path, "truncated grid record");
            Meta m;
            m.grid_size = get<uint64_t>(file, at);
            m.file_size = get<uint64_t>(file, at + 8);
            m.type = get<uint32_t>(file, at + 32);
            const uint32_t name_size = get<uint32_t>(file, at + 136);
            m.codec = get<uint16_t>(file, at + 168);
            at += kFileMeta;
            if (!in_range(file, at, name_size)) return bad(path, "truncated grid name");
            m.name.assign(file.data() + at, name_size ? name_size - 1 : 0);
            at += name_size;
            metas.push_back(m);
        }
        for (const Meta& m : metas) {
            if (!in_range(file, at, m.file_size)) return bad(path, "truncated grid");
            if (m.type == kGridTypeFloat) {
                if (m.grid_size < kGridData + kTreeData || m.grid_size > (uint64_t(1) << 40))
                    return bad(path, "bad grid size");
                grid.resize(m.grid_size);
                if (m.codec == kCodecNone) {
                    if (m.file_size != m.grid_size) return bad(path, "grid size and stored size differ");
                    std::memcpy(grid.data(), file.data() + at, m.grid_size);
                } else if (m.codec == kCodecZip) {
                    if (m.file_size < 8) return bad(path, "truncated ZIP grid");
                    const uint64_t zsize = get<uint64_t>(file, at);
                    if (zsize > m.file_size - 8) return bad(path, "ZIP size beyond the grid");
                    uLongf out = (uLongf)m.grid_size;
                    if (uncompress(reinterpret_cast<Bytef*>(grid.data()), &out,
                                   reinterpret_cast<const Bytef*>(file.data() + at + 8), (uLong)zsize) != Z_OK ||
                        out != m.grid_size)
                        return bad(path, "ZIP grid does not inflate to its size");
                } else {
                    return bad(path, "BLOSC-compressed grids are not supported");
                }
                name = m.name;
                return MCPT_OK;
            }
            at += m.file_size;
        }
    }
    return bad(path, "no float grid");
}

int parse_grid(const char* path, const std::vector<char>& b, Grid& g) {
    g.buf = &b;
    const uint64_t magic = get<uint64_t>(b, 0);
    if (magic != kMagicOld && magic != kMagicGrid) return bad(path, "grid buffer without the grid magic");
    if (get<uint64_t>(b, 32) != b.size()) return bad(path, "grid buffer size differs from its record");
    if (get<uint32_t>(b, 636) != kGridTypeFloat) return bad(path, "grid is not a float grid");
    for (int i = 0; i < 9; i++) g.mat[i] = get<double>(b, 296 + 88 + 8 * i);
    for (int i = 0; i < 3; i++) g.vec[i] = get<double>(b, 296 + 88 + 144 + 8 * i);
    g.tree = kGridData;
    const int64_t root_off = get<int64_t>(b, g.tree + 24);
    if (root_off < (int64_t)kTreeData || !in_range(b, g.tree + (size_t)root_off, kRootData))
        return bad(path, "root outside the grid buffer");
    g.root = g.tree + (size_t)root_off;
    for (int c = 0; c < 3; c++) {
        g.bmin[c] = get<int32_t>(b, g.root + 4 * c);
        g.bmax[c] = get<int32_t>(b, g.root + 12 + 4 * c);
    }
    g.background = get<float>(b, g.root + 28);
    g.maximum = get<float>(b, g.root + 36);
    return MCPT_OK;
}

bool bit(const std::vector<char>& b, size_t mask, int n) {
    return (get<uint64_t>(b, mask + 8 * (size_t)(n >> 6)) >> (n & 63)) & 1u;
}

// a node's child at byte offset `off` from the node, checked against the buffer
bool child_at(const std::vector<char>& b, size_t node, int64_t off, size_t need, size_t& out) {
    if (off <= 0 || (uint64_t)off > b.size()) return false;
    out = node + (size_t)off;
    return in_range(b, out, need);
}

int densify_leaf(const char* path, const std::vector<char>& b, size_t leaf, const int o[3], const Dense& d) {
    if (!in_range(b, leaf, kLeafValues + 512 * 4)) return bad(path, "leaf outside the grid buffer");
    for (int x = 0; x < 8; x++)
        for (int y = 0; y < 8; y++)
            for (int z = 0; z < 8; z++) {
                const int p[3] = {o[0] + x, o[1] + y, o[2] + z};
                d.fill(p, 1, get<float>(b, leaf + kLeafValues + 4 * (size_t)((x << 6) | (y << 3) | z)));
            }
    return MCPT_OK;
}

// an internal node of 2^log2 children per axis, each child covering 2^child_total voxels per axis
int densify_internal(const char* path, const std::vector<char>& b, size_t node, int log2, int child_total,
                     const int o[3], const Dense& d) {
    const size_t mask_bytes = log2 == 5 ? kUpperMask : kLowerMask;
    const size_t table = log2 == 5 ? kUpperTable : kLowerTable;
    const int dim = 1 << log2, span = 1 << child_total;
    if (!in_range(b, node, table + 8 * (size_t)dim * dim * dim)) return bad(path, "node outside the grid buffer");
    const size_t child_mask = node + 32 + mask_bytes;
    for (int x = 0; x < dim; x++)
        for (int y = 0; y < dim; y++)
            for (int z = 0; z < dim; z++) {
                const int n = (x << (2 * log2)) | (y << log2) | z;
                const int c[3] = {o[0] + x * span, o[1] + y * span, o[2] + z * span};
                bool overlaps = true;
                for (int a = 0; a < 3; a++)
                    overlaps = overlaps && c[a] + span > d.lo[a] && c[a] < d.lo[a] + d.n[a];
                if (!overlaps) continue;
                const size_t entry = table + 8 * (size_t)n;
                if (bit(b, child_mask, n)) {
                    size_t ch;
                    if (child_total == 3) {
                        if (!child_at(b, node, get<int64_t>(b, node + entry), kLeafValues + 2048, ch))
                            return bad(path, "leaf offset outside the grid buffer");
                        MCPT_TRY(densify_leaf(path, b, ch, c, d));
                    } else {
                        if (!child_at(b, node, get<int64_t>(b, node + entry), kLowerTable, ch))
                            return bad(path, "node offset outside the grid buffer");
                        MCPT_TRY(densify_internal(path, b, ch, 4, 3, c, d));
                    }
                } else {
                    d.fill(c, span, get<float>(b, node + entry));
                }
            }
    return MCPT_OK;
}

// NanoVDB root key of the 4096^3 block holding ijk (CoordToKey)
uint64_t nvdb_root_key(int i, int j, int k) {
    return (uint64_t)((uint32_t)k >> 12) | ((uint64_t)((uint32_t)j >> 12) << 21) |
           ((uint64_t)((uint32_t)i >> 12) << 42);
}

}  // namespace

int read_nvdb(const char* path, mcpt_density_grid* out, float* density, float* max_density) {
    std::vector<char> buf;
    std::string name;  // the grid's name in the file
    MCPT_TRY(read_buffer(path, buf, name));
    Grid g;
    MCPT_TRY(parse_grid(path, buf, g));
    if (g.background != 0.0f) return bad(path, "the grid's background is not 0 (not a fog volume)");
    if (g.mat[1] != 0.0 || g.mat[2] != 0.0 || g.mat[3] != 0.0 || g.mat[5] != 0.0 || g.mat[6] != 0.0 ||
        g.mat[7] != 0.0 || !(g.mat[0] > 0.0 && g.mat[4] > 0.0 && g.mat[8] > 0.0))
        return bad(path, "the grid's index-to-world map is not an axis-aligned scale and translation");
    int64_t n[3];
    for (int c = 0; c < 3; c++) {
        n[c] = (int64_t)g.bmax[c] - g.bmin[c] + 1;
        if (n[c] < 1 || n[c] > (1 << 20)) return bad(path, "empty or oversized index bounding box");
    }
    if (n[0] * n[1] * n[2] > (int64_t(1) << 34)) return bad(path, "dense grid above 16 Gi voxels");
    for (int c = 0; c < 3; c++) {
        out->dims[c] = (int32_t)n[c];
        out->voxel_size[c] = (float)g.mat[4 * c];
        out->origin[c] = (float)(g.vec[c] + g.mat[4 * c] * g.bmin[c]);
    }
    out->reserved = 0;
    out->density = density;
    if (max_density) *max_density = g.maximum;
    if (!density) return MCPT_OK;
    const size_t total = (size_t)(n[0] * n[1] * n[2]);
    std::fill(density, density + total, 0.0f);
    Dense d{{g.bmin[0], g.bmin[1], g.bmin[2]}, {(int)n[0], (int)n[1], (int)n[2]}, density};
    const uint32_t tiles = get<uint32_t>(buf, g.root + 24);
    if (!in_range(buf, g.root + kRootData, (size_t)tiles * kRootTile)) return bad(path, "root tiles outside the grid");
    for (uint32_t t = 0; t < tiles; t++) {
        const size_t tile = g.root + kRootData + (size_t)t * kRootTile;
        const uint64_t key = get<uint64_t>(buf, tile);
        // the block's origin from its key (21 bits per axis, sign-extended)
        int o[3];
        const uint32_t kx = (uint32_t)(key >> 42) & 0x1FFFFFu, ky = (uint32_t)(key >> 21) & 0x1FFFFFu,
                       kz = (uint32_t)key & 0x1FFFFFu;
        o[0] = (int)(kx << 12);
        o[1] = (int)(ky << 12);
        o[2] = (int)(kz << 12);
        if (nvdb_root_key(o[0], o[1], o[2]) != key) return bad(path, "root tile key out of range");
        const int64_t child = get<int64_t>(buf, tile + 8);
        if (child) {
            size_t up;
            if (!child_at(buf, g.root, child, kUpperTable, up)) return bad(path, "upper node outside the grid buffer");
            MCPT_TRY(densify_internal(path, buf, up, 5, 7, o, d));
        } else {
            d.fill(o, 4096, get<float>(buf, tile + 20));
        }
    }
    for (size_t i = 0; i < total; i++)
        if (!(density[i] >= 0.0f) || density[i] == __builtin_inff())
            return bad(path, "negative or non-finite density");
    return MCPT_OK;
}

}  // namespace mcpt
