// cache.hip — the radiance cache: a BF16 MLP on the matrix cores, trained
// online with Adam on the tracer's own paths, and the wavefront render whose
// paths end in it. The reference has no learned component; the network and
// the loss are the ones include/mcpt.h declares.
//
// Matrix products use v_mfma_f32_16x16x32_bf16 (gfx950): D(16x16, f32) +=
// A(16x32, bf16) * B(32x16, bf16), with lane l = r + 16 g (r < 16, g < 4)
// holding A[r][8g..8g+7], B[8g..8g+7][r] and D[4g..4g+3][r]. Layers are
// evaluated transposed, X_{l+1}^T = relu(W_l X_l^T): A is a 16-row block of
// W_l, B a block of 16 samples' inputs. A wave owns 16 samples and keeps
// their activations in registers from layer to layer (k_mlp_infer).
//
// Training (mcpt_mlp_train, and every frame of mcpt_render_cached): one
// kernel runs forward and backward per tile of samples (k_mlp_grad) and
// keeps the weight gradients of its tiles in registers; each block writes
// its partial sums and a second kernel adds the blocks in a fixed order, so
// gradients do not depend on scheduling. Adam then updates FP32 master
// weights and writes their BF16 image.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "scene.hpp"
#include "shade.hpp"

namespace mcpt {
namespace {

constexpr int kBlock = 256;
constexpr int kIn = MCPT_MLP_INPUTS;           // 64
constexpr int kW = MCPT_MLP_WIDTH;             // 64
constexpr int kRow = 72;                       // bf16 row stride in LDS and in the weight image (144 B)
constexpr int kOutRows = 16;                   // output layer rows padded to one MFMA tile
constexpr int kImgRows = MCPT_MLP_HIDDEN * kW + kOutRows;
constexpr int kImg = kImgRows * kRow;          // bf16 elements of the weight image
constexpr int kParams = MCPT_MLP_PARAMS;
constexpr int kMaxGradBlocks = 256;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

static_assert(kIn == 64 && kW == 64, "the tiling assumes 64-wide layers");

// row of parameter (layer, out) in the weight image, and the flat index
MCPT_HDI int img_row(int layer, int o) { return layer * kW + o; }
MCPT_HDI int param_index(int layer, int o, int in) { return layer * kW * kIn + o * kIn + in; }

struct Query {  // one cache lookup (64 bytes)
    float f[MCPT_MLP_FEATURES];
    float thr[3];
    int path;
    float pad;
};
static_assert(sizeof(Query) == 64, "query record is 64 bytes");

// ------------------------------------------------------------ encoding
// input f of the 64 (order as in include/mcpt.h and oracle/mlp.py)
MCPT_DEV float enc_feature(int f, const float* x) {
    if (f < 36) {
        const int a = f / 12, r = f % 12, k = r >> 1;
        const float arg = float(1 << k) * MCPT_PI_F * x[a];
        return (r & 1) ? mf_cos(arg) : mf_sin(arg);
    }
    if (f < 52) {
        const bool dir = f < 44;
        const int r = f - (dir ? 36 : 44), c = r >> 2, k = (r >> 1) & 1;
        const float arg = float(1 << k) * MCPT_PI_F * x[(dir ? 3 : 5) + c];
        return (r & 1) ? mf_cos(arg) : mf_sin(arg);
    }
    if (f == 52) return x[7];
    if (f < 56) return x[8 + (f - 53)];
    if (f == 56) return 1.0f;
    if (f < 60) return x[f - 57];
    if (f < 62) return x[3 + (f - 60)];
    return x[5 + (f - 62)];
}

// unit vector -> (theta/pi, phi/(2 pi) + 1/2)
MCPT_DEV void sph2(const vec3& d, float* out) {
    const float z = fminf(fmaxf(d.z, -1.0f), 1.0f);
    out[0] = mf_acos(z) / MCPT_PI_F;
    out[1] = mf_atan2(d.y, d.x) / (2.0f * MCPT_PI_F) + 0.5f;
}

// ------------------------------------------------------------- kernels
// Inference keeps each sample's activations in registers from the encoding
// to the output. A wave owns 16 samples and computes all 64 rows of every
// layer for them (4 row blocks x 2 K chunks = 8 MFMAs per layer). Its D
// registers (lane r + 16g: rows 16i + 4g + e of sample r) become the next
// layer's B operand (lane r + 16g: inputs 8g..8g+7 of each 32-input chunk)
// without moving: B element e of chunk kc is taken to be row
// 16(2kc + e/4) + 4g + e%4, and the next layer's weight columns are stored
// in that order (infer_inh). Layer 0's inputs are computed in registers the
// same way, each lane encoding 16 of its sample's 64 inputs (infer_in0).
// So waves exchange nothing, need no barrier after the weights are staged,
// and read only weights (16 bytes per MFMA) from LDS.
constexpr int kInferBlock = 512;

// the input of layer 0 that logical input k stands for: lane group g of
// chunk kc supplies inputs j = 8kc + e of its group (position axis g's
// octaves, then direction / normal component g's; group 3: the normal's
// second component, then the raw features)
MCPT_HDI constexpr int infer_in0(int k) {
    const int kc = k >> 5, g = (k >> 3) & 3, j = 8 * kc + (k & 7);
    if (g < 3) return j < 12 ? 12 * g + j : 36 + 4 * g + (j - 12);
    return j < 4 ? 48 + j : 52 + (j - 4);
}
// the row of the previous layer that logical input k of layers 1..4 stands for
MCPT_HDI constexpr int infer_inh(int k) {
    const int kc = k >> 5, g = (k >> 3) & 3, e = k & 7;
    return 16 * (2 * kc + (e >> 2)) + 4 * g + (e & 3);
}

// the image column that holds logical input m of layer 0 (inverse of
// infer_in0) and of layers 1..4 (inverse of infer_inh)
MCPT_HDI constexpr int image_col0(int m) {
    int g = 3, j = m < 52 ? m - 48 : 4 + (m - 52);
    if (m < 36) {
        g = m / 12;
        j = m % 12;
    } else if (m < 48) {
        g = (m - 36) / 4;
        j = 12 + (m - 36) % 4;
    }
    return 32 * (j >> 3) + 8 * g + (j & 7);
}
MCPT_HDI constexpr int image_colh(int m) {
    const int q = m >> 4, g = (m >> 2) & 3, e2 = m & 3;
    return 32 * (q >> 1) + 8 * g + 4 * (q & 1) + e2;
}
MCPT_HDI constexpr int image_col(int layer, int m) { return layer == 0 ? image_col0(m) : image_colh(m); }
MCPT_HDI constexpr int image_logical(int layer, int k) { return layer == 0 ? infer_in0(k) : infer_inh(k); }

// each image column map is a permutation with the stated inverse
constexpr bool image_cols_invert() {
    for (int k = 0; k < kIn; k++)
        if (image_col0(infer_in0(k)) != k || image_colh(infer_inh(k)) != k) return false;
    return true;
}
static_assert(image_cols_invert(), "image_col0 / image_colh invert infer_in0 / infer_inh");

// The weight image keeps each layer's columns in operand order (column k
// of layer l holds logical input image_logical(l, k)), so a block stages it
// with plain 16-byte copies.
MCPT_DEV void load_weights_infer(const __bf16* wimg, __bf16* w) {
    const uint4* src = reinterpret_cast<const uint4*>(wimg);
    uint4* dst = reinterpret_cast<uint4*>(w);
    for (int i = threadIdx.x; i < kImg / 8; i += blockDim.x) dst[i] = src[i];
}

// sin and cos of 2^k pi x for k < n: the hardware sine and cosine of pi x
// (their argument is in turns), then the double-angle identities
// sin 2a = (2 sin a) cos a, cos 2a = 1 - (2 sin a) sin a: three operations
// an octave
MCPT_DEV void octaves(float x, int n, float* out) {
    float s = __builtin_amdgcn_sinf(0.5f * x), c = __builtin_amdgcn_cosf(0.5f * x);
    for (int k = 0; k < n; k++) {
        out[2 * k] = s;
        out[2 * k + 1] = c;
        const float t = s + s;
        c = __builtin_fmaf(-t, s, 1.0f);
        s = t * out[2 * k + 1];
    }
}

// lane group g's 16 inputs of a sample as two B chunks, from its features
// f0..f3 (fa), f4..f7 (fb), f8..f10 (fc.xyz). Every lane runs the same
// code: groups 0..2 take six octaves of position axis g and two of
// direction / normal component g; group 3 takes the first two octaves of
// f6 (the same arithmetic) and then the raw features. The features stay in
// named registers (an array picked by a lane index would go to scratch).
MCPT_DEV void encode_lane(const float4& fa, const float4& fb, const float4& fc, int g, bf16x8& b0, bf16x8& b1) {
    const float xa = g == 0 ? fa.x : g == 1 ? fa.y : g == 2 ? fa.z : fb.z;
    const float xb = g == 0 ? fa.w : g == 1 ? fb.x : fb.y;
    float v[16];
    octaves(xa, 6, v);
    octaves(xb, 2, v + 12);
    if (g == 3) {
        v[4] = fb.w;
        v[5] = fc.x;
        v[6] = fc.y;
        v[7] = fc.z;
        v[8] = 1.0f;
        v[9] = fa.x;
        v[10] = fa.y;
        v[11] = fa.z;
        v[12] = fa.w;
        v[13] = fb.x;
        v[14] = fb.y;
        v[15] = fb.z;
    }
    for (int e = 0; e < 8; e++) {
        b0[e] = (__bf16)v[e];
        b1[e] = (__bf16)v[8 + e];
    }
}

typedef short i16x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
// bf16(max(a, 0)), bf16(max(b, 0)) packed: round first (one packed
// conversion), then clamp as 16-bit integers (a negative bf16 is a negative
// integer, so an integer max with 0 maps it to +0). Rounding and clamping
// commute for every non-NaN input.
MCPT_DEV uint32_t relu_pk(float a, float b) {
    const i16x2 v = __builtin_bit_cast(i16x2, __builtin_convertvector(f32x2{a, b}, bf16x2));
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(v, i16x2(0)));
}
// the 16 rows a lane holds after a layer, as the next layer's B chunks
MCPT_DEV void relu_bf16(const f32x4 d[4], bf16x8& b0, bf16x8& b1) {
    b0 = __builtin_bit_cast(bf16x8, u32x4{relu_pk(d[0][0], d[0][1]), relu_pk(d[0][2], d[0][3]),
                                          relu_pk(d[1][0], d[1][1]), relu_pk(d[1][2], d[1][3])});
    b1 = __builtin_bit_cast(bf16x8, u32x4{relu_pk(d[2][0], d[2][1]), relu_pk(d[2][2], d[2][3]),
                                          relu_pk(d[3][0], d[3][1]), relu_pk(d[3][2], d[3][3])});
}

// bf16(x > 0 ? d : 0) of the 16 rows a lane holds, for x the layer's
// input (a ReLU output: +0 or positive): each packed half of bf16(d) times
// min(x bits, 1), which is 1 where x > 0 and 0 where x is +0.
MCPT_DEV uint32_t mask_pk(float a, float b, uint32_t x) {
    const i16x2 v = __builtin_bit_cast(i16x2, __builtin_convertvector(f32x2{a, b}, bf16x2));
    return __builtin_bit_cast(uint32_t, v * __builtin_elementwise_min(__builtin_bit_cast(i16x2, x), i16x2(1)));
}
MCPT_DEV void masked_bf16(const f32x4 d[4], const bf16x8& x0, const bf16x8& x1, bf16x8& b0, bf16x8& b1) {
    const u32x4 m0 = __builtin_bit_cast(u32x4, x0), m1 = __builtin_bit_cast(u32x4, x1);
    b0 = __builtin_bit_cast(bf16x8, u32x4{mask_pk(d[0][0], d[0][1], m0[0]), mask_pk(d[0][2], d[0][3], m0[1]),
                                          mask_pk(d[1][0], d[1][1], m0[2]), mask_pk(d[1][2], d[1][3], m0[3])});
    b1 = __builtin_bit_cast(bf16x8, u32x4{mask_pk(d[2][0], d[2][1], m1[0]), mask_pk(d[2][2], d[2][3], m1[1]),
                                          mask_pk(d[3][0], d[3][1], m1[2]), mask_pk(d[3][2], d[3][3], m1[3])});
}

// Inference over queries. ADD: add thr * max(y, 0) to the path's sample (the
// cached render); otherwise write y (n*3). A wave takes kInferGroups groups
// of 16 queries at a time (independent MFMA chains, so one group's
// conversions overlap the other's products) and loads the next groups'
// features while it computes these. Queries past n read the last one's
// features and write nothing.
constexpr int kInferGroups = 2;
template <bool ADD>
__global__ __launch_bounds__(kInferBlock) void k_mlp_infer(const __bf16* wimg, const Query* q, const int* count_ptr,
                                                          int count, float* yout, float4* smp) {
    constexpr int G = kInferGroups;
    __shared__ __attribute__((aligned(16))) __bf16 w[kImg];
    const int n = count_ptr ? *count_ptr : count;
    load_weights_infer(wimg, w);
    __syncthreads();
    const int lane = lane_id(), r = lane & 15, g = lane >> 4;
    const int waves = (int)(gridDim.x * blockDim.x) >> 6;
    const int units = (n + 16 * G - 1) / (16 * G);
    int unit = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    if (unit >= units) return;
    // features f[0..11] of this lane's sample in each group, and with ADD
    // the sample's thr[1..2], path (the record's last 16 bytes)
    float4 pf[G][3], pt[G];
    auto fetch = [&](int u) {
        for (int j = 0; j < G; j++) {
            const float4* rec = reinterpret_cast<const float4*>(q + min(16 * (G * u + j) + r, n - 1));
            pf[j][0] = rec[0];
            pf[j][1] = rec[1];
            pf[j][2] = rec[2];
            if (ADD) pt[j] = rec[3];
        }
    };
    fetch(unit);
    while (true) {
        // the lane's weight offset, opaque to the compiler on every pass:
        // the A operands are read from LDS where they are used instead of
        // being held in 136 registers, which leaves room for more waves
        int wo = r * kRow + 8 * g;
        asm volatile("" : "+v"(wo));
        bf16x8 b[G][2];
        float thr0[G];
        float4 thr12[G];
        for (int j = 0; j < G; j++) {
            encode_lane(pf[j][0], pf[j][1], pf[j][2], g, b[j][0], b[j][1]);
            thr0[j] = pf[j][2].w;
            thr12[j] = pt[j];
        }
        const int next = unit + waves;
        if (next < units) fetch(next);
        for (int l = 0; l < MCPT_MLP_HIDDEN; l++) {
            const __bf16* Wl = w + img_row(l, 0) * kRow + wo;
            f32x4 d[G][4];
            for (int i = 0; i < 4; i++) {
                const bf16x8 a0 = *reinterpret_cast<const bf16x8*>(Wl + 16 * i * kRow);
                const bf16x8 a1 = *reinterpret_cast<const bf16x8*>(Wl + 16 * i * kRow + 32);
                for (int j = 0; j < G; j++) {
                    d[j][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b[j][0], f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
                    d[j][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b[j][1], d[j][i], 0, 0, 0);
                }
            }
            for (int j = 0; j < G; j++) relu_bf16(d[j], b[j][0], b[j][1]);
        }
        const __bf16* W4 = w + img_row(MCPT_MLP_HIDDEN, 0) * kRow + wo;
        const bf16x8 a0 = *reinterpret_cast<const bf16x8*>(W4);
        const bf16x8 a1 = *reinterpret_cast<const bf16x8*>(W4 + 32);
        for (int j = 0; j < G; j++) {
            f32x4 y = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b[j][0], f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
            y = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b[j][1], y, 0, 0, 0);
            const int s = 16 * (G * unit + j) + r;
            if (g == 0 && s < n) {  // lanes of group 0 hold outputs 0..3 of sample r
                if (ADD) {
                    const int path = __float_as_int(thr12[j].z);
                    float4 v = smp[path];
                    v.x = v.x + thr0[j] * fmaxf(y[0], 0.0f);
                    v.y = v.y + thr12[j].x * fmaxf(y[1], 0.0f);
                    v.z = v.z + thr12[j].y * fmaxf(y[2], 0.0f);
                    smp[path] = v;
                } else {
                    yout[3 * s] = y[0];
                    yout[3 * s + 1] = y[1];
                    yout[3 * s + 2] = y[2];
                }
            }
        }
        if (next >= units) break;
        unit = next;
    }
}

// Forward + backward over samples [first, first + n) of (feat, targ), tiles
// of kGradTile samples block-strided; per-block weight gradients and loss
// sums. A wave owns 16 samples of a tile and runs the forward pass the way
// k_mlp_infer does, keeping every layer's input in registers (the MFMA
// operand layout). The input gradients run backwards in registers the same
// way, with the transposed weights as A operands (wb: W_l^T with the
// permuted columns, w4b: W4^T). A weight gradient dW_l = sum over samples of
// dZ_l^T X_l needs the samples along K, so each layer's dZ_l and X_l are
// written transposed into LDS (one row per unit, one column per sample of
// the tile) and the block's waves split dW_l's 16x16 tiles among
// themselves, accumulating in registers across tiles. bfloat16 roundings:
// every activation, dY, and every masked input gradient, as the oracle.
constexpr int kGradWaves = 8;
constexpr int kGradBlock = 64 * kGradWaves;
constexpr int kGradTile = 16 * kGradWaves;  // samples per block iteration
constexpr int kTRow = kGradTile + 8;        // bf16 row stride of the transposed tiles
constexpr int kW4Row = 40;                  // W4^T: 32 columns (outputs, zero-padded) + pad
static_assert(kGradTile % 32 == 0, "dW takes samples in K chunks of 32");
// The transposed weight image the training kernel stages as is: W_l^T for
// l = 1..3 (row = input unit, columns in operand order over the outputs,
// kRow apart), then W4^T (row = input unit, the outputs zero-padded to 32,
// kW4Row apart).
constexpr int kTImgB = (MCPT_MLP_HIDDEN - 1) * kW * kRow;
constexpr int kTImg = kTImgB + kW * kW4Row;
static_assert(kTImgB % 8 == 0 && kTImg % 8 == 0, "16-byte staging copies");

// the row (unit) that element e of chunk kc of lane group g carries in the
// B-operand registers of layer l's input
MCPT_HDI int operand_row(int l, int kc, int g, int e) {
    const int k = 32 * kc + 8 * g + e;
    return l == 0 ? infer_in0(k) : infer_inh(k);
}

__global__ __launch_bounds__(kGradBlock) void k_mlp_grad(const __bf16* wimg, const __bf16* wimg_t, const float* feat,
                                                        const float* targ,
                                                        int first, int n, int n_total_for_mean, float* partial,
                                                        float* partial_loss) {
    __shared__ __attribute__((aligned(16))) __bf16 wf[kImg];                          // forward, infer layout
    __shared__ __attribute__((aligned(16))) __bf16 wt[kTImg];                          // W_l^T, l = 1..3; W4^T
    __shared__ __attribute__((aligned(16))) __bf16 xt[2][kW * kTRow];                  // X_l^T
    __shared__ __attribute__((aligned(16))) __bf16 dt[2][kW * kTRow];                  // dZ_l^T
    __shared__ float lsum[kGradWaves];
    load_weights_infer(wimg, wf);
    for (int i = threadIdx.x; i < kTImg / 8; i += kGradBlock)
        reinterpret_cast<uint4*>(wt)[i] = reinterpret_cast<const uint4*>(wimg_t)[i];
    __syncthreads();
    const __bf16* w4b = wt + kTImgB;
    const int lane = lane_id(), wave = (int)threadIdx.x >> 6, r = lane & 15, g = lane >> 4;
    const int col = 16 * wave + r;  // this lane's sample column in the transposed tiles
    const float inv_count = 1.0f / float(3 * n_total_for_mean);
    // dW tiles of this wave: hidden layers, output rows 16 ob.., inputs 16 nt0.., 16 (nt0 + 1)..
    const int ob = wave & 3, nt0 = 2 * (wave >> 2);
    f32x4 gh[MCPT_MLP_HIDDEN][2];
    for (int l = 0; l < MCPT_MLP_HIDDEN; l++) gh[l][0] = gh[l][1] = f32x4{0.f, 0.f, 0.f, 0.f};
    f32x4 go = f32x4{0.f, 0.f, 0.f, 0.f};  // output layer (waves 0..3): outputs 4g + e, inputs 16 wave + r
    float loss = 0.0f;
    const int ntiles = (n + kGradTile - 1) / kGradTile;
    // features and targets of this lane's sample of a tile, loaded one tile
    // ahead (samples past n read the last one's and contribute nothing)
    float4 fa, fb, fc;
    float tg[MCPT_MLP_OUTPUTS];
    auto fetch = [&](int tile) {
        const size_t i = (size_t)first + min(tile * kGradTile + col, n - 1);
        const float* fs = feat + i * MCPT_MLP_FEATURES;
        fa = make_float4(fs[0], fs[1], fs[2], fs[3]);
        fb = make_float4(fs[4], fs[5], fs[6], fs[7]);
        fc = make_float4(fs[8], fs[9], fs[10], 0.0f);
        for (int e = 0; e < MCPT_MLP_OUTPUTS; e++) tg[e] = targ[i * 3 + e];
    };
    if ((int)blockIdx.x < ntiles) fetch(blockIdx.x);
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int s = tile * kGradTile + col;
        const bool valid = s < n;
        // ---- forward, inputs of every layer kept (act[l] = input of layer l)
        bf16x8 act[MCPT_MLP_HIDDEN + 1][2];
        encode_lane(fa, fb, fc, g, act[0][0], act[0][1]);
        float t3[MCPT_MLP_OUTPUTS];
        for (int e = 0; e < MCPT_MLP_OUTPUTS; e++) t3[e] = tg[e];
        if (tile + (int)gridDim.x < ntiles) fetch(tile + gridDim.x);
#pragma unroll
        for (int l = 0; l < MCPT_MLP_HIDDEN; l++) {
            const __bf16* Wl = wf + img_row(l, 0) * kRow;
            f32x4 d[4];
            for (int i = 0; i < 4; i++) {
                const bf16x8 a0 = *reinterpret_cast<const bf16x8*>(Wl + (16 * i + r) * kRow + 8 * g);
                const bf16x8 a1 = *reinterpret_cast<const bf16x8*>(Wl + (16 * i + r) * kRow + 32 + 8 * g);
                d[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, act[l][0], f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
                d[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, act[l][1], d[i], 0, 0, 0);
            }
            relu_bf16(d, act[l + 1][0], act[l + 1][1]);
        }
        const __bf16* W4 = wf + img_row(MCPT_MLP_HIDDEN, 0) * kRow;
        f32x4 y = __builtin_amdgcn_mfma_f32_16x16x32_bf16(*reinterpret_cast<const bf16x8*>(W4 + r * kRow + 8 * g),
                                                         act[MCPT_MLP_HIDDEN][0], f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0,
                                                         0);
        y = __builtin_amdgcn_mfma_f32_16x16x32_bf16(*reinterpret_cast<const bf16x8*>(W4 + r * kRow + 32 + 8 * g),
                                                   act[MCPT_MLP_HIDDEN][1], y, 0, 0, 0);
        // ---- loss and dY (lanes of group 0 hold outputs 0..3 of sample r)
        bf16x8 dz0, dz1;  // the gradient at the current layer's output, B layout
        for (int e = 0; e < 8; e++) dz0[e] = dz1[e] = (__bf16)0.0f;
        if (g == 0) {
            for (int e = 0; e < MCPT_MLP_OUTPUTS; e++) {
                float d = 0.0f;
                if (valid) {
                    const float t = t3[e];
                    const float den = y[e] * y[e] + 0.01f;
                    const float diff = y[e] - t;
                    loss += diff * diff / den;
                    d = 2.0f * diff / den * inv_count;
                }
                dz0[e] = (__bf16)d;
            }
        }
        // ---- backward: stage l writes dZ_l^T and X_l^T, then dW_l from LDS
        // and dZ_{l-1} in registers
#pragma unroll
        for (int l = MCPT_MLP_HIDDEN; l >= 0; l--) {
            __bf16* X = xt[l & 1];
            __bf16* D = dt[l & 1];
            for (int kc = 0; kc < 2; kc++)
                for (int e = 0; e < 8; e++) X[operand_row(l, kc, g, e) * kTRow + col] = act[l][kc][e];
            if (l == MCPT_MLP_HIDDEN) {  // dY: rows 0..15 (outputs, then zeros)
                for (int e = 0; e < 8; e++)
                    if (8 * g + e < kOutRows) D[(8 * g + e) * kTRow + col] = dz0[e];
            } else {
                for (int kc = 0; kc < 2; kc++)
                    for (int e = 0; e < 8; e++) D[operand_row(1, kc, g, e) * kTRow + col] = (kc ? dz1 : dz0)[e];
            }
            __syncthreads();
            if (l == MCPT_MLP_HIDDEN) {
                if (wave < 4) {  // dW4: outputs 0..15 x inputs 16 wave.., K = the tile's samples
                    for (int kc = 0; kc < kGradTile / 32; kc++) {
                        const bf16x8 a = *reinterpret_cast<const bf16x8*>(D + r * kTRow + 32 * kc + 8 * g);
                        const bf16x8 b =
                            *reinterpret_cast<const bf16x8*>(X + (16 * wave + r) * kTRow + 32 * kc + 8 * g);
                        go = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, go, 0, 0, 0);
                    }
                }
            } else {
                for (int kc = 0; kc < kGradTile / 32; kc++) {
                    const bf16x8 a = *reinterpret_cast<const bf16x8*>(D + (16 * ob + r) * kTRow + 32 * kc + 8 * g);
                    for (int j = 0; j < 2; j++) {
                        const bf16x8 b =
                            *reinterpret_cast<const bf16x8*>(X + (16 * (nt0 + j) + r) * kTRow + 32 * kc + 8 * g);
                        gh[l][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, gh[l][j], 0, 0, 0);
                    }
                }
            }
            if (l == 0) break;
            // dX_l = W_l^T dZ_l, masked where layer l's input was clamped to 0
            f32x4 d[4];
            for (int i = 0; i < 4; i++) {
                if (l == MCPT_MLP_HIDDEN) {
                    const bf16x8 a = *reinterpret_cast<const bf16x8*>(w4b + (16 * i + r) * kW4Row + 8 * g);
                    d[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, dz0, f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
                } else {
                    const __bf16* Wt = wt + (l - 1) * kW * kRow;
                    const bf16x8 a0 = *reinterpret_cast<const bf16x8*>(Wt + (16 * i + r) * kRow + 8 * g);
                    const bf16x8 a1 = *reinterpret_cast<const bf16x8*>(Wt + (16 * i + r) * kRow + 32 + 8 * g);
                    d[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, dz0, f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
                    d[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, dz1, d[i], 0, 0, 0);
                }
            }
            masked_bf16(d, act[l][0], act[l][1], dz0, dz1);
        }
        __syncthreads();  // the next tile's writes after every read of this one
    }
    // partial sums of this block, in the flat parameter order
    float* out = partial + (size_t)blockIdx.x * kParams;
    for (int l = 0; l < MCPT_MLP_HIDDEN; l++)
        for (int j = 0; j < 2; j++)
            for (int e = 0; e < 4; e++) out[param_index(l, 16 * ob + 4 * g + e, 16 * (nt0 + j) + r)] = gh[l][j][e];
    if (wave < 4 && g == 0)
        for (int e = 0; e < MCPT_MLP_OUTPUTS; e++) out[param_index(MCPT_MLP_HIDDEN, e, 16 * wave + r)] = go[e];
    // loss: lanes r of group 0 hold per-sample sums
    float ls = loss;
    for (int off = 32; off > 0; off >>= 1) ls += __shfl_xor(ls, off);
    if (lane == 0) lsum[wave] = ls;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.0f;
        for (int w = 0; w < kGradWaves; w++) t += lsum[w];
        partial_loss[blockIdx.x] = t;
    }
}

// grad[p] = sum over the gradient blocks: a block takes 64 consecutive
// parameters (one wave reads 256 contiguous bytes of a block's partials)
// and 16 threads per parameter each add a contiguous sixteenth of the
// blocks in block order, then the sixteen sums are added in order, so the
// result does not depend on scheduling; the loss in block 0 by strided sums
// and a fixed-order tree
constexpr int kReduceSplit = 16;
constexpr int kReduceParams = 64;  // parameters per block
constexpr int kReduceBlock = kReduceSplit * kReduceParams;
__global__ __launch_bounds__(kReduceBlock) void k_mlp_reduce(const float* partial, const float* partial_loss,
                                                             int nblk, float* grad, float* loss_out,
                                                             float loss_scale) {
    __shared__ float part[kReduceSplit][kReduceParams];
    const int j = threadIdx.x % kReduceParams, q = threadIdx.x / kReduceParams;
    const int p = blockIdx.x * kReduceParams + j;
    const int b0 = (int)((int64_t)nblk * q / kReduceSplit), b1 = (int)((int64_t)nblk * (q + 1) / kReduceSplit);
    float s = 0.0f;
    if (p < kParams)
        for (int b = b0; b < b1; b++) s += partial[(size_t)b * kParams + p];
    part[q][j] = s;
    __syncthreads();
    if (q == 0 && p < kParams) {
        float t = part[0][j];
        for (int k = 1; k < kReduceSplit; k++) t += part[k][j];
        grad[p] = t;
    }
    if (blockIdx.x == 0) {  // the loss: strided sums, then a fixed-order tree in LDS
        __shared__ float ls[kReduceBlock];
        float l = 0.0f;
        for (int b = threadIdx.x; b < nblk; b += kReduceBlock) l += partial_loss[b];
        ls[threadIdx.x] = l;
        __syncthreads();
        for (int w = kReduceBlock / 2; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) ls[threadIdx.x] = ls[threadIdx.x] + ls[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) loss_out[0] = ls[0] * loss_scale;
    }
}

// the position of parameter (layer, o, in) in the transposed image
// (layers 1..4)
MCPT_HDI int timg_index(int layer, int o, int in) {
    return layer < MCPT_MLP_HIDDEN ? (layer - 1) * kW * kRow + in * kRow + image_colh(o) : kTImgB + in * kW4Row + o;
}

// Adam on the FP32 master weights, then their BF16 images (operand order
// and transposed); with decay > 0 also the exponential average of the
// weights and its image
__global__ __launch_bounds__(kBlock) void k_adam(float* master, float* m, float* v, const float* grad, __bf16* wimg,
                                                __bf16* wimg_t, float lr, float bc1, float bc2, float* ema,
                                                __bf16* wimg_ema, float decay) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= kParams) return;
    const float g = grad[p];
    const float mm = 0.9f * m[p] + 0.1f * g;
    const float vv = 0.99f * v[p] + 0.01f * g * g;
    m[p] = mm;
    v[p] = vv;
    const float wv = master[p] - lr * (mm / bc1) / (sqrtf(vv / bc2) + 1e-8f);
    master[p] = wv;
    const int layer = p / (kW * kIn), o = (p % (kW * kIn)) / kIn, in = p % kIn;
    const int at = img_row(layer, o) * kRow + image_col(layer, in);
    wimg[at] = (__bf16)wv;
    if (layer > 0) wimg_t[timg_index(layer, o, in)] = (__bf16)wv;
    if (decay > 0.0f) {
        const float a = decay * ema[p] + (1.0f - decay) * wv;
        ema[p] = a;
        wimg_ema[at] = (__bf16)a;
    }
}

// BF16 image of the master weights, columns in operand order (padding
// stays zero)
__global__ __launch_bounds__(kBlock) void k_pack(const float* master, __bf16* wimg) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= kImg) return;
    const int row = i / kRow, col = i % kRow;
    const int layer = row / kW, o = row % kW;
    float val = 0.0f;
    if (col < kIn && (layer < MCPT_MLP_HIDDEN || o < MCPT_MLP_OUTPUTS))
        val = master[param_index(layer, o, image_logical(layer, col))];
    wimg[i] = (__bf16)val;
}
// the transposed image (padding stays zero)
__global__ __launch_bounds__(kBlock) void k_pack_t(const float* master, __bf16* wimg_t) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= kTImg) return;
    float val = 0.0f;
    if (i < kTImgB) {
        const int layer = 1 + i / (kW * kRow), in = (i / kRow) % kW, k = i % kRow;
        if (k < kW) val = master[param_index(layer, infer_inh(k), in)];
    } else {
        const int in = (i - kTImgB) / kW4Row, o = (i - kTImgB) % kW4Row;
        if (o < MCPT_MLP_OUTPUTS) val = master[param_index(MCPT_MLP_HIDDEN, o, in)];
    }
    wimg_t[i] = (__bf16)val;
}

int blocks_for(int64_t n) { return (int)((n + kBlock - 1) / kBlock); }

// inference grid: no more blocks than the queries fill, and at most four
// times what fits on the device at once (occupancy): blocks that start late
// keep the SIMDs busy while the first ones drain (measured 5% faster than
// one resident grid), and every block stages the weights once
int infer_blocks(const Ctx* ctx, int64_t n) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_mlp_infer<false>, kInferBlock, 0) != hipSuccess)
        per_cu = 1;
    const int per_block = 16 * kInferGroups * (kInferBlock / 64);  // queries of one pass of a block
    const int64_t need = (n + per_block - 1) / per_block;
    return (int)std::max<int64_t>(1, std::min<int64_t>(need, (int64_t)std::max(1, per_cu) * ctx->cu_count * 4));
}

// ----------------------------------------------------- cached wavefront
// Counters, one 128-byte line each (see render.hip): C_NS0/C_NS1 are u64
// per ray queue, queue length in the low half and the length of the shadow
// queue filled in the same bounce in the high half.
enum { C_NS0 = 0, C_NS1 = 32, C_WORK = 64, C_QCOUNT = 96, C_NCOUNTERS = 128 };

MCPT_DEV unsigned long long* ns_slot(int* c, int q) {
    return reinterpret_cast<unsigned long long*>(c + (q ? C_NS1 : C_NS0));
}
MCPT_DEV int queue_len(const int* c, int q) { return c[q ? C_NS1 : C_NS0]; }
MCPT_DEV int shadow_len(const int* c, int q) { return c[(q ? C_NS1 : C_NS0) + 1]; }

struct alignas(16) VertexRec {  // a diffuse vertex of a training path (80 bytes)
    float f[MCPT_MLP_FEATURES];
    float s[3];  // sample before the vertex
    float t[3];  // throughput arriving at the vertex
    float pad[3];
};

struct CacheWs {
    RayItem* q[2];
    mcpt_packed_hit* hits;
    ShadowItem* sq;
    float4* thr;
    float4* smp;
    uint64_t* rng;
    int* pix;    // pixel of each path
    int* tslot;  // training slot of an image path, -1 for none
    int* tpix;   // pixel of a training slot
    Query* queries;
    VertexRec* vrec;  // train paths * max_depth
    int* nv;          // vertices recorded per training path
    int* offs;        // exclusive prefix of nv (train paths + 1)
    float* tfeat;     // compacted training batch
    float* ttarg;
    int* counters;
};

struct Frame {  // the per-frame constants of the shade kernels
    float bmin[3], bext[3];  // scene box for position features
    int max_depth;
    int depth;
    int term_depth;
};

MCPT_DEV void vertex_features(const Frame& fr, const vec3& p, const vec3& rd, const vec3& n, const vec3& albedo,
                              float* f) {
    f[0] = (p.x - fr.bmin[0]) / fr.bext[0];
    f[1] = (p.y - fr.bmin[1]) / fr.bext[1];
    f[2] = (p.z - fr.bmin[2]) / fr.bext[2];
    sph2(-rd, f + 3);
    sph2(n, f + 5);
    f[7] = 1.0f;  // diffuse
    f[8] = albedo.x;
    f[9] = albedo.y;
    f[10] = albedo.z;
}

// image paths: spp per pixel, path p = pixel + npix * k for sample
// sample0 + k (generateCameraRaysKernel, :3342-3374)
__global__ __launch_bounds__(kBlock) void k_cc_camera(mcpt_camera cam, uint32_t sample0, int spp, uint64_t seed,
                                                     CacheWs ws) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    const int W = cam.width, npix = cam.width * cam.height, n = npix * spp;
    if (p == 0) {
        *ns_slot(ws.counters, 0) = (unsigned long long)(uint32_t)n;
        ws.counters[C_QCOUNT] = 0;
    }
    if (p >= n) return;
    const int pixel = p % npix;
    Rng rng;
    mcpt_pcg32_for_path(&rng.g, (uint32_t)pixel, sample0 + (uint32_t)(p / npix), seed);
    vec3 o, d;
    float time;
    camera_ray(cam, pixel % W, pixel / W, rng, o, d, time);
    RayItem it;
    it.ot = make_float4(o.x, o.y, o.z, time);
    it.dp = make_float4(d.x, d.y, d.z, __int_as_float(p));
    ws.q[0][p] = it;
    ws.thr[p] = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
    ws.smp[p] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    ws.rng[p] = rng.g.state;
    ws.pix[p] = pixel;
    ws.tslot[p] = -1;
}

// training slots: image paths drawn at random that are traced without the
// cache and recorded; a path drawn twice keeps its larger slot
__global__ __launch_bounds__(kBlock) void k_cc_pick(uint32_t frame, uint64_t seed, int n, int npaths, CacheWs ws) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    mcpt_pcg32 pick;
    mcpt_pcg32_for_path(&pick, (uint32_t)i, frame, seed ^ 0x5851F42D4C957F2DULL);
    const int p = (int)(mcpt_pcg32_next(&pick) % (uint32_t)npaths);
    ws.tpix[i] = p;
    ws.nv[i] = 0;
    atomicMax(&ws.tslot[p], i);
}

__global__ void k_cc_bounce(int* c, int cur) {
    if (threadIdx.x == 0) {
        *ns_slot(c, 1 - cur) = 0;
        c[C_WORK] = 0;
    }
}

// the 16-byte hit record; prim_index keeps the leaf's code (triangle, or
// ~sphere)
__device__ __forceinline__ mcpt_packed_hit cc_hit(const Trav& s, const BvhView& v) {
    mcpt_packed_hit h;
    h.t = s.t;
    h.prim_type = s.leaf >= 0 ? 1 : -1;
    h.prim_index = s.leaf >= 0 ? v.leaf_prim[s.leaf] : -1;
    h.did_hit = s.leaf >= 0 ? 1 : 0;
    return h;
}

// persistent extend (extendKernelPersistent, :3416-3468) with deferred
// leaf tests; a queue shorter than group_below is k_cc_extend_group's
__global__ __launch_bounds__(kBlock) void k_cc_extend(SceneView sv, CacheWs ws, int cur, int refill,
                                                     int group_below) {
    extern __shared__ int lds_stack[];
    const int total = queue_len(ws.counters, cur);
    if (total < group_below) return;
    if ((int)blockIdx.x * kBlock >= total) return;  // blocks the queue does not need
    uint32_t ns = 0, nl = 0;
    const RayItem* q = ws.q[cur];
    persistent_closest<false, false>(
        sv.bvh, &ws.counters[C_WORK], total, refill, lds_stack + threadIdx.x, kBlock,
        [&](int i, Trav& s, const BvhView&) {
            const RayItem it = q[i];
            trav_begin(s, vec3(it.ot.x, it.ot.y, it.ot.z), vec3(it.dp.x, it.dp.y, it.dp.z), 1e30f);
            s.time = it.ot.w;
        },
        [&](int i, const Trav& s, const BvhView& v) { ws.hits[i] = cc_hit(s, v); }, ns, nl);
}

// a short queue: 16 lanes per ray (group_closest)
__global__ __launch_bounds__(kBlock) void k_cc_extend_group(SceneView sv, CacheWs ws, int cur, int group_below,
                                                           int cap, int cap_phys) {
    extern __shared__ int lds[];
    const int total = queue_len(ws.counters, cur);
    if (total >= group_below || (int)blockIdx.x * (kBlock / kGroup) >= total) return;
    int* base = lds + (threadIdx.x / kGroup) * group_lds_ints(cap_phys);
    const RayItem* q = ws.q[cur];
    BvhView bv = sv.bvh;
    group_closest(
        bv, &ws.counters[C_WORK], total, reinterpret_cast<int2*>(base), base + 2 * cap_phys, cap, cap_phys,
        [&](int i, Trav& s, const BvhView&) {
            const RayItem it = q[i];
            trav_begin(s, vec3(it.ot.x, it.ot.y, it.ot.z), vec3(it.dp.x, it.dp.y, it.dp.z), 1e30f);
            s.time = it.ot.w;
        },
        [&](int i, const Trav& s, const BvhView& v) { ws.hits[i] = cc_hit(s, v); });
}

// shadeKernel (:3470-3696). A training path records every diffuse vertex
// (features, sample and throughput on arrival) and is shaded as usual; on
// other paths a diffuse hit at depth >= term_depth ends the path in a cache
// query.
__global__ __launch_bounds__(kBlock) void k_cc_shade(SceneView sv, CacheWs ws, int cur, Frame fr, vec3 bg,
                                                    float rr_cap) {
    const int slot = blockIdx.x * kBlock + threadIdx.x;
    const bool active = slot < queue_len(ws.counters, cur);
    Bounce out;
    out.next = false;
    out.shadow = false;
    bool query = false;
    Query qr;
    int path = 0;
    float time = 0.f;
    if (active) {
        const RayItem it = ws.q[cur][slot];
        const mcpt_packed_hit hit = ws.hits[slot];
        path = __float_as_int(it.dp.w);
        time = it.ot.w;
        const vec3 ro(it.ot.x, it.ot.y, it.ot.z), rd(it.dp.x, it.dp.y, it.dp.z);
        const float4 thr = ws.thr[path], smp = ws.smp[path];
        PathState ps;
        ps.thr = vec3(thr.x, thr.y, thr.z);
        ps.prev_pdf = thr.w;
        ps.smp = vec3(smp.x, smp.y, smp.z);
        ps.prev_spec = smp.w != 0.0f;
        bool diffuse = false;
        HitInfo h;
        vec3 albedo;
        if (hit.did_hit) {
            const mcpt_material& m = sv.mats[prim_material(sv, hit.prim_index)];
            diffuse = m.type == MCPT_MAT_LAMBERTIAN;
            if (diffuse) {
                rebuild_hit(sv, hit.prim_index, ro, rd, hit.t, time, h, m.texture > 0);
                albedo = sample_albedo(sv, m, h.u, h.v);
            }
        }
        const int ts = ws.tslot[path];
        if (ts >= 0 && diffuse) {
            const int k = ws.nv[ts];
            VertexRec& v = ws.vrec[(size_t)ts * fr.max_depth + k];
            vertex_features(fr, h.p, rd, h.normal, albedo, v.f);
            v.s[0] = ps.smp.x;
            v.s[1] = ps.smp.y;
            v.s[2] = ps.smp.z;
            v.t[0] = ps.thr.x;
            v.t[1] = ps.thr.y;
            v.t[2] = ps.thr.z;
            ws.nv[ts] = k + 1;
        }
        if (ts < 0 && diffuse && fr.depth >= fr.term_depth) {
            query = true;
            vertex_features(fr, h.p, rd, h.normal, albedo, qr.f);
            qr.thr[0] = ps.thr.x;
            qr.thr[1] = ps.thr.y;
            qr.thr[2] = ps.thr.z;
            qr.path = path;
            qr.pad = 0.0f;
        } else {
            Rng rng;
            rng.g.state = ws.rng[path];
            rng.g.inc = ((uint64_t)(uint32_t)ws.pix[path] << 1u) | 1u;
            shade_bounce(sv, bg, rr_cap, fr.depth, ro, rd, hit.did_hit != 0, hit.t, hit.prim_index, ps, rng, out,
                         time);
            ws.thr[path] = make_float4(ps.thr.x, ps.thr.y, ps.thr.z, ps.prev_pdf);
            ws.smp[path] = make_float4(ps.smp.x, ps.smp.y, ps.smp.z, ps.prev_spec ? 1.0f : 0.0f);
            ws.rng[path] = rng.g.state;
        }
    }
    int nslot, sslot;
    wave_push2(ns_slot(ws.counters, 1 - cur), out.next, out.shadow, nslot, sslot);
    if (out.next) {
        RayItem it;
        it.ot = make_float4(out.no.x, out.no.y, out.no.z, time);
        it.dp = make_float4(out.nd.x, out.nd.y, out.nd.z, __int_as_float(path));
        ws.q[1 - cur][nslot] = it;
    }
    if (out.shadow) {
        ShadowItem si;
        si.om = make_float4(out.so.x, out.so.y, out.so.z, out.smax);
        si.dp = make_float4(out.sd.x, out.sd.y, out.sd.z, __int_as_float(path));
        si.c = make_float4(out.contrib.x, out.contrib.y, out.contrib.z, time);
        ws.sq[sslot] = si;
    }
    const int qslot = wave_push(&ws.counters[C_QCOUNT], query);
    if (query) ws.queries[qslot] = qr;
}

// shadowKernel (:3698-3728)
__global__ __launch_bounds__(kBlock) void k_cc_shadow(SceneView sv, CacheWs ws, int cur) {
    extern __shared__ int lds_stack[];
    const int slot = blockIdx.x * kBlock + threadIdx.x;
    if (slot >= shadow_len(ws.counters, 1 - cur)) return;
    const ShadowItem si = ws.sq[slot];
    const int path = __float_as_int(si.dp.w);
    const float tr = shadow_transmittance(sv, vec3(si.om.x, si.om.y, si.om.z), vec3(si.dp.x, si.dp.y, si.dp.z),
                                          si.om.w, lds_stack + threadIdx.x, kBlock, si.c.w);
    if (tr > 0.0f) {
        float4 v = ws.smp[path];
        v.x = v.x + si.c.x * tr;
        v.y = v.y + si.c.y * tr;
        v.z = v.z + si.c.z * tr;
        ws.smp[path] = v;
    }
}

// exclusive prefix of nv into offs (one block; offs[n] = total)
__global__ __launch_bounds__(1024) void k_cc_scan(const int* nv, int n, int* offs) {
    __shared__ int part[1024];
    const int per = (n + 1023) / 1024;
    const int b = threadIdx.x * per, e = min(n, b + per);
    int s = 0;
    for (int i = b; i < e; i++) s += nv[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {  // inclusive Hillis-Steele scan
        const int v = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    int run = part[threadIdx.x] - s;
    for (int i = b; i < e; i++) {
        offs[i] = run;
        run += nv[i];
    }
    if (threadIdx.x == 1023) offs[n] = part[1023];
}

// training batch: features and target (radiance gathered after the vertex
// over the throughput that reached it) of every recorded vertex, in path
// order
__global__ __launch_bounds__(kBlock) void k_cc_targets(CacheWs ws, int n, int max_depth) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int k = ws.nv[i], base = ws.offs[i];
    if (k == 0) return;
    const float4 fin = ws.smp[ws.tpix[i]];
    for (int v = 0; v < k; v++) {
        const VertexRec r = ws.vrec[(size_t)i * max_depth + v];
        float* f = ws.tfeat + (size_t)(base + v) * MCPT_MLP_FEATURES;
        for (int i = 0; i < MCPT_MLP_FEATURES; i++) f[i] = r.f[i];
        const float fs[3] = {fin.x, fin.y, fin.z};
        float* t = ws.ttarg + (size_t)(base + v) * 3;
        for (int c = 0; c < 3; c++) t[c] = r.t[c] > 0.0f ? (fs[c] - r.s[c]) / r.t[c] : 0.0f;
    }
}

// finalizeSampleKernel (:3740-3755): the frame's samples of a pixel added
// in sample order, as the plain render adds them
__global__ __launch_bounds__(kBlock) void k_cc_finalize(const float4* smp, float* accum, int npix, int spp) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= npix) return;
    float a0 = accum[3 * p], a1 = accum[3 * p + 1], a2 = accum[3 * p + 2];
    for (int k = 0; k < spp; k++) {
        const float4 s = smp[(size_t)k * npix + p];
        const vec3 c = clamp_sample(vec3(s.x, s.y, s.z));
        a0 = a0 + c.x;
        a1 = a1 + c.y;
        a2 = a2 + c.z;
    }
    accum[3 * p] = a0;
    accum[3 * p + 1] = a1;
    accum[3 * p + 2] = a2;
}

}  // namespace
}  // namespace mcpt

using namespace mcpt;

struct mcpt_mlp_s {
    Ctx* ctx = nullptr;
    DevBuf<float> master, m, v, grad, partial, partial_loss, loss, ema;
    DevBuf<__bf16> wimg, wimg_t, wimg_ema;  // BF16 images of the weights (also transposed) and of their average
    DevBuf<float> io;  // staging for the host-array entry points
    int step = 0;
    // cached-render workspace, grown on demand
    void* ws_mem = nullptr;
    size_t ws_bytes = 0;
    int64_t ws_paths = 0, ws_train = 0;
    int ws_depth = 0;
};

namespace mcpt {
namespace {

// He-uniform initial weights: w = (2u - 1) sqrt(6 / fan_in), u from the
// PCG32 stream keyed by (layer, 0, seed)
void init_params(uint64_t seed, std::vector<float>& p) {
    p.assign(kParams, 0.0f);
    int k = 0;
    for (int l = 0; l <= MCPT_MLP_HIDDEN; l++) {
        const int rows = l < MCPT_MLP_HIDDEN ? kW : MCPT_MLP_OUTPUTS, cols = l == 0 ? kIn : kW;
        mcpt_pcg32 rng;
        mcpt_pcg32_for_path(&rng, (uint32_t)l, 0u, seed);
        const double scale = std::sqrt(6.0 / cols);
        for (int i = 0; i < rows * cols; i++) {
            const float u = (float)((double)mcpt_pcg32_next(&rng) * 0x1p-32);
            p[k++] = (float)((2.0 * (double)u - 1.0) * scale);
        }
    }
}

// gradient of the mean loss over samples [first, first + n) of device
// arrays; result in mlp->grad and mlp->loss
int mlp_gradient_dev(mcpt_mlp_s* mlp, const float* feat, const float* targ, int first, int n, hipStream_t st) {
    const int ntiles = (n + kGradTile - 1) / kGradTile;
    const int nblk = std::max(1, std::min(ntiles, kMaxGradBlocks));
    k_mlp_grad<<<nblk, kGradBlock, 0, st>>>(mlp->wimg.ptr, mlp->wimg_t.ptr, feat, targ, first, n, n, mlp->partial.ptr,
                                            mlp->partial_loss.ptr);
    k_mlp_reduce<<<(kParams + kReduceParams - 1) / kReduceParams, kReduceBlock, 0, st>>>(mlp->partial.ptr, mlp->partial_loss.ptr, nblk,
                                                         mlp->grad.ptr, mlp->loss.ptr, 1.0f / float(3 * n));
    MCPT_HIP(hipGetLastError());
    return MCPT_OK;
}

int mlp_adam(mcpt_mlp_s* mlp, float lr, hipStream_t st, float ema_decay = 0.0f) {
    mlp->step++;
    const float bc1 = (float)(1.0 - std::pow(0.9, mlp->step)), bc2 = (float)(1.0 - std::pow(0.99, mlp->step));
    k_adam<<<blocks_for(kParams), kBlock, 0, st>>>(mlp->master.ptr, mlp->m.ptr, mlp->v.ptr, mlp->grad.ptr,
                                                   mlp->wimg.ptr, mlp->wimg_t.ptr, lr, bc1, bc2, mlp->ema.ptr,
                                                   mlp->wimg_ema.ptr,
                                                   ema_decay);
    MCPT_HIP(hipGetLastError());
    return MCPT_OK;
}

int ensure_io(mcpt_mlp_s* mlp, size_t floats) {
    if (mlp->io.count >= floats) return MCPT_OK;
    return mlp->io.alloc(mlp->ctx, floats);
}

// carve the cached-render workspace for `paths` image paths, `train`
// training paths of up to `depth` bounces
int ensure_cache_ws(mcpt_mlp_s* mlp, int64_t paths, int64_t train, int depth, CacheWs& ws) {
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const int64_t cap = paths;
    const int64_t nrec = train * depth;
    const size_t parts[] = {al(cap * sizeof(RayItem)),      al(cap * sizeof(RayItem)),   al(cap * sizeof(mcpt_packed_hit)),
                            al(cap * sizeof(ShadowItem)),   al(cap * sizeof(float4)),    al(cap * sizeof(float4)),
                            al(cap * sizeof(uint64_t)),     al(cap * sizeof(int)),       al(cap * sizeof(int)),
                            al(train * sizeof(int)),        al(paths * sizeof(Query)),
                            al(nrec * sizeof(VertexRec)),   al(train * sizeof(int)),     al((train + 1) * sizeof(int)),
                            al(nrec * MCPT_MLP_FEATURES * sizeof(float)), al(nrec * 3 * sizeof(float)),
                            al(C_NCOUNTERS * sizeof(int))};
    size_t total = 0;
    for (size_t x : parts) total += x;
    if (mlp->ws_paths < paths || mlp->ws_train < train || mlp->ws_depth < depth) {
        if (mlp->ws_mem) dev_free(mlp->ctx, mlp->ws_mem, mlp->ws_bytes);
        mlp->ws_mem = nullptr;
        mlp->ws_bytes = 0;
        mlp->ws_paths = mlp->ws_train = 0;
        mlp->ws_depth = 0;
        MCPT_TRY(dev_alloc(mlp->ctx, &mlp->ws_mem, total));
        mlp->ws_bytes = total;
        mlp->ws_paths = paths;
        mlp->ws_train = train;
        mlp->ws_depth = depth;
    }
    // carve with the allocated capacities (the layout only depends on them)
    return MCPT_OK;
}

void carve_cache_ws(mcpt_mlp_s* mlp, CacheWs& ws) {
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const int64_t paths = mlp->ws_paths, train = mlp->ws_train;
    const int64_t cap = paths, nrec = train * mlp->ws_depth;
    char* b = (char*)mlp->ws_mem;
    ws.q[0] = (RayItem*)b, b += al(cap * sizeof(RayItem));
    ws.q[1] = (RayItem*)b, b += al(cap * sizeof(RayItem));
    ws.hits = (mcpt_packed_hit*)b, b += al(cap * sizeof(mcpt_packed_hit));
    ws.sq = (ShadowItem*)b, b += al(cap * sizeof(ShadowItem));
    ws.thr = (float4*)b, b += al(cap * sizeof(float4));
    ws.smp = (float4*)b, b += al(cap * sizeof(float4));
    ws.rng = (uint64_t*)b, b += al(cap * sizeof(uint64_t));
    ws.pix = (int*)b, b += al(cap * sizeof(int));
    ws.tslot = (int*)b, b += al(cap * sizeof(int));
    ws.tpix = (int*)b, b += al(train * sizeof(int));
    ws.queries = (Query*)b, b += al(paths * sizeof(Query));
    ws.vrec = (VertexRec*)b, b += al(nrec * sizeof(VertexRec));
    ws.nv = (int*)b, b += al(train * sizeof(int));
    ws.offs = (int*)b, b += al((train + 1) * sizeof(int));
    ws.tfeat = (float*)b, b += al(nrec * MCPT_MLP_FEATURES * sizeof(float));
    ws.ttarg = (float*)b, b += al(nrec * 3 * sizeof(float));
    ws.counters = (int*)b;
}

}  // namespace
}  // namespace mcpt

extern "C" {

int mcpt_mlp_create(mcpt_ctx ctx, uint64_t seed, mcpt_mlp* out) {
    if (!ctx || !out) return set_error(MCPT_ERR_INVALID, "bad arguments");
    *out = nullptr;
    Ctx* c = &ctx->c;
    MCPT_HIP(hipSetDevice(c->device));
    auto* mlp = new mcpt_mlp_s();
    mlp->ctx = c;
    c->children++;  // mcpt_mlp_destroy gives it back, also on the error path below
    int s = MCPT_OK;
    if (s >= 0) s = mlp->master.alloc(c, kParams);
    if (s >= 0) s = mlp->m.alloc(c, kParams);
    if (s >= 0) s = mlp->v.alloc(c, kParams);
    if (s >= 0) s = mlp->grad.alloc(c, kParams);
    if (s >= 0) s = mlp->partial.alloc(c, (size_t)kMaxGradBlocks * kParams);
    if (s >= 0) s = mlp->partial_loss.alloc(c, kMaxGradBlocks);
    if (s >= 0) s = mlp->loss.alloc(c, 4);
    if (s >= 0) s = mlp->wimg.alloc(c, kImg);
    if (s >= 0) s = mlp->wimg_t.alloc(c, kTImg);
    if (s >= 0) s = mlp->ema.alloc(c, kParams);
    if (s >= 0) s = mlp->wimg_ema.alloc(c, kImg);
    if (s < 0) {
        mcpt_mlp_destroy(mlp);
        return s;
    }
    std::vector<float> p;
    init_params(seed, p);
    *out = mlp;
    return mcpt_mlp_set_params(mlp, p.data());
}

int mcpt_mlp_destroy(mcpt_mlp mlp) {
    if (!mlp) return MCPT_OK;
    Ctx* c = mlp->ctx;
    for (DevBuf<float>* b : {&mlp->master, &mlp->m, &mlp->v, &mlp->grad, &mlp->partial, &mlp->partial_loss,
                             &mlp->loss, &mlp->io, &mlp->ema})
        b->release(c);
    mlp->wimg.release(c);
    mlp->wimg_t.release(c);
    mlp->wimg_ema.release(c);
    if (mlp->ws_mem) dev_free(c, mlp->ws_mem, mlp->ws_bytes);
    delete mlp;
    c->children--;
    return MCPT_OK;
}

int mcpt_mlp_get_params(mcpt_mlp mlp, float* params) {
    if (!mlp || !params) return set_error(MCPT_ERR_INVALID, "bad arguments");
    MCPT_HIP(hipStreamSynchronize(mlp->ctx->stream));
    MCPT_HIP(hipMemcpy(params, mlp->master.ptr, kParams * sizeof(float), hipMemcpyDeviceToHost));
    return MCPT_OK;
}

int mcpt_mlp_set_params(mcpt_mlp mlp, const float* params) {
    if (!mlp || !params) return set_error(MCPT_ERR_INVALID, "bad arguments");
    hipStream_t st = mlp->ctx->stream;
    MCPT_HIP(hipMemcpyAsync(mlp->master.ptr, params, kParams * sizeof(float), hipMemcpyHostToDevice, st));
    MCPT_HIP(hipMemsetAsync(mlp->m.ptr, 0, kParams * sizeof(float), st));
    MCPT_HIP(hipMemsetAsync(mlp->v.ptr, 0, kParams * sizeof(float), st));
    k_pack<<<blocks_for(kImg), kBlock, 0, st>>>(mlp->master.ptr, mlp->wimg.ptr);
    k_pack_t<<<blocks_for(kTImg), kBlock, 0, st>>>(mlp->master.ptr, mlp->wimg_t.ptr);
    MCPT_HIP(hipMemcpyAsync(mlp->ema.ptr, mlp->master.ptr, kParams * sizeof(float), hipMemcpyDeviceToDevice, st));
    MCPT_HIP(hipMemcpyAsync(mlp->wimg_ema.ptr, mlp->wimg.ptr, kImg * sizeof(__bf16), hipMemcpyDeviceToDevice, st));
    MCPT_HIP(hipGetLastError());
    MCPT_HIP(hipStreamSynchronize(st));
    mlp->step = 0;
    return MCPT_OK;
}

int mcpt_mlp_forward(mcpt_mlp mlp, const float* features, int32_t n, float* rgb) {
    if (!mlp || n < 0 || (n > 0 && (!features || !rgb))) return set_error(MCPT_ERR_INVALID, "bad arguments");
    if (n == 0) return MCPT_OK;
    hipStream_t st = mlp->ctx->stream;
    // queries (16 floats each) then outputs (3 each)
    MCPT_TRY(ensure_io(mlp, (size_t)n * 19));
    std::vector<Query> q(n);
    for (int i = 0; i < n; i++) {
        std::memset(&q[i], 0, sizeof(Query));
        std::memcpy(q[i].f, features + (size_t)i * MCPT_MLP_FEATURES, sizeof(q[i].f));
    }
    Query* dq = reinterpret_cast<Query*>(mlp->io.ptr);
    float* dy = mlp->io.ptr + (size_t)n * 16;
    MCPT_HIP(hipMemcpyAsync(dq, q.data(), (size_t)n * sizeof(Query), hipMemcpyHostToDevice, st));
    k_mlp_infer<false><<<infer_blocks(mlp->ctx, n), kInferBlock, 0, st>>>(mlp->wimg.ptr, dq, nullptr, n, dy, nullptr);
    MCPT_HIP(hipGetLastError());
    MCPT_HIP(hipMemcpyAsync(rgb, dy, (size_t)n * 3 * sizeof(float), hipMemcpyDeviceToHost, st));
    MCPT_HIP(hipStreamSynchronize(st));
    return MCPT_OK;
}

int mcpt_mlp_gradient(mcpt_mlp mlp, const float* features, const float* targets, int32_t n, float* loss,
                      float* grad) {
    if (!mlp || n < 1 || !features || !targets) return set_error(MCPT_ERR_INVALID, "bad arguments");
    hipStream_t st = mlp->ctx->stream;
    MCPT_TRY(ensure_io(mlp, (size_t)n * 14));
    float* df = mlp->io.ptr;
    float* dt = df + (size_t)n * MCPT_MLP_FEATURES;
    MCPT_HIP(hipMemcpyAsync(df, features, (size_t)n * MCPT_MLP_FEATURES * sizeof(float), hipMemcpyHostToDevice, st));
    MCPT_HIP(hipMemcpyAsync(dt, targets, (size_t)n * 3 * sizeof(float), hipMemcpyHostToDevice, st));
    MCPT_TRY(mlp_gradient_dev(mlp, df, dt, 0, n, st));
    if (grad) MCPT_HIP(hipMemcpyAsync(grad, mlp->grad.ptr, kParams * sizeof(float), hipMemcpyDeviceToHost, st));
    float l = 0.0f;
    MCPT_HIP(hipMemcpyAsync(&l, mlp->loss.ptr, sizeof(float), hipMemcpyDeviceToHost, st));
    MCPT_HIP(hipStreamSynchronize(st));
    if (loss) *loss = l;
    return MCPT_OK;
}

int mcpt_mlp_train(mcpt_mlp mlp, const float* features, const float* targets, int32_t n, int32_t batch, float lr,
                   float* loss) {
    if (!mlp || n < 1 || batch < 1 || !features || !targets || !(lr > 0.0f))
        return set_error(MCPT_ERR_INVALID, "bad arguments");
    hipStream_t st = mlp->ctx->stream;
    MCPT_TRY(ensure_io(mlp, (size_t)n * 14));
    float* df = mlp->io.ptr;
    float* dt = df + (size_t)n * MCPT_MLP_FEATURES;
    MCPT_HIP(hipMemcpyAsync(df, features, (size_t)n * MCPT_MLP_FEATURES * sizeof(float), hipMemcpyHostToDevice, st));
    MCPT_HIP(hipMemcpyAsync(dt, targets, (size_t)n * 3 * sizeof(float), hipMemcpyHostToDevice, st));
    for (int first = 0; first < n; first += batch) {
        const int cnt = std::min(batch, n - first);
        MCPT_TRY(mlp_gradient_dev(mlp, df, dt, first, cnt, st));
        MCPT_TRY(mlp_adam(mlp, lr, st));
    }
    float l = 0.0f;
    MCPT_HIP(hipMemcpyAsync(&l, mlp->loss.ptr, sizeof(float), hipMemcpyDeviceToHost, st));
    MCPT_HIP(hipStreamSynchronize(st));
    if (loss) *loss = l;
    return MCPT_OK;
}

int mcpt_render_cached(mcpt_scene scene, const mcpt_camera* cam, const mcpt_render_opts* o, mcpt_film film,
                       mcpt_mlp mlp, const mcpt_cache_opts* co, mcpt_cache_stats* stats) {
    if (!scene || !cam || !o || !film || !mlp) return set_error(MCPT_ERR_INVALID, "bad arguments");
    if (!scene->built) return set_error(MCPT_ERR_INVALID, "scene has no BVH (call mcpt_lbvh_build)");
    MCPT_TRY(scene->check_textures());
    if (film->w != cam->width || film->h != cam->height)
        return set_error(MCPT_ERR_INVALID, "film %dx%d does not match camera %dx%d", film->w, film->h, cam->width,
                         cam->height);
    if (o->mode == MCPT_MODE_MEGAKERNEL) return set_error(MCPT_ERR_INVALID, "the cached render is wavefront only");
    if (o->spp < 1 || o->max_depth < 1 || o->sample_begin < 0 || o->sample_begin + o->spp > 0x40000000)
        return set_error(MCPT_ERR_INVALID, "spp, max_depth or sample_begin out of range");
    if (scene->ctx != mlp->ctx) return set_error(MCPT_ERR_INVALID, "scene and MLP belong to different contexts");
    const int npix = cam->width * cam->height, depth = o->max_depth;
    const int fspp = std::min(o->spp, (co && co->frame_spp > 0) ? co->frame_spp : 1);
    if ((int64_t)npix * fspp > (1 << 28)) return set_error(MCPT_ERR_INVALID, "frame_spp x pixels exceeds 256 Mi paths");
    const int train_paths = std::min(npix * fspp, (co && co->train_paths > 0) ? co->train_paths : 65536);
    const int batch = (co && co->train_batch > 0) ? co->train_batch : 16384;
    const int max_steps = (co && co->train_steps > 0) ? co->train_steps : 4;
    const int term_depth = (co && co->terminate_depth > 0) ? co->terminate_depth : 1;
    const float lr = (co && co->lr > 0.0f) ? co->lr : 0.005f;
    const float decay = (co && co->ema_decay > 0.0f) ? co->ema_decay : 0.0f;
    if (!(decay < 1.0f)) return set_error(MCPT_ERR_INVALID, "ema_decay must be < 1");
    if (batch > (1 << 24)) return set_error(MCPT_ERR_INVALID, "train_batch must be <= 16 Mi");
    Ctx* ctx = scene->ctx;
    hipStream_t st = ctx->stream;
    CacheWs ws;
    MCPT_TRY(ensure_cache_ws(mlp, (int64_t)npix * fspp, train_paths, depth, ws));
    carve_cache_ws(mlp, ws);
    const SceneView sv = scene->view();
    const size_t lds = (size_t)kBlock * stack_entries(scene->max_depth) * sizeof(int);
    const int refill = o->refill_lanes > 0 ? std::min(64, o->refill_lanes) : 32;
    int per_cu = 0;
    MCPT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_cc_extend, kBlock, lds));
    const int persistent_blocks = std::max(1, per_cu) * ctx->cu_count;
    // short queues: 16 lanes per ray (mcpt_render's default threshold)
    const int gb = o->group_below < 0 ? 0 : o->group_below > 0 ? o->group_below : 32768;
    const int gcap_phys = group_stack_pairs(stack_entries(scene->max_depth));
    const size_t glds = (size_t)(kBlock / kGroup) * group_lds_ints(gcap_phys) * sizeof(int);
    int group_blocks = 0;
    if (gb > 0) {
        MCPT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_cc_extend_group, kBlock, glds));
        group_blocks = std::min(std::max(1, per_cu) * ctx->cu_count, (gb + kBlock / kGroup - 1) / (kBlock / kGroup));
    }
    Frame fr;
    {
        float b[6];
        MCPT_HIP(hipMemcpy(b, scene->root_box.ptr, sizeof(b), hipMemcpyDeviceToHost));
        for (int k = 0; k < 3; k++) {
            fr.bmin[k] = b[k];
            fr.bext[k] = b[3 + k] - b[k] > 0.0f ? b[3 + k] - b[k] : 1.0f;
        }
    }
    fr.max_depth = depth;
    const vec3 bg = load3(cam->background);
    mcpt_cache_stats acc;
    std::memset(&acc, 0, sizeof(acc));
    hipEvent_t ev[6];
    for (auto& e : ev) MCPT_HIP(hipEventCreate(&e));
    auto elapsed = [&](hipEvent_t a, hipEvent_t b) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, a, b);
        return (double)ms;
    };
    for (int f0 = 0; f0 < o->spp; f0 += fspp) {
        const int fs = std::min(fspp, o->spp - f0);  // samples in this frame
        const int npaths = npix * fs;
        const int grid = blocks_for(npaths);
        const uint32_t sample = (uint32_t)(o->sample_begin + f0);
        // a network that has not been trained yet is not queried: the frame
        // is traced without the cache
        fr.term_depth = mlp->step > 0 ? term_depth : depth;
        MCPT_HIP(hipEventRecord(ev[0], st));
        k_cc_camera<<<grid, kBlock, 0, st>>>(*cam, sample, fs, o->seed, ws);
        k_cc_pick<<<blocks_for(train_paths), kBlock, 0, st>>>(sample, o->seed, train_paths, npaths, ws);
        int cur = 0;
        for (int d = 0; d < depth; d++) {
            fr.depth = d;
            k_cc_bounce<<<1, 64, 0, st>>>(ws.counters, cur);
            k_cc_extend<<<std::min(grid, persistent_blocks), kBlock, lds, st>>>(sv, ws, cur, refill, gb);
            if (gb > 0)
                k_cc_extend_group<<<std::min(grid, group_blocks), kBlock, glds, st>>>(sv, ws, cur, gb, kGroupCap,
                                                                                    gcap_phys);
            k_cc_shade<<<grid, kBlock, 0, st>>>(sv, ws, cur, fr, bg, cam->rr_cap);
            k_cc_shadow<<<grid, kBlock, lds, st>>>(sv, ws, cur);
            cur = 1 - cur;
        }
        MCPT_HIP(hipEventRecord(ev[1], st));
        k_mlp_infer<true><<<infer_blocks(ctx, npaths), kInferBlock, 0, st>>>(
            decay > 0.0f ? mlp->wimg_ema.ptr : mlp->wimg.ptr, ws.queries, ws.counters + C_QCOUNT, 0, nullptr,
            ws.smp);
        MCPT_HIP(hipEventRecord(ev[2], st));
        k_cc_finalize<<<blocks_for(npix), kBlock, 0, st>>>(ws.smp, film->accum.ptr, npix, fs);
        // training samples of this frame's recorded paths, for the next frame
        k_cc_scan<<<1, 1024, 0, st>>>(ws.nv, train_paths, ws.offs);
        k_cc_targets<<<blocks_for(train_paths), kBlock, 0, st>>>(ws, train_paths, depth);
        MCPT_HIP(hipGetLastError());
        int n_train = 0, nq = 0;
        MCPT_HIP(hipMemcpyAsync(&n_train, ws.offs + train_paths, sizeof(int), hipMemcpyDeviceToHost, st));
        MCPT_HIP(hipMemcpyAsync(&nq, ws.counters + C_QCOUNT, sizeof(int), hipMemcpyDeviceToHost, st));
        MCPT_HIP(hipEventRecord(ev[3], st));
        MCPT_HIP(hipEventSynchronize(ev[3]));
        const int steps = std::min(max_steps, (n_train + batch - 1) / batch);
        for (int k = 0; k < steps; k++) {
            const int first = k * batch, cnt = std::min(batch, n_train - first);
            MCPT_TRY(mlp_gradient_dev(mlp, ws.tfeat, ws.ttarg, first, cnt, st));
            MCPT_TRY(mlp_adam(mlp, lr, st, decay));
        }
        MCPT_HIP(hipEventRecord(ev[4], st));
        MCPT_HIP(hipEventSynchronize(ev[4]));
        float l = 0.f;
        if (steps) MCPT_HIP(hipMemcpy(&l, mlp->loss.ptr, sizeof(float), hipMemcpyDeviceToHost));
        acc.render_ms += elapsed(ev[0], ev[1]);
        acc.train_trace_ms += elapsed(ev[2], ev[3]);
        acc.infer_ms += elapsed(ev[1], ev[2]);
        acc.train_ms += elapsed(ev[3], ev[4]);
        acc.total_ms += elapsed(ev[0], ev[4]);
        acc.train_samples += (uint64_t)std::min(n_train, steps * batch);
        acc.queries += (uint64_t)nq;
        acc.paths += (uint64_t)npaths;
        if (steps) acc.loss = l;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    if (stats) *stats = acc;
    return MCPT_OK;
}

}  // extern "C"
