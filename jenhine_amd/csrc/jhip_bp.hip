// Voxel-driven normalised backprojection for gfx950.
//
// Semantics (BPAgent/BPCore.cpp:108-205): for every voxel
//   vol = sum_v bilinear(q_v, su, sv) / sum_v bilinear(uW_v, su, sv),
//   su = U/W + 0.5, sv = V/W + 0.5, (U, V, W) = P_v (x, y, z, 1),
// where q_v = filtered_v * uW_v row-wise and a sample whose coordinate falls
// outside [0, W-1] x [0, V-1] contributes to neither sum.
//
// Design:
//  * per-view geometry is set up on the device in FP64 and rounded once to
//    float coefficients in voxel-index space;
//  * each thread owns one (x, y) voxel column and ZC consecutive z voxels,
//    keeping ZC accumulators and ZC counts in registers across all views of
//    a launch, so the volume is written once per launch (one-shot) or
//    read-modify-written once per view chunk (streaming);
//  * projections are staged so that one bilinear sample is one aligned
//    gather: FP32 quads (16 B, a pixel's 2 x 2 neighbourhood), BF16 / FP16
//    quads (8 B) or, for geometry class 0, BF16 / FP16 column strips (8 B =
//    four rows of one column), where one 16-byte gather serves two
//    consecutive z samples when the row step is below 2 pixels (the staging
//    section below gives the layouts);
//  * when u and w do not depend on z (circular orbits) the u-side of the
//    bilinear weights and the count weight are hoisted out of the z loop;
//  * 64 x TY voxel columns per workgroup, workgroups remapped so that each of
//    the 8 XCDs owns a contiguous z-range of tiles (its L2 then holds the
//    detector rows of that z-range).
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "jhip_internal.h"

#include "jhip_bp_plan.h"

namespace jh {

// ---------------------------------------------------------------------------
// Per-view set-up, FP64.
// ---------------------------------------------------------------------------

// Index-space coefficients: U(i,j,k) = cu0 + cui*i + cuj*j + cuk*k with the
// centred axes a[i] = (i - (n-1)/2) * pitch of BPCore.cpp:47-66.
__global__ void bp_geometry_kernel(const double* __restrict__ ptm, int N, int X, int Y, int Z,
                                   double pxy, double pz, Geo* __restrict__ geo)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    const double* P = ptm + 12 * (size_t)v;
    const double ax0 = -((double)X - 1.0) * 0.5 * pxy;
    const double ay0 = -((double)Y - 1.0) * 0.5 * pxy;
    const double az0 = -((double)Z - 1.0) * 0.5 * pz;
    Geo g;
    for (int r = 0; r < 3; ++r) {
        const double* p = P + 4 * r;
        g.c[4 * r + 0] = (float)(p[0] * ax0 + p[1] * ay0 + p[2] * az0 + p[3]);
        g.c[4 * r + 1] = (float)(p[0] * pxy);
        g.c[4 * r + 2] = (float)(p[1] * pxy);
        g.c[4 * r + 3] = (float)(p[2] * pz);
    }
    geo[v] = g;
}

// Detector weight uW across u (BPCore.cpp:68-100), FP64 then rounded.
__global__ void bp_uweight_kernel(const double* __restrict__ ptm, int N, int W, float* __restrict__ uw)
{
    const int v = blockIdx.y;
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N || u >= W) return;
    const double* P = ptm + 12 * (size_t)v;
    const double uu = P[3] / P[11];
    const long u0 = uu > 0.0 ? (long)uu : 0;
    const double pi = 3.14159265358979323846;
    double w = 1.0;
    if (u0 > (long)(W / 2)) {
        const long nT = u0 < (long)W ? ((long)W - u0) * 2 : 0;
        const long i = u - (W - nT);
        if (i >= 0 && i < nT) w = cos((double)i * pi / (double)nT) * 0.5 + 0.5;
    } else {
        long nT = u0 * 2;
        if (nT > W) nT = W;
        if (u < nT) w = cos((double)u * pi / (double)nT) * -0.5 + 0.5;
    }
    uw[(size_t)v * W + u] = (float)w;
}

// ---------------------------------------------------------------------------
// Staging: q = filtered * uW in the backprojector's layout.
//
// Frame layout: V rows of pixels; the data a bilinear sample whose integer
// part is (u, v) needs is one aligned gather, with q[v][W] = 0 and
// q[V][u] = 0 (so a sample at u <= W-1, v <= V-1 needs no clamping: at the
// far edges the weight of the padding is exactly 0).  A gather's cost in
// the texture path follows the cache lines each wave instruction touches more
// than its bytes.
//  * FP32 (16 B): (q[v][u], q[v][u+1] - q[v][u], q[v+1][u], q[v+1][u+1] -
//    q[v+1][u]).  Each row interpolation is one FMA, q + fu d, and d is the
//    difference the interpolation would form anyway (same result to the bit).
//  * BF16 / FP16 quad (8 B): (q[v][u], q[v][u+1], q[v+1][u], q[v+1][u+1]),
//    each value rounded once.
//  * BF16 / FP16 column strip (8 B, the default for geometry class 0): the
//    four rows q[v..v+3][u] of column u, rows padded to W + 1 pixels.  A
//    16-byte gather at (u, v) reads columns u and u+1 of four rows: the
//    sample's 2 x 2 neighbourhood, and that of the next z sample of the
//    column when the row step is below 2 pixels (the backprojector's pair
//    path), which halves the gathers; neighbouring lanes' windows overlap,
//    so a wave touches fewer cache lines than with 16-byte-per-pixel data.
//    Storing differences instead makes the rounding error scale with |d|,
//    which exceeds |q| on ramp-filtered noise, and a BF16 weight for v_dot2
//    would move the sample by up to 2^-9 pixel; both measurably raise the
//    error.
// ---------------------------------------------------------------------------

__device__ inline uint32_t f32_to_bf16_rne(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (u >> 16) | 0x40u;  // quiet NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return u >> 16;
}

typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));

__device__ inline uint32_t pack_f16x2(float a, float b)
{
    // a and b are FP32 values, each rounded once more here: the empty asm
    // keeps the compiler from folding the product that made them into a
    // mixed-precision multiply that rounds straight to FP16
    asm volatile("" : "+v"(a), "+v"(b));
    half2_t h;
    h.x = (_Float16)a;
    h.y = (_Float16)b;
    return __builtin_bit_cast(uint32_t, h);
}

// ZFilterAgent's 1-2-1 smoothing along v (FilterAgent/ZFilterAgent.cpp:49-71)
// of column u at row r of one frame, in the reference's operation order:
// (0.5 in[r] + 0.25 in[max(r-1, 0)]) + 0.25 in[min(r+1, V-1)].
__device__ __forceinline__ float zfilter_at(const float* __restrict__ frame, int W, int V, int r, int u)
{
#pragma clang fp contract(off)
    const float a = frame[(size_t)r * W + u];
    const float b = frame[(size_t)(r > 0 ? r - 1 : 0) * W + u];
    const float c = frame[(size_t)(r + 1 < V ? r + 1 : V - 1) * W + u];
    const float t = a * 0.5f + b * 0.25f;
    return t + c * 0.25f;
}

// CTNumCore::Process (CTNumAgent/CTNumCore.cpp:22-36): x * scale - 1000 with
// scale = norm0 * 1000 / muWater formed on the host in float as the reference
// does; above 300 the value becomes min(8191, (x - 300) norm1 + 300).
__device__ __forceinline__ float ct_number(float x, float scale, float norm1)
{
#pragma clang fp contract(off)
    float y = x * scale;
    y = y - 1000.0f;
    if (y > 300.0f) {
        float z = (y - 300.0f) * norm1;
        z = z + 300.0f;
        y = fminf(8191.0f, z);
    }
    return y;
}

__global__ void __launch_bounds__(256) zfilter_kernel(const float* __restrict__ in, float* __restrict__ out, int W,
                                                      int V, size_t total)
{
    const size_t frame = (size_t)W * V;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const size_t f = t / frame;
        const int pix = (int)(t - f * frame);
        const int r = pix / W;
        out[t] = zfilter_at(in + f * frame, W, V, r, pix - r * W);
    }
}

// One thread per column and run of kStageRun staged rows:
// blockIdx.x * 64 + threadIdx.x = u, (blockIdx.y * 4 + threadIdx.y) *
// kStageRun = first row, blockIdx.z = frame.  Layouts: FP32 quads (16 B),
// BF16 / FP16 quads (8 B) or, with strip, BF16 / FP16 column strips (8 B, rows
// padded to W + 1 pixels; thread u = W - 1 also writes the zero padding).
// The thread slides a window of input rows down its run in registers, so an
// input row is loaded (kStageRun + NR + 1) / kStageRun times (NR = 2 for
// quads, 4 for strips) rather than NR + 2 times; every staged value is the
// same function of the same inputs as with one thread per staged pixel.
constexpr int kStageRun = 16;

template <int QT, bool strip>
__global__ void __launch_bounds__(256) bp_stage_kernel(const float* __restrict__ filt,
                                                       const float* __restrict__ uw, int W, int V, int zf,
                                                       float qs, void* __restrict__ out)
// uw == NULL packs the frames without weighting; zf != 0 applies the v
// smoothing of ZFilterAgent to the frames first (before the uW weight, as in
// the reference chain Filter -> ZFilter -> BPCore::InitView).  qs (a power of
// two) scales the staged values; uW * qs and x * (uW qs) = (x uW) qs are
// exact in the normal range, so the values differ from qs = 1 by qs alone.
{
    // every staged value is rounded as written: q = RN(x uW), and the FP32
    // quad's differences RN(q[u+1] - q[u]) are not fused with the products
#pragma clang fp contract(off)
    static_assert(!(strip && QT == JH_F32), "FP32 is staged as quads");
    constexpr int NR = strip ? 4 : 2;          // rows of the quad or the strip
    constexpr int NX = kStageRun + NR + 1;     // window: rows row0-1 .. row0+kStageRun+NR-1
    const int u = blockIdx.x * 64 + threadIdx.x;
    const int row0 = (blockIdx.y * 4 + threadIdx.y) * kStageRun;
    if (u >= W || row0 >= V) return;
    const int rowpx = W + (strip ? 1 : 0);
    const size_t view = blockIdx.z;
    const float* fr = filt + view * (size_t)W * V;
    const float* w = uw ? uw + view * W : nullptr;
    const bool right = u + 1 < W;
    const int u1 = right ? u + 1 : u;
    const float wa = (w ? w[u] : 1.0f) * qs;
    const float wb = (w && right ? w[u + 1] : 1.0f) * qs;
    // rows row0-1+j of columns u (and u+1 for quads), clamped into the frame
    float xa[NX], xb[NX];
#pragma unroll
    for (int j = 0; j < NX; ++j) {
        const int r = min(max(row0 - 1 + j, 0), V - 1);
        xa[j] = fr[(size_t)r * W + u];
        xb[j] = strip ? 0.0f : fr[(size_t)r * W + u1];
    }
    // value of row row0+i at column u or u+1, 0 beyond the frame
    auto val = [&](const float* x, int i, float wgt, bool inside) {
        if (row0 + i >= V || !inside) return 0.0f;
        float v = x[i + 1];
        if (zf) {
#pragma clang fp contract(off)
            // ZFilterAgent.cpp:49-71, as zfilter_at: the clamped neighbours
            // are the clamped loads above
            const float tt = v * 0.5f + x[i] * 0.25f;
            v = tt + x[i + 2] * 0.25f;
        }
        return v * wgt;
    };
    auto pack = [&](float lo, float hi) {
        return QT == JH_BF16 ? (f32_to_bf16_rne(lo) | (f32_to_bf16_rne(hi) << 16)) : pack_f16x2(lo, hi);
    };
    size_t t = (view * V + row0) * (size_t)rowpx + u;
#pragma unroll
    for (int r = 0; r < kStageRun; ++r) {
        if (row0 + r < V) {
            if constexpr (QT == JH_F32) {
                const float a0 = val(xa, r, wa, true), b0 = val(xb, r, wb, right);
                const float a1 = val(xa, r + 1, wa, true), b1 = val(xb, r + 1, wb, right);
                reinterpret_cast<float4*>(out)[t] = make_float4(a0, b0 - a0, a1, b1 - a1);
            } else if constexpr (strip) {
                // column u, rows row .. row+3
                u32x2_t o;
                o.x = pack(val(xa, r, wa, true), val(xa, r + 1, wa, true));
                o.y = pack(val(xa, r + 2, wa, true), val(xa, r + 3, wa, true));
                reinterpret_cast<u32x2_t*>(out)[t] = o;
                if (u == W - 1) reinterpret_cast<u32x2_t*>(out)[t + 1] = u32x2_t{0u, 0u};
            } else {
                // rows row, row+1 of columns u, u+1
                u32x2_t o;
                o.x = pack(val(xa, r, wa, true), val(xb, r, wb, right));
                o.y = pack(val(xa, r + 1, wa, true), val(xb, r + 1, wb, right));
                reinterpret_cast<u32x2_t*>(out)[t] = o;
            }
        }
        t += rowpx;
    }
}

// Count-weight pairs (uW[u], uW[u+1]) per view; uW[W] := uW[W-1] (its weight
// is always 0).
__global__ void bp_uwpair_kernel(const float* __restrict__ uw, int N, int W, float2* __restrict__ uwp)
{
    const size_t n = (size_t)N * W;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
        const int u = (int)(t % W);
        uwp[t] = make_float2(uw[t], u + 1 < W ? uw[t + 1] : uw[t]);
    }
}

// ---------------------------------------------------------------------------
// Backprojection kernels.
// ---------------------------------------------------------------------------

constexpr int kBufferFlags = 0x00020000;  // raw buffer, 32-bit dwords (gfx9 family)

// One quad load per sample; row r (0 = top, 1 = bottom) interpolated along u
// with the weight w of the sample's column (weight()).
//
// The BF16 and FP16 samplers interpolate a row with one v_dot2_f32_bf16 /
// v_dot2_f32_f16 of its packed pair (q[u], q[u+1]) and the packed weights
// (1 - w, w), accumulated in FP32.  w is the fraction fu rounded to a
// multiple of 1/256 (BF16) or 1/2048 (FP16): both weights are then exact in
// the 16-bit type and sum to one, and the position along u is off by at most
// 1/512 (1/4096) pixel -- the reference's hardware bilinear texture fetch
// (cudaFilterModeLinear) quantises both fractions to 1/256.  The FP32 sampler
// uses fu itself.  weight() also returns the fraction it applies, for the
// count weight, so that numerator and count interpolate alike.
template <int QT>
struct Sampler;

template <>
struct Sampler<JH_F32> {
    typedef u32x4_t raw_t;
    typedef float wt_t;
    static constexpr int bytes = 16;   // bytes per sample load
    __device__ static raw_t load(__amdgpu_buffer_rsrc_t r, uint32_t off)
    {
        return __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
    }
    __device__ static wt_t weight(float fu, float& fe)
    {
        fe = fu;
        return fu;
    }
    // (q, d): q + fu d
    __device__ static float lerp(raw_t p, int r, wt_t fu)
    {
        return r == 0 ? fmaf(fu, __uint_as_float(p.y), __uint_as_float(p.x))
                      : fmaf(fu, __uint_as_float(p.w), __uint_as_float(p.z));
    }
    // acc + the sample at row fraction fv
    __device__ static float accum(float acc, raw_t p, float fv, wt_t fu)
    {
        const float top = lerp(p, 0, fu), bot = lerp(p, 1, fu);
        return acc + fmaf(fv, bot - top, top);
    }
    // the sample alone (the distance-weighted form scales it before adding)
    __device__ static float sample(raw_t p, float fv, wt_t fu)
    {
        const float top = lerp(p, 0, fu), bot = lerp(p, 1, fu);
        return fmaf(fv, bot - top, top);
    }
    // bilinear coefficients (q00, q01 - q00, q10 - q00, (q11 - q10) - (q01 - q00))
    __device__ static float4 abcd(raw_t p)
    {
        const float a = __uint_as_float(p.x), b = __uint_as_float(p.y);
        return make_float4(a, b, __uint_as_float(p.z) - a, __uint_as_float(p.w) - b);
    }
};

// Packed-pair arithmetic of the 16-bit storage types.
// Row interpolation of the 16-bit storage types, one v_dot2 per row pair
// (q[u], q[u+1]), accumulated in FP32.  Half16<QT>::Wt is the weight of a
// sample's column, formed once per view; accum() adds one sample interpolated
// from its top and bottom row pairs.
template <int QT>
struct Half16;

// BF16: the row is interpolated from its nearer column,
//   q[u] + g (q[u+1] - q[u])   (fu <= 1/2, g = fu), or
//   q[u+1] + g (q[u] - q[u+1]) (fu > 1/2, g = 1 - fu),
// as one dot of the pair with (-g, g) (or (g, -g)) accumulated onto the nearer
// value, which is exact in FP32.  g is rounded to BF16 (relative error
// 2^-9, g <= 1/2): the position along u is off by at most 2^-10 pixel.
template <>
struct Half16<JH_BF16> {
    struct Wt {
        uint32_t w;    // (weight of q[u], weight of q[u+1]), BF16 pair
        uint32_t sel;  // v_perm selector moving the nearer value's BF16 bits into an FP32 word
    };
    // x.lo w.lo + x.hi w.hi + c
    __device__ static float dot(uint32_t x, uint32_t w, float c)
    {
        return __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2_t, x), __builtin_bit_cast(bf16x2_t, w), c,
                                               false);
    }
    __device__ static Wt weight(float fu, float& fe)
    {
        const bool far = fu > 0.5f;
        const float g = far ? 1.0f - fu : fu;  // exact, in [0, 1/2]: finite
        const uint32_t gu = __float_as_uint(g);
        const uint32_t gb = (gu + 0x7fffu + ((gu >> 16) & 1u)) >> 16;  // round to nearest even
        const float ge = __uint_as_float(gb << 16);
        fe = far ? 1.0f - ge : ge;
        const uint32_t neg = gb | 0x8000u;
        Wt r;
        r.w = far ? (gb | (neg << 16)) : (neg | (gb << 16));
        r.sel = far ? 0x03020c0cu : 0x01000c0cu;
        return r;
    }
    __device__ static float row(uint32_t x, Wt w)
    {
        return dot(x, w.w, __uint_as_float(__builtin_amdgcn_perm(x, x, w.sel)));
    }
    // acc + (top + fv (bot - top)) for the row pairs rt (top) and rb (bottom)
    __device__ static float accum(float acc, uint32_t rt, uint32_t rb, Wt w, float fv)
    {
        const float top = row(rt, w), bot = row(rb, w);
        return acc + fmaf(fv, bot - top, top);
    }
    __device__ static float sample(uint32_t rt, uint32_t rb, Wt w, float fv)
    {
        const float top = row(rt, w), bot = row(rb, w);
        return fmaf(fv, bot - top, top);
    }
    __device__ static float lo(uint32_t w) { return __uint_as_float(w << 16); }
    __device__ static float hi(uint32_t w) { return __uint_as_float(w & 0xffff0000u); }
};

// FP16: the pair's weights are (1 - w, w), w the fraction rounded to a
// multiple of 1/2048, so both are exact in FP16 and sum to one; the position
// along u is off by at most 1/4096 pixel.  bot - top is one dot of the bottom
// pair with (1 - w, w) accumulated into one of the top pair with -(1 - w, w),
// and top is accumulated into acc directly (the accumulator is rounded twice
// per sample).
template <>
struct Half16<JH_F16> {
    struct Wt {
        uint32_t w, nw;  // (1 - w, w) and its negation, FP16 pairs
    };
    __device__ static float dot(uint32_t x, uint32_t w, float c)
    {
        return __builtin_amdgcn_fdot2(__builtin_bit_cast(half2_t, x), __builtin_bit_cast(half2_t, w), c, false);
    }
    __device__ static Wt weight(float fu, float& fe)
    {
        const float w = __builtin_rintf(fu * 2048.0f) * (1.0f / 2048.0f);
        fe = w;
        const uint32_t p = __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(1.0f - w, w));
        return Wt{p, p ^ 0x80008000u};
    }
    __device__ static float accum(float acc, uint32_t rt, uint32_t rb, Wt w, float fv)
    {
        const float d = dot(rt, w.nw, dot(rb, w.w, 0.0f));
        return fmaf(fv, d, dot(rt, w.w, acc));
    }
    __device__ static float sample(uint32_t rt, uint32_t rb, Wt w, float fv)
    {
        const float d = dot(rt, w.nw, dot(rb, w.w, 0.0f));
        return fmaf(fv, d, dot(rt, w.w, 0.0f));
    }
    __device__ static float lo(uint32_t w) { return (float)__builtin_bit_cast(half2_t, w).x; }
    __device__ static float hi(uint32_t w) { return (float)__builtin_bit_cast(half2_t, w).y; }
};

// BF16 / FP16 quads: staged pixel (u, v) holds the row pairs
// (q[v][u], q[v][u+1]) and (q[v+1][u], q[v+1][u+1]), low half first.
template <int QT>
struct QuadSampler {
    typedef Half16<QT> H;
    typedef u32x2_t raw_t;
    typedef typename H::Wt wt_t;
    static constexpr int bytes = 8;
    __device__ static raw_t load(__amdgpu_buffer_rsrc_t r, uint32_t off)
    {
        return __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
    }
    __device__ static wt_t weight(float fu, float& fe) { return H::weight(fu, fe); }
    __device__ static float accum(float acc, raw_t p, float fv, wt_t w)
    {
        // the words are copied out of the vector first: a bit cast taken
        // directly on the element p.y was compiled as one of p.x
        const uint32_t px = p.x, py = p.y;
        return H::accum(acc, px, py, w, fv);
    }
    __device__ static float sample(raw_t p, float fv, wt_t w)
    {
        const uint32_t px = p.x, py = p.y;
        return H::sample(px, py, w, fv);
    }
    __device__ static float4 abcd(raw_t p)
    {
        const uint32_t px = p.x, py = p.y;
        const float q00 = H::lo(px), q01 = H::hi(px), q10 = H::lo(py), q11 = H::hi(py);
        const float b = q01 - q00;
        return make_float4(q00, b, q10 - q00, (q11 - q10) - b);
    }
};
template <>
struct Sampler<JH_BF16> : QuadSampler<JH_BF16> {};
template <>
struct Sampler<JH_F16> : QuadSampler<JH_F16> {};

// Column strips (BF16 / FP16): staged pixel (u, v) holds the four rows
// v .. v+3 of column u as two 32-bit words, w0 = (q[v][u], q[v+1][u]) and
// w1 = (q[v+2][u], q[v+3][u]) (low half first), 8 bytes; rows are padded to
// W + 1 pixels (the padding column is 0).  One 16-byte gather at (u, v)
// returns p = (w0(u), w1(u), w0(u+1), w1(u+1)): columns u and u+1 of four
// rows.  Row r's pair (q[v+r][u], q[v+r][u+1]) is one v_perm_b32 of a word
// of each column; it is the word the quad layout stores for that row, so the
// two layouts interpolate identical operands.
constexpr int kBF16Strip = 4;  // sampler ids of the strip layout (kernel template argument QT)
constexpr int kF16Strip = 5;

template <int QT>
struct StripSampler {
    typedef Half16<QT == kBF16Strip ? JH_BF16 : JH_F16> H;
    typedef u32x4_t raw_t;
    typedef typename H::Wt wt_t;
    static constexpr int bytes = 16;
    __device__ static raw_t load(__amdgpu_buffer_rsrc_t r, uint32_t off)
    {
        return __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
    }
    __device__ static wt_t weight(float fu, float& fe) { return H::weight(fu, fe); }
    // row r (0..3) of the window as the pair (column u, column u+1)
    template <int R>
    __device__ static uint32_t rowpair(raw_t p)
    {
        constexpr uint32_t sel = (R & 1) ? 0x07060302u : 0x05040100u;
        return R < 2 ? __builtin_amdgcn_perm(p.z, p.x, sel) : __builtin_amdgcn_perm(p.w, p.y, sel);
    }
    __device__ static float accum(float acc, raw_t p, float fv, wt_t w)
    {
        return H::accum(acc, rowpair<0>(p), rowpair<1>(p), w, fv);
    }
    __device__ static float sample(raw_t p, float fv, wt_t w)
    {
        return H::sample(rowpair<0>(p), rowpair<1>(p), w, fv);
    }
    __device__ static float4 abcd(raw_t p)
    {
        const float q00 = H::lo(p.x), q01 = H::lo(p.z), q10 = H::hi(p.x), q11 = H::hi(p.z);
        const float b = q01 - q00;
        return make_float4(q00, b, q10 - q00, (q11 - q10) - b);
    }
};
template <>
struct Sampler<kBF16Strip> : StripSampler<kBF16Strip> {};
template <>
struct Sampler<kF16Strip> : StripSampler<kF16Strip> {};

template <int QT>
constexpr bool is_strip() { return QT == kBF16Strip || QT == kF16Strip; }

struct BPArgs {
    const void* q;       // views view0..view0+nviews-1, quad layout
    const Geo* geo;      // all views
    const float2* uwp;   // all views, N x W count-weight pairs
    const float* colcnt; // geometry class 0: X x Y count weights of the launch's views (bp_colcount_kernel)
    float* vol;          // one-shot output
    float* acc;          // streaming accumulators
    float* cnt;
    int W, V, X, Y, Z;
    int view0, nviews;
    int ntx, nty, ntz;
    int tz0;             // first z tile of the launch (z-slab launches)
    uint32_t frame_bytes;
    int ct;              // apply the CT-number map to normalised output
    float ct_scale, ct_norm1;
    int lds_px;          // LDS-staged fetch: pixels per coefficient buffer (two buffers)
    const uint32_t* sbox;  // LDS-staged strips: rectangle origins u0 | v0 << 16 per tile and view
    int nbox_views;        // views per tile in sbox (the plan's N)
    int sl_stride, sl_rows;  // LDS-staged strips: chunks per rectangle row (odd), strip rows
    int pxb;             // bytes per staged pixel (16: FP32 quads; 8: BF16/FP16 quads and strips)
    int row_px;          // staged pixels per detector row (W, or W + 1 for strips)
    float wscale;        // distance-weighted outputs: vol = wscale * acc
    float qscale;        // normalised outputs are multiplied by it (2^f16exp for FP16 storage)
};

constexpr int kZB = 8;  // samples whose loads are issued together
constexpr int kModeStaged = 3;  // geometry class 0 with the LDS-staged fetch

// Output forms of the backprojection kernel.
constexpr int kOutNormalised = 0;  // vol = acc / count (one-shot)
constexpr int kOutAccumulate = 1;  // acc += ..., count += ... (streaming)
constexpr int kOutRaw = 2;         // vol = acc (the un-normalised operator, FP32 storage only)
// The FDK distance-weighted forms: every sample is scaled by 1/w^2 (w the
// third homogeneous coordinate of the voxel in the view) and there is no
// count: vol = wscale * acc (one-shot) or acc += ... (streaming).
constexpr int kOutFdk = 3;
constexpr int kOutFdkAccumulate = 4;

constexpr bool distance_weighted(int out) { return out == kOutFdk || out == kOutFdkAccumulate; }

// acc + s (count-normalised forms) or acc + s / w^2 (iw2 = 1/w^2)
template <bool WT>
__device__ __forceinline__ float add_sample(float acc, float s, float iw2)
{
    if constexpr (WT) return fmaf(s, iw2, acc);
    else return acc + s;
}

// The samples of one all-in-range column in pairs over column strips: one
// 16-byte gather at row floor(sv_k) serves samples k and k+1; the second
// sample's rows start HI or HI + 1 rows down.  Each sample interpolates the
// same row pairs with the same operations as the quad layout's one gather per
// sample, so the results are identical.  fetch(off) returns the 16-byte
// window at byte offset off = row * rowb + ub (from memory, or from LDS).
template <int ZC, int QT, int HI, bool WT, class F>
__device__ __forceinline__ void bp_column_pairs(float* acc, const F& fetch, uint32_t rowb, uint32_t ub, float sv0,
                                                float dsv, typename Sampler<QT>::wt_t wq, float iw2)
{
    typedef Sampler<QT> S;
    constexpr int NP = ZC / 2;
    uint32_t off[NP];
    float fa[NP], fb[NP];
    bool e[NP];
#pragma unroll
    for (int m = 0; m < NP; ++m) {
        const float sa = fmaf((float)(2 * m), dsv, sv0);
        const float sb = fmaf((float)(2 * m + 1), dsv, sv0);
        const int ra = (int)sa, rb = (int)sb;
        fa[m] = __builtin_amdgcn_fractf(sa);
        fb[m] = __builtin_amdgcn_fractf(sb);
        e[m] = (rb - ra) != HI;
        off[m] = __umul24((uint32_t)ra, rowb) + ub;
    }
    u32x4_t p[NP];
#pragma unroll
    for (int m = 0; m < NP; ++m) p[m] = fetch(off[m]);
#pragma unroll
    for (int m = 0; m < NP; ++m) {
        const uint32_t r0 = S::template rowpair<0>(p[m]);
        const uint32_t r1 = S::template rowpair<1>(p[m]);
        const uint32_t r2 = S::template rowpair<2>(p[m]);
        if constexpr (WT) acc[2 * m] = fmaf(S::H::sample(r0, r1, wq, fa[m]), iw2, acc[2 * m]);
        else acc[2 * m] = S::H::accum(acc[2 * m], r0, r1, wq, fa[m]);
        uint32_t rt, rb;
        if constexpr (HI == 0) {
            rt = e[m] ? r1 : r0;
            rb = e[m] ? r2 : r1;
        } else {
            const uint32_t r3 = S::template rowpair<3>(p[m]);
            rt = e[m] ? r2 : r1;
            rb = e[m] ? r3 : r2;
        }
        if constexpr (WT) acc[2 * m + 1] = fmaf(S::H::sample(rt, rb, wq, fb[m]), iw2, acc[2 * m + 1]);
        else acc[2 * m + 1] = S::H::accum(acc[2 * m + 1], rt, rb, wq, fb[m]);
    }
}

// One tile: 64 blockDim.y voxel columns (8 x 8 per wave), ZC voxels deep,
// all views of the launch.
template <int MODE, int ZC, int QT, int OUT>
__device__ __forceinline__ void bp_tile(const BPArgs& a, unsigned tile)
{
    typedef Sampler<QT> S;
    constexpr bool WT = distance_weighted(OUT);
    // acc + the sample of raw p (scaled by iw2 = 1/w^2 in the weighted forms)
    auto add = [](float acc, typename S::raw_t p, float fv, typename S::wt_t wq, float iw2) {
        if constexpr (WT) return fmaf(S::sample(p, fv, wq), iw2, acc);
        else return S::accum(acc, p, fv, wq);
    };
    const int tx = tile % a.ntx;
    const unsigned rest = tile / a.ntx;
    const int tyi = rest % a.nty;
    const int tz = a.tz0 + (int)(rest / a.nty);
    // Each wave covers an 8 x 8 patch of voxel columns; the blockDim.y waves
    // of a workgroup tile 8 x 8 (1 wave), 16 x 8 (2) or 16 x 16 (4).  A wave's
    // 64 samples of one view then fall on a compact detector footprint: a
    // gather's texture-path cost follows the cache lines the wave touches,
    // which a 64 x 1 row of voxels spreads over ~4x more lines.
    const int px = blockDim.y >= 2 ? 2 : 1;
    const int wx = threadIdx.y % px, wy = threadIdx.y / px;
    const int i = (tx * px + wx) * 8 + (threadIdx.x & 7);
    const int j = (tyi * (blockDim.y / px) + wy) * 8 + (threadIdx.x >> 3);
    const int k0 = tz * ZC;
    const float fi = (float)min(i, a.X - 1);
    const float fj = (float)min(j, a.Y - 1);
    const float fk0 = (float)k0;
    const float Wm1 = (float)(a.W - 1), Vm1 = (float)(a.V - 1);
    const uint32_t rowb = (uint32_t)a.row_px * (uint32_t)a.pxb;

    float acc[ZC], cnt[ZC];
#pragma unroll
    for (int kk = 0; kk < ZC; ++kk) { acc[kk] = 0.0f; cnt[kk] = 0.0f; }
    // MODE 0: the per-sample count weights of the views whose column is only
    // partly inside the detector rows are kept in LDS, one slot per thread and
    // sample (touched by that thread only, so no barrier): such views are rare,
    // and ZC registers held across the view loop would cost occupancy.
    __shared__ float cnt_part[MODE == 0 && !WT ? ZC * 256 : 1];
    const int tid = threadIdx.y * 64 + threadIdx.x;
    if constexpr (MODE == 0 && !WT) {
#pragma unroll
        for (int kk = 0; kk < ZC; ++kk) cnt_part[kk * 256 + tid] = 0.0f;
    }
    // The streaming forms continue the running sums of the earlier launches
    // (loaded here, stored at the end), so every voxel's additions run in
    // view order whatever the launch boundaries: the result equals one launch
    // over all views to the bit.  The count term kept per voxel is the
    // samples' count in geometry classes 1 and 2, and the weight of the
    // samples outside the detector rows in class 0 (whose count is the
    // column's running count weight less that).
    if constexpr (OUT == kOutAccumulate || OUT == kOutFdkAccumulate) {
        // every lane loads (clamped into the volume) and selects: a branch per
        // sample here made the compiler hold ~18 more registers through the
        // view loop and keep fewer gathers in flight
        const size_t plane = (size_t)a.X * a.Y;
        const bool col = i < a.X && j < a.Y;
        const size_t base = ((size_t)k0 * a.Y + min(j, a.Y - 1)) * a.X + min(i, a.X - 1);
#pragma unroll
        for (int kk = 0; kk < ZC; ++kk) {
            const bool ok = col && k0 + kk < a.Z;
            const size_t idx = base + (size_t)min(kk, a.Z - 1 - k0) * plane;
            const float x = a.acc[idx];
            acc[kk] = ok ? x : 0.0f;
            if constexpr (!WT) {
                const float c = a.cnt[idx];
                if constexpr (MODE == 0) cnt_part[kk * 256 + tid] = ok ? c : 0.0f;
                else cnt[kk] = ok ? c : 0.0f;
            }
        }
    }

    for (int v = 0; v < a.nviews; ++v) {
        const Geo g = geo_of(a.geo, a.view0 + v);
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<char*>(reinterpret_cast<const char*>(a.q)) + (size_t)v * a.frame_bytes, (short)0,
            (int)a.frame_bytes, kBufferFlags);
        const float2* uwv = a.uwp + (size_t)(a.view0 + v) * a.W;
        const float Uc = fmaf(g.c[3], fk0, fmaf(g.c[1], fi, fmaf(g.c[2], fj, g.c[0])));
        const float Vc = fmaf(g.c[7], fk0, fmaf(g.c[5], fi, fmaf(g.c[6], fj, g.c[4])));
        const float Wc = fmaf(g.c[11], fk0, fmaf(g.c[9], fi, fmaf(g.c[10], fj, g.c[8])));
        if (MODE == 0) {
            // u and w are constant along the column: hoist the u side.
            // 1/w by v_rcp_f32 (1 ulp; the forward projector takes the same
            // reciprocal, so the pair stays adjoint)
            const float rw = __builtin_amdgcn_rcpf(Wc);
            const float su = fmaf(Uc, rw, 0.5f);
            const float iw2 = rw * rw;  // distance weight (weighted forms)
            if (su >= 0.0f && su <= Wm1) {
                const int iu = (int)su;
                float fe;
                const typename S::wt_t wq = S::weight(__builtin_amdgcn_fractf(su), fe);
                const float sv0 = fmaf(Vc, rw, 0.5f);
                const float dsv = g.c[7] * rw;
                const uint32_t ub = __umul24((uint32_t)iu, (uint32_t)a.pxb);
                // sv is monotonic in k, so the column is inside the detector
                // rows iff its first and last samples are.
                const float svl = fmaf((float)(ZC - 1), dsv, sv0);
                if (fminf(sv0, svl) >= 0.0f && fmaxf(sv0, svl) <= Vm1) {
                    // every sample counts: the count weight is added once.
                    // All gathers of the column are issued before any is
                    // consumed: this keeps ZC / 2 of them in flight per wave,
                    // so their latency hides behind the VALU work (which
                    // bounds the kernel: permutes, dots, integer and
                    // conversion instructions issue every ~4 clocks).
                    // BF16 / FP16 strips: one b128 gather per pair of samples
                    // (k, k+1) when the row step lies clear of an integer
                    // (margin 1/64 against the rounding of sv): sample k
                    // reads rows 0-1 of its strip, sample k+1 rows d-(d+1),
                    // d = floor(sv_k+1) - floor(sv_k), which is hi or hi+1
                    bool paired = false;
                    if constexpr (is_strip<QT>()) {
                        const auto gather = [&](uint32_t o) { return S::load(rs, o); };
                        if (dsv >= 0.0f && dsv <= 0.984375f) {
                            bp_column_pairs<ZC, QT, 0, WT>(acc, gather, rowb, ub, sv0, dsv, wq, iw2);
                            paired = true;
                        } else if (dsv >= 1.015625f && dsv <= 1.984375f) {
                            bp_column_pairs<ZC, QT, 1, WT>(acc, gather, rowb, ub, sv0, dsv, wq, iw2);
                            paired = true;
                        }
                    }
                    if (!paired) {
                        uint32_t off[ZC];
                        float fv[ZC];
#pragma unroll
                        for (int k = 0; k < ZC; ++k) {
                            const float sv = fmaf((float)k, dsv, sv0);
                            fv[k] = __builtin_amdgcn_fractf(sv);
                            off[k] = __umul24((uint32_t)(int)sv, rowb) + ub;
                        }
                        typename S::raw_t p[ZC];
#pragma unroll
                        for (int k = 0; k < ZC; ++k) p[k] = S::load(rs, off[k]);
#pragma unroll
                        for (int k = 0; k < ZC; ++k) acc[k] = add(acc[k], p[k], fv[k], wq, iw2);
                    }
                    // every sample counts: the column's count weight of this
                    // view is in a.colcnt already
                } else {
                    // the count weight of the samples outside the detector
                    // rows is taken off a.colcnt at the end
                    float cw = 0.0f;
                    if constexpr (!WT) {
                        const float2 uwq = uwv[iu];
                        cw = fmaf(fe, uwq.y - uwq.x, uwq.x);
                    }
#pragma unroll
                    for (int kb = 0; kb < ZC; kb += kZB) {
                        uint32_t off[kZB];
                        float fv[kZB];
                        bool in[kZB];
#pragma unroll
                        for (int q = 0; q < kZB; ++q) {
                            const float sv = fmaf((float)(kb + q), dsv, sv0);
                            const float t = __builtin_amdgcn_fmed3f(sv, 0.0f, Vm1);
                            in[q] = (t == sv);
                            fv[q] = __builtin_amdgcn_fractf(t);
                            off[q] = __umul24((uint32_t)(int)t, rowb) + ub;
                        }
                        typename S::raw_t p[kZB];
#pragma unroll
                        for (int q = 0; q < kZB; ++q) p[q] = S::load(rs, off[q]);
#pragma unroll
                        for (int q = 0; q < kZB; ++q) {
                            const float sum = add(acc[kb + q], p[q], fv[q], wq, iw2);
                            acc[kb + q] = in[q] ? sum : acc[kb + q];
                            if constexpr (!WT) cnt_part[(kb + q) * 256 + tid] += in[q] ? 0.0f : cw;
                        }
                    }
                }
            }
        } else {
            // MODE 1: w constant along the column, u and v linear in z;
            // MODE 2: general, one reciprocal per sample.
            const float rw = __builtin_amdgcn_rcpf(Wc);
            const float su0 = fmaf(Uc, rw, 0.5f), dsu = g.c[3] * rw;
            const float sv0 = fmaf(Vc, rw, 0.5f), dsv = g.c[7] * rw;
#pragma unroll
            for (int kb = 0; kb < ZC; kb += kZB) {
                uint32_t off[kZB];
                float fe[kZB], fv[kZB], iw2[kZB];
                typename S::wt_t wq[kZB];
                bool in[kZB];
                float2 uwq[kZB];
#pragma unroll
                for (int q = 0; q < kZB; ++q) {
                    const float fk = (float)(kb + q);
                    float su, sv;
                    if (MODE == 1) {
                        su = fmaf(fk, dsu, su0);
                        sv = fmaf(fk, dsv, sv0);
                        iw2[q] = rw * rw;
                    } else {
                        const float r = __builtin_amdgcn_rcpf(fmaf(fk, g.c[11], Wc));
                        su = fmaf(fmaf(fk, g.c[3], Uc), r, 0.5f);
                        sv = fmaf(fmaf(fk, g.c[7], Vc), r, 0.5f);
                        iw2[q] = r * r;
                    }
                    const float tu = __builtin_amdgcn_fmed3f(su, 0.0f, Wm1);
                    const float t = __builtin_amdgcn_fmed3f(sv, 0.0f, Vm1);
                    in[q] = (t == sv) && (tu == su);
                    const int iu = (int)tu;
                    wq[q] = S::weight(__builtin_amdgcn_fractf(tu), fe[q]);
                    fv[q] = __builtin_amdgcn_fractf(t);
                    off[q] = __umul24((uint32_t)(int)t, rowb) + __umul24((uint32_t)iu, (uint32_t)a.pxb);
                    uwq[q] = uwv[iu];
                }
                typename S::raw_t p[kZB];
#pragma unroll
                for (int q = 0; q < kZB; ++q) p[q] = S::load(rs, off[q]);
#pragma unroll
                for (int q = 0; q < kZB; ++q) {
                    const float sum = add(acc[kb + q], p[q], fv[q], wq[q], iw2[q]);
                    acc[kb + q] = in[q] ? sum : acc[kb + q];
                    if constexpr (!WT) {
                        const float cw = fmaf(fe[q], uwq[q].y - uwq[q].x, uwq[q].x);
                        cnt[kb + q] += in[q] ? cw : 0.0f;
                    }
                }
            }
        }
    }

    if (i >= a.X || j >= a.Y) return;
    if constexpr (MODE == 0 && OUT == kOutNormalised) {
        const float c = a.colcnt[(size_t)j * a.X + i];
#pragma unroll
        for (int kk = 0; kk < ZC; ++kk) cnt[kk] = c - cnt_part[kk * 256 + tid];
    }
    const size_t plane = (size_t)a.X * a.Y;
    size_t idx = ((size_t)k0 * a.Y + j) * a.X + i;
#pragma unroll
    for (int kk = 0; kk < ZC; ++kk) {
        if (k0 + kk < a.Z) {
            if constexpr (OUT == kOutNormalised || OUT == kOutFdk) {
                const float v = (OUT == kOutFdk ? a.wscale * acc[kk] : acc[kk] / cnt[kk]) * a.qscale;
                a.vol[idx] = a.ct ? ct_number(v, a.ct_scale, a.ct_norm1) : v;
            } else if constexpr (OUT == kOutAccumulate) {
                a.acc[idx] = acc[kk];
                if constexpr (MODE == 0) a.cnt[idx] = cnt_part[kk * 256 + tid];
                else a.cnt[idx] = cnt[kk];
            } else if constexpr (OUT == kOutFdkAccumulate) {
                a.acc[idx] = acc[kk];
            } else {
                a.vol[idx] = acc[kk];
            }
        }
        idx += plane;
    }
}


// ---------------------------------------------------------------------------
// LDS-staged fetch (geometry class 0: u and w constant along z).
//
// Per view the workgroup computes the detector footprint of its voxel block
// from the block's corners -- u/w and v/w are linear-fractional in the voxel
// coordinates, so their extremes over the block are at its corners -- widened
// by one pixel against rounding, and copies that rectangle of quads once into
// LDS as FP32 bilinear coefficients (a, b, c, d):
//   value(fu, fv) = a + fu b + fv (c + fu d).
// Every sample is then one ds_read_b128 and three FMAs; the per-sample
// gathers no longer pass through the texture addresser, which bounds the
// gather form.  Sample positions are computed exactly as in the gather form.
//
// The staging is software-pipelined over the views, with one barrier per
// view: while the workgroup samples view v from one LDS buffer, its global
// loads of view v+1's footprint are in flight (issued after the previous
// barrier, held in registers), and they are converted into the other buffer
// after the sampling.  Footprints are computed two views ahead by wave 0 into
// a ring of three boxes.  A footprint larger than the buffer falls back to
// gathers for that view.
// ---------------------------------------------------------------------------

constexpr int kLdsPre = 8;    // footprint pixels per thread held in registers across a view

// acc + the sample with bilinear coefficients c = (a, b, c, d) at (fu, fv),
// rounded at the accumulator in the same places as the gather form's
// Sampler<QT>::accum: once, or twice (top, then fv (bot - top)) for FP16.
template <int QT>
__device__ __forceinline__ float lds_accum(float acc, float4 c, float fu, float fv)
{
    const float top = fmaf(fu, c.y, c.x), d = fmaf(fu, c.w, c.z);
    if constexpr (QT == JH_F16 || QT == kF16Strip) return fmaf(fv, d, acc + top);
    else return acc + fmaf(fv, d, top);
}
constexpr int kLdsUw = 256;   // widest footprint whose count-weight pairs are staged too

// Workgroup-shared LDS: two coefficient buffers of a.lds_px pixels each
// (dynamic shared memory, sized by the host per tile shape).
extern __shared__ float4 lds_dyn[];

template <int ZC, int QT, int OUT>
__device__ __forceinline__ void bp_tile_lds(const BPArgs& a, unsigned tile)
{
    static_assert(!distance_weighted(OUT), "the distance-weighted forms gather");
    typedef Sampler<QT> S;
    __shared__ int box[3][4];  // u0, v0, width, pixels (pixels > lds_px: not staged)
    // count-weight pairs uwp[u0 .. u0 + width) of the staged view, so that the
    // sampling loop issues no vector-memory load (a wait on one would also
    // wait for the footprint prefetch in flight)
    __shared__ float2 uwl[2][kLdsUw];
    const int tx = tile % a.ntx;
    const unsigned rest = tile / a.ntx;
    const int tyi = rest % a.nty;
    const int tz = a.tz0 + (int)(rest / a.nty);
    const int px = blockDim.y >= 2 ? 2 : 1;
    const int wx = threadIdx.y % px, wy = threadIdx.y / px;
    const int ex = px * 8, ey = (blockDim.y / px) * 8;
    const int i = tx * ex + wx * 8 + (threadIdx.x & 7);
    const int j = tyi * ey + wy * 8 + (threadIdx.x >> 3);
    const int k0 = tz * ZC;
    const float fi = (float)min(i, a.X - 1);
    const float fj = (float)min(j, a.Y - 1);
    const float fk0 = (float)k0;
    // corner columns of the block (clamped like the threads' own columns)
    const float cx0 = (float)min(tx * ex, a.X - 1), cx1 = (float)min(tx * ex + ex - 1, a.X - 1);
    const float cy0 = (float)min(tyi * ey, a.Y - 1), cy1 = (float)min(tyi * ey + ey - 1, a.Y - 1);
    const float Wm1 = (float)(a.W - 1), Vm1 = (float)(a.V - 1);
    const uint32_t rowb = (uint32_t)a.row_px * (uint32_t)a.pxb;
    const int nthr = 64 * blockDim.y;
    const int tid = threadIdx.y * 64 + threadIdx.x;
    const int cap = a.lds_px;
    const int n = a.nviews;

    auto rsrc = [&](int vv) {
        return __builtin_amdgcn_make_buffer_rsrc(
            const_cast<char*>(reinterpret_cast<const char*>(a.q)) + (size_t)vv * a.frame_bytes, (short)0,
            (int)a.frame_bytes, kBufferFlags);
    };

    // Footprint of the block in view vv into box[slot]: all of wave 0, one
    // lane per corner column in every group of 4 (reciprocal instead of
    // division: the one-pixel margin absorbs the difference).
    auto footprint = [&](int vv, int slot) {
        if (threadIdx.y != 0) return;
        const Geo g = geo_of(a.geo, a.view0 + vv);
        const int c = threadIdx.x & 3;
        const float cx = (c & 1) ? cx1 : cx0, cy = (c & 2) ? cy1 : cy0;
        const float U = fmaf(g.c[3], fk0, fmaf(g.c[1], cx, fmaf(g.c[2], cy, g.c[0])));
        const float Vv = fmaf(g.c[7], fk0, fmaf(g.c[5], cx, fmaf(g.c[6], cy, g.c[4])));
        const float Ww = fmaf(g.c[11], fk0, fmaf(g.c[9], cx, fmaf(g.c[10], cy, g.c[8])));
        const float r = __builtin_amdgcn_rcpf(Ww);
        const float su = fmaf(U, r, 0.5f);
        const float s0 = fmaf(Vv, r, 0.5f);
        const float sl = fmaf((float)(ZC - 1), g.c[7] * r, s0);
        float umin = su, umax = su, vmin = fminf(s0, sl), vmax = fmaxf(s0, sl);
#pragma unroll
        for (int m = 1; m <= 2; m <<= 1) {
            umin = fminf(umin, __shfl_xor(umin, m));
            umax = fmaxf(umax, __shfl_xor(umax, m));
            vmin = fminf(vmin, __shfl_xor(vmin, m));
            vmax = fmaxf(vmax, __shfl_xor(vmax, m));
        }
        if (threadIdx.x == 0) {
            const int u0 = (int)fminf(fmaxf(floorf(umin) - 1.0f, 0.0f), Wm1);
            const int u1 = (int)fminf(fmaxf(floorf(umax) + 1.0f, 0.0f), Wm1);
            const int v0 = (int)fminf(fmaxf(floorf(vmin) - 1.0f, 0.0f), Vm1);
            const int v1 = (int)fminf(fmaxf(floorf(vmax) + 1.0f, 0.0f), Vm1);
            box[slot][0] = u0;
            box[slot][1] = v0;
            box[slot][2] = u1 - u0 + 1;
            box[slot][3] = (u1 - u0 + 1) * (v1 - v0 + 1);
        }
    };

    // Pixel b of a footprint box: row b / width (exact: (b + 0.5) / width is
    // at least 0.5 / width away from an integer, far beyond float rounding).
    auto pixel_offset = [&](int b, int u0, int v0, int wdt, float rinv) {
        const int r = (int)(((float)b + 0.5f) * rinv);
        const int c = b - r * wdt;
        return (uint32_t)(v0 + r) * rowb + (uint32_t)(u0 + c) * (uint32_t)a.pxb;
    };

    // a view is staged when its footprint and count-weight row fit
    auto staged = [&](int slot) { return box[slot][3] <= cap && box[slot][2] <= kLdsUw; };

    typename S::raw_t pre[kLdsPre];
    float2 preuw = make_float2(0.0f, 0.0f);
    // issue the global loads of view vv's footprint (box slot) into pre[]
    // (every lane issues every load, at a pixel clamped into the footprint:
    // a load under a divergent branch makes the compiler wait for all loads
    // in flight before it, which would serialise the prefetch)
    auto issue = [&](int vv, int slot) {
        if (!staged(slot)) return;
        const int npx = box[slot][3];
        const int u0 = box[slot][0], v0 = box[slot][1], wdt = box[slot][2];
        preuw = a.uwp[(size_t)(a.view0 + vv) * a.W + u0 + min(tid, wdt - 1)];
        const float rinv = 1.0f / (float)wdt;
        const auto rs = rsrc(vv);
#pragma unroll
        for (int m = 0; m < kLdsPre; ++m) {
            const int b = min(tid + m * nthr, npx - 1);
            pre[m] = S::load(rs, pixel_offset(b, u0, v0, wdt, rinv));
        }
    };
    // convert pre[] (and any pixels beyond kLdsPre per thread) into buffer buf
    auto commit = [&](int vv, int slot, int buf) {
        if (!staged(slot)) return;
        const int npx = box[slot][3];
        if (tid < box[slot][2]) uwl[buf][tid] = preuw;
        if (box[slot][2] > nthr) {  // small workgroups: the rest of the row synchronously
            for (int b = tid + nthr; b < box[slot][2]; b += nthr)
                uwl[buf][b] = a.uwp[(size_t)(a.view0 + vv) * a.W + box[slot][0] + b];
        }
        float4* dst = lds_dyn + buf * cap;
#pragma unroll
        for (int m = 0; m < kLdsPre; ++m) {
            const int b = tid + m * nthr;
            if (b < npx) dst[b] = S::abcd(pre[m]);
        }
        if (npx > kLdsPre * nthr) {  // small workgroups: the rest synchronously
            const int u0 = box[slot][0], v0 = box[slot][1], wdt = box[slot][2];
            const float rinv = 1.0f / (float)wdt;
            const auto rs = rsrc(vv);
            for (int b = tid + kLdsPre * nthr; b < npx; b += nthr)
                dst[b] = S::abcd(S::load(rs, pixel_offset(b, u0, v0, wdt, rinv)));
        }
    };

    // part: the count weight of the samples outside the detector rows (the
    // count is the column's count weight from bp_colcount_kernel less it, as
    // in the gather form); the streaming form continues the running sums
    float acc[ZC], part[ZC];
#pragma unroll
    for (int kk = 0; kk < ZC; ++kk) { acc[kk] = 0.0f; part[kk] = 0.0f; }
    if constexpr (OUT == kOutAccumulate) {
        // clamped loads and selects (see bp_tile)
        const size_t plane = (size_t)a.X * a.Y;
        const bool col = i < a.X && j < a.Y;
        const size_t base = ((size_t)k0 * a.Y + min(j, a.Y - 1)) * a.X + min(i, a.X - 1);
#pragma unroll
        for (int kk = 0; kk < ZC; ++kk) {
            const bool ok = col && k0 + kk < a.Z;
            const size_t idx = base + (size_t)min(kk, a.Z - 1 - k0) * plane;
            const float x = a.acc[idx], c = a.cnt[idx];
            acc[kk] = ok ? x : 0.0f;
            part[kk] = ok ? c : 0.0f;
        }
    }

    if (n > 0) {
        footprint(0, 0);
        if (n > 1) footprint(1, 1);
        __syncthreads();
        issue(0, 0);
        commit(0, 0, 0);
        if (n > 1) issue(1, 1);
        __syncthreads();
    }
    for (int v = 0; v < n; ++v) {
        const int slot = v % 3, buf = v & 1;
        if (v + 2 < n) footprint(v + 2, (v + 2) % 3);

        const Geo g = geo_of(a.geo, a.view0 + v);
        const float Uc = fmaf(g.c[3], fk0, fmaf(g.c[1], fi, fmaf(g.c[2], fj, g.c[0])));
        const float Vc = fmaf(g.c[7], fk0, fmaf(g.c[5], fi, fmaf(g.c[6], fj, g.c[4])));
        const float Wc = fmaf(g.c[11], fk0, fmaf(g.c[9], fi, fmaf(g.c[10], fj, g.c[8])));
        const float rw = __builtin_amdgcn_rcpf(Wc);
        const float su = fmaf(Uc, rw, 0.5f);
        if (su >= 0.0f && su <= Wm1) {
            const int iu = (int)su;
            // fu: the u fraction the gather form applies for this storage type
            float fu;
            const typename S::wt_t wq = S::weight(__builtin_amdgcn_fractf(su), fu);
            const float sv0 = fmaf(Vc, rw, 0.5f);
            const float dsv = g.c[7] * rw;
            const float svl = fmaf((float)(ZC - 1), dsv, sv0);
            const bool all_in = fminf(sv0, svl) >= 0.0f && fmaxf(sv0, svl) <= Vm1;
            const int u0 = box[slot][0], v0 = box[slot][1], wdt = box[slot][2];
            if (staged(slot)) {
                const float2 uwq = uwl[buf][iu - u0];
                const float cw = fmaf(fu, uwq.y - uwq.x, uwq.x);
                // byte offset of sample row iv: iv * wdt * 16 + lb (mod 2^32)
                const uint32_t lrow = (uint32_t)wdt * 16u;
                const uint32_t lb = (uint32_t)(iu - u0) * 16u - (uint32_t)v0 * lrow;
                const char* lbase = reinterpret_cast<const char*>(lds_dyn + buf * cap);
                if (all_in) {
                    // kZB reads in flight per wave before any is consumed: at
                    // the occupancy the LDS budget allows, LDS latency is
                    // otherwise exposed
#pragma unroll
                    for (int kb = 0; kb < ZC; kb += kZB) {
                        float4 c4[kZB];
                        float fv[kZB];
#pragma unroll
                        for (int q = 0; q < kZB; ++q) {
                            const float sv = fmaf((float)(kb + q), dsv, sv0);
                            fv[q] = __builtin_amdgcn_fractf(sv);
                            c4[q] = *reinterpret_cast<const float4*>(lbase + __umul24((uint32_t)(int)sv, lrow) + lb);
                        }
                        // keep the scheduler from interleaving the reads with
                        // their uses (it pairs them under register pressure)
                        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                        for (int q = 0; q < kZB; ++q) acc[kb + q] = lds_accum<QT>(acc[kb + q], c4[q], fu, fv[q]);
                    }
                } else {
#pragma unroll
                    for (int k = 0; k < ZC; ++k) {
                        const float sv = fmaf((float)k, dsv, sv0);
                        const float t = __builtin_amdgcn_fmed3f(sv, 0.0f, Vm1);
                        const bool in = (t == sv);
                        const float fv = __builtin_amdgcn_fractf(t);
                        const float4 c4 =
                            *reinterpret_cast<const float4*>(lbase + __umul24((uint32_t)(int)t, lrow) + lb);
                        const float sum = lds_accum<QT>(acc[k], c4, fu, fv);
                        acc[k] = in ? sum : acc[k];
                        part[k] += in ? 0.0f : cw;
                    }
                }
            } else {
                // footprint larger than the LDS buffer: gathers for this view
                const float2 uwq = a.uwp[(size_t)(a.view0 + v) * a.W + iu];
                const float cw = fmaf(fu, uwq.y - uwq.x, uwq.x);
                const auto rs = rsrc(v);
                const uint32_t ub = __umul24((uint32_t)iu, (uint32_t)a.pxb);
#pragma unroll
                for (int kb = 0; kb < ZC; kb += kZB) {
                    uint32_t off[kZB];
                    float fv[kZB];
                    bool in[kZB];
#pragma unroll
                    for (int q = 0; q < kZB; ++q) {
                        const float sv = fmaf((float)(kb + q), dsv, sv0);
                        const float t = __builtin_amdgcn_fmed3f(sv, 0.0f, Vm1);
                        in[q] = (t == sv);
                        fv[q] = __builtin_amdgcn_fractf(t);
                        off[q] = __umul24((uint32_t)(int)t, rowb) + ub;
                    }
                    typename S::raw_t p[kZB];
#pragma unroll
                    for (int q = 0; q < kZB; ++q) p[q] = S::load(rs, off[q]);
#pragma unroll
                    for (int q = 0; q < kZB; ++q) {
                        const float sum = S::accum(acc[kb + q], p[q], fv[q], wq);
                        acc[kb + q] = in[q] ? sum : acc[kb + q];
                        part[kb + q] += in[q] ? 0.0f : cw;
                    }
                }
            }
        }
        // view v+1's footprint into the other buffer (its last reader was view
        // v-1, before the previous barrier), then one barrier: the buffer and
        // box v+2 are ready and every read of this view's buffer is done
        if (v + 1 < n) commit(v + 1, (v + 1) % 3, buf ^ 1);
        __syncthreads();
        if (v + 2 < n) issue(v + 2, (v + 2) % 3);
    }

    if (i >= a.X || j >= a.Y) return;
    const float c = OUT == kOutNormalised ? a.colcnt[(size_t)j * a.X + i] : 0.0f;
    const size_t plane = (size_t)a.X * a.Y;
    size_t idx = ((size_t)k0 * a.Y + j) * a.X + i;
#pragma unroll
    for (int kk = 0; kk < ZC; ++kk) {
        if (k0 + kk < a.Z) {
            if (OUT == kOutNormalised) {
                const float v = acc[kk] / (c - part[kk]) * a.qscale;
                a.vol[idx] = a.ct ? ct_number(v, a.ct_scale, a.ct_norm1) : v;
            } else if (OUT == kOutAccumulate) {
                a.acc[idx] = acc[kk];
                a.cnt[idx] = part[kk];
            } else {
                a.vol[idx] = acc[kk];
            }
        }
        idx += plane;
    }
}

// Count weights of geometry class 0 per voxel column (u and w do not depend
// on z): the sum over the launch's views of the interpolated uW at the
// column's u, in view order and with the sampler's u weight -- the same
// operations as the backprojector's own sample positions.  The
// backprojector takes off the samples that fall outside the detector rows;
// every other sample of the column counts.  One thread per column, shared by
// all the column's z tiles (so the backprojector loads no count weight per
// view).  With cont the sum continues the column's running sum in colcnt
// (the streaming form), else it starts from zero.
template <int QT>
__global__ void __launch_bounds__(256) bp_colcount_kernel(const Geo* __restrict__ geo, const float2* __restrict__ uwp,
                                                          int view0, int nviews, int X, int Y, int W,
                                                          float* __restrict__ colcnt, int cont)
{
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int j = blockIdx.y * 4 + threadIdx.y;
    if (i >= X || j >= Y) return;
    const float fi = (float)i, fj = (float)j, Wm1 = (float)(W - 1);
    float c = cont ? colcnt[(size_t)j * X + i] : 0.0f;
    for (int v = 0; v < nviews; ++v) {
        const Geo g = geo[view0 + v];
        // as bp_tile with c[3] = c[11] = 0 (fmaf(0, k0, x) = x)
        const float Uc = fmaf(g.c[1], fi, fmaf(g.c[2], fj, g.c[0]));
        const float Wc = fmaf(g.c[9], fi, fmaf(g.c[10], fj, g.c[8]));
        const float rw = __builtin_amdgcn_rcpf(Wc);
        const float su = fmaf(Uc, rw, 0.5f);
        if (su >= 0.0f && su <= Wm1) {
            float fe;
            (void)Sampler<QT>::weight(__builtin_amdgcn_fractf(su), fe);
            const float2 uwq = uwp[(size_t)(view0 + v) * W + (int)su];
            c += fmaf(fe, uwq.y - uwq.x, uwq.x);
        }
    }
    colcnt[(size_t)j * X + i] = c;
}

// ---------------------------------------------------------------------------
// LDS-staged column strips (geometry class 0, BF16 / FP16 column strips,
// count-normalised forms, workgroups of 16 x 16 voxel columns).
//
// Per view the workgroup copies a rectangle of strip pixels around its
// block's detector footprint -- strip rows [v0, v0 + rows), pixels from u0 in
// 16-byte chunks of two pixels, `stride` chunks per row (the largest
// footprint over the plan's views, measured on the device, rounded up to an
// odd number of chunks so that consecutive LDS rows start in other banks) --
// from the staged frame into LDS unchanged, and samples there with the
// operations of the gather form: a sample's 16-byte window is one
// ds_read2_b64, and one window serves a pair of z samples where the gather
// form pairs them, so the volume is the gather form's to the bit.
//
// Every thread copies a fixed set of the rectangle's chunks (offsets formed
// once per workgroup, plus the view's origin) through registers: view v + 1's
// loads are issued after view v - 1 was written to LDS and written after view
// v is sampled, into the buffer view v - 1 used, so two buffers, one set of
// registers and one barrier per view suffice, and no load is waited for
// while a view is sampled.  (buffer_load ... lds, which needs no registers,
// was measured at ~70 texture-data clocks per 64-lane 16-byte copy, 2.5x a
// gather, and the copies then cost as much as the gathers they replace.)
// The rectangle origins (u0, v0) of every tile and view come from a table
// built once per tile shape on the device, read with scalar loads.  Partial
// columns (samples outside the detector rows) keep their count corrections
// in registers.
// What this buys: each footprint pixel passes the texture path once per
// workgroup instead of once per sampling lane (the gather form's texture data
// path spends ~28 clocks per 64-lane 16-byte gather), and the window reads run
// from LDS; what it costs: the window addressing inside the rectangle adds
// VALU work, which bounds both forms.
// ---------------------------------------------------------------------------

// The strip rectangle a block of voxel columns [cx0, cx1] x [cy0, cy1], z
// voxels fk0 .. fk0 + ZC - 1, samples in a view: first pixel u0, first strip
// row v0, 16-byte chunks per row and strip rows.  Evaluated by the 4 lanes of
// every lane group of 4 (one per corner column; the result is the group's):
// u/w and v/w are linear-fractional in the voxel coordinates, so their
// extremes over the block are at its corners; the reciprocal is the one the
// samples take, and one pixel of margin on each side covers the difference.
// A window is pixels u, u+1 (the padding column W included) of one strip row.
struct StripBox {
    int u0, v0, cpr, rows;
};

template <int ZC>
__device__ __forceinline__ StripBox strip_footprint(const Geo& g, float cx0, float cx1, float cy0, float cy1, float fk0,
                                                    int W, int V)
{
    const int c = threadIdx.x & 3;
    const float cx = (c & 1) ? cx1 : cx0, cy = (c & 2) ? cy1 : cy0;
    const float U = fmaf(g.c[3], fk0, fmaf(g.c[1], cx, fmaf(g.c[2], cy, g.c[0])));
    const float Vv = fmaf(g.c[7], fk0, fmaf(g.c[5], cx, fmaf(g.c[6], cy, g.c[4])));
    const float Ww = fmaf(g.c[11], fk0, fmaf(g.c[9], cx, fmaf(g.c[10], cy, g.c[8])));
    const float r = __builtin_amdgcn_rcpf(Ww);
    const float su = fmaf(U, r, 0.5f);
    const float s0 = fmaf(Vv, r, 0.5f);
    const float sl = fmaf((float)(ZC - 1), g.c[7] * r, s0);
    float umin = su, umax = su, vmin = fminf(s0, sl), vmax = fmaxf(s0, sl);
#pragma unroll
    for (int m = 1; m <= 2; m <<= 1) {
        umin = fminf(umin, __shfl_xor(umin, m));
        umax = fmaxf(umax, __shfl_xor(umax, m));
        vmin = fminf(vmin, __shfl_xor(vmin, m));
        vmax = fmaxf(vmax, __shfl_xor(vmax, m));
    }
    const float Wm1 = (float)(W - 1), Vm1 = (float)(V - 1);
    StripBox b;
    b.u0 = (int)fminf(fmaxf(floorf(umin) - 1.0f, 0.0f), Wm1);
    const int u1 = (int)fminf(fmaxf(floorf(umax) + 2.0f, 0.0f), (float)W);
    b.v0 = (int)fminf(fmaxf(floorf(vmin) - 1.0f, 0.0f), Vm1);
    const int v1 = (int)fminf(fmaxf(floorf(vmax) + 1.0f, 0.0f), Vm1);
    b.cpr = (u1 - b.u0 + 2) >> 1;
    b.rows = v1 - b.v0 + 1;
    return b;
}

// The rectangle origins of every tile of a tile shape (workgroups of ex x ey
// voxel columns, ZC deep) in every view, box[tile * N + view] = u0 | v0 << 16,
// and the largest rectangle: chunks per row into ext[0], strip rows into
// ext[1] (atomic max, one per wave).  Lane group g of the grid takes tile
// g / N, view g % N.
template <int ZC>
__global__ void strip_box_kernel(const Geo* __restrict__ geo, int N, int ntx, int nty, int ntz, int ex, int ey, int X,
                                 int Y, int W, int V, uint32_t* __restrict__ box, int* __restrict__ ext)
{
    const size_t grp = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) >> 2;
    const size_t total = (size_t)ntx * nty * ntz * N;
    const size_t gi = grp < total ? grp : total - 1;  // the lanes of a group stay together
    const int vv = (int)(gi % N);
    const size_t tile = gi / N;
    const int tx = (int)(tile % ntx), tyi = (int)((tile / ntx) % nty), tz = (int)(tile / ((size_t)ntx * nty));
    const float cx0 = (float)min(tx * ex, X - 1), cx1 = (float)min(tx * ex + ex - 1, X - 1);
    const float cy0 = (float)min(tyi * ey, Y - 1), cy1 = (float)min(tyi * ey + ey - 1, Y - 1);
    const StripBox b = strip_footprint<ZC>(geo[vv], cx0, cx1, cy0, cy1, (float)(tz * ZC), W, V);
    if ((threadIdx.x & 3) == 0 && grp < total) box[gi] = (uint32_t)b.u0 | ((uint32_t)b.v0 << 16);
    // the wave's largest rectangle, then one atomic per wave
    int mc = grp < total ? b.cpr : 0, mr = grp < total ? b.rows : 0;
#pragma unroll
    for (int s = 4; s < 64; s <<= 1) {
        mc = max(mc, __shfl_xor(mc, s));
        mr = max(mr, __shfl_xor(mr, s));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(ext, mc);
        atomicMax(ext + 1, mr);
    }
}

// Copy instructions per thread and view: the rectangle holds at most
// kStripRounds(ZC) x 256 chunks (dental-scale footprints of 16 x 16 columns:
// about 19 chunks x (1.44 ZC + 7) rows).
constexpr int kStripRounds(int zc) { return zc >= 32 ? 5 : zc >= 16 ? 3 : 2; }

template <int ZC, int QT, int OUT>
__device__ __forceinline__ void bp_tile_strip_lds(const BPArgs& a, unsigned tile)
{
    static_assert(is_strip<QT>() && (OUT == kOutNormalised || OUT == kOutAccumulate),
                  "column strips, count-normalised forms");
    typedef Sampler<QT> S;
    constexpr int MR = kStripRounds(ZC);
    const int tx = tile % a.ntx;
    const unsigned rest = tile / a.ntx;
    const int tyi = rest % a.nty;
    const int tz = a.tz0 + (int)(rest / a.nty);
    // 16 x 16 voxel columns: four 8 x 8 waves (as bp_tile with four waves)
    const int wx = threadIdx.y & 1, wy = threadIdx.y >> 1;
    const int i = tx * 16 + wx * 8 + (threadIdx.x & 7);
    const int j = tyi * 16 + wy * 8 + (threadIdx.x >> 3);
    const int k0 = tz * ZC;
    const float fi = (float)min(i, a.X - 1);
    const float fj = (float)min(j, a.Y - 1);
    const float fk0 = (float)k0;
    const float Wm1 = (float)(a.W - 1), Vm1 = (float)(a.V - 1);
    const uint32_t rowb = (uint32_t)a.row_px * 8u;  // strip pixels are 8 bytes
    const int n = a.nviews;
    const int stride = a.sl_stride;                   // chunks per rectangle row (odd)
    const uint32_t lrow = (uint32_t)stride * 16u;     // LDS bytes per rectangle row
    const int rect_chunks = stride * a.sl_rows;       // at most MR x 256 (host-checked)
    constexpr uint32_t kBufBytes = MR * 256 * 16;  // two buffers
    unsigned char* const lbuf = reinterpret_cast<unsigned char*>(lds_dyn);
    const int tid = threadIdx.y * 64 + threadIdx.x;
    const __attribute__((address_space(4))) uint32_t* boxes =
        (const __attribute__((address_space(4))) uint32_t*)a.sbox +
        ((size_t)tile + (size_t)a.tz0 * a.ntx * a.nty) * a.nbox_views + a.view0;

    // this thread's chunks c = tid + 256 r of the rectangle, row c / stride,
    // chunk c % stride (chunks past the rectangle copy its last chunk again,
    // to the same place)
    uint32_t goff[MR], loff[MR];
#pragma unroll
    for (int r = 0; r < MR; ++r) {
        const int c = min(tid + 256 * r, rect_chunks - 1);
        const int row = c / stride, col = c - row * stride;
        goff[r] = (uint32_t)row * rowb + (uint32_t)col * 16u;
        loff[r] = (uint32_t)row * lrow + (uint32_t)col * 16u;
    }
    // (views past the last load the last again: every view issues MR loads,
    // so the wait before a store counts statically)
    auto load = [&](int vv, u32x4_t (&regs)[MR]) {
        vv = min(vv, n - 1);
        const uint32_t bx = boxes[vv];
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<char*>(reinterpret_cast<const char*>(a.q)) + (size_t)vv * a.frame_bytes, (short)0,
            (int)a.frame_bytes, kBufferFlags);
        // (the origin goes into the vector offset: the range check of a raw
        // buffer does not count the scalar offset)
        const uint32_t base = (bx >> 16) * rowb + (bx & 0xffffu) * 8u;
#pragma unroll
        for (int r = 0; r < MR; ++r) regs[r] = __builtin_amdgcn_raw_buffer_load_b128(rs, goff[r] + base, 0, 0);
    };
    auto store = [&](int vv, const u32x4_t (&regs)[MR]) {
        unsigned char* const dst = lbuf + (vv & 1) * kBufBytes;
#pragma unroll
        for (int r = 0; r < MR; ++r) *reinterpret_cast<u32x4_t*>(dst + loff[r]) = regs[r];
    };

    float acc[ZC], cpart[ZC];
#pragma unroll
    for (int kk = 0; kk < ZC; ++kk) { acc[kk] = 0.0f; cpart[kk] = 0.0f; }
    // the streaming form continues the running sums (clamped loads and
    // selects, see bp_tile)
    if constexpr (OUT == kOutAccumulate) {
        const size_t plane = (size_t)a.X * a.Y;
        const bool col = i < a.X && j < a.Y;
        const size_t base = ((size_t)k0 * a.Y + min(j, a.Y - 1)) * a.X + min(i, a.X - 1);
#pragma unroll
        for (int kk = 0; kk < ZC; ++kk) {
            const bool ok = col && k0 + kk < a.Z;
            const size_t idx = base + (size_t)min(kk, a.Z - 1 - k0) * plane;
            const float x = a.acc[idx], c = a.cnt[idx];
            acc[kk] = ok ? x : 0.0f;
            cpart[kk] = ok ? c : 0.0f;
        }
    }

    // sample view v from its buffer
    auto sample = [&](int v) {
        const Geo g = geo_of(a.geo, a.view0 + v);
        const float Uc = fmaf(g.c[3], fk0, fmaf(g.c[1], fi, fmaf(g.c[2], fj, g.c[0])));
        const float Vc = fmaf(g.c[7], fk0, fmaf(g.c[5], fi, fmaf(g.c[6], fj, g.c[4])));
        const float Wc = fmaf(g.c[11], fk0, fmaf(g.c[9], fi, fmaf(g.c[10], fj, g.c[8])));
        const float rw = __builtin_amdgcn_rcpf(Wc);
        const float su = fmaf(Uc, rw, 0.5f);
        if (!(su >= 0.0f && su <= Wm1)) return;
        const int iu = (int)su;
        float fe;
        const typename S::wt_t wq = S::weight(__builtin_amdgcn_fractf(su), fe);
        const float sv0 = fmaf(Vc, rw, 0.5f);
        const float dsv = g.c[7] * rw;
        const float svl = fmaf((float)(ZC - 1), dsv, sv0);
        const bool all_in = fminf(sv0, svl) >= 0.0f && fmaxf(sv0, svl) <= Vm1;
        const uint32_t bx = boxes[v];
        // the window of detector row `row` at byte o = row * lrow + lb of the
        // LDS (lb holds the buffer, the column and the rectangle origin)
        const uint32_t lb = (uint32_t)(v & 1) * kBufBytes + (uint32_t)(iu - (int)(bx & 0xffffu)) * 8u -
                            (bx >> 16) * lrow;
        const auto fetch = [&](uint32_t o) -> u32x4_t {
            const uint2* q = reinterpret_cast<const uint2*>(lbuf + o);
            const uint2 x = q[0], y = q[1];
            return (u32x4_t){x.x, x.y, y.x, y.y};
        };
        if (all_in) {
            if (dsv >= 0.0f && dsv <= 0.984375f) {
                bp_column_pairs<ZC, QT, 0, false>(acc, fetch, lrow, lb, sv0, dsv, wq, 0.0f);
            } else if (dsv >= 1.015625f && dsv <= 1.984375f) {
                bp_column_pairs<ZC, QT, 1, false>(acc, fetch, lrow, lb, sv0, dsv, wq, 0.0f);
            } else {
                uint32_t off[ZC];
                float fv[ZC];
#pragma unroll
                for (int k = 0; k < ZC; ++k) {
                    const float sv = fmaf((float)k, dsv, sv0);
                    fv[k] = __builtin_amdgcn_fractf(sv);
                    off[k] = __umul24((uint32_t)(int)sv, lrow) + lb;
                }
                typename S::raw_t p[ZC];
#pragma unroll
                for (int k = 0; k < ZC; ++k) p[k] = fetch(off[k]);
#pragma unroll
                for (int k = 0; k < ZC; ++k) acc[k] = S::accum(acc[k], p[k], fv[k], wq);
            }
            // every sample counts: the column's count weight is in a.colcnt
        } else {
            // the count weight of the samples outside the detector rows is
            // taken off a.colcnt at the end (this load is waited for, and
            // with it the copies in flight: partial views are rare)
            const float2 uwq = a.uwp[(size_t)(a.view0 + v) * a.W + iu];
            const float cw = fmaf(fe, uwq.y - uwq.x, uwq.x);
#pragma unroll
            for (int kb = 0; kb < ZC; kb += kZB) {
                uint32_t off[kZB];
                float fv[kZB];
                bool in[kZB];
#pragma unroll
                for (int q = 0; q < kZB; ++q) {
                    const float sv = fmaf((float)(kb + q), dsv, sv0);
                    const float t = __builtin_amdgcn_fmed3f(sv, 0.0f, Vm1);
                    in[q] = (t == sv);
                    fv[q] = __builtin_amdgcn_fractf(t);
                    off[q] = __umul24((uint32_t)(int)t, lrow) + lb;
                }
                typename S::raw_t p[kZB];
#pragma unroll
                for (int q = 0; q < kZB; ++q) p[q] = fetch(off[q]);
#pragma unroll
                for (int q = 0; q < kZB; ++q) {
                    const float sum = S::accum(acc[kb + q], p[q], fv[q], wq);
                    acc[kb + q] = in[q] ? sum : acc[kb + q];
                    cpart[kb + q] += in[q] ? 0.0f : cw;
                }
            }
        }
    };
    u32x4_t regs[MR];
    if (n > 0) {
        load(0, regs);
        store(0, regs);
        load(1, regs);
    }
    for (int v = 0; v < n; ++v) {
        // buffer v & 1 holds view v (every wave stored its part before the
        // barrier), and every wave is done with view v - 1, whose buffer
        // view v + 1 takes
        __syncthreads();
        sample(v);
        store(v + 1, regs);
        load(v + 2, regs);
    }

    if (i >= a.X || j >= a.Y) return;
    const float c = OUT == kOutNormalised ? a.colcnt[(size_t)j * a.X + i] : 0.0f;
    const size_t plane = (size_t)a.X * a.Y;
    size_t idx = ((size_t)k0 * a.Y + j) * a.X + i;
#pragma unroll
    for (int kk = 0; kk < ZC; ++kk) {
        if (k0 + kk < a.Z) {
            if constexpr (OUT == kOutNormalised) {
                const float v = acc[kk] / (c - cpart[kk]) * a.qscale;
                a.vol[idx] = a.ct ? ct_number(v, a.ct_scale, a.ct_norm1) : v;
            } else {
                a.acc[idx] = acc[kk];
                a.cnt[idx] = cpart[kk];
            }
        }
        idx += plane;
    }
}

// The LDS-staged strip kernel: bp_kernel's tile schedule, 16 x 16 voxel
// columns (64 x 4 threads) per workgroup.
template <int ZC, int QT, int OUT>
__global__ void __launch_bounds__(256) bp_strip_lds_kernel(BPArgs a)
{
    const unsigned ntiles = (unsigned)a.ntx * a.nty * a.ntz;
    const unsigned t8 = gridDim.x >> 3;  // gridDim.x = 8 * ceil(ntiles / 8)
    const unsigned tile = (blockIdx.x & 7u) * t8 + (blockIdx.x >> 3);
    if (tile >= ntiles) return;  // uniform per workgroup
    bp_tile_strip_lds<ZC, QT, OUT>(a, tile);
}

// Tile schedule.  Workgroups are dispatched round-robin over the 8 XCDs, so
// workgroup b runs on XCD b % 8.  XCD x takes the contiguous tile range
// [x T8, (x+1) T8): a z-slab, whose detector rows then stay in that XCD's L2.
// (A persistent grid that sweeps the tiles in rounds was measured slower in
// every tile shape on dental512 and made the compiler emit ~45 % more vector
// instructions, so each workgroup runs exactly one tile.)
// ZC = 16 in geometry class 0 is held to five waves per SIMD (at most 96
// VGPRs; the BF16 strip kernel otherwise takes 97 and runs four): measured
// 2 % faster on dental512, with no spills.  Six waves (80 VGPRs) was slower.
template <int MODE, int ZC, int QT, int OUT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(ZC == 16 && MODE == 0 ? 5 : 1)))
bp_kernel(BPArgs a)
{
    const unsigned ntiles = (unsigned)a.ntx * a.nty * a.ntz;
    const unsigned t8 = gridDim.x >> 3;  // gridDim.x = 8 * ceil(ntiles / 8)
    const unsigned tile = (blockIdx.x & 7u) * t8 + (blockIdx.x >> 3);
    if (tile >= ntiles) return;  // uniform per workgroup
    if constexpr (MODE == kModeStaged) {
        if constexpr (!distance_weighted(OUT)) bp_tile_lds<ZC, QT, OUT>(a, tile);
    } else {
        bp_tile<MODE, ZC, QT, OUT>(a, tile);
    }
}

__global__ void ctnum_kernel(float* __restrict__ x, size_t n, float scale, float norm1)
{
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
        x[t] = ct_number(x[t], scale, norm1);
}

// vol = acc / count (or wscale * acc) of the streaming sums: the count is
// cnt, or in geometry class 0 (colsum given) the column's running count
// weight less cnt -- the operations of the one-launch epilogue.  acc, cnt
// and vol start at a whole plane (so t % plane is the column).
__global__ void bp_finalize_kernel(const float* __restrict__ acc, const float* __restrict__ cnt,
                                   const float* __restrict__ colsum, size_t plane, float* __restrict__ vol,
                                   size_t n, int ct, float ct_scale, float ct_norm1, int weighted, float wscale,
                                   float qscale)
{
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
        float v;
        if (weighted) v = wscale * acc[t] * qscale;
        else v = acc[t] / (colsum ? colsum[t % plane] - cnt[t] : cnt[t]) * qscale;
        vol[t] = ct ? ct_number(v, ct_scale, ct_norm1) : v;
    }
}

__global__ void fill_zero_kernel(float* __restrict__ p, size_t n)
{
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
        p[t] = 0.0f;
}

float bp_stage_scale(const jh_bp_plan* p, int dtype)
{
    return dtype == JH_F16 ? std::ldexp(1.0f, -p->f16exp) : 1.0f;
}

static unsigned grid_for(size_t n, unsigned block)
{
    return (unsigned)std::min<size_t>((n + block - 1) / block, (size_t)65536);
}

typedef void (*bp_kernel_t)(BPArgs);

template <int MODE, int ZC, int OUT>
static bp_kernel_t pick_q(int dtype, int strip)
{
    if (OUT == kOutRaw) return bp_kernel<MODE, ZC, JH_F32, kOutRaw>;
    switch (dtype) {
    case JH_F32: return bp_kernel<MODE, ZC, JH_F32, OUT>;
    case JH_BF16: return strip ? bp_kernel<MODE, ZC, kBF16Strip, OUT> : bp_kernel<MODE, ZC, JH_BF16, OUT>;
    default: return strip ? bp_kernel<MODE, ZC, kF16Strip, OUT> : bp_kernel<MODE, ZC, JH_F16, OUT>;
    }
}

template <int MODE, int OUT>
static bp_kernel_t pick_zc(int zc, int dtype, int strip)
{
    switch (zc) {
    case 8: return pick_q<MODE, 8, OUT>(dtype, strip);
    case 16: return pick_q<MODE, 16, OUT>(dtype, strip);
    default: return pick_q<MODE, 32, OUT>(dtype, strip);
    }
}

// LDS-staged column strips (count-normalised forms only; nullptr otherwise).
template <int OUT>
static bp_kernel_t pick_strip_lds(int zc, int dtype)
{
    if constexpr (OUT == kOutNormalised || OUT == kOutAccumulate) {
        const bool bf = dtype == JH_BF16;
        switch (zc) {
        case 8: return bf ? bp_strip_lds_kernel<8, kBF16Strip, OUT> : bp_strip_lds_kernel<8, kF16Strip, OUT>;
        case 16: return bf ? bp_strip_lds_kernel<16, kBF16Strip, OUT> : bp_strip_lds_kernel<16, kF16Strip, OUT>;
        default: return bf ? bp_strip_lds_kernel<32, kBF16Strip, OUT> : bp_strip_lds_kernel<32, kF16Strip, OUT>;
        }
    }
    return nullptr;
}

// The LDS-staged strip kernel's rectangle table for the plan's tile shape
// (16 x 16 voxel columns, zc deep): built once per shape on the device with
// the kernel's own arithmetic (origins of every tile and view, the largest
// rectangle), kept in the plan.  True when the kernel can run this shape:
// workgroups of 4 waves, a rectangle of at most kStripRounds(zc) x 256 chunks,
// 16-bit origins; false also where the table cannot be built now (the stream
// is capturing a graph): the launch then gathers.
static bool strip_table(jh_bp_plan* p, hipStream_t stream)
{
    if (p->ty != 4 || p->W > 65535 || p->V > 65535) return false;
    if (p->sl_zc == p->zc) return p->sl_stride > 0;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    JH_HIP(hipStreamIsCapturing(stream, &cs));
    if (cs != hipStreamCaptureStatusNone) return false;
    const int ex = 16, ey = 16;
    const int ntx = (p->X + ex - 1) / ex, nty = (p->Y + ey - 1) / ey, ntz = (p->Z + p->zc - 1) / p->zc;
    const size_t entries = (size_t)ntx * nty * ntz * p->N;
    const size_t lanes = entries * 4;
    require(lanes / 256 < (1ull << 31), "backprojection footprint grid too large");
    p->sl_zc = -1;
    p->sl_stride = 0;
    p->sbox.release();
    p->sbox.alloc(entries * sizeof(uint32_t));
    DevBuf d(2 * sizeof(int));
    const int zero[2] = {0, 0};
    JH_HIP(hipMemcpyAsync(d.get(), zero, sizeof(zero), hipMemcpyHostToDevice, stream));
    const unsigned grid = (unsigned)((lanes + 255) / 256);
    const Geo* geo = p->geo.as<Geo>();
    uint32_t* box = p->sbox.as<uint32_t>();
    if (p->zc >= 32)
        strip_box_kernel<32><<<grid, 256, 0, stream>>>(geo, p->N, ntx, nty, ntz, ex, ey, p->X, p->Y, p->W, p->V, box,
                                                       d.as<int>());
    else if (p->zc >= 16)
        strip_box_kernel<16><<<grid, 256, 0, stream>>>(geo, p->N, ntx, nty, ntz, ex, ey, p->X, p->Y, p->W, p->V, box,
                                                       d.as<int>());
    else
        strip_box_kernel<8><<<grid, 256, 0, stream>>>(geo, p->N, ntx, nty, ntz, ex, ey, p->X, p->Y, p->W, p->V, box,
                                                      d.as<int>());
    check_launch("strip_box_kernel");
    int h[2] = {0, 0};
    JH_HIP(hipMemcpyAsync(h, d.get(), sizeof(h), hipMemcpyDeviceToHost, stream));
    JH_HIP(hipStreamSynchronize(stream));
    const int stride = h[0] | 1;  // odd: consecutive LDS rows start in other banks
    p->sl_zc = p->zc;
    p->sl_rows = h[1];
    p->sl_stride = (h[0] >= 1 && h[1] >= 1 && (size_t)stride * h[1] <= (size_t)kStripRounds(p->zc) * 256) ? stride : 0;
    if (p->sl_stride == 0) p->sbox.release();
    return p->sl_stride > 0;
}

template <int OUT>
static bp_kernel_t pick_kernel(int mode, int zc, int dtype, int fetch, int strip)
{
    switch (mode) {
    case 0:
        // the distance-weighted forms have no LDS-staged fetch: they gather
        if constexpr (!distance_weighted(OUT))
            if (fetch == JH_FETCH_LDS) {
                if (strip && dtype != JH_F32)
                    if (bp_kernel_t k = pick_strip_lds<OUT>(zc, dtype)) return k;
                return pick_zc<kModeStaged, OUT>(zc, dtype, strip);
            }
        return pick_zc<0, OUT>(zc, dtype, strip);
    case 1: return pick_zc<1, OUT>(zc, dtype, strip);
    default: return pick_zc<2, OUT>(zc, dtype, strip);
    }
}

// Staged layout of one view: bytes per pixel (FP32 quad 16, BF16 / FP16 quad
// or column strip 8) and pixels per row (column strips carry a zero padding
// column).
static int stage_pixel_bytes(const jh_bp_plan* p, int dtype)
{
    (void)p;
    return dtype == JH_F32 ? 16 : 8;
}

static bool stage_strip(const jh_bp_plan* p, int dtype) { return dtype != JH_F32 && p->strip; }

static size_t stage_frame_bytes(const jh_bp_plan* p, int dtype)
{
    return (size_t)(p->W + (stage_strip(p, dtype) ? 1 : 0)) * p->V * stage_pixel_bytes(p, dtype);
}

// acc, cnt, colsum: the streaming sums (default: the plan's).
// tz0, ntz: the launch's z tiles (a z-slab of the streaming forms; the column
// count sums are continued only by the slab that starts at z = 0); ntz < 0:
// every tile from tz0 on.
static void launch_bp(jh_bp_plan* p, const void* q, int dtype, int view0, int nviews, float* vol,
                      int out, hipStream_t stream, float* acc = nullptr, float* cnt = nullptr,
                      float* colsum = nullptr, int tz0 = 0, int ntz = -1)
{
    BPArgs a;
    a.q = q;
    a.geo = p->geo.as<Geo>();
    a.uwp = p->uwp.as<float2>();
    a.colcnt = p->colcnt.as<float>();
    a.vol = vol;
    a.acc = acc ? acc : p->acc.as<float>();
    a.cnt = cnt ? cnt : p->cnt.as<float>();
    a.W = p->W; a.V = p->V; a.X = p->X; a.Y = p->Y; a.Z = p->Z;
    a.view0 = view0; a.nviews = nviews;
    a.ct = p->ct;
    a.ct_scale = p->ct_scale;
    a.ct_norm1 = p->ct_norm1;
    a.wscale = p->wscale;
    a.qscale = dtype == JH_F16 ? std::ldexp(1.0f, p->f16exp) : 1.0f;
    // workgroup extent: 8 x 8 voxel columns per wave (see bp_tile)
    const int ex = p->ty >= 2 ? 16 : 8, ey = p->ty >= 4 ? 16 : 8;
    a.ntx = (p->X + ex - 1) / ex;
    a.nty = (p->Y + ey - 1) / ey;
    const int ntz_all = (p->Z + p->zc - 1) / p->zc;
    require(tz0 >= 0 && tz0 < ntz_all && (ntz < 0 || tz0 + ntz <= ntz_all), "backprojection: z tiles outside the volume");
    a.tz0 = tz0;
    a.ntz = ntz < 0 ? ntz_all - tz0 : ntz;
    const bool slab = a.ntz < ntz_all;
    require(!slab || out == kOutAccumulate || out == kOutFdkAccumulate, "backprojection: z-slabs are streaming only");
    const size_t fb = stage_frame_bytes(p, dtype);
    a.pxb = stage_pixel_bytes(p, dtype);
    a.row_px = p->W + (stage_strip(p, dtype) ? 1 : 0);
    require(fb < (1ull << 31), "backprojection frame too large");
    a.frame_bytes = (uint32_t)fb;
    const size_t ntiles = (size_t)a.ntx * a.nty * a.ntz;
    require(ntiles < (1ull << 31), "backprojection grid too large");
    const size_t grid = (ntiles + 7) / 8 * 8;
    const int st = stage_strip(p, dtype) ? 1 : 0;
    // LDS-staged fetch.  Column strips take the strip kernel, whose two
    // buffers hold the largest footprint of the tile shape (rounded up to the
    // copy's 256-chunk rounds); where that exceeds the LDS the launch gathers.
    // Other layouts take the coefficient kernel: two buffers sized for the
    // block footprints of this tile depth (a larger footprint gathers), 64 zc
    // pixels each, capped so that both, with the kernel's static box ring and
    // count-weight rows (4 KiB + 48 B), stay within 64 KiB.
    int fetch = p->fetch;
    bool strip_lds = p->mode == 0 && fetch == JH_FETCH_LDS && st && (out == kOutNormalised || out == kOutAccumulate);
    if (strip_lds && !strip_table(p, stream)) {
        strip_lds = false;
        fetch = JH_FETCH_GATHER;
    }
    const bool staged = p->mode == 0 && fetch == JH_FETCH_LDS && !distance_weighted(out) && !strip_lds;
    bp_kernel_t k = out == kOutNormalised      ? pick_kernel<kOutNormalised>(p->mode, p->zc, dtype, fetch, st)
                    : out == kOutAccumulate    ? pick_kernel<kOutAccumulate>(p->mode, p->zc, dtype, fetch, st)
                    : out == kOutFdk           ? pick_kernel<kOutFdk>(p->mode, p->zc, dtype, fetch, st)
                    : out == kOutFdkAccumulate ? pick_kernel<kOutFdkAccumulate>(p->mode, p->zc, dtype, fetch, st)
                                               : pick_kernel<kOutRaw>(p->mode, p->zc, dtype, fetch, st);
    a.lds_px = staged ? std::min(64 * p->zc, (65536 - 4096 - 64) / 32) : 0;
    a.sbox = strip_lds ? p->sbox.as<uint32_t>() : nullptr;
    a.nbox_views = p->N;
    a.sl_stride = strip_lds ? p->sl_stride : 0;
    a.sl_rows = strip_lds ? p->sl_rows : 0;
    // (two rectangle buffers of kStripRounds(zc) x 256 chunks of 16 bytes)
    const size_t lds_bytes = strip_lds ? 2 * (size_t)kStripRounds(p->zc) * 256 * 16
                             : staged  ? 2 * (size_t)a.lds_px * sizeof(float4) : 0;
    if (p->mode == 0 && (out == kOutNormalised || (out == kOutAccumulate && tz0 == 0))) {
        // the count weights of the launch's views per voxel column; the
        // streaming form continues the running sums (once per view: in the
        // launch whose tiles start at z = 0)
        float* cc = out == kOutAccumulate ? (colsum ? colsum : p->colsum.as<float>()) : p->colcnt.as<float>();
        const int cont = out == kOutAccumulate ? 1 : 0;
        const dim3 cg((unsigned)((p->X + 63) / 64), (unsigned)((p->Y + 3) / 4));
        const dim3 cb(64, 4);
        if (dtype == JH_F32)
            bp_colcount_kernel<JH_F32><<<cg, cb, 0, stream>>>(a.geo, a.uwp, view0, nviews, p->X, p->Y, p->W, cc, cont);
        else if (dtype == JH_BF16)
            bp_colcount_kernel<JH_BF16><<<cg, cb, 0, stream>>>(a.geo, a.uwp, view0, nviews, p->X, p->Y, p->W, cc, cont);
        else
            bp_colcount_kernel<JH_F16><<<cg, cb, 0, stream>>>(a.geo, a.uwp, view0, nviews, p->X, p->Y, p->W, cc, cont);
        check_launch("bp_colcount_kernel");
    }
    // (the strip kernel's workgroups carry one more wave, its copy wave)
    hipLaunchKernelGGL(k, dim3((unsigned)grid), dim3(64, p->ty), lds_bytes, stream, a);
    check_launch("bp_kernel");
}

}  // namespace jh

using namespace jh;

extern "C" {

jh_status jh_bp_create(jh_bp_plan** plan, int device, const double* h_ptm, int nviews, int W,
                       int V, int X, int Y, int Z, float pxy, float pz)
{
    return guard([&] {
        require(plan != nullptr, "jh_bp_create: plan is NULL");
        *plan = nullptr;
        require(h_ptm != nullptr && nviews > 0, "jh_bp_create: no projection matrices");
        require(W >= 2 && V >= 2, "jh_bp_create: detector must be at least 2 x 2");
        require(X > 0 && Y > 0 && Z > 0, "jh_bp_create: empty volume");
        require((size_t)W * V < (1ull << 31), "jh_bp_create: detector too large");
        require(pxy > 0.0f && pz > 0.0f, "jh_bp_create: pitches must be positive");
        int n = 0;
        JH_HIP(hipGetDeviceCount(&n));
        require(device >= 0 && device < n, "jh_bp_create: no such device");
        DeviceScope ds(device);
        auto* p = new jh_bp_plan;
        try {
            p->device = device;
            p->N = nviews; p->W = W; p->V = V; p->X = X; p->Y = Y; p->Z = Z;
            p->pxy = pxy; p->pz = pz;
            p->zc = 32;
            p->ty = 4;
            p->fetch = JH_FETCH_GATHER;
            p->strip = 0;
            p->zfilter = 0;
            p->ct = 0;
            p->ct_scale = 1.0f;
            p->ct_norm1 = 1.0f;
            p->weight = JH_WEIGHT_COUNT;
            p->wscale = 2.0f / (float)nviews;
            p->f16exp = 0;
            p->acc_scale = 1.0f;
            p->acc_dtype = -1;
            p->sl_zc = -1;
            p->sl_stride = 0;
            p->sl_rows = 0;
            p->ptm.alloc(sizeof(double) * 12 * nviews);
            p->geo.alloc(sizeof(Geo) * nviews);
            p->uw.alloc(sizeof(float) * (size_t)nviews * W);
            JH_HIP(hipMemcpy(p->ptm.get(), h_ptm, sizeof(double) * 12 * nviews, hipMemcpyHostToDevice));
            bp_geometry_kernel<<<(nviews + 63) / 64, 64>>>(p->ptm.as<double>(), nviews, X, Y, Z, pxy, pz,
                                                            p->geo.as<Geo>());
            check_launch("bp_geometry_kernel");
            bp_uweight_kernel<<<dim3((W + 255) / 256, nviews), 256>>>(p->ptm.as<double>(), nviews, W,
                                                                       p->uw.as<float>());
            check_launch("bp_uweight_kernel");
            p->uwp.alloc(sizeof(float2) * (size_t)nviews * W);
            bp_uwpair_kernel<<<grid_for((size_t)nviews * W, 256), 256>>>(p->uw.as<float>(), nviews, W,
                                                                          p->uwp.as<float2>());
            check_launch("bp_uwpair_kernel");
            std::vector<Geo> g(nviews);
            JH_HIP(hipMemcpy(g.data(), p->geo.get(), sizeof(Geo) * nviews, hipMemcpyDeviceToHost));
            bool zinv_u = true, zinv_w = true;
            for (const Geo& e : g) {
                if (e.c[3] != 0.0f) zinv_u = false;
                if (e.c[11] != 0.0f) zinv_w = false;
            }
            p->mode = (zinv_u && zinv_w) ? 0 : (zinv_w ? 1 : 2);
            if (p->mode == 0) p->colcnt.alloc(sizeof(float) * (size_t)X * Y);
            // BF16 / FP16 layout: column strips for geometry class 0 (the
            // backprojector's pair path needs them; where the row step rules
            // pairs out they measure as fast as quads), quads otherwise
            p->strip = p->mode == 0 ? 1 : 0;
        } catch (...) {
            delete p;
            throw;
        }
        *plan = p;
    });
}

jh_status jh_bp_uweights(jh_bp_plan* p, float* h_out)
{
    return guard([&] {
        require(p != nullptr && h_out != nullptr, "jh_bp_uweights: bad arguments");
        DeviceScope ds(p->device);
        JH_HIP(hipMemcpy(h_out, p->uw.get(), sizeof(float) * (size_t)p->N * p->W, hipMemcpyDeviceToHost));
    });
}

jh_status jh_bp_stage_bytes(jh_bp_plan* p, int dtype, int nviews, size_t* bytes)
{
    return guard([&] {
        require(p != nullptr && bytes != nullptr && nviews >= 0, "jh_bp_stage_bytes: bad arguments");
        require(dtype == JH_F32 || dtype == JH_BF16 || dtype == JH_F16, "jh_bp_stage_bytes: unknown dtype");
        *bytes = stage_frame_bytes(p, dtype) * nviews;
    });
}

static void stage(jh_bp_plan* p, const float* d_filtered, int view0, int nviews, int dtype, void* d_q,
                  bool weighted, jh_stream_t stream)
{
    {
        require(p != nullptr && d_filtered != nullptr && d_q != nullptr, "jh_bp_stage: bad arguments");
        require(view0 >= 0 && nviews >= 0 && view0 + nviews <= p->N, "jh_bp_stage: view range outside the scan");
        require(dtype == JH_F32 || dtype == JH_BF16 || dtype == JH_F16, "jh_bp_stage: unknown dtype");
        if (nviews == 0) return;
        DeviceScope ds(p->device);
        require(p->V <= 65535, "jh_bp_stage: more than 65535 detector rows");
        hipStream_t s = as_stream(stream);
        const int zf = weighted ? p->zfilter : 0;
        const float qs = bp_stage_scale(p, dtype);
        const size_t fin = (size_t)p->W * p->V, fout = stage_frame_bytes(p, dtype);
        const int st = stage_strip(p, dtype) ? 1 : 0;
        for (int c0 = 0; c0 < nviews; c0 += 65535) {  // grid z: frames
            const int n = std::min(nviews - c0, 65535);
            const float* uw = weighted ? p->uw.as<float>() + (size_t)(view0 + c0) * p->W : nullptr;
            const float* in = d_filtered + (size_t)c0 * fin;
            void* out = static_cast<char*>(d_q) + (size_t)c0 * fout;
            const dim3 grid((unsigned)((p->W + 63) / 64), (unsigned)((p->V + 4 * kStageRun - 1) / (4 * kStageRun)),
                            (unsigned)n);
            const dim3 block(64, 4);
            if (dtype == JH_F32) bp_stage_kernel<JH_F32, false><<<grid, block, 0, s>>>(in, uw, p->W, p->V, zf, qs, out);
            else if (dtype == JH_BF16 && st)
                bp_stage_kernel<JH_BF16, true><<<grid, block, 0, s>>>(in, uw, p->W, p->V, zf, qs, out);
            else if (dtype == JH_BF16)
                bp_stage_kernel<JH_BF16, false><<<grid, block, 0, s>>>(in, uw, p->W, p->V, zf, qs, out);
            else if (st) bp_stage_kernel<JH_F16, true><<<grid, block, 0, s>>>(in, uw, p->W, p->V, zf, qs, out);
            else bp_stage_kernel<JH_F16, false><<<grid, block, 0, s>>>(in, uw, p->W, p->V, zf, qs, out);
        }
        check_launch("bp_stage_kernel");
    }
}

jh_status jh_bp_stage(jh_bp_plan* p, const float* d_filtered, int view0, int nviews, int dtype,
                      void* d_q, jh_stream_t stream)
{
    return guard([&] { stage(p, d_filtered, view0, nviews, dtype, d_q, true, stream); });
}

jh_status jh_bp_pack(jh_bp_plan* p, const float* d_frames, int view0, int nviews, int dtype, void* d_q,
                     jh_stream_t stream)
{
    return guard([&] { stage(p, d_frames, view0, nviews, dtype, d_q, false, stream); });
}

jh_status jh_bp_backproject(jh_bp_plan* p, const void* d_q, int view0, int nviews, float* d_out,
                            jh_stream_t stream)
{
    return guard([&] {
        require(p != nullptr && d_q != nullptr && d_out != nullptr, "jh_bp_backproject: bad arguments");
        require(view0 >= 0 && nviews >= 0 && view0 + nviews <= p->N,
                "jh_bp_backproject: view range outside the scan");
        DeviceScope ds(p->device);
        launch_bp(p, d_q, JH_F32, view0, nviews, d_out, kOutRaw, as_stream(stream));
    });
}

jh_status jh_bp_run(jh_bp_plan* p, const void* d_q, int dtype, float* d_vol, jh_stream_t stream)
{
    return guard([&] {
        require(p != nullptr && d_q != nullptr && d_vol != nullptr, "jh_bp_run: bad arguments");
        require(dtype == JH_F32 || dtype == JH_BF16 || dtype == JH_F16, "jh_bp_run: unknown dtype");
        DeviceScope ds(p->device);
        launch_bp(p, d_q, dtype, 0, p->N, d_vol, p->weight == JH_WEIGHT_FDK ? kOutFdk : kOutNormalised,
                  as_stream(stream));
    });
}

static void ensure_accumulators(jh_bp_plan* p)
{
    const size_t n = (size_t)p->X * p->Y * p->Z;
    if (!p->acc.get()) p->acc.alloc(sizeof(float) * n);
    if (!p->cnt.get()) p->cnt.alloc(sizeof(float) * n);
    if (p->mode == 0 && !p->colsum.get()) p->colsum.alloc(sizeof(float) * (size_t)p->X * p->Y);
}

jh_status jh_bp_reset(jh_bp_plan* p, jh_stream_t stream)
{
    return guard([&] {
        require(p != nullptr, "jh_bp_reset: plan is NULL");
        DeviceScope ds(p->device);
        ensure_accumulators(p);
        const size_t n = (size_t)p->X * p->Y * p->Z;
        hipStream_t s = as_stream(stream);
        fill_zero_kernel<<<grid_for(n, 256), 256, 0, s>>>(p->acc.as<float>(), n);
        fill_zero_kernel<<<grid_for(n, 256), 256, 0, s>>>(p->cnt.as<float>(), n);
        if (p->mode == 0) {
            const size_t nc = (size_t)p->X * p->Y;
            fill_zero_kernel<<<grid_for(nc, 256), 256, 0, s>>>(p->colsum.as<float>(), nc);
        }
        check_launch("fill_zero_kernel");
        p->acc_scale = 1.0f;
        p->acc_dtype = -1;
    });
}

jh_status jh_bp_accumulate(jh_bp_plan* p, const void* d_q, int dtype, int view0, int nviews,
                           jh_stream_t stream)
{
    return guard([&] {
        require(p != nullptr && d_q != nullptr, "jh_bp_accumulate: bad arguments");
        require(dtype == JH_F32 || dtype == JH_BF16 || dtype == JH_F16, "jh_bp_accumulate: unknown dtype");
        require(view0 >= 0 && nviews >= 0 && view0 + nviews <= p->N, "jh_bp_accumulate: view range outside the scan");
        require(p->acc.get() != nullptr, "jh_bp_accumulate: call jh_bp_reset first");
        if (nviews == 0) return;
        // the accumulators hold one output scale: every chunk since the
        // reset must be staged alike
        const float sc = dtype == JH_F16 ? std::ldexp(1.0f, p->f16exp) : 1.0f;
        require(p->acc_dtype < 0 || sc == p->acc_scale,
                "jh_bp_accumulate: chunks of differently scaled storage since the last reset");
        p->acc_scale = sc;
        p->acc_dtype = dtype;
        DeviceScope ds(p->device);
        launch_bp(p, d_q, dtype, view0, nviews, nullptr,
                  p->weight == JH_WEIGHT_FDK ? kOutFdkAccumulate : kOutAccumulate, as_stream(stream));
    });
}

jh_status jh_bp_accumulate_slab(jh_bp_plan* p, const void* d_q, int dtype, int view0, int nviews, int z0, int nz,
                                jh_stream_t stream)
{
    return guard([&] {
        require(p != nullptr && d_q != nullptr, "jh_bp_accumulate_slab: bad arguments");
        require(dtype == JH_F32 || dtype == JH_BF16 || dtype == JH_F16, "jh_bp_accumulate_slab: unknown dtype");
        require(view0 >= 0 && nviews >= 0 && view0 + nviews <= p->N,
                "jh_bp_accumulate_slab: view range outside the scan");
        require(z0 >= 0 && nz > 0 && z0 + nz <= p->Z, "jh_bp_accumulate_slab: slice range outside the volume");
        require(z0 % p->zc == 0 && (nz % p->zc == 0 || z0 + nz == p->Z),
                "jh_bp_accumulate_slab: slab not aligned to the backprojector's z tile (jh_bp_get_tile)");
        require(p->acc.get() != nullptr, "jh_bp_accumulate_slab: call jh_bp_reset first");
        if (nviews == 0) return;
        const float sc = dtype == JH_F16 ? std::ldexp(1.0f, p->f16exp) : 1.0f;
        require(p->acc_dtype < 0 || sc == p->acc_scale,
                "jh_bp_accumulate_slab: chunks of differently scaled storage since the last reset");
        p->acc_scale = sc;
        p->acc_dtype = dtype;
        DeviceScope ds(p->device);
        launch_bp(p, d_q, dtype, view0, nviews, nullptr,
                  p->weight == JH_WEIGHT_FDK ? kOutFdkAccumulate : kOutAccumulate, as_stream(stream), nullptr,
                  nullptr, nullptr, z0 / p->zc, (nz + p->zc - 1) / p->zc);
    });
}

jh_status jh_bp_finalize(jh_bp_plan* p, float* d_vol, int z0, int nz, jh_stream_t stream)
{
    return guard([&] {
        require(p != nullptr && d_vol != nullptr, "jh_bp_finalize: bad arguments");
        require(z0 >= 0 && nz >= 0 && z0 + nz <= p->Z, "jh_bp_finalize: slice range outside the volume");
        require(p->acc.get() != nullptr, "jh_bp_finalize: call jh_bp_reset first");
        if (nz == 0) return;
        DeviceScope ds(p->device);
        const size_t plane = (size_t)p->X * p->Y;
        const size_t n = plane * nz;
        const bool wt = p->weight == JH_WEIGHT_FDK;
        bp_finalize_kernel<<<grid_for(n, 256), 256, 0, as_stream(stream)>>>(
            p->acc.as<float>() + plane * z0, p->cnt.as<float>() + plane * z0,
            p->mode == 0 && !wt ? p->colsum.as<float>() : nullptr, plane, d_vol, n, p->ct, p->ct_scale, p->ct_norm1,
            wt ? 1 : 0, p->wscale, p->acc_scale);
        check_launch("bp_finalize_kernel");
    });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Allocation-free streaming accumulation (jh_bp_accumulate_noalloc).
//
// The accumulator and count volumes (kNoAllocVoxels voxels each) and the
// running column count weights (kNoAllocColumns columns) live in one
// file-scope device array, `scratch`, instead of plan-owned buffers: every GPU
// the library runs on holds its own copy (72 MiB) until the process ends.  Host code never passes the array's name to a kernel (in host code it
// is a stand-in, not a device address): the device address is looked up once
// per device with hipGetSymbolAddress, with that device current, and kept.
// ---------------------------------------------------------------------------

namespace jh {

constexpr size_t kNoAllocVoxels = size_t(1) << 23;   // 8 Mi voxels: e.g. 256 x 256 x 128
constexpr size_t kNoAllocColumns = size_t(1) << 21;  // 2 Mi voxel columns: e.g. 1024 x 2048

static __device__ __align__(256) float scratch[2 * kNoAllocVoxels + kNoAllocColumns];

constexpr int kMaxDevices = 64;

struct NoAllocState {
    std::mutex m;
    float* ptr[kMaxDevices] = {};    // device address of `scratch`, per device
    // the open accumulation sequence of each device (host-side bookkeeping)
    bool open[kMaxDevices] = {};
    size_t nvox[kMaxDevices] = {};
    float scale[kMaxDevices] = {};
    int weight[kMaxDevices] = {};
    int mode[kMaxDevices] = {};      // geometry class: what the count volume holds
};

static NoAllocState& noalloc_state()
{
    static NoAllocState s;
    return s;
}

// Device address of `scratch` on the current device `dev` (looked up once).
static float* scratch_on(int dev)
{
    NoAllocState& s = noalloc_state();
    require(dev >= 0 && dev < kMaxDevices, "jh_bp_accumulate_noalloc: device index beyond the supported range");
    if (!s.ptr[dev]) {
        void* ptr = nullptr;
        JH_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        require(ptr != nullptr, "jh_bp_accumulate_noalloc: no device address for the scratch array");
        s.ptr[dev] = static_cast<float*>(ptr);
    }
    return s.ptr[dev];
}

}  // namespace jh

extern "C" {

jh_status jh_bp_accumulate_noalloc(jh_bp_plan* p, const void* d_q, int dtype, int view0, int nviews, int flags,
                                   float* d_vol, jh_stream_t stream)
{
    return guard([&] {
        require(p != nullptr, "jh_bp_accumulate_noalloc: plan is NULL");
        require(dtype == JH_F32 || dtype == JH_BF16 || dtype == JH_F16, "jh_bp_accumulate_noalloc: unknown dtype");
        require((flags & ~(JH_NOALLOC_FIRST | JH_NOALLOC_LAST)) == 0, "jh_bp_accumulate_noalloc: unknown flags");
        require(view0 >= 0 && nviews >= 0 && view0 + nviews <= p->N,
                "jh_bp_accumulate_noalloc: view range outside the scan");
        require(nviews == 0 || d_q != nullptr, "jh_bp_accumulate_noalloc: no staged views");
        require(!(flags & JH_NOALLOC_LAST) || d_vol != nullptr, "jh_bp_accumulate_noalloc: no output volume");
        const size_t nvox = (size_t)p->X * p->Y * p->Z;
        require(nvox <= kNoAllocVoxels,
                "jh_bp_accumulate_noalloc: volume of " + std::to_string(nvox) + " voxels exceeds the " +
                    std::to_string(kNoAllocVoxels) + "-voxel scratch accumulators");
        const size_t ncol = (size_t)p->X * p->Y;
        require(ncol <= kNoAllocColumns, "jh_bp_accumulate_noalloc: " + std::to_string(ncol) +
                                             " voxel columns exceed the " + std::to_string(kNoAllocColumns) +
                                             "-column scratch count weights");
        DeviceScope ds(p->device);
        NoAllocState& st = noalloc_state();
        std::lock_guard<std::mutex> lock(st.m);
        float* const acc = scratch_on(p->device);
        float* const cnt = acc + kNoAllocVoxels;
        float* const colsum = cnt + kNoAllocVoxels;
        const int dev = p->device;
        const float sc = dtype == JH_F16 ? std::ldexp(1.0f, p->f16exp) : 1.0f;
        const bool wt = p->weight == JH_WEIGHT_FDK;
        hipStream_t s = as_stream(stream);
        if (flags & JH_NOALLOC_FIRST) {
            // the array keeps the previous call's contents: clear it here
            // (a fill kernel, so that a captured graph holds no memset node)
            fill_zero_kernel<<<grid_for(nvox, 256), 256, 0, s>>>(acc, nvox);
            if (!wt) fill_zero_kernel<<<grid_for(nvox, 256), 256, 0, s>>>(cnt, nvox);
            if (!wt && p->mode == 0) fill_zero_kernel<<<grid_for(ncol, 256), 256, 0, s>>>(colsum, ncol);
            check_launch("fill_zero_kernel");
            st.open[dev] = true;
            st.nvox[dev] = nvox;
            st.scale[dev] = sc;
            st.weight[dev] = p->weight;
            st.mode[dev] = p->mode;
        } else {
            require(st.open[dev], "jh_bp_accumulate_noalloc: no open sequence on this device (pass JH_NOALLOC_FIRST)");
            require(st.nvox[dev] == nvox && st.weight[dev] == p->weight && st.mode[dev] == p->mode &&
                        (nviews == 0 || st.scale[dev] == sc),
                    "jh_bp_accumulate_noalloc: the call does not continue the open sequence (volume, weighting, "
                    "geometry class or storage scale differ)");
        }
        if (nviews > 0)
            launch_bp(p, d_q, dtype, view0, nviews, nullptr, wt ? kOutFdkAccumulate : kOutAccumulate, s, acc, cnt,
                      colsum);
        if (flags & JH_NOALLOC_LAST) {
            bp_finalize_kernel<<<grid_for(nvox, 256), 256, 0, s>>>(
                acc, cnt, !wt && p->mode == 0 ? colsum : nullptr, ncol, d_vol, nvox, p->ct, p->ct_scale, p->ct_norm1,
                wt ? 1 : 0, p->wscale, st.scale[dev]);
            check_launch("bp_finalize_kernel");
            st.open[dev] = false;
        }
    });
}

jh_status jh_bp_set_tile(jh_bp_plan* p, int zc, int ty)
{
    return guard([&] {
        require(p != nullptr, "jh_bp_set_tile: plan is NULL");
        require(zc == 0 || zc == 8 || zc == 16 || zc == 32, "jh_bp_set_tile: zc must be 8, 16 or 32");
        require(ty == 0 || ty == 1 || ty == 2 || ty == 4, "jh_bp_set_tile: ty must be 1, 2 or 4");
        if (zc) p->zc = zc;
        if (ty) p->ty = ty;
    });
}

jh_status jh_bp_set_layout(jh_bp_plan* p, int layout)
{
    return guard([&] {
        require(p != nullptr, "jh_bp_set_layout: plan is NULL");
        require(layout == JH_LAYOUT_QUAD || layout == JH_LAYOUT_STRIP, "jh_bp_set_layout: unknown layout");
        p->strip = layout == JH_LAYOUT_STRIP ? 1 : 0;
    });
}

jh_status jh_bp_get_layout(jh_bp_plan* p, int* layout)
{
    return guard([&] {
        require(p != nullptr && layout != nullptr, "jh_bp_get_layout: bad arguments");
        *layout = p->strip ? JH_LAYOUT_STRIP : JH_LAYOUT_QUAD;
    });
}

jh_status jh_bp_set_fetch(jh_bp_plan* p, int fetch)
{
    return guard([&] {
        require(p != nullptr, "jh_bp_set_fetch: plan is NULL");
        require(fetch == JH_FETCH_GATHER || fetch == JH_FETCH_LDS, "jh_bp_set_fetch: unknown fetch form");
        p->fetch = fetch;
    });
}

jh_status jh_bp_get_fetch(jh_bp_plan* p, int* fetch)
{
    return guard([&] {
        require(p != nullptr && fetch != nullptr, "jh_bp_get_fetch: bad arguments");
        *fetch = p->fetch;
    });
}

jh_status jh_bp_fetch_in_use(jh_bp_plan* p, int dtype, jh_stream_t stream, int* fetch)
{
    return guard([&] {
        require(p != nullptr && fetch != nullptr, "jh_bp_fetch_in_use: bad arguments");
        require(dtype == JH_F32 || dtype == JH_BF16 || dtype == JH_F16, "jh_bp_fetch_in_use: unknown dtype");
        DeviceScope ds(p->device);
        int f = p->mode == 0 ? p->fetch : JH_FETCH_GATHER;
        if (f == JH_FETCH_LDS && stage_strip(p, dtype) && !strip_table(p, as_stream(stream))) f = JH_FETCH_GATHER;
        *fetch = f;
    });
}

jh_status jh_bp_set_zfilter(jh_bp_plan* p, int enable)
{
    return guard([&] {
        require(p != nullptr, "jh_bp_set_zfilter: plan is NULL");
        p->zfilter = enable ? 1 : 0;
    });
}

jh_status jh_bp_set_f16_exponent(jh_bp_plan* p, int exponent)
{
    return guard([&] {
        require(p != nullptr, "jh_bp_set_f16_exponent: plan is NULL");
        require(exponent >= -64 && exponent <= 64, "jh_bp_set_f16_exponent: exponent outside [-64, 64]");
        p->f16exp = exponent;
    });
}

jh_status jh_bp_get_f16_exponent(jh_bp_plan* p, int* exponent)
{
    return guard([&] {
        require(p != nullptr && exponent != nullptr, "jh_bp_get_f16_exponent: bad arguments");
        *exponent = p->f16exp;
    });
}

jh_status jh_bp_set_weight(jh_bp_plan* p, int weight, float scale)
{
    return guard([&] {
        require(p != nullptr, "jh_bp_set_weight: plan is NULL");
        require(weight == JH_WEIGHT_COUNT || weight == JH_WEIGHT_FDK, "jh_bp_set_weight: unknown weighting");
        require(std::isfinite(scale) && scale >= 0.0f, "jh_bp_set_weight: scale must be finite and >= 0");
        p->weight = weight;
        p->wscale = scale > 0.0f ? scale : 2.0f / (float)p->N;
    });
}

jh_status jh_bp_get_weight(jh_bp_plan* p, int* weight, float* scale)
{
    return guard([&] {
        require(p != nullptr, "jh_bp_get_weight: plan is NULL");
        if (weight) *weight = p->weight;
        if (scale) *scale = p->wscale;
    });
}

jh_status jh_bp_set_ctnum(jh_bp_plan* p, int enable, float norm0, float norm1, float muWater)
{
    return guard([&] {
        require(p != nullptr, "jh_bp_set_ctnum: plan is NULL");
        require(!enable || (muWater != 0.0f && std::isfinite(norm0) && std::isfinite(norm1) && std::isfinite(muWater)),
                "jh_bp_set_ctnum: bad CT-number parameters");
        p->ct = enable ? 1 : 0;
        p->ct_scale = norm0 * 1000.f / muWater;  // in float, as CTNumCore.cpp:24 forms it
        p->ct_norm1 = norm1;
    });
}

jh_status jh_zfilter_run(const float* d_in, float* d_out, int W, int V, int nviews, jh_stream_t stream)
{
    return guard([&] {
        require(d_in != nullptr && d_out != nullptr && d_in != d_out, "jh_zfilter_run: bad buffers");
        require(W >= 1 && V >= 1 && nviews >= 0, "jh_zfilter_run: bad frame size");
        const size_t total = (size_t)W * V * nviews;
        if (total == 0) return;
        zfilter_kernel<<<grid_for(total, 256), 256, 0, as_stream(stream)>>>(d_in, d_out, W, V, total);
        check_launch("zfilter_kernel");
    });
}

jh_status jh_ctnum_run(float* d_inout, size_t n, float norm0, float norm1, float muWater, jh_stream_t stream)
{
    return guard([&] {
        require(d_inout != nullptr || n == 0, "jh_ctnum_run: buffer is NULL");
        require(muWater != 0.0f, "jh_ctnum_run: muWater is 0");
        if (n == 0) return;
        const float scale = norm0 * 1000.f / muWater;  // in float, as CTNumCore.cpp:24
        ctnum_kernel<<<grid_for(n, 256), 256, 0, as_stream(stream)>>>(d_inout, n, scale, norm1);
        check_launch("ctnum_kernel");
    });
}

jh_status jh_bp_autotune(jh_bp_plan* p, const void* d_q, int dtype, int nviews, float* d_vol, int reps,
                         jh_stream_t stream, float* h_ms)
{
    return guard([&] {
        require(p != nullptr && d_q != nullptr && d_vol != nullptr, "jh_bp_autotune: bad arguments");
        require(dtype == JH_F32 || dtype == JH_BF16 || dtype == JH_F16, "jh_bp_autotune: unknown dtype");
        require(nviews >= 1 && nviews <= p->N && reps >= 1, "jh_bp_autotune: bad view count or repetitions");
        DeviceScope ds(p->device);
        hipStream_t s = as_stream(stream);
        hipEvent_t e0, e1;
        JH_HIP(hipEventCreate(&e0));
        JH_HIP(hipEventCreate(&e1));
        const int zcs[3] = {8, 16, 32}, tys[3] = {1, 2, 4};
        float best = 1e30f;
        int bz = p->zc, bt = p->ty, bf = p->fetch;
        int n = 0;
        try {
            for (int f = 0; f < 2; ++f)
                for (int zi = 0; zi < 3; ++zi)
                    for (int ti = 0; ti < 3; ++ti, ++n) {
                        if (f == JH_FETCH_LDS && p->mode != 0) {  // staged fetch: geometry class 0 only
                            if (h_ms) h_ms[n] = -1.0f;
                            continue;
                        }
                        p->fetch = f;
                        p->zc = zcs[zi];
                        p->ty = tys[ti];
                        if (f == JH_FETCH_LDS && stage_strip(p, dtype) && p->weight != JH_WEIGHT_FDK &&
                            !strip_table(p, s)) {
                            // the strip footprint exceeds the LDS at this tile: the launch
                            // would gather, which the gather entry already times
                            if (h_ms) h_ms[n] = -1.0f;
                            continue;
                        }
                        const int out = p->weight == JH_WEIGHT_FDK ? kOutFdk : kOutNormalised;
                        launch_bp(p, d_q, dtype, 0, nviews, d_vol, out, s);  // warm-up
                        JH_HIP(hipEventRecord(e0, s));
                        for (int r = 0; r < reps; ++r) launch_bp(p, d_q, dtype, 0, nviews, d_vol, out, s);
                        JH_HIP(hipEventRecord(e1, s));
                        JH_HIP(hipEventSynchronize(e1));
                        float ms = 0.0f;
                        JH_HIP(hipEventElapsedTime(&ms, e0, e1));
                        ms /= (float)reps;
                        if (h_ms) h_ms[n] = ms;
                        if (ms < best) { best = ms; bz = p->zc; bt = p->ty; bf = p->fetch; }
                    }
        } catch (...) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            p->zc = bz; p->ty = bt; p->fetch = bf;
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        p->zc = bz;
        p->ty = bt;
        p->fetch = bf;
        // the LDS-staged strips' table for the chosen shape, built now (a
        // launch under graph capture cannot build it)
        if (p->fetch == JH_FETCH_LDS && p->mode == 0 && stage_strip(p, dtype)) (void)strip_table(p, s);
    });
}

jh_status jh_bp_get_tile(jh_bp_plan* p, int* zc, int* ty)
{
    return guard([&] {
        require(p != nullptr && zc != nullptr && ty != nullptr, "jh_bp_get_tile: bad arguments");
        *zc = p->zc;
        *ty = p->ty;
    });
}

jh_status jh_bp_geometry_mode(jh_bp_plan* p, int* mode)
{
    return guard([&] {
        require(p != nullptr && mode != nullptr, "jh_bp_geometry_mode: bad arguments");
        *mode = p->mode;
    });
}

jh_status jh_bp_destroy(jh_bp_plan* p)
{
    return guard([&] {
        if (!p) return;
        DeviceScope ds(p->device);
        delete p;
    });
}

}  // extern "C"
