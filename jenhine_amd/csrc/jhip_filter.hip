// Row ramp filter with pre-filter weight, as a Toeplitz GEMM on gfx950.
//
// The reference filters each detector row by FFT (FilterAgent/FilterCore.cpp:
// 189-222): buf = [x*pw, 0 ..., edge*ramp], y = irFFT(H * rFFT(buf)) / L,
// out = y[0..W).  H is real and even, so that is a circular convolution with
// the real even kernel h = irFFT(H) / L, and for n < W
//   out[n] = sum_{m<W} h[|n-m|] (x*pw)[m] + edge * r[n],
//   r[n]   = sum_{i<R} (i/R) h[(n + R - i) mod L],   edge = (x0+x1+x2)/3,
// with R = (L-W)/2.  Writing edge*r[n] = x0*r[n]/3 + x1*r[n]/3 + x2*r[n]/3
// makes the whole filter one GEMM per frame:
//   OUT (rows x W) = A (rows x K) * B (K x W),
//   A = [x*pw | x0 x1 x2 | 0-pad],  B = [T ; r/3 ; r/3 ; r/3 ; 0],
//   T[m][n] = h[|n-m|],  K = W + 3 rounded up to the K tile.
// A is formed on the fly from the frame and the pre-weight map; B is built
// once per plan (FP64 on the host, rounded once).
//
// Arithmetic forms:
//  * JH_FILTER_FP32: FP32 on MFMA (v_mfma_f32_16x16x4_f32);
//  * JH_FILTER_MFMA_BF16X3: MFMA 16x16x32 BF16 with hi/lo split operands,
//    hi*hi + hi*lo + lo*hi products accumulated in FP32;
//  * JH_FILTER_MFMA_BF16X6: hi/mid/lo split, the 6 products of total order
//    <= 2 (hh, hm, mh, hl, lh, mm);
//  * JH_FILTER_FP32_VALU: FP32 FMA on the vector ALUs (LDS-tiled SIMT GEMM),
//    kept as a cross-check of the MFMA forms.
//
// Pre-filter weight maps (FilterAgent/PreFilterWeightGenerator.cpp:10-124)
// are built on the device in FP64, one per refresh block of 9 views
// (FilterAgent.cpp:101-126).
#include <cmath>
#include <cstring>

#include "jhip_bp_plan.h"
#include "jhip_filter_plan.h"
#include "jhip_internal.h"


namespace jh {

// ---------------------------------------------------------------------------
// Pre-filter weight maps, FP64.
// ---------------------------------------------------------------------------

// Inverse TM of every refresh block's source view, one thread per block.
__global__ void pw_inverse_kernel(const double* __restrict__ ptm, const int* __restrict__ src_view,
                                  int nb, double dso, double* __restrict__ inv)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    tm_inverse_fp64(ptm + 12 * (size_t)src_view[b], dso, inv + 16 * (size_t)b);
}

// Distance between the back-projections of detector point (u, v) at the two
// pseudo-depths w = 0 and w = 1.
__device__ inline double pw_distance(const double* T, double u, double v)
{
    const double t = u * T[12] + v * T[13];
    const double n0 = t + T[15];
    const double n1 = t + (T[14] + T[15]);
    double s = 0.0;
    for (int r = 0; r < 3; ++r) {
        const double q = u * T[4 * r] + v * T[4 * r + 1];
        const double a0 = (q + T[4 * r + 3]) / n0;
        const double a1 = (q + (T[4 * r + 2] + T[4 * r + 3])) / n1;
        const double d = a0 - a1;
        s += d * d;
    }
    return sqrt(s);
}

// Minimum of the distance over the full L x V map of each block.
__global__ void pw_min_kernel(const double* __restrict__ inv, int L, int V, int W, int onRight,
                              unsigned long long* __restrict__ minbits)
{
    const int b = blockIdx.y;
    const double* T = inv + 16 * (size_t)b;
    const size_t n = (size_t)L * V;
    double m = INFINITY;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
        const int v = (int)(t / L), u = (int)(t - (size_t)v * L);
        const double uu = onRight ? (double)u - (double)W / 2.0 : (double)u;
        const double d = pw_distance(T, uu, (double)v);
        m = fmin(m, d);
    }
    // wave reduction then one atomic per wave (positive doubles order as integers)
    for (int off = 32; off > 0; off >>= 1) m = fmin(m, __shfl_down(m, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMin(&minbits[b], (unsigned long long)__double_as_longlong(m));
}

__global__ void pw_write_kernel(const double* __restrict__ inv, const unsigned long long* __restrict__ minbits,
                                int W, int V, int onRight, float* __restrict__ out)
{
    const int b = blockIdx.y;
    const double* T = inv + 16 * (size_t)b;
    const double mn = __longlong_as_double((long long)minbits[b]);
    const size_t n = (size_t)W * V;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
        const int v = (int)(t / W), u = (int)(t - (size_t)v * W);
        const double uu = onRight ? (double)u - (double)W / 2.0 : (double)u;
        out[(size_t)b * n + t] = (float)(mn / pw_distance(T, uu, (double)v));
    }
}

// ---------------------------------------------------------------------------
// FP32 Toeplitz GEMM (vector ALUs).  Tile 64 rows x 64 columns x 16 k,
// 256 threads, 4 x 4 outputs per thread.
// ---------------------------------------------------------------------------

struct FiltArgs {
    const float* x;      // rows x W (frames of V rows)
    const float* pw;     // nblocks x V x W
    const int* blk;      // block of every view of the scan
    const float* B;      // Kp x Wout
    const void* Bhl;     // MFMA operand
    float* out;          // rows x Wout
    int W, V, rows, Kp, view0;
    int Wout;            // output width
    int edge0;           // first of the three edge columns (0, or W - 3 with yMirroring)
    // fused staging (out unused): BF16/FP16 column strips of the launch's
    // views, q = filtered * uW, rows padded to W + 1 pixels
    void* q;             // NULL: write the filtered FP32 rows to out
    const float* uw;     // uW of the launch's first view onwards, Wout per view
    int qtype;           // JH_BF16 or JH_F16
    float qs;            // staged values are q * qs (qs = 2^-f16exp for FP16, exact; else 1)
    int tiled;           // x is the pre-weighted A operand in the MFMA tile order (pre_operand_kernel)
    int bstride;         // float4s per 16-row block of that operand
    float bscale;        // split-FP16 form: factor of the packed filter matrix (a power of two)
    // generated split-FP16 operand (see filter_mfma_kernel)
    const unsigned short* Bgen;
    const unsigned short* R3;
    int gen_stride, gen_window, gen_nT;
};

// Element (r, m) of the A operand.
__device__ inline float a_elem(const FiltArgs& a, int r, int m, const float* xr, const float* pwr)
{
    if (m < a.W) return xr[m] * pwr[m];
    if (m < a.W + 3) return xr[m - a.W + a.edge0];
    return 0.0f;
}

__global__ void __launch_bounds__(256) filter_fp32_kernel(FiltArgs a)
{
    __shared__ float As[16][64 + 4];
    __shared__ float Bs[16][64];
    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    const int row0 = blockIdx.y * 64, col0 = blockIdx.x * 64;
    float c[4][4] = {};
    // each thread stages 4 A elements (rows lr, k lk..) and 4 B elements
    const int lr = tid >> 2;         // 0..63
    const int lk = (tid & 3) * 4;    // 0,4,8,12
    const int r = min(row0 + lr, a.rows - 1);
    const int view = r / a.V, vrow = r - view * a.V;
    const float* xr = a.x + (size_t)r * a.W;
    const float* pwr = a.pw + ((size_t)a.blk[a.view0 + view] * a.V + vrow) * a.W;
    const int bk = tid >> 4, bn = (tid & 15) * 4;
    for (int k0 = 0; k0 < a.Kp; k0 += 16) {
#pragma unroll
        for (int q = 0; q < 4; ++q) As[lk + q][lr] = a_elem(a, r, k0 + lk + q, xr, pwr);
        {
            const float* bp = a.B + (size_t)(k0 + bk) * a.Wout + col0 + bn;
#pragma unroll
            for (int q = 0; q < 4; ++q) Bs[bk][bn + q] = (col0 + bn + q < a.Wout) ? bp[q] : 0.0f;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
            float av[4], bv[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) { av[q] = As[kk][ty * 4 + q]; bv[q] = Bs[kk][tx * 4 + q]; }
#pragma unroll
            for (int p = 0; p < 4; ++p)
#pragma unroll
                for (int q = 0; q < 4; ++q) c[p][q] = fmaf(av[p], bv[q], c[p][q]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const int rr = row0 + ty * 4 + p;
        if (rr >= a.rows) continue;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int cc = col0 + tx * 4 + q;
            if (cc < a.Wout) a.out[(size_t)rr * a.Wout + cc] = c[p][q];
        }
    }
}

// ---------------------------------------------------------------------------
// Preprocessing written straight into the MFMA kernel's A operand.
//
// Element (r, k) of the operand, r a row of the launch (view r / V): the
// preprocessed pixel times the pre-weight for k < W, the unweighted pixel
// edge0 + k - W for W <= k < W + 3 (the truncation ramp's edge columns), 0
// beyond -- the values the MFMA kernel forms from the preprocessed frames and
// the pre-weight maps (load_a8), computed by the same operations.  Layout: the
// floats (r, k) of 16-row block r16 and k step s = k / 32 that lane
// l = 16 g + (r % 16), g = (k % 32) / 8, supplies, in two 16-byte halves h =
// (k % 8) / 4: float4 index r16 bstride + (2 s + h) 64 + l, bstride = 128
// nsteps.  The MFMA kernel then reads 1 KB contiguous per wave load and no
// pre-weight.  Rows beyond the launch are 0.
//
// Split-FP16 form: the filter scales every row r by 2^e_r, chosen from the
// row's largest magnitude, before splitting it into FP16 parts; each block
// then also holds the row maxima of its 128-column k tiles, tile-major (16
// floats per tile: ntk = ceil(Kp / 128) tiles, 4 ntk more float4s, so
// bstride = 128 nsteps + 4 ntk).
// ---------------------------------------------------------------------------

struct PreOpArgs {
    const uint16_t* raw;  // nviews frames of inW x inH
    const uint16_t* dark;
    const float* gain;    // W x V (NULL: none)
    const float* pw;      // nblocks x V x W pre-weight maps
    const int* blk;       // block of every view of the scan
    float4* out;
    int inW, inH, bUp, bLeft;
    int W, V, rows, nsteps, view0, edge0;
    int bstride;          // float4s per 16-row block of the operand
    int rowmax;           // split-FP16 form: write the row maxima of each k tile
    float bf, a0, a1, a2;
    int doBH;
    int vec;              // rows of counts, dark, gain and pre-weights allow 16-byte loads
};

// Operand elements (r, kb .. kb+7) of a launch row r < rows (0 beyond K).
__device__ inline void pre_operand_elems(const PreOpArgs& a, int r, int kb, float (&e)[8])
{
#pragma unroll
    for (int j = 0; j < 8; ++j) e[j] = 0.0f;
    if (kb < a.W + 3) {
        const int view = r / a.V, vrow = r - view * a.V;
        const uint16_t* raw = a.raw + (size_t)view * a.inW * a.inH;
        const size_t rowin = (size_t)(vrow + a.bUp) * a.inW + a.bLeft;
        const size_t rowout = (size_t)vrow * a.W;
        const float* pw = a.pw + ((size_t)a.blk[a.view0 + view] * a.V + vrow) * a.W;
        if (kb + 8 <= a.W && a.vec) {
            // 16-byte loads: 8 counts, 8 dark values, 8 gains, 8 pre-weights
            const uint4 rv = *reinterpret_cast<const uint4*>(raw + rowin + kb);
            uint4 dv = make_uint4(0u, 0u, 0u, 0u);
            if (a.dark) dv = *reinterpret_cast<const uint4*>(a.dark + rowin + kb);
            float g[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
            if (a.gain) {
                const float4 g0 = *reinterpret_cast<const float4*>(a.gain + rowout + kb);
                const float4 g1 = *reinterpret_cast<const float4*>(a.gain + rowout + kb + 4);
                g[0] = g0.x; g[1] = g0.y; g[2] = g0.z; g[3] = g0.w; g[4] = g1.x; g[5] = g1.y; g[6] = g1.z; g[7] = g1.w;
            }
            const float4 w0 = *reinterpret_cast<const float4*>(pw + kb);
            const float4 w1 = *reinterpret_cast<const float4*>(pw + kb + 4);
            const float w[8] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w};
            const uint32_t rw[4] = {rv.x, rv.y, rv.z, rv.w}, dw[4] = {dv.x, dv.y, dv.z, dv.w};
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const uint16_t rc = (uint16_t)(rw[j >> 1] >> (16 * (j & 1)));
                const uint16_t dc = (uint16_t)(dw[j >> 1] >> (16 * (j & 1)));
                // the steps of load_corrected and pre_log, on the loaded values
                int v = (int)rc;
                if (a.dark) v -= (int)dc;
                float f = (float)v;
                f = f < 1.0f ? 1.0f : f;
                float d = a.bf;
                if (a.gain) d = d * g[j];
                e[j] = pre_log_div(f, d, a.a0, a.a1, a.a2, a.doBH) * w[j];
            }
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int k = kb + j;
                const int c = k < a.W ? k : a.edge0 + k - a.W;
                if (k < a.W + 3) {
                    const float p = pre_log(load_corrected(raw, a.dark, rowin + c), a.bf, a.gain, rowout + c, a.a0,
                                            a.a1, a.a2, a.doBH);
                    e[j] = k < a.W ? p * pw[k] : p;
                }
            }
        }
    }
}

constexpr int kOpPitch = 132;  // LDS row pitch (floats) of the operand tile

// Workgroup: 16 rows x 128 k (four k steps) of the operand.  Loads: thread t
// takes row t / 16 and the 8 columns 8 (t % 16) ..: a wave reads four rows of
// 256 contiguous bytes of raw counts per instruction.  The values go through
// LDS and leave in the operand order, 1 KB contiguous per wave store.
__global__ void __launch_bounds__(256) pre_operand_kernel(PreOpArgs a)
{
    __shared__ float tile[16 * kOpPitch];
    const int t = threadIdx.x;
    const int i = t >> 4, cg = t & 15;
    const int r = blockIdx.y * 16 + i;
    const int kb = blockIdx.x * 128 + 8 * cg;
    float e[8];
    if (r < a.rows) {
        pre_operand_elems(a, r, kb, e);
    } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) e[j] = 0.0f;
    }
    float* dstt = tile + i * kOpPitch + 8 * cg;
#pragma unroll
    for (int j = 0; j < 8; ++j) dstt[j] = e[j];
    float4* blk = a.out + (size_t)blockIdx.y * a.bstride;
    if (a.rowmax) {
        // the row's largest magnitude in this k tile (the 16 lanes of row i)
        float m = 0.0f;
#pragma unroll
        for (int j = 0; j < 8; ++j) m = fmaxf(m, fabsf(e[j]));
#pragma unroll
        for (int x = 1; x < 16; x <<= 1) m = fmaxf(m, __shfl_xor(m, x));
        if (cg == 0) reinterpret_cast<float*>(blk + (size_t)a.nsteps * 128)[blockIdx.x * 16 + i] = m;
    }
    __syncthreads();
    // operand order: k step s = 4 blockIdx.x + t / 64, lane l = t % 64, halves h
    const int sl = t >> 6, l = t & 63;
    const int s = blockIdx.x * 4 + sl;
    if (s >= a.nsteps) return;
    const float* src = tile + (l & 15) * kOpPitch + 32 * sl + 8 * (l >> 4);
    float4* dst = blk + (size_t)s * 128 + l;
    dst[0] = make_float4(src[0], src[1], src[2], src[3]);
    dst[64] = make_float4(src[4], src[5], src[6], src[7]);
}

// ---------------------------------------------------------------------------
// Split-FP16 operand rows: every row r of the A operand is scaled by 2^e_r,
// e_r chosen from the row's largest magnitude m so that m 2^e_r lies in
// [2^14, 2^15) (FP16's largest finite value is 65504), and each scaled value
// v is split into hi = RN16(v) and lo = RN16(v - hi); v - hi is exact in
// FP32, so hi + lo keep 22 significant bits.  The results are multiplied by
// 2^-e_r and by the filter matrix's own power-of-two factor, exactly.
// ---------------------------------------------------------------------------

// Exponent e_r of a row's scale 2^e_r from its largest magnitude m (m >= 0,
// not NaN): m 2^e_r in [2^14, 2^15); 100 for m < 2^-85 (and 0); -114 for an
// infinite m (the row's results are then infinite or NaN in every form).
__device__ inline int row_scale_exp(float m)
{
    const int E = (int)((__float_as_uint(m) >> 23) & 0xffu);  // m = f 2^(E - 126), f in [0.5, 1)
    return E == 0 ? 100 : min(141 - E, 100);
}

// 2^e for -126 <= e <= 127.
__device__ inline float exp2i(int e) { return __uint_as_float((uint32_t)(e + 127) << 23); }

// Eight operand values of one row times the row's scale s, as FP16 hi and lo
// parts, packed two per word in k order.  The scaled value v = x s is exact
// (s a power of two); hi = RN16(v), and v - hi is exact in FP32.
__device__ inline void split_f16(const float (&e)[8], float s, uint32_t (&hi)[4], uint32_t (&lo)[4])
{
#pragma clang fp contract(off)
#pragma unroll
    for (int h = 0; h < 4; ++h) {
        const float v0 = e[2 * h] * s, v1 = e[2 * h + 1] * s;
        const _Float16 h0 = (_Float16)v0, h1 = (_Float16)v1;
        const _Float16 l0 = (_Float16)(v0 - (float)h0), l1 = (_Float16)(v1 - (float)h1);
        hi[h] = (uint32_t)__builtin_bit_cast(unsigned short, h0) | ((uint32_t)__builtin_bit_cast(unsigned short, h1) << 16);
        lo[h] = (uint32_t)__builtin_bit_cast(unsigned short, l0) | ((uint32_t)__builtin_bit_cast(unsigned short, l1) << 16);
    }
}

// ---------------------------------------------------------------------------
// Toeplitz GEMM on MFMA.
//
// Workgroup: 256 threads = 4 waves, tile 128 rows x 128 columns; wave w owns
// rows 32w..32w+31 (two 16-row blocks) and all 8 column blocks of 16.  K step
// 32.  In every form lane l supplies, for its row (l % 16) and column
// (l % 16), the 8 consecutive k values 32 s + 8 (l / 16) + 0..7:
//  * BF16 forms use v_mfma_f32_16x16x32_bf16, whose operand layout is exactly
//    that; the operands are split into P = 2 (hi, lo) or P = 3 (hi, mid, lo)
//    BF16 parts and every product of parts whose orders sum to <= P-1 is
//    accumulated in FP32 (3 or 6 products), smallest first;
//  * the split-FP16 form uses v_mfma_f32_16x16x32_f16 on hi / lo FP16 parts
//    of the row-scaled operand and of the scaled filter matrix (3 products,
//    smallest first); the exact power-of-two factors are applied to the FP32
//    results;
//  * the FP32 form issues 8 v_mfma_f32_16x16x4_f32, sub-step q taking
//    k = 32 s + 8 (l / 16) + q (a permutation of k applied to both operands).
// D layout (16x16 family): lane l holds D[4 (l / 16) + i][l % 16], i = 0..3.
// The B operand is packed on the host per (k step, column block, lane) --
// part-major per (k step, column block) in the split-FP16 form -- and copied
// into LDS once per workgroup and k step, one step ahead into the other of
// two buffers (global_load_lds, no registers).
// ---------------------------------------------------------------------------

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
// Native 16-byte vector for the B prefetch: HIP's uint4 is a struct whose
// copies go through memory and would keep the prefetch in scratch.
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));

constexpr int kModeF32 = 0, kModeX3 = 2, kModeX6 = 3, kModeH3 = 4;

template <int MODE>
struct FiltMode;
template <>
struct FiltMode<kModeF32> { static constexpr int parts = 1; static constexpr int lane_bytes = 32; };
template <>
struct FiltMode<kModeX3> { static constexpr int parts = 2; static constexpr int lane_bytes = 32; };
template <>
struct FiltMode<kModeX6> { static constexpr int parts = 3; static constexpr int lane_bytes = 48; };
template <>
struct FiltMode<kModeH3> { static constexpr int parts = 2; static constexpr int lane_bytes = 32; };

__device__ inline float bf16_round(float f)
{
    uint32_t u = __float_as_uint(f);
    u = (u + 0x7fffu + ((u >> 16) & 1u)) & 0xffff0000u;
    return __uint_as_float(u);
}

__device__ inline __bf16 to_bf16(float rounded) { return __builtin_bit_cast(__bf16, (unsigned short)(__float_as_uint(rounded) >> 16)); }

// A operand element (r, m): pre-weighted input, or one of the three edge
// columns (unweighted x0, x1, x2), or padding.
__device__ inline void load_a8(const FiltArgs& a, const float* xr, const float* pwr, int kb, float e[8])
{
    if (kb + 8 <= a.W) {
        const float4 x0 = *reinterpret_cast<const float4*>(xr + kb);
        const float4 x1 = *reinterpret_cast<const float4*>(xr + kb + 4);
        const float4 w0 = *reinterpret_cast<const float4*>(pwr + kb);
        const float4 w1 = *reinterpret_cast<const float4*>(pwr + kb + 4);
        e[0] = x0.x * w0.x; e[1] = x0.y * w0.y; e[2] = x0.z * w0.z; e[3] = x0.w * w0.w;
        e[4] = x1.x * w1.x; e[5] = x1.y * w1.y; e[6] = x1.z * w1.z; e[7] = x1.w * w1.w;
    } else {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int m = kb + q;
            e[q] = m < a.W ? xr[m] * pwr[m] : (m < a.W + 3 ? xr[m - a.W + a.edge0] : 0.0f);
        }
    }
}

// 16-bit pattern of x rounded to nearest-even: BF16 (as the backprojector's
// staging) or FP16 (the empty asm keeps the product that made x from being
// folded into a multiply that rounds straight to FP16).
__device__ inline uint32_t round16(float x, int qtype)
{
    if (qtype == JH_BF16) {
        uint32_t u = __float_as_uint(x);
        if ((u & 0x7fffffffu) > 0x7f800000u) return ((u >> 16) | 0x40u) & 0xffffu;  // quiet NaN
        u += 0x7fffu + ((u >> 16) & 1u);
        return u >> 16;
    }
    asm volatile("" : "+v"(x));
    return __builtin_bit_cast(unsigned short, (_Float16)x);
}

constexpr int kStagePitch = 136;  // LDS row pitch (16-bit values) of the staged tile: 4 rows apart = 16 banks
// LDS of the generated-operand launch at most: three workgroups per CU
constexpr size_t kGenMaxLds = 53248;

// Fused staging epilogue of filter_mfma_kernel: the workgroup's 128 x 128
// tile of filtered values, times the backprojector's uW and rounded to 16
// bits, goes through LDS and is written as column strips -- pixel (v, u)
// holds rows v .. v+3 of column u, 0 beyond the view's last row, rows padded
// to W + 1 pixels (the padding column is 0) -- bit-identical to
// jh_filter_run followed by jh_bp_stage.  A strip whose rows reach into the
// next row tile is written in 16-bit slots, each by the workgroup that owns
// the slot's row; a slot whose row lies beyond the view or the launch is
// written as 0 by the strip's owner.
__device__ void stage_strips_epilogue(const FiltArgs& a, const f32x4_t (&acc)[2][8], unsigned char* lds, int lane,
                                      int wave, int ctile, int rtile)
{
    const int T0 = rtile * 128, C0 = ctile * 128;
    const int vt = T0 / a.V, vrow0 = T0 - vt * a.V;  // view and row (in its view) of the tile's first row
    uint16_t* t16 = reinterpret_cast<uint16_t*>(lds);
    __syncthreads();  // every wave is done with the B tile
    // the lane's 8 columns' uW, reloaded when its row enters another view
    int cview = -1;
    float uwc[8];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int rl = wave * 32 + rb * 16 + 4 * (lane >> 4) + q, row = T0 + rl;
            int view = vt, vrow = vrow0 + rl;
            while (vrow >= a.V) { vrow -= a.V; ++view; }
            const bool live = row < a.rows;
            if (live && view != cview) {
                cview = view;
                const float* uwv = a.uw + (size_t)view * a.Wout;
#pragma unroll
                for (int c = 0; c < 8; ++c) {
                    const int col = C0 + c * 16 + (lane & 15);
                    // uW * qs is exact (qs a power of two), and so is
                    // acc * (uW qs) = (acc uW) qs in the normal range
                    uwc[c] = col < a.Wout ? uwv[col] * a.qs : 0.0f;
                }
            }
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const int cl = c * 16 + (lane & 15);
                uint32_t h = 0;
                if (live && C0 + cl < a.Wout) h = round16(acc[rb][c][q] * uwc[c], a.qtype);
                t16[rl * kStagePitch + cl] = (uint16_t)h;
            }
        }
    __syncthreads();
    const int cl = threadIdx.x & 127, half = threadIdx.x >> 7;
    const int col = C0 + cl;
    if (col >= a.Wout) return;
    const bool padcol = col == a.Wout - 1;  // also writes the zero padding column
    const size_t rowpx = (size_t)a.Wout + 1;
    // strip pixel (v, col) of launch row v (frames are contiguous: v = view V + row)
    char* const q = static_cast<char*>(a.q) + (size_t)col * 8;
    int vl = half * 64;
    int vend = (vt + 1) * a.V;  // first row of the view after row T0 + vl
    {
        int vrow = vrow0 + vl;
        while (vrow >= a.V) { vrow -= a.V; vend += a.V; }
    }
    // window of the staged values of rows vl .. vl+3 (0 where not this tile's)
    uint32_t w1 = t16[vl * kStagePitch + cl], w2 = t16[(vl + 1) * kStagePitch + cl];
    uint32_t w3 = t16[(vl + 2) * kStagePitch + cl];
    // strips whose four rows lie in this tile, in the view of the run's first
    // row and in the launch: whole 8-byte pixels, no per-slot decisions
    const int R0 = T0 + vl;
    const int nfast = max(0, min(min(64, 125 - vl), min(vend - 3 - R0, a.rows - 3 - R0)));
    int k = 0;
    for (; k < nfast; ++k, ++vl) {
        const uint32_t w4 = t16[(vl + 3) * kStagePitch + cl];
        char* px = q + (size_t)(T0 + vl) * rowpx * 8;
        *reinterpret_cast<uint2*>(px) = make_uint2(w1 | (w2 << 16), w3 | (w4 << 16));
        if (padcol) *reinterpret_cast<uint2*>(px + 8) = make_uint2(0u, 0u);
        w1 = w2; w2 = w3; w3 = w4;
    }
    for (; k < 64; ++k, ++vl) {
        const int R = T0 + vl;
        if (R >= a.rows) break;
        if (R >= vend) vend += a.V;
        const uint32_t w4 = vl + 3 < 128 ? t16[(vl + 3) * kStagePitch + cl] : 0u;
        const uint32_t hv[4] = {w1, w2, w3, w4};
        bool all = true;
        uint32_t mine = 0;  // bit s: this workgroup writes slot s
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int r = R + s;
            const bool beyond = r >= vend || r >= a.rows;  // written as 0 here
            const bool here = vl + s < 128;                 // else the next tile's owner writes it
            if (beyond || here) mine |= 1u << s;
            else all = false;
        }
        uint32_t v4[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) v4[s] = (R + s >= vend || R + s >= a.rows) ? 0u : hv[s];
        char* px = q + (size_t)R * rowpx * 8;
        if (all) {
            *reinterpret_cast<uint2*>(px) = make_uint2(v4[0] | (v4[1] << 16), v4[2] | (v4[3] << 16));
        } else {
#pragma unroll
            for (int s = 0; s < 4; ++s)
                if (mine & (1u << s)) *reinterpret_cast<uint16_t*>(px + 2 * s) = (uint16_t)v4[s];
        }
        if (padcol) *reinterpret_cast<uint2*>(px + 8) = make_uint2(0u, 0u);
        w1 = w2; w2 = w3; w3 = w4;
    }
    // this tile's first rows in the last strips of the previous tile (same view only)
    if (half == 0 && T0 > 0) {
#pragma unroll
        for (int rl = 0; rl < 3; ++rl) {
            const int r = T0 + rl;
            if (r >= a.rows) break;
            int vrow = vrow0 + rl;
            while (vrow >= a.V) vrow -= a.V;
#pragma unroll
            for (int s = rl + 1; s < 4; ++s)
                if (vrow >= s)  // row r - s is in r's view
                    *reinterpret_cast<uint16_t*>(q + (size_t)(r - s) * rowpx * 8 + 2 * s) = t16[rl * kStagePitch + cl];
        }
    }
}

// 16 bytes per lane from global memory straight into LDS (global_load_lds):
// lane i's bytes land at wave_base + 16 i; wave_base is the wave's LDS base.
__device__ __forceinline__ void copy16_to_lds(const void* src, unsigned char* wave_base)
{
    __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void*)wave_base, 16, 0, 0);
}

// s_waitcnt vmcnt(0) (expcnt and lgkmcnt left at their maxima): every copy
// into LDS has landed.
__device__ __forceinline__ void wait_lds_copies() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// TILED: A is the pre-weighted operand (a.tiled); the split-BF16 forms then
// load the next k step's A while the current one computes.
//
// GEN (split-FP16 form, W a multiple of 32, Wout of 128): the B operand is
// not copied from the packed matrix step by step but read from the Toeplitz
// diagonals kept in LDS.  B[k][n] = d(n + off - k) for k < W, so lane l's 8
// values of k step s and column block c are d at 8 consecutive descending
// arguments, i.e. 8 consecutive entries of the reversed diagonal, starting at
// i0 = 32 s + 127 - 16 c - (l % 16) + 8 (l / 16) within the workgroup's
// window.  i0 mod 8 = (7 - l) mod 8 is fixed per lane, so of 8 copies of the
// window, shifted by 0..7 entries, copy (7 - l) mod 8 serves lane l with
// one aligned 16-byte read per part (the copies padded 8 banks apart).  The copies hold the same FP16 hi / lo
// values as the packed matrix, so the product is the packed form's to the
// bit.  The last k step (rows W .. W + 31) holds the edge rows r / 3 at
// k = W .. W + 2 and zeros.  Each workgroup reads (W + 128) x 32 bytes of
// copies once instead of 16 KB per k step, and the k loop has no barrier.
template <int MODE, bool TILED, bool GEN = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) filter_mfma_kernel(FiltArgs a)
{
    static_assert(!GEN || MODE == kModeH3, "the generated operand is the split-FP16 form's");
    typedef FiltMode<MODE> M;
    constexpr int P = M::parts;
    constexpr int LB = M::lane_bytes;                 // B bytes per lane per column block
    constexpr int TILE_BYTES = 8 * 64 * LB;           // B bytes per k step per workgroup
    constexpr int CHUNKS = TILE_BYTES / 16 / 256;     // 16-byte chunks per thread
    // the B tiles of two k steps: step s reads buffer s & 1 while the B tile
    // of step s + 1 is copied from memory straight into the other
    // (global_load_lds, no registers); after the last step, the staged tile
    // of the fused epilogue (128 rows at kStagePitch 16-bit values)
    constexpr int LDS_BYTES = 2 * TILE_BYTES > 128 * kStagePitch * 2 ? 2 * TILE_BYTES : 128 * kStagePitch * 2;
    __shared__ __attribute__((aligned(16))) unsigned char lds_static[GEN ? 16 : LDS_BYTES];
    // GEN: the copies of the diagonals, then the edge rows, then (after the
    // k loop) the staged tile, sized by the launch
    extern __shared__ __attribute__((aligned(16))) unsigned char filt_lds_dyn[];
    unsigned char* const lds = GEN ? filt_lds_dyn : lds_static;
    constexpr bool H16 = MODE == kModeH3;
    __shared__ float rowfac[H16 ? 128 : 1];  // split-FP16: the result factor of each row

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int ncb8 = (a.Wout + 127) / 128 * 8;         // packed column blocks (multiple of 8)
    // tile of this workgroup: the launch is 1-D and the hardware hands
    // workgroup L to XCD L % 8, so XCD x takes a contiguous run of tiles in
    // column-fastest order -- the column tiles of a row tile share one L2
    // and read its operand rows from memory once
    int ctile, rtile;
    {
        const int ntiles = gridDim.x, q = ntiles / 8, r = ntiles % 8;
        const int L = blockIdx.x, x = L % 8, i = L / 8;
        const int t = (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + i;
        const int nct = (a.Wout + 127) / 128;
        rtile = t / nct;
        ctile = t - rtile * nct;
    }
    const int cb0 = ctile * 8;
    const int rowTile = rtile * 128 + wave * 32;
    const int kq = 8 * (lane >> 4);
    const int nsteps = a.Kp / 32;

    const float* xr[2];
    const float* pwr[2];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) {
        const int r = min(rowTile + rb * 16 + (lane & 15), a.rows - 1);
        const int view = r / a.V, vrow = r - view * a.V;
        xr[rb] = a.x + (size_t)r * a.W;
        pwr[rb] = a.pw + ((size_t)a.blk[a.view0 + view] * a.V + vrow) * a.W;
    }

    f32x4_t acc[2][8];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int c = 0; c < 8; ++c) acc[rb][c] = (f32x4_t){0.f, 0.f, 0.f, 0.f};

    // split-FP16: each row's scale from its largest magnitude -- from the
    // operand's k-tile maxima, or over all of K from the FP32 rows (the four
    // lanes l, l ^ 16, l ^ 32, l ^ 48 of a row share its k groups)
    float rscale[2] = {1.0f, 1.0f};
    if constexpr (H16) {
#pragma unroll
        for (int rb = 0; rb < 2; ++rb) {
            float m = 0.0f;
            if constexpr (TILED) {
                const int rs = min(rowTile + rb * 16, (a.rows - 1) / 16 * 16);
                const float* tm = reinterpret_cast<const float*>(reinterpret_cast<const float4*>(a.x) +
                                                                 (size_t)(rs / 16) * a.bstride + nsteps * 128);
                const int ntk = (a.Kp + 127) / 128;
                for (int kt = lane >> 4; kt < ntk; kt += 4) m = fmaxf(m, tm[kt * 16 + (lane & 15)]);
            } else {
                for (int s = 0; s < nsteps; ++s) {
                    float e[8];
                    load_a8(a, xr[rb], pwr[rb], 32 * s + kq, e);
#pragma unroll
                    for (int q = 0; q < 8; ++q) m = fmaxf(m, fabsf(e[q]));
                }
            }
            m = fmaxf(m, __shfl_xor(m, 16));
            m = fmaxf(m, __shfl_xor(m, 32));
            const int se = row_scale_exp(m);
            rscale[rb] = exp2i(se);
            if (lane < 16) rowfac[wave * 32 + rb * 16 + lane] = exp2i(-se) * a.bscale;
        }
    }

    const u32x4_t* gB = reinterpret_cast<const u32x4_t*>(a.Bhl);
    auto gsrc = [&](int s) { return gB + ((size_t)s * ncb8 + cb0) * 64 * LB / 16; };
    // B tile of step s into LDS buffer buf: 16-byte chunk threadIdx.x + 256 c
    // of the tile; each wave's 64 chunks land contiguously at its LDS base
    auto copy_b = [&](int s, int buf) {
        const u32x4_t* src = gsrc(s) + threadIdx.x;
#pragma unroll
        for (int c = 0; c < CHUNKS; ++c)
            copy16_to_lds(src + c * 256, lds + buf * TILE_BYTES + (c * 256 + (threadIdx.x & ~63u)) * 16);
    };
    // LDS copies: part p (hi, lo) of shift sg at p * 8 gS + sg * gS halves.
    // gS = the window + 16 halves: the window (W + 128 halves) is a multiple
    // of 64 banks, and lanes 0..7 of a 16-byte read -- one per shift, at the
    // same offset -- would all start on one bank; 8 banks apart they do not
    // (3x fewer bank-conflict cycles, measured)
    const int gL = a.gen_window;                 // halves per copy a workgroup needs
    const int gS = gL + 16;                      // halves between copies
    const unsigned short* const gl = reinterpret_cast<const unsigned short*>(lds);
    const int gsig = (7 - lane) & 7;             // the copy serving this lane
    if constexpr (GEN) {
        // this column tile's window of the 16 copies (8 shifts x hi, lo), and
        // its 128 columns of the edge rows
        unsigned short* const glw = reinterpret_cast<unsigned short*>(lds);
        const int i0g = a.Wout - 128 * (ctile + 1);
        const int per = gL / 8;
        for (int idx = threadIdx.x; idx < 16 * per; idx += 256) {
            const int cp = idx / per, t8 = idx - cp * per;  // cp = 2 sg + p in the global layout
            *reinterpret_cast<u32x4_t*>(glw + (cp & 1) * 8 * gS + (cp >> 1) * gS + 8 * t8) =
                *reinterpret_cast<const u32x4_t*>(a.Bgen + (size_t)cp * a.gen_stride + i0g + 8 * t8);
        }
        for (int idx = threadIdx.x; idx < 256; idx += 256)
            glw[16 * gS + idx] = a.R3[(size_t)(idx >> 7) * a.Wout + 128 * ctile + (idx & 127)];
        __syncthreads();
    } else {
        copy_b(0, 0);
        wait_lds_copies();
        __syncthreads();
    }

    // the 8 A values lane supplies for row block rb at k step s: formed from
    // the preprocessed frame and the pre-weight map, or read from the
    // pre-weighted operand (two 16-byte loads; the 16-row block exists in the
    // operand iff it starts inside the launch)
    auto load_a = [&](int rb, int s, int kb, float (&e)[8]) {
        if (a.tiled) {
            const int rs = rtile * 128 + wave * 32 + rb * 16;
            float4 x0 = make_float4(0.f, 0.f, 0.f, 0.f), x1 = x0;
            if (rs < a.rows) {
                const float4* src = reinterpret_cast<const float4*>(a.x) + (size_t)(rs / 16) * a.bstride + s * 128 + lane;
                x0 = src[0];
                x1 = src[64];
            }
            e[0] = x0.x; e[1] = x0.y; e[2] = x0.z; e[3] = x0.w;
            e[4] = x1.x; e[5] = x1.y; e[6] = x1.z; e[7] = x1.w;
        } else {
            load_a8(a, xr[rb], pwr[rb], kb, e);
        }
    };

    // the operand's A of step s for row block rb; rows past the launch read
    // the last block (their results are never stored), so the loads carry
    // no condition and the wait for one step's A does not wait for the next
    auto load_tiled = [&](int rb, int s, float4 (&x)[2]) {
        const int rs = min(rtile * 128 + wave * 32 + rb * 16, (a.rows - 1) / 16 * 16);
        const float4* src = reinterpret_cast<const float4*>(a.x) + (size_t)(rs / 16) * a.bstride + s * 128 + lane;
        x[0] = src[0];
        x[1] = src[64];
    };
    float4 acur[2][2];
    if constexpr (TILED && MODE != kModeF32) {
#pragma unroll
        for (int rb = 0; rb < 2; ++rb) load_tiled(rb, 0, acur[rb]);
    }

    for (int s = 0; s < nsteps; ++s) {
        const unsigned char* bt = lds + (s & 1) * TILE_BYTES;
        const int kb = 32 * s + kq;
        if (MODE == kModeF32) {
            float ea[2][8];
#pragma unroll
            for (int rb = 0; rb < 2; ++rb) load_a(rb, s, kb, ea[rb]);
            // unconditional (the last step recopies its own tile into the idle
            // buffer), so that the waits for A do not also wait for the copy
            copy_b(min(s + 1, nsteps - 1), (s + 1) & 1);
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const float4* bp = reinterpret_cast<const float4*>(bt + (c * 64 + lane) * LB);
                const float4 b0 = bp[0], b1 = bp[1];
                const float eb[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
#pragma unroll
                for (int rb = 0; rb < 2; ++rb)
#pragma unroll
                    for (int q = 0; q < 8; ++q)
                        acc[rb][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(ea[rb][q], eb[q], acc[rb][c], 0, 0, 0);
            }
        } else if constexpr (H16) {
            f16x8_t ap[2][2];
            float4 anext[2][2];
            if constexpr (TILED) {
#pragma unroll
                for (int rb = 0; rb < 2; ++rb) load_tiled(rb, min(s + 1, nsteps - 1), anext[rb]);
#pragma unroll
                for (int rb = 0; rb < 2; ++rb) {
                    const float e[8] = {acur[rb][0].x, acur[rb][0].y, acur[rb][0].z, acur[rb][0].w,
                                        acur[rb][1].x, acur[rb][1].y, acur[rb][1].z, acur[rb][1].w};
                    uint32_t hi[4], lo[4];
                    split_f16(e, rscale[rb], hi, lo);
                    ap[rb][0] = __builtin_bit_cast(f16x8_t, (u32x4_t){hi[0], hi[1], hi[2], hi[3]});
                    ap[rb][1] = __builtin_bit_cast(f16x8_t, (u32x4_t){lo[0], lo[1], lo[2], lo[3]});
                }
            } else {
#pragma unroll
                for (int rb = 0; rb < 2; ++rb) {
                    float e[8];
                    load_a8(a, xr[rb], pwr[rb], kb, e);
                    uint32_t hi[4], lo[4];
                    split_f16(e, rscale[rb], hi, lo);
                    ap[rb][0] = __builtin_bit_cast(f16x8_t, (u32x4_t){hi[0], hi[1], hi[2], hi[3]});
                    ap[rb][1] = __builtin_bit_cast(f16x8_t, (u32x4_t){lo[0], lo[1], lo[2], lo[3]});
                }
            }
            if constexpr (!GEN) copy_b(min(s + 1, nsteps - 1), (s + 1) & 1);
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                f16x8_t b0, b1;
                if constexpr (GEN) {
                    if (s < a.gen_nT) {
                        const int t = 32 * s + 127 - 16 * c - (lane & 15) + 8 * (lane >> 4) - gsig;
                        b0 = *reinterpret_cast<const f16x8_t*>(gl + gsig * gS + t);
                        b1 = *reinterpret_cast<const f16x8_t*>(gl + 8 * gS + gsig * gS + t);
                    } else {
                        // rows W .. W + 31: lanes 0..15 hold k = W .. W + 7, of
                        // which W .. W + 2 carry r / 3; everything else is 0
                        const bool edge = lane < 16;
                        const uint32_t h = edge ? gl[16 * gS + 16 * c + lane] : 0u;
                        const uint32_t l = edge ? gl[16 * gS + 128 + 16 * c + lane] : 0u;
                        b0 = __builtin_bit_cast(f16x8_t, (u32x4_t){h | (h << 16), h, 0u, 0u});
                        b1 = __builtin_bit_cast(f16x8_t, (u32x4_t){l | (l << 16), l, 0u, 0u});
                    }
                } else {
                    const f16x8_t* bp = reinterpret_cast<const f16x8_t*>(bt + c * 64 * LB) + lane;
                    b0 = bp[0];
                    b1 = bp[64];
                }
#pragma unroll
                for (int rb = 0; rb < 2; ++rb) {
                    acc[rb][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ap[rb][1], b0, acc[rb][c], 0, 0, 0);
                    acc[rb][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ap[rb][0], b1, acc[rb][c], 0, 0, 0);
                    acc[rb][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ap[rb][0], b0, acc[rb][c], 0, 0, 0);
                }
            }
            if constexpr (TILED) {
#pragma unroll
                for (int rb = 0; rb < 2; ++rb) { acur[rb][0] = anext[rb][0]; acur[rb][1] = anext[rb][1]; }
            }
        } else {
            bf16x8_t ap[2][P];
            float4 anext[2][2];
            if constexpr (TILED) {
#pragma unroll
                for (int rb = 0; rb < 2; ++rb) load_tiled(rb, min(s + 1, nsteps - 1), anext[rb]);
            }
#pragma unroll
            for (int rb = 0; rb < 2; ++rb) {
                float e[8];
                if constexpr (TILED) {
                    const float4 x0 = acur[rb][0], x1 = acur[rb][1];
                    e[0] = x0.x; e[1] = x0.y; e[2] = x0.z; e[3] = x0.w;
                    e[4] = x1.x; e[5] = x1.y; e[6] = x1.z; e[7] = x1.w;
                } else {
                    load_a(rb, s, kb, e);
                }
                if constexpr (P == 3) {
                    // hi / mid / lo by truncation: each part is the top 16
                    // bits of the remainder (exact in BF16, same sign), the
                    // remainders are exact, and the three parts sum to the
                    // FP32 value exactly; one v_perm packs a part of two k
                    u32x4_t w[3];
#pragma unroll
                    for (int h = 0; h < 4; ++h) {
                        float r0 = e[2 * h], r1 = e[2 * h + 1];
#pragma unroll
                        for (int p = 0; p < 3; ++p) {
                            const uint32_t b0 = __float_as_uint(r0), b1 = __float_as_uint(r1);
                            w[p][h] = __builtin_amdgcn_perm(b1, b0, 0x07060302u);
                            if (p < 2) {
                                r0 = r0 - __uint_as_float(b0 & 0xffff0000u);
                                r1 = r1 - __uint_as_float(b1 & 0xffff0000u);
                            }
                        }
                    }
#pragma unroll
                    for (int p = 0; p < 3; ++p) ap[rb][p] = __builtin_bit_cast(bf16x8_t, w[p]);
                } else {
#pragma unroll
                    for (int q = 0; q < 8; ++q) {
                        float r = e[q];
#pragma unroll
                        for (int p = 0; p < P; ++p) {
                            const float h = bf16_round(r);
                            ap[rb][p][q] = to_bf16(h);
                            r = r - h;
                        }
                    }
                }
            }
            // unconditional (the last step recopies its own tile into the idle
            // buffer), so that the waits for A do not also wait for the copy
            copy_b(min(s + 1, nsteps - 1), (s + 1) & 1);
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const bf16x8_t* bp = reinterpret_cast<const bf16x8_t*>(bt + (c * 64 + lane) * LB);
                bf16x8_t bq[P];
#pragma unroll
                for (int p = 0; p < P; ++p) bq[p] = bp[p];
#pragma unroll
                for (int rb = 0; rb < 2; ++rb) {
                    if (P == 3) {
                        acc[rb][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[rb][1], bq[1], acc[rb][c], 0, 0, 0);
                        acc[rb][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[rb][0], bq[2], acc[rb][c], 0, 0, 0);
                        acc[rb][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[rb][2], bq[0], acc[rb][c], 0, 0, 0);
                    }
                    acc[rb][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[rb][1], bq[0], acc[rb][c], 0, 0, 0);
                    acc[rb][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[rb][0], bq[1], acc[rb][c], 0, 0, 0);
                    acc[rb][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[rb][0], bq[0], acc[rb][c], 0, 0, 0);
                }
            }
            if constexpr (TILED) {
#pragma unroll
                for (int rb = 0; rb < 2; ++rb) { acur[rb][0] = anext[rb][0]; acur[rb][1] = anext[rb][1]; }
            }
        }
        if constexpr (!GEN) {
            wait_lds_copies();
            __syncthreads();
        }
    }
    if constexpr (GEN) __syncthreads();  // the staged tile reuses the copies' LDS
    if constexpr (H16) {
        // the row factors 2^-e_r * bscale (exact powers of two) of rows
        // 4 (l / 16) + q of each row block (written by this wave)
#pragma unroll
        for (int rb = 0; rb < 2; ++rb)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float rf = rowfac[wave * 32 + rb * 16 + 4 * (lane >> 4) + q];
#pragma unroll
                for (int c = 0; c < 8; ++c) acc[rb][c][q] *= rf;
            }
    }
    if (a.q != nullptr) {
        stage_strips_epilogue(a, acc, lds, lane, wave, ctile, rtile);
        return;
    }
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const int col = (cb0 + c) * 16 + (lane & 15);
            if (col >= a.Wout) continue;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int rr = rowTile + rb * 16 + 4 * (lane >> 4) + q;
                if (rr < a.rows) a.out[(size_t)rr * a.Wout + col] = acc[rb][c][q];
            }
        }
}

static unsigned short bf16_rne_bits(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u = u + 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

static float bf16_to_float(unsigned short b)
{
    uint32_t u = (uint32_t)b << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// x (finite, |x| < 65520) rounded to the nearest FP16 value, ties to even,
// subnormals included: the value, and its bit pattern in *bits.
static double f16_rne(double x, unsigned short* bits)
{
    const unsigned short sign = std::signbit(x) ? 0x8000u : 0u;
    const double ax = std::fabs(x);
    if (ax == 0.0) {
        *bits = sign;
        return x;
    }
    int e;
    std::frexp(ax, &e);                   // ax = f 2^e, f in [0.5, 1)
    const int qe = std::max(e - 11, -24);  // quantum: 11 significant bits, or the subnormal step 2^-24
    const double n = std::nearbyint(std::ldexp(ax, -qe));  // ties to even (default rounding mode)
    const double v = std::ldexp(n, qe);
    unsigned short b;
    if (v < 0x1p-14) {
        b = (unsigned short)n;  // subnormal: v = n 2^-24
    } else {
        int ev;
        const double f = std::frexp(v, &ev);  // v = f 2^ev; FP16 exponent field ev + 14
        b = (unsigned short)(((ev + 14) << 10) | (int)std::ldexp(2.0 * f - 1.0, 10));
    }
    *bits = (unsigned short)(sign | b);
    return sign ? -v : v;
}

}  // namespace jh

using namespace jh;

namespace jh {

void filter_plan_shape(const jh_filter_plan* p, int* device, int* W, int* V, int* nviews)
{
    *device = p->device;
    *W = p->W == p->Wout && !p->mirror ? p->W : -1;  // the batch graph takes the plain form only
    *V = p->V;
    *nviews = p->N;
}

}  // namespace jh

namespace {

// FilterCore constructor with widthOut and yMirroring (FilterCore.cpp:13-61)
// plus the FilterAgent's pre-weight generator and schedule
// (FilterAgent.cpp:10-56, 101-126).
jh_status create_filter(jh_filter_plan** plan, int device, int W, int V, int Wout, int mirror, int nviews,
                        const double* h_ptm, float dso, float halfSampleRate, float cutStart, float cutEnd,
                        const float* adjPts, const float* adjDB, int nAdj, int window, int mode)
{
    return guard([&] {
        require(plan != nullptr, "jh_filter_create: plan is NULL");
        *plan = nullptr;
        require(W >= 3 && V >= 1 && nviews >= 1, "jh_filter_create: bad frame size or view count");
        require(Wout >= 1 && (mirror ? 2 * Wout : Wout) <= fft_length(W),
                "jh_filter_create: output columns beyond the FFT length");
        require(h_ptm != nullptr, "jh_filter_create: no projection matrices");
        require(dso > 1.0f && halfSampleRate > 0.0f, "jh_filter_create: bad DSO or sample rate");
        require(nAdj >= 0 && (nAdj == 0 || (adjPts && adjDB)), "jh_filter_create: bad adjust points");
        require(window == JH_WINDOW_REFERENCE || window == JH_WINDOW_SHEPP_LOGAN, "jh_filter_create: unknown window");
        require(mode == JH_FILTER_FP32 || mode == JH_FILTER_MFMA_BF16X3 || mode == JH_FILTER_MFMA_BF16X6 ||
                    mode == JH_FILTER_FP32_VALU || mode == JH_FILTER_MFMA_F16X3,
                "jh_filter_create: unknown mode");
        require(W % 4 == 0 || mode == JH_FILTER_FP32_VALU, "jh_filter_create: the MFMA forms need W % 4 == 0");
        int ndev = 0;
        JH_HIP(hipGetDeviceCount(&ndev));
        require(device >= 0 && device < ndev, "jh_filter_create: no such device");
        DeviceScope ds(device);
        auto* p = new jh_filter_plan;
        try {
            p->device = device;
            p->W = W; p->V = V; p->N = nviews; p->mode = mode;
            p->Wout = Wout; p->mirror = mirror ? 1 : 0;
            const int off = p->mirror ? Wout : 0;  // FilterCore.cpp:28
            p->L = fft_length(W);
            p->K = W + 3;
            p->Kp = (p->K + 31) / 32 * 32;
            p->bscale = 1.0f;
            p->bstride = p->Kp * 4 + (mode == JH_FILTER_MFMA_F16X3 ? 4 * operand_ktiles(p) : 0);
            std::vector<double> pts(adjPts, adjPts + nAdj), dbs(adjDB, adjDB + nAdj);
            p->H = design_filter(W, halfSampleRate, cutStart, cutEnd, pts, dbs, window);
            // convolution kernel h = irFFT(H) / L (real, even)
            const int L = p->L;
            const double pi = 3.14159265358979323846;
            std::vector<double> h(L);
            std::vector<double> cs(L);
            for (int i = 0; i < L; ++i) cs[i] = std::cos(2.0 * pi * i / L);
            for (int n = 0; n < L; ++n) {
                double s = p->H[0] + ((n & 1) ? -p->H[L / 2] : p->H[L / 2]);
                for (int k = 1; k < L / 2; ++k) s += 2.0 * p->H[k] * cs[(size_t)k * n % L];
                h[n] = s / L;
            }
            // h at a circular offset d (any sign)
            auto hc = [&](long d) { return h[(size_t)(((d % L) + L) % L)]; };
            // truncation ramp response at output column n (FilterCore.cpp:45-59,
            // 205-215): ramp i/R at L-R+i, or (R-1-i)/R at W+i with yMirroring
            const int R = (L - W) / 2;
            std::vector<double> rr(Wout, 0.0);
            for (int n = 0; n < Wout; ++n) {
                const long g = n + off;
                double s = 0.0;
                for (int i = 0; i < R; ++i)
                    s += p->mirror ? ((double)(R - 1 - i) / R) * hc(g - (W + i)) : ((double)i / R) * hc(g - (L - R + i));
                rr[n] = s;
            }
            std::vector<float> Bh((size_t)p->Kp * Wout, 0.0f);
            for (int m = 0; m < W; ++m)
                for (int n = 0; n < Wout; ++n) Bh[(size_t)m * Wout + n] = (float)hc((long)n + off - m);
            for (int e = 0; e < 3; ++e)
                for (int n = 0; n < Wout; ++n) Bh[(size_t)(W + e) * Wout + n] = (float)(rr[n] / 3.0);
            // the operator's infinity norm over its inputs (the pre-weight is
            // at most 1): |out[n]| <= gain_bound * max|in|
            p->gain_bound = 0.0;
            for (int n = 0; n < Wout; ++n) {
                double s = std::fabs(rr[n]);
                for (int m = 0; m < W; ++m) s += std::fabs(hc((long)n + off - m));
                p->gain_bound = std::max(p->gain_bound, s);
            }
            p->B.alloc(sizeof(float) * Bh.size());
            JH_HIP(hipMemcpy(p->B.get(), Bh.data(), sizeof(float) * Bh.size(), hipMemcpyHostToDevice));
            if (mode != JH_FILTER_FP32_VALU) {
                // operand packed per (k step s, column block cb, lane l): the 8
                // values B[32 s + 8 (l / 16) + q][16 cb + l % 16], q = 0..7, as
                // FP32, or as P BF16 parts (hi, [mid,] lo) split from the FP64 value
                const int ncb8 = (Wout + 127) / 128 * 8, nst = p->Kp / 32;
                const int P = mode == JH_FILTER_MFMA_BF16X6 ? 3 : (mode == JH_FILTER_MFMA_BF16X3 ? 2 : 0);
                auto bval = [&](int k, int col) -> double {
                    if (col >= Wout) return 0.0;
                    if (k < W) return hc((long)col + off - k);
                    if (k < W + 3) return rr[col] / 3.0;
                    return 0.0;
                };
                const size_t lanes = (size_t)nst * ncb8 * 64;
                if (mode == JH_FILTER_MFMA_F16X3) {
                    // hi / lo FP16 parts of B 2^eb, eb such that max|B| 2^eb
                    // lies in [2^13, 2^14): hi + lo keep 22 significant bits
                    // (lo may be subnormal: absolute error <= 2^-25)
                    double bmax = 0.0;
                    for (int k = 0; k < p->K; ++k)
                        for (int col = 0; col < Wout; ++col) bmax = std::max(bmax, std::fabs(bval(k, col)));
                    int e = 0;
                    if (bmax > 0.0) std::frexp(bmax, &e);
                    const int eb = 14 - e;
                    p->bscale = (float)std::ldexp(1.0, -eb);
                    // part-major per (k step, column block): the 64 lanes' hi
                    // parts, then their lo parts (each LDS read of a part is
                    // 64 lanes x 16 contiguous bytes)
                    std::vector<unsigned short> pk(lanes * 16);
                    for (int s = 0; s < nst; ++s)
                        for (int cb = 0; cb < ncb8; ++cb)
                            for (int l = 0; l < 64; ++l) {
                                unsigned short* dst = &pk[((size_t)s * ncb8 + cb) * 64 * 16 + l * 8];
                                for (int q = 0; q < 8; ++q) {
                                    const double v = std::ldexp(bval(32 * s + 8 * (l >> 4) + q, 16 * cb + (l & 15)), eb);
                                    const double hi = f16_rne(v, &dst[q]);
                                    f16_rne(v - hi, &dst[64 * 8 + q]);
                                }
                            }
                    p->Bhl.alloc(sizeof(unsigned short) * pk.size());
                    JH_HIP(hipMemcpy(p->Bhl.get(), pk.data(), sizeof(unsigned short) * pk.size(), hipMemcpyHostToDevice));
                    // the generated operand (filter_mfma_kernel GEN): the
                    // reversed diagonal R[i] = B[k][n] at n + off - k =
                    // Wout - 1 + off - i, as the same hi / lo parts, in 8
                    // copies shifted by 0..7 entries; and the edge rows
                    const size_t gen_lds = std::max<size_t>((size_t)128 * kStagePitch * 2,
                                                            ((size_t)16 * (W + 144) + 256) * 2);
                    if (W % 32 == 0 && Wout % 128 == 0 && gen_lds <= kGenMaxLds) {
                        const int stride = (Wout + W + 8 + 7) / 8 * 8;
                        std::vector<unsigned short> rh((size_t)stride + 8, 0), rl((size_t)stride + 8, 0);
                        for (int i = 0; i < Wout + W - 1; ++i) {
                            const double v = std::ldexp(hc((long)Wout - 1 + off - i), eb);
                            const double hi = f16_rne(v, &rh[i]);
                            f16_rne(v - hi, &rl[i]);
                        }
                        std::vector<unsigned short> g((size_t)16 * stride, 0);
                        for (int sg = 0; sg < 8; ++sg)
                            for (int t = 0; t < stride; ++t) {
                                g[(size_t)(2 * sg) * stride + t] = rh[(size_t)t + sg];
                                g[(size_t)(2 * sg + 1) * stride + t] = rl[(size_t)t + sg];
                            }
                        std::vector<unsigned short> r3((size_t)2 * Wout, 0);
                        for (int col = 0; col < Wout; ++col) {
                            const double v = std::ldexp(rr[col] / 3.0, eb);
                            const double hi = f16_rne(v, &r3[col]);
                            f16_rne(v - hi, &r3[(size_t)Wout + col]);
                        }
                        p->Bgen.alloc(sizeof(unsigned short) * g.size());
                        JH_HIP(hipMemcpy(p->Bgen.get(), g.data(), sizeof(unsigned short) * g.size(), hipMemcpyHostToDevice));
                        p->R3.alloc(sizeof(unsigned short) * r3.size());
                        JH_HIP(hipMemcpy(p->R3.get(), r3.data(), sizeof(unsigned short) * r3.size(), hipMemcpyHostToDevice));
                        p->gen_stride = stride;
                        p->gen_window = W + 128;
                        p->gen_lds = (int)gen_lds;
                        p->use_gen = 1;
                    }
                } else if (P == 0) {
                    std::vector<float> pk(lanes * 8);
                    for (int s = 0; s < nst; ++s)
                        for (int cb = 0; cb < ncb8; ++cb)
                            for (int l = 0; l < 64; ++l)
                                for (int q = 0; q < 8; ++q)
                                    pk[(((size_t)s * ncb8 + cb) * 64 + l) * 8 + q] =
                                        (float)bval(32 * s + 8 * (l >> 4) + q, 16 * cb + (l & 15));
                    p->Bhl.alloc(sizeof(float) * pk.size());
                    JH_HIP(hipMemcpy(p->Bhl.get(), pk.data(), sizeof(float) * pk.size(), hipMemcpyHostToDevice));
                } else {
                    std::vector<unsigned short> pk(lanes * 8 * P);
                    for (int s = 0; s < nst; ++s)
                        for (int cb = 0; cb < ncb8; ++cb)
                            for (int l = 0; l < 64; ++l) {
                                unsigned short* dst = &pk[(((size_t)s * ncb8 + cb) * 64 + l) * 8 * P];
                                for (int q = 0; q < 8; ++q) {
                                    double r = bval(32 * s + 8 * (l >> 4) + q, 16 * cb + (l & 15));
                                    for (int part = 0; part < P; ++part) {
                                        const unsigned short b = bf16_rne_bits((float)r);
                                        dst[part * 8 + q] = b;
                                        r -= (double)bf16_to_float(b);
                                    }
                                }
                            }
                    p->Bhl.alloc(sizeof(unsigned short) * pk.size());
                    JH_HIP(hipMemcpy(p->Bhl.get(), pk.data(), sizeof(unsigned short) * pk.size(), hipMemcpyHostToDevice));
                }
            }
            // pre-weight refresh schedule (FilterAgent.cpp:101-126)
            const int per = 9;
            p->nblocks = (nviews + per - 1) / per;
            std::vector<int> src(p->nblocks);
            for (int b = 0; b < p->nblocks; ++b) src[b] = std::min(b * per + per / 2, nviews - 1);
            p->view_block.resize(nviews);
            for (int v = 0; v < nviews; ++v) p->view_block[v] = v / per;
            p->blk.alloc(sizeof(int) * nviews);
            JH_HIP(hipMemcpy(p->blk.get(), p->view_block.data(), sizeof(int) * nviews, hipMemcpyHostToDevice));
            DevBuf dptm(sizeof(double) * 12 * nviews), dsrc(sizeof(int) * p->nblocks);
            DevBuf dinv(sizeof(double) * 16 * p->nblocks), dmin(sizeof(unsigned long long) * p->nblocks);
            JH_HIP(hipMemcpy(dptm.get(), h_ptm, sizeof(double) * 12 * nviews, hipMemcpyHostToDevice));
            JH_HIP(hipMemcpy(dsrc.get(), src.data(), sizeof(int) * p->nblocks, hipMemcpyHostToDevice));
            std::vector<unsigned long long> inf(p->nblocks, 0x7ff0000000000000ull);
            JH_HIP(hipMemcpy(dmin.get(), inf.data(), sizeof(unsigned long long) * p->nblocks, hipMemcpyHostToDevice));
            pw_inverse_kernel<<<(p->nblocks + 63) / 64, 64>>>(dptm.as<double>(), dsrc.as<int>(), p->nblocks, (double)dso,
                                                              dinv.as<double>());
            check_launch("pw_inverse_kernel");
            pw_min_kernel<<<dim3(256, p->nblocks), 256>>>(dinv.as<double>(), L, V, W, p->mirror,
                                                          dmin.as<unsigned long long>());
            check_launch("pw_min_kernel");
            p->pw.alloc(sizeof(float) * (size_t)p->nblocks * V * W);
            pw_write_kernel<<<dim3(256, p->nblocks), 256>>>(dinv.as<double>(), dmin.as<unsigned long long>(), W, V,
                                                             p->mirror, p->pw.as<float>());
            check_launch("pw_write_kernel");
            JH_HIP(hipDeviceSynchronize());
        } catch (...) {
            delete p;
            throw;
        }
        *plan = p;
    });
}


bool stage_fusable(const jh_filter_plan* p, const jh_bp_plan* bp, int dtype)
{
    return p != nullptr && bp != nullptr && p->mode != JH_FILTER_FP32_VALU && (dtype == JH_BF16 || dtype == JH_F16) &&
           bp->strip && !bp->zfilter && bp->device == p->device && p->Wout == bp->W && p->V == bp->V &&
           p->N == bp->N;
}

// One launch of the filter over views view0 .. view0+nviews-1 of the scan:
// FP32 rows to d_out, or (d_q != NULL) fused staging into column strips.
void run_filter(jh_filter_plan* p, const float* d_in, float* d_out, int view0, int nviews, jh_stream_t stream,
                void* d_q, const float* uw, int qtype, int tiled = 0, float qs = 1.0f)
{
    require(view0 >= 0 && nviews >= 0 && view0 + nviews <= p->N, "jh_filter_run: view range outside the scan");
    if (nviews == 0) return;
    DeviceScope ds(p->device);
    FiltArgs a;
    a.x = d_in;
    a.pw = p->pw.as<float>();
    a.blk = p->blk.as<int>();
    a.B = p->B.as<float>();
    a.Bhl = p->Bhl.get();
    a.out = d_out;
    a.W = p->W; a.V = p->V;
    a.Wout = p->Wout;
    a.edge0 = p->mirror ? p->W - 3 : 0;  // FilterCore.cpp:205-215
    a.rows = p->V * nviews;
    a.Kp = p->Kp;
    a.view0 = view0;
    a.q = d_q;
    a.uw = uw;
    a.qtype = qtype;
    a.qs = qs;
    a.tiled = tiled;
    a.bstride = p->bstride;
    a.bscale = p->bscale;
    a.Bgen = reinterpret_cast<const unsigned short*>(p->Bgen.get());
    a.R3 = reinterpret_cast<const unsigned short*>(p->R3.get());
    a.gen_stride = p->gen_stride;
    a.gen_window = p->gen_window;
    a.gen_nT = p->W / 32;
    require((size_t)p->V * nviews < (1u << 31), "jh_filter_run: too many rows");
    hipStream_t st = as_stream(stream);
    if (p->mode == JH_FILTER_FP32_VALU) {
        require(d_q == nullptr && !tiled, "jh_filter_run: the vector-ALU filter does not stage");
        dim3 grid((p->Wout + 63) / 64, (a.rows + 63) / 64);
        require(grid.y <= 65535u, "jh_filter_run: too many rows for one launch");
        filter_fp32_kernel<<<grid, 256, 0, st>>>(a);
        check_launch("filter_fp32_kernel");
    } else {
        // one workgroup per 128 x 128 output tile, 1-D (the kernel maps it to
        // tiles XCD by XCD)
        const size_t ntiles = (size_t)((p->Wout + 127) / 128) * ((a.rows + 127) / 128);
        require(ntiles <= 2147483647u, "jh_filter_run: too many rows for one launch");
        const dim3 grid((unsigned)ntiles);
        if (p->mode == JH_FILTER_FP32) {
            if (tiled) filter_mfma_kernel<kModeF32, true><<<grid, 256, 0, st>>>(a);
            else filter_mfma_kernel<kModeF32, false><<<grid, 256, 0, st>>>(a);
        } else if (p->mode == JH_FILTER_MFMA_BF16X3) {
            if (tiled) filter_mfma_kernel<kModeX3, true><<<grid, 256, 0, st>>>(a);
            else filter_mfma_kernel<kModeX3, false><<<grid, 256, 0, st>>>(a);
        } else if (p->mode == JH_FILTER_MFMA_F16X3 && p->use_gen) {
            if (tiled) filter_mfma_kernel<kModeH3, true, true><<<grid, 256, p->gen_lds, st>>>(a);
            else filter_mfma_kernel<kModeH3, false, true><<<grid, 256, p->gen_lds, st>>>(a);
        } else if (p->mode == JH_FILTER_MFMA_F16X3) {
            if (tiled) filter_mfma_kernel<kModeH3, true><<<grid, 256, 0, st>>>(a);
            else filter_mfma_kernel<kModeH3, false><<<grid, 256, 0, st>>>(a);
        } else {
            if (tiled) filter_mfma_kernel<kModeX6, true><<<grid, 256, 0, st>>>(a);
            else filter_mfma_kernel<kModeX6, false><<<grid, 256, 0, st>>>(a);
        }
        check_launch("filter_mfma_kernel");
    }
}

// Bytes of the pre-weighted A operand of nviews views: 16-row blocks of Kp
// floats (plus the k-tile row maxima in the split-FP16 form).
size_t operand_bytes(const jh_filter_plan* p, int nviews)
{
    const size_t rows = (size_t)p->V * nviews;
    return (rows + 15) / 16 * (size_t)p->bstride * 16;
}


}  // namespace

extern "C" {

jh_status jh_filter_create(jh_filter_plan** plan, int device, int W, int V, int nviews,
                           const double* h_ptm, float dso, float halfSampleRate, float cutStart,
                           float cutEnd, const float* adjPts, const float* adjDB, int nAdj,
                           int window, int mode)
{
    return create_filter(plan, device, W, V, W, 0, nviews, h_ptm, dso, halfSampleRate, cutStart, cutEnd, adjPts,
                         adjDB, nAdj, window, mode);
}

jh_status jh_filter_create_ex(jh_filter_plan** plan, int device, int Win, int V, int Wout, int yMirroring,
                              int nviews, const double* h_ptm, float dso, float halfSampleRate, float cutStart,
                              float cutEnd, const float* adjPts, const float* adjDB, int nAdj, int window,
                              int mode)
{
    return create_filter(plan, device, Win, V, Wout, yMirroring, nviews, h_ptm, dso, halfSampleRate, cutStart,
                         cutEnd, adjPts, adjDB, nAdj, window, mode);
}

jh_status jh_filter_shape(const jh_filter_plan* p, int* Win, int* Wout, int* yMirroring)
{
    return guard([&] {
        require(p != nullptr, "jh_filter_shape: plan is NULL");
        if (Win) *Win = p->W;
        if (Wout) *Wout = p->Wout;
        if (yMirroring) *yMirroring = p->mirror;
    });
}

jh_status jh_filter_output_bound(const jh_filter_plan* p, double inBound, double* outBound)
{
    return guard([&] {
        require(p != nullptr && outBound != nullptr, "jh_filter_output_bound: bad arguments");
        require(inBound >= 0.0 && std::isfinite(inBound), "jh_filter_output_bound: bad input bound");
        *outBound = inBound * p->gain_bound;
    });
}

jh_status jh_filter_response(jh_filter_plan* p, double* h_H, int* L)
{
    return guard([&] {
        require(p != nullptr && h_H != nullptr && L != nullptr, "jh_filter_response: bad arguments");
        std::memcpy(h_H, p->H.data(), sizeof(double) * p->H.size());
        *L = p->L;
    });
}

jh_status jh_filter_preweight(jh_filter_plan* p, int view, float* d_out, jh_stream_t stream)
{
    return guard([&] {
        require(p != nullptr && d_out != nullptr, "jh_filter_preweight: bad arguments");
        require(view >= 0 && view < p->N, "jh_filter_preweight: view outside the scan");
        DeviceScope ds(p->device);
        const size_t n = (size_t)p->V * p->W;
        JH_HIP(hipMemcpyAsync(d_out, p->pw.as<float>() + (size_t)p->view_block[view] * n, sizeof(float) * n,
                              hipMemcpyDeviceToDevice, as_stream(stream)));
    });
}

jh_status jh_filter_run(jh_filter_plan* p, const float* d_in, float* d_out, int view0, int nviews,
                        jh_stream_t stream)
{
    return guard([&] {
        require(p != nullptr && d_in != nullptr && d_out != nullptr, "jh_filter_run: bad arguments");
        require(d_in != d_out, "jh_filter_run: in-place filtering is not supported");
        run_filter(p, d_in, d_out, view0, nviews, stream, nullptr, nullptr, 0);
    });
}

jh_status jh_filter_stage_fusable(const jh_filter_plan* p, const jh_bp_plan* bp, int dtype, int* fusable)
{
    return guard([&] {
        require(fusable != nullptr, "jh_filter_stage_fusable: bad arguments");
        *fusable = stage_fusable(p, bp, dtype) ? 1 : 0;
    });
}

jh_status jh_filter_run_staged(jh_filter_plan* p, jh_bp_plan* bp, const float* d_in, int view0, int nviews,
                               int dtype, void* d_q, jh_stream_t stream)
{
    return guard([&] {
        require(p != nullptr && bp != nullptr && d_in != nullptr && d_q != nullptr,
                "jh_filter_run_staged: bad arguments");
        require(stage_fusable(p, bp, dtype),
                "jh_filter_run_staged: the plans do not fuse (MFMA filter writing the backprojector's W columns, "
                "BF16/FP16 column strips, no ZFilter, same device, frame size and view count)");
        run_filter(p, d_in, nullptr, view0, nviews, stream, d_q, bp->uw.as<float>() + (size_t)view0 * bp->W, dtype, 0,
                   bp_stage_scale(bp, dtype));
    });
}

jh_status jh_filter_operand_bytes(const jh_filter_plan* p, int nviews, size_t* bytes)
{
    return guard([&] {
        require(p != nullptr && bytes != nullptr && nviews >= 0, "jh_filter_operand_bytes: bad arguments");
        require(p->mode != JH_FILTER_FP32_VALU, "jh_filter_operand_bytes: the vector-ALU filter takes no operand");
        *bytes = operand_bytes(p, nviews);
    });
}

jh_status jh_pre_run_operand(jh_pre_plan* pre, jh_filter_plan* p, const uint16_t* d_raw, const uint16_t* d_dark,
                             const float* d_gain, int view0, int nviews, float* d_a, jh_stream_t stream)
{
    return guard([&] {
        require(pre != nullptr && p != nullptr && d_raw != nullptr && d_a != nullptr,
                "jh_pre_run_operand: bad arguments");
        require(view0 >= 0 && nviews >= 0 && view0 + nviews <= p->N, "jh_pre_run_operand: view range outside the scan");
        require(p->mode != JH_FILTER_FP32_VALU, "jh_pre_run_operand: the vector-ALU filter takes no operand");
        require(pre->device == p->device, "jh_pre_run_operand: plans on different devices");
        const int roiW = pre->inW - pre->bLeft - pre->bRight, roiH = pre->inH - pre->bUp - pre->bDown;
        require(roiW == pre->outW && roiH == pre->outH, "jh_pre_run_operand: resizing preprocessing takes jh_pre_run");
        require(pre->outW == p->W && pre->outH == p->V, "jh_pre_run_operand: frame size differs from the filter's input");
        if (nviews == 0) return;
        const size_t rows = (size_t)p->V * nviews;
        require((rows + 15) / 16 <= 65535u, "jh_pre_run_operand: more than 1048560 rows in one call");
        DeviceScope ds(p->device);
        PreOpArgs a;
        a.raw = d_raw;
        a.dark = d_dark;
        a.gain = d_gain;
        a.pw = p->pw.as<float>();
        a.blk = p->blk.as<int>();
        a.out = reinterpret_cast<float4*>(d_a);
        a.inW = pre->inW; a.inH = pre->inH; a.bUp = pre->bUp; a.bLeft = pre->bLeft;
        a.W = p->W; a.V = p->V; a.rows = (int)rows; a.nsteps = p->Kp / 32; a.view0 = view0;
        a.edge0 = p->mirror ? p->W - 3 : 0;
        a.bstride = p->bstride;
        a.rowmax = p->mode == JH_FILTER_MFMA_F16X3 ? 1 : 0;
        a.bf = pre->brightField; a.a0 = pre->bh[0]; a.a1 = pre->bh[1]; a.a2 = pre->bh[2];
        a.doBH = pre->doBH ? 1 : 0;
        // 16-byte loads need 8-count alignment of every row start (and of the frames and tables)
        a.vec = (pre->inW % 8 == 0 && pre->bLeft % 8 == 0 && p->W % 8 == 0 &&
                 reinterpret_cast<uintptr_t>(d_raw) % 16 == 0 &&
                 (d_dark == nullptr || reinterpret_cast<uintptr_t>(d_dark) % 16 == 0) &&
                 (d_gain == nullptr || reinterpret_cast<uintptr_t>(d_gain) % 16 == 0)) ? 1 : 0;
        const unsigned nblk = (unsigned)((rows + 15) / 16);
        const dim3 grid((unsigned)((p->Kp + 127) / 128), nblk);
        pre_operand_kernel<<<grid, 256, 0, as_stream(stream)>>>(a);
        check_launch("pre_operand_kernel");
    });
}

jh_status jh_filter_run_operand_staged(jh_filter_plan* p, jh_bp_plan* bp, const float* d_a, int view0, int nviews,
                                       int dtype, void* d_q, jh_stream_t stream)
{
    return guard([&] {
        require(p != nullptr && bp != nullptr && d_a != nullptr && d_q != nullptr,
                "jh_filter_run_operand_staged: bad arguments");
        require(stage_fusable(p, bp, dtype),
                "jh_filter_run_operand_staged: the plans do not fuse (MFMA filter writing the backprojector's W "
                "columns, BF16/FP16 column strips, no ZFilter, same device, frame size and view count)");
        run_filter(p, d_a, nullptr, view0, nviews, stream, d_q, bp->uw.as<float>() + (size_t)view0 * bp->W, dtype,
                   1, bp_stage_scale(bp, dtype));
    });
}

jh_status jh_filter_set_generated(jh_filter_plan* p, int enable)
{
    return guard([&] {
        require(p != nullptr, "jh_filter_set_generated: plan is NULL");
        require(!enable || p->Bgen.get() != nullptr,
                "jh_filter_set_generated: the plan has no generated operand (split-FP16 form, W % 32 == 0, "
                "output width % 128 == 0)");
        p->use_gen = enable ? 1 : 0;
    });
}

jh_status jh_filter_get_generated(jh_filter_plan* p, int* enabled)
{
    return guard([&] {
        require(p != nullptr && enabled != nullptr, "jh_filter_get_generated: bad arguments");
        *enabled = p->use_gen;
    });
}

jh_status jh_filter_destroy(jh_filter_plan* p)
{
    return guard([&] {
        if (!p) return;
        DeviceScope ds(p->device);
        delete p;
    });
}

}  // extern "C"
