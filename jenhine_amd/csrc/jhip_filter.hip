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
// Two arithmetic forms:
//  * JH_FILTER_FP32: FP32 FMA on the vector ALUs (LDS-tiled SIMT GEMM);
//  * JH_FILTER_MFMA_BF16X3: MFMA 16x16x32 BF16 with hi/lo split operands,
//    hi*hi + hi*lo + lo*hi products accumulated in FP32.
//
// Pre-filter weight maps (FilterAgent/PreFilterWeightGenerator.cpp:10-124)
// are built on the device in FP64, one per refresh block of 9 views
// (FilterAgent.cpp:101-126).
#include <cmath>
#include <cstring>

#include "jhip_internal.h"

struct jh_filter_plan {
    int device;
    int W, V, N, L, K, Kp;
    int mode;
    int nblocks;
    std::vector<int> view_block;  // block index of every view
    std::vector<double> H;        // L/2 + 1
    jh::DevBuf B;                 // Kp x W float (FP32 form)
    jh::DevBuf Bhl;               // MFMA form: packed hi/lo BF16 operand
    jh::DevBuf pw;                // nblocks x V x W float
    jh::DevBuf blk;               // N int: block of every view (device copy)
};

namespace jh {

// ---------------------------------------------------------------------------
// Pre-filter weight maps, FP64.
// ---------------------------------------------------------------------------

// TM of TransformMatrix/MatrixConvertor.cpp:8-27 and its inverse (Gauss-Jordan
// with partial pivoting), one thread per refresh block.
__global__ void pw_inverse_kernel(const double* __restrict__ ptm, const int* __restrict__ src_view,
                                  int nb, double dso, double* __restrict__ inv)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const double* P = ptm + 12 * (size_t)src_view[b];
    const double far_ = (1.0 + 1.0 / dso) * dso, near_ = (1.0 - 1.0 / dso) * dso;
    const double s20 = (far_ + near_) / (far_ - near_);
    const double s23 = ((far_ + near_) - 2.0 * far_ * near_ / dso) / (far_ - near_) / dso;
    double A[4][8];
    for (int c = 0; c < 4; ++c) { A[0][c] = P[c]; A[1][c] = P[4 + c]; A[3][c] = P[8 + c]; }
    A[2][0] = s20 * P[8]; A[2][1] = s20 * P[9]; A[2][2] = 0.0; A[2][3] = s23 * P[11];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) A[r][4 + c] = (r == c) ? 1.0 : 0.0;
    for (int k = 0; k < 4; ++k) {
        int p = k;
        for (int r = k + 1; r < 4; ++r)
            if (fabs(A[r][k]) > fabs(A[p][k])) p = r;
        if (p != k)
            for (int c = 0; c < 8; ++c) { double t = A[k][c]; A[k][c] = A[p][c]; A[p][c] = t; }
        const double d = A[k][k];
        for (int c = 0; c < 8; ++c) A[k][c] /= d;
        for (int r = 0; r < 4; ++r) {
            if (r == k) continue;
            const double f = A[r][k];
            for (int c = 0; c < 8; ++c) A[r][c] -= f * A[k][c];
        }
    }
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) inv[16 * (size_t)b + 4 * r + c] = A[r][4 + c];
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
    const float* B;      // Kp x W
    const void* Bhl;     // MFMA operand
    float* out;          // rows x W
    int W, V, rows, Kp, view0;
};

// Element (r, m) of the A operand.
__device__ inline float a_elem(const FiltArgs& a, int r, int m, const float* xr, const float* pwr)
{
    if (m < a.W) return xr[m] * pwr[m];
    if (m < a.W + 3) return xr[m - a.W];
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
            const float* bp = a.B + (size_t)(k0 + bk) * a.W + col0 + bn;
#pragma unroll
            for (int q = 0; q < 4; ++q) Bs[bk][bn + q] = (col0 + bn + q < a.W) ? bp[q] : 0.0f;
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
            if (cc < a.W) a.out[(size_t)rr * a.W + cc] = c[p][q];
        }
    }
}

// ---------------------------------------------------------------------------
// MFMA Toeplitz GEMM with split BF16 (hi/lo).
//
// v_mfma_f32_16x16x32_bf16: lane l supplies A[row = l%16][k = 8*(l/16) .. +7]
// and B[k = 8*(l/16) .. +7][col = l%16]; it receives D[row = 4*(l/16) + i]
// [col = l%16], i = 0..3.
// Workgroup: 256 threads = 4 waves, tile 64 rows x 64 columns; wave w owns
// rows 16*w..16*w+15 and all 4 column blocks of 16.  K step 32.
// B operand layout (built on the host): for each (k-step s, column block cb,
// lane l) 16 bf16 values = hi[8] then lo[8] of B[32 s + 8 (l/16) + 0..7]
// [16 cb + l%16], i.e. one 32-byte load per lane per column block.
// ---------------------------------------------------------------------------

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

__device__ inline void split_bf16(float f, __bf16& hi, __bf16& lo)
{
    // hi = round-to-nearest-even BF16 of f, lo = BF16 of the remainder
    uint32_t u = __float_as_uint(f);
    uint32_t r = (u + 0x7fffu + ((u >> 16) & 1u)) & 0xffff0000u;
    const float fh = __uint_as_float(r);
    const float rem = f - fh;
    uint32_t v = __float_as_uint(rem);
    v = (v + 0x7fffu + ((v >> 16) & 1u)) & 0xffff0000u;
    hi = __builtin_bit_cast(__bf16, (unsigned short)(r >> 16));
    lo = __builtin_bit_cast(__bf16, (unsigned short)(v >> 16));
}

__global__ void __launch_bounds__(256) filter_mfma_kernel(FiltArgs a)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int row0 = blockIdx.y * 64 + wave * 16;
    const int col0 = blockIdx.x * 64;
    const int ncb = (a.W + 15) / 16;  // column blocks of 16 in the whole matrix
    const int cb0 = col0 / 16;
    // this lane's A row
    const int r = min(row0 + (lane & 15), a.rows - 1);
    const int view = r / a.V, vrow = r - view * a.V;
    const float* xr = a.x + (size_t)r * a.W;
    const float* pwr = a.pw + ((size_t)a.blk[a.view0 + view] * a.V + vrow) * a.W;
    const int kq = 8 * (lane >> 4);
    f32x4_t acc[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[c] = (f32x4_t){0.f, 0.f, 0.f, 0.f};
    const bf16x8_t* Bp = reinterpret_cast<const bf16x8_t*>(a.Bhl);
    const int nsteps = a.Kp / 32;
    for (int s = 0; s < nsteps; ++s) {
        bf16x8_t ahi, alo;
        const int kb = 32 * s + kq;
        if (kb + 8 <= a.W) {
            const float4 x0 = *reinterpret_cast<const float4*>(xr + kb);
            const float4 x1 = *reinterpret_cast<const float4*>(xr + kb + 4);
            const float4 w0 = *reinterpret_cast<const float4*>(pwr + kb);
            const float4 w1 = *reinterpret_cast<const float4*>(pwr + kb + 4);
            const float e[8] = {x0.x * w0.x, x0.y * w0.y, x0.z * w0.z, x0.w * w0.w,
                                x1.x * w1.x, x1.y * w1.y, x1.z * w1.z, x1.w * w1.w};
#pragma unroll
            for (int q = 0; q < 8; ++q) { __bf16 h, l; split_bf16(e[q], h, l); ahi[q] = h; alo[q] = l; }
        } else {
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int m = kb + q;
                float e = 0.0f;
                if (m < a.W) e = xr[m] * pwr[m];
                else if (m < a.W + 3) e = xr[m - a.W];
                __bf16 h, l;
                split_bf16(e, h, l);
                ahi[q] = h; alo[q] = l;
            }
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int cb = cb0 + c;
            if (cb >= ncb) break;
            const bf16x8_t* bp = Bp + (((size_t)s * ncb + cb) * 64 + lane) * 2;
            const bf16x8_t bhi = bp[0];
            const bf16x8_t blo = bp[1];
            acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo, bhi, acc[c], 0, 0, 0);
            acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, blo, acc[c], 0, 0, 0);
            acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, bhi, acc[c], 0, 0, 0);
        }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int col = col0 + 16 * c + (lane & 15);
        if (col >= a.W) continue;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int rr = blockIdx.y * 64 + wave * 16 + 4 * (lane >> 4) + q;
            if (rr < a.rows) a.out[(size_t)rr * a.W + col] = acc[c][q];
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

}  // namespace jh

using namespace jh;

extern "C" {

jh_status jh_filter_create(jh_filter_plan** plan, int device, int W, int V, int nviews,
                           const double* h_ptm, float dso, float halfSampleRate, float cutStart,
                           float cutEnd, const float* adjPts, const float* adjDB, int nAdj,
                           int window, int mode)
{
    return guard([&] {
        require(plan != nullptr, "jh_filter_create: plan is NULL");
        *plan = nullptr;
        require(W >= 3 && V >= 1 && nviews >= 1, "jh_filter_create: bad frame size or view count");
        require(h_ptm != nullptr, "jh_filter_create: no projection matrices");
        require(dso > 1.0f && halfSampleRate > 0.0f, "jh_filter_create: bad DSO or sample rate");
        require(nAdj >= 0 && (nAdj == 0 || (adjPts && adjDB)), "jh_filter_create: bad adjust points");
        require(window == JH_WINDOW_REFERENCE || window == JH_WINDOW_SHEPP_LOGAN, "jh_filter_create: unknown window");
        require(mode == JH_FILTER_FP32 || mode == JH_FILTER_MFMA_BF16X3, "jh_filter_create: unknown mode");
        require(W % 8 == 0 || mode == JH_FILTER_FP32, "jh_filter_create: the MFMA form needs W % 8 == 0");
        int ndev = 0;
        JH_HIP(hipGetDeviceCount(&ndev));
        require(device >= 0 && device < ndev, "jh_filter_create: no such device");
        DeviceScope ds(device);
        auto* p = new jh_filter_plan;
        try {
            p->device = device;
            p->W = W; p->V = V; p->N = nviews; p->mode = mode;
            p->L = fft_length(W);
            p->K = W + 3;
            p->Kp = (p->K + 31) / 32 * 32;
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
            const int R = (L - W) / 2;
            std::vector<double> rr(W, 0.0);
            for (int n = 0; n < W; ++n) {
                double s = 0.0;
                for (int i = 0; i < R; ++i) s += ((double)i / R) * h[(n + R - i) % L];
                rr[n] = s;
            }
            std::vector<float> Bh((size_t)p->Kp * W, 0.0f);
            for (int m = 0; m < W; ++m)
                for (int n = 0; n < W; ++n) Bh[(size_t)m * W + n] = (float)h[std::abs(n - m)];
            for (int e = 0; e < 3; ++e)
                for (int n = 0; n < W; ++n) Bh[(size_t)(W + e) * W + n] = (float)(rr[n] / 3.0);
            p->B.alloc(sizeof(float) * Bh.size());
            JH_HIP(hipMemcpy(p->B.get(), Bh.data(), sizeof(float) * Bh.size(), hipMemcpyHostToDevice));
            if (mode == JH_FILTER_MFMA_BF16X3) {
                const int ncb = (W + 15) / 16, nst = p->Kp / 32;
                std::vector<unsigned short> pk((size_t)nst * ncb * 64 * 16);
                for (int s = 0; s < nst; ++s)
                    for (int cb = 0; cb < ncb; ++cb)
                        for (int l = 0; l < 64; ++l) {
                            unsigned short* dst = &pk[(((size_t)s * ncb + cb) * 64 + l) * 16];
                            const int col = 16 * cb + (l & 15);
                            for (int q = 0; q < 8; ++q) {
                                const int k = 32 * s + 8 * (l >> 4) + q;
                                // hi/lo from the FP64 value for the best split
                                double bv = 0.0;
                                if (col < W && k < W) bv = h[std::abs(col - k)];
                                else if (col < W && k < W + 3) bv = rr[col] / 3.0;
                                const unsigned short hb = bf16_rne_bits((float)bv);
                                const unsigned short lb = bf16_rne_bits((float)(bv - (double)bf16_to_float(hb)));
                                dst[q] = hb;
                                dst[8 + q] = lb;
                            }
                        }
                p->Bhl.alloc(sizeof(unsigned short) * pk.size());
                JH_HIP(hipMemcpy(p->Bhl.get(), pk.data(), sizeof(unsigned short) * pk.size(), hipMemcpyHostToDevice));
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
            pw_min_kernel<<<dim3(256, p->nblocks), 256>>>(dinv.as<double>(), L, V, W, 0, dmin.as<unsigned long long>());
            check_launch("pw_min_kernel");
            p->pw.alloc(sizeof(float) * (size_t)p->nblocks * V * W);
            pw_write_kernel<<<dim3(256, p->nblocks), 256>>>(dinv.as<double>(), dmin.as<unsigned long long>(), W, V, 0,
                                                             p->pw.as<float>());
            check_launch("pw_write_kernel");
            JH_HIP(hipDeviceSynchronize());
        } catch (...) {
            delete p;
            throw;
        }
        *plan = p;
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
        require(view0 >= 0 && nviews >= 0 && view0 + nviews <= p->N, "jh_filter_run: view range outside the scan");
        require(d_in != d_out, "jh_filter_run: in-place filtering is not supported");
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
        a.rows = p->V * nviews;
        a.Kp = p->Kp;
        a.view0 = view0;
        require((size_t)p->V * nviews < (1u << 31), "jh_filter_run: too many rows");
        dim3 grid((p->W + 63) / 64, (a.rows + 63) / 64);
        require(grid.y <= 65535u * 16u, "jh_filter_run: too many rows for one launch");
        if (p->mode == JH_FILTER_FP32) {
            filter_fp32_kernel<<<grid, 256, 0, as_stream(stream)>>>(a);
            check_launch("filter_fp32_kernel");
        } else {
            filter_mfma_kernel<<<grid, 256, 0, as_stream(stream)>>>(a);
            check_launch("filter_mfma_kernel");
        }
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
