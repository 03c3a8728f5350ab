// Offset-detector composite on gfx950 (CompositeAgent/CompositeCore.cpp).
//
// A half-fan scan puts the detector to one side of the central ray.  The
// reference widens every preprocessed view to 2W columns before filtering:
// the direct frame fills one half and the other half is the conjugate view
// (v + N/2) % N re-sampled through the iso-centre plane (JEngine.cpp:213-238,
// InputAgent.cpp:47-85).  Per output pixel (c, r) of the conjugate half the
// reference forms the virtual detector point (u1, v1, 0, 1), u1 = c - W when
// the detector is on the right, else c + W, back-projects it with the
// inverse transform matrix of the direct view and projects it with the
// conjugate view's matrix (CompositeCore.cpp:66-104).  Because the
// pseudo-depth coordinate is 0, that chain is a 3 x 3 homography per view:
//   (U, V, S) = G (u1, v1, 1),  G = P_conj * TM_direct^-1 [columns 0, 1, 3],
// composed once per view in FP64 on the device and rounded to float; the
// per-pixel work is then 3 multiply-add chains and two divisions in FP32.
//
// The conjugate frame is sampled bilinearly at (U/S, clamp(V/S, 0, V-1));
// a sample outside the frame leaves 0 (the remap semantics of the
// backprojector).  The 40 conjugate columns next to the direct half are
// blended towards the mean of the 5 direct pixels at the shared border:
// out = conj * maskConj[i] + edge * maskDirc[i] (CompositeCore.cpp:106-148).
// The masks are formed on the host in double and rounded once.
#include <cmath>
#include <vector>

#include "jhip_filter_plan.h"
#include "jhip_internal.h"

struct jh_composite_plan {
    int device;
    int W, V, N;
    int detRight;
    jh::DevBuf G;       // N x 9 float: the per-view homography
    float maskDirc[40];
    float maskConj[40];
};

namespace jh {

namespace {

constexpr int kTransWidth = 40;  // CompositeCore.cpp:41
constexpr int kEdgeWidth = 5;    // CompositeCore.cpp:42

// G of every view in FP64: TM^-1 of the view, times the conjugate view's
// matrix, keeping the columns that multiply u1, v1 and the homogeneous 1.
__global__ void composite_setup_kernel(const double* __restrict__ ptm, int N, double dso, float* __restrict__ G)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    double inv[16];
    tm_inverse_fp64(ptm + 12 * (size_t)v, dso, inv);
    const double* Pc = ptm + 12 * (size_t)((v + N / 2) % N);
    const int cols[3] = {0, 1, 3};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 4; ++k) s += Pc[4 * i + k] * inv[4 * k + cols[j]];
            G[9 * (size_t)v + 3 * i + j] = (float)s;
        }
}

struct CompArgs {
    const float* dirc;
    const float* conj;
    long long inStride;  // elements between consecutive input frames
    float* out;          // nviews x V x 2W
    const float* G;      // N x 9
    int W, V, view0, nviews, detRight;
    float md[kTransWidth];
    float mc[kTransWidth];
};

#pragma clang fp contract(off)
// Bilinear sample of frame img (W x V) at (su, sv); false outside the frame.
__device__ __forceinline__ bool remap_at(const float* __restrict__ img, int W, int V, float su, float sv, float& val)
{
    if (!(su >= 0.0f && su <= (float)(W - 1) && sv >= 0.0f && sv <= (float)(V - 1))) return false;
    int x0 = (int)su, y0 = (int)sv;
    x0 = min(x0, W - 2);
    y0 = min(y0, V - 2);
    const float fx = su - (float)x0, fy = sv - (float)y0;
    const float* r0 = img + (size_t)y0 * W + x0;
    const float* r1 = r0 + W;
    const float top = r0[0] + fx * (r0[1] - r0[0]);
    const float bot = r1[0] + fx * (r1[1] - r1[0]);
    val = top + fy * (bot - top);
    return true;
}

// Pixel (r, c) of a frame's conjugate half (output column conjOff + c):
// dirc and conj are the frame's direct and conjugate frames, g its
// homography.
__device__ __forceinline__ float conj_value_at(const CompArgs& a, const float* __restrict__ dirc,
                                               const float* __restrict__ conj, const float* __restrict__ g, int r,
                                               int c)
{
    const int W = a.W, V = a.V;
    const float u1 = a.detRight ? (float)c - (float)W : (float)(c + W);
    const float v1 = (float)r;
    const float U = fmaf(g[0], u1, fmaf(g[1], v1, g[2]));
    const float Vv = fmaf(g[3], u1, fmaf(g[4], v1, g[5]));
    const float S = fmaf(g[6], u1, fmaf(g[7], v1, g[8]));
    const float u2 = U / S;
    float v2 = Vv / S;
    if (v2 < 0.0f) v2 = 0.0f;
    if (v2 > (float)(V - 1)) v2 = (float)(V - 1);
    float val = 0.0f;
    remap_at(conj, W, V, u2, v2, val);
    const int conjOff = a.detRight ? 0 : W;
    const int transOff = a.detRight ? W - kTransWidth : W;
    const int oc = conjOff + c;
    const float* drow = dirc + (size_t)r * W;
    if (oc >= transOff && oc < transOff + kTransWidth) {
        const int i = oc - transOff;
        const float* e = a.detRight ? drow : drow + W - kEdgeWidth;
        float sum = e[0];
#pragma unroll
        for (int k = 1; k < kEdgeWidth; ++k) sum = sum + e[k];
        const float edge = sum / (float)kEdgeWidth;
        const float x = val * a.mc[i];
        const float y = edge * a.md[i];
        val = x + y;
    }
    return val;
}

// Pixel (r, c) of frame f's conjugate half.
__device__ __forceinline__ float conj_value(const CompArgs& a, int f, int r, int c)
{
    return conj_value_at(a, a.dirc + (size_t)f * a.inStride, a.conj + (size_t)f * a.inStride,
                         a.G + 9 * (size_t)(a.view0 + f), r, c);
}

// Output column oc (0 .. 2W-1) of row r of composite frame f.
__device__ __forceinline__ float composite_value(const CompArgs& a, int f, int r, int oc)
{
    const int W = a.W;
    const int dOff = a.detRight ? W : 0;
    if (oc >= dOff && oc < dOff + W) return a.dirc[(size_t)f * a.inStride + (size_t)r * W + (oc - dOff)];
    return conj_value(a, f, r, oc - (a.detRight ? 0 : W));
}

// One thread per pixel of the conjugate half; the direct half is copied by
// the same thread (column c of the direct frame).
__global__ void __launch_bounds__(256) composite_kernel(CompArgs a)
{
    const int W = a.W, V = a.V, W2 = 2 * W;
    const size_t total = (size_t)a.nviews * V * W;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const int f = (int)(t / ((size_t)V * W));
        const int rem = (int)(t - (size_t)f * V * W);
        const int r = rem / W, c = rem - r * W;
        float* o = a.out + ((size_t)f * V + r) * W2;
        o[(a.detRight ? 0 : W) + c] = conj_value(a, f, r, c);
        o[(a.detRight ? W : 0) + c] = a.dirc[(size_t)f * a.inStride + (size_t)r * W + c];
    }
}

// The composite written straight into the filter's pre-weighted operand (the
// layout of pre_operand_kernel, for a filter whose input is the 2W-wide
// composite): element (R, k) of operand row R = (view0 - aview0 + f) V + r is
// the composite pixel times the pre-weight for k < 2W, the unweighted
// composite pixel edge0 + k - 2W for the three edge columns, 0 beyond.  The
// launch writes the rows of its views only (each float4 of the operand holds
// one row), so the two halves of a scan can be written by two launches.
// Workgroup: 16 operand rows x 128 k, through LDS as in pre_operand_kernel.
struct CompOpArgs {
    CompArgs c;
    float4* out;
    const float* pw;  // the filter's pre-weight maps, nblocks x V x 2W
    const int* blk;   // block of every view of the scan
    int R0, R1;       // operand rows of this launch's views
    int blockRow0;    // first 16-row block of the launch
    int nsteps, aview0, edge0;
    int bstride;      // float4s per 16-row block of the operand
    int rowmax;       // split-FP16 filter: write the row maxima of the k tile
};

constexpr int kCompOpPitch = 132;
constexpr int kCompOpThreads = 128;  // 16 rows x 128 columns: 16 rows per thread

__global__ void __launch_bounds__(kCompOpThreads) composite_operand_kernel(CompOpArgs a)
{
    __shared__ float tile[16 * kCompOpPitch];
    const int W2 = 2 * a.c.W, V = a.c.V;
    const int t = threadIdx.x;
    const int rb = a.blockRow0 + blockIdx.y;
    // compute phase: consecutive lanes take consecutive columns of one row
    // (coalesced frame reads); thread t holds rows t / 128 + kRowStep j
    constexpr int kRowStep = kCompOpThreads / 128;
    const int kc = t & 127, k = blockIdx.x * 128 + kc;
    // one row per wave: scalar geometry, pre-weight row and view block.  The
    // wave's rows step by kRowStep; (f, r) of the first row at or after R0
    // comes from one division and is advanced from there.
    const int rows = a.R1 - a.R0;
    const int vbase = a.aview0 + a.R0 / V;  // scan view of launch row R0
    int rel = __builtin_amdgcn_readfirstlane(rb * 16 + (t >> 7) - a.R0);
    const int relS = rel < 0 ? rel + (kRowStep - 1 - rel) / kRowStep * kRowStep : rel;  // first row >= R0
    int f = relS / V, r = relS - f * V;
    // the frame's pointers, advanced per view; the column this lane reads
    const int W = a.c.W;
    const int dOff = a.c.detRight ? W : 0, cOff = a.c.detRight ? 0 : W;
    const int kk = k < W2 ? k : a.edge0 + k - W2;  // composite column of operand column k
    const bool direct = kk >= dOff && kk < dOff + W;
    const int fc = min(f, a.c.nviews - 1);  // rows past the launch's views are not read
    const float* dircF = a.c.dirc + (size_t)fc * a.c.inStride;
    const float* conjF = a.c.conj + (size_t)fc * a.c.inStride;
    const float* g = a.c.G + 9 * (size_t)(a.c.view0 + fc);
    const float* pwF = a.pw + (size_t)a.blk[vbase + fc] * V * W2;
#pragma unroll
    for (int j = 0; j < 16 / kRowStep; ++j) {
        float e = 0.0f;
        if (rel >= 0 && rel < rows && k < W2 + 3) {
            e = direct ? dircF[(size_t)r * W + (kk - dOff)] : conj_value_at(a.c, dircF, conjF, g, r, kk - cOff);
            if (k < W2) e = e * pwF[(size_t)r * W2 + k];
        }
        tile[((t >> 7) + kRowStep * j) * kCompOpPitch + kc] = e;
        if (rel >= 0) {
            r += kRowStep;
            if (r >= V) {
                while (r >= V) { r -= V; ++f; }
                if (f < a.c.nviews) {
                    dircF = a.c.dirc + (size_t)f * a.c.inStride;
                    conjF = a.c.conj + (size_t)f * a.c.inStride;
                    g = a.c.G + 9 * (size_t)(a.c.view0 + f);
                    pwF = a.pw + (size_t)a.blk[vbase + f] * V * W2;
                }
            }
        }
        rel += kRowStep;
    }
    __syncthreads();
    if (a.rowmax) {
        // the k tile's largest magnitude of each of the launch's rows: thread
        // t scans 16 columns of row t % 16, then the 8 threads of a row meet
        static_assert(kCompOpThreads == 128, "two waves of 16 rows x 8 column groups");
        __shared__ float wmax[2][16];
        const int i = t & 15, c0 = (t >> 4) * 16;
        float m = 0.0f;
#pragma unroll
        for (int c = 0; c < 16; ++c) m = fmaxf(m, fabsf(tile[i * kCompOpPitch + c0 + c]));
        m = fmaxf(m, __shfl_xor(m, 16));
        m = fmaxf(m, __shfl_xor(m, 32));
        if ((t & 63) < 16) wmax[t >> 6][i] = m;
        __syncthreads();
        const int Rl = rb * 16 + i;
        if (t < 16 && Rl >= a.R0 && Rl < a.R1)
            reinterpret_cast<float*>(a.out + (size_t)rb * a.bstride + (size_t)a.nsteps * 128)[blockIdx.x * 16 + i] =
                fmaxf(wmax[0][i], wmax[1][i]);
    }
    // write phase: the 512 float4s of the block (4 k steps x 2 halves x 64
    // lanes of the MFMA operand order), 512 / kCompOpThreads per thread
#pragma unroll
    for (int i = 0; i < 512 / kCompOpThreads; ++i) {
        const int idx = t + i * kCompOpThreads;
        const int sl = idx >> 7, h = (idx >> 6) & 1, l = idx & 63;
        const int s = blockIdx.x * 4 + sl;
        const int Rl = rb * 16 + (l & 15);
        if (s >= a.nsteps || Rl < a.R0 || Rl >= a.R1) continue;
        const float* src = tile + (l & 15) * kCompOpPitch + 32 * sl + 8 * (l >> 4) + 4 * h;
        a.out[(size_t)rb * a.bstride + ((size_t)s * 2 + h) * 64 + l] = make_float4(src[0], src[1], src[2], src[3]);
    }
}
#pragma clang fp contract(on)

unsigned grid_1d(size_t n)
{
    return (unsigned)std::min<size_t>((n + 255) / 256, (size_t)65536);
}

}  // namespace

}  // namespace jh

using namespace jh;

extern "C" {

int jh_detector_on_right(const double* ptm0, int numUsedDetsU)
{
    if (!ptm0) return 0;
    // JEngine.cpp:153-161, in the reference's float arithmetic
    const float originU = (float)ptm0[3] / (float)ptm0[11];
    return originU < (float)numUsedDetsU / 2.0f ? 1 : 0;
}

jh_status jh_composite_create(jh_composite_plan** plan, int device, int W, int V, int nviews, const double* h_ptm,
                              float dso, int detectorOnRight)
{
    return guard([&] {
        require(plan != nullptr, "jh_composite_create: plan is NULL");
        *plan = nullptr;
        require(W >= kTransWidth && W >= kEdgeWidth && V >= 2, "jh_composite_create: frame narrower than 40 columns");
        require(nviews >= 2 && nviews % 2 == 0, "jh_composite_create: the view count must be even (conjugate pairs)");
        require(h_ptm != nullptr, "jh_composite_create: no projection matrices");
        require(dso > 1.0f, "jh_composite_create: bad DSO");
        require((size_t)nviews * V * 2 * W < ((size_t)1 << 40), "jh_composite_create: scan too large");
        int ndev = 0;
        JH_HIP(hipGetDeviceCount(&ndev));
        require(device >= 0 && device < ndev, "jh_composite_create: no such device");
        DeviceScope ds(device);
        auto* p = new jh_composite_plan;
        try {
            p->device = device;
            p->W = W; p->V = V; p->N = nviews;
            p->detRight = detectorOnRight < 0 ? jh_detector_on_right(h_ptm, W) : (detectorOnRight ? 1 : 0);
            const double pi = 3.14159265358979323846;
            for (int i = 0; i < kTransWidth; ++i) {
                const double c = std::cos((double)i / kTransWidth * pi);
                const double md = p->detRight ? 0.5 - 0.5 * c : 0.5 + 0.5 * c;
                p->maskDirc[i] = (float)md;
                p->maskConj[i] = (float)(1.0 - md);
            }
            DevBuf dptm(sizeof(double) * 12 * nviews);
            JH_HIP(hipMemcpy(dptm.get(), h_ptm, sizeof(double) * 12 * nviews, hipMemcpyHostToDevice));
            p->G.alloc(sizeof(float) * 9 * nviews);
            composite_setup_kernel<<<(nviews + 63) / 64, 64>>>(dptm.as<double>(), nviews, (double)dso, p->G.as<float>());
            check_launch("composite_setup_kernel");
            JH_HIP(hipDeviceSynchronize());
        } catch (...) {
            delete p;
            throw;
        }
        *plan = p;
    });
}

jh_status jh_composite_side(const jh_composite_plan* p, int* detectorOnRight)
{
    return guard([&] {
        require(p != nullptr && detectorOnRight != nullptr, "jh_composite_side: bad arguments");
        *detectorOnRight = p->detRight;
    });
}

jh_status jh_composite_run(jh_composite_plan* p, const float* d_dirc, const float* d_conj, long long inStride,
                           float* d_out, int view0, int nviews, jh_stream_t stream)
{
    return guard([&] {
        require(p != nullptr && d_dirc != nullptr && d_conj != nullptr && d_out != nullptr,
                "jh_composite_run: bad arguments");
        require(view0 >= 0 && nviews >= 0 && view0 + nviews <= p->N, "jh_composite_run: view range outside the scan");
        require(nviews <= 1 || inStride >= (long long)p->W * p->V, "jh_composite_run: input frames overlap");
        if (nviews == 0) return;
        DeviceScope ds(p->device);
        CompArgs a;
        a.dirc = d_dirc;
        a.conj = d_conj;
        a.inStride = inStride;
        a.out = d_out;
        a.G = p->G.as<float>();
        a.W = p->W; a.V = p->V;
        a.view0 = view0;
        a.nviews = nviews;
        a.detRight = p->detRight;
        for (int i = 0; i < kTransWidth; ++i) { a.md[i] = p->maskDirc[i]; a.mc[i] = p->maskConj[i]; }
        const size_t total = (size_t)nviews * p->V * p->W;
        composite_kernel<<<grid_1d(total), 256, 0, as_stream(stream)>>>(a);
        check_launch("composite_kernel");
    });
}

jh_status jh_composite_run_operand(jh_composite_plan* p, jh_filter_plan* fp, const float* d_dirc, const float* d_conj,
                                   long long inStride, int view0, int nviews, float* d_a, int aview0,
                                   jh_stream_t stream)
{
    return guard([&] {
        require(p != nullptr && fp != nullptr && d_dirc != nullptr && d_conj != nullptr && d_a != nullptr,
                "jh_composite_run_operand: bad arguments");
        require(view0 >= 0 && nviews >= 0 && view0 + nviews <= p->N,
                "jh_composite_run_operand: view range outside the scan");
        require(aview0 >= 0 && aview0 <= view0, "jh_composite_run_operand: the operand starts after the views");
        require(nviews <= 1 || inStride >= (long long)p->W * p->V, "jh_composite_run_operand: input frames overlap");
        require(fp->mode != JH_FILTER_FP32_VALU, "jh_composite_run_operand: the vector-ALU filter takes no operand");
        require(fp->W == 2 * p->W && fp->V == p->V && fp->N == p->N && fp->device == p->device,
                "jh_composite_run_operand: the filter does not take this composite (2W columns, same views)");
        if (nviews == 0) return;
        const size_t R0 = (size_t)(view0 - aview0) * p->V, R1 = R0 + (size_t)nviews * p->V;
        require(R1 < (1u << 31), "jh_composite_run_operand: too many rows");
        const size_t b0 = R0 / 16, b1 = (R1 - 1) / 16;
        require(b1 - b0 + 1 <= 65535u, "jh_composite_run_operand: more than 65535 row blocks in one call");
        DeviceScope ds(p->device);
        CompOpArgs a;
        a.c.dirc = d_dirc;
        a.c.conj = d_conj;
        a.c.inStride = inStride;
        a.c.out = nullptr;
        a.c.G = p->G.as<float>();
        a.c.W = p->W; a.c.V = p->V;
        a.c.view0 = view0;
        a.c.nviews = nviews;
        a.c.detRight = p->detRight;
        for (int i = 0; i < kTransWidth; ++i) { a.c.md[i] = p->maskDirc[i]; a.c.mc[i] = p->maskConj[i]; }
        a.out = reinterpret_cast<float4*>(d_a);
        a.pw = fp->pw.as<float>();
        a.blk = fp->blk.as<int>();
        a.R0 = (int)R0; a.R1 = (int)R1;
        a.blockRow0 = (int)b0;
        a.nsteps = fp->Kp / 32;
        a.aview0 = aview0;
        a.edge0 = fp->mirror ? fp->W - 3 : 0;
        a.bstride = fp->bstride;
        a.rowmax = fp->mode == JH_FILTER_MFMA_F16X3 ? 1 : 0;
        const dim3 grid((unsigned)((fp->Kp + 127) / 128), (unsigned)(b1 - b0 + 1));
        composite_operand_kernel<<<grid, kCompOpThreads, 0, as_stream(stream)>>>(a);
        check_launch("composite_operand_kernel");
    });
}

jh_status jh_composite_destroy(jh_composite_plan* p)
{
    return guard([&] {
        if (!p) return;
        DeviceScope ds(p->device);
        delete p;
    });
}

}  // extern "C"
