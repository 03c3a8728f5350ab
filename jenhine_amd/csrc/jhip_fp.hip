// Forward projector: the exact transpose of the backprojection operator.
//
// The backprojector (jhip_bp.hip) applies, per view, the bilinear sampling
// operator B_v: acc(voxel) += bilinear(q_v, su, sv), with su, sv computed in
// float from the voxel-index coefficients of the view, and a sample outside
// [0, W-1] x [0, V-1] contributing nothing.  This file applies B_v^T:
//   y_v(iu, iv) = sum over voxels of x(voxel) * weight of pixel (iu, iv) in
//                 that voxel's bilinear sample,
// optionally multiplied by uW_v(iu) (the transpose of the weighted staging).
//
// For geometries whose u and w do not depend on z (mode 0, circular orbits)
// it is a deterministic gather (fp_block_kernel): a workgroup covers a block
// of detector columns and walks the wedge of voxel columns whose samples can
// fall in them, recomputing su and sv with the backprojector's own float
// expressions (including the z tiling), so membership and weights are those
// of the backprojector.  Other geometries use a scatter with float atomics
// (same weights, summation order not fixed).
#include <cmath>

#include "jhip_bp_plan.h"

namespace jh {

// vol[z][y][x] -> xt[y][x][z] through LDS tiles of 32 x 32 (x, z).
__global__ void __launch_bounds__(256) fp_transpose_kernel(const float* __restrict__ vol,
                                                           float* __restrict__ xt, int X, int Y, int Z)
{
    __shared__ float tile[32][33];
    const int x0 = blockIdx.x * 32, z0 = blockIdx.y * 32, y = blockIdx.z;
    for (int r = threadIdx.y; r < 32; r += 8) {
        const int z = z0 + r, x = x0 + threadIdx.x;
        if (z < Z && x < X) tile[r][threadIdx.x] = vol[((size_t)z * Y + y) * X + x];
    }
    __syncthreads();
    for (int r = threadIdx.y; r < 32; r += 8) {
        const int x = x0 + r, z = z0 + threadIdx.x;
        if (z < Z && x < X) xt[((size_t)y * X + x) * Z + z] = tile[threadIdx.x][r];
    }
}

struct FPArgs {
    const float* xt;   // z-fastest volume
    const float* vol;  // z-slowest volume (scatter form)
    const Geo* geo;
    const float* uw;   // N x W, or NULL (unweighted)
    float* y;          // nviews x V x W
    int W, V, X, Y, Z;
    int view0, zc;
};

// su of voxel column (i, j) exactly as the backprojector computes it.
__device__ inline float bp_su(const Geo& g, float fi, float fj, float& rw)
{
    const float Uc = fmaf(g.c[3], 0.0f, fmaf(g.c[1], fi, fmaf(g.c[2], fj, g.c[0])));
    const float Wc = fmaf(g.c[11], 0.0f, fmaf(g.c[9], fi, fmaf(g.c[10], fj, g.c[8])));
    rw = __builtin_amdgcn_rcpf(Wc);
    return fmaf(Uc, rw, 0.5f);
}

// Block form of the mode-0 gather: one workgroup per view and block of kFpCols
// detector columns [u0, u0 + kFpCols), rows in passes of 64 x blockDim.y.
// The workgroup walks the wedge of voxel columns whose su lies in
// [u0 - 1, u0 + kFpCols) once (each voxel column feeds detector columns
// floor(su) and floor(su) + 1); every lane owns one detector row and adds,
// per voxel column, the column's vertical profile at its row -- the samples
// k whose rows floor(sv_k), floor(sv_k) + 1 include it, with the
// backprojector's own sample expressions (z tiling included) -- times
// 1 - fu and fu into its LDS row.  Lanes own their rows, so the LDS needs no
// atomics and the summation order is fixed.
constexpr int kFpCols = 16;

__global__ void __launch_bounds__(512) fp_block_kernel(FPArgs a)
{
    __shared__ float accl[512 * (kFpCols + 1)];  // row r of the pass: accl[r * 17 + c], c < kFpCols
    const int v = blockIdx.y;
    const int u0 = blockIdx.x * kFpCols;
    const int lane_row = threadIdx.y * 64 + threadIdx.x;
    const int rows_per_pass = 64 * blockDim.y;
    const Geo g = a.geo[a.view0 + v];
    const float Wm1 = (float)(a.W - 1), Vm1 = (float)(a.V - 1);
    const int zmask = ~(a.zc - 1);  // the backprojector's z tile start (zc is a power of two)
    // wedge boundaries (c1 - t c9) i + (c2 - t c10) j + (c0 - t c8) = 0 at
    // su = t + 0.5 for t = u0 - 1.5 and t = u0 + kFpCols - 0.5
    const double t0 = (double)u0 - 1.5, t1 = (double)(u0 + kFpCols) - 0.5;
    const double ai0 = g.c[1] - t0 * g.c[9], aj0 = g.c[2] - t0 * g.c[10], a00 = g.c[0] - t0 * g.c[8];
    const double ai1 = g.c[1] - t1 * g.c[9], aj1 = g.c[2] - t1 * g.c[10], a01 = g.c[0] - t1 * g.c[8];
    const double tc = (double)u0 + 0.5 * kFpCols - 0.5;
    const bool major_i = fabs(g.c[2] - tc * g.c[10]) >= fabs(g.c[1] - tc * g.c[9]);
    const int nmaj = major_i ? a.X : a.Y, nmin = major_i ? a.Y : a.X;
    // minor-axis bound on line b: -(a_maj m + a_0) / a_min, linear in m
    const double am0 = major_i ? aj0 : ai0, ak0 = major_i ? ai0 : aj0;
    const double am1 = major_i ? aj1 : ai1, ak1 = major_i ? ai1 : aj1;
    const double s0 = am0 != 0.0 ? -ak0 / am0 : 0.0, c0 = am0 != 0.0 ? -a00 / am0 : -1e30;
    const double s1 = am1 != 0.0 ? -ak1 / am1 : 0.0, c1 = am1 != 0.0 ? -a01 / am1 : 1e30;
    for (int rp = 0; rp < a.V; rp += rows_per_pass) {
        const int iv = rp + lane_row;
        float* const mine = accl + lane_row * (kFpCols + 1);
#pragma unroll
        for (int c = 0; c < kFpCols; ++c) mine[c] = 0.0f;
        const float fiv = (float)iv;
        for (int m = 0; m < nmaj; ++m) {
            const double b0 = am0 != 0.0 ? fma((double)m, s0, c0) : -1e30;
            const double b1 = am1 != 0.0 ? fma((double)m, s1, c1) : 1e30;
            const int n0 = (int)fmax(0.0, floor(fmin(b0, b1)) - 2.0);
            const int n1 = (int)fmin((double)(nmin - 1), ceil(fmax(b0, b1)) + 2.0);
            for (int n = n0; n <= n1; ++n) {
                const int i = major_i ? m : n, j = major_i ? n : m;
                const float fi = (float)i, fj = (float)j;
                float rw;
                const float su = bp_su(g, fi, fj, rw);
                if (!(su >= 0.0f && su <= Wm1)) continue;
                const int iu0 = (int)su;
                if (iu0 < u0 - 1 || iu0 >= u0 + kFpCols) continue;
                const float fu = __builtin_amdgcn_fractf(su);
                const float dsv = g.c[7] * rw;
                const float Vc0 = fmaf(g.c[5], fi, fmaf(g.c[6], fj, g.c[4]));
                // samples k whose rows floor(sv_k) .. floor(sv_k) + 1 may
                // include this lane's row: sv_k in [iv - 1, iv + 1), with a
                // margin of one sample against the rounding of sv
                const float svA = fmaf(Vc0, rw, 0.5f);
                int k0 = 0, k1 = a.Z - 1;
                if (fabsf(dsv) > 1e-6f) {
                    // k with sv_k in [iv - 1, iv + 1): sv_k = svA + k dsv up to
                    // the rounding of the backprojector's per-tile expressions
                    // (a few 1e-6 of a row on these sizes); the margin of
                    // 1/64 sample also absorbs the reciprocal's ulp
                    const float r = __builtin_amdgcn_rcpf(dsv);
                    const float ka = (fiv - 1.0f - svA) * r, kb = (fiv + 1.0f - svA) * r;
                    k0 = (int)fmaxf(0.0f, ceilf(fminf(ka, kb) - 0.015625f));
                    k1 = (int)fminf((float)(a.Z - 1), floorf(fmaxf(ka, kb) + 0.015625f));
                }
                if (iv >= a.V) k1 = -1;
                const float* xc = a.xt + ((size_t)j * a.X + i) * a.Z;
                float prof = 0.0f;
                // candidates in pairs: the loads of a pair (at addresses
                // clamped into the column) are in flight together, then the
                // exact membership test picks the samples that count
                for (int kb = k0; kb <= k1; kb += 2) {
                    float xv[2];
#pragma unroll
                    for (int q = 0; q < 2; ++q) xv[q] = xc[min(kb + q, a.Z - 1)];
#pragma unroll
                    for (int q = 0; q < 2; ++q) {
                        const int k = kb + q;
                        const int kt = k & zmask;
                        const float Vc = fmaf(g.c[7], (float)kt, Vc0);
                        const float sv0 = fmaf(Vc, rw, 0.5f);
                        const float sv = fmaf((float)(k - kt), dsv, sv0);
                        const float tt = __builtin_amdgcn_fmed3f(sv, 0.0f, Vm1);
                        const int ivk = (int)tt;
                        const float fv = __builtin_amdgcn_fractf(tt);
                        const bool take = k <= k1 && tt == sv && (ivk == iv || ivk == iv - 1);
                        const float wv = ivk == iv ? 1.0f - fv : fv;
                        prof = take ? fmaf(wv, xv[q], prof) : prof;
                    }
                }
                // detector columns iu0 (weight 1 - fu) and iu0 + 1 (fu) of the block
                if (iu0 >= u0) mine[iu0 - u0] = fmaf(1.0f - fu, prof, mine[iu0 - u0]);
                if (iu0 + 1 < u0 + kFpCols) mine[iu0 + 1 - u0] = fmaf(fu, prof, mine[iu0 + 1 - u0]);
            }
        }
        if (iv < a.V) {
            float* yr = a.y + ((size_t)v * a.V + iv) * a.W;
#pragma unroll
            for (int c = 0; c < kFpCols; ++c) {
                const int u = u0 + c;
                if (u < a.W) yr[u] = mine[c] * (a.uw ? a.uw[(size_t)(a.view0 + v) * a.W + u] : 1.0f);
            }
        }
    }
}

// Scatter form for any geometry: one thread per voxel column and z tile, the
// backprojector's sample expressions, float atomics into y.
__global__ void __launch_bounds__(256) fp_scatter_kernel(FPArgs a, int nviews)
{
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k0 = blockIdx.z * a.zc;
    if (i >= a.X || j >= a.Y) return;
    const float fi = (float)i, fj = (float)j, fk0 = (float)k0;
    const float Wm1 = (float)(a.W - 1), Vm1 = (float)(a.V - 1);
    const int kn = min(a.zc, a.Z - k0);
    for (int v = 0; v < nviews; ++v) {
        const Geo g = a.geo[a.view0 + v];
        float* yv = a.y + (size_t)v * a.V * a.W;
        const float* uwv = a.uw ? a.uw + (size_t)(a.view0 + v) * a.W : nullptr;
        const float Uc = fmaf(g.c[3], fk0, fmaf(g.c[1], fi, fmaf(g.c[2], fj, g.c[0])));
        const float Vc = fmaf(g.c[7], fk0, fmaf(g.c[5], fi, fmaf(g.c[6], fj, g.c[4])));
        const float Wc = fmaf(g.c[11], fk0, fmaf(g.c[9], fi, fmaf(g.c[10], fj, g.c[8])));
        const float rw = __builtin_amdgcn_rcpf(Wc);
        const float su0 = fmaf(Uc, rw, 0.5f), dsu = g.c[3] * rw;
        const float sv0 = fmaf(Vc, rw, 0.5f), dsv = g.c[7] * rw;
        for (int kk = 0; kk < kn; ++kk) {
            const float fk = (float)kk;
            float su, sv;
            if (g.c[11] == 0.0f) {
                su = fmaf(fk, dsu, su0);
                sv = fmaf(fk, dsv, sv0);
            } else {
                const float r = __builtin_amdgcn_rcpf(fmaf(fk, g.c[11], Wc));
                su = fmaf(fmaf(fk, g.c[3], Uc), r, 0.5f);
                sv = fmaf(fmaf(fk, g.c[7], Vc), r, 0.5f);
            }
            const float tu = __builtin_amdgcn_fmed3f(su, 0.0f, Wm1);
            const float tv = __builtin_amdgcn_fmed3f(sv, 0.0f, Vm1);
            if (tu != su || tv != sv) continue;
            const int iu = (int)tu, iv = (int)tv;
            const float fu = __builtin_amdgcn_fractf(tu), fv = __builtin_amdgcn_fractf(tv);
            const float x = a.vol[((size_t)(k0 + kk) * a.Y + j) * a.X + i];
            const float w00 = (1.0f - fu) * (1.0f - fv), w01 = fu * (1.0f - fv);
            const float w10 = (1.0f - fu) * fv, w11 = fu * fv;
            const float s0 = uwv ? uwv[iu] : 1.0f;
            atomicAdd(&yv[(size_t)iv * a.W + iu], w00 * x * s0);
            if (iu + 1 < a.W) atomicAdd(&yv[(size_t)iv * a.W + iu + 1], w01 * x * (uwv ? uwv[iu + 1] : 1.0f));
            if (iv + 1 < a.V) {
                atomicAdd(&yv[(size_t)(iv + 1) * a.W + iu], w10 * x * s0);
                if (iu + 1 < a.W)
                    atomicAdd(&yv[(size_t)(iv + 1) * a.W + iu + 1], w11 * x * (uwv ? uwv[iu + 1] : 1.0f));
            }
        }
    }
}

__global__ void fp_zero_kernel(float* __restrict__ p, size_t n)
{
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
        p[t] = 0.0f;
}

}  // namespace jh

using namespace jh;

extern "C" {

jh_status jh_fp_project(jh_bp_plan* p, const float* d_vol, float* d_proj, int view0, int nviews,
                        int weighted, jh_stream_t stream)
{
    return guard([&] {
        require(p != nullptr && d_vol != nullptr && d_proj != nullptr, "jh_fp_project: bad arguments");
        require(view0 >= 0 && nviews >= 0 && view0 + nviews <= p->N, "jh_fp_project: view range outside the scan");
        require(nviews <= 65535, "jh_fp_project: at most 65535 views per call");
        if (nviews == 0) return;
        DeviceScope ds(p->device);
        hipStream_t s = as_stream(stream);
        FPArgs a;
        a.vol = d_vol;
        a.geo = p->geo.as<Geo>();
        a.uw = weighted ? p->uw.as<float>() : nullptr;
        a.y = d_proj;
        a.W = p->W; a.V = p->V; a.X = p->X; a.Y = p->Y; a.Z = p->Z;
        a.view0 = view0;
        a.zc = p->zc;
        if (p->mode == 0) {
            if (!p->xt.get()) p->xt.alloc(sizeof(float) * (size_t)p->X * p->Y * p->Z);
            a.xt = p->xt.as<float>();
            require(p->Y <= 65535, "jh_fp_project: Y too large");
            fp_transpose_kernel<<<dim3((p->X + 31) / 32, (p->Z + 31) / 32, p->Y), dim3(32, 8), 0, s>>>(
                d_vol, p->xt.as<float>(), p->X, p->Y, p->Z);
            check_launch("fp_transpose_kernel");
            // rows in passes of 64 x 8
            fp_block_kernel<<<dim3((p->W + kFpCols - 1) / kFpCols, nviews), dim3(64, 8), 0, s>>>(a);
            check_launch("fp_block_kernel");
        } else {
            a.xt = nullptr;
            const size_t n = (size_t)nviews * p->V * p->W;
            fp_zero_kernel<<<(unsigned)std::min<size_t>((n + 255) / 256, 65536), 256, 0, s>>>(d_proj, n);
            check_launch("fp_zero_kernel");
            fp_scatter_kernel<<<dim3((p->X + 63) / 64, (p->Y + 3) / 4, (p->Z + p->zc - 1) / p->zc), dim3(64, 4), 0,
                                s>>>(a, nviews);
            check_launch("fp_scatter_kernel");
        }
    });
}

}  // extern "C"
