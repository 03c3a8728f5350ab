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
// it is deterministic (fp_column_kernel): each wave owns one detector column,
// walks the wedge of voxel columns whose samples can fall in it and spreads
// their z samples over its lanes, recomputing su and sv with the
// backprojector's own float expressions (including the z tiling), so
// membership and weights are those of the backprojector.  Other geometries use a scatter with float atomics
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

// Column form of the mode-0 transpose.  A workgroup of nw waves covers the
// detector columns [c0, c0 + nw) of one view; wave w owns column c = c0 + w,
// accumulated in LDS (V + 1 floats: row V takes the zero-weight share of
// samples at sv = V - 1 and is never written out).  The wave walks the wedge
// of voxel columns whose su can lie in [c - 1, c + 1), eight candidates on
// each of eight major-axis lines per round, and keeps those that the
// backprojector's own su puts at floor(su) = c (weight 1 - fu on column c) or
// c - 1 (weight fu).  Each such member's samples are then spread over the
// lanes, 64 consecutive z per step: lane k adds x_k (1 - fv_k) w_u to row
// floor(sv_k) and x_k fv_k w_u to the next row, in two phases (every lane's
// first row, then every lane's second row), with sv_k the backprojector's
// expression (z tiling included).  The rows of one phase are distinct when
// the sample spacing |dsv| is at least 1 + 2^-8 rows (the margin covers the
// rounding of the per-tile expressions); below that the lanes take turns, in
// ceil((1 + 2^-8) / |dsv|) passes.  The adds are plain LDS reads and writes
// (LDS float atomics run lane by lane: 22x slower here).  A wave owns its
// column, so nothing else touches it and the summation order is fixed.
constexpr float kFpRowMargin = 1.00390625f;

__device__ __forceinline__ void lds_order()
{
    // the two phases touch different rows within a lane, so only this
    // barrier keeps the compiler from moving one phase's LDS accesses across
    // the other's; the LDS executes one wave's accesses in order
    asm volatile("" ::: "memory");
}

// A member of a wave's column: voxel column j X + i, the weight wu its su
// puts on the column, and its 1 / w and V coefficient at z = 0.
struct FpMember {
    int col;
    float wu, rw, vc0;
};

__device__ __forceinline__ FpMember fp_pick(int cix, float wu, float rw, float vc0, int l)
{
    FpMember m;
    m.col = __builtin_amdgcn_readlane(cix, l);
    m.wu = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wu), l));
    m.rw = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(rw), l));
    m.vc0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(vc0), l));
    return m;
}

// The member's samples at the per-lane byte offsets koff (clamped into the
// column, so every lane loads and no load needs a branch; 32-bit unsigned
// byte offsets, so the load takes the column address as a scalar base).
template <int NC>
__device__ __forceinline__ void fp_load(const float* __restrict__ xt, int Z, const FpMember& m,
                                        const unsigned* koff, float* xv)
{
    const char* xc = (const char*)(xt + (size_t)m.col * Z);
#pragma unroll
    for (int i = 0; i < NC; ++i) xv[i] = *(const float*)(xc + koff[i]);
}

// The same through a buffer resource over the whole volume (< 4 GiB): the
// column's byte offset rides in the scalar offset, the lane's in the vector
// offset, so a member costs no address arithmetic.  The offsets stay inside
// the volume, which is also the resource's range.
template <int NC>
__device__ __forceinline__ void fp_load_buf(__amdgpu_buffer_rsrc_t xr, unsigned soff, const unsigned* koff,
                                            float* xv)
{
#pragma unroll
    for (int i = 0; i < NC; ++i)
        xv[i] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(xr, koff[i], soff, 0));
}

// Adds NC chunks of 64 samples of one member to the wave's column.  fkt /
// fkk: each lane's sample z tile start and offset in it, as floats; kval:
// sample inside the volume; c7v: the V coefficient of z, in a vector
// register (the other operands of its FMA are scalar).  FORM 0: every sample
// of the member lies at least a row inside [0, V - 1] and inside the volume,
// so no lane is masked; FORM 1: samples tested; FORM 2: as 1, the lanes
// taking turns in `passes` passes (sample spacing below a row).
//
// In forms 0 and 1 the spacing is at least 1 + 2^-8 rows, so the first rows
// floor(sv_k) of all the member's samples are distinct, and so are the second
// rows: the chunks' first-row reads issue together, then their writes, then
// the same for the second rows -- two LDS round trips per member.
template <int NC, int FORM>
__device__ __forceinline__ void fp_apply(float* col, const float* xv, const float* fkt, const float* fkk,
                                         const bool* kval, float c7v, const FpMember& m, float dsv, float Vm1,
                                         int lane, int passes)
{
    int iv[NC];
    float lo[NC], hi[NC];
    bool ok[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        const float Vc = fmaf(c7v, fkt[i], m.vc0);
        const float sv0 = fmaf(Vc, m.rw, 0.5f);
        const float sv = fmaf(fkk[i], dsv, sv0);
        const float tt = FORM == 0 ? sv : __builtin_amdgcn_fmed3f(sv, 0.0f, Vm1);
        ok[i] = FORM == 0 || (kval[i] && tt == sv);
        iv[i] = (int)tt;
        const float fv = __builtin_amdgcn_fractf(tt);
        const float xw = m.wu * xv[i];
        lo[i] = fmaf(-fv, xw, xw);
        hi[i] = fv * xw;
    }
    if constexpr (FORM < 2) {
        float o[NC];
#pragma unroll
        for (int i = 0; i < NC; ++i)
            if (ok[i]) o[i] = col[iv[i]];
#pragma unroll
        for (int i = 0; i < NC; ++i)
            if (ok[i]) col[iv[i]] = o[i] + lo[i];
        lds_order();
#pragma unroll
        for (int i = 0; i < NC; ++i)
            if (ok[i]) o[i] = col[iv[i] + 1];
#pragma unroll
        for (int i = 0; i < NC; ++i)
            if (ok[i]) col[iv[i] + 1] = o[i] + hi[i];
        lds_order();
    } else {
        // one instruction stream for all lanes: a pass's first-row updates
        // are written before any lane reads its second row
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            for (int p = 0; p < passes; ++p) {
                if (ok[i] && lane % passes == p) col[iv[i]] += lo[i];
                lds_order();
            }
            for (int p = 0; p < passes; ++p) {
                if (ok[i] && lane % passes == p) col[iv[i] + 1] += hi[i];
                lds_order();
            }
        }
    }
}

// single: |dsv| >= 1 + 2^-8; interior: FORM 0 applies (both wave-uniform,
// from the candidate round's ballots).
template <int NC>
__device__ __forceinline__ void fp_member(float* col, const float* xv, const float* fkt, const float* fkk,
                                          const bool* kval, float c7v, const FpMember& m, float Vm1, int lane,
                                          bool single, bool interior)
{
    const float dsv = c7v * m.rw;
    if (single && interior) {
        fp_apply<NC, 0>(col, xv, fkt, fkk, kval, c7v, m, dsv, Vm1, lane, 1);
    } else if (single) {
        fp_apply<NC, 1>(col, xv, fkt, fkk, kval, c7v, m, dsv, Vm1, lane, 1);
    } else {
        const float adsv = fabsf(dsv);
        const int passes = __builtin_amdgcn_readfirstlane(
            adsv > 0.0f ? min(64, (int)ceilf(kFpRowMargin / adsv)) : 64);
        fp_apply<NC, 2>(col, xv, fkt, fkk, kval, c7v, m, dsv, Vm1, lane, passes);
    }
}

// NI > 0: Z <= 64 NI and a volume below 4 GiB, the per-lane sample
// constants held in registers and the next member's samples loaded while the
// current one is added; NI == 0: any size, in chunks of 4 x 64 samples.
template <int NI>
__global__ void __launch_bounds__(512) fp_column_kernel(FPArgs a)
{
    extern __shared__ float colbuf[];
    const int lane = threadIdx.x, nw = blockDim.y;
    const int v = blockIdx.y;
    const int c0 = blockIdx.x * nw;
    const int c = c0 + (int)threadIdx.y;
    const int rows = a.V + 1;
    float* const col = colbuf + threadIdx.y * rows;
    for (int r = lane; r < rows; r += 64) col[r] = 0.0f;
    const Geo g = geo_of(a.geo, a.view0 + v);
    float c7v;
    asm("v_mov_b32 %0, %1" : "=v"(c7v) : "s"(g.c[7]));
    const float Wm1 = (float)(a.W - 1), Vm1 = (float)(a.V - 1);
    const int zmask = ~(a.zc - 1);  // the backprojector's z tile start (zc is a power of two)
    constexpr int NR = NI > 0 ? NI : 4;
    float fkt[NR], fkk[NR];
    bool kval[NR];
    unsigned koff[NR];
    if constexpr (NI > 0) {
#pragma unroll
        for (int i = 0; i < NI; ++i) {
            const int k = 64 * i + lane, kt = k & zmask;
            fkt[i] = (float)kt;
            fkk[i] = (float)(k - kt);
            kval[i] = k < a.Z;
            koff[i] = 4u * (unsigned)min(k, a.Z - 1);
        }
    }
    // raw buffer over the z-fastest volume; dword 3: 32-bit data format
    const __amdgpu_buffer_rsrc_t xr = __builtin_amdgcn_make_buffer_rsrc(
        (void*)a.xt, (short)0, (int)(unsigned)((size_t)a.X * a.Y * a.Z * sizeof(float)), 0x00020000);
    const unsigned zbytes = 4u * (unsigned)a.Z;
    // the last sample's z tile start and offset; full: every lane of every
    // chunk holds a sample
    const int klast = a.Z - 1;
    const float ftlast = (float)(klast & zmask), fklast = (float)(klast - (klast & zmask));
    const bool full = NI > 0 && a.Z == 64 * NI;
    if (c < a.W) {
        // wedge boundaries (c1 - t c9) i + (c2 - t c10) j + (c0 - t c8) = 0
        // at su = t + 0.5 for t = c - 1.5 and t = c + 0.5
        const double t0 = (double)c - 1.5, t1 = (double)c + 0.5;
        const double ai0 = g.c[1] - t0 * g.c[9], aj0 = g.c[2] - t0 * g.c[10], a00 = g.c[0] - t0 * g.c[8];
        const double ai1 = g.c[1] - t1 * g.c[9], aj1 = g.c[2] - t1 * g.c[10], a01 = g.c[0] - t1 * g.c[8];
        const double tc = (double)c - 0.5;
        const bool major_i = fabs(g.c[2] - tc * g.c[10]) >= fabs(g.c[1] - tc * g.c[9]);
        const int nmaj = major_i ? a.X : a.Y, nmin = major_i ? a.Y : a.X;
        // minor-axis bound on line b: -(a_maj m + a_0) / a_min, linear in m
        const double am0 = major_i ? aj0 : ai0, ak0 = major_i ? ai0 : aj0;
        const double am1 = major_i ? aj1 : ai1, ak1 = major_i ? ai1 : aj1;
        const double s0 = am0 != 0.0 ? -ak0 / am0 : 0.0, q0 = am0 != 0.0 ? -a00 / am0 : -1e30;
        const double s1 = am1 != 0.0 ? -ak1 / am1 : 0.0, q1 = am1 != 0.0 ? -a01 / am1 : 1e30;
        const int lm = lane >> 3, ls = lane & 7;
        for (int mb = 0; mb < nmaj; mb += 8) {
            const int m = mb + lm;
            int n0 = 1, n1 = 0;
            if (m < nmaj) {
                const double b0 = am0 != 0.0 ? fma((double)m, s0, q0) : -1e30;
                const double b1 = am1 != 0.0 ? fma((double)m, s1, q1) : 1e30;
                // clamped to [0, nmin] and [-1, nmin - 1], so n0 + ls + t
                // cannot overflow however steep the lines
                n0 = (int)fmax(0.0, fmin((double)nmin, floor(fmin(b0, b1)) - 2.0));
                n1 = (int)fmax(-1.0, fmin((double)(nmin - 1), ceil(fmax(b0, b1)) + 2.0));
            }
            for (int t = 0;; t += 8) {
                const int n = n0 + ls + t;
                bool mem = n <= n1;
                if (__ballot(mem) == 0) break;
                float wu = 0.0f, rw = 0.0f, vc0 = 0.0f;
                int cix = 0;
                bool single = false, interior = false;
                if (mem) {
                    const int i = major_i ? m : n, j = major_i ? n : m;
                    const float fi = (float)i, fj = (float)j;
                    const float su = bp_su(g, fi, fj, rw);
                    const bool in = su >= 0.0f && su <= Wm1;
                    const int iu0 = in ? (int)su : -2;
                    const float fu = __builtin_amdgcn_fractf(su);
                    mem = iu0 == c || iu0 == c - 1;
                    wu = iu0 == c ? 1.0f - fu : fu;
                    vc0 = fmaf(g.c[5], fi, fmaf(g.c[6], fj, g.c[4]));
                    cix = j * a.X + i;
                    const float dsv = c7v * rw;
                    single = fabsf(dsv) >= kFpRowMargin;
                    // the first and last samples' sv (the samples' own
                    // expressions), a row inside the detector: sv is linear
                    // in k up to rounding far below a row
                    const float svf = fmaf(vc0, rw, 0.5f);
                    const float svl = fmaf(fklast, dsv, fmaf(fmaf(c7v, ftlast, vc0), rw, 0.5f));
                    interior = full && fminf(svf, svl) >= 1.0f && fmaxf(svf, svl) <= Vm1 - 1.0f;
                }
                unsigned long long mask = __ballot(mem);
                if (mask == 0) continue;
                const unsigned long long msingle = __ballot(single), minterior = __ballot(interior);
                if constexpr (NI > 0) {
                    // two members in flight: the next one's samples load
                    // while the current one is added.  The buffers
                    // alternate (no copy waits for a load), and the next
                    // load is issued on every path (after the last member it
                    // reloads that member's column), so the wait for the
                    // current samples never counts the next ones.
                    int la = __builtin_ctzll(mask), lb;
                    mask &= mask - 1;
                    FpMember ma = fp_pick(cix, wu, rw, vc0, la), mb;
                    float xa[NI], xb[NI];
                    fp_load_buf<NI>(xr, (unsigned)ma.col * zbytes, koff, xa);
                    for (;;) {
                        bool more = mask != 0;
                        lb = more ? __builtin_ctzll(mask) : la;
                        mask &= mask - 1;
                        mb = fp_pick(cix, wu, rw, vc0, lb);
                        fp_load_buf<NI>(xr, (unsigned)mb.col * zbytes, koff, xb);
                        fp_member<NI>(col, xa, fkt, fkk, kval, c7v, ma, Vm1, lane, (msingle >> la) & 1,
                                      (minterior >> la) & 1);
                        if (!more) break;
                        more = mask != 0;
                        la = more ? __builtin_ctzll(mask) : lb;
                        mask &= mask - 1;
                        ma = fp_pick(cix, wu, rw, vc0, la);
                        fp_load_buf<NI>(xr, (unsigned)ma.col * zbytes, koff, xa);
                        fp_member<NI>(col, xb, fkt, fkk, kval, c7v, mb, Vm1, lane, (msingle >> lb) & 1,
                                      (minterior >> lb) & 1);
                        if (!more) break;
                    }
                } else {
                    while (mask) {
                        const int l = __builtin_ctzll(mask);
                        mask &= mask - 1;
                        const FpMember cur = fp_pick(cix, wu, rw, vc0, l);
                        for (int kb = 0; kb < a.Z; kb += 256) {
#pragma unroll
                            for (int i = 0; i < 4; ++i) {
                                const int k = kb + 64 * i + lane, kt = k & zmask;
                                fkt[i] = (float)kt;
                                fkk[i] = (float)(k - kt);
                                kval[i] = k < a.Z;
                                koff[i] = 4u * (unsigned)min(k, a.Z - 1);
                            }
                            float xa[4];
                            fp_load<4>(a.xt, a.Z, cur, koff, xa);
                            fp_member<4>(col, xa, fkt, fkk, kval, c7v, cur, Vm1, lane, (msingle >> l) & 1, false);
                        }
                    }
                }
            }
        }
    }
    __syncthreads();
    // rows of the nw columns, column fastest
    const int lg = nw == 8 ? 3 : nw == 4 ? 2 : nw == 2 ? 1 : 0;
    const int nthreads = 64 * nw;
    const float* uwv = a.uw ? a.uw + (size_t)(a.view0 + v) * a.W : nullptr;
    for (int e = lane + 64 * (int)threadIdx.y; e < (a.V << lg); e += nthreads) {
        const int r = e >> lg, cc = e & (nw - 1), u = c0 + cc;
        if (u < a.W) a.y[((size_t)v * a.V + r) * a.W + u] = colbuf[cc * rows + r] * (uwv ? uwv[u] : 1.0f);
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
            // nw waves (detector columns) per workgroup, each with V + 1 floats of
            // LDS.  Four: a workgroup's waves wait at the write-out for its
            // slowest column (measured at 256^3 from 360 views of 512^2: 8 waves
            // 6.41 ms, 4 waves 6.00, 2 waves 6.15)
            int nw = 4;
            while (nw > 1 && (size_t)nw * (p->V + 1) * sizeof(float) > 65536) nw /= 2;
            const size_t lds = (size_t)nw * (p->V + 1) * sizeof(float);
            require(lds <= 65536, "jh_fp_project: detector too tall for the column projector (V <= 16383)");
            const dim3 grid((p->W + nw - 1) / nw, nviews), block(64, nw);
            const bool small = (size_t)p->X * p->Y * p->Z * sizeof(float) < ((size_t)1 << 32);
            if (!small) fp_column_kernel<0><<<grid, block, lds, s>>>(a);
            else if (p->Z <= 64) fp_column_kernel<1><<<grid, block, lds, s>>>(a);
            else if (p->Z <= 128) fp_column_kernel<2><<<grid, block, lds, s>>>(a);
            else if (p->Z <= 256) fp_column_kernel<4><<<grid, block, lds, s>>>(a);
            else if (p->Z <= 512) fp_column_kernel<8><<<grid, block, lds, s>>>(a);
            else fp_column_kernel<0><<<grid, block, lds, s>>>(a);
            check_launch("fp_column_kernel");
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
