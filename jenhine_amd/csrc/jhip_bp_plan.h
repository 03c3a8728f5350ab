// Backprojection plan shared by the backprojector and its adjoint.
#pragma once

#include "jhip_internal.h"

namespace jh {

// Per-view geometry in voxel-index space, rounded once from FP64:
// U(i,j,k) = c[0] + c[1] i + c[2] j + c[3] k, V likewise with c[4..7],
// W with c[8..11].
struct Geo {
    float c[12];
};

// A view's geometry through the constant address space: scalar loads, which
// the compiler otherwise makes only where it can prove that no store of the
// kernel reaches the array.  A vector load in the view loop would make each
// view wait for every gather or copy in flight.
__device__ __forceinline__ Geo geo_of(const Geo* geo, int view)
{
    const __attribute__((address_space(4))) float* p = (const __attribute__((address_space(4))) float*)(geo + view);
    Geo g;
#pragma unroll
    for (int k = 0; k < 12; ++k) g.c[k] = p[k];
    return g;
}

}  // namespace jh

struct jh_bp_plan {
    int device;
    int N, W, V, X, Y, Z;
    float pxy, pz;
    int mode;        // 0: u and w constant along z, 1: w constant along z, 2: general
    int zc, ty;      // backprojector tile: z voxels per thread, waves per workgroup
    int fetch;       // JH_FETCH_GATHER or JH_FETCH_LDS (geometry class 0 only)
    int strip;       // BF16/FP16 staged as four-row strips (JH_LAYOUT_STRIP) or quads
    int zfilter;     // staging applies the 1-2-1 v smoothing first
    int ct;          // normalised output goes through the CT-number map
    float ct_scale, ct_norm1;
    int weight;      // JH_WEIGHT_COUNT (the reference's normalisation) or JH_WEIGHT_FDK
    float wscale;    // JH_WEIGHT_FDK: vol = wscale * sum of samples / w^2
    int f16exp;      // FP16 storage holds q * 2^-f16exp; outputs are scaled back by 2^f16exp
    float acc_scale; // output scale of what the accumulators hold (2^f16exp of FP16 chunks, else 1)
    int acc_dtype;   // storage type of the chunks accumulated since the last reset (-1: none)
    int sl_zc;       // LDS-staged strips: the zc the rectangle table (sbox) was built for (-1: none)
    int sl_stride;   // its chunks per rectangle row (odd; 0: the shape does not fit the LDS)
    int sl_rows;     // its strip rows per rectangle
    jh::DevBuf ptm;  // N x 12 double
    jh::DevBuf geo;  // N x Geo
    jh::DevBuf uw;   // N x W float
    jh::DevBuf uwp;  // N x W float2 count-weight pairs
    jh::DevBuf colcnt;  // geometry class 0: X x Y count weights of a launch's views
    jh::DevBuf acc;  // X*Y*Z float (streaming form, allocated on first use)
    jh::DevBuf cnt;  // X*Y*Z float: count (geometry classes 1, 2) or outside-row weight (class 0)
    jh::DevBuf colsum;  // geometry class 0, streaming form: X x Y running count weights
    jh::DevBuf xt;   // X*Y*Z float, z-fastest copy of the volume for the projector
    jh::DevBuf sbox; // LDS-staged strips: rectangle origins per tile and view (ntiles x N uint32)
};
