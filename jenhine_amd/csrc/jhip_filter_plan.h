// Row-filter plan shared by the filter and the kernels that write its
// pre-weighted operand.
#pragma once

#include <vector>

#include "jhip_internal.h"

struct jh_filter_plan {
    int device;
    int W, V, N, L, K, Kp;  // W: input width
    int Wout, mirror;       // output width, yMirroring
    int mode;
    int nblocks;
    double gain_bound;            // max over output columns of sum_k |B[k][n]| (FP64): |out| <= gain_bound * max|in|
    float bscale;                 // split-FP16 form: the packed matrix is B / bscale (bscale a power of two), else 1
    int bstride;                  // float4s per 16-row block of the A operand (Kp * 4, plus 4 per k tile in FP16)
    std::vector<int> view_block;  // block index of every view
    std::vector<double> H;        // L/2 + 1
    jh::DevBuf B;                 // Kp x Wout float (vector-ALU form)
    jh::DevBuf Bhl;               // MFMA forms: packed FP32, BF16-part or FP16-part operand
    // split-FP16 form, generated operand: the Toeplitz part of the operand
    // as 8 shifted copies of its FP16 hi / lo diagonals (gen_stride halves
    // each), the edge rows' r / 3 per column (hi / lo, Wout each)
    jh::DevBuf Bgen, R3;
    int gen_stride = 0;           // halves per copy
    int gen_window = 0;           // halves per copy a workgroup keeps in LDS (W + 128)
    int gen_lds = 0;              // LDS bytes of the generated form's launch
    int use_gen = 0;              // generated operand in use (else packed)
    jh::DevBuf pw;                // nblocks x V x W float
    jh::DevBuf blk;               // N int: block of every view (device copy)
};

// 128-column k tiles of the A operand (the split-FP16 form keeps the row
// maxima of each).
inline int operand_ktiles(const jh_filter_plan* p) { return (p->Kp + 127) / 128; }
