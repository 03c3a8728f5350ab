/*
 * jhip.h -- C-ABI of the MI355X (gfx950) FDK reconstruction back end.
 *
 * libjhip.so exports plain C entry points: plain pointers and sizes, no
 * exceptions, no C++ or torch types.  Device buffers passed in are owned by
 * the caller; plans own their internal tables and accumulators.  Every call
 * that takes a stream enqueues work on it and returns without synchronising,
 * except the *_create calls (which finish their set-up before returning) and
 * the calls that copy tables back to the host.  A failing call returns a
 * non-zero jh_status and leaves a message for jh_last_error() on the calling
 * thread.  A plan may be used from one host thread at a time.
 *
 * Each entry point names the reference interface it replaces (paths are
 * relative to the root of jay6987/JEnhine).
 */
#ifndef JHIP_H
#define JHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int jh_status;
#define JH_OK 0
#define JH_ERR_INVALID 1 /* bad argument or shape */
#define JH_ERR_OOM 2     /* device allocation failed (message starts "HIP failed to malloc ") */
#define JH_ERR_HIP 3     /* HIP runtime error */

/* Storage type of backprojection input (filtered, uW-weighted projections). */
#define JH_F32 0
#define JH_BF16 1
#define JH_F16 2

/* Row-filter arithmetic (all forms: Toeplitz GEMM, FP32 accumulation). */
#define JH_FILTER_FP32 0        /* FP32 operands on MFMA (v_mfma_f32_16x16x4_f32) */
#define JH_FILTER_MFMA_BF16X3 1 /* hi/lo split BF16 on MFMA, 3 products */
#define JH_FILTER_MFMA_BF16X6 2 /* hi/mid/lo split BF16 on MFMA, 6 products */
#define JH_FILTER_FP32_VALU 3   /* FP32 FMA on the vector ALUs (cross-check) */
#define JH_FILTER_MFMA_F16X3 4  /* hi/lo split FP16 on MFMA, 3 products: every row of the
                                   pre-weighted input and the filter matrix scaled by a power
                                   of two so that hi + lo keep 22 significant bits */

/* Filter windows. */
#define JH_FETCH_GATHER 0 /* backprojector sample fetch forms (jh_bp_set_fetch) */
#define JH_FETCH_LDS 1

#define JH_LAYOUT_QUAD 0  /* staged BF16/FP16 pixel: its 2 x 2 neighbourhood (8 B) */
#define JH_LAYOUT_STRIP 1 /* staged BF16/FP16 pixel: four rows of its column (8 B) */

#define JH_WEIGHT_COUNT 0 /* backprojection normalised by the back-projected uW (the reference) */
#define JH_WEIGHT_FDK 1   /* FDK distance weight 1/w^2, no count (jh_bp_set_weight) */

#define JH_WINDOW_REFERENCE 0   /* FilterCore.cpp:96-134 cut-off semantics */
#define JH_WINDOW_SHEPP_LOGAN 1 /* extension: sinc(f / (2 f_N)), zero above f_N */

/* ProjectionMatrixModifier operations (TransformMatrix/ProjectionMatrixModifier.cpp:5-43). */
#define JH_PTM_MOVE_U 0
#define JH_PTM_MOVE_V 1
#define JH_PTM_REVERSE_X 2
#define JH_PTM_REVERSE_Y 3
#define JH_PTM_MOVE_Z 4

typedef void* jh_stream_t; /* a hipStream_t (NULL = the default stream) */

typedef struct jh_pre_plan jh_pre_plan;
typedef struct jh_filter_plan jh_filter_plan;
typedef struct jh_composite_plan jh_composite_plan;
typedef struct jh_bp_plan jh_bp_plan;
typedef struct jh_graph jh_graph;

/* Last error message of the calling thread ("" when none).
 * Replaces Common/Exception.h:12-41 (ThrowExceptionAndLog) across the ABI. */
const char* jh_last_error(void);

/* Library version, major*10000 + minor*100 + patch. */
int jh_version(void);

/* Number of visible devices (0 on a host without GPUs). */
int jh_device_count(void);

/* ---------------------------------------------------------------------- */
/* Geometry                                                                */
/* ---------------------------------------------------------------------- */

/* Apply one ProjectionMatrixModifier operation, in FP64 on the device, to n
 * row-major 3x4 projection matrices in device memory (d_ptm, n*12 doubles).
 * Replaces TransformMatrix/ProjectionMatrixModifier.cpp:5-43 (float, host)
 * as called from JEngine/JEngine.cpp:139-150. */
jh_status jh_ptm_modify(double* d_ptm, int n, int op, double value, jh_stream_t stream);

/* ---------------------------------------------------------------------- */
/* Preprocessing: u16 -> line integrals                                    */
/* ---------------------------------------------------------------------- */

/* Replaces PreProcessingAgent/PreProcessingCore.cpp:15-65 (constructor):
 * input frame inW x inH, borders (up, down, left, right), output frame
 * outW x outH (crop, or linear resize when the cropped size differs),
 * bright field and beam-hardening coefficients bh[3] (skipped iff {1,0,0}). */
jh_status jh_pre_create(jh_pre_plan** plan, int device, int inW, int inH, int bUp, int bDown,
                        int bLeft, int bRight, int outW, int outH, float brightField,
                        const float* bh);

/* Replaces PreProcessingCore::Process (PreProcessingCore.cpp:67-86) for
 * nviews frames at once: d_raw (nviews x inH x inW u16) -> d_out
 * (nviews x outH x outW f32).  d_dark (inH x inW u16, subtracted as
 * integers) and d_gain (outH x outW f32, divisor BrightField*gain) are
 * optional extensions; NULL gives the reference chain exactly. */
jh_status jh_pre_run(jh_pre_plan* plan, const uint16_t* d_raw, const uint16_t* d_dark,
                     const float* d_gain, float* d_out, int nviews, jh_stream_t stream);

/* An upper bound of |preprocessed value| for gain maps within [gainMin,
 * gainMax] (no gain map: 1, 1): the corrected count lies in [1, 65535], so
 * -ln(f / (BrightField g)) and its beam-hardening polynomial are bounded
 * over that range.  Used to choose the FP16 storage exponent
 * (jh_bp_set_f16_exponent).  No counterpart in the reference. */
jh_status jh_pre_output_bound(const jh_pre_plan* plan, double gainMin, double gainMax, double* bound);

jh_status jh_pre_destroy(jh_pre_plan* plan);

/* ---------------------------------------------------------------------- */
/* Row ramp filter with pre-filter weight                                  */
/* ---------------------------------------------------------------------- */

/* Replaces FilterAgent/FilterCore.cpp:13-61 (design), the per-view pre-weight
 * of FilterAgent/PreFilterWeightGenerator.cpp:10-124 with its refresh
 * schedule (FilterAgent.cpp:101-126), for frames of W x V.  h_ptm: nviews
 * row-major 3x4 matrices on the host (FP64).  Pre-weight maps are built on
 * the device in FP64.  adjPts/adjDB: nAdj dB-adjust knots (may be NULL when
 * nAdj = 0).  Output width = W, yMirroring = false (the oracle chain). */
jh_status jh_filter_create(jh_filter_plan** plan, int device, int W, int V, int nviews,
                           const double* h_ptm, float dso, float halfSampleRate, float cutStart,
                           float cutEnd, const float* adjPts, const float* adjDB, int nAdj,
                           int window, int mode);

/* Replaces the FilterCore constructor with its widthOut and yMirroring
 * arguments (FilterAgent/FilterCore.cpp:13-61) as FilterAgent builds it
 * (FilterAgent.cpp:10-56): frames of Win x V in, Wout x V out.  With
 * yMirroring the truncation ramp (R-1-i)/R follows the data, its edge is the
 * mean of the last three inputs and the output is columns [Wout, 2 Wout) of
 * the filtered row; the pre-weight u axis is shifted by Win / 2
 * (PreFilterWeightGenerator.cpp:36-37).  JEngine's offset-detector chain
 * uses Win = 2 NumUsedDetsU, Wout = NumUsedDetsU, yMirroring =
 * detectorOnTheRight (JEngine.cpp:223-238).  jh_filter_create is the form
 * Win = Wout = W, yMirroring = 0.  Requires (yMirroring ? 2 : 1) * Wout <= L. */
jh_status jh_filter_create_ex(jh_filter_plan** plan, int device, int Win, int V, int Wout, int yMirroring,
                              int nviews, const double* h_ptm, float dso, float halfSampleRate, float cutStart,
                              float cutEnd, const float* adjPts, const float* adjDB, int nAdj, int window,
                              int mode);

/* Input width, output width and yMirroring of a filter plan (any may be NULL). */
jh_status jh_filter_shape(const jh_filter_plan* plan, int* Win, int* Wout, int* yMirroring);

/* Copies the frequency response H[0..L/2] (L = FFT length of the reference,
 * FilterCore.cpp:254-260) to the host; *L receives L. */
jh_status jh_filter_response(jh_filter_plan* plan, double* h_H, int* L);

/* Copies the pre-weight map used for view `view` (V rows x Win columns) to
 * device memory d_out. */
jh_status jh_filter_preweight(jh_filter_plan* plan, int view, float* d_out, jh_stream_t stream);

/* Replaces FilterCore::ProcessRow (FilterCore.cpp:189-222) over all rows of
 * nviews frames: d_in is nviews x V x Win f32, d_out nviews x V x Wout;
 * frame i is view (view0 + i) of the scan (selects its pre-weight map). */
jh_status jh_filter_run(jh_filter_plan* plan, const float* d_in, float* d_out, int view0,
                        int nviews, jh_stream_t stream);

/* Whether jh_filter_run_staged can take the plan pair: an MFMA filter whose
 * Wout is the backprojector's W (same V and view count, same device), and a
 * backprojector staging `dtype` (JH_BF16 / JH_F16) as column strips without
 * the ZFilter.  *fusable = 1 or 0. */
jh_status jh_filter_stage_fusable(const jh_filter_plan* plan, const jh_bp_plan* bp, int dtype, int* fusable);

/* jh_filter_run followed by jh_bp_stage in one pass, with bit-identical
 * output: views view0..view0+nviews-1 are filtered (FilterCore::ProcessRow,
 * FilterCore.cpp:189-222), weighted by uW and written in the backprojector's
 * staged layout (BPCore::InitView, BPCore.cpp:147-151) to d_q
 * (jh_bp_stage_bytes bytes); the filtered FP32 frames are not written.
 * Requires jh_filter_stage_fusable. */
jh_status jh_filter_run_staged(jh_filter_plan* plan, jh_bp_plan* bp, const float* d_in, int view0, int nviews,
                               int dtype, void* d_q, jh_stream_t stream);

/* The pre-weighted filter operand: preprocessing (PreProcessingCore::Process,
 * PreProcessingCore.cpp:67-86, without resizing) written straight into the
 * MFMA filter's A operand -- each pixel times its pre-weight (FilterCore's
 * ProcessRow input, FilterCore.cpp:189-222, with FilterAgent's weighting),
 * plus the three unweighted edge columns -- in the order the MFMA kernel
 * loads it; jh_filter_run_operand_staged then filters and stages it.  The
 * pair is bit-identical to jh_pre_run, jh_filter_run_staged.  Operand size:
 * jh_filter_operand_bytes (16-row blocks of Kp floats).  Requires a
 * non-resizing preprocessing plan whose output frame is the filter's input
 * frame, an MFMA filter, and V * nviews <= 1048560. */
jh_status jh_filter_operand_bytes(const jh_filter_plan* plan, int nviews, size_t* bytes);
jh_status jh_pre_run_operand(jh_pre_plan* pre, jh_filter_plan* plan, const uint16_t* d_raw, const uint16_t* d_dark,
                             const float* d_gain, int view0, int nviews, float* d_a, jh_stream_t stream);
jh_status jh_filter_run_operand_staged(jh_filter_plan* plan, jh_bp_plan* bp, const float* d_a, int view0,
                                       int nviews, int dtype, void* d_q, jh_stream_t stream);

/* outBound = inBound * max_n sum_k |B[k][n]|: a bound of the filtered
 * values for inputs bounded by inBound (the pre-weight is at most 1; B is
 * the Toeplitz operand with the truncation-ramp rows, in FP64).  No
 * counterpart in the reference. */
jh_status jh_filter_output_bound(const jh_filter_plan* plan, double inBound, double* outBound);

/* Split-FP16 form (JH_FILTER_MFMA_F16X3): the filter matrix operand read from
 * its Toeplitz diagonals kept in LDS (enable = 1; the default where W is a
 * multiple of 32 and the output width of 128) or copied from the packed
 * matrix per k step (enable = 0).  Both give the same result to the bit.
 * Enabling fails where the plan has no generated operand.  No counterpart in
 * the reference (FilterCore filters by FFT, FilterCore.cpp:189-222). */
jh_status jh_filter_set_generated(jh_filter_plan* plan, int enable);
jh_status jh_filter_get_generated(jh_filter_plan* plan, int* enabled);

jh_status jh_filter_destroy(jh_filter_plan* plan);

/* ---------------------------------------------------------------------- */
/* Offset-detector composite                                               */
/* ---------------------------------------------------------------------- */

/* JEngine/JEngine.cpp:153-161: 1 when the iso-centre projects left of the
 * detector middle in the first matrix (ptm0: 12 doubles, row-major). */
int jh_detector_on_right(const double* ptm0, int numUsedDetsU);

/* Replaces CompositeAgent/CompositeCore.cpp:13-65 (constructor) for frames
 * of W x V and nviews (even) views of a full rotation; h_ptm: nviews
 * row-major 3x4 matrices on the host.  detectorOnRight < 0 decides it from
 * the first matrix as jh_detector_on_right does.  The per-view maps from the
 * direct view's detector, through the iso-centre plane (pseudo-depth 0),
 * into the conjugate view (v + nviews/2) % nviews are composed on the device
 * in FP64. */
jh_status jh_composite_create(jh_composite_plan** plan, int device, int W, int V, int nviews,
                              const double* h_ptm, float dso, int detectorOnRight);

/* The detector side a composite plan uses. */
jh_status jh_composite_side(const jh_composite_plan* plan, int* detectorOnRight);

/* Replaces CompositeCore::Process (CompositeCore.cpp:66-150) for nviews
 * views: frame i is view (view0 + i), its direct frame at
 * d_dirc + i * inStride and its conjugate frame (view (view0 + i + N/2) % N,
 * the pairing of JEngine/InputAgent.cpp:47-85) at d_conj + i * inStride
 * (elements; W x V f32 each).  d_out: nviews x V x 2W f32, the direct frame
 * in the right half when the detector is on the right, else in the left. */
jh_status jh_composite_run(jh_composite_plan* plan, const float* d_dirc, const float* d_conj, long long inStride,
                           float* d_out, int view0, int nviews, jh_stream_t stream);

/* The composite (as jh_composite_run) written straight into the MFMA filter's
 * pre-weighted operand, the way jh_pre_run_operand writes preprocessing:
 * views view0 .. view0 + nviews - 1 land at operand rows
 * (view0 - aview0) * V .. of d_a, an operand of the filter launch that starts
 * at view aview0, so the two halves of a scan (different conjugate buffers)
 * fill one operand with two calls.  The filter takes frames of 2W columns
 * (plan->W == 2 * composite W, same V and view count, same device, an MFMA
 * mode).  d_a holds at least jh_filter_operand_bytes(plan, view0 + nviews -
 * aview0) bytes; a call writes only the float4s of its own rows, so calls for
 * disjoint views may share a 16-row block.  jh_filter_run_operand_staged(plan,
 * bp, d_a, aview0, ...) then gives bit-identically what jh_composite_run,
 * jh_filter_run_staged give. */
jh_status jh_composite_run_operand(jh_composite_plan* plan, jh_filter_plan* filter, const float* d_dirc,
                                   const float* d_conj, long long inStride, int view0, int nviews, float* d_a,
                                   int aview0, jh_stream_t stream);

jh_status jh_composite_destroy(jh_composite_plan* plan);

/* ---------------------------------------------------------------------- */
/* Backprojection                                                          */
/* ---------------------------------------------------------------------- */

/* Replaces BPAgent/BPCore.cpp:15-101 and BPCUDAAgent/BPCUDACore.cpp:17-127
 * (constructors): nviews projection matrices (h_ptm, host, FP64, row-major
 * 3x4), detector W x V, volume X x Y x Z (memory order vol[z][y][x]) at
 * pitches pxy, pz.  Per-view geometry and detector weights are set up on the
 * device in FP64. */
jh_status jh_bp_create(jh_bp_plan** plan, int device, const double* h_ptm, int nviews, int W,
                       int V, int X, int Y, int Z, float pxy, float pz);

/* Copies the per-view detector weights uW (nviews x W, BPCore.cpp:68-100)
 * to the host. */
jh_status jh_bp_uweights(jh_bp_plan* plan, float* h_out);

/* Bytes of staged backprojection input for nviews views in `dtype`. */
jh_status jh_bp_stage_bytes(jh_bp_plan* plan, int dtype, int nviews, size_t* bytes);

/* Replaces BPCore::InitView's projection weighting (BPCore.cpp:147-151):
 * q = filtered * uW for views view0..view0+nviews-1, written in the
 * backprojector's storage layout and type (d_q, jh_bp_stage_bytes bytes). */
jh_status jh_bp_stage(jh_bp_plan* plan, const float* d_filtered, int view0, int nviews, int dtype,
                      void* d_q, jh_stream_t stream);

/* One-shot normalised backprojection of all views: d_vol (Z x Y x X f32) =
 * sum_v bilinear(q_v) / sum_v bilinear(uW_v), accumulated in FP32.  The u
 * weight of a BF16 sample is taken from the nearer column and rounded to BF16
 * (u off by at most 2^-10 pixel), of an FP16 sample rounded to a multiple of
 * 1/2048 (2^-12 pixel); FP32 samples use the exact fraction.  The count
 * weights use the same u weight as the samples.  Replaces the whole
 * BPCore InitShot/InitView/ProcessSlice/DoneSlice sequence
 * (BPCore.cpp:108-205) and BPCUDACore's per-view loop
 * (BPCUDAAgent/BPCUDAAgent.cpp:231-287). */
jh_status jh_bp_run(jh_bp_plan* plan, const void* d_q, int dtype, float* d_vol, jh_stream_t stream);

/* Streaming form, used by the agent: zero the accumulators (BPCore::InitShot,
 * BPCore.cpp:108-113; BPCUDACore::InitShot_Synced, BPCUDACore.cpp:134-138). */
jh_status jh_bp_reset(jh_bp_plan* plan, jh_stream_t stream);

/* Accumulate views view0..view0+nviews-1 (d_q holds exactly those views) into
 * the plan's accumulator and count volumes.  Each launch continues the running
 * sums of the earlier ones in view order, so the finished volume does not
 * depend on how the views were split into calls: it equals jh_bp_run's to
 * the bit when the views are accumulated in order.  Replaces
 * BPCore::ProcessSlice (BPCore.cpp:157-196) and BPCUDACore::DeployBackProj
 * (BPCUDACore.cu:89-114). */
jh_status jh_bp_accumulate(jh_bp_plan* plan, const void* d_q, int dtype, int view0, int nviews,
                           jh_stream_t stream);

/* jh_bp_accumulate restricted to slices z0..z0+nz-1: z0 a multiple of the
 * backprojector's z tile (jh_bp_get_tile), nz too unless the slab ends at Z.
 * The per-column count sums of geometry class 0 are continued only by the
 * call whose slab starts at z0 = 0, so a set of views goes into that slab
 * exactly once, before any slice is finalised.  Each voxel still takes its
 * views in order, so views accumulated in order -- whole-volume calls first,
 * then the remaining views slab by slab -- give jh_bp_run's volume to the bit.
 * A slab can then be finalised and downloaded while the next one is
 * accumulated.  No counterpart in the reference (BPCUDAAgent accumulates every
 * view over the whole volume, BPCUDAAgent.cpp:231-268). */
jh_status jh_bp_accumulate_slab(jh_bp_plan* plan, const void* d_q, int dtype, int view0, int nviews, int z0,
                                int nz, jh_stream_t stream);

/* Write slices z0..z0+nz-1 of acc/count to d_vol (nz x Y x X f32).
 * Replaces BPCore::DoneSlice (BPCore.cpp:198-205) and
 * BPCUDACore::DeployUpdateOutput + Div2 (BPCUDACore.cu:68-87,177-215). */
jh_status jh_bp_finalize(jh_bp_plan* plan, float* d_vol, int z0, int nz, jh_stream_t stream);

/* Packs frames (nviews x V x W f32) into the backprojector's layout without
 * the uW weighting, for use with jh_bp_backproject.  No counterpart in the
 * reference. */
jh_status jh_bp_pack(jh_bp_plan* plan, const float* d_frames, int view0, int nviews, int dtype,
                     void* d_q, jh_stream_t stream);

/* Un-normalised backprojection B of views view0..view0+nviews-1 of an FP32
 * packed or staged buffer: d_out (Z x Y x X) = sum_v bilinear(q_v).  This is
 * the accumulation of BPCore::ProcessSlice (BPCore.cpp:170-192) without the
 * count division, i.e. the exact transpose of jh_fp_project. */
jh_status jh_bp_backproject(jh_bp_plan* plan, const void* d_q, int view0, int nviews, float* d_out,
                            jh_stream_t stream);

/* Forward projector, the exact transpose of jh_bp_backproject: d_proj
 * (nviews x V x W f32, plain row-major frames) = B^T d_vol for views
 * view0..view0+nviews-1, multiplied per pixel column by uW when weighted != 0
 * (then it is the transpose of staging followed by B).  The reference has no
 * forward projector (its call was removed, JEngine/JEngine.cpp:27). */
jh_status jh_fp_project(jh_bp_plan* plan, const float* d_vol, float* d_proj, int view0, int nviews,
                        int weighted, jh_stream_t stream);

/* Allocation-free streaming accumulation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, hosts that forbid hipMalloc).  It is
 * jh_bp_reset / jh_bp_accumulate / jh_bp_finalize in one entry point whose
 * accumulator and count volumes live in a static device array of the
 * library (two volumes of at most 8 Mi = 8388608 voxels, e.g. 256 x 256 x
 * 128, and the columns' count weights, at most 2 Mi = 2097152 columns; 72 MiB
 * held by every device the library runs on) instead of the plan's buffers; it
 * calls no allocating routine.  flags:
 *   JH_NOALLOC_FIRST  clear the static accumulators, then accumulate;
 *   JH_NOALLOC_LAST   after accumulating, write the finished volume (as
 *                     jh_bp_finalize, Z x Y x X f32) to d_vol.
 * views view0 .. view0+nviews-1 of d_q (staged as for jh_bp_accumulate) are
 * accumulated; nviews may be 0.  Calls between FIRST and LAST continue one
 * sequence and must keep the volume size, weighting, geometry class and
 * storage scale.
 * Every call on a device shares that device's array and the array keeps the
 * last call's contents: sequences that overlap on one device (from several
 * threads or streams) race and give wrong volumes.  A volume above the limit
 * is refused with JH_ERR_INVALID. */
#define JH_NOALLOC_FIRST 1
#define JH_NOALLOC_LAST 2
jh_status jh_bp_accumulate_noalloc(jh_bp_plan* plan, const void* d_q, int dtype, int view0, int nviews, int flags,
                                   float* d_vol, jh_stream_t stream);

/* Tile shape of the backprojector: z voxels per thread (8, 16 or 32) and
 * waves per workgroup (1, 2 or 4; each wave covers 8 x 8 voxel columns, so a
 * workgroup covers 8 x 8, 16 x 8 or 16 x 16).  0 keeps the current value.  No
 * counterpart in the reference (its launch shape is fixed at
 * BPCUDACore.cu:99-112); used by the autotuner. */
jh_status jh_bp_set_tile(jh_bp_plan* plan, int zc, int ty);

/* How the backprojector fetches detector samples: JH_FETCH_GATHER (one
 * gather per sample through the vector memory path) or JH_FETCH_LDS (each
 * workgroup stages its detector footprint of the view into LDS once, then
 * samples from LDS; geometry class 0 only, other classes gather).  The
 * reference samples a texture (BPCUDACore.cu:62); no counterpart. */
jh_status jh_bp_set_fetch(jh_bp_plan* plan, int fetch);
jh_status jh_bp_get_fetch(jh_bp_plan* plan, int* fetch);

/* The fetch form a count-normalised launch on storage type dtype takes now at
 * the plan's tile shape: JH_FETCH_LDS only where it applies (geometry class
 * 0; for BF16 / FP16 column strips, workgroups of 4 waves and a footprint
 * that fits the LDS, whose table this call builds if needed), else
 * JH_FETCH_GATHER.  Synchronises the stream when it builds the table. */
jh_status jh_bp_fetch_in_use(jh_bp_plan* plan, int dtype, jh_stream_t stream, int* fetch);

/* Staged layout of BF16/FP16 projections, 8 bytes per pixel either way:
 * JH_LAYOUT_QUAD (a pixel holds the 2 x 2 neighbourhood of a bilinear
 * sample) or JH_LAYOUT_STRIP (a pixel holds four rows v .. v+3 of its
 * column, rows padded with one zero pixel; a 16-byte gather then reads
 * columns u and u+1 of four rows and serves two consecutive z samples of a
 * voxel column).  jh_bp_create picks strips for geometry class 0 and quads
 * otherwise.  Changing it changes the format jh_bp_stage writes (and
 * jh_bp_stage_bytes): stage again after a change.  FP32 always uses its
 * 16-byte quad.  No counterpart in the reference (textures,
 * BPCUDACore.cu:62). */
jh_status jh_bp_set_layout(jh_bp_plan* plan, int layout);
jh_status jh_bp_get_layout(jh_bp_plan* plan, int* layout);

/* Fuses ZFilterAgent (FilterAgent/ZFilterAgent.cpp:49-71) into jh_bp_stage:
 * with enable != 0 the filtered frames are smoothed along v, out row r =
 * (0.5 in[r] + 0.25 in[max(r-1,0)]) + 0.25 in[min(r+1,V-1)], before the uW
 * weight -- the reference's production chain Filter -> ZFilter -> BP
 * (JEngine/JEngine.cpp:240-245).  jh_bp_pack is not affected. */
jh_status jh_bp_set_zfilter(jh_bp_plan* plan, int enable);

/* Fuses CTNumAgent (CTNumAgent/CTNumCore.cpp:22-36) into the normalised
 * output of jh_bp_run and jh_bp_finalize: with enable != 0 each value x
 * becomes x * (norm0 * 1000 / muWater) - 1000 and, above 300,
 * min(8191, (x - 300) * norm1 + 300) (JEngine.cpp:297-305; MuWater defaults to
 * 0.02, Config/ConfigHelper.cpp:70). */
jh_status jh_bp_set_ctnum(jh_bp_plan* plan, int enable, float norm0, float norm1, float muWater);

/* FP16 storage range: jh_bp_stage and the fused filter staging store
 * q * 2^-exponent in FP16, and the outputs of FP16 storage (jh_bp_run,
 * jh_bp_finalize of FP16 chunks) are multiplied back by 2^exponent.  Both
 * scalings are exact, so a result whose staged values are within FP16's
 * normal range is bit-identical to exponent 0; a host picks the exponent so
 * that its bound of |q| (jh_pre_output_bound, jh_filter_output_bound; uW <=
 * 1) maps to at most 2^15, which rules out overflow (FP16's largest finite
 * value is 65504).  Set it before staging: chunks accumulated since the last
 * jh_bp_reset must share one exponent.  Range [-64, 64]; default 0.  No
 * counterpart in the reference (it stores FP32 textures). */
jh_status jh_bp_set_f16_exponent(jh_bp_plan* plan, int exponent);
jh_status jh_bp_get_f16_exponent(jh_bp_plan* plan, int* exponent);

/* Voxel weighting of the backprojection outputs (jh_bp_run,
 * jh_bp_accumulate + jh_bp_finalize).  JH_WEIGHT_COUNT (the default) is the
 * reference's: vol = sum_v bilinear(q_v) / sum_v bilinear(uW_v)
 * (BPAgent/BPCore.cpp:147-151,198-205).  JH_WEIGHT_FDK is the textbook FDK
 * distance weight that normalisation replaces: vol = scale * sum_v
 * bilinear(q_v) / w_v^2, w_v = P_v[2] . (x, y, z, 1) the voxel's third
 * homogeneous coordinate (1 at the iso-centre for the reference's
 * normalised matrices), a sample off the detector adding nothing (no count;
 * BPCUDAAgent/BPCUDACore.cu:62 accumulates unweighted samples).  scale = 0
 * selects 2 / nviews, which matches the count form at the iso-centre of a
 * full orbit, where the conjugate views' uW sum to one.  The weighted form
 * always gathers (JH_FETCH_LDS applies to the count form only).  No
 * counterpart in the reference as an option. */
jh_status jh_bp_set_weight(jh_bp_plan* plan, int weight, float scale);
jh_status jh_bp_get_weight(jh_bp_plan* plan, int* weight, float* scale);

/* Autotuner: times the normalised backprojection of views 0..nviews-1 of
 * d_q (reps launches each, after one warm-up) for every fetch form x tile
 * shape zc in {8, 16, 32} x ty in {1, 2, 4}, and keeps the fastest in the
 * plan.  h_ms (18 floats, may be NULL) receives the milliseconds per launch
 * in the order fetch, zc, ty (outer to inner); -1 marks a combination that
 * does not apply (the LDS fetch outside geometry class 0, or column strips
 * whose footprint at that tile exceeds the LDS, where the launch would
 * gather).  The per-view cost
 * depends on the view angle, so nviews should cover the whole orbit.
 * Writes d_vol; synchronises the stream.  No counterpart in the reference. */
jh_status jh_bp_autotune(jh_bp_plan* plan, const void* d_q, int dtype, int nviews, float* d_vol, int reps,
                         jh_stream_t stream, float* h_ms);

/* Current tile shape. */
jh_status jh_bp_get_tile(jh_bp_plan* plan, int* zc, int* ty);

/* Geometry class picked at create time: 0 = u and w constant along z,
 * 1 = w constant along z, 2 = general. */
jh_status jh_bp_geometry_mode(jh_bp_plan* plan, int* mode);

jh_status jh_bp_destroy(jh_bp_plan* plan);

/* Stand-alone ZFilterAgent::ProcessTask (FilterAgent/ZFilterAgent.cpp:49-71)
 * over nviews frames of V rows x W (d_in != d_out), on the current device. */
jh_status jh_zfilter_run(const float* d_in, float* d_out, int W, int V, int nviews, jh_stream_t stream);

/* Stand-alone CTNumCore::Process (CTNumAgent/CTNumCore.cpp:22-36) in place
 * over n values, on the current device. */
jh_status jh_ctnum_run(float* d_inout, size_t n, float norm0, float norm1, float muWater, jh_stream_t stream);

/* ---------------------------------------------------------------------- */
/* Batches: many scans captured into one hipGraph                          */
/* ---------------------------------------------------------------------- */

/* A batch of independent scans reconstructed by one graph launch.  The
 * reference reconstructs one scan per run of its agent graph
 * (JEngine/JEngine.cpp:185-313, recon-part loop :400-425); the batch chains,
 * per scan, the entry points that replace its agents -- jh_pre_run
 * (PreProcessingAgent), jh_filter_run (FilterAgent), jh_bp_stage +
 * jh_bp_accumulate per chunk of views (BPCUDAAgent::ProcessView loop,
 * BPCUDAAgent.cpp:231-287) and jh_bp_finalize (DeployUpdateOutput). */
jh_status jh_graph_create(jh_graph** graph, int device);

/* Adds one scan.  The plans must describe the same detector and view count
 * (pre output = filter frame = backprojector detector) and outlive the
 * graph; each backprojection plan and each d_vol may appear once.  d_raw:
 * nviews x inH x inW u16 (read at every launch), d_dark/d_gain as in
 * jh_pre_run (may be NULL), d_vol: Z x Y x X f32 output.  dtype: storage
 * of the staged projections.  viewsPerChunk: views per chunk, 0 = chosen by
 * jh_graph_autotune (128 if it is not run).  *index (may be NULL) receives
 * the scan's index. */
jh_status jh_graph_add_scan(jh_graph* graph, jh_pre_plan* pre, jh_filter_plan* filter, jh_bp_plan* bp,
                            const uint16_t* d_raw, const uint16_t* d_dark, const float* d_gain, int dtype,
                            int viewsPerChunk, float* d_vol, int* index);

/* Autotuner for a batch: per scan, the backprojector tile (jh_bp_autotune
 * over the whole orbit on a zero-filled staging buffer; the timing depends
 * on the geometry and the tile, not on the values) and, for scans added with
 * viewsPerChunk = 0, the smallest chunk of {16, 32, ..., 256, min(N, 512)}
 * whose accumulation time per view is within 3 % of the best.  Writes the
 * scans' d_vol; synchronises the stream. */
jh_status jh_graph_autotune(jh_graph* graph, int reps, jh_stream_t stream);

/* The chunk and tile a scan will run with. */
jh_status jh_graph_scan_tuning(jh_graph* graph, int index, int* viewsPerChunk, int* zc, int* ty);

/* Assigns the scans to nbranches (1..16) parallel branches (longest first,
 * to the least-loaded branch), allocates one set of chunk buffers per
 * branch and captures the whole batch into one hipGraph.  No scans can be
 * added afterwards. */
jh_status jh_graph_instantiate(jh_graph* graph, int nbranches);

/* Replays the batch on `stream`: every scan's volume is recomputed from its
 * input buffers as they are at launch time.  Launches must not overlap
 * (they share the chunk buffers and the plans' accumulators). */
jh_status jh_graph_launch(jh_graph* graph, jh_stream_t stream);

jh_status jh_graph_destroy(jh_graph* graph);

#ifdef __cplusplus
}
#endif

#endif /* JHIP_H */
