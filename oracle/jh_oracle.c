/*
 * CPU oracle for the FDK back end: a portable C restatement of the
 * reference's CPU path  PreProcessingCore -> FilterCore -> BPCore
 * (PreProcessingAgent/PreProcessingCore.cpp, FilterAgent/FilterCore.cpp,
 * FilterAgent/PreFilterWeightGenerator.cpp, BPAgent/BPCore.cpp), in two
 * numeric forms:
 *   *_f32  FP32-ref: float arithmetic in the reference's operation order;
 *   *_f64  FP64: the same semantics in double precision.
 *
 * TEST INFRASTRUCTURE ONLY.  The shipped HIP library never links, loads or
 * calls this code; only tests/, the smoke check and the CPU-baseline leg of
 * bench.py do, and only to check results against.  Compiled with -ffp-contract=off so
 * that no multiply-add is fused behind the restated operation order.
 *
 * Semantics chosen where the reference leaves them to a closed library
 * (parity unpinned by any reference test, documented here and pinned by
 * invariants in tests/):
 *   - ippiRemap linear edge band: a sample outside [0,W-1]x[0,V-1] is skipped;
 *   - IPP bilinear form: top = p00 + fx*(p01-p00), ..., val = top + fy*(bot-top);
 *   - MKL real FFT rounding: radix-2 with double-evaluated twiddles.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>

#define R float
#define SFX _f32
#define COS cosf
#define SIN sinf
#define LOG logf
#define POW powf
#define SQRT sqrtf
#define FABS fabsf
#define PI_R ((float)3.141592653589793L)
#include "jh_oracle_impl.h"
#undef R
#undef SFX
#undef COS
#undef SIN
#undef LOG
#undef POW
#undef SQRT
#undef FABS
#undef PI_R

#define R double
#define SFX _f64
#define COS cos
#define SIN sin
#define LOG log
#define POW pow
#define SQRT sqrt
#define FABS fabs
#define PI_R 3.141592653589793
#include "jh_oracle_impl.h"

/* FFT length of the row filter: the smallest power of two >= 2W-1
 * (FilterAgent/FilterCore.cpp:254-260). */
int jo_fft_length(int W)
{
    int L = 1;
    while (L < W * 2 - 1) L *= 2;
    return L;
}

/* ------------------------------------------------------------------------ */
/* Multi-frame drivers (threaded over frames / rows, as the reference agents */
/* spread frames and 240-row tasks over threads).                            */
/* ------------------------------------------------------------------------ */

typedef struct {
    const uint16_t* raw; const uint16_t* dark; const float* gain;
    int inW, inH, bUp, bDown, bLeft, bRight, outW, outH;
    double bf; const double* bh; void* out; int nviews, fp64, nthreads, tid;
} pre_args;

static void* pre_worker(void* p)
{
    pre_args* a = (pre_args*)p;
    const size_t nin = (size_t)a->inW * a->inH, nout = (size_t)a->outW * a->outH;
    if (a->fp64) {
        double* work = (double*)malloc(sizeof(double) * nin);
        for (int v = a->tid; v < a->nviews; v += a->nthreads)
            jo_preprocess_frame_f64(a->raw + v * nin, a->dark, a->gain, a->inW, a->inH, a->bUp, a->bDown,
                                    a->bLeft, a->bRight, a->outW, a->outH, a->bf, a->bh,
                                    (double*)a->out + v * nout, work);
        free(work);
    } else {
        float* work = (float*)malloc(sizeof(float) * nin);
        float bh[3] = {(float)a->bh[0], (float)a->bh[1], (float)a->bh[2]};
        for (int v = a->tid; v < a->nviews; v += a->nthreads)
            jo_preprocess_frame_f32(a->raw + v * nin, a->dark, a->gain, a->inW, a->inH, a->bUp, a->bDown,
                                    a->bLeft, a->bRight, a->outW, a->outH, (float)a->bf, bh,
                                    (float*)a->out + v * nout, work);
        free(work);
    }
    return NULL;
}

/* Preprocess nviews frames.  out is float (fp64 = 0) or double (fp64 = 1). */
int jo_preprocess(const uint16_t* raw, const uint16_t* dark, const float* gain,
                  int inW, int inH, int bUp, int bDown, int bLeft, int bRight,
                  int outW, int outH, double brightField, const double* bh,
                  void* out, int nviews, int fp64, int nthreads)
{
    if (inW - bLeft - bRight < 1 || inH - bUp - bDown < 1 || outW < 1 || outH < 1) return -1;
    if (nthreads < 1) nthreads = 1;
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * nthreads);
    pre_args* args = (pre_args*)malloc(sizeof(pre_args) * nthreads);
    for (int t = 0; t < nthreads; ++t) {
        pre_args a = {raw, dark, gain, inW, inH, bUp, bDown, bLeft, bRight, outW, outH,
                      brightField, bh, out, nviews, fp64, nthreads, t};
        args[t] = a;
        pthread_create(&th[t], NULL, pre_worker, &args[t]);
    }
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    free(th); free(args);
    return 0;
}

typedef struct {
    const void* in; const void* pw; const void* H; int Win, Wout, mirror, V, nviews, fp64, nthreads, tid;
    const int* pw_index; int L; void* out;
} filt_args;

static void* filt_worker(void* p)
{
    filt_args* a = (filt_args*)p;
    const size_t fin = (size_t)a->Win * a->V, fout = (size_t)a->Wout * a->V, pwfr = (size_t)a->L * a->V;
    for (int v = a->tid; v < a->nviews; v += a->nthreads) {
        const int b = a->pw_index ? a->pw_index[v] : 0;
        if (a->fp64) {
            const double* pw = a->pw ? (const double*)a->pw + (size_t)b * pwfr : NULL;
            jo_filter_rows_ex_f64((const double*)a->in + v * fin, pw, (const double*)a->H, a->Win, a->Wout,
                                  a->mirror, a->V, (double*)a->out + v * fout);
        } else {
            const float* pw = a->pw ? (const float*)a->pw + (size_t)b * pwfr : NULL;
            jo_filter_rows_ex_f32((const float*)a->in + v * fin, pw, (const float*)a->H, a->Win, a->Wout,
                                  a->mirror, a->V, (float*)a->out + v * fout);
        }
    }
    return NULL;
}

/* Filter nviews frames of V rows x Win into V rows x Wout (FilterCore with
 * widthOut and yMirroring, FilterAgent/FilterCore.cpp:13-61,189-222).  pw
 * holds nblocks pre-weight maps (L x V each) and pw_index[view] picks the
 * map for each view (NULL -> 0); pw may be NULL (no pre-weight). */
int jo_filter_ex(const void* in, const void* pw, const int* pw_index, const void* H,
                 int Win, int Wout, int mirror, int V, int nviews, void* out, int fp64, int nthreads)
{
    if (Win < 3 || V < 1 || Wout < 1 || Wout * (mirror ? 2 : 1) > jo_fft_length(Win)) return -1;
    if (nthreads < 1) nthreads = 1;
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * nthreads);
    filt_args* args = (filt_args*)malloc(sizeof(filt_args) * nthreads);
    for (int t = 0; t < nthreads; ++t) {
        filt_args a = {in, pw, H, Win, Wout, mirror, V, nviews, fp64, nthreads, t, pw_index, jo_fft_length(Win), out};
        args[t] = a;
        pthread_create(&th[t], NULL, filt_worker, &args[t]);
    }
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    free(th); free(args);
    return 0;
}

/* The oracle chain's filter (Win = Wout = W, yMirroring = false). */
int jo_filter(const void* in, const void* pw, const int* pw_index, const void* H,
              int W, int V, int nviews, void* out, int fp64, int nthreads)
{
    return jo_filter_ex(in, pw, pw_index, H, W, W, 0, V, nviews, out, fp64, nthreads);
}

typedef struct {
    const void* pre; const void* ptm; int N, W, H, detRight, fp64, nthreads, tid; double dso; void* out;
} comp_args;

static void* comp_worker(void* p)
{
    comp_args* a = (comp_args*)p;
    const size_t fin = (size_t)a->W * a->H, fout = 2 * fin;
    for (int v = a->tid; v < a->N; v += a->nthreads) {
        const int c = (v + a->N / 2) % a->N;
        if (a->fp64) {
            const double* pre = (const double*)a->pre;
            const double* P = (const double*)a->ptm;
            jo_composite_frame_f64(pre + v * fin, pre + c * fin, P + 12 * v, P + 12 * c, a->W, a->H, a->dso,
                                   a->detRight, (double*)a->out + v * fout);
        } else {
            const float* pre = (const float*)a->pre;
            const float* P = (const float*)a->ptm;
            jo_composite_frame_f32(pre + v * fin, pre + c * fin, P + 12 * v, P + 12 * c, a->W, a->H,
                                   (float)a->dso, a->detRight, (float*)a->out + v * fout);
        }
    }
    return NULL;
}

/* Composite of every view of a full rotation (CompositeAgent.cpp:50-85 with
 * the frame pairing of JEngine/InputAgent.cpp:47-85): view v is composed with
 * its conjugate view (v + N/2) % N.  pre: N frames H x W; out: N frames
 * H x 2W; ptm: N x 12 in the precision of the form. */
int jo_composite(const void* pre, const void* ptm, int N, int W, int H, double dso, int detRight, void* out,
                 int fp64, int nthreads)
{
    if (W < 40 || H < 2 || N < 2) return -1;
    if (nthreads < 1) nthreads = 1;
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * nthreads);
    comp_args* args = (comp_args*)malloc(sizeof(comp_args) * nthreads);
    for (int t = 0; t < nthreads; ++t) {
        comp_args a = {pre, ptm, N, W, H, detRight, fp64, nthreads, t, dso, out};
        args[t] = a;
        pthread_create(&th[t], NULL, comp_worker, &args[t]);
    }
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    free(th); free(args);
    return 0;
}

/* Pre-weight refresh schedule of FilterAgent/FilterAgent.cpp:101-126: views
 * are handled in blocks of 9 and block b uses the map generated from view
 * min(9b + 4, N - 1).  Fills src_view[b] for b < ceil(N/9) and view_block[v]
 * for every view; returns the number of blocks. */
int jo_preweight_schedule(int nviews, int* src_view, int* view_block)
{
    const int per = 9;
    const int nb = (nviews + per - 1) / per;
    for (int b = 0; b < nb; ++b) {
        int v = b * per + per / 2;
        src_view[b] = (v < nviews - 1) ? v : nviews - 1;
    }
    for (int v = 0; v < nviews; ++v) view_block[v] = v / per;
    return nb;
}
