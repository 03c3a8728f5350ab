/*
 * Type-generic body of the CPU oracle.  jh_oracle.c includes this file twice:
 * once with R=float  (SFX=_f32, "FP32-ref": float arithmetic in the
 * reference's operation order) and once with R=double (SFX=_f64, the FP64
 * oracle: the same semantics evaluated in double precision).
 *
 * This is test infrastructure.  Nothing in the shipped library calls it; only
 * the tests, the smoke check and the CPU-baseline leg of the benchmark load it.
 */

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, SFX)

/* ------------------------------------------------------------------------ */
/* Geometry                                                                  */
/* ------------------------------------------------------------------------ */

/* Centred voxel axis a[i] = (i - (n-1)/2) * pitch.
 * Follows BPAgent/BPCore.cpp:47-66 (start at -(n-1)/2, increment by one,
 * then one multiply by the pitch). */
void FN(jo_axis)(int n, R pitch, R* out)
{
    R x = -((R)n - (R)1) / (R)2;
    for (int i = 0; i < n; ++i) { out[i] = x; x = x + (R)1; }
    for (int i = 0; i < n; ++i) out[i] = out[i] * pitch;
}

/* Per-view raised-cosine detector weight across u.
 * Follows BPAgent/BPCore.cpp:68-100: u0 = (size_t)(u/w) of the iso-centre
 * (0,0,0,1); taper of width 2(W-u0) at the right end when u0 > W/2, else
 * of width 2*u0 at the left end.  A negative u0 (which the reference leaves
 * to an unsigned conversion) is taken as 0 here. */
void FN(jo_uweights)(const R* ptm, int nviews, int W, R* out)
{
    for (int iv = 0; iv < nviews; ++iv) {
        const R* P = ptm + 12 * (size_t)iv;
        R* w = out + (size_t)iv * W;
        /* MatrixMultiplier(3,4,1) of P with (0,0,0,1): u = P3, w = P11 */
        R uu = P[3] / P[11];
        long u0 = (uu > (R)0) ? (long)uu : 0;
        for (int i = 0; i < W; ++i) w[i] = (R)1;
        if (u0 > (long)(W / 2)) {
            long nT = (u0 < (long)W) ? ((long)W - u0) * 2 : 0;
            R* p = w + (W - nT);
            for (long i = 0; i < nT; ++i)
                p[i] = COS((R)i * PI_R / (R)nT) * (R)0.5 + (R)0.5;
        } else {
            long nT = u0 * 2;
            if (nT > W) nT = W;
            for (long i = 0; i < nT; ++i)
                w[i] = COS((R)i * PI_R / (R)nT) * (R)(-0.5) + (R)0.5;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Preprocessing                                                             */
/* ------------------------------------------------------------------------ */

/* IPP-style linear resize of a float ROI with pixel-centre mapping
 * x_src = (x_dst + 0.5) * sw/dw - 0.5 and replicated borders.
 * Follows Performance/ImgResizer.cpp:10-59 (pinned by
 * UnitTests/UTPerformance/UTImgResizer.cpp:15-53).  At equal sizes it is an
 * exact crop. */
void FN(jo_resize_linear)(const R* src, int srcStride, int sw, int sh,
                          R* dst, int dstStride, int dw, int dh)
{
    const R sx = (R)sw / (R)dw, sy = (R)sh / (R)dh;
    for (int y = 0; y < dh; ++y) {
        R fy = ((R)y + (R)0.5) * sy - (R)0.5;
        if (fy < 0) fy = 0;
        if (fy > (R)(sh - 1)) fy = (R)(sh - 1);
        int y0 = (int)fy; if (y0 > sh - 2) y0 = sh - 2; if (y0 < 0) y0 = 0;
        R ty = fy - (R)y0;
        int y1 = (sh > 1) ? y0 + 1 : y0;
        for (int x = 0; x < dw; ++x) {
            R fx = ((R)x + (R)0.5) * sx - (R)0.5;
            if (fx < 0) fx = 0;
            if (fx > (R)(sw - 1)) fx = (R)(sw - 1);
            int x0 = (int)fx; if (x0 > sw - 2) x0 = sw - 2; if (x0 < 0) x0 = 0;
            R tx = fx - (R)x0;
            int x1 = (sw > 1) ? x0 + 1 : x0;
            const R* r0 = src + (size_t)y0 * srcStride;
            const R* r1 = src + (size_t)y1 * srcStride;
            R a = r0[x0] + tx * (r0[x1] - r0[x0]);
            R b = r1[x0] + tx * (r1[x1] - r1[x0]);
            dst[(size_t)y * dstStride + x] = a + ty * (b - a);
        }
    }
}

/* One frame of preprocessing.
 * Follows PreProcessingAgent/PreProcessingCore.cpp:67-86 and the steps
 * GainCorr.cpp:10-13, NLog.cpp:9-13, BeamHardeningCorr.cpp:17-27:
 *   u16 -> float, LowerBound(1), crop/resize, /BrightField, ln, *-1,
 *   [BH: buf = p*a2; buf += a1; buf *= p; buf += a0; p *= buf].
 * Extension (reduces to the reference when dark == NULL and gain == NULL):
 *   the dark frame is subtracted as integers before the float conversion and
 *   the divisor is BrightField*gain[pixel].  The BH step is skipped iff
 *   bh == {1,0,0} (PreProcessingCore.cpp:56).  work must hold inW*inH. */
void FN(jo_preprocess_frame)(const uint16_t* raw, const uint16_t* dark,
                             const float* gain, int inW, int inH,
                             int bUp, int bDown, int bLeft, int bRight,
                             int outW, int outH, R brightField, const R* bh,
                             R* out, R* work)
{
    const size_t n = (size_t)inW * inH;
    for (size_t i = 0; i < n; ++i) {
        int32_t v = (int32_t)raw[i];
        if (dark) v -= (int32_t)dark[i];
        R f = (R)v;
        if (f < (R)1) f = (R)1;
        work[i] = f;
    }
    const int rw = inW - bLeft - bRight, rh = inH - bUp - bDown;
    FN(jo_resize_linear)(work + (size_t)bUp * inW + bLeft, inW, rw, rh, out, outW, outW, outH);
    const size_t m = (size_t)outW * outH;
    for (size_t i = 0; i < m; ++i) {
        R d = brightField;
        if (gain) d = d * (R)gain[i];
        out[i] = out[i] / d;
    }
    for (size_t i = 0; i < m; ++i) out[i] = LOG(out[i]);
    for (size_t i = 0; i < m; ++i) out[i] = out[i] * (R)(-1);
    if (!(bh[0] == (R)1 && bh[1] == (R)0 && bh[2] == (R)0)) {
        for (size_t i = 0; i < m; ++i) {
            R p = out[i];
            R buf = p * bh[2];
            buf = buf + bh[1];
            buf = buf * p;
            buf = buf + bh[0];
            out[i] = p * buf;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Real FFT in CCS layout                                                    */
/* ------------------------------------------------------------------------ */

/* Complex FFT, in place, sign -1 (forward) or +1 (backward, unscaled).
 * Radix-2 for powers of two, direct DFT otherwise; twiddles are evaluated
 * in double and rounded to R. */
static void FN(cfft)(int n, R* re, R* im, int sign)
{
    if (n <= 1) return;
    if ((n & (n - 1)) != 0) {
        R* tr = (R*)malloc(sizeof(R) * n);
        R* ti = (R*)malloc(sizeof(R) * n);
        for (int k = 0; k < n; ++k) {
            R sr = 0, si = 0;
            for (int j = 0; j < n; ++j) {
                long long kj = ((long long)k * j) % n;
                double a = sign * 2.0 * 3.141592653589793238462643 * (double)kj / (double)n;
                R c = (R)cos(a), s = (R)sin(a);
                sr += re[j] * c - im[j] * s;
                si += re[j] * s + im[j] * c;
            }
            tr[k] = sr; ti[k] = si;
        }
        memcpy(re, tr, sizeof(R) * n); memcpy(im, ti, sizeof(R) * n);
        free(tr); free(ti);
        return;
    }
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) { R t = re[i]; re[i] = re[j]; re[j] = t; t = im[i]; im[i] = im[j]; im[j] = t; }
    }
    for (int len = 2; len <= n; len <<= 1) {
        const int half = len >> 1;
        for (int k = 0; k < half; ++k) {
            double a = sign * 2.0 * 3.141592653589793238462643 * (double)k / (double)len;
            R wr = (R)cos(a), wi = (R)sin(a);
            for (int i = k; i < n; i += len) {
                R xr = re[i + half], xi = im[i + half];
                R tr = xr * wr - xi * wi;
                R ti = xr * wi + xi * wr;
                re[i + half] = re[i] - tr; im[i + half] = im[i] - ti;
                re[i] = re[i] + tr; im[i] = im[i] + ti;
            }
        }
    }
}

/* Forward real FFT, unscaled, output in CCS layout
 * [R0, I0, R1, I1, ..., R_{n/2}, I_{n/2}] of length floor(n/2)*2+2.
 * Follows Performance/RealFFT.cpp:13-21,39-42 and RealFFT.h:11-17. */
void FN(jo_rfft)(int n, const R* x, R* ccs)
{
    R* re = (R*)malloc(sizeof(R) * n);
    R* im = (R*)malloc(sizeof(R) * n);
    for (int i = 0; i < n; ++i) { re[i] = x[i]; im[i] = 0; }
    FN(cfft)(n, re, im, -1);
    for (int k = 0; k <= n / 2; ++k) { ccs[2 * k] = re[k]; ccs[2 * k + 1] = im[k]; }
    ccs[1] = 0;
    if ((n & 1) == 0) ccs[2 * (n / 2) + 1] = 0;
    free(re); free(im);
}

/* Backward real FFT from CCS, scaled by 1/n (RealFFT.cpp:19,44-47). */
void FN(jo_irfft)(int n, const R* ccs, R* x)
{
    R* re = (R*)malloc(sizeof(R) * n);
    R* im = (R*)malloc(sizeof(R) * n);
    for (int k = 0; k <= n / 2; ++k) { re[k] = ccs[2 * k]; im[k] = ccs[2 * k + 1]; }
    im[0] = 0;
    if ((n & 1) == 0) im[n / 2] = 0;
    for (int k = n / 2 + 1; k < n; ++k) { re[k] = re[n - k]; im[k] = -im[n - k]; }
    FN(cfft)(n, re, im, +1);
    const R s = (R)1 / (R)n;
    for (int i = 0; i < n; ++i) x[i] = re[i] * s;
    free(re); free(im);
}

/* ------------------------------------------------------------------------ */
/* Filter design                                                             */
/* ------------------------------------------------------------------------ */

/* 4x4 linear solve with partial pivoting (LAPACK ?gesv semantics, row-major,
 * one right-hand side).  Used by the spline (FilterCore.cpp:280-303). */
static void FN(solve4)(R* A, R* b)
{
    int piv[4] = {0, 1, 2, 3};
    for (int k = 0; k < 4; ++k) {
        int p = k; R best = FABS(A[k * 4 + k]);
        for (int i = k + 1; i < 4; ++i) { R v = FABS(A[i * 4 + k]); if (v > best) { best = v; p = i; } }
        if (p != k) {
            for (int j = 0; j < 4; ++j) { R t = A[k * 4 + j]; A[k * 4 + j] = A[p * 4 + j]; A[p * 4 + j] = t; }
            R t = b[k]; b[k] = b[p]; b[p] = t;
            int ti = piv[k]; piv[k] = piv[p]; piv[p] = ti;
        }
        if (A[k * 4 + k] == (R)0) continue;
        R inv = (R)1 / A[k * 4 + k];
        for (int i = k + 1; i < 4; ++i) {
            R l = A[i * 4 + k] * inv;
            A[i * 4 + k] = l;
            for (int j = k + 1; j < 4; ++j) A[i * 4 + j] -= l * A[k * 4 + j];
        }
    }
    for (int i = 0; i < 4; ++i) for (int k = 0; k < i; ++k) b[i] -= A[i * 4 + k] * b[k];
    for (int i = 3; i >= 0; --i) {
        for (int k = i + 1; k < 4; ++k) b[i] -= A[i * 4 + k] * b[k];
        b[i] = b[i] / A[i * 4 + i];
    }
}

/* Piecewise-cubic smooth interpolation, FilterCore.cpp:262-318: knot slopes
 * are non-zero only at local extrema; each segment is the cubic through the
 * two knots with those slopes; queries outside every segment return 1. */
static void FN(smooth_interp)(int n, const R* x0, const R* y0, int nq, const R* xq, R* yq)
{
    R* yPi = (R*)calloc(n, sizeof(R));
    for (int i = 1; i < n - 1; ++i) {
        if ((y0[i] - y0[i - 1]) * (y0[i + 1] - y0[i]) < 0) {
            R l = x0[i] - x0[i - 1];
            R r = x0[i + 1] - x0[i];
            R dl = (y0[i] - y0[i - 1]) / l;
            R dr = (y0[i + 1] - y0[i]) / r;
            yPi[i] = (dl * r + dr * l) / (l + r);
        }
    }
    R* c4 = (R*)malloc(sizeof(R) * 4 * (n - 1));
    for (int i = 0; i < n - 1; ++i) {
        R xl = x0[i], xr = x0[i + 1];
        R X[16];
        X[0] = xl * xl * xl; X[1] = xl * xl; X[2] = xl; X[3] = 1;
        X[4] = xr * xr * xr; X[5] = xr * xr; X[6] = xr; X[7] = 1;
        X[8] = 3 * xl * xl;  X[9] = 2 * xl;  X[10] = 1; X[11] = 0;
        X[12] = 3 * xr * xr; X[13] = 2 * xr; X[14] = 1; X[15] = 0;
        R Y[4] = {y0[i], y0[i + 1], yPi[i], yPi[i + 1]};
        FN(solve4)(X, Y);
        memcpy(c4 + 4 * i, Y, sizeof(R) * 4);
    }
    for (int i = 0; i < nq; ++i) {
        yq[i] = (R)1;
        for (int j = 0; j < n - 1; ++j) {
            if (xq[i] >= x0[j] && xq[i] <= x0[j + 1]) {
                const R* c = c4 + 4 * j;
                yq[i] = xq[i] * (xq[i] * (xq[i] * c[0] + c[1]) + c[2]) + c[3];
                break;
            }
        }
    }
    free(c4); free(yPi);
}

/* Real frequency response H[k], k = 0..L/2, including the 2*pi*f_N gain.
 * Follows FilterAgent/FilterCore.cpp:13-43 (design), :69-138 (window and
 * dB-adjust spline), :224-242 (Ram-Lak kernel -> real part of its rFFT).
 * window 0 = the reference's cut-off semantics (cut-offs 0/0 give the cosine
 * window, else a raised-cosine taper from start to end); window 1 = the
 * Shepp-Logan extension sinc(f/(2 f_N)) for f <= f_N, zero above. */
void FN(jo_filter_design)(int W, R halfSampleRate, R cutStart, R cutEnd,
                          const R* adjPts, const R* adjDB, int nAdj, int window,
                          R* H)
{
    int L = 1;
    while (L < W * 2 - 1) L *= 2;
    const int nf = L / 2 + 1;
    /* Ram-Lak in the space domain, then its real spectrum */
    R* rl = (R*)calloc(L, sizeof(R));
    rl[0] = (R)0.25;
    for (int i = 1; i <= L / 2; i += 2)
        rl[i] = rl[L - i] = (R)(-1) / (PI_R * PI_R * (R)i * (R)i);
    R* ccs = (R*)malloc(sizeof(R) * (L + 2));
    FN(jo_rfft)(L, rl, ccs);
    R* freq = (R*)malloc(sizeof(R) * nf);
    for (int i = 0; i < nf; ++i) freq[i] = (R)i * halfSampleRate * (R)2 / (R)L;
    /* knots: [0, pts..., f_N] with levels [0 dB, dB..., 0 dB] */
    const int nk = nAdj + 2;
    R* x0 = (R*)malloc(sizeof(R) * nk);
    R* y0 = (R*)malloc(sizeof(R) * nk);
    x0[0] = 0; y0[0] = 0;
    for (int i = 0; i < nAdj; ++i) { x0[i + 1] = adjPts[i]; y0[i + 1] = adjDB[i]; }
    x0[nk - 1] = halfSampleRate; y0[nk - 1] = 0;
    for (int i = 0; i < nk; ++i) y0[i] = POW((R)10, y0[i] / (R)20);
    FN(smooth_interp)(nk, x0, y0, nf, freq, H);
    if (window == 1) {
        for (int i = 0; i < nf; ++i) {
            R f = freq[i];
            if (f <= halfSampleRate) {
                R a = PI_R * f / ((R)2 * halfSampleRate);
                H[i] *= (a == (R)0) ? (R)1 : SIN(a) / a;
            } else {
                H[i] = 0;
            }
        }
    } else if (cutStart == (R)0 && cutEnd == (R)0) {
        for (int i = 0; i < nf; ++i) {
            R f = freq[i];
            if (f < halfSampleRate) H[i] *= COS(f / halfSampleRate * (PI_R * (R)0.5));
            else H[i] = 0;
        }
    } else {
        for (int i = 0; i < nf; ++i) {
            R f = freq[i];
            if (f <= cutStart) continue;
            else if (f < cutEnd)
                H[i] *= COS((f - cutStart) / (cutEnd - cutStart) * PI_R) * (R)0.5 + (R)0.5;
            else H[i] = 0;
        }
    }
    const R gain = PI_R * halfSampleRate * (R)2;
    for (int i = 0; i < nf; ++i) H[i] = (H[i] * ccs[2 * i]) * gain;
    free(rl); free(ccs); free(freq); free(x0); free(y0);
}

/* ------------------------------------------------------------------------ */
/* Pre-filter weight                                                         */
/* ------------------------------------------------------------------------ */

/* Pre-filter weight map, L columns x V rows.
 * Follows FilterAgent/PreFilterWeightGenerator.cpp:10-124 with
 * TransformMatrix/MatrixConvertor.cpp:8-27: TM = PTM rows 0,1, pseudo-depth
 * row (s*P20, s*P21, 0, s3*P23), PTM row 2; invert it; back-project
 * (u, v, w in {0,1}, 1); d = |X0 - X1|; weight = min(d)/d over the map. */
void FN(jo_prefilter_weight)(const R* ptm, int W, int V, int detectorOnRight, R DSO, R* out)
{
    int L = 1;
    while (L < W * 2 - 1) L *= 2;
    const R far_ = ((R)1 + (R)1 / DSO) * DSO;
    const R near_ = ((R)1 - (R)1 / DSO) * DSO;
    const R s20 = (far_ + near_) / (far_ - near_);
    const R s23 = ((far_ + near_) - (R)2 * far_ * near_ / DSO) / (far_ - near_) / DSO;
    R tm[16];
    memcpy(tm, ptm, 8 * sizeof(R));
    tm[8] = s20 * ptm[8]; tm[9] = s20 * ptm[9]; tm[10] = 0; tm[11] = s23 * ptm[11];
    memcpy(tm + 12, ptm + 8, 4 * sizeof(R));
    /* inverse by 4 solves against the identity (sgesv with B = I) */
    R inv[16];
    for (int c = 0; c < 4; ++c) {
        R A[16], b[4] = {0, 0, 0, 0};
        memcpy(A, tm, sizeof(A));
        b[c] = 1;
        FN(solve4)(A, b);
        for (int r = 0; r < 4; ++r) inv[r * 4 + c] = b[r];
    }
#define T(r, c) inv[(r) * 4 + (c)]
    const R w0 = 0, w1 = 1;
    R minv = (R)INFINITY;
    for (int iv = 0; iv < V; ++iv) {
        const R v = (R)iv;
        R* pw = out + (size_t)iv * L;
        for (int i = 0; i < L; ++i) {
            R u = (R)i;
            if (detectorOnRight) u = u - (R)W / (R)2;
            R t = u * T(3, 0); t = t + v * T(3, 1);
            R n0 = t + (w0 * T(3, 2) + T(3, 3));
            R n1 = t + (w1 * T(3, 2) + T(3, 3));
            R acc;
            {
                R q = u * T(0, 0); q = q + v * T(0, 1);
                R a0 = (q + (w0 * T(0, 2) + T(0, 3))) / n0;
                R a1 = (q + (w1 * T(0, 2) + T(0, 3))) / n1;
                R d = a0 - a1; acc = d * d;
            }
            {
                R q = u * T(1, 0); q = q + v * T(1, 1);
                R a0 = (q + (w0 * T(1, 2) + T(1, 3))) / n0;
                R a1 = (q + (w1 * T(1, 2) + T(1, 3))) / n1;
                R d = a0 - a1; d = d * d; acc = acc + d;
            }
            {
                R q = u * T(2, 0); q = q + v * T(2, 1);
                R a0 = (q + (w0 * T(2, 2) + T(2, 3))) / n0;
                R a1 = (q + (w1 * T(2, 2) + T(2, 3))) / n1;
                R d = a0 - a1; d = d * d; acc = acc + d;
            }
            pw[i] = SQRT(acc);
            if (pw[i] < minv) minv = pw[i];
        }
    }
#undef T
    const size_t n = (size_t)L * V;
    for (size_t i = 0; i < n; ++i) out[i] = minv / out[i];
}

/* ------------------------------------------------------------------------ */
/* Row filter                                                                */
/* ------------------------------------------------------------------------ */

/* Filters V rows of width Win into rows of width Wout.  Follows
 * FilterAgent/FilterCore.cpp:13-61 (ramp) and :189-222 (ProcessRow with a
 * pre-weight): copy, zero-pad to L = CalFFTLength(Win), multiply by the
 * pre-weight row (L values, may be NULL = all ones), write the truncation
 * ramp with the UNweighted edge mean, rFFT, multiply each CCS bin by H, irFFT
 * (1/L), keep Wout samples from offset (yMirroring ? Wout : 0).
 *   yMirroring = 0: ramp i/R into the last R = (L-Win)/2 samples, edge =
 *                   (x[0] + x[1] + x[2]) / 3;
 *   yMirroring = 1: ramp (R-1-i)/R right after the data, edge =
 *                   (x[Win-1] + x[Win-2] + x[Win-3]) / 3.
 * The oracle chain uses Win = Wout, yMirroring = 0; the offset-detector
 * chain of JEngine/JEngine.cpp:223-238 uses Win = 2 Wout, yMirroring =
 * detectorOnTheRight. */
void FN(jo_filter_rows_ex)(const R* in, const R* pw, const R* H, int Win, int Wout, int mirror, int V, R* out)
{
    int L = 1;
    while (L < Win * 2 - 1) L *= 2;
    const int Rn = (L - Win) / 2;
    const int off = mirror ? Wout : 0;
    R* buf = (R*)malloc(sizeof(R) * L);
    R* ccs = (R*)malloc(sizeof(R) * (L + 2));
    R* ramp = (R*)malloc(sizeof(R) * (Rn > 0 ? Rn : 1));
    for (int i = 0; i < Rn; ++i) ramp[i] = mirror ? (R)(Rn - 1 - i) / (R)Rn : (R)i / (R)Rn;
    for (int r = 0; r < V; ++r) {
        const R* x = in + (size_t)r * Win;
        for (int i = 0; i < Win; ++i) buf[i] = x[i];
        for (int i = Win; i < L; ++i) buf[i] = 0;
        if (pw) { const R* p = pw + (size_t)r * L; for (int i = 0; i < L; ++i) buf[i] = buf[i] * p[i]; }
        if (mirror) {
            R edge = (x[Win - 1] + x[Win - 2] + x[Win - 3]) / (R)3;
            for (int i = 0; i < Rn; ++i) buf[Win + i] = ramp[i] * edge;
        } else {
            R edge = (x[0] + x[1] + x[2]) / (R)3;
            for (int i = 0; i < Rn; ++i) buf[L - Rn + i] = ramp[i] * edge;
        }
        FN(jo_rfft)(L, buf, ccs);
        for (int k = 0; k <= L / 2; ++k) { ccs[2 * k] *= H[k]; ccs[2 * k + 1] *= H[k]; }
        FN(jo_irfft)(L, ccs, buf);
        memcpy(out + (size_t)r * Wout, buf + off, sizeof(R) * Wout);
    }
    free(buf); free(ccs); free(ramp);
}

/* The oracle chain's filter: Win = Wout = W, yMirroring = false. */
void FN(jo_filter_rows)(const R* in, const R* pw, const R* H, int W, int V, R* out)
{
    FN(jo_filter_rows_ex)(in, pw, H, W, W, 0, V, out);
}

/* ------------------------------------------------------------------------ */
/* Backprojection                                                            */
/* ------------------------------------------------------------------------ */

/* Bilinear remap sample with the chosen IPP semantics: a sample whose
 * source coordinate lies outside [0, n-1] in either axis is skipped (the
 * destination keeps 0).  Performance/ImgEditor.cpp:92-113 calls
 * ippiRemap_32f_C1R with IPPI_INTER_LINEAR; its exact edge band is not
 * pinned by any reference test. */
static inline int FN(remap_sample)(const R* img, int W, int V, R su, R sv, R* val)
{
    if (!(su >= (R)0 && su <= (R)(W - 1) && sv >= (R)0 && sv <= (R)(V - 1))) return 0;
    int x0 = (int)su, y0 = (int)sv;
    if (x0 > W - 2) x0 = W - 2;
    if (y0 > V - 2) y0 = V - 2;
    R fx = su - (R)x0, fy = sv - (R)y0;
    const R* r0 = img + (size_t)y0 * W;
    const R* r1 = r0 + W;
    R top = r0[x0] + fx * (r0[x0 + 1] - r0[x0]);
    R bot = r1[x0] + fx * (r1[x0 + 1] - r1[x0]);
    *val = top + fy * (bot - top);
    return 1;
}

/* Normalised voxel-driven backprojection over slices [z0, z0+nz).
 * Follows BPAgent/BPCore.cpp:108-205 and BPAgent.cpp:60-92:
 *   per view: scale PTM columns by the axes, XY planes
 *     U = X*P0 + (Y*P1 + P3), V = X*P4 + (Y*P5 + P7), Wd = X*P8 + (Y*P9 + P11);
 *     q = projection * uW (row-wise), weight image = uW rows;
 *   per slice: su = (U + Z*P2)/(Wd + Z*P10) + 0.5, sv likewise;
 *     acc += remap(q), count += remap(uW);
 *   vol = acc / count (count 0 gives NaN, as the reference).
 * proj holds N frames of V x W filtered projections (not yet weighted).
 * acc/count outputs are optional (NULL).  Slices are spread over nthreads in
 * tasks of 10 slices; every voxel accumulates its views in view order. */
typedef struct {
    const R* proj; const R* ptm; int N, W, V, X, Y, Z; R pxy, pz; int z0, nz;
    R* vol; R* acc; R* cnt; const R* uw; const R *ax, *ay, *az;
    int nthreads, tid;
    int fdkw; R scale;  /* distance-weighted form (jo_bp_fdkw) */
} FN(bp_args);

static void* FN(bp_worker)(void* p)
{
    FN(bp_args)* a = (FN(bp_args)*)p;
    const int W = a->W, V = a->V, X = a->X, Y = a->Y;
    const size_t XY = (size_t)X * Y, WV = (size_t)W * V;
    R* U = (R*)malloc(sizeof(R) * XY);
    R* Vv = (R*)malloc(sizeof(R) * XY);
    R* Wd = (R*)malloc(sizeof(R) * XY);
    R* q = (R*)malloc(sizeof(R) * WV);
    R* ones = (R*)malloc(sizeof(R) * WV);
    R* xs = (R*)malloc(sizeof(R) * X * 3);
    R* ys = (R*)malloc(sizeof(R) * Y * 3);
    /* tasks of 10 slices, round-robin over threads */
    for (int t0 = a->tid * 10; t0 < a->nz; t0 += 10 * a->nthreads) {
        const int t1 = (t0 + 10 < a->nz) ? t0 + 10 : a->nz;
        for (int k = t0; k < t1; ++k)
            for (size_t i = 0; i < XY; ++i) { a->acc[(size_t)k * XY + i] = 0; a->cnt[(size_t)k * XY + i] = 0; }
        for (int iv = 0; iv < a->N; ++iv) {
            const R* P = a->ptm + 12 * (size_t)iv;
            const R* uw = a->uw + (size_t)iv * W;
            for (int i = 0; i < X; ++i) { xs[i] = a->ax[i] * P[0]; xs[X + i] = a->ax[i] * P[4]; xs[2 * X + i] = a->ax[i] * P[8]; }
            for (int j = 0; j < Y; ++j) { ys[j] = a->ay[j] * P[1]; ys[Y + j] = a->ay[j] * P[5]; ys[2 * Y + j] = a->ay[j] * P[9]; }
            for (int j = 0; j < Y; ++j) {
                const R cu = ys[j] + P[3], cv = ys[Y + j] + P[7], cw = ys[2 * Y + j] + P[11];
                for (int i = 0; i < X; ++i) {
                    U[(size_t)j * X + i] = xs[i] + cu;
                    Vv[(size_t)j * X + i] = xs[X + i] + cv;
                    Wd[(size_t)j * X + i] = xs[2 * X + i] + cw;
                }
            }
            const R* pr = a->proj + (size_t)iv * WV;
            for (int r = 0; r < V; ++r)
                for (int i = 0; i < W; ++i) {
                    q[(size_t)r * W + i] = pr[(size_t)r * W + i] * uw[i];
                    ones[(size_t)r * W + i] = uw[i];
                }
            for (int k = t0; k < t1; ++k) {
                const int iz = a->z0 + k;
                const R zu = a->az[iz] * P[2], zv = a->az[iz] * P[6], zw = a->az[iz] * P[10];
                R* acc = a->acc + (size_t)k * XY;
                R* cnt = a->cnt + (size_t)k * XY;
                for (size_t i = 0; i < XY; ++i) {
                    R uu = U[i] + zu, vv = Vv[i] + zv, ww = Wd[i] + zw;
                    uu = uu / ww; vv = vv / ww;
                    uu = uu + (R)0.5; vv = vv + (R)0.5;
                    R s = 0, c = 0;
                    if (a->fdkw) {
                        /* distance weight 1/w^2 of the sample, no count */
                        if (FN(remap_sample)(q, W, V, uu, vv, &s)) {
                            const R iw = (R)1 / ww;
                            acc[i] = acc[i] + s * (iw * iw);
                        }
                        continue;
                    }
                    if (FN(remap_sample)(q, W, V, uu, vv, &s)) FN(remap_sample)(ones, W, V, uu, vv, &c);
                    acc[i] = acc[i] + s;
                    cnt[i] = cnt[i] + c;
                }
            }
        }
        for (int k = t0; k < t1; ++k)
            for (size_t i = 0; i < XY; ++i)
                a->vol[(size_t)k * XY + i] = a->fdkw ? a->scale * a->acc[(size_t)k * XY + i]
                                                     : a->acc[(size_t)k * XY + i] / a->cnt[(size_t)k * XY + i];
    }
    free(U); free(Vv); free(Wd); free(q); free(ones); free(xs); free(ys);
    return NULL;
}

static int FN(bp_run)(const R* proj, const R* ptm, int N, int W, int V, int X, int Y, int Z,
                      R pxy, R pz, int z0, int nz, R* vol, R* acc, R* cnt, int nthreads, int fdkw, R scale)
{
    if (z0 < 0 || nz < 0 || z0 + nz > Z || W < 2 || V < 2) return -1;
    const size_t XY = (size_t)X * Y;
    R* accb = acc ? acc : (R*)malloc(sizeof(R) * XY * (nz ? nz : 1));
    R* cntb = cnt ? cnt : (R*)malloc(sizeof(R) * XY * (nz ? nz : 1));
    R* uw = (R*)malloc(sizeof(R) * (size_t)N * W);
    R* ax = (R*)malloc(sizeof(R) * X);
    R* ay = (R*)malloc(sizeof(R) * Y);
    R* az = (R*)malloc(sizeof(R) * Z);
    FN(jo_uweights)(ptm, N, W, uw);
    FN(jo_axis)(X, pxy, ax); FN(jo_axis)(Y, pxy, ay); FN(jo_axis)(Z, pz, az);
    if (nthreads < 1) nthreads = 1;
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * nthreads);
    FN(bp_args)* args = (FN(bp_args)*)malloc(sizeof(FN(bp_args)) * nthreads);
    for (int t = 0; t < nthreads; ++t) {
        FN(bp_args) a = {proj, ptm, N, W, V, X, Y, Z, pxy, pz, z0, nz, vol, accb, cntb, uw, ax, ay, az, nthreads, t,
                         fdkw, scale};
        args[t] = a;
        pthread_create(&th[t], NULL, FN(bp_worker), &args[t]);
    }
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    free(th); free(args); free(uw); free(ax); free(ay); free(az);
    if (!acc) free(accb);
    if (!cnt) free(cntb);
    return 0;
}

int FN(jo_bp)(const R* proj, const R* ptm, int N, int W, int V, int X, int Y, int Z,
              R pxy, R pz, int z0, int nz, R* vol, R* acc, R* cnt, int nthreads)
{
    return FN(bp_run)(proj, ptm, N, W, V, X, Y, Z, pxy, pz, z0, nz, vol, acc, cnt, nthreads, 0, (R)0);
}

/* The FDK distance-weighted form of the same backprojection, the weighting
 * the reference replaces by the count normalisation (BPCore.cpp:147-151,
 * 198-205; BPCUDACore.cu:62 accumulates unweighted samples):
 *   acc += remap(q) * (1/w)^2 with w = Wd + Z*P10 the sample's third
 *   homogeneous coordinate (1 at the iso-centre for the reference's
 *   normalised matrices), a sample outside the detector adding nothing;
 *   vol = scale * acc.  No count volume (cnt stays 0). */
int FN(jo_bp_fdkw)(const R* proj, const R* ptm, int N, int W, int V, int X, int Y, int Z,
                   R pxy, R pz, int z0, int nz, R scale, R* vol, R* acc, int nthreads)
{
    const size_t n = (size_t)X * Y * (nz > 0 ? nz : 1);
    R* cnt = (R*)calloc(n, sizeof(R));
    const int rc = FN(bp_run)(proj, ptm, N, W, V, X, Y, Z, pxy, pz, z0, nz, vol, acc, cnt, nthreads, 1, scale);
    free(cnt);
    return rc;
}


/* ------------------------------------------------------------------------ */
/* Un-normalised backprojection and its transpose (differentiation form)     */
/* ------------------------------------------------------------------------ */

/* Sample coordinate of voxel (i, j, k) in view P, in the operation order of
 * BPCore (BPCore.cpp:120-145, 207-220). */
static inline void FN(voxel_uv)(const R* P, const R* ax, const R* ay, const R* az, int i, int j, int k,
                                R* su, R* sv)
{
    R U = ax[i] * P[0] + (ay[j] * P[1] + P[3]);
    R Vv = ax[i] * P[4] + (ay[j] * P[5] + P[7]);
    R Wd = ax[i] * P[8] + (ay[j] * P[9] + P[11]);
    U = U + az[k] * P[2]; Vv = Vv + az[k] * P[6]; Wd = Wd + az[k] * P[10];
    *su = U / Wd + (R)0.5;
    *sv = Vv / Wd + (R)0.5;
}

/* acc = sum_v remap(proj_v * (weighted ? uW_v : 1)) over the whole volume,
 * i.e. the accumulation of BPCore::ProcessSlice without the count division.
 * Single-threaded; for small test volumes. */
void FN(jo_bp_acc)(const R* proj, const R* ptm, int N, int W, int V, int X, int Y, int Z, R pxy, R pz,
                   int weighted, R* acc)
{
    R* uw = (R*)malloc(sizeof(R) * (size_t)N * W);
    R* ax = (R*)malloc(sizeof(R) * X); R* ay = (R*)malloc(sizeof(R) * Y); R* az = (R*)malloc(sizeof(R) * Z);
    R* q = (R*)malloc(sizeof(R) * (size_t)W * V);
    FN(jo_uweights)(ptm, N, W, uw);
    FN(jo_axis)(X, pxy, ax); FN(jo_axis)(Y, pxy, ay); FN(jo_axis)(Z, pz, az);
    memset(acc, 0, sizeof(R) * (size_t)X * Y * Z);
    for (int v = 0; v < N; ++v) {
        const R* P = ptm + 12 * (size_t)v;
        for (int r = 0; r < V; ++r)
            for (int u = 0; u < W; ++u)
                q[(size_t)r * W + u] = proj[((size_t)v * V + r) * W + u] * (weighted ? uw[(size_t)v * W + u] : (R)1);
        for (int k = 0; k < Z; ++k)
            for (int j = 0; j < Y; ++j)
                for (int i = 0; i < X; ++i) {
                    R su, sv, s;
                    FN(voxel_uv)(P, ax, ay, az, i, j, k, &su, &sv);
                    if (FN(remap_sample)(q, W, V, su, sv, &s)) acc[((size_t)k * Y + j) * X + i] += s;
                }
    }
    free(uw); free(ax); free(ay); free(az); free(q);
}

/* Transpose of jo_bp_acc: out_v(u, r) = (weighted ? uW_v(u) : 1) *
 * sum over voxels of x * (bilinear weight of pixel (u, r) in the voxel's
 * sample), with the same edge rule (samples outside [0,W-1]x[0,V-1] skipped;
 * the clamped stencil of remap_sample). */
void FN(jo_fp)(const R* x, const R* ptm, int N, int W, int V, int X, int Y, int Z, R pxy, R pz,
               int weighted, R* out)
{
    R* uw = (R*)malloc(sizeof(R) * (size_t)N * W);
    R* ax = (R*)malloc(sizeof(R) * X); R* ay = (R*)malloc(sizeof(R) * Y); R* az = (R*)malloc(sizeof(R) * Z);
    FN(jo_uweights)(ptm, N, W, uw);
    FN(jo_axis)(X, pxy, ax); FN(jo_axis)(Y, pxy, ay); FN(jo_axis)(Z, pz, az);
    memset(out, 0, sizeof(R) * (size_t)N * W * V);
    for (int v = 0; v < N; ++v) {
        const R* P = ptm + 12 * (size_t)v;
        R* o = out + (size_t)v * W * V;
        for (int k = 0; k < Z; ++k)
            for (int j = 0; j < Y; ++j)
                for (int i = 0; i < X; ++i) {
                    R su, sv;
                    FN(voxel_uv)(P, ax, ay, az, i, j, k, &su, &sv);
                    if (!(su >= (R)0 && su <= (R)(W - 1) && sv >= (R)0 && sv <= (R)(V - 1))) continue;
                    int x0 = (int)su, y0 = (int)sv;
                    if (x0 > W - 2) x0 = W - 2;
                    if (y0 > V - 2) y0 = V - 2;
                    const R fx = su - (R)x0, fy = sv - (R)y0;
                    const R val = x[((size_t)k * Y + j) * X + i];
                    o[(size_t)y0 * W + x0] += ((R)1 - fx) * ((R)1 - fy) * val;
                    o[(size_t)y0 * W + x0 + 1] += fx * ((R)1 - fy) * val;
                    o[(size_t)(y0 + 1) * W + x0] += ((R)1 - fx) * fy * val;
                    o[(size_t)(y0 + 1) * W + x0 + 1] += fx * fy * val;
                }
        if (weighted)
            for (int r = 0; r < V; ++r)
                for (int u = 0; u < W; ++u) o[(size_t)r * W + u] *= uw[(size_t)v * W + u];
    }
    free(uw); free(ax); free(ay); free(az);
}

/* ZFilterAgent::ProcessTask (FilterAgent/ZFilterAgent.cpp:49-71): per frame of
 * V rows x W, out row i = (0.5 in[i] + 0.25 in[max(i-1, 0)]) + 0.25
 * in[min(i+1, V-1)], in that operation order (Mul, then Add of a scaled
 * neighbour, then Add of the other). */
void FN(jo_zfilter)(const R* in, int nframes, int W, int V, R* out)
{
    for (int f = 0; f < nframes; ++f) {
        const R* src = in + (size_t)f * W * V;
        R* dst = out + (size_t)f * W * V;
        for (int i = 0; i < V; ++i) {
            const R* a = src + (size_t)i * W;
            const R* b = src + (size_t)(i > 0 ? i - 1 : 0) * W;
            const R* c = src + (size_t)(i + 1 < V ? i + 1 : V - 1) * W;
            R* o = dst + (size_t)i * W;
            for (int u = 0; u < W; ++u) {
                R t = a[u] * (R)0.5;
                const R bb = b[u] * (R)0.25;
                t = t + bb;
                const R cc = c[u] * (R)0.25;
                o[u] = t + cc;
            }
        }
    }
}

/* CTNumCore::Process (CTNumAgent/CTNumCore.cpp:22-36): x * (norm0 * 1000 /
 * muWater) - 1000, then values above 300 become min(8191, (x - 300) norm1 +
 * 300).  The factor is formed in R as the reference forms it in float. */
void FN(jo_ctnum)(R* inout, size_t n, R norm0, R norm1, R muWater)
{
    const R s = norm0 * (R)1000 / muWater;
    for (size_t t = 0; t < n; ++t) {
        R x = inout[t] * s;
        x = x - (R)1000;
        if (x > (R)300) {
            R y = (x - (R)300) * norm1;
            y = y + (R)300;
            x = y < (R)8191 ? y : (R)8191;
        }
        inout[t] = x;
    }
}

/* ------------------------------------------------------------------------ */
/* Offset-detector composite                                                 */
/* ------------------------------------------------------------------------ */

/* Transition masks of CompositeCore (CompositeAgent/CompositeCore.cpp:
 * 202-231) over transWidth = 40 columns: dirc[i] = 0.5 -/+ 0.5 cos(i pi / 40)
 * (minus when the detector is on the right), conj[i] = 1 - dirc[i]. */
void FN(jo_composite_masks)(int detRight, R* dirc, R* conj)
{
    const int tw = 40;
    for (int i = 0; i < tw; ++i) {
        const R c = COS((R)i / (R)tw * PI_R);
        dirc[i] = detRight ? c * (R)-0.5 + (R)0.5 : c * (R)0.5 + (R)0.5;
        conj[i] = (R)1 - dirc[i];
    }
}

/* CompositeCore::Process (CompositeAgent/CompositeCore.cpp:13-150) for one
 * view: out (H rows x 2W) = [conj half | direct half] (direct half on the
 * right when detRight, else on the left).
 *   - TM of the direct view (MatrixConvertor.cpp:8-27), inverted by 4 solves
 *     against the identity (LAPACK sgesv, CompositeCore.cpp:194-199);
 *   - every pixel (c, r) of the conj half is the detector point
 *     (u1, v1, 0, 1) with u1 = c - W (detRight) or c + W, v1 = r
 *     (GenerateUVW1Grid, :152-183); xyz = TM^-1 (u1, v1, 0, 1) and
 *     (U, V, W) = P_conj xyz (the two sgemm calls, sums in k order);
 *   - u2 = U / W, v2 = V / W clamped to [0, H-1]; bilinear remap of the
 *     conjugate frame at (u2, v2), a sample outside the frame leaves 0;
 *   - the 40 conj columns next to the direct half: out * maskConj +
 *     edge * maskDirc, edge = mean of the 5 direct pixels at the shared
 *     border (first 5 when detRight, else last 5);
 *   - the direct frame is copied into its half. */
void FN(jo_composite_frame)(const R* dirc, const R* conj, const R* ptm_d, const R* ptm_c, int W, int H,
                            R DSO, int detRight, R* out)
{
    const int W2 = 2 * W, tw = 40, ew = 5;
    for (size_t i = 0; i < (size_t)W2 * H; ++i) out[i] = 0;
    const R far_ = ((R)1 + (R)1 / DSO) * DSO;
    const R near_ = ((R)1 - (R)1 / DSO) * DSO;
    const R s20 = (far_ + near_) / (far_ - near_);
    const R s23 = ((far_ + near_) - (R)2 * far_ * near_ / DSO) / (far_ - near_) / DSO;
    R tm[16];
    memcpy(tm, ptm_d, 8 * sizeof(R));
    tm[8] = s20 * ptm_d[8]; tm[9] = s20 * ptm_d[9]; tm[10] = 0; tm[11] = s23 * ptm_d[11];
    memcpy(tm + 12, ptm_d + 8, 4 * sizeof(R));
    R inv[16];
    for (int c = 0; c < 4; ++c) {
        R A[16], b[4] = {0, 0, 0, 0};
        memcpy(A, tm, sizeof(A));
        b[c] = 1;
        FN(solve4)(A, b);
        for (int r = 0; r < 4; ++r) inv[r * 4 + c] = b[r];
    }
    const int conjOff = detRight ? 0 : W;
    const int dircOff = detRight ? W : 0;
    const int transOff = detRight ? W - tw : W;
    R md[40], mc[40];
    FN(jo_composite_masks)(detRight, md, mc);
    for (int r = 0; r < H; ++r) {
        R* o = out + (size_t)r * W2;
        for (int c = 0; c < W; ++c) {
            const R g[4] = {detRight ? (R)c - (R)W : (R)(c + W), (R)r, (R)0, (R)1};
            R xyz[4], uvw[3];
            for (int i = 0; i < 4; ++i) {
                R t = inv[i * 4] * g[0];
                for (int k = 1; k < 4; ++k) t = t + inv[i * 4 + k] * g[k];
                xyz[i] = t;
            }
            for (int i = 0; i < 3; ++i) {
                R t = ptm_c[i * 4] * xyz[0];
                for (int k = 1; k < 4; ++k) t = t + ptm_c[i * 4 + k] * xyz[k];
                uvw[i] = t;
            }
            const R u2 = uvw[0] / uvw[2];
            R v2 = uvw[1] / uvw[2];
            if (v2 < (R)0) v2 = (R)0;
            if (v2 > (R)(H - 1)) v2 = (R)(H - 1);
            R val;
            if (FN(remap_sample)(conj, W, H, u2, v2, &val)) o[conjOff + c] = val;
        }
        const R* d = dirc + (size_t)r * W;
        const R* e = detRight ? d : d + W - ew;
        R sum = e[0];
        for (int i = 1; i < ew; ++i) sum = sum + e[i];
        const R edge = sum / (R)ew;
        for (int i = 0; i < tw; ++i) {
            const R a = o[transOff + i] * mc[i];
            const R b = edge * md[i];
            o[transOff + i] = a + b;
        }
        for (int c = 0; c < W; ++c) o[dircOff + c] = d[c];
    }
}

#undef FN
#undef CAT
#undef CAT_
