/*
 * Synthetic scan data for the FDK back end: seeded ellipsoid phantoms, their
 * analytic line integrals through any projection matrix (FP64), per-pixel
 * gain/dark frames and raw uint16 counts I = Poisson(I0*g*exp(-p)) + d.
 *
 * All randomness comes from a counter-based generator (SplitMix64 of
 * seed, stream and index), so every array is reproducible bit for bit,
 * independent of the thread count.
 *
 * Conventions (reference PTM convention, row-major 3x4):
 *   (u, v, w) = P (x, y, z, 1);   pixel (j, k) is sampled by the reference
 *   backprojector at s = (u/w + 0.5, v/w + 0.5) (BPAgent/BPCore.cpp:207-220),
 *   so the ray of pixel (j, k) passes through PTM coordinates (j-0.5, k-0.5).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ---- counter-based RNG ---------------------------------------------- */

static inline uint64_t smix(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

typedef struct { uint64_t s; } rng_t;

static inline rng_t rng_at(uint64_t seed, uint64_t stream, uint64_t idx)
{
    rng_t r;
    r.s = smix(smix(seed ^ smix(stream + 0x1234567ull)) + idx);
    return r;
}

static inline double rng_u01(rng_t* r)
{
    r->s = smix(r->s);
    return (double)(r->s >> 11) * (1.0 / 9007199254740992.0);
}

/* ---- phantom ---------------------------------------------------------- */

/* Ellipsoid record: cx cy cz | a b c | R[9] (columns = ellipsoid axes in
 * world coordinates) | mu    (16 doubles). */
#define EREC 16

static void rot_from_quat(double qw, double qx, double qy, double qz, double* R)
{
    double n = sqrt(qw * qw + qx * qx + qy * qy + qz * qz);
    qw /= n; qx /= n; qy /= n; qz /= n;
    R[0] = 1 - 2 * (qy * qy + qz * qz); R[1] = 2 * (qx * qy - qz * qw);     R[2] = 2 * (qx * qz + qy * qw);
    R[3] = 2 * (qx * qy + qz * qw);     R[4] = 1 - 2 * (qx * qx + qz * qz); R[5] = 2 * (qy * qz - qx * qw);
    R[6] = 2 * (qx * qz - qy * qw);     R[7] = 2 * (qy * qz + qx * qw);     R[8] = 1 - 2 * (qx * qx + qy * qy);
}

/* n random ellipsoids: centres uniform in frac * (hx, hy, hz), semi-axes
 * uniform in [ax_lo, ax_hi], attenuation uniform in [mu_lo, mu_hi], uniform
 * random rotation.  With tooth != 0 one more ellipsoid of semi-axes
 * (3, 3, 5) mm and attenuation tooth_mu is appended, centred uniformly in
 * 0.4 * (hx, hy, hz).  Returns the number of records written. */
int jh_synth_phantom(uint64_t seed, int n, double hx, double hy, double hz, double frac,
                     double ax_lo, double ax_hi, double mu_lo, double mu_hi,
                     int tooth, double tooth_mu, double* out)
{
    for (int i = 0; i < n + (tooth ? 1 : 0); ++i) {
        rng_t r = rng_at(seed, 1, (uint64_t)i);
        double* e = out + (size_t)i * EREC;
        const int is_tooth = (i == n);
        const double f = is_tooth ? 0.4 : frac;
        e[0] = (2 * rng_u01(&r) - 1) * f * hx;
        e[1] = (2 * rng_u01(&r) - 1) * f * hy;
        e[2] = (2 * rng_u01(&r) - 1) * f * hz;
        if (is_tooth) { e[3] = 3; e[4] = 3; e[5] = 5; }
        else for (int k = 0; k < 3; ++k) e[3 + k] = ax_lo + (ax_hi - ax_lo) * rng_u01(&r);
        /* uniform random unit quaternion (Shoemake) */
        double u1 = rng_u01(&r), u2 = rng_u01(&r), u3 = rng_u01(&r);
        double a = sqrt(1 - u1), b = sqrt(u1);
        rot_from_quat(b * cos(2 * M_PI * u3), a * sin(2 * M_PI * u2), a * cos(2 * M_PI * u2),
                      b * sin(2 * M_PI * u3), e + 6);
        e[15] = is_tooth ? tooth_mu : mu_lo + (mu_hi - mu_lo) * rng_u01(&r);
    }
    return n + (tooth ? 1 : 0);
}

/* Point-sampled attenuation volume vol[z][y][x] at the voxel centres
 * a[i] = (i - (n-1)/2) * pitch (overlapping ellipsoids add). */
void jh_synth_voxelize(const double* ell, int ne, int X, int Y, int Z, double pxy, double pz, float* vol)
{
#pragma omp parallel for schedule(dynamic, 1)
    for (int k = 0; k < Z; ++k) {
        const double z = (k - (Z - 1) / 2.0) * pz;
        for (int j = 0; j < Y; ++j) {
            const double y = (j - (Y - 1) / 2.0) * pxy;
            for (int i = 0; i < X; ++i) {
                const double x = (i - (X - 1) / 2.0) * pxy;
                double mu = 0;
                for (int q = 0; q < ne; ++q) {
                    const double* e = ell + (size_t)q * EREC;
                    const double dx = x - e[0], dy = y - e[1], dz = z - e[2];
                    const double* R = e + 6;
                    const double l0 = (R[0] * dx + R[3] * dy + R[6] * dz) / e[3];
                    const double l1 = (R[1] * dx + R[4] * dy + R[7] * dz) / e[4];
                    const double l2 = (R[2] * dx + R[5] * dy + R[8] * dz) / e[5];
                    if (l0 * l0 + l1 * l1 + l2 * l2 <= 1.0) mu += e[15];
                }
                vol[((size_t)k * Y + j) * X + i] = (float)mu;
            }
        }
    }
}

/* ---- projection geometry helpers ------------------------------------- */

static int inv3(const double* M, double* Mi)
{
    const double a = M[0], b = M[1], c = M[2], d = M[3], e = M[4], f = M[5], g = M[6], h = M[7], k = M[8];
    const double A = e * k - f * h, B = -(d * k - f * g), C = d * h - e * g;
    const double det = a * A + b * B + c * C;
    if (det == 0) return -1;
    const double id = 1.0 / det;
    Mi[0] = A * id; Mi[1] = -(b * k - c * h) * id; Mi[2] = (b * f - c * e) * id;
    Mi[3] = B * id; Mi[4] = (a * k - c * g) * id;  Mi[5] = -(a * f - c * d) * id;
    Mi[6] = C * id; Mi[7] = -(a * h - b * g) * id; Mi[8] = (a * e - b * d) * id;
    return 0;
}

/* Analytic line integrals p[view][k][j] = sum_e mu_e * chord_e along the ray
 * of pixel (j, k), for N views of V rows x W columns.  Rays and chords are
 * evaluated in double; every ellipsoid is culled per view to the detector
 * box of its projected bounding box. */
int jh_synth_lineint(const double* ptm, int N, int W, int V, const double* ell, int ne, float* out)
{
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : bad)
    for (int iv = 0; iv < N; ++iv) {
        const double* P = ptm + 12 * (size_t)iv;
        const double M[9] = {P[0], P[1], P[2], P[4], P[5], P[6], P[8], P[9], P[10]};
        double Mi[9];
        float* img = out + (size_t)iv * W * V;
        memset(img, 0, sizeof(float) * (size_t)W * V);
        if (inv3(M, Mi)) { bad++; continue; }
        /* source: M C = -p4 */
        const double p4[3] = {P[3], P[7], P[11]};
        double S[3];
        for (int r = 0; r < 3; ++r) S[r] = -(Mi[3 * r] * p4[0] + Mi[3 * r + 1] * p4[1] + Mi[3 * r + 2] * p4[2]);
        for (int q = 0; q < ne; ++q) {
            const double* e = ell + (size_t)q * EREC;
            const double* R = e + 6;
            /* detector box from the 8 corners of the oriented bounding box */
            double umin = 1e300, umax = -1e300, vmin = 1e300, vmax = -1e300;
            int behind = 0;
            for (int c = 0; c < 8; ++c) {
                const double s0 = (c & 1 ? 1 : -1) * e[3], s1 = (c & 2 ? 1 : -1) * e[4], s2 = (c & 4 ? 1 : -1) * e[5];
                const double X = e[0] + R[0] * s0 + R[1] * s1 + R[2] * s2;
                const double Yc = e[1] + R[3] * s0 + R[4] * s1 + R[5] * s2;
                const double Zc = e[2] + R[6] * s0 + R[7] * s1 + R[8] * s2;
                const double uu = P[0] * X + P[1] * Yc + P[2] * Zc + P[3];
                const double vv = P[4] * X + P[5] * Yc + P[6] * Zc + P[7];
                const double ww = P[8] * X + P[9] * Yc + P[10] * Zc + P[11];
                if (ww <= 0) { behind = 1; break; }
                const double su = uu / ww + 0.5, sv = vv / ww + 0.5;
                if (su < umin) umin = su; if (su > umax) umax = su;
                if (sv < vmin) vmin = sv; if (sv > vmax) vmax = sv;
            }
            int j0 = 0, j1 = W - 1, k0 = 0, k1 = V - 1;
            if (!behind) {
                j0 = (int)fmax(0.0, floor(umin) - 1); j1 = (int)fmin(W - 1.0, ceil(umax) + 1);
                k0 = (int)fmax(0.0, floor(vmin) - 1); k1 = (int)fmin(V - 1.0, ceil(vmax) + 1);
            }
            /* source in the ellipsoid frame, scaled to the unit sphere */
            const double ds[3] = {S[0] - e[0], S[1] - e[1], S[2] - e[2]};
            double o[3];
            for (int a = 0; a < 3; ++a) o[a] = (R[a] * ds[0] + R[3 + a] * ds[1] + R[6 + a] * ds[2]) / e[3 + a];
            const double oo = o[0] * o[0] + o[1] * o[1] + o[2] * o[2];
            for (int k = k0; k <= k1; ++k) {
                for (int j = j0; j <= j1; ++j) {
                    const double pu = j - 0.5, pv = k - 0.5;
                    double d[3];
                    for (int r = 0; r < 3; ++r) d[r] = Mi[3 * r] * pu + Mi[3 * r + 1] * pv + Mi[3 * r + 2];
                    const double dn = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                    double l[3];
                    for (int a = 0; a < 3; ++a) l[a] = (R[a] * d[0] + R[3 + a] * d[1] + R[6 + a] * d[2]) / e[3 + a];
                    const double A = l[0] * l[0] + l[1] * l[1] + l[2] * l[2];
                    const double B = o[0] * l[0] + o[1] * l[1] + o[2] * l[2];
                    const double disc = B * B - A * (oo - 1.0);
                    if (disc <= 0) continue;
                    /* chord length in world units: (t2 - t1) |d| */
                    img[(size_t)k * W + j] += (float)(e[15] * 2.0 * sqrt(disc) / A * dn);
                }
            }
        }
    }
    return bad ? -1 : 0;
}

/* ---- detector frames and raw counts ----------------------------------- */

/* Per-pixel gain g ~ U(g_lo, g_hi) and integer dark d ~ U{d_lo..d_hi}. */
void jh_synth_gain_dark(uint64_t seed, int npix, float g_lo, float g_hi, int d_lo, int d_hi,
                        float* gain, uint16_t* dark)
{
    for (int i = 0; i < npix; ++i) {
        rng_t r = rng_at(seed, 2, (uint64_t)i);
        gain[i] = (float)(g_lo + (g_hi - g_lo) * rng_u01(&r));
        int d = d_lo + (int)floor(rng_u01(&r) * (d_hi - d_lo + 1));
        if (d > d_hi) d = d_hi;
        dark[i] = (uint16_t)d;
    }
}

/* Poisson sample: multiplication method for small means, PTRS (Hormann,
 * "The transformed rejection method for generating Poisson random
 * variables", 1993) for mean >= 10. */
static double log_factorial(double k)
{
    return lgamma(k + 1.0);
}

static long poisson(rng_t* r, double lam)
{
    if (lam <= 0) return 0;
    if (lam < 10) {
        const double L = exp(-lam);
        long k = 0; double p = 1.0;
        do { ++k; p *= rng_u01(r); } while (p > L);
        return k - 1;
    }
    const double slam = sqrt(lam), loglam = log(lam);
    const double b = 0.931 + 2.53 * slam;
    const double a = -0.059 + 0.02483 * b;
    const double invalpha = 1.1239 + 1.1328 / (b - 3.4);
    const double vr = 0.9277 - 3.6224 / (b - 2);
    for (;;) {
        const double U = rng_u01(r) - 0.5;
        const double V = rng_u01(r);
        const double us = 0.5 - fabs(U);
        const double k = floor((2 * a / us + b) * U + lam + 0.43);
        if (us >= 0.07 && V <= vr) return (long)k;
        if (k < 0 || (us < 0.013 && V > us)) continue;
        if (log(V) + log(invalpha) - log(a / (us * us) + b) <= -lam + k * loglam - log_factorial(k))
            return (long)k;
    }
}

/* raw = clamp(Poisson(I0 * g * exp(-p)) + d, 0, 65535) for N frames of npix.
 * gain/dark may be NULL (g = 1, d = 0).  noise = 0 gives the rounded mean. */
void jh_synth_raw(uint64_t seed, const float* p, int N, int npix, double I0,
                  const float* gain, const uint16_t* dark, int noise, uint16_t* raw)
{
#pragma omp parallel for schedule(dynamic, 1)
    for (int iv = 0; iv < N; ++iv) {
        for (int i = 0; i < npix; ++i) {
            const size_t idx = (size_t)iv * npix + i;
            const double mean = I0 * (gain ? gain[i] : 1.0) * exp(-(double)p[idx]);
            long c;
            if (noise) { rng_t r = rng_at(seed, 3, idx); c = poisson(&r, mean); }
            else c = lround(mean);
            if (dark) c += dark[i];
            if (c < 0) c = 0;
            if (c > 65535) c = 65535;
            raw[idx] = (uint16_t)c;
        }
    }
}

/* Seeded uniform floats in [lo, hi) (random test inputs). */
void jh_synth_uniform(uint64_t seed, uint64_t stream, size_t n, float lo, float hi, float* out)
{
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)n; ++i) {
        rng_t r = rng_at(seed, 100 + stream, (uint64_t)i);
        out[i] = (float)(lo + (hi - lo) * rng_u01(&r));
    }
}

/* Seeded standard normal floats (Box-Muller on counter-based uniforms). */
void jh_synth_normal(uint64_t seed, uint64_t stream, size_t n, float* out)
{
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)n; ++i) {
        rng_t r = rng_at(seed, 200 + stream, (uint64_t)i);
        double u1 = rng_u01(&r), u2 = rng_u01(&r);
        if (u1 < 1e-300) u1 = 1e-300;
        out[i] = (float)(sqrt(-2 * log(u1)) * cos(2 * M_PI * u2));
    }
}
