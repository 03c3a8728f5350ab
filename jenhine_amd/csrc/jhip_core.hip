// Core of libjhip.so: error state, device queries, FP64 projection-matrix
// modifiers, the host filter design and the fused preprocessing kernel.
#include <cmath>
#include <cstring>

#include "jhip_internal.h"

namespace jh {

static thread_local std::string g_last_error;
std::atomic<size_t> g_device_bytes{0};

void set_last_error(const std::string& msg) { g_last_error = msg; }

// ---------------------------------------------------------------------------
// Filter design (host, FP64).  Restates FilterAgent/FilterCore.cpp:13-43
// (design), :69-138 (dB-adjust spline + window) and :224-242 (Ram-Lak kernel
// h[0] = 1/4, h[odd n] = -1/(pi^2 n^2), then the real part of its spectrum).
// ---------------------------------------------------------------------------

static void solve4_pivot(double A[16], double b[4])
{
    for (int k = 0; k < 4; ++k) {
        int p = k;
        for (int i = k + 1; i < 4; ++i)
            if (std::fabs(A[i * 4 + k]) > std::fabs(A[p * 4 + k])) p = i;
        if (p != k) {
            for (int j = 0; j < 4; ++j) std::swap(A[k * 4 + j], A[p * 4 + j]);
            std::swap(b[k], b[p]);
        }
        if (A[k * 4 + k] == 0.0) continue;
        for (int i = k + 1; i < 4; ++i) {
            const double l = A[i * 4 + k] / A[k * 4 + k];
            for (int j = k; j < 4; ++j) A[i * 4 + j] -= l * A[k * 4 + j];
            b[i] -= l * b[k];
        }
    }
    for (int i = 3; i >= 0; --i) {
        double s = b[i];
        for (int j = i + 1; j < 4; ++j) s -= A[i * 4 + j] * b[j];
        b[i] = s / A[i * 4 + i];
    }
}

std::vector<double> design_filter(int W, double hsr, double cutStart, double cutEnd,
                                  const std::vector<double>& adjPts,
                                  const std::vector<double>& adjDB, int window)
{
    const int L = fft_length(W);
    const int nf = L / 2 + 1;
    const double pi = 3.14159265358979323846;
    // Real spectrum of the Ram-Lak kernel: h even, so Re F[k] = sum h[n] cos.
    std::vector<double> ramlak(nf);
    for (int k = 0; k < nf; ++k) {
        double s = 0.25;
        for (int n = 1; n <= L / 2; n += 2) {
            const double h = -1.0 / (pi * pi * (double)n * (double)n);
            const double c = std::cos(2.0 * pi * (double)((long long)k * n % L) / (double)L);
            s += (n == L - n) ? h * c : 2.0 * h * c;
        }
        ramlak[k] = s;
    }
    std::vector<double> freq(nf);
    for (int i = 0; i < nf; ++i) freq[i] = (double)i * hsr * 2.0 / (double)L;
    // dB-adjust spline: knots [0, pts..., f_N], levels [0, dB..., 0] dB
    std::vector<double> x0{0.0}, y0{0.0};
    for (size_t i = 0; i < adjPts.size(); ++i) { x0.push_back(adjPts[i]); y0.push_back(adjDB[i]); }
    x0.push_back(hsr);
    y0.push_back(0.0);
    for (double& v : y0) v = std::pow(10.0, v / 20.0);
    const size_t nk = x0.size();
    std::vector<double> slope(nk, 0.0);
    for (size_t i = 1; i + 1 < nk; ++i) {
        if ((y0[i] - y0[i - 1]) * (y0[i + 1] - y0[i]) < 0) {
            const double l = x0[i] - x0[i - 1], r = x0[i + 1] - x0[i];
            const double dl = (y0[i] - y0[i - 1]) / l, dr = (y0[i + 1] - y0[i]) / r;
            slope[i] = (dl * r + dr * l) / (l + r);
        }
    }
    std::vector<double> cub(4 * (nk - 1));
    for (size_t i = 0; i + 1 < nk; ++i) {
        const double xl = x0[i], xr = x0[i + 1];
        double A[16] = {xl * xl * xl, xl * xl, xl, 1, xr * xr * xr, xr * xr, xr, 1,
                        3 * xl * xl, 2 * xl, 1, 0, 3 * xr * xr, 2 * xr, 1, 0};
        double b[4] = {y0[i], y0[i + 1], slope[i], slope[i + 1]};
        solve4_pivot(A, b);
        for (int j = 0; j < 4; ++j) cub[4 * i + j] = b[j];
    }
    std::vector<double> H(nf);
    for (int i = 0; i < nf; ++i) {
        const double f = freq[i];
        double y = 1.0;
        for (size_t j = 0; j + 1 < nk; ++j) {
            if (f >= x0[j] && f <= x0[j + 1]) {
                const double* c = &cub[4 * j];
                y = f * (f * (f * c[0] + c[1]) + c[2]) + c[3];
                break;
            }
        }
        double win;
        if (window == JH_WINDOW_SHEPP_LOGAN) {
            if (f <= hsr) {
                const double a = pi * f / (2.0 * hsr);
                win = (a == 0.0) ? 1.0 : std::sin(a) / a;
            } else {
                win = 0.0;
            }
        } else if (cutStart == 0.0 && cutEnd == 0.0) {
            win = (f < hsr) ? std::cos(f / hsr * (pi * 0.5)) : 0.0;
        } else {
            if (f <= cutStart) win = 1.0;
            else if (f < cutEnd) win = std::cos((f - cutStart) / (cutEnd - cutStart) * pi) * 0.5 + 0.5;
            else win = 0.0;
        }
        H[i] = y * win * ramlak[i] * (pi * hsr * 2.0);
    }
    return H;
}

// ---------------------------------------------------------------------------
// Projection-matrix modifiers, FP64 on the device.
// Restates TransformMatrix/ProjectionMatrixModifier.cpp:5-43.
// ---------------------------------------------------------------------------

__global__ void ptm_modify_kernel(double* ptm, int n, int op, double val)
{
    // each element is one product and one sum, rounded separately (no
    // contraction), so the device and the FP64 host restatement agree to the bit
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double* P = ptm + 12 * (size_t)i;
    switch (op) {
    case JH_PTM_MOVE_U:
        for (int j = 0; j < 4; ++j) P[j] -= P[8 + j] * val;
        break;
    case JH_PTM_MOVE_V:
        for (int j = 0; j < 4; ++j) P[4 + j] -= P[8 + j] * val;
        break;
    case JH_PTM_REVERSE_X:
        for (int r = 0; r < 3; ++r) P[4 * r] = -P[4 * r];
        break;
    case JH_PTM_REVERSE_Y:
        for (int r = 0; r < 3; ++r) P[4 * r + 1] = -P[4 * r + 1];
        break;
    case JH_PTM_MOVE_Z:
        for (int r = 0; r < 3; ++r) P[4 * r + 3] += P[4 * r + 2] * val;
        break;
    }
}

// ---------------------------------------------------------------------------
// Fused preprocessing (load_corrected, pre_log): one thread per output
// pixel.
// ---------------------------------------------------------------------------

// Grid: blockIdx.z = frame, (blockIdx.y * 4 + threadIdx.y) = output row,
// blockIdx.x * 256 + i * 64 + threadIdx.x (i < 4) = output columns: each
// load and store instruction of a wave covers 64 consecutive pixels, and no
// thread divides an index.
__global__ void __launch_bounds__(256) preprocess_kernel(
    const uint16_t* __restrict__ raw, const uint16_t* __restrict__ dark,
    const float* __restrict__ gain, float* __restrict__ out, int inW, int inH, int bUp,
    int bLeft, int roiW, int roiH, int outW, int outH, int resize, float bf, float a0, float a1,
    float a2, int doBH)
{
#pragma clang fp contract(off)
    const int y = blockIdx.y * 4 + threadIdx.y;
    if (y >= outH) return;
    const size_t view = blockIdx.z;
    const uint16_t* r = raw + view * (size_t)inW * inH;
    float* o = out + (view * outH + y) * (size_t)outW;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int x = blockIdx.x * 256 + i * 64 + threadIdx.x;
        if (x >= outW) break;
        const int pix = y * outW + x;
        float f;
        if (!resize) {
            f = load_corrected(r, dark, (size_t)(y + bUp) * inW + (x + bLeft));
        } else {
            // linear resize, pixel-centre mapping, replicated border
            // (Performance/ImgResizer.cpp:10-59)
            const float sx = (float)roiW / (float)outW, sy = (float)roiH / (float)outH;
            float fx = ((float)x + 0.5f) * sx - 0.5f;
            float fy = ((float)y + 0.5f) * sy - 0.5f;
            fx = fminf(fmaxf(fx, 0.0f), (float)(roiW - 1));
            fy = fminf(fmaxf(fy, 0.0f), (float)(roiH - 1));
            int x0 = (int)fx, y0 = (int)fy;
            x0 = min(max(x0, 0), max(roiW - 2, 0));
            y0 = min(max(y0, 0), max(roiH - 2, 0));
            const int x1 = roiW > 1 ? x0 + 1 : x0, y1 = roiH > 1 ? y0 + 1 : y0;
            const float tx = fx - (float)x0, ty = fy - (float)y0;
            const size_t r0 = (size_t)(y0 + bUp) * inW + bLeft, r1 = (size_t)(y1 + bUp) * inW + bLeft;
            const float p00 = load_corrected(r, dark, r0 + x0), p01 = load_corrected(r, dark, r0 + x1);
            const float p10 = load_corrected(r, dark, r1 + x0), p11 = load_corrected(r, dark, r1 + x1);
            const float a = p00 + tx * (p01 - p00);
            const float b = p10 + tx * (p11 - p10);
            f = a + ty * (b - a);
        }
        o[x] = pre_log(f, bf, gain, (size_t)pix, a0, a1, a2, doBH);
    }
}

// The same without resizing, 8 pixels per thread with 16-byte loads and
// stores (rows of counts, dark, gain and output aligned to 8 pixels):
// blockIdx.x * 512 + 8 threadIdx.x = first column.
__global__ void __launch_bounds__(256) preprocess_vec_kernel(
    const uint16_t* __restrict__ raw, const uint16_t* __restrict__ dark,
    const float* __restrict__ gain, float* __restrict__ out, int inW, int inH, int bUp,
    int bLeft, int outW, int outH, float bf, float a0, float a1, float a2, int doBH)
{
#pragma clang fp contract(off)
    const int y = blockIdx.y * 4 + threadIdx.y;
    const int x = (blockIdx.x * 64 + threadIdx.x) * 8;
    if (y >= outH || x >= outW) return;
    const size_t view = blockIdx.z;
    const size_t in = (size_t)(y + bUp) * inW + (x + bLeft);
    const uint4 rv = *reinterpret_cast<const uint4*>(raw + view * (size_t)inW * inH + in);
    uint4 dv = make_uint4(0u, 0u, 0u, 0u);
    if (dark) dv = *reinterpret_cast<const uint4*>(dark + in);
    float g[8] = {1.f, 1.f, 1.f, 1.f, 1.f, 1.f, 1.f, 1.f};
    if (gain) {
        const float4 g0 = *reinterpret_cast<const float4*>(gain + (size_t)y * outW + x);
        const float4 g1 = *reinterpret_cast<const float4*>(gain + (size_t)y * outW + x + 4);
        g[0] = g0.x; g[1] = g0.y; g[2] = g0.z; g[3] = g0.w; g[4] = g1.x; g[5] = g1.y; g[6] = g1.z; g[7] = g1.w;
    }
    const uint32_t rw[4] = {rv.x, rv.y, rv.z, rv.w}, dw[4] = {dv.x, dv.y, dv.z, dv.w};
    float o[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        // the steps of load_corrected and pre_log on the loaded values
        int v = (int)(uint16_t)(rw[j >> 1] >> (16 * (j & 1)));
        if (dark) v -= (int)(uint16_t)(dw[j >> 1] >> (16 * (j & 1)));
        float f = (float)v;
        f = f < 1.0f ? 1.0f : f;
        float d = bf;
        if (gain) d = d * g[j];
        o[j] = pre_log_div(f, d, a0, a1, a2, doBH);
    }
    float4* dst = reinterpret_cast<float4*>(out + (view * outH + y) * (size_t)outW + x);
    dst[0] = make_float4(o[0], o[1], o[2], o[3]);
    dst[1] = make_float4(o[4], o[5], o[6], o[7]);
}

}  // namespace jh

using namespace jh;

extern "C" {

const char* jh_last_error(void) { return g_last_error.c_str(); }

int jh_version(void) { return 100; }

int jh_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

jh_status jh_ptm_modify(double* d_ptm, int n, int op, double value, jh_stream_t stream)
{
    return guard([&] {
        require(d_ptm != nullptr && n >= 0, "jh_ptm_modify: bad arguments");
        require(op >= JH_PTM_MOVE_U && op <= JH_PTM_MOVE_Z, "jh_ptm_modify: unknown op");
        if (n == 0) return;
        ptm_modify_kernel<<<(n + 127) / 128, 128, 0, as_stream(stream)>>>(d_ptm, n, op, value);
        check_launch("ptm_modify_kernel");
    });
}

jh_status jh_pre_output_bound(const jh_pre_plan* p, double gainMin, double gainMax, double* bound)
{
    return guard([&] {
        require(p != nullptr && bound != nullptr, "jh_pre_output_bound: bad arguments");
        require(gainMin > 0.0 && gainMax >= gainMin && std::isfinite(gainMax), "jh_pre_output_bound: bad gain range");
        // the corrected count f lies in [1, 65535] (raw - dark clamped below
        // at 1; a resize interpolates between such values), so
        // p = -ln(f / (BF g)) lies in [ln(BF gmin / 65535), ln(BF gmax)]
        const double bf = (double)p->brightField;
        const double lo = std::log(bf * gainMin / 65535.0), hi = std::log(bf * gainMax);
        double m = std::max(std::fabs(lo), std::fabs(hi));
        if (p->doBH) {
            // |p (a0 + a1 p + a2 p^2)| over [lo, hi]: the cubic's extremes are
            // at the ends or where 3 a2 p^2 + 2 a1 p + a0 = 0
            const double a0 = p->bh[0], a1 = p->bh[1], a2 = p->bh[2];
            auto c = [&](double x) { return std::fabs(x * (a0 + a1 * x + a2 * x * x)); };
            m = std::max(c(lo), c(hi));
            const double A = 3.0 * a2, B = 2.0 * a1, C = a0;
            if (A != 0.0) {
                const double d = B * B - 4.0 * A * C;
                if (d >= 0.0) {
                    for (double x : {(-B + std::sqrt(d)) / (2.0 * A), (-B - std::sqrt(d)) / (2.0 * A)})
                        if (x > lo && x < hi) m = std::max(m, c(x));
                }
            } else if (B != 0.0) {
                const double x = -C / B;
                if (x > lo && x < hi) m = std::max(m, c(x));
            }
        }
        *bound = m * (1.0 + 1e-5);  // float evaluation of the chain
    });
}

jh_status jh_pre_create(jh_pre_plan** plan, int device, int inW, int inH, int bUp, int bDown,
                        int bLeft, int bRight, int outW, int outH, float brightField,
                        const float* bh)
{
    return guard([&] {
        require(plan != nullptr, "jh_pre_create: plan is NULL");
        *plan = nullptr;
        require(inW > 0 && inH > 0 && bUp >= 0 && bDown >= 0 && bLeft >= 0 && bRight >= 0,
                "jh_pre_create: bad frame or border size");
        require(inW - bLeft - bRight > 0 && inH - bUp - bDown > 0, "jh_pre_create: borders cover the frame");
        require(outW > 0 && outH > 0, "jh_pre_create: bad output size");
        require(brightField > 0.0f, "jh_pre_create: bright field must be positive");
        int n = 0;
        JH_HIP(hipGetDeviceCount(&n));
        require(device >= 0 && device < n, "jh_pre_create: no such device");
        auto* p = new jh_pre_plan;
        p->device = device;
        p->inW = inW; p->inH = inH; p->bUp = bUp; p->bDown = bDown; p->bLeft = bLeft; p->bRight = bRight;
        p->outW = outW; p->outH = outH;
        p->brightField = brightField;
        const float def[3] = {1.0f, 0.0f, 0.0f};
        const float* b = bh ? bh : def;
        for (int i = 0; i < 3; ++i) p->bh[i] = b[i];
        p->doBH = !(b[0] == 1.0f && b[1] == 0.0f && b[2] == 0.0f);
        *plan = p;
    });
}

jh_status jh_pre_run(jh_pre_plan* p, const uint16_t* d_raw, const uint16_t* d_dark,
                     const float* d_gain, float* d_out, int nviews, jh_stream_t stream)
{
    return guard([&] {
        require(p != nullptr, "jh_pre_run: plan is NULL");
        require(d_raw != nullptr && d_out != nullptr && nviews >= 0, "jh_pre_run: bad arguments");
        if (nviews == 0) return;
        DeviceScope ds(p->device);
        const int roiW = p->inW - p->bLeft - p->bRight, roiH = p->inH - p->bUp - p->bDown;
        const int resize = (roiW != p->outW || roiH != p->outH) ? 1 : 0;
        const size_t fin = (size_t)p->inW * p->inH, fout = (size_t)p->outW * p->outH;
        const bool vec = !resize && p->inW % 8 == 0 && p->bLeft % 8 == 0 && p->outW % 8 == 0 &&
                         reinterpret_cast<uintptr_t>(d_raw) % 16 == 0 && reinterpret_cast<uintptr_t>(d_out) % 16 == 0 &&
                         (d_dark == nullptr || reinterpret_cast<uintptr_t>(d_dark) % 16 == 0) &&
                         (d_gain == nullptr || reinterpret_cast<uintptr_t>(d_gain) % 16 == 0);
        for (int c0 = 0; c0 < nviews; c0 += 65535) {  // grid z: frames
            const int n = std::min(nviews - c0, 65535);
            if (vec) {
                const dim3 grid((unsigned)((p->outW + 511) / 512), (unsigned)((p->outH + 3) / 4), (unsigned)n);
                preprocess_vec_kernel<<<grid, dim3(64, 4), 0, as_stream(stream)>>>(
                    d_raw + (size_t)c0 * fin, d_dark, d_gain, d_out + (size_t)c0 * fout, p->inW, p->inH, p->bUp,
                    p->bLeft, p->outW, p->outH, p->brightField, p->bh[0], p->bh[1], p->bh[2], p->doBH ? 1 : 0);
                continue;
            }
            const dim3 grid((unsigned)((p->outW + 255) / 256), (unsigned)((p->outH + 3) / 4), (unsigned)n);
            preprocess_kernel<<<grid, dim3(64, 4), 0, as_stream(stream)>>>(
                d_raw + (size_t)c0 * fin, d_dark, d_gain, d_out + (size_t)c0 * fout, p->inW, p->inH, p->bUp,
                p->bLeft, roiW, roiH, p->outW, p->outH, resize, p->brightField, p->bh[0], p->bh[1], p->bh[2],
                p->doBH ? 1 : 0);
        }
        check_launch("preprocess_kernel");
    });
}

jh_status jh_pre_destroy(jh_pre_plan* p)
{
    return guard([&] { delete p; });
}

}  // extern "C"
