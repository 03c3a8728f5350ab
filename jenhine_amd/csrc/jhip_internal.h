// Internal declarations shared by the translation units of libjhip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/jhip.h"

namespace jh {

// Error raised inside the library and turned into a jh_status at the ABI.
struct Error {
    int code;
    std::string msg;
};

void set_last_error(const std::string& msg);

[[noreturn]] inline void fail(int code, const std::string& msg) { throw Error{code, msg}; }

inline void check_hip(hipError_t e, const char* what)
{
    if (e != hipSuccess) fail(JH_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define JH_HIP(expr) ::jh::check_hip((expr), #expr)

inline void require(bool ok, const std::string& msg)
{
    if (!ok) fail(JH_ERR_INVALID, msg);
}

// Runs f(); converts library errors and std exceptions to a status.
template <class F>
jh_status guard(F&& f)
{
    try {
        f();
        set_last_error("");
        return JH_OK;
    } catch (const Error& e) {
        set_last_error(e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return JH_ERR_OOM;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return JH_ERR_INVALID;
    } catch (...) {
        set_last_error("unknown error");
        return JH_ERR_INVALID;
    }
}

// Bytes currently held by the library's plans, and the optional cap set by
// the environment variable JH_MAX_DEVICE_BYTES (a test hook that lets a
// host exercise its out-of-memory path on any device).
extern std::atomic<size_t> g_device_bytes;

inline size_t device_byte_cap()
{
    const char* s = std::getenv("JH_MAX_DEVICE_BYTES");
    return s ? (size_t)std::strtoull(s, nullptr, 10) : SIZE_MAX;
}

// Owned device allocation.  An allocation failure raises JH_ERR_OOM with the
// message prefix "HIP failed to malloc " (the host's part-retry keys on it).
class DevBuf {
public:
    DevBuf() = default;
    explicit DevBuf(size_t bytes) { alloc(bytes); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept
    {
        if (this != &o) { release(); p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t bytes)
    {
        release();
        if (bytes == 0) return;
        if (g_device_bytes.load() + bytes > device_byte_cap())
            fail(JH_ERR_OOM, "HIP failed to malloc " + std::to_string(bytes) + " bytes (cap)");
        if (hipMalloc(&p_, bytes) != hipSuccess || p_ == nullptr) {
            (void)hipGetLastError();
            p_ = nullptr;
            fail(JH_ERR_OOM, "HIP failed to malloc " + std::to_string(bytes) + " bytes");
        }
        n_ = bytes;
        g_device_bytes += bytes;
    }
    void release()
    {
        if (p_) {
            (void)hipFree(p_);
            g_device_bytes -= n_;
        }
        p_ = nullptr;
        n_ = 0;
    }
    template <class T>
    T* as() const { return static_cast<T*>(p_); }
    void* get() const { return p_; }
    size_t size() const { return n_; }

private:
    void* p_ = nullptr;
    size_t n_ = 0;
};

// Makes `device` current for the lifetime of the scope.
class DeviceScope {
public:
    explicit DeviceScope(int device)
    {
        JH_HIP(hipGetDevice(&prev_));
        if (device != prev_) JH_HIP(hipSetDevice(device));
        changed_ = device != prev_;
    }
    ~DeviceScope()
    {
        if (changed_) (void)hipSetDevice(prev_);
    }

private:
    int prev_ = 0;
    bool changed_ = false;
};

inline hipStream_t as_stream(jh_stream_t s) { return static_cast<hipStream_t>(s); }

inline void check_launch(const char* what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) fail(JH_ERR_HIP, std::string("launch of ") + what + ": " + hipGetErrorString(e));
}


// Host filter design (FP64): H[k], k = 0..L/2, incl. the 2*pi*f_N gain.
std::vector<double> design_filter(int W, double hsr, double cutStart, double cutEnd,
                                  const std::vector<double>& adjPts,
                                  const std::vector<double>& adjDB, int window);

// Factor of the values a backprojection plan stages in `dtype`: 2^-f16exp for
// FP16 (exact), 1 otherwise.
float bp_stage_scale(const jh_bp_plan* p, int dtype);

// Device and frame shape of a filter plan (W x V frames, nviews views).
void filter_plan_shape(const jh_filter_plan* p, int* device, int* W, int* V, int* nviews);

inline int fft_length(int W)
{
    int L = 1;
    while (L < W * 2 - 1) L *= 2;
    return L;
}

// Inverse of the 4 x 4 transform matrix of a projection matrix, in FP64:
// TM = PTM rows 0, 1, the pseudo-depth row (s20 P20, s20 P21, 0, s23 P23)
// and PTM row 2 (TransformMatrix/MatrixConvertor.cpp:8-27), inverted by
// Gauss-Jordan elimination with partial pivoting.  inv: row-major 4 x 4.
__device__ inline void tm_inverse_fp64(const double* P, double dso, double* inv)
{
    const double far_ = (1.0 + 1.0 / dso) * dso, near_ = (1.0 - 1.0 / dso) * dso;
    const double s20 = (far_ + near_) / (far_ - near_);
    const double s23 = ((far_ + near_) - 2.0 * far_ * near_ / dso) / (far_ - near_) / dso;
    double A[4][8];
    for (int c = 0; c < 4; ++c) { A[0][c] = P[c]; A[1][c] = P[4 + c]; A[3][c] = P[8 + c]; }
    A[2][0] = s20 * P[8]; A[2][1] = s20 * P[9]; A[2][2] = 0.0; A[2][3] = s23 * P[11];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) A[r][4 + c] = (r == c) ? 1.0 : 0.0;
    for (int k = 0; k < 4; ++k) {
        int p = k;
        for (int r = k + 1; r < 4; ++r)
            if (fabs(A[r][k]) > fabs(A[p][k])) p = r;
        if (p != k)
            for (int c = 0; c < 8; ++c) { double t = A[k][c]; A[k][c] = A[p][c]; A[p][c] = t; }
        const double d = A[k][k];
        for (int c = 0; c < 8; ++c) A[k][c] /= d;
        for (int r = 0; r < 4; ++r) {
            if (r == k) continue;
            const double f = A[r][k];
            for (int c = 0; c < 8; ++c) A[r][c] -= f * A[k][c];
        }
    }
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) inv[4 * r + c] = A[r][4 + c];
}

// Fused preprocessing of one pixel.  Restates
// PreProcessingAgent/PreProcessingCore.cpp:67-86 with the steps of
// GainCorr.cpp:10-13, NLog.cpp:9-13 and BeamHardeningCorr.cpp:17-27, in the
// reference's float operation order (no contraction):
//   v = raw - dark (integers, extension) ; f = max(float(v), 1)
//   crop/resize ; f /= BrightField*gain ; p = -ln f ;
//   BH: b = p*a2; b += a1; b *= p; b += a0; p *= b
// Shared by the preprocessing kernel and the kernel that writes the
// preprocessed frames straight into the filter's operand.
__device__ inline float load_corrected(const uint16_t* raw, const uint16_t* dark, size_t idx)
{
    int v = (int)raw[idx];
    if (dark) v -= (int)dark[idx];
    float f = (float)v;
    return f < 1.0f ? 1.0f : f;
}

// -ln(f / d) and the beam-hardening polynomial, d = BrightField * gain
__device__ inline float pre_log_div(float f, float d, float a0, float a1, float a2, int doBH)
{
#pragma clang fp contract(off)
    f = f / d;
    float p = logf(f);
    p = p * -1.0f;
    if (doBH) {
        float b = p * a2;
        b = b + a1;
        b = b * p;
        b = b + a0;
        p = p * b;
    }
    return p;
}

// pre_log_div with d = bf * gain[pix] (bf alone when gain is NULL)
__device__ inline float pre_log(float f, float bf, const float* gain, size_t pix, float a0, float a1, float a2,
                                int doBH)
{
#pragma clang fp contract(off)
    float d = bf;
    if (gain) d = d * gain[pix];
    return pre_log_div(f, d, a0, a1, a2, doBH);
}

}  // namespace jh

// Opaque plan types of the C-ABI.
struct jh_pre_plan {
    int device;
    int inW, inH, bUp, bDown, bLeft, bRight, outW, outH;
    float brightField;
    float bh[3];
    bool doBH;
};
