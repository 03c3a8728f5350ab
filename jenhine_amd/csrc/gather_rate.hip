// gather_rate: clocks per 64-lane 16-byte buffer gather per CU on the device,
// for the lane-address patterns of the backprojector's column-strip gathers
// and for reference patterns.  The backprojector's vector-memory roof uses
// the resulting rate (bench.py, roofline "gups_vmem_issue").
//
// Every workgroup sweeps a 64 KiB region of its own, so the data come from
// the caches; eight workgroups of four waves run per CU.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

#define CHECK(x)                                                                          \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess) {                                                           \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                  \
            std::exit(1);                                                                 \
        }                                                                                 \
    } while (0)

constexpr int kIters = 256;
constexpr unsigned kRegion = 65536;

__global__ void __launch_bounds__(256) gather(const char* __restrict__ buf, unsigned bytes, int pattern, unsigned* out)
{
    const int lane = threadIdx.x & 63;
    const unsigned wave = blockIdx.x * 4 + (threadIdx.x >> 6);
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<char*>(buf), (short)0, (int)bytes, 0x00020000);
    unsigned off;
    switch (pattern) {
    case 0: off = lane * 16; break;                                        // 1 KiB contiguous
    case 1: off = 0; break;                                                // every lane the same 16 bytes
    case 2: off = (lane >> 2) * 256 + (lane & 3) * 8; break;               // quads 8 bytes apart
    case 3: off = (lane & 7) * 8 + (lane >> 3) * 6152; break;              // 8 x 8 columns of a strip frame (768 wide)
    case 4: off = ((lane & 7) * 10 / 8) * 8 + (lane >> 3) * 6152; break;  // the same, 1.25 pixels per column
    default: off = lane * 64; break;                                       // one 64-byte line per lane
    }
    // offsets stay below base + kRegion + 64 KiB <= bytes
    const unsigned base = (wave * 4096u) % (bytes - 2 * kRegion);
    u32x4 acc = {0u, 0u, 0u, 0u};
    for (int it = 0; it < kIters; ++it) {
        const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(rs, base + ((it * 1024u) % kRegion) + off, 0, 0);
        acc += v;
    }
    if ((acc.x ^ acc.y ^ acc.z ^ acc.w) == 0x9e3779b9u) out[0] = acc.x;  // keeps the loads
}

int main()
{
    const unsigned bytes = 64u << 20;
    char* buf = nullptr;
    unsigned* out = nullptr;
    CHECK(hipMalloc(&buf, bytes));
    CHECK(hipMalloc(&out, 4));
    CHECK(hipMemset(buf, 1, bytes));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    int cus = 0, khz = 0;
    CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0));
    CHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeClockRate, 0));
    const int blocks = cus * 8;
    const char* names[] = {"contiguous 1 KiB", "all lanes one address", "quads 8 B apart", "strip 8x8 columns",
                           "strip 8x8, 1.25 px/column", "one line per lane"};
    std::printf("{\"cus\": %d, \"clock_mhz\": %d, \"patterns\": {", cus, khz / 1000);
    for (int p = 0; p < 6; ++p) {
        gather<<<blocks, 256>>>(buf, bytes, p, out);
        CHECK(hipGetLastError());
        const int reps = 20;
        CHECK(hipEventRecord(e0));
        for (int r = 0; r < reps; ++r) gather<<<blocks, 256>>>(buf, bytes, p, out);
        CHECK(hipEventRecord(e1));
        CHECK(hipEventSynchronize(e1));
        float ms = 0.0f;
        CHECK(hipEventElapsedTime(&ms, e0, e1));
        ms /= reps;
        const double per_cu = (double)blocks * 4 * kIters / cus;
        const double clocks = ms * 1e-3 * khz * 1e3 / per_cu;
        std::printf("%s\"%s\": %.2f", p ? ", " : "", names[p], clocks);
    }
    std::printf("}, \"unit\": \"clocks per 16-byte gather instruction per CU\"}\n");
    return 0;
}
