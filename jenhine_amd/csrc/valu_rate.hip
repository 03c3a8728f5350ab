// valu_rate: wave-level VALU instruction rates of the instruction types of
// the backprojector's sampling loop (v_fma_f32, v_perm_b32, v_dot2_f32_bf16,
// three-input integer add, v_add_f32, float <-> int conversions, v_fract_f32,
// v_med3_f32, v_mad_u32_u24, v_cmp_eq_f32), each as a stream of one
// instruction per step of independent chains.  Run under
// rocprofv3 --pmc with GRBM_GUI_ACTIVE, the streams give the clocks per wave
// instruction of each type (bench.py --summarize-valu), from which the
// roofline weighs the backprojector's VALU instruction mix (bench.py,
// roofline "gups_valu").
//
// Every lane runs kChains independent dependency chains for kIters rounds;
// 2048 workgroups of 4 waves (8 waves per SIMD) keep every SIMD busy, and
// each kernel runs ~20 ms after a warm-up launch, long enough for the clock
// to settle.  Rate = wave instructions of the chains / kernel time (the loop
// control is scalar).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));

#define CHECK(x)                                                                          \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess) {                                                           \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                  \
            std::exit(1);                                                                 \
        }                                                                                 \
    } while (0)

constexpr int kChains = 8;
constexpr int kIters = 4096 * 40;
constexpr int kBlocks = 256 * 8;

// OP: 0 FMA, 1 v_perm_b32, 2 v_dot2_f32_bf16, 3 three-input integer add
// (v_xad_u32), 4 v_add_f32, 5 v_cvt_i32_f32 (with an and), 6 v_fract_f32,
// 7 v_med3_f32, 8 v_mad_u32_u24, 9 v_cmp_eq_f32 (into vcc; the chains are
// the compared values, updated every 64 steps)
template <int OP>
__global__ void __launch_bounds__(256) stream(const unsigned* __restrict__ in, unsigned* __restrict__ out)
{
    const unsigned t = threadIdx.x + blockIdx.x * 256;
    unsigned x[kChains];
    for (int c = 0; c < kChains; ++c) x[c] = in[(t + c) & 1023];
    const unsigned s0 = in[1024 + (t & 63)], s1 = in[1100 + (t & 63)];
    for (int it = 0; it < kIters; ++it) {
#pragma unroll
        for (int c = 0; c < kChains; ++c) {
            if constexpr (OP == 0)
                x[c] = __float_as_uint(fmaf(__uint_as_float(x[c]), __uint_as_float(s0), __uint_as_float(s1)));
            else if constexpr (OP == 1)
                x[c] = __builtin_amdgcn_perm(x[c], s0, s1);
            else if constexpr (OP == 2)
                x[c] = __float_as_uint(__builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2_t, s0),
                                                                       __builtin_bit_cast(bf16x2_t, s1),
                                                                       __uint_as_float(x[c]), false));
            else if constexpr (OP == 3)
                x[c] = (x[c] ^ s1) + s0;
            else if constexpr (OP == 4)
                x[c] = __float_as_uint(__uint_as_float(x[c]) + __uint_as_float(s0));
            else if constexpr (OP == 5)
                x[c] = (unsigned)(int)__uint_as_float(x[c] & 0x4fffffffu);
            else if constexpr (OP == 6)
                x[c] = __float_as_uint(__builtin_amdgcn_fractf(__uint_as_float(x[c])));
            else if constexpr (OP == 7)
                x[c] = __float_as_uint(__builtin_amdgcn_fmed3f(__uint_as_float(x[c]), __uint_as_float(s0),
                                                               __uint_as_float(s1)));
            else if constexpr (OP == 8)
                x[c] = __umul24(x[c], s0) + s1;
            else
                asm volatile("v_cmp_eq_f32 vcc, %0, %1" : : "v"(x[c]), "v"(s0) : "vcc");
        }
        if constexpr (OP == 9) {
            if ((it & 63) == 0) {
#pragma unroll
                for (int c = 0; c < kChains; ++c) x[c] += s0;
            }
        }
    }
    unsigned r = 0;
    for (int c = 0; c < kChains; ++c) r ^= x[c];
    out[t] = r;
}

int main()
{
    unsigned *in, *out;
    CHECK(hipMalloc(&in, 4096 * sizeof(unsigned)));
    CHECK(hipMalloc(&out, (size_t)kBlocks * 256 * sizeof(unsigned)));
    static unsigned h[4096];
    for (int i = 0; i < 4096; ++i) h[i] = 0x3f800000u + ((i * 2654435761u) >> 12);
    CHECK(hipMemcpy(in, h, sizeof(h), hipMemcpyHostToDevice));
    const char* names[] = {"fma_f32", "perm_b32", "dot2_f32_bf16", "xad_u32", "add_f32", "cvt",
                           "fract_f32", "med3_f32", "mad_u32_u24", "cmp_f32"};
    const double per_step[] = {1, 1, 1, 1, 1, 2, 1, 1, 1, 1};  // wave instructions per chain step
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    std::printf("{\"waves\": %d, \"chains\": %d, \"iters\": %d, \"streams\": {", kBlocks * 4, kChains, kIters);
    for (int op = 0; op < 10; ++op) {
        float ms = 0.0f;
        for (int rep = 0; rep < 2; ++rep) {  // the first launch warms the clock up
            CHECK(hipEventRecord(e0));
            switch (op) {
            case 0: stream<0><<<kBlocks, 256>>>(in, out); break;
            case 1: stream<1><<<kBlocks, 256>>>(in, out); break;
            case 2: stream<2><<<kBlocks, 256>>>(in, out); break;
            case 3: stream<3><<<kBlocks, 256>>>(in, out); break;
            case 4: stream<4><<<kBlocks, 256>>>(in, out); break;
            case 5: stream<5><<<kBlocks, 256>>>(in, out); break;
            case 6: stream<6><<<kBlocks, 256>>>(in, out); break;
            case 7: stream<7><<<kBlocks, 256>>>(in, out); break;
            case 8: stream<8><<<kBlocks, 256>>>(in, out); break;
            default: stream<9><<<kBlocks, 256>>>(in, out); break;
            }
            CHECK(hipGetLastError());
            CHECK(hipEventRecord(e1));
            CHECK(hipEventSynchronize(e1));
            CHECK(hipEventElapsedTime(&ms, e0, e1));
        }
        const double steps = (double)kBlocks * 4 * kIters * kChains;
        std::printf("%s\"%s\": {\"ms\": %.3f, \"wave_instr_per_s\": %.4e}", op ? ", " : "", names[op], ms,
                    steps * per_step[op] / (ms * 1e-3));
    }
    std::printf("}}\n");
    return 0;
}
