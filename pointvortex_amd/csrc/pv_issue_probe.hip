// Issue cost of single VALU instructions on the device, the figures behind the pair kernels'
// instruction-count floors.  Each kernel runs one instruction kind (kOps) in 8 independent
// chains per thread for kIter iterations on 16 blocks of 4 waves per CU, so the SIMDs always have an instruction of
// the chain ready; its cost is the SIMD cycles per wave64 instruction:
//     cycles = time x clock x (4 SIMDs x CUs) / (waves x kIter x 8 x instructions per chain).
// The clock is the one the kernel ran at: every block's first thread reads the core cycle
// counter (clock64) and the constant-rate counter (wall_clock64) at its start and end, and the
// clock is their summed ratio times the constant rate.  The cost at the device's nominal clock
// is printed beside it.  Built as pointvortex_amd/lib/pv_issue_probe by pointvortex_amd/build.py.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <utility>
#include <vector>

namespace {

constexpr int kIter = 4096;
constexpr int kChains = 8;
constexpr int kBlock = 256;

struct Times {
    unsigned long long core0, core1, wall0, wall1;
};

template <int OP>
__global__ __launch_bounds__(kBlock) void probe(double *out, Times *times, double b, double c,
                                                float bf, float cf) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    Times tm{};
    if (threadIdx.x == 0) {
        tm.core0 = clock64();
        tm.wall0 = wall_clock64();
    }
    // Per-thread operands in VGPRs.  A kernel argument sits in an SGPR, which for FP32
    // instructions costs more: the "SGPR operand" rows use b and c themselves.
    double a[kChains], bv[kChains], cv[kChains];
    float f[kChains], g[kChains], fb[kChains], fc[kChains];
#pragma unroll
    for (int k = 0; k < kChains; ++k) {
        bv[k] = b + 1e-9 * (t & 7) * k;
        cv[k] = c * (1 + (t & 3));
        fb[k] = bf + 1e-6f * (float)((t & 7) * k);
        fc[k] = cf * (float)(1 + (t & 3));
        a[k] = 1.0 + 1e-3 * (t + k);
        f[k] = 1.0f + 1e-3f * (float)(t + k);
        g[k] = 2.0f - 1e-3f * (float)(t + k);
    }
    for (int i = 0; i < kIter; ++i) {
#pragma unroll
        for (int k = 0; k < kChains; ++k) {
            if constexpr (OP == 0) a[k] = __builtin_fma(a[k], bv[k], cv[k]);
            else if constexpr (OP == 1) a[k] = __builtin_fma(bv[k], cv[k], a[k]);
            else if constexpr (OP == 2) a[k] = a[k] * bv[k];
            else if constexpr (OP == 3) a[k] = a[k] + bv[k];
            else if constexpr (OP == 4) a[k] = __builtin_amdgcn_rcp(a[k]);
            else if constexpr (OP == 5) f[k] = __builtin_fmaf(f[k], fb[k], fc[k]);
            else if constexpr (OP == 6) f[k] = __builtin_fmaf(fb[k], fc[k], f[k]);
            else if constexpr (OP == 7) f[k] = f[k] * fb[k];
            else if constexpr (OP == 8) f[k] = f[k] + fb[k];
            else if constexpr (OP == 9) f[k] = __builtin_fmaf(f[k], bf, cf);
            else if constexpr (OP == 10) f[k] = f[k] * bf;
            else if constexpr (OP == 11) f[k] = __builtin_amdgcn_rcpf(f[k]);
            else if constexpr (OP == 12) {
                // The add keeps the round trip from folding away (a float widened and narrowed
                // again is itself).
                f[k] = (float)a[k];
                a[k] = (double)f[k] + bv[k];
            } else if constexpr (OP == 13)
                f[k] = __int_as_float(
                    __builtin_amdgcn_mov_dpp(__float_as_int(f[k]), 0x13C, 0xF, 0xF, true));
            else if constexpr (OP == 14) {
                a[k] = __builtin_fma(a[k], bv[k], cv[k]);
                f[k] = __builtin_fmaf(f[k], fb[k], fc[k]);
            } else if constexpr (OP == 15) {
                f[k] = __builtin_fmaf(f[k], fb[k], fc[k]);
                g[k] = __builtin_fmaf(g[k], fb[k], fc[k]);
            }
        }
    }
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < kChains; ++k)
        s += a[k] + (double)f[k] + (double)g[k];
    out[t] = s;
    if (threadIdx.x == 0) {
        tm.core1 = clock64();
        tm.wall1 = wall_clock64();
        times[blockIdx.x] = tm;
    }
}

struct Op {
    const char *name;
    int per_chain;   // instructions per chain step
};

const Op kOps[] = {
    {"v_fma_f64 (VGPR operands)", 1},
    {"v_fmac_f64 (accumulator form)", 1},
    {"v_mul_f64", 1},
    {"v_add_f64", 1},
    {"v_rcp_f64", 1},
    {"v_fma_f32 (VGPR operands)", 1},
    {"v_fmac_f32 (accumulator form)", 1},
    {"v_mul_f32 (VGPR)", 1},
    {"v_add_f32 (VGPR)", 1},
    {"v_fma_f32, SGPR operands", 1},
    {"v_mul_f32, SGPR operand", 1},
    {"v_rcp_f32", 1},
    {"cvt_f32_f64, cvt_f64_f32, add_f64", 3},
    {"v_mov_b32 DPP wave_ror:1", 1},
    {"v_fma_f64 + v_fma_f32 interleaved", 2},
    {"2 chains of v_fma_f32", 2},
};
constexpr int kNumOps = sizeof(kOps) / sizeof(kOps[0]);

void check(hipError_t e, const char *what) {
    if (e != hipSuccess) {
        std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
        std::exit(1);
    }
}

template <int OP>
void run(double *out, Times *times_d, int blocks, std::vector<Times> &times, float &ms) {
    hipEvent_t e0, e1;
    check(hipEventCreate(&e0), "hipEventCreate");
    check(hipEventCreate(&e1), "hipEventCreate");
    hipLaunchKernelGGL(probe<OP>, dim3(blocks), dim3(kBlock), 0, 0, out, times_d, 0.999, 1e-3,
                       0.999f, 1e-3f);   // warm-up
    check(hipEventRecord(e0), "hipEventRecord");
    hipLaunchKernelGGL(probe<OP>, dim3(blocks), dim3(kBlock), 0, 0, out, times_d, 0.999, 1e-3,
                       0.999f, 1e-3f);
    check(hipGetLastError(), "launch probe");
    check(hipEventRecord(e1), "hipEventRecord");
    check(hipEventSynchronize(e1), "hipEventSynchronize");
    check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    check(hipMemcpy(times.data(), times_d, times.size() * sizeof(Times), hipMemcpyDeviceToHost),
          "copy block times");
    hipEventDestroy(e0);
    hipEventDestroy(e1);
}

template <int... OPS>
void run_all(std::integer_sequence<int, OPS...>, double *out, Times *times_d, int blocks,
             std::vector<std::vector<Times>> &times, std::vector<float> &ms) {
    (run<OPS>(out, times_d, blocks, times[OPS], ms[OPS]), ...);
}

} // namespace

int main() {
    hipDeviceProp_t p;
    check(hipGetDeviceProperties(&p, 0), "hipGetDeviceProperties");
    int wall_khz = 0;
    check(hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, 0),
          "wall clock rate");
    const int cus = p.multiProcessorCount;
    const double nominal_ghz = p.clockRate / 1e6;
    const int blocks = cus * 16;
    double *out = nullptr;
    Times *times_d = nullptr;
    check(hipMalloc(&out, (size_t)blocks * kBlock * sizeof(double)), "hipMalloc");
    check(hipMalloc(&times_d, (size_t)blocks * sizeof(Times)), "hipMalloc");
    std::vector<std::vector<Times>> times(kNumOps, std::vector<Times>(blocks));
    std::vector<float> ms(kNumOps);
    run_all(std::make_integer_sequence<int, kNumOps>{}, out, times_d, blocks, times, ms);
    std::printf("device %s, %d CUs, nominal clock %.3f GHz, constant counter %.1f MHz\n",
                p.gcnArchName, cus, nominal_ghz, wall_khz / 1e3);
    std::printf("%-36s %9s %10s %14s %16s\n", "instruction", "ms", "clock GHz", "cycles/instr",
                "at nominal clock");
    const double waves = (double)blocks * (kBlock / 64);
    const double simds = cus * 4.0;
    for (int k = 0; k < kNumOps; ++k) {
        double core = 0.0, wall = 0.0;
        for (const Times &t : times[k]) {
            core += (double)(t.core1 - t.core0);
            wall += (double)(t.wall1 - t.wall0);
        }
        const double ghz = wall > 0.0 ? core / wall * wall_khz / 1e6 : 0.0;
        const double instr = waves * kIter * kChains * kOps[k].per_chain;
        const double cyc = ms[k] * 1e-3 * ghz * 1e9 * simds / instr;
        const double cyc_nominal = ms[k] * 1e-3 * nominal_ghz * 1e9 * simds / instr;
        std::printf("%-36s %9.3f %10.3f %14.2f %16.2f\n", kOps[k].name, ms[k], ghz, cyc,
                    cyc_nominal);
    }
    hipFree(out);
    hipFree(times_d);
    return 0;
}
