// Integer-VALU throughput probe for gfx950: measures the issue rate of
// v_mad_u64_u32 (the MAC32 of the field arithmetic) and of the companion
// carry/add instructions, so roofline fractions are quoted against a measured
// peak rather than an assumed one.  Each lane runs 8 independent dependency
// chains of the probed instruction; no memory traffic inside the timed loop.
//
//   valu_peak [iters]  ->  one JSON line per instruction:
//     {"inst": ..., "ops_per_s": ..., "per_cu_per_clk": ...}
//   valu_peak copy <MiB>  ->  one plain device-to-device copy of that size
//     (16-byte loads and stores), the calibration kernel for the FETCH_SIZE /
//     WRITE_SIZE counters: its algorithmic traffic is exactly MiB read + MiB
//     written.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CHK(x)                                                                      \
    do {                                                                            \
        hipError_t e_ = (x);                                                        \
        if (e_ != hipSuccess) {                                                     \
            std::fprintf(stderr, "%s:%d %s\n", __FILE__, __LINE__, hipGetErrorString(e_)); \
            std::exit(1);                                                           \
        }                                                                           \
    } while (0)

typedef unsigned long long u64;
typedef unsigned int u32;

template <int OP>
__global__ void __launch_bounds__(256) probe(u32* out, int iters, u32 seed) {
    u32 a = seed ^ threadIdx.x, b = seed * 3u + blockIdx.x;
    u64 acc[8];
    u32 h[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        acc[i] = (u64)(a + i) * 0x9E3779B9u;
        h[i] = a + 7 * i;
    }
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                u64 cc;
                if (OP == 0) {
                    asm volatile("v_mad_u64_u32 %0, %1, %2, %3, %0" : "+v"(acc[i]), "=s"(cc) : "v"(a), "v"(b));
                } else if (OP == 1) {
                    asm volatile("v_add_co_u32 %0, %1, %0, %2" : "+v"(h[i]), "=s"(cc) : "v"(a));
                } else if (OP == 2) {
                    asm volatile("v_mul_lo_u32 %0, %0, %1" : "+v"(h[i]) : "v"(b));
                } else if (OP == 3) {
                    asm volatile("v_mul_hi_u32 %0, %0, %1" : "+v"(h[i]) : "v"(b));
                } else if (OP == 4) {
                    asm volatile("v_lshl_add_u64 %0, %0, 0, %1" : "+v"(acc[i]) : "v"(acc[(i + 1) & 7]));
                } else if (OP == 5) {
                    asm volatile("v_add_u32 %0, %0, %1" : "+v"(h[i]) : "v"(b));
                } else if (OP == 6) {
                    asm volatile("v_mad_u32_u24 %0, %0, %1, %2" : "+v"(h[i]) : "v"(b), "v"(a));
                }
            }
        }
    }
    u32 s = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) s += (u32)acc[i] ^ (u32)(acc[i] >> 32) ^ h[i];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// Latency probe: ONE dependent v_mad_u64_u32 chain per lane, one wave per
// SIMD; cycles per step = dependent-issue latency.
__global__ void __launch_bounds__(64) probe_latency(u32* out, int iters, u32 seed) {
    u32 a = seed ^ threadIdx.x, b = seed * 3u + blockIdx.x;
    u64 acc = (u64)a * 0x9E3779B9u;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int r = 0; r < 32; ++r) {
            u64 cc;
            asm volatile("v_mad_u64_u32 %0, %1, %2, %3, %0" : "+v"(acc), "=s"(cc) : "v"(a), "v"(b));
        }
    }
    out[blockIdx.x * blockDim.x + threadIdx.x] = (u32)acc ^ (u32)(acc >> 32);
}

static void run_latency(int iters, int cus, double clk_ghz) {
    u32* out;
    const int blocks = cus * 4;   // one 64-lane wave per SIMD
    CHK(hipMalloc(&out, (size_t)blocks * 64 * 4));
    hipEvent_t e0, e1;
    CHK(hipEventCreate(&e0));
    CHK(hipEventCreate(&e1));
    hipLaunchKernelGGL(probe_latency, dim3(blocks), dim3(64), 0, 0, out, 8, 1u);
    CHK(hipDeviceSynchronize());
    CHK(hipEventRecord(e0, 0));
    hipLaunchKernelGGL(probe_latency, dim3(blocks), dim3(64), 0, 0, out, iters, 1u);
    CHK(hipEventRecord(e1, 0));
    CHK(hipEventSynchronize(e1));
    float ms = 0;
    CHK(hipEventElapsedTime(&ms, e0, e1));
    const double steps = (double)iters * 32.0;
    std::printf("{\"inst\": \"v_mad_u64_u32_dependent_chain\", \"ms\": %.3f, \"cycles_per_step\": %.2f}\n", ms,
                ms * 1e-3 * clk_ghz * 1e9 / steps);
    CHK(hipFree(out));
    CHK(hipEventDestroy(e0));
    CHK(hipEventDestroy(e1));
}

template <int OP>
static void run(const char* name, int iters, int blocks, double clk_ghz, int cus) {
    u32* out;
    CHK(hipMalloc(&out, (size_t)blocks * 256 * 4));
    hipEvent_t e0, e1;
    CHK(hipEventCreate(&e0));
    CHK(hipEventCreate(&e1));
    hipLaunchKernelGGL(probe<OP>, dim3(blocks), dim3(256), 0, 0, out, 8, 1u);   // warm-up
    CHK(hipDeviceSynchronize());
    CHK(hipEventRecord(e0, 0));
    hipLaunchKernelGGL(probe<OP>, dim3(blocks), dim3(256), 0, 0, out, iters, 1u);
    CHK(hipEventRecord(e1, 0));
    CHK(hipEventSynchronize(e1));
    float ms = 0;
    CHK(hipEventElapsedTime(&ms, e0, e1));
    const double ops = (double)blocks * 256.0 * iters * 32.0;
    const double rate = ops / (ms * 1e-3);
    std::printf("{\"inst\": \"%s\", \"ms\": %.3f, \"ops_per_s\": %.4e, \"per_cu_per_clk\": %.2f}\n", name, ms, rate,
                rate / (cus * clk_ghz * 1e9));
    CHK(hipFree(out));
    CHK(hipEventDestroy(e0));
    CHK(hipEventDestroy(e1));
}

__global__ void __launch_bounds__(256) copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

static int run_copy(int mib) {
    const size_t bytes = (size_t)mib << 20, n = bytes / 16;
    uint4 *src, *dst;
    CHK(hipMalloc(&src, bytes));
    CHK(hipMalloc(&dst, bytes));
    CHK(hipMemset(src, 1, bytes));
    hipEvent_t e0, e1;
    CHK(hipEventCreate(&e0));
    CHK(hipEventCreate(&e1));
    hipLaunchKernelGGL(copy_kernel, dim3(4096), dim3(256), 0, 0, src, dst, n);   // warm-up
    CHK(hipDeviceSynchronize());
    CHK(hipEventRecord(e0, 0));
    hipLaunchKernelGGL(copy_kernel, dim3(4096), dim3(256), 0, 0, src, dst, n);
    CHK(hipEventRecord(e1, 0));
    CHK(hipEventSynchronize(e1));
    float ms = 0;
    CHK(hipEventElapsedTime(&ms, e0, e1));
    std::printf("{\"copy_bytes\": %zu, \"ms\": %.4f, \"gbps\": %.1f}\n", bytes, ms, 2.0 * bytes / (ms * 1e-3) / 1e9);
    CHK(hipFree(src));
    CHK(hipFree(dst));
    return 0;
}

int main(int argc, char** argv) {
    if (argc > 2 && argv[1][0] == 'c') return run_copy(std::atoi(argv[2]));
    const int iters = argc > 1 ? std::atoi(argv[1]) : 4096;
    hipDeviceProp_t p;
    CHK(hipGetDeviceProperties(&p, 0));
    const double clk = p.clockRate / 1e6;   // GHz
    const int cus = p.multiProcessorCount;
    std::printf("{\"device\": \"%s\", \"arch\": \"%s\", \"cus\": %d, \"clock_ghz\": %.3f}\n", p.name, p.gcnArchName,
                cus, clk);
    const int blocks = cus * 8;   // 8 x 256 lanes per CU = 8 waves per SIMD
    run<0>("v_mad_u64_u32", iters, blocks, clk, cus);
    run<1>("v_add_co_u32", iters, blocks, clk, cus);
    run<2>("v_mul_lo_u32", iters, blocks, clk, cus);
    run<3>("v_mul_hi_u32", iters, blocks, clk, cus);
    run<4>("v_lshl_add_u64", iters, blocks, clk, cus);
    run<5>("v_add_u32", iters, blocks, clk, cus);
    run<6>("v_mad_u32_u24", iters, blocks, clk, cus);
    run_latency(iters, cus, clk);
    return 0;
}
