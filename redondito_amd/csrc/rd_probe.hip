// FP64 throughput ceilings of the current device: register-only loops of
// v_mfma_f64_16x16x4_f64 (four independent accumulators per wave), of
// v_mfma_f64_4x4x4_4b_f64 (eight) and of v_fma_f64 (sixteen independent
// chains per lane), one double stored per thread at the end.  They set the
// measured ceilings beside the published FP64 peak for the Legendre kernels,
// which run on the 16x16x4 FP64 MFMA.
#include "rd_internal.h"

namespace rd {

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void probe_mfma_kernel(double *out, int iters, double a0)
{
    v4d acc[4];
#pragma unroll
    for (int i = 0; i < 4; i++) acc[i] = (v4d){0.0, 0.0, 0.0, 0.0};
    const double a = a0 + threadIdx.x * 1e-9, b = a0 - threadIdx.x * 1e-9;
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int i = 0; i < 4; i++) acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[i], 0, 0, 0);
    }
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < 4; i++) s += acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3];
    out[(long)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// v_mfma_f64_4x4x4_4b_f64: four 4x4x4 blocks per instruction, 512 flops;
// eight independent accumulators per wave
__global__ __launch_bounds__(256) void probe_mfma4_kernel(double *out, int iters, double a0)
{
    double acc[8];
#pragma unroll
    for (int i = 0; i < 8; i++) acc[i] = 0.0;
    const double a = a0 + threadIdx.x * 1e-9, b = a0 - threadIdx.x * 1e-9;
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int i = 0; i < 8; i++) acc[i] = __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, acc[i], 0, 0, 0);
    }
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < 8; i++) s += acc[i];
    out[(long)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void probe_valu_kernel(double *out, int iters, double a0)
{
    double x[16];
#pragma unroll
    for (int i = 0; i < 16; i++) x[i] = a0 + i * 1e-3 + threadIdx.x * 1e-9;
    const double m = 0.999999, c = 1e-7;
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int i = 0; i < 16; i++) x[i] = __builtin_fma(x[i], m, c);
    }
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < 16; i++) s += x[i];
    out[(long)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

}  // namespace

}  // namespace rd

using namespace rd;

namespace {

// Times the register-only loops: kind 0 the 16x16x4 MFMA, 1 the FMA, 2 the
// 4x4x4 MFMA, each over eight workgroups of four waves per CU.
int time_probes(const int *kinds, double *const *res, int n)
{
    int dev = 0, ncu = 0;
    RD_HIP(hipGetDevice(&dev));
    RD_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    // eight workgroups of four waves per CU: every SIMD holds several waves
    const int grid = 8 * ncu, iters = 4096, reps = 5;
    double *d = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = RD_OK;
    auto check = [&](hipError_t e) {
        if (e != hipSuccess && rc == RD_OK) {
            fprintf(stderr, "redondito: %s\n", hipGetErrorString(e));
            rc = RD_ERR_HIP;
        }
        return rc == RD_OK;
    };
    if (check(hipMalloc(&d, sizeof(double) * grid * 256)) && check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) &&
        check(hipEventCreate(&e0)) && check(hipEventCreate(&e1))) {
        // flops per thread and iteration: 4 MFMAs of 2048 flops over 64 lanes;
        // 64 FMAs of 2 flops; 8 MFMAs of 512 flops over 64 lanes
        const double fl[3] = {4 * 32.0, 64 * 2.0, 8 * 8.0};
        for (int q = 0; q < n && rc == RD_OK; q++) {
            const int k = kinds[q];
            auto launch = [&]() {
                if (k == 0)
                    hipLaunchKernelGGL(probe_mfma_kernel, dim3(grid), dim3(256), 0, st, d, iters, 1.0);
                else if (k == 1)
                    hipLaunchKernelGGL(probe_valu_kernel, dim3(grid), dim3(256), 0, st, d, iters, 1.0);
                else
                    hipLaunchKernelGGL(probe_mfma4_kernel, dim3(grid), dim3(256), 0, st, d, iters, 1.0);
            };
            launch();
            if (!check(hipGetLastError()) || !check(hipStreamSynchronize(st))) break;
            check(hipEventRecord(e0, st));
            for (int r = 0; r < reps; r++) launch();
            check(hipEventRecord(e1, st));
            if (!check(hipEventSynchronize(e1))) break;
            float ms = 0.0f;
            check(hipEventElapsedTime(&ms, e0, e1));
            *res[q] = fl[k] * iters * (double)grid * 256.0 / (ms / reps * 1e-3) / 1e12;
        }
    }
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (st) hipStreamDestroy(st);
    if (d) hipFree(d);
    return rc;
}

}  // namespace

extern "C" int rd_fp64_ceilings(double *mfma_tflops, double *valu_tflops)
{
    if (!mfma_tflops || !valu_tflops) return RD_ERR_ARG;
    const int kinds[2] = {0, 1};
    double *res[2] = {mfma_tflops, valu_tflops};
    return time_probes(kinds, res, 2);
}

extern "C" int rd_fp64_mfma4x4_ceiling(double *tflops)
{
    if (!tflops) return RD_ERR_ARG;
    const int kinds[1] = {2};
    double *res[1] = {tflops};
    return time_probes(kinds, res, 1);
}
