// svmi_calib.hip -- FP64 issue-rate calibration of the device (no reference
// counterpart).  The sweep kernels apply small complex matrices with plain
// v_fma_f64; a fused 16x16 block could instead run as FP64 matrix-core
// products (v_mfma_f64_16x16x4f64).  These two kernels measure both
// instruction streams at full occupancy with eight independent accumulator
// chains per lane (vector) or per wave (matrix), so neither is limited by
// the latency of one dependent chain.
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/svmi_debug.h"

namespace svmi {

namespace {

typedef double calib_d4 __attribute__((ext_vector_type(4)));

constexpr int kCalibChains = 8;

__global__ void __launch_bounds__(256) k_calib_vfma(double* out, int iters, double b, double c)
{
    double a[kCalibChains];
#pragma unroll
    for (int j = 0; j < kCalibChains; j++) a[j] = double(threadIdx.x) + j;
    for (int i = 0; i < iters; i++)
#pragma unroll
        for (int j = 0; j < kCalibChains; j++) a[j] = __builtin_fma(a[j], b, c);
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < kCalibChains; j++) s += a[j];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ void __launch_bounds__(256) k_calib_mfma(double* out, int iters, double b, double c)
{
    calib_d4 acc[kCalibChains];
#pragma unroll
    for (int j = 0; j < kCalibChains; j++) acc[j] = calib_d4{double(threadIdx.x), 1.0 + j, 2.0, 3.0};
    const double x = b + threadIdx.x, y = c;
    for (int i = 0; i < iters; i++)
#pragma unroll
        for (int j = 0; j < kCalibChains; j++) acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, acc[j], 0, 0, 0);
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < kCalibChains; j++) s += acc[j][0] + acc[j][1] + acc[j][2] + acc[j][3];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

}  // namespace

}  // namespace svmi

extern "C" svmi_status svmi_debug_fp64_rates(int device, int waves_per_simd, double* vfma_tflops,
                                             double* mfma_tflops)
{
    using namespace svmi;
    if (!vfma_tflops || !mfma_tflops || waves_per_simd < 1 || waves_per_simd > 8) return SVMI_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return SVMI_ERR_DEVICE;
    if (hipSetDevice(device) != hipSuccess) return SVMI_ERR_DEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return SVMI_ERR_DEVICE;
    // 256-thread workgroups (one wave per SIMD), waves_per_simd of them per CU
    const unsigned blocks = unsigned(prop.multiProcessorCount) * unsigned(waves_per_simd);
    double* out = nullptr;
    if (hipMalloc(&out, size_t(blocks) * 256 * sizeof(double)) != hipSuccess) return SVMI_ERR_OOM;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    svmi_status st = SVMI_OK;
    const int iv = 1 << 15, im = 1 << 12;
    for (int rep = 0; rep < 2 && st == SVMI_OK; rep++) {   // the first pass warms up
        float ms_v = 0.f, ms_m = 0.f;
        hipLaunchKernelGGL(k_calib_vfma, dim3(blocks), dim3(256), 0, nullptr, out, iv, 1.0000001, 1e-9);
        (void)hipEventRecord(e0, nullptr);
        hipLaunchKernelGGL(k_calib_vfma, dim3(blocks), dim3(256), 0, nullptr, out, iv, 1.0000001, 1e-9);
        (void)hipEventRecord(e1, nullptr);
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms_v, e0, e1) != hipSuccess)
            st = SVMI_ERR_DEVICE;
        (void)hipEventRecord(e0, nullptr);
        hipLaunchKernelGGL(k_calib_mfma, dim3(blocks), dim3(256), 0, nullptr, out, im, 1.0000001, 1e-9);
        (void)hipEventRecord(e1, nullptr);
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms_m, e0, e1) != hipSuccess)
            st = SVMI_ERR_DEVICE;
        // vector: 2 flops per FMA per lane; matrix: 16 x 16 x 4 multiply-adds per wave
        const double threads = double(blocks) * 256.0;
        *vfma_tflops = 2.0 * kCalibChains * iv * threads / (ms_v * 1e-3) / 1e12;
        *mfma_tflops = 2.0 * 16 * 16 * 4 * double(kCalibChains) * im * (threads / 64.0) / (ms_m * 1e-3) / 1e12;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(out);
    return st;
}
