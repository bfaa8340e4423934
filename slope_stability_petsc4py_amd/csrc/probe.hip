// Device probes: the FP64 vector-FMA and FP64 matrix-core (v_mfma_f64_16x16x4) rates this GPU
// attains, measured beside the kernels so their roofline fractions can be read against both the
// published peaks and what the chip delivers.  Each lane (wavefront) runs independent dependency
// chains long enough to hide the pipeline latency; the results are written out so nothing is
// eliminated.  No reference counterpart (measurement only).
#include "ssx_common.h"

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int kFmaChains = 16;
constexpr int kMfmaChains = 4;

__global__ void __launch_bounds__(256) k_probe_fma(double *__restrict__ out, double a, double b, int64_t iters) {
    double acc[kFmaChains];
#pragma unroll
    for (int k = 0; k < kFmaChains; ++k) acc[k] = threadIdx.x * 1e-3 + k;
    for (int64_t i = 0; i < iters; ++i) {
#pragma unroll
        for (int k = 0; k < kFmaChains; ++k) acc[k] = fma(acc[k], a, b);
    }
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < kFmaChains; ++k) s += acc[k];
    out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ void __launch_bounds__(256) k_probe_mfma(double *__restrict__ out, double a, double b, int64_t iters) {
    d4 acc[kMfmaChains];
#pragma unroll
    for (int k = 0; k < kMfmaChains; ++k) acc[k] = d4{1e-3 * threadIdx.x, 0.0, 0.0, (double)k};
    const double x = a * (1.0 + 1e-6 * threadIdx.x), y = b;
    for (int64_t i = 0; i < iters; ++i) {
#pragma unroll
        for (int k = 0; k < kMfmaChains; ++k) acc[k] = __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, acc[k], 0, 0, 0);
    }
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < kMfmaChains; ++k) s += acc[k][0] + acc[k][1] + acc[k][2] + acc[k][3];
    out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

}  // namespace

extern "C" {

int ssx_probe_fp64(int mode, int64_t blocks, int64_t iters, double *out, double *flops, void *stream) {
    SSX_REQUIRE((mode == 0 || mode == 1) && blocks > 0 && blocks <= 0x7fffffff && iters > 0 && out && flops,
                "ssx_probe_fp64: bad arguments");
    hipStream_t st = ssx::as_stream(stream);
    if (mode == 0) {
        hipLaunchKernelGGL(k_probe_fma, dim3((unsigned)blocks), dim3(256), 0, st, out, 1.0000001, 1e-9, iters);
        *flops = 2.0 * kFmaChains * 256.0 * (double)blocks * (double)iters;
    } else {
        hipLaunchKernelGGL(k_probe_mfma, dim3((unsigned)blocks), dim3(256), 0, st, out, 1.0000001, 1e-9, iters);
        *flops = 2.0 * 16 * 16 * 4 * kMfmaChains * 4.0 * (double)blocks * (double)iters;   // 4 wavefronts per block
    }
    SSX_CHECK_LAUNCH("ssx_probe_fp64");
    return SSX_OK;
}

}  // extern "C"
