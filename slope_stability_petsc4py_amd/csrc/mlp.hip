// Per-node MLP of the learned continuation predictor: bf16 GEMMs on the matrix cores with FP32
// accumulation (v_mfma_f32_16x16x16_bf16), fused bias + tanh, the matching backward pieces and an
// Adam step on FP32 master weights.  No reference counterpart: the reference's predictors are the
// secant and reduced-Newton families (continuation/indirect.py:550-1626); this network fills the
// predictor slot of continuation/indirect.py:2703-2830.
//
// GEMM: C (M x N, FP32, row stride ldc) = op(A) (M x K) * op(B) (K x N) with bf16 operands given by
// element strides, so A^T / B^T need no copies.  One wavefront computes one 16 x 16 tile; lanes
// load 4 consecutive k of their row (A) / column (B) per 16-wide k step, out-of-range reads are
// zero.  Epilogue: + bias[j], optional tanh, optional bf16 copy of the result.
#include "ssx_common.h"

#include <hip/hip_bf16.h>

namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short f2bf(float x) {
    // round to nearest even
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u) return (short)(u >> 16);  // inf / nan
    u += 0x7fffu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

__device__ __forceinline__ float bf2f(short b) { return __uint_as_float(((uint32_t)(uint16_t)b) << 16); }

__global__ void __launch_bounds__(64) k_gemm_bf16(int64_t M, int64_t N, int64_t K, const short *__restrict__ A,
                                                  int64_t a_si, int64_t a_sk, const short *__restrict__ B,
                                                  int64_t b_sk, int64_t b_sj, float *__restrict__ C, int64_t ldc,
                                                  const float *__restrict__ bias, int act, short *__restrict__ Cb,
                                                  float beta) {
    const int lane = threadIdx.x;
    const int64_t i0 = (int64_t)blockIdx.x * 16, j0 = (int64_t)blockIdx.y * 16;
    const int rc = lane & 15, kg = lane >> 4;
    const int64_t ia = i0 + rc, jb = j0 + rc;
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    // Out-of-range operands: the index is clamped into range (every load is in bounds) and the
    // value is masked to zero, with no branch around the load.
    const int64_t iac = ia < M ? ia : M - 1, jbc = jb < N ? jb : N - 1;
    const short ma = ia < M ? (short)-1 : (short)0, mb = jb < N ? (short)-1 : (short)0;
    for (int64_t k0 = 0; k0 < K; k0 += 16) {
        s4 a, b;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int64_t k = k0 + 4 * kg + t;
            const int64_t kc = k < K ? k : K - 1;
            const short mk = k < K ? (short)-1 : (short)0;
            a[t] = (short)(A[iac * a_si + kc * a_sk] & ma & mk);
            b[t] = (short)(B[kc * b_sk + jbc * b_sj] & mb & mk);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int64_t i = i0 + 4 * kg + r, j = j0 + rc;
        if (i < M && j < N) {
            float v = acc[r];
            if (bias) v += bias[j];
            if (act == 1) v = tanhf(v);
            if (beta != 0.f) v += beta * C[i * ldc + j];
            C[i * ldc + j] = v;
            if (Cb) Cb[i * ldc + j] = f2bf(v);
        }
    }
}

// dZ = dY * (1 - Y^2) (tanh backward), also a bf16 copy for the next GEMMs.
__global__ void k_tanh_bwd(int64_t n, const float *__restrict__ dY, const float *__restrict__ Y, float *__restrict__ dZ,
                           short *__restrict__ dZb) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float v = dY[i] * (1.f - Y[i] * Y[i]);
    dZ[i] = v;
    if (dZb) dZb[i] = f2bf(v);
}

// db[j] = sum_i dZ[i][j]: one workgroup per column, fixed-order tree reduction.
__global__ void __launch_bounds__(256) k_colsum(int64_t M, int64_t N, const float *__restrict__ dZ, float *__restrict__ db) {
    __shared__ float red[256];
    const int64_t j = blockIdx.x;
    float acc = 0.f;
    for (int64_t i = threadIdx.x; i < M; i += 256) acc += dZ[i * N + j];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[j] = red[0];
}

// Adam on FP32 master weights; writes the bf16 copy used by the forward GEMMs.
__global__ void k_adam(int64_t n, float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2,
                       short *__restrict__ wb) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (wb) wb[i] = f2bf(wi);
}

__global__ void k_f32_to_bf16(int64_t n, const float *__restrict__ x, short *__restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
}

__global__ void k_bf16_to_f32(int64_t n, const short *__restrict__ x, float *__restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = bf2f(x[i]);
}

// Masked mean-squared error: g = 2 (p - t) mask / count; per-block partial losses.
__global__ void __launch_bounds__(256) k_mse(int64_t n, const float *__restrict__ p, const float *__restrict__ t,
                                             const float *__restrict__ mask, float inv_count, float *__restrict__ g,
                                             float *__restrict__ part) {
    __shared__ float red[256];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (i < n) {
        const float d = (p[i] - t[i]) * (mask ? mask[i] : 1.f);
        g[i] = 2.f * d * inv_count;
        l = d * d * inv_count;
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

}  // namespace

extern "C" {

int ssx_gemm_bf16(int64_t M, int64_t N, int64_t K, const void *A, int64_t a_si, int64_t a_sk, const void *B,
                  int64_t b_sk, int64_t b_sj, float *C, int64_t ldc, const float *bias, int act, void *Cb, float beta,
                  void *stream) {
    SSX_REQUIRE(M >= 0 && N >= 0 && K >= 0 && ldc >= N, "ssx_gemm_bf16: bad sizes");
    SSX_REQUIRE(act == 0 || act == 1, "ssx_gemm_bf16: act must be 0 (none) or 1 (tanh)");
    if (M == 0 || N == 0) return SSX_OK;
    dim3 grid((unsigned)((M + 15) / 16), (unsigned)((N + 15) / 16));
    hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(64), 0, ssx::as_stream(stream), M, N, K,
                       reinterpret_cast<const short *>(A), a_si, a_sk, reinterpret_cast<const short *>(B), b_sk, b_sj,
                       C, ldc, bias, act, reinterpret_cast<short *>(Cb), beta);
    SSX_CHECK_LAUNCH("ssx_gemm_bf16");
    return SSX_OK;
}

int ssx_tanh_bwd(int64_t n, const float *dY, const float *Y, float *dZ, void *dZb, void *stream) {
    if (n <= 0) return SSX_OK;
    hipLaunchKernelGGL(k_tanh_bwd, dim3(ssx::grid_for(n, 256)), dim3(256), 0, ssx::as_stream(stream), n, dY, Y, dZ,
                       reinterpret_cast<short *>(dZb));
    SSX_CHECK_LAUNCH("ssx_tanh_bwd");
    return SSX_OK;
}

int ssx_colsum_f32(int64_t M, int64_t N, const float *dZ, float *db, void *stream) {
    if (N <= 0) return SSX_OK;
    hipLaunchKernelGGL(k_colsum, dim3((unsigned)N), dim3(256), 0, ssx::as_stream(stream), M, N, dZ, db);
    SSX_CHECK_LAUNCH("ssx_colsum_f32");
    return SSX_OK;
}

int ssx_adam_f32(int64_t n, float *w, const float *g, float *m, float *v, float lr, float b1, float b2, float eps,
                 int step, void *wb, void *stream) {
    SSX_REQUIRE(step >= 1, "ssx_adam_f32: step counts from 1");
    if (n <= 0) return SSX_OK;
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3(ssx::grid_for(n, 256)), dim3(256), 0, ssx::as_stream(stream), n, w, g, m, v, lr,
                       b1, b2, eps, c1, c2, reinterpret_cast<short *>(wb));
    SSX_CHECK_LAUNCH("ssx_adam_f32");
    return SSX_OK;
}

int ssx_f32_to_bf16(int64_t n, const float *x, void *y, void *stream) {
    if (n <= 0) return SSX_OK;
    hipLaunchKernelGGL(k_f32_to_bf16, dim3(ssx::grid_for(n, 256)), dim3(256), 0, ssx::as_stream(stream), n, x,
                       reinterpret_cast<short *>(y));
    SSX_CHECK_LAUNCH("ssx_f32_to_bf16");
    return SSX_OK;
}

int ssx_bf16_to_f32(int64_t n, const void *x, float *y, void *stream) {
    if (n <= 0) return SSX_OK;
    hipLaunchKernelGGL(k_bf16_to_f32, dim3(ssx::grid_for(n, 256)), dim3(256), 0, ssx::as_stream(stream), n,
                       reinterpret_cast<const short *>(x), y);
    SSX_CHECK_LAUNCH("ssx_bf16_to_f32");
    return SSX_OK;
}

// loss_part: >= ceil(n/256) floats; the caller sums them.
int ssx_mse_f32(int64_t n, const float *p, const float *t, const float *mask, float inv_count, float *g,
                float *loss_part, void *stream) {
    if (n <= 0) return SSX_OK;
    hipLaunchKernelGGL(k_mse, dim3(ssx::grid_for(n, 256)), dim3(256), 0, ssx::as_stream(stream), n, p, t, mask,
                       inv_count, g, loss_part);
    SSX_CHECK_LAUNCH("ssx_mse_f32");
    return SSX_OK;
}

}  // extern "C"
