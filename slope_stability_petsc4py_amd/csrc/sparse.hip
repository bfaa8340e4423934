// CSR operator applies and the vector kernels of the Krylov/multigrid solvers on gfx950.
//
// Reference behaviour: the PETSc MatMult / VecDot / VecAXPY calls of linear/solver.py:3431-3451 and
// linear/deflated_fgmres.py:174-295 (classical Gram-Schmidt over the Arnoldi basis), and the
// smoother applies of linear/solver.py:155-162,417-436.
//
// SpMV: a group of G lanes per row (G a power of two picked by the host from the mean row
// length), lanes stride the row's nonzeros so each load instruction of the group reads G
// consecutive values/indices; partial sums are combined with cross-lane shuffles.  FP64 values for
// the Newton operator, FP32 values for the preconditioner; the vector precision matches the values.
#include "ssx_common.h"

#include <mutex>

namespace {

// Row dot products of CSR rows with x, G lanes per row.
//  * scalar form: lane l takes entries k0 + l, k0 + l + G, ... (one value and one index per load);
//  * vec4 form: lane l takes the aligned quadruples starting at ka + 4 l, ka + 4 l + 4 G, ... with
//    one 16-byte index load and one (FP32) or two (FP64) 16-byte value loads, the up to 3 entries
//    before the first multiple of 4 (head) and after the last (tail) by lanes 0..2.  Needs G >= 4
//    and 16-byte aligned index/value arrays (checked by the launchers).
// Both sum each lane's entries in a fixed order, so results are reproducible run to run.
template <typename T>
struct Quad;
template <>
struct Quad<float> {
    __device__ __forceinline__ static void load(const float *p, float v[4]) {
        const float4 q = *reinterpret_cast<const float4 *>(p);
        v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
    }
};
template <>
struct Quad<double> {
    __device__ __forceinline__ static void load(const double *p, double v[4]) {
        const double2 a = *reinterpret_cast<const double2 *>(p);
        const double2 b = *reinterpret_cast<const double2 *>(p + 2);
        v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
    }
};

template <typename T, int G, bool V4>
__device__ __forceinline__ T row_dot(int32_t k0, int32_t k1, int lane, const int32_t *__restrict__ indices,
                                     const T *__restrict__ val, const T *__restrict__ x) {
    T acc = 0;
    if constexpr (!V4) {
        for (int32_t k = k0 + lane; k < k1; k += G) acc += val[k] * x[indices[k]];
    } else {
        static_assert(G >= 4, "vec4 rows need at least 4 lanes");
        const int32_t ka = ((k0 + 3) & ~3) < k1 ? ((k0 + 3) & ~3) : k1;
        const int32_t kb = ka + ((k1 - ka) & ~3);
        if (k0 + lane < ka) acc += val[k0 + lane] * x[indices[k0 + lane]];
        for (int32_t k = ka + 4 * lane; k < kb; k += 4 * G) {
            const int4 c = *reinterpret_cast<const int4 *>(indices + k);
            T v[4];
            Quad<T>::load(val + k, v);
            acc += v[0] * x[c.x];
            acc += v[1] * x[c.y];
            acc += v[2] * x[c.z];
            acc += v[3] * x[c.w];
        }
        if (kb + lane < k1) acc += val[kb + lane] * x[indices[kb + lane]];
    }
    return acc;
}

// Row form: aligned quadruples (16-byte index and value loads) where the index and value arrays are
// 16-byte aligned and a row has at least 4 lanes (13-18% faster than one entry per lane on the 3D
// P2/P4 operators); one entry per lane otherwise.
__host__ inline bool vec4_ok(const void *indices, const void *val, int lanes) {
    return lanes >= 4 && ((uintptr_t)indices % 16 == 0) && ((uintptr_t)val % 16 == 0);
}

template <typename T, int G, bool V4>
__global__ void __launch_bounds__(256) k_spmv(int64_t n_rows, const int32_t *__restrict__ indptr,
                                              const int32_t *__restrict__ indices, const T *__restrict__ val,
                                              const T *__restrict__ x, T *__restrict__ y, T alpha, T beta) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t row = tid / G;
    const int lane = (int)(tid % G);
    T acc = 0;
    if (row < n_rows) acc = row_dot<T, G, V4>(indptr[row], indptr[row + 1], lane, indices, val, x);
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) acc += __shfl_down(acc, off, G);
    if (row < n_rows && lane == 0) y[row] = beta == (T)0 ? alpha * acc : alpha * acc + beta * y[row];
}

template <typename T>
int launch_spmv(int64_t n_rows, const int32_t *indptr, const int32_t *indices, const T *val, const T *x, T *y,
                T alpha, T beta, int lanes, hipStream_t st) {
    const int64_t threads = n_rows * lanes;
    const dim3 grid(ssx::grid_for(threads, 256)), block(256);
    const bool v4 = vec4_ok(indices, val, lanes);
#define SSX_SPMV_CASE(G)                                                                                          \
    case G:                                                                                                       \
        if (v4)                                                                                                   \
            hipLaunchKernelGGL((k_spmv<T, G, (G >= 4)>), grid, block, 0, st, n_rows, indptr, indices, val, x, y,  \
                               alpha, beta);                                                                      \
        else                                                                                                      \
            hipLaunchKernelGGL((k_spmv<T, G, false>), grid, block, 0, st, n_rows, indptr, indices, val, x, y,     \
                               alpha, beta);                                                                      \
        break;
    switch (lanes) {
        SSX_SPMV_CASE(1)
        SSX_SPMV_CASE(2)
        SSX_SPMV_CASE(4)
        SSX_SPMV_CASE(8)
        SSX_SPMV_CASE(16)
        SSX_SPMV_CASE(32)
        SSX_SPMV_CASE(64)
        default: ssx::set_error("ssx_spmv: lanes per row must be a power of two <= 64"); return SSX_EINVAL;
    }
#undef SSX_SPMV_CASE
    SSX_CHECK_LAUNCH("ssx_spmv");
    return SSX_OK;
}

// y = a*x + b*y
template <typename T>
__global__ void k_axpby(int64_t n, T a, const T *__restrict__ x, T b, T *__restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = b == (T)0 ? a * x[i] : a * x[i] + b * y[i];
}

// Precision casts for vectors (FP64 <-> FP32) at the preconditioner boundary.
__global__ void k_d2f(int64_t n, const double *__restrict__ x, float *__restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = (float)x[i];
}
__global__ void k_f2d(int64_t n, const float *__restrict__ x, double *__restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = (double)x[i];
}

// out[j] = sum_i V[j*ld + i] * x[i] for j < k (the classical Gram-Schmidt projections of one
// Arnoldi step): stage 1 is a (chunk, column) grid writing one partial per workgroup, stage 2 sums
// each column's partials in chunk order -- deterministic, and the grid has k * n_chunk workgroups
// instead of k.
constexpr int kDotChunk = 1024;   // vector entries per stage-1 workgroup (more where n needs over kDotMaxChunks)
constexpr int kDotMaxChunks = 64;

__global__ void __launch_bounds__(256) k_dot_cols_part(int64_t n, const double *__restrict__ V, int64_t ld,
                                                       const double *__restrict__ x, int64_t chunk,
                                                       double *__restrict__ part) {
    __shared__ double red[256];
    const int c = blockIdx.x, j = blockIdx.y;
    const double *v = V + (int64_t)j * ld;
    const int64_t i0 = (int64_t)c * chunk;
    const int64_t i1 = i0 + chunk < n ? i0 + chunk : n;
    double acc = 0.0;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) acc += v[i] * x[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(int64_t)j * gridDim.x + c] = red[0];
}

__global__ void k_dot_cols_sum(int k, int n_chunk, const double *__restrict__ part, double *__restrict__ out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= k) return;
    double acc = 0.0;
    for (int c = 0; c < n_chunk; ++c) acc += part[(int64_t)j * n_chunk + c];
    out[j] = acc;
}

// Two-stage deterministic dot for long vectors: partials per workgroup, then one final sum.
__global__ void __launch_bounds__(256) k_dot_partial(int64_t n, const double *__restrict__ x,
                                                     const double *__restrict__ y, double *__restrict__ part) {
    __shared__ double red[256];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        acc += x[i] * y[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) k_sum(int n, const double *__restrict__ part, double *__restrict__ out) {
    __shared__ double red[256];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) acc += part[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

// x -= V h  (x[i] -= sum_j V[j*ld+i] h[j]); and x += V h with sign = +1.
__global__ void k_gemv_cols(int64_t n, int k, const double *__restrict__ V, int64_t ld, const double *__restrict__ h,
                            double sign, double *__restrict__ x) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double acc = 0.0;
    for (int j = 0; j < k; ++j) acc += V[(int64_t)j * ld + i] * h[j];
    x[i] += sign * acc;
}

// One damped-Jacobi / Chebyshev step kernel: x_new = x + a*(x - x_prev) + c * dinv .* (b - A x)
// with the residual r = b - A x supplied; writes the new iterate and keeps the previous one.
template <typename T>
__global__ void k_cheb_update(int64_t n, const T *__restrict__ dinv, const T *__restrict__ r, T c, T a,
                              T *__restrict__ x, T *__restrict__ d) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const T di = a * d[i] + c * dinv[i] * r[i];
    d[i] = di;
    x[i] += di;
}

// One fused Chebyshev(Jacobi) step: per row i (G lanes, as in the SpMV)
//   r_i = b_i - (A x_in)_i  (x_in == nullptr: the zero vector, r = b)
//   d_i = a d_i + c dinv_i r_i  (first: d_i = c dinv_i r_i, the old d not read)
//   x_out_i = x_in_i + d_i
// x_in and x_out are different buffers (every row reads all of x_in while rows write x_out).
template <typename T, int G, bool V4>
__global__ void __launch_bounds__(256) k_cheb_step(int64_t n_rows, const int32_t *__restrict__ indptr,
                                                   const int32_t *__restrict__ indices, const T *__restrict__ val,
                                                   const T *__restrict__ b, const T *__restrict__ x_in,
                                                   T *__restrict__ x_out, T *__restrict__ d,
                                                   const T *__restrict__ dinv, T c, T a, int first) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t row = tid / G;
    const int lane = (int)(tid % G);
    T acc = 0;
    if (row < n_rows && x_in != nullptr) acc = row_dot<T, G, V4>(indptr[row], indptr[row + 1], lane, indices, val, x_in);
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) acc += __shfl_down(acc, off, G);
    if (row < n_rows && lane == 0) {
        const T r = b[row] - acc;
        const T di = first ? c * dinv[row] * r : a * d[row] + c * dinv[row] * r;
        d[row] = di;
        x_out[row] = (x_in != nullptr ? x_in[row] : (T)0) + di;
    }
}

template <typename T>
int launch_cheb_step(int64_t n_rows, const int32_t *indptr, const int32_t *indices, const T *val, const T *b,
                     const T *x_in, T *x_out, T *d, const T *dinv, T c, T a, int first, int lanes, hipStream_t st) {
    const dim3 grid(ssx::grid_for(n_rows * lanes, 256)), block(256);
    const bool v4 = vec4_ok(indices, val, lanes);
#define SSX_CHEB_CASE(G)                                                                                          \
    case G:                                                                                                       \
        if (v4)                                                                                                   \
            hipLaunchKernelGGL((k_cheb_step<T, G, (G >= 4)>), grid, block, 0, st, n_rows, indptr, indices, val,   \
                               b, x_in, x_out, d, dinv, c, a, first);                                             \
        else                                                                                                      \
            hipLaunchKernelGGL((k_cheb_step<T, G, false>), grid, block, 0, st, n_rows, indptr, indices, val, b,   \
                               x_in, x_out, d, dinv, c, a, first);                                                \
        break;
    switch (lanes) {
        SSX_CHEB_CASE(1)
        SSX_CHEB_CASE(2)
        SSX_CHEB_CASE(4)
        SSX_CHEB_CASE(8)
        SSX_CHEB_CASE(16)
        SSX_CHEB_CASE(32)
        SSX_CHEB_CASE(64)
        default: ssx::set_error("ssx_cheb_step: lanes per row must be a power of two <= 64"); return SSX_EINVAL;
    }
#undef SSX_CHEB_CASE
    SSX_CHECK_LAUNCH("ssx_cheb_step");
    return SSX_OK;
}

// Extract the inverse diagonal of a CSR matrix (rows with a zero diagonal get 0): kDiagLanes lanes
// per row search the row's columns together (coalesced), the one that holds the diagonal
// contributes its value to a shuffle sum (the others contribute 0).
constexpr int kDiagLanes = 16;

template <typename T>
__global__ void __launch_bounds__(256) k_inv_diag(int64_t n_rows, int64_t col0, const int32_t *__restrict__ indptr,
                                                  const int32_t *__restrict__ indices, const T *__restrict__ val,
                                                  T *__restrict__ dinv) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t r = tid / kDiagLanes;
    const int lane = (int)(tid % kDiagLanes);
    T d = 0;
    if (r < n_rows) {
        const int32_t k1 = indptr[r + 1];
        for (int32_t k = indptr[r] + lane; k < k1; k += kDiagLanes)
            if ((int64_t)indices[k] == r + col0) d = val[k];
    }
#pragma unroll
    for (int off = kDiagLanes / 2; off > 0; off >>= 1) d += __shfl_down(d, off, kDiagLanes);
    if (r < n_rows && lane == 0) dinv[r] = d != (T)0 ? (T)1 / d : (T)0;
}

}  // namespace

extern "C" {

int ssx_spmv_f64(int64_t n_rows, const int32_t *indptr, const int32_t *indices, const double *val, const double *x,
                 double *y, double alpha, double beta, int lanes, void *stream) {
    SSX_REQUIRE(n_rows >= 0, "ssx_spmv_f64: n_rows < 0");
    if (n_rows == 0) return SSX_OK;
    return launch_spmv<double>(n_rows, indptr, indices, val, x, y, alpha, beta, lanes, ssx::as_stream(stream));
}

int ssx_spmv_f32(int64_t n_rows, const int32_t *indptr, const int32_t *indices, const float *val, const float *x,
                 float *y, float alpha, float beta, int lanes, void *stream) {
    SSX_REQUIRE(n_rows >= 0, "ssx_spmv_f32: n_rows < 0");
    if (n_rows == 0) return SSX_OK;
    return launch_spmv<float>(n_rows, indptr, indices, val, x, y, alpha, beta, lanes, ssx::as_stream(stream));
}

int ssx_axpby_f64(int64_t n, double a, const double *x, double b, double *y, void *stream) {
    if (n <= 0) return SSX_OK;
    hipLaunchKernelGGL(k_axpby<double>, dim3(ssx::grid_for(n, 256)), dim3(256), 0, ssx::as_stream(stream), n, a, x, b, y);
    SSX_CHECK_LAUNCH("ssx_axpby_f64");
    return SSX_OK;
}

int ssx_axpby_f32(int64_t n, float a, const float *x, float b, float *y, void *stream) {
    if (n <= 0) return SSX_OK;
    hipLaunchKernelGGL(k_axpby<float>, dim3(ssx::grid_for(n, 256)), dim3(256), 0, ssx::as_stream(stream), n, a, x, b, y);
    SSX_CHECK_LAUNCH("ssx_axpby_f32");
    return SSX_OK;
}

int ssx_cast_f64_f32(int64_t n, const double *x, float *y, void *stream) {
    if (n <= 0) return SSX_OK;
    hipLaunchKernelGGL(k_d2f, dim3(ssx::grid_for(n, 256)), dim3(256), 0, ssx::as_stream(stream), n, x, y);
    SSX_CHECK_LAUNCH("ssx_cast_f64_f32");
    return SSX_OK;
}

int ssx_cast_f32_f64(int64_t n, const float *x, double *y, void *stream) {
    if (n <= 0) return SSX_OK;
    hipLaunchKernelGGL(k_f2d, dim3(ssx::grid_for(n, 256)), dim3(256), 0, ssx::as_stream(stream), n, x, y);
    SSX_CHECK_LAUNCH("ssx_cast_f32_f64");
    return SSX_OK;
}

int ssx_dot_cols(int64_t n, int k, const double *V, int64_t ld, const double *x, double *part, double *out,
                 void *stream) {
    SSX_REQUIRE(n >= 0 && k >= 0 && ld >= n, "ssx_dot_cols: bad sizes");
    SSX_REQUIRE(k <= 65535, "ssx_dot_cols: k above 65535 columns");
    if (k == 0) return SSX_OK;
    SSX_REQUIRE(V && x && part && out, "ssx_dot_cols: null pointer");
    int64_t n_chunk = (n + kDotChunk - 1) / kDotChunk;
    if (n_chunk < 1) n_chunk = 1;
    if (n_chunk > kDotMaxChunks) n_chunk = kDotMaxChunks;
    const int64_t chunk = (n + n_chunk - 1) / n_chunk;
    hipStream_t st = ssx::as_stream(stream);
    hipLaunchKernelGGL(k_dot_cols_part, dim3((unsigned)n_chunk, (unsigned)k), dim3(256), 0, st, n, V, ld, x,
                       chunk > 0 ? chunk : 1, part);
    hipLaunchKernelGGL(k_dot_cols_sum, dim3((unsigned)((k + 63) / 64)), dim3(64), 0, st, k, (int)n_chunk, part, out);
    SSX_CHECK_LAUNCH("ssx_dot_cols");
    return SSX_OK;
}

// out[0] = x . y ; `part` is caller scratch of at least 1024 doubles.
int ssx_dot(int64_t n, const double *x, const double *y, double *part, double *out, void *stream) {
    SSX_REQUIRE(n >= 0, "ssx_dot: n < 0");
    const int nb = (int)(n / 256 + 1 < 1024 ? n / 256 + 1 : 1024);
    hipStream_t st = ssx::as_stream(stream);
    hipLaunchKernelGGL(k_dot_partial, dim3(nb), dim3(256), 0, st, n, x, y, part);
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, st, nb, part, out);
    SSX_CHECK_LAUNCH("ssx_dot");
    return SSX_OK;
}

int ssx_gemv_cols(int64_t n, int k, const double *V, int64_t ld, const double *h, double sign, double *x,
                  void *stream) {
    SSX_REQUIRE(n >= 0 && k >= 0 && ld >= n, "ssx_gemv_cols: bad sizes");
    if (n == 0 || k == 0) return SSX_OK;
    hipLaunchKernelGGL(k_gemv_cols, dim3(ssx::grid_for(n, 256)), dim3(256), 0, ssx::as_stream(stream), n, k, V, ld, h,
                       sign, x);
    SSX_CHECK_LAUNCH("ssx_gemv_cols");
    return SSX_OK;
}

int ssx_cheb_step_f32(int64_t n_rows, const int32_t *indptr, const int32_t *indices, const float *val,
                      const float *b, const float *x_in, float *x_out, float *d, const float *dinv, float c, float a,
                      int first, int lanes, void *stream) {
    SSX_REQUIRE(n_rows >= 0, "ssx_cheb_step_f32: n_rows < 0");
    if (n_rows == 0) return SSX_OK;
    SSX_REQUIRE(indptr && indices && val && b && x_out && d && dinv && x_out != x_in,
                "ssx_cheb_step_f32: null pointer or x_out == x_in");
    return launch_cheb_step<float>(n_rows, indptr, indices, val, b, x_in, x_out, d, dinv, c, a, first, lanes,
                                   ssx::as_stream(stream));
}

int ssx_cheb_step_f64(int64_t n_rows, const int32_t *indptr, const int32_t *indices, const double *val,
                      const double *b, const double *x_in, double *x_out, double *d, const double *dinv, double c,
                      double a, int first, int lanes, void *stream) {
    SSX_REQUIRE(n_rows >= 0, "ssx_cheb_step_f64: n_rows < 0");
    if (n_rows == 0) return SSX_OK;
    SSX_REQUIRE(indptr && indices && val && b && x_out && d && dinv && x_out != x_in,
                "ssx_cheb_step_f64: null pointer or x_out == x_in");
    return launch_cheb_step<double>(n_rows, indptr, indices, val, b, x_in, x_out, d, dinv, c, a, first, lanes,
                                    ssx::as_stream(stream));
}

int ssx_cheb_update_f32(int64_t n, const float *dinv, const float *r, float c, float a, float *x, float *d,
                        void *stream) {
    if (n <= 0) return SSX_OK;
    hipLaunchKernelGGL(k_cheb_update<float>, dim3(ssx::grid_for(n, 256)), dim3(256), 0, ssx::as_stream(stream), n, dinv,
                       r, c, a, x, d);
    SSX_CHECK_LAUNCH("ssx_cheb_update_f32");
    return SSX_OK;
}

int ssx_cheb_update_f64(int64_t n, const double *dinv, const double *r, double c, double a, double *x, double *d,
                        void *stream) {
    if (n <= 0) return SSX_OK;
    hipLaunchKernelGGL(k_cheb_update<double>, dim3(ssx::grid_for(n, 256)), dim3(256), 0, ssx::as_stream(stream), n,
                       dinv, r, c, a, x, d);
    SSX_CHECK_LAUNCH("ssx_cheb_update_f64");
    return SSX_OK;
}

int ssx_inv_diag_f32(int64_t n_rows, int64_t col0, const int32_t *indptr, const int32_t *indices, const float *val,
                     float *dinv, void *stream) {
    if (n_rows <= 0) return SSX_OK;
    hipLaunchKernelGGL(k_inv_diag<float>, dim3(ssx::grid_for(n_rows * kDiagLanes, 256)), dim3(256), 0, ssx::as_stream(stream),
                       n_rows, col0, indptr, indices, val, dinv);
    SSX_CHECK_LAUNCH("ssx_inv_diag_f32");
    return SSX_OK;
}

int ssx_inv_diag_f64(int64_t n_rows, int64_t col0, const int32_t *indptr, const int32_t *indices, const double *val,
                     double *dinv, void *stream) {
    if (n_rows <= 0) return SSX_OK;
    hipLaunchKernelGGL(k_inv_diag<double>, dim3(ssx::grid_for(n_rows * kDiagLanes, 256)), dim3(256), 0, ssx::as_stream(stream),
                       n_rows, col0, indptr, indices, val, dinv);
    SSX_CHECK_LAUNCH("ssx_inv_diag_f64");
    return SSX_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------------------------
// Galerkin coarse operators of the p-multigrid, Ac = P^T A P (replaces the PETSc MatPtAP of
// linear/solver.py:510-513 feeding the coarse solve).  Dense form: one lane per nonzero a_ij of
// the fine operator scatters its products p_iI a_ij p_jJ (P has a few entries per fine row:
// p_ptr/p_idx/p_w, CSR) into the coarse matrix.  CSR form: one wavefront per coarse row gathers
// them (below).
//
// Both sum in 64-bit fixed point: each product is rounded to an integer multiple of 2^-f
// and added with an integer atomic.  Integer addition is associative, so every coarse entry is the
// same bit pattern whatever order the atomics land in (FP64 atomics give a different rounding at
// every run).  The scale comes from max|a| (an atomicMax on the IEEE bits of |a|, which order the
// same way as the values, also order-independent) and a bound `w_bound` on the weights supplied by
// the caller: every partial sum of a coarse entry is at most sum |p_iI| |a_ij| |p_jJ| <= max|a| *
// w_bound, and f is chosen so that this stays below 2^61.  The resolution, max|a| w_bound 2^-61,
// is far below the FP64 rounding of the large entries.
// ---------------------------------------------------------------------------------------------
namespace {
using u64 = unsigned long long;

__global__ void k_zero_u64(u64 *w) { w[0] = 0ull; }

__global__ void k_absmax_bits(int64_t n, const double *__restrict__ val, u64 *__restrict__ amax) {
    u64 m = 0ull;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const u64 b = (u64)__double_as_longlong(fabs(val[i]));
        m = b > m ? b : m;
    }
    // wave maximum, then one atomic per wavefront
    for (int off = 32; off > 0; off >>= 1) {
        const u64 o = __shfl_down(m, off, 64);
        m = o > m ? o : m;
    }
    if ((threadIdx.x & 63) == 0) atomicMax(amax, m);
}

// 2^f with max|a| * w_bound * 2^f < 2^61 (f may be negative); 1 when A is zero.
__device__ __forceinline__ double fixed_scale(const u64 *amax, double w_bound) {
    const double bound = __longlong_as_double((long long)amax[0]) * w_bound;
    if (!(bound > 0.0)) return 1.0;
    int e;
    frexp(bound, &e);   // bound < 2^e
    return ldexp(1.0, 61 - e);
}

__device__ __forceinline__ void add_fixed(u64 *dst, double v, double scale) {
    atomicAdd(dst, (u64)__double2ll_rn(v * scale));
}

__global__ void k_galerkin_dense(int64_t n_rows, const int32_t *__restrict__ indptr, const int32_t *__restrict__ indices,
                                 const double *__restrict__ val, const int32_t *__restrict__ p_ptr,
                                 const int32_t *__restrict__ p_idx, const double *__restrict__ p_w, int64_t n_c,
                                 const u64 *__restrict__ amax, double w_bound, u64 *__restrict__ Ac) {
    const int64_t row = (int64_t)blockIdx.x * blockDim.y + threadIdx.y;
    if (row >= n_rows) return;
    const int32_t pi0 = p_ptr[row], pi1 = p_ptr[row + 1];
    if (pi0 == pi1) return;
    const double scale = fixed_scale(amax, w_bound);
    for (int32_t k = indptr[row] + threadIdx.x; k < indptr[row + 1]; k += blockDim.x) {
        const int32_t col = indices[k];
        const double a = val[k];
        const int32_t pj0 = p_ptr[col], pj1 = p_ptr[col + 1];
        for (int32_t s = pi0; s < pi1; ++s) {
            const double ai = a * p_w[s];
            const int64_t I = p_idx[s];
            for (int32_t t = pj0; t < pj1; ++t) add_fixed(&Ac[(int64_t)p_idx[t] * n_c + I], ai * p_w[t], scale);
        }
    }
}

// Coarse operator onto a given coarse CSR pattern (columns sorted within a row), gathered per coarse
// row: one wavefront per row I takes the fine rows i with P[i][I] != 0 (the transpose of P,
// pt_ptr/pt_idx/pt_w), and for every nonzero a_ij of those rows and every P[j][J] adds
// P[i][I] a_ij P[j][J] to entry (I, J), found by binary search in the row's column list staged in
// LDS, with a 64-bit integer atomic in LDS.  The products and their fixed-point rounding are the
// scatter form's, the sums are exact integers, and each coarse entry is written once: no global
// atomics, the same bits at every call.  Products outside the pattern are dropped: for nested
// spaces (P1 inside P2 on one mesh, P the interpolation) P^T A P is the coarse stiffness matrix,
// whose sparsity is the coarse node adjacency, so those products sum to zero up to rounding.
__global__ void __launch_bounds__(256) k_galerkin_rows(
    int64_t n_c, const int32_t *__restrict__ indptr, const int32_t *__restrict__ indices,
    const double *__restrict__ val, const int32_t *__restrict__ p_ptr, const int32_t *__restrict__ p_idx,
    const double *__restrict__ p_w, const int32_t *__restrict__ pt_ptr, const int32_t *__restrict__ pt_idx,
    const double *__restrict__ pt_w, const int32_t *__restrict__ c_indptr, const int32_t *__restrict__ c_indices,
    int max_row, const u64 *__restrict__ amax, double w_bound, double *__restrict__ c_val) {
    extern __shared__ u64 smem[];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    u64 *acc = smem + (int64_t)wv * max_row;
    int32_t *cols = reinterpret_cast<int32_t *>(smem + 4 * (int64_t)max_row) + (int64_t)wv * max_row;
    const int64_t I = (int64_t)blockIdx.x * 4 + wv;
    const bool live = I < n_c;
    const int32_t r0 = live ? c_indptr[I] : 0;
    const int32_t len = live ? c_indptr[I + 1] - r0 : 0;
    const int32_t L = len < max_row ? len : max_row;   // max_row is the pattern's longest row
    for (int p = lane; p < L; p += 64) {
        cols[p] = c_indices[r0 + p];
        acc[p] = 0ull;
    }
    __syncthreads();
    const double scale = fixed_scale(amax, w_bound);
    if (live) {
        for (int32_t c = pt_ptr[I]; c < pt_ptr[I + 1]; ++c) {
            const int32_t i = pt_idx[c];
            const double wI = pt_w[c];
            for (int32_t k = indptr[i] + lane; k < indptr[i + 1]; k += 64) {
                const int32_t j = indices[k];
                const double ai = val[k] * wI;
                for (int32_t t = p_ptr[j]; t < p_ptr[j + 1]; ++t) {
                    const int32_t J = p_idx[t];
                    int32_t lo = 0, hi = L;     // first position with cols[pos] >= J
                    while (lo < hi) {
                        const int32_t mid = lo + ((hi - lo) >> 1);
                        if (cols[mid] < J) lo = mid + 1; else hi = mid;
                    }
                    if (lo < L && cols[lo] == J) atomicAdd(&acc[lo], (u64)__double2ll_rn(ai * p_w[t] * scale));
                }
            }
        }
    }
    __syncthreads();
    const double inv = 1.0 / scale;   // a power of two: exact
    for (int p = lane; p < L; p += 64) c_val[r0 + p] = (double)(long long)acc[p] * inv;
}

// In place: fixed-point sums -> FP64.
__global__ void k_fixed_to_f64(int64_t n, u64 *__restrict__ buf, const u64 *__restrict__ amax, double w_bound) {
    const double inv = 1.0 / fixed_scale(amax, w_bound);   // a power of two: exact
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const long long q = (long long)buf[i];
        reinterpret_cast<double *>(buf)[i] = (double)q * inv;
    }
}

void galerkin_prologue(int64_t nnz, const double *val, u64 *amax, hipStream_t st) {
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, st, amax);
    const int64_t want = (nnz + 255) / 256;
    const unsigned grid = (unsigned)(want < 4096 ? (want > 0 ? want : 1) : 4096);
    hipLaunchKernelGGL(k_absmax_bits, dim3(grid), dim3(256), 0, st, nnz, val, amax);
}
}  // namespace

extern "C" int ssx_galerkin_dense(int64_t n_rows, const int32_t *indptr, const int32_t *indices, const double *val,
                                  int64_t nnz, const int32_t *p_ptr, const int32_t *p_idx, const double *p_w, int64_t n_c,
                                  double w_bound, double *Ac, void *work, void *stream) {
    SSX_REQUIRE(n_rows >= 0 && n_c >= 0 && nnz >= 0, "ssx_galerkin_dense: bad sizes");
    SSX_REQUIRE(w_bound > 0.0, "ssx_galerkin_dense: w_bound must be positive");
    if (n_rows == 0 || n_c == 0) return SSX_OK;
    SSX_REQUIRE(indptr && indices && val && p_ptr && p_idx && p_w && Ac && work, "ssx_galerkin_dense: null pointer");
    hipStream_t st = ssx::as_stream(stream);
    u64 *amax = reinterpret_cast<u64 *>(work);
    galerkin_prologue(nnz, val, amax, st);
    dim3 block(16, 16);
    hipLaunchKernelGGL(k_galerkin_dense, dim3(ssx::grid_for(n_rows, 16)), block, 0, st, n_rows, indptr, indices, val,
                       p_ptr, p_idx, p_w, n_c, amax, w_bound, reinterpret_cast<u64 *>(Ac));
    hipLaunchKernelGGL(k_fixed_to_f64, dim3(ssx::grid_for(n_c * n_c < 4096 * 256 ? n_c * n_c : 4096 * 256, 256)),
                       dim3(256), 0, st, n_c * n_c, reinterpret_cast<u64 *>(Ac), amax, w_bound);
    SSX_CHECK_LAUNCH("ssx_galerkin_dense");
    return SSX_OK;
}

extern "C" int ssx_galerkin_csr(int64_t n_rows, const int32_t *indptr, const int32_t *indices, const double *val,
                                int64_t nnz, const int32_t *p_ptr, const int32_t *p_idx, const double *p_w,
                                const int32_t *pt_ptr, const int32_t *pt_idx, const double *pt_w, int64_t n_c,
                                const int32_t *c_indptr, const int32_t *c_indices, int64_t c_nnz, int max_c_row,
                                double w_bound, double *c_val, void *work, void *stream) {
    SSX_REQUIRE(n_rows >= 0 && n_c >= 0 && c_nnz >= 0 && nnz >= 0, "ssx_galerkin_csr: bad sizes");
    SSX_REQUIRE(w_bound > 0.0, "ssx_galerkin_csr: w_bound must be positive");
    SSX_REQUIRE(max_c_row >= 0 && max_c_row <= 1024, "ssx_galerkin_csr: coarse rows of %d entries (at most 1024)",
                max_c_row);
    if (n_rows == 0 || n_c == 0 || c_nnz == 0) return SSX_OK;
    SSX_REQUIRE(indptr && indices && val && p_ptr && p_idx && p_w && pt_ptr && pt_idx && pt_w && c_indptr &&
                    c_indices && c_val && work,
                "ssx_galerkin_csr: null pointer");
    hipStream_t st = ssx::as_stream(stream);
    u64 *amax = reinterpret_cast<u64 *>(work);
    galerkin_prologue(nnz, val, amax, st);
    const int L = max_c_row > 0 ? max_c_row : 1;
    const size_t lds = (size_t)4 * L * (sizeof(u64) + sizeof(int32_t));
    hipLaunchKernelGGL(k_galerkin_rows, dim3(ssx::grid_for(n_c, 4)), dim3(256), lds, st, n_c, indptr, indices, val,
                       p_ptr, p_idx, p_w, pt_ptr, pt_idx, pt_w, c_indptr, c_indices, L, amax, w_bound, c_val);
    SSX_CHECK_LAUNCH("ssx_galerkin_csr");
    return SSX_OK;
}

// ---------------------------------------------------------------------------------------------
// Multi-vector SpMV and the A-orthonormalisation of the deflation basis.
//
// SpMM: Y_j = A X_j for k vectors stored as rows (X_j = X + j*ldx); a group of G lanes per row as
// in the SpMV, KC vectors per pass, so the matrix is read once per KC vectors.
//
// A-orthonormalisation (linear/orthogonalize.py:124-170, the MATLAB A_orthogonalize): columns are
// taken newest first; column i is v = b_i - sum_j sgn_j w_j (w_j . A b_i) with the raw product
// A b_i, its signed A-norm is v . A b_i, and it is kept (w_i = v / sqrt|v . A b_i|, sgn_i = sign)
// when |v . A b_i| > tol, else dropped (w_i = 0, sgn_i = 0).  Here every A b_i comes from one SpMM
// pass before the sweep, the per-column work is four launches with no host round trip, and the
// scaling of w_i is carried as scale_i = 1/sqrt|nv_i| until one final pass (the stored rows are
// unscaled during the sweep, so the coefficient of row j is scale_j^2 sgn_j (row_j . A b_i)).
// A W is formed alongside by the same recurrence from the raw products (A w_i = scale_i (A b_i -
// sum_j h_j A row_j)), which replaces a second SpMV per kept column.
// ---------------------------------------------------------------------------------------------
namespace {

constexpr int kSpmmKC = 8;

template <int G>
__global__ void __launch_bounds__(256) k_spmm_f64(int64_t n_rows, const int32_t *__restrict__ indptr,
                                                  const int32_t *__restrict__ indices, const double *__restrict__ val,
                                                  int kc, const double *__restrict__ X, int64_t ldx,
                                                  double *__restrict__ Y, int64_t ldy) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t row = tid / G;
    const int lane = (int)(tid % G);
    double acc[kSpmmKC];
#pragma unroll
    for (int j = 0; j < kSpmmKC; ++j) acc[j] = 0.0;
    if (row < n_rows) {
        const int32_t k1 = indptr[row + 1];
        for (int32_t k = indptr[row] + lane; k < k1; k += G) {
            const double a = val[k];
            const int64_t c = indices[k];
#pragma unroll
            for (int j = 0; j < kSpmmKC; ++j)
                if (j < kc) acc[j] += a * X[(int64_t)j * ldx + c];
        }
    }
#pragma unroll
    for (int j = 0; j < kSpmmKC; ++j)
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1) acc[j] += __shfl_down(acc[j], off, G);
    if (row < n_rows && lane == 0) {
#pragma unroll
        for (int j = 0; j < kSpmmKC; ++j)
            if (j < kc) Y[(int64_t)j * ldy + row] = acc[j];
    }
}

int launch_spmm(int64_t n_rows, const int32_t *indptr, const int32_t *indices, const double *val, int k,
                const double *X, int64_t ldx, double *Y, int64_t ldy, int lanes, hipStream_t st) {
    // `lanes` is the SpMV's choice; where that reads quadruples (about mean/12 lanes) the SpMM, whose
    // entries each gather kc vectors, spreads a row over 4x as many lanes (about mean/3)
    if (vec4_ok(indices, val, lanes)) lanes = lanes * 4 < 64 ? lanes * 4 : 64;
    const dim3 grid(ssx::grid_for(n_rows * lanes, 256)), block(256);
    for (int j0 = 0; j0 < k; j0 += kSpmmKC) {
        const int kc = k - j0 < kSpmmKC ? k - j0 : kSpmmKC;
        const double *Xj = X + (int64_t)j0 * ldx;
        double *Yj = Y + (int64_t)j0 * ldy;
        switch (lanes) {
            case 1: hipLaunchKernelGGL((k_spmm_f64<1>), grid, block, 0, st, n_rows, indptr, indices, val, kc, Xj, ldx, Yj, ldy); break;
            case 2: hipLaunchKernelGGL((k_spmm_f64<2>), grid, block, 0, st, n_rows, indptr, indices, val, kc, Xj, ldx, Yj, ldy); break;
            case 4: hipLaunchKernelGGL((k_spmm_f64<4>), grid, block, 0, st, n_rows, indptr, indices, val, kc, Xj, ldx, Yj, ldy); break;
            case 8: hipLaunchKernelGGL((k_spmm_f64<8>), grid, block, 0, st, n_rows, indptr, indices, val, kc, Xj, ldx, Yj, ldy); break;
            case 16: hipLaunchKernelGGL((k_spmm_f64<16>), grid, block, 0, st, n_rows, indptr, indices, val, kc, Xj, ldx, Yj, ldy); break;
            case 32: hipLaunchKernelGGL((k_spmm_f64<32>), grid, block, 0, st, n_rows, indptr, indices, val, kc, Xj, ldx, Yj, ldy); break;
            case 64: hipLaunchKernelGGL((k_spmm_f64<64>), grid, block, 0, st, n_rows, indptr, indices, val, kc, Xj, ldx, Yj, ldy); break;
            default: ssx::set_error("ssx_spmm_f64: lanes per row must be a power of two <= 64"); return SSX_EINVAL;
        }
    }
    SSX_CHECK_LAUNCH("ssx_spmm_f64");
    return SSX_OK;
}

// W[i] = B[kmax - 1 - i] (newest column first)
__global__ void k_copy_rows_rev(int64_t n, int kmax, const double *__restrict__ B, int64_t ldb, double *__restrict__ W) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * kmax) return;
    const int64_t i = t / n, r = t - i * n;
    W[t] = B[(kmax - 1 - i) * ldb + r];
}

// h[j] = scale_j^2 sgn_j sum_c part[j*n_chunk + c]
__global__ void k_aorth_coeff(int k, int n_chunk, const double *__restrict__ part, const double *__restrict__ scale,
                              const double *__restrict__ sgn, double *__restrict__ h) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= k) return;
    double acc = 0.0;
    for (int c = 0; c < n_chunk; ++c) acc += part[(int64_t)j * n_chunk + c];
    h[j] = scale[j] * scale[j] * sgn[j] * acc;
}

// row i of W and AW: v = W_i - sum_j h_j W_j, aw = AW_i - sum_j h_j AW_j (AW_i holds A b_i on entry);
// block partials of v . (A b_i) into part
constexpr int kAorthBlocks = 1024;

__global__ void __launch_bounds__(256) k_aorth_update(int64_t n, int i, double *__restrict__ W,
                                                      double *__restrict__ AW, const double *__restrict__ h,
                                                      double *__restrict__ part) {
    __shared__ double red[256];
    double acc = 0.0;
    double *wi = W + (int64_t)i * n;
    double *awi = AW + (int64_t)i * n;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        double v = wi[r];
        const double ab = awi[r];
        double aw = ab;
        for (int j = 0; j < i; ++j) {
            const double hj = h[j];
            v -= hj * W[(int64_t)j * n + r];
            aw -= hj * AW[(int64_t)j * n + r];
        }
        wi[r] = v;
        awi[r] = aw;
        acc += v * ab;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// nv_i = sum of the partials (fixed order); keep test, sign and scale of column i
__global__ void __launch_bounds__(256) k_aorth_finish(int nb, const double *__restrict__ part, int i, double tol,
                                                      double *__restrict__ nv, double *__restrict__ sgn,
                                                      double *__restrict__ scale) {
    __shared__ double red[256];
    double acc = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) acc += part[b];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double v = red[0];
        const double a = fabs(v);
        const bool keep = a > tol;
        nv[i] = v;
        sgn[i] = keep ? (v > 0.0 ? 1.0 : -1.0) : 0.0;
        scale[i] = keep ? 1.0 / sqrt(a) : 0.0;
    }
}

__global__ void k_aorth_scale(int64_t n, int kmax, const double *__restrict__ scale, double *__restrict__ W,
                              double *__restrict__ AW) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * kmax) return;
    const double s = scale[t / n];
    W[t] *= s;
    AW[t] *= s;
}

}  // namespace

extern "C" {

int ssx_spmm_f64(int64_t n_rows, const int32_t *indptr, const int32_t *indices, const double *val, int k,
                 const double *X, int64_t ldx, double *Y, int64_t ldy, int lanes, void *stream) {
    SSX_REQUIRE(n_rows >= 0 && k >= 0 && ldx >= n_rows && ldy >= n_rows, "ssx_spmm_f64: bad sizes");
    if (n_rows == 0 || k == 0) return SSX_OK;
    SSX_REQUIRE(indptr && indices && val && X && Y, "ssx_spmm_f64: null pointer");
    return launch_spmm(n_rows, indptr, indices, val, k, X, ldx, Y, ldy, lanes, ssx::as_stream(stream));
}

}  // extern "C"

namespace {

// Body of the A-orthonormalisation; `work` holds at least 3*kmax + max(kmax * kDotMaxChunks,
// kAorthBlocks) doubles.  Every scratch word is written before it is read (h[j] by the
// coefficient kernel of step j+1, scale/sgn/nv of column i by its finish kernel, the partials by
// the kernel that produces them), so the scratch needs no clearing.
int a_orthogonalize(int64_t n, int kmax, const double *B, int64_t ldb, const int32_t *indptr, const int32_t *indices,
                    const double *val, int lanes, double tol, double *W, double *AW, double *sgn, double *work,
                    hipStream_t st, const char *name) {
    double *h = work;                 // kmax
    double *scale = work + kmax;      // kmax
    double *nv = work + 2 * kmax;     // kmax
    double *part = work + 3 * kmax;   // max(kmax * kDotMaxChunks, kAorthBlocks)
    const int64_t total = n * (int64_t)kmax;
    hipLaunchKernelGGL(k_copy_rows_rev, dim3(ssx::grid_for(total, 256)), dim3(256), 0, st, n, kmax, B, ldb, W);
    SSX_CHECK_LAUNCH(name);
    const int rc = launch_spmm(n, indptr, indices, val, kmax, W, n, AW, n, lanes, st);
    if (rc != SSX_OK) return rc;
    int64_t n_chunk = (n + kDotChunk - 1) / kDotChunk;
    if (n_chunk > kDotMaxChunks) n_chunk = kDotMaxChunks;
    const int64_t chunk = (n + n_chunk - 1) / n_chunk;
    const int nb = (int)(n / 256 + 1 < kAorthBlocks ? n / 256 + 1 : kAorthBlocks);
    for (int i = 0; i < kmax; ++i) {
        if (i > 0) {
            hipLaunchKernelGGL(k_dot_cols_part, dim3((unsigned)n_chunk, (unsigned)i), dim3(256), 0, st, n, W, n,
                               AW + (int64_t)i * n, chunk, part);
            hipLaunchKernelGGL(k_aorth_coeff, dim3((unsigned)((i + 63) / 64)), dim3(64), 0, st, i, (int)n_chunk, part,
                               scale, sgn, h);
        }
        hipLaunchKernelGGL(k_aorth_update, dim3((unsigned)nb), dim3(256), 0, st, n, i, W, AW, h, part);
        hipLaunchKernelGGL(k_aorth_finish, dim3(1), dim3(256), 0, st, nb, part, i, tol, nv, sgn, scale);
    }
    hipLaunchKernelGGL(k_aorth_scale, dim3(ssx::grid_for(total, 256)), dim3(256), 0, st, n, kmax, scale, W, AW);
    SSX_CHECK_LAUNCH(name);
    return SSX_OK;
}

// Scratch of the allocation-free A-orthonormalisation: one copy per GPU, held until the process
// ends.  Its device address comes from hipGetSymbolAddress, once per device, with that device
// current; host code uses only that address (the array's name is a host-side stand-in).
constexpr int kStaticKmax = 64;
constexpr int64_t kStaticWork =
    3 * kStaticKmax + (kStaticKmax * kDotMaxChunks > kAorthBlocks ? kStaticKmax * kDotMaxChunks : kAorthBlocks);
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

}  // namespace

static __device__ __align__(256) double scratch[kStaticWork];

namespace {
// Device address of `scratch` on the current device (looked up on the first call per device).
int scratch_address(double **out, const char *name) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        ssx::set_error("%s: hipGetDevice failed", name);
        return SSX_ERUNTIME;
    }
    if (dev < 0 || dev >= kMaxDevices) {
        ssx::set_error("%s: device %d beyond the %d devices the scratch table holds", name, dev, kMaxDevices);
        return SSX_ERUNTIME;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr) {
            ssx::set_error("%s: hipGetSymbolAddress(scratch) failed: %s", name, hipGetErrorString(e));
            return SSX_ERUNTIME;
        }
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[dev];
    return SSX_OK;
}
}  // namespace

extern "C" {

int ssx_a_orthogonalize(int64_t n, int kmax, const double *B, int64_t ldb, const int32_t *indptr,
                        const int32_t *indices, const double *val, int lanes, double tol, double *W, double *AW,
                        double *sgn, double *work, void *stream) {
    SSX_REQUIRE(n > 0 && kmax >= 0 && ldb >= n, "ssx_a_orthogonalize: bad sizes");
    SSX_REQUIRE(kmax <= 4096, "ssx_a_orthogonalize: at most 4096 basis columns");
    if (kmax == 0) return SSX_OK;
    SSX_REQUIRE(B && indptr && indices && val && W && AW && sgn && work, "ssx_a_orthogonalize: null pointer");
    return a_orthogonalize(n, kmax, B, ldb, indptr, indices, val, lanes, tol, W, AW, sgn, work,
                           ssx::as_stream(stream), "ssx_a_orthogonalize");
}

int ssx_a_orthogonalize_noalloc(int64_t n, int kmax, const double *B, int64_t ldb, const int32_t *indptr,
                                const int32_t *indices, const double *val, int lanes, double tol, double *W,
                                double *AW, double *sgn, void *stream) {
    SSX_REQUIRE(n > 0 && kmax >= 0 && ldb >= n, "ssx_a_orthogonalize_noalloc: bad sizes");
    SSX_REQUIRE(kmax <= kStaticKmax, "ssx_a_orthogonalize_noalloc: at most %d basis columns (the static scratch)",
                kStaticKmax);
    if (kmax == 0) return SSX_OK;
    SSX_REQUIRE(B && indptr && indices && val && W && AW && sgn, "ssx_a_orthogonalize_noalloc: null pointer");
    double *work = nullptr;
    const int rc = scratch_address(&work, "ssx_a_orthogonalize_noalloc");
    if (rc != SSX_OK) return rc;
    return a_orthogonalize(n, kmax, B, ldb, indptr, indices, val, lanes, tol, W, AW, sgn, work,
                           ssx::as_stream(stream), "ssx_a_orthogonalize_noalloc");
}

}  // extern "C"
