// Coarse-level solve of the p-multigrid: y = Ainv x with the explicit inverse of the (symmetric)
// Galerkin coarse operator, stored once per coarse rebuild as the upper triangle of 64 x 64 tiles.
//
// A plain dense matrix-vector product reads all n^2 entries per apply; the coarse solve runs once
// per Krylov iteration and was the largest single kernel of a solve (11.5k coarse dofs: 532 MB of
// FP32 per apply).  With the symmetric tile store every stored tile (I, J), I <= J, serves both
// y_I += A_IJ x_J and y_J += A_IJ^T x_I, so an apply reads n^2/2 entries.  Each workgroup takes one
// tile, stages it in LDS and writes its one or two 64-value partial results to part[I][J] (the
// contribution of tile column J to row block I); a second launch sums part[I][:] in a fixed order
// -- deterministic, no atomics.
//
// Reference behaviour: the coarse solve of the manual p-multigrid V-cycle (linear/solver.py:97-160,
// a Hypre BoomerAMG coarse solve there; here the coarse problem is solved exactly).
#include "ssx_common.h"

namespace {

constexpr int kT = 64;          // tile edge
constexpr int kTP = kT + 1;     // LDS row stride (column reads without bank conflicts)

__device__ __forceinline__ void tile_coords(int64_t t, int64_t &I, int64_t &J) {
    // t enumerates the upper triangle column by column: J = 0: (0,0); J = 1: (0,1), (1,1); ...
    int64_t j = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((j + 1) * (j + 2) / 2 <= t) ++j;
    while (j * (j + 1) / 2 > t) --j;
    J = j;
    I = t - j * (j + 1) / 2;
}

// tiles[t][r][c] = A[(I*T + r) * n + J*T + c] (zero outside n x n); A is symmetric, so its storage
// order (row- or column-major) does not matter.
template <typename Tin, typename Tout>
__global__ void k_sym_pack(int64_t n, int64_t n_tiles, const Tin *__restrict__ A, Tout *__restrict__ tiles) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n_tiles * kT * kT) return;
    const int64_t t = g / (kT * kT);
    const int rc = (int)(g - t * kT * kT);
    const int r = rc / kT, c = rc - (rc / kT) * kT;
    int64_t I, J;
    tile_coords(t, I, J);
    const int64_t i = I * kT + r, j = J * kT + c;
    const bool in = i < n && j < n;
    const int64_t off = in ? i * n + j : 0;
    tiles[g] = in ? (Tout)A[off] : (Tout)0;
}

template <typename T>
struct Vec4;
template <>
struct Vec4<float> {
    using type = float4;
};
template <>
struct Vec4<double> {
    using type = double4;
};

template <typename T>
__global__ void __launch_bounds__(256) k_sym_tiles_gemv(int64_t n, int64_t nt, const T *__restrict__ tiles,
                                                       const T *__restrict__ x, T *__restrict__ part) {
    using V = typename Vec4<T>::type;
    __shared__ T L[kT * kTP];
    __shared__ T xs[2 * kT];
    const int64_t t = blockIdx.x;
    int64_t I, J;
    tile_coords(t, I, J);
    // the tile: 1024 four-wide loads, four per thread, coalesced
    const V *src = reinterpret_cast<const V *>(tiles + t * kT * kT);
    V v[kT * kT / 4 / 256];
#pragma unroll
    for (int u = 0; u < kT * kT / 4 / 256; ++u) v[u] = src[threadIdx.x + 256 * u];
    if (threadIdx.x < 2 * kT) {
        const int64_t idx = threadIdx.x < kT ? J * kT + threadIdx.x : I * kT + (threadIdx.x - kT);
        const bool in = idx < n;
        xs[threadIdx.x] = in ? x[in ? idx : 0] : (T)0;
    }
#pragma unroll
    for (int u = 0; u < kT * kT / 4 / 256; ++u) {
        const int k = 4 * (threadIdx.x + 256 * u);
        T *d = L + (k / kT) * kTP + (k % kT);
        d[0] = v[u].x;
        d[1] = v[u].y;
        d[2] = v[u].z;
        d[3] = v[u].w;
    }
    __syncthreads();
    // two lanes per output (64 rows against x_J, then 64 columns against x_I): lane q sums entries
    // q, q + 2, ... and the pair is combined with one shuffle
    const int tid = threadIdx.x;
    const int o = tid >> 1, q = tid & 1;
    T acc = 0;
    if (o < kT) {
#pragma unroll 8
        for (int c = q; c < kT; c += 2) acc += L[o * kTP + c] * xs[c];
    } else {
        const int c = o - kT;
#pragma unroll 8
        for (int r = q; r < kT; r += 2) acc += L[r * kTP + c] * xs[kT + r];
    }
    acc += __shfl_xor(acc, 1, 2);
    if (q == 0) {
        if (o < kT)
            part[(I * nt + J) * kT + o] = acc;
        else if (I != J)
            part[(J * nt + I) * kT + (o - kT)] = acc;
    }
}

// y for row block I: 256 threads = 64 rows x 4 groups of tile columns (J = g, g + 4, ...); the four
// group sums are combined in LDS in a fixed order
template <typename T>
__global__ void __launch_bounds__(256) k_sym_tiles_sum(int64_t n, int64_t nt, const T *__restrict__ part,
                                                      T *__restrict__ y) {
    __shared__ T red[4][kT];
    const int64_t I = blockIdx.x;
    const int r = threadIdx.x & (kT - 1), g = threadIdx.x >> 6;
    T acc = 0;
#pragma unroll 4
    for (int64_t J = g; J < nt; J += 4) acc += part[(I * nt + J) * kT + r];
    red[g][r] = acc;
    __syncthreads();
    const int64_t i = I * kT + threadIdx.x;
    if (threadIdx.x < kT && i < n) y[i] = (red[0][threadIdx.x] + red[1][threadIdx.x]) + (red[2][threadIdx.x] + red[3][threadIdx.x]);
}

template <typename T>
int sym_gemv(int64_t n, const T *tiles, const T *x, T *part, T *y, void *stream) {
    SSX_REQUIRE(n >= 0, "ssx_sym_tiles_gemv: n < 0");
    if (n == 0) return SSX_OK;
    SSX_REQUIRE(tiles && x && part && y, "ssx_sym_tiles_gemv: null pointer");
    const int64_t nt = (n + kT - 1) / kT;
    const int64_t n_tiles = nt * (nt + 1) / 2;
    SSX_REQUIRE(n_tiles <= 0x7fffffff, "ssx_sym_tiles_gemv: coarse size too large");
    hipStream_t st = ssx::as_stream(stream);
    hipLaunchKernelGGL((k_sym_tiles_gemv<T>), dim3((unsigned)n_tiles), dim3(256), 0, st, n, nt, tiles, x, part);
    hipLaunchKernelGGL((k_sym_tiles_sum<T>), dim3((unsigned)nt), dim3(256), 0, st, n, nt, part, y);
    SSX_CHECK_LAUNCH("ssx_sym_tiles_gemv");
    return SSX_OK;
}

// ---------------------------------------------------------------------------------------------
// Explicit inverse of the symmetric positive definite coarse operator, straight into the symmetric
// tile store above: A = L L^T (blocked right-looking Cholesky on 64 x 64 tiles), W = L^-1 (blocked
// right-looking triangular inverse), then A^-1 = W^T W for the upper tiles only.  About n^3 flops
// in all (n^3/3 per stage), in T = float or double, every tile product done by one workgroup from
// LDS (tile_mm).  The matrix is padded to a multiple of 64 with an identity block (its inverse is
// then the identity there, and the stored tiles are zeroed outside n x n).
//
// Work (elements of T): M (n_pad^2, the operator, overwritten by L), W (n_pad^2), and the inverses
// of the 64 x 64 diagonal blocks of L (nb tiles).  `info` receives 0, or 1 + the index of the
// first non-positive pivot (the caller then retries in FP64 or with a diagonal shift).
// Replaces the dense factorisation of the coarse level (Hypre/PETSc coarse solvers there,
// linear/solver.py:466-513).
// ---------------------------------------------------------------------------------------------
constexpr int kH = 32;            // kk extent of one LDS staging pass of tile_mm

// acc (4 x 4 per thread, rows ty + 16 i, columns tx + 16 j, ty = tid / 16, tx = tid % 16) +=
// (SUB: -=) op(A) op(B) for one 64 x 64 x 64 tile product; A, B point at tile origins with
// leading dimensions lda, ldb; op(A)[r][kk] = TA ? A[kk][r] : A[r][kk], op(B)[kk][c] =
// TB ? B[c][kk] : B[kk][c].  All 256 threads call it; it stages kH values of kk at a time.
template <typename T, bool TA, bool TB, bool SUB>
__device__ __forceinline__ void tile_mm(T (&acc)[4][4], const T *__restrict__ A, int64_t lda,
                                        const T *__restrict__ B, int64_t ldb, T *As, T *Bs) {
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    for (int h = 0; h < kT; h += kH) {
        __syncthreads();   // the previous pass's reads are done
        for (int e = tid; e < kH * kT; e += 256) {
            int kk, r;
            if constexpr (TA) { r = e & (kT - 1); kk = e >> 6; }
            else { kk = e & (kH - 1); r = e >> 5; }
            As[kk * kTP + r] = TA ? A[(int64_t)(h + kk) * lda + r] : A[(int64_t)r * lda + h + kk];
            int kb, c;
            if constexpr (TB) { kb = e & (kH - 1); c = e >> 5; }
            else { c = e & (kT - 1); kb = e >> 6; }
            Bs[kb * kTP + c] = TB ? B[(int64_t)c * ldb + h + kb] : B[(int64_t)(h + kb) * ldb + c];
        }
        __syncthreads();
#pragma unroll 4
        for (int kk = 0; kk < kH; ++kk) {
            T a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[kk * kTP + ty + 16 * i];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[kk * kTP + tx + 16 * j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = SUB ? fma(-a[i], b[j], acc[i][j]) : fma(a[i], b[j], acc[i][j]);
        }
    }
}

template <typename T>
__device__ __forceinline__ void tile_load(T (&acc)[4][4], const T *__restrict__ C, int64_t ld) {
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = C[(int64_t)(ty + 16 * i) * ld + tx + 16 * j];
}

template <typename T>
__device__ __forceinline__ void tile_store(const T (&acc)[4][4], T *__restrict__ C, int64_t ld) {
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) C[(int64_t)(ty + 16 * i) * ld + tx + 16 * j] = acc[i][j];
}

// M (n_pad x n_pad, row-major) = A (n x n, FP64, symmetric) padded with the identity; W = identity.
template <typename T>
__global__ void k_spd_load(int64_t n, int64_t n_pad, const double *__restrict__ A, T *__restrict__ M,
                           T *__restrict__ W, int *__restrict__ info) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g == 0) *info = 0;
    if (g >= n_pad * n_pad) return;
    const int64_t i = g / n_pad, j = g - (g / n_pad) * n_pad;
    const bool in = i < n && j < n;
    M[g] = in ? (T)A[in ? j * n + i : 0] : (T)(i == j ? 1 : 0);
    W[g] = (T)(i == j ? 1 : 0);
}

// Diagonal block k: L_kk (Cholesky in LDS, written to M) and its inverse (to Dinv[k]).  One
// barrier per column in each phase:
//   factorisation: right-looking without the column scaling.  After step j the trailing entries
//     hold A - sum_{m <= j} a_m a_m^T / d_m (a_m the unscaled column m, d_m its pivot), so a step
//     reads column j and writes only entries right of it; the columns are scaled by 1 / sqrt(d_m)
//     at the end;
//   inverse X = L^-1, columns right to left: X[i][j] = -(1 / L[j][j]) sum_{m=j+1..i} X[i][m] L[m][j].
//     X is written transposed into the upper triangle (X[i][m] at S[m][i], its diagonal on S's
//     diagonal; L's diagonal is kept in sq[]), so a step writes row j and reads only rows below it.
//     Row i's sum is split over the four lanes 4 i + h (m = j + 1 + h, j + 5 + h, ...) and combined
//     by two cross-lane exchanges.
// A non-positive pivot is reported as its (1-based) row in info and replaced by 1.
template <typename T>
__global__ void __launch_bounds__(256) k_spd_potrf(int64_t n_pad, int k, T *__restrict__ M, T *__restrict__ Dinv,
                                                  int *__restrict__ info) {
    __shared__ T S[kT * kTP];
    __shared__ T dg[kT], sq[kT];
    const int tid = threadIdx.x;
    T *Mk = M + (int64_t)k * kT * n_pad + (int64_t)k * kT;
    for (int e = tid; e < kT * kT; e += 256) S[(e >> 6) * kTP + (e & 63)] = Mk[(int64_t)(e >> 6) * n_pad + (e & 63)];
    for (int j = 0; j < kT; ++j) {
        __syncthreads();
        T d = S[j * kTP + j];   // the same for every thread
        if (!(d > (T)0)) {
            if (tid == 0 && *info == 0) *info = k * kT + j + 1;
            d = (T)1;
        }
        if (tid == 0) dg[j] = d;
        // this thread: column l of rows w, w + 4, ..., w + 60; its 33 operands are loaded before any
        // of its 16 stores (the stores could alias later loads as far as the compiler can tell, so
        // an update loop would wait for each store before the next loads)
        const int l = tid & 63, w = tid >> 6;
        if (l > j) {
            const T b = S[l * kTP + j] / d;
            T a[kT / 4], c[kT / 4];
#pragma unroll
            for (int r = 0; r < kT / 4; ++r) {
                a[r] = S[(w + 4 * r) * kTP + j];
                c[r] = S[(w + 4 * r) * kTP + l];
            }
#pragma unroll
            for (int r = 0; r < kT / 4; ++r)
                if (l <= w + 4 * r) S[(w + 4 * r) * kTP + l] = c[r] - a[r] * b;
        }
    }
    __syncthreads();
    if (tid < kT) sq[tid] = sqrt(dg[tid]);
    __syncthreads();
    for (int e = tid; e < kT * kT; e += 256) {
        const int i = e >> 6, l = e & 63;
        if (l < i) S[i * kTP + l] /= sq[l];
    }
    __syncthreads();
    for (int e = tid; e < kT * kT; e += 256) {
        const int i = e >> 6, l = e & 63;
        Mk[(int64_t)i * n_pad + l] = l < i ? S[i * kTP + l] : (l == i ? sq[i] : (T)0);
    }
    const int ri = tid >> 2, h = tid & 3;
    for (int j = kT - 1; j >= 0; --j) {
        const T inv = (T)1 / sq[j];
        T y = 0;
        if (ri > j)
            for (int m = j + 1 + h; m <= ri; m += 4) y += S[m * kTP + ri] * S[m * kTP + j];
        y += __shfl_xor(y, 1);
        y += __shfl_xor(y, 2);
        if (ri > j && h == 0) S[j * kTP + ri] = -inv * y;
        if (tid == 0) S[j * kTP + j] = inv;
        __syncthreads();
    }
    T *D = Dinv + (int64_t)k * kT * kT;
    for (int e = tid; e < kT * kT; e += 256) {
        const int i = e >> 6, l = e & 63;
        D[e] = l <= i ? S[l * kTP + i] : (T)0;
    }
}

// Panel: L_ik = A_ik L_kk^-T for block rows i > k.
template <typename T>
__global__ void __launch_bounds__(256) k_spd_trsm(int64_t n_pad, int k, T *__restrict__ M, const T *__restrict__ Dinv) {
    __shared__ T As[kH * kTP], Bs[kH * kTP];
    const int64_t i = (int64_t)k + 1 + blockIdx.x;
    T *Aik = M + i * kT * n_pad + (int64_t)k * kT;
    T acc[4][4] = {};
    tile_mm<T, false, true, false>(acc, Aik, n_pad, Dinv + (int64_t)k * kT * kT, kT, As, Bs);
    tile_store(acc, Aik, n_pad);
}

// lower-triangle tile index t -> (i, j), 0 <= j <= i
__device__ __forceinline__ void lower_coords(int64_t t, int64_t &i, int64_t &j) {
    int64_t r = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((r + 1) * (r + 2) / 2 <= t) ++r;
    while (r * (r + 1) / 2 > t) --r;
    i = r;
    j = t - r * (r + 1) / 2;
}

// Trailing update: A_ij -= L_ik L_jk^T for k < j <= i.
template <typename T>
__global__ void __launch_bounds__(256) k_spd_syrk(int64_t n_pad, int k, T *__restrict__ M) {
    __shared__ T As[kH * kTP], Bs[kH * kTP];
    int64_t ii, jj;
    lower_coords(blockIdx.x, ii, jj);
    const int64_t i = k + 1 + ii, j = k + 1 + jj;
    T *Cij = M + i * kT * n_pad + j * kT;
    T acc[4][4];
    tile_load(acc, Cij, n_pad);
    tile_mm<T, false, true, true>(acc, M + i * kT * n_pad + (int64_t)k * kT, n_pad,
                                  M + j * kT * n_pad + (int64_t)k * kT, n_pad, As, Bs);
    tile_store(acc, Cij, n_pad);
}

// Triangular inverse, step k (W starts as the identity): row block k, W_kj = L_kk^-1 W_kj (j <= k) ...
template <typename T>
__global__ void __launch_bounds__(256) k_spd_trtri_row(int64_t n_pad, int k, T *__restrict__ W,
                                                      const T *__restrict__ Dinv) {
    __shared__ T As[kH * kTP], Bs[kH * kTP];
    T *Wkj = W + (int64_t)k * kT * n_pad + (int64_t)blockIdx.x * kT;
    T acc[4][4] = {};
    tile_mm<T, false, false, false>(acc, Dinv + (int64_t)k * kT * kT, kT, Wkj, n_pad, As, Bs);
    tile_store(acc, Wkj, n_pad);
}

// ... then W_ij -= L_ik W_kj for i > k, j <= k.
template <typename T>
__global__ void __launch_bounds__(256) k_spd_trtri_update(int64_t n_pad, int k, const T *__restrict__ M,
                                                         T *__restrict__ W) {
    __shared__ T As[kH * kTP], Bs[kH * kTP];
    const int64_t i = (int64_t)k + 1 + blockIdx.x / (k + 1), j = blockIdx.x % (k + 1);
    T *Wij = W + i * kT * n_pad + j * kT;
    T acc[4][4];
    tile_load(acc, Wij, n_pad);
    tile_mm<T, false, false, true>(acc, M + i * kT * n_pad + (int64_t)k * kT, n_pad,
                                   W + (int64_t)k * kT * n_pad + j * kT, n_pad, As, Bs);
    tile_store(acc, Wij, n_pad);
}

// A^-1 = W^T W: the stored upper tile (I, J), I <= J, is sum_{k >= J} W_kI^T W_kJ, written to the
// symmetric tile store (zero outside n x n).
template <typename T, typename Tout>
__global__ void __launch_bounds__(256) k_spd_lauum_pack(int64_t n, int64_t n_pad, int nb, const T *__restrict__ W,
                                                       Tout *__restrict__ tiles) {
    __shared__ T As[kH * kTP], Bs[kH * kTP];
    int64_t I, J;
    tile_coords(blockIdx.x, I, J);
    T acc[4][4] = {};
    for (int64_t k = J; k < nb; ++k)
        tile_mm<T, true, false, false>(acc, W + k * kT * n_pad + I * kT, n_pad, W + k * kT * n_pad + J * kT, n_pad,
                                       As, Bs);
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    Tout *out = tiles + (int64_t)blockIdx.x * kT * kT;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int r = ty + 16 * i, c = tx + 16 * j;
            const bool in = I * kT + r < n && J * kT + c < n;
            out[r * kT + c] = in ? (Tout)acc[i][j] : (Tout)0;
        }
}

template <typename T>
int spd_inverse(int64_t n, const double *A, T *work, void *tiles, bool t_f64, int *info, hipStream_t st) {
    const int64_t nb = (n + kT - 1) / kT, n_pad = nb * kT;
    T *M = work, *W = work + n_pad * n_pad, *Dinv = work + 2 * n_pad * n_pad;
    hipLaunchKernelGGL((k_spd_load<T>), dim3(ssx::grid_for(n_pad * n_pad, 256)), dim3(256), 0, st, n, n_pad, A, M,
                       W, info);
    for (int k = 0; k < nb; ++k) {
        hipLaunchKernelGGL((k_spd_potrf<T>), dim3(1), dim3(256), 0, st, n_pad, k, M, Dinv, info);
        const int64_t m = nb - k - 1;
        if (m > 0) {
            hipLaunchKernelGGL((k_spd_trsm<T>), dim3((unsigned)m), dim3(256), 0, st, n_pad, k, M, Dinv);
            hipLaunchKernelGGL((k_spd_syrk<T>), dim3((unsigned)(m * (m + 1) / 2)), dim3(256), 0, st, n_pad, k, M);
        }
    }
    for (int k = 0; k < nb; ++k) {
        hipLaunchKernelGGL((k_spd_trtri_row<T>), dim3((unsigned)(k + 1)), dim3(256), 0, st, n_pad, k, W, Dinv);
        const int64_t m = nb - k - 1;
        if (m > 0)
            hipLaunchKernelGGL((k_spd_trtri_update<T>), dim3((unsigned)(m * (k + 1))), dim3(256), 0, st, n_pad, k, M,
                               W);
    }
    const unsigned n_tiles = (unsigned)(nb * (nb + 1) / 2);
    if (t_f64)
        hipLaunchKernelGGL((k_spd_lauum_pack<T, double>), dim3(n_tiles), dim3(256), 0, st, n, n_pad, (int)nb, W,
                           (double *)tiles);
    else
        hipLaunchKernelGGL((k_spd_lauum_pack<T, float>), dim3(n_tiles), dim3(256), 0, st, n, n_pad, (int)nb, W,
                           (float *)tiles);
    SSX_CHECK_LAUNCH("ssx_spd_inverse_tiles");
    return SSX_OK;
}

}  // namespace

extern "C" {

int ssx_sym_tiles_sizes(int64_t n, int64_t *tile_elems, int64_t *part_elems) {
    SSX_REQUIRE(n >= 0 && tile_elems && part_elems, "ssx_sym_tiles_sizes: bad arguments");
    const int64_t nt = (n + kT - 1) / kT;
    *tile_elems = nt * (nt + 1) / 2 * kT * kT;
    *part_elems = nt * nt * kT;
    return SSX_OK;
}

int ssx_sym_tiles_pack(int64_t n, const void *A, int a_f64, void *tiles, int t_f64, void *stream) {
    SSX_REQUIRE(n >= 0, "ssx_sym_tiles_pack: n < 0");
    if (n == 0) return SSX_OK;
    SSX_REQUIRE(A && tiles, "ssx_sym_tiles_pack: null pointer");
    const int64_t nt = (n + kT - 1) / kT;
    const int64_t n_tiles = nt * (nt + 1) / 2;
    const int64_t total = n_tiles * kT * kT;
    const dim3 grid(ssx::grid_for(total, 256)), block(256);
    hipStream_t st = ssx::as_stream(stream);
    if (a_f64 && t_f64)
        hipLaunchKernelGGL((k_sym_pack<double, double>), grid, block, 0, st, n, n_tiles, (const double *)A, (double *)tiles);
    else if (a_f64)
        hipLaunchKernelGGL((k_sym_pack<double, float>), grid, block, 0, st, n, n_tiles, (const double *)A, (float *)tiles);
    else if (!t_f64)
        hipLaunchKernelGGL((k_sym_pack<float, float>), grid, block, 0, st, n, n_tiles, (const float *)A, (float *)tiles);
    else
        hipLaunchKernelGGL((k_sym_pack<float, double>), grid, block, 0, st, n, n_tiles, (const float *)A, (double *)tiles);
    SSX_CHECK_LAUNCH("ssx_sym_tiles_pack");
    return SSX_OK;
}

int ssx_spd_inverse_sizes(int64_t n, int64_t *work_elems) {
    SSX_REQUIRE(n >= 0 && work_elems, "ssx_spd_inverse_sizes: bad arguments");
    const int64_t nb = (n + kT - 1) / kT, n_pad = nb * kT;
    *work_elems = 2 * n_pad * n_pad + nb * kT * kT;
    return SSX_OK;
}

int ssx_spd_inverse_tiles(int64_t n, const double *A, int f64, void *work, void *tiles, int t_f64, int *info,
                          void *stream) {
    SSX_REQUIRE(n >= 0, "ssx_spd_inverse_tiles: n < 0");
    if (n == 0) return SSX_OK;
    SSX_REQUIRE(A && work && tiles && info, "ssx_spd_inverse_tiles: null pointer");
    const int64_t nb = (n + kT - 1) / kT;
    SSX_REQUIRE(nb * (nb + 1) / 2 <= 0x7fffffff && nb * nb <= 0x7fffffff, "ssx_spd_inverse_tiles: n too large");
    hipStream_t st = ssx::as_stream(stream);
    if (f64) return spd_inverse<double>(n, A, (double *)work, tiles, t_f64 != 0, info, st);
    return spd_inverse<float>(n, A, (float *)work, tiles, t_f64 != 0, info, st);
}

int ssx_sym_tiles_gemv_f32(int64_t n, const float *tiles, const float *x, float *part, float *y, void *stream) {
    return sym_gemv<float>(n, tiles, x, part, y, stream);
}

int ssx_sym_tiles_gemv_f64(int64_t n, const double *tiles, const double *x, double *part, double *y, void *stream) {
    return sym_gemv<double>(n, tiles, x, part, y, stream);
}

}  // extern "C"
