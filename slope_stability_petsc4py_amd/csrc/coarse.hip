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

int ssx_sym_tiles_gemv_f32(int64_t n, const float *tiles, const float *x, float *part, float *y, void *stream) {
    return sym_gemv<float>(n, tiles, x, part, y, stream);
}

int ssx_sym_tiles_gemv_f64(int64_t n, const double *tiles, const double *x, double *part, double *y, void *stream) {
    return sym_gemv<double>(n, tiles, x, part, y, stream);
}

}  // extern "C"
