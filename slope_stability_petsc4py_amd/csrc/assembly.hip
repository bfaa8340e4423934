// Fixed-pattern finite-element assembly on gfx950: strain gather E = B u, internal force
// F = B^T (w S), and tangent values K = sum B^T D B written into the fixed CSR pattern, with the
// regularised operator K_r = (1-r) K + r K_el and an optional FP32 copy produced in the same pass.
//
// Reference behaviour: cython/assemble_tangent_values_3d.c:216-273 (strain), :275-327 (force),
// :111-214 (tangent rows); fem/distributed_tangent.py:1559-1612 (constrained diagonal = 1),
// constitutive/problem.py:1595-1644 (K_r on the same pattern).
//
// HBM layout: dphi is (dim, n_int, n_p) (one (n_int, n_p) slab per spatial derivative, the
// reference's dphi1/dphi2/dphi3), w is (n_int), E/S are point-major (n_int, n_strain), DS is
// point-major (n_int, n_strain^2, column-major per point).  Element DOF tables are (n_elem, dim*n_p).
#include "ssx_common.h"

namespace {

// 3D strain order [e11 e22 e33 g12 g23 g13]; 2D [e11 e22 g12].
//
// Code-generation rule for every kernel in this library: no divergent multi-way branch may choose
// between addresses (or offsets) that are then dereferenced.  Select VALUES with branch-free
// arithmetic instead.  hipcc (AMD clang 22, gfx950) has been seen to place the address of one arm
// of such a branch under the wrong execution mask, leaving lanes of another arm dereferencing an
// unassigned register (a GPU page fault); the LLVM IR was correct, the machine code was not.
template <int DIM>
struct BCol;

template <>
struct BCol<3> {
    static constexpr int NS = 6;
};

template <>
struct BCol<2> {
    static constexpr int NS = 3;
};

// Row ci of B^T for a node with gradient (x, y, z), as coefficients on the strain rows:
// ci = 0: (x, 0, 0, y, 0, z); ci = 1: (0, y, 0, x, z, 0); ci = 2: (0, 0, z, 0, y, x).
// Branch-free (0/1 weights, exact), so the same six D entries are read for every ci.
__device__ __forceinline__ void brow3(int ci, double x, double y, double z, double b[6]) {
    const double c0 = ci == 0 ? 1.0 : 0.0, c1 = ci == 1 ? 1.0 : 0.0, c2 = ci == 2 ? 1.0 : 0.0;
    b[0] = c0 * x;
    b[1] = c1 * y;
    b[2] = c2 * z;
    b[3] = c0 * y + c1 * x;
    b[4] = c1 * z + c2 * y;
    b[5] = c0 * z + c2 * x;
}

// 2D: ci = 0: (x, 0, y); ci = 1: (0, y, x).
__device__ __forceinline__ void brow2(int ci, double x, double y, double b[3]) {
    const double c0 = ci == 0 ? 1.0 : 0.0, c1 = 1.0 - c0;
    b[0] = c0 * x;
    b[1] = c1 * y;
    b[2] = c0 * y + c1 * x;
}

// ------------------------------------------------------------------------------------------
// Strain gather: one lane per quadrature point.  A workgroup owns SB consecutive points (128, 64 or
// 32 as the element grows, so the LDS per workgroup stays near 20-30 KB): their gradient rows (three contiguous slabs of SB*NP values) are
// staged in LDS with coalesced
// loads (row stride NP+1 against bank conflicts), each lane then gathers its element's
// displacements (u is small and stays in L2) and the block's strains leave through LDS in one
// coalesced pass over the contiguous range E[g0*NS .. (g0+np)*NS).
// ------------------------------------------------------------------------------------------
template <int NP>
struct StrainBlock {
    static constexpr int value = NP > 15 ? 32 : (NP > 6 ? 64 : 128);
};

template <int DIM, int NP, int kSBlock = StrainBlock<NP>::value>
__global__ void __launch_bounds__(kSBlock) k_strain(const double *__restrict__ dphi, const int32_t *__restrict__ edofs,
                                                    const double *__restrict__ u, int64_t n_int, int n_q,
                                                    double *__restrict__ E) {
    constexpr int NS = BCol<DIM>::NS;
    constexpr int RS = NP + 1;
    static_assert(DIM * RS >= NS + 1, "the strain staging reuses the gradient area");
    // gradient rows of the block's points (component c at c*kSBlock*RS), then -- after a barrier --
    // the block's strains (row stride NS + 1) in the same area: one LDS area keeps more
    // workgroups resident per CU
    __shared__ double sg[DIM * kSBlock * RS];
    const int t = threadIdx.x;
    const int64_t g0 = (int64_t)blockIdx.x * kSBlock;
    const int np = (int)(n_int - g0 < kSBlock ? n_int - g0 : kSBlock);
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
        const double *src = dphi + ((int64_t)c * n_int + g0) * NP;
        for (int i = t; i < np * NP; i += kSBlock) sg[c * kSBlock * RS + (i / NP) * RS + i % NP] = src[i];
    }
    __syncthreads();
    double a[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) a[k] = 0.0;
    if (t < np) {
        const int64_t g = g0 + t;
        const int64_t e = g / n_q;
        const int32_t *dof = edofs + e * (DIM * NP);
        const double *gx = sg + t * RS;
        const double *gy = sg + kSBlock * RS + t * RS;
        if constexpr (DIM == 3) {
            const double *gz = sg + 2 * kSBlock * RS + t * RS;
#pragma unroll
            for (int j = 0; j < NP; ++j) {
                const double ux = u[dof[3 * j]], uy = u[dof[3 * j + 1]], uz = u[dof[3 * j + 2]];
                const double x = gx[j], y = gy[j], z = gz[j];
                a[0] += x * ux;
                a[1] += y * uy;
                a[2] += z * uz;
                a[3] += y * ux + x * uy;
                a[4] += z * uy + y * uz;
                a[5] += z * ux + x * uz;
            }
        } else {
#pragma unroll
            for (int j = 0; j < NP; ++j) {
                const double ux = u[dof[2 * j]], uy = u[dof[2 * j + 1]];
                const double x = gx[j], y = gy[j];
                a[0] += x * ux;
                a[1] += y * uy;
                a[2] += y * ux + x * uy;
            }
        }
    }
    __syncthreads();   // gradients read: the area now takes the strains
    if (t < np) {
        double *o = sg + t * (NS + 1);
#pragma unroll
        for (int k = 0; k < NS; ++k) o[k] = a[k];
    }
    __syncthreads();
    for (int i = t; i < np * NS; i += kSBlock) E[g0 * NS + i] = sg[(i / NS) * (NS + 1) + i % NS];
}

// ------------------------------------------------------------------------------------------
// Internal force, stage 1: element-local vector f_e[a] = sum_q w_q B_a(q)^T S_q, one lane per
// (element, node) producing the node's DIM components from one read of each point's stress.
// Stage 2 gathers f_e into rows through the row-slot lists, so every row is written by one
// lane in a fixed (element) order: deterministic, no atomics.
// ------------------------------------------------------------------------------------------
// NQ > 0: the rule's point count at compile time (the point loop unrolls, so every point's loads
// are in flight together); NQ == 0: the runtime n_q.
template <int DIM, int NP, int NQ = 0>
__global__ void __launch_bounds__(256) k_force_elem(const double *__restrict__ dphi, const double *__restrict__ S,
                                                    const double *__restrict__ w, int64_t n_elem, int64_t n_int,
                                                    int n_q_rt, double *__restrict__ fe) {
    constexpr int NLD = DIM * NP;
    constexpr int NS = BCol<DIM>::NS;
    const int n_q = NQ > 0 ? NQ : n_q_rt;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_elem * NP) return;
    const int64_t e = t / NP;
    const int node = (int)(t - e * NP);
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;
#pragma unroll
    for (int q = 0; q < n_q; ++q) {
        const int64_t g = e * n_q + q;
        const double gx = dphi[g * NP + node];
        const double gy = dphi[(n_int + g) * NP + node];
        const double wg = w[g];
        const double *s = S + g * NS;
        if constexpr (DIM == 3) {
            const double gz = dphi[(2 * n_int + g) * NP + node];
            const double s0 = s[0], s1 = s[1], s2 = s[2], s3 = s[3], s4 = s[4], s5 = s[5];
            f0 += wg * (gx * s0 + gy * s3 + gz * s5);
            f1 += wg * (gx * s3 + gy * s1 + gz * s4);
            f2 += wg * (gx * s5 + gy * s4 + gz * s2);
        } else {
            const double s0 = s[0], s1 = s[1], s2 = s[2];
            f0 += wg * (gx * s0 + gy * s2);
            f1 += wg * (gy * s1 + gx * s2);
        }
    }
    double *o = fe + e * NLD + DIM * node;
    o[0] = f0;
    o[1] = f1;
    if constexpr (DIM == 3) o[2] = f2;
}

__global__ void __launch_bounds__(256) k_force_gather(const double *__restrict__ fe, const int32_t *__restrict__ row_slot_ptr,
                                                      const int32_t *__restrict__ slot_elem,
                                                      const uint8_t *__restrict__ slot_lrow, int64_t n_rows, int nld,
                                                      double *__restrict__ F) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_rows) return;
    double acc = 0.0;
    const int32_t s1 = row_slot_ptr[r + 1];
    for (int32_t s = row_slot_ptr[r]; s < s1; ++s) acc += fe[(int64_t)slot_elem[s] * nld + slot_lrow[s]];
    F[r] = acc;  // rows without slots (constrained) get 0, as the reference zeroes them
}

// Stage 2 by mesh node: lane (node, c) sums fe[e*nld + dim*lnode + c] over the node's element
// slots (element order, the order of the row slots, so the sums are the row-slot gather's bit for
// bit); the dim lanes of a node read dim adjacent entries of each element vector instead of one
// scattered entry per row.  Constrained rows get 0.
__global__ void __launch_bounds__(256) k_force_gather_nodes(const double *__restrict__ fe,
                                                            const int32_t *__restrict__ ns_ptr,
                                                            const int32_t *__restrict__ ns_elem,
                                                            const uint8_t *__restrict__ ns_lnode,
                                                            const uint8_t *__restrict__ row_free, int64_t n_nodes,
                                                            int dim, int nld, double *__restrict__ F) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_nodes * dim) return;
    const int64_t node = r / dim;
    const int c = (int)(r - node * dim);
    double acc = 0.0;
    int32_t s = ns_ptr[node];
    const int32_t s1 = ns_ptr[node + 1];
    // four slots' loads in flight at a time (one slot's metadata, then its entry, one slot after
    // another, kept a single chain of dependent loads in flight); the adds stay in slot order, so
    // the sums are the same bit for bit
    for (; s + 4 <= s1; s += 4) {
        int64_t o[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) o[k] = (int64_t)ns_elem[s + k] * nld + dim * (int)ns_lnode[s + k] + c;
        double v[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = fe[o[k]];
#pragma unroll
        for (int k = 0; k < 4; ++k) acc += v[k];
    }
    for (; s < s1; ++s) acc += fe[(int64_t)ns_elem[s] * nld + dim * (int)ns_lnode[s] + c];
    F[r] = row_free[r] ? acc : 0.0;
}

// ------------------------------------------------------------------------------------------
// Tangent values on the fixed pattern.  One workgroup of ROWS lanes owns ROWS consecutive rows,
// i.e. one contiguous run of the CSR value array; with USE_LDS that run is accumulated in LDS and
// written once, fully coalesced, with the regularisation and the FP32 copy fused in.  Each lane
// walks its row's slots in element order; per slot it sums the point contributions of its
// 3*n_p (or 2*n_p) columns in registers and adds them to the row once.
// ------------------------------------------------------------------------------------------
constexpr int kRows = 64;

template <int DIM, int NP, bool USE_LDS>
__global__ void __launch_bounds__(kRows) k_tangent_rows(
    const double *__restrict__ dphi, const double *__restrict__ DS, const double *__restrict__ w,
    const int32_t *__restrict__ row_slot_ptr, const int32_t *__restrict__ slot_elem,
    const uint8_t *__restrict__ slot_lrow, const int32_t *__restrict__ slot_pos, const int32_t *__restrict__ indptr,
    const uint8_t *__restrict__ row_free, int64_t n_rows, int64_t n_int, int n_q, double r,
    const double *__restrict__ elastic, double *__restrict__ out, float *__restrict__ out32) {
    constexpr int NLD = DIM * NP;
    constexpr int NS = BCol<DIM>::NS;
    extern __shared__ double lds[];
    const int64_t r0 = (int64_t)blockIdx.x * kRows;
    const int64_t r1 = r0 + kRows < n_rows ? r0 + kRows : n_rows;
    const int64_t v0 = indptr[r0];
    const int64_t L = (int64_t)indptr[r1] - v0;
    double *buf = USE_LDS ? lds : out + v0;
    for (int64_t i = threadIdx.x; i < L; i += kRows) buf[i] = 0.0;
    __syncthreads();

    const int64_t row = r0 + threadIdx.x;
    if (row < r1) {
        if (!row_free[row]) {
            buf[indptr[row] - v0] = 1.0;  // constrained row: only its diagonal, set to 1
        } else {
            const int32_t s_end = row_slot_ptr[row + 1];
            for (int32_t s = row_slot_ptr[row]; s < s_end; ++s) {
                const int64_t e = slot_elem[s];
                const int a = slot_lrow[s];
                const int ni = a / DIM, ci = a % DIM;
                double acc[NLD];
#pragma unroll
                for (int b = 0; b < NLD; ++b) acc[b] = 0.0;
                for (int q = 0; q < n_q; ++q) {
                    const int64_t g = e * n_q + q;
                    const double *gx = dphi + g * NP;
                    const double *gy = dphi + (n_int + g) * NP;
                    const double *gz = DIM == 3 ? dphi + (2 * n_int + g) * NP : gx;
                    const double *D = DS + g * (NS * NS);
                    const double wg = w[g];
                    // t_m = w * sum_s B_row[s] D[s, m]  (D column-major: D[s + NS*m])
                    double t[NS];
                    if constexpr (DIM == 3) {
                        const double xi = gx[ni], yi = gy[ni], zi = gz[ni];
                        double b[6];
                        brow3(ci, xi, yi, zi, b);
#pragma unroll
                        for (int m = 0; m < 6; ++m) {
                            const double *Dm = D + 6 * m;
                            double v = b[0] * Dm[0];
#pragma unroll
                            for (int s2 = 1; s2 < 6; ++s2) v += b[s2] * Dm[s2];
                            t[m] = wg * v;
                        }
#pragma unroll
                        for (int j = 0; j < NP; ++j) {
                            const double xj = gx[j], yj = gy[j], zj = gz[j];
                            acc[3 * j + 0] += t[0] * xj + t[3] * yj + t[5] * zj;
                            acc[3 * j + 1] += t[3] * xj + t[1] * yj + t[4] * zj;
                            acc[3 * j + 2] += t[5] * xj + t[4] * yj + t[2] * zj;
                        }
                    } else {
                        const double xi = gx[ni], yi = gy[ni];
                        double b[3];
                        brow2(ci, xi, yi, b);
#pragma unroll
                        for (int m = 0; m < 3; ++m) {
                            const double *Dm = D + 3 * m;
                            t[m] = wg * (b[0] * Dm[0] + b[1] * Dm[1] + b[2] * Dm[2]);
                        }
#pragma unroll
                        for (int j = 0; j < NP; ++j) {
                            const double xj = gx[j], yj = gy[j];
                            acc[2 * j + 0] += t[0] * xj + t[2] * yj;
                            acc[2 * j + 1] += t[2] * xj + t[1] * yj;
                        }
                    }
                }
                const int32_t *pos = slot_pos + (int64_t)s * NLD;
                int32_t pv[NLD];
#pragma unroll
                for (int b = 0; b < NLD; ++b) pv[b] = pos[b];
                if constexpr (USE_LDS) {
                    // constrained columns (p < 0) go to the dummy word buf[L]: no branch per column
#pragma unroll
                    for (int b = 0; b < NLD; ++b) {
                        const int64_t idx = pv[b] >= 0 ? (int64_t)pv[b] - v0 : L;
                        buf[idx] += acc[b];
                    }
                } else {
#pragma unroll
                    for (int b = 0; b < NLD; ++b)
                        if (pv[b] >= 0) buf[pv[b] - v0] += acc[b];
                }
            }
        }
    }
    __syncthreads();
    const double one_minus_r = 1.0 - r;
    for (int64_t i = threadIdx.x; i < L; i += kRows) {
        double v = buf[i];
        if (elastic) v = one_minus_r * v + r * elastic[v0 + i];
        if (USE_LDS) out[v0 + i] = v;
        else buf[i] = v;
        if (out32) out32[v0 + i] = (float)v;
    }
}

template <int DIM, int NP>
int launch_tangent(const double *dphi, const double *DS, const double *w, const int32_t *row_slot_ptr,
                   const int32_t *slot_elem, const uint8_t *slot_lrow, const int32_t *slot_pos, const int32_t *indptr,
                   const uint8_t *row_free, int64_t n_rows, int64_t n_int, int n_q, double r, const double *elastic,
                   double *out, float *out32, int64_t max_block_nnz, hipStream_t st) {
    const unsigned grid = ssx::grid_for(n_rows, kRows);
    const size_t lds_bytes = ((size_t)max_block_nnz + 1) * sizeof(double);  // + the dummy word
    static int max_lds = -1;
    if (max_lds < 0) {
        int dev = 0, v = 65536;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
            v = 65536;
        max_lds = v;
    }
    bool lds_ok = max_block_nnz >= 0 && lds_bytes <= (size_t)max_lds;
    if (lds_ok && lds_bytes > 65536)
        lds_ok = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_tangent_rows<DIM, NP, true>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes) == hipSuccess;
    if (lds_ok) {
        hipLaunchKernelGGL((k_tangent_rows<DIM, NP, true>), dim3(grid), dim3(kRows), lds_bytes, st, dphi, DS, w,
                           row_slot_ptr, slot_elem, slot_lrow, slot_pos, indptr, row_free, n_rows, n_int, n_q, r,
                           elastic, out, out32);
    } else {
        hipLaunchKernelGGL((k_tangent_rows<DIM, NP, false>), dim3(grid), dim3(kRows), 0, st, dphi, DS, w, row_slot_ptr,
                           slot_elem, slot_lrow, slot_pos, indptr, row_free, n_rows, n_int, n_q, r, elastic, out,
                           out32);
    }
    SSX_CHECK_LAUNCH("ssx_tangent_rows");
    return SSX_OK;
}

template <int DIM, int NP>
int launch_strain(const double *dphi, const int32_t *edofs, const double *u, int64_t n_int, int n_q, double *E,
                  hipStream_t st) {
    constexpr int SB = StrainBlock<NP>::value;
    hipLaunchKernelGGL((k_strain<DIM, NP>), dim3(ssx::grid_for(n_int, SB)), dim3(SB), 0, st, dphi, edofs, u, n_int,
                       n_q, E);
    SSX_CHECK_LAUNCH("ssx_strain");
    return SSX_OK;
}

template <int DIM, int NP>
int launch_force_elem(const double *dphi, const double *S, const double *w, int64_t n_elem, int64_t n_int, int n_q,
                      double *fe, hipStream_t st) {
    if (DIM == 3 && NP == 10 && n_q == 11)
        hipLaunchKernelGGL((k_force_elem<DIM, NP, 11>), dim3(ssx::grid_for(n_elem * NP, 256)), dim3(256), 0, st, dphi,
                           S, w, n_elem, n_int, n_q, fe);
    else
        hipLaunchKernelGGL((k_force_elem<DIM, NP>), dim3(ssx::grid_for(n_elem * NP, 256)), dim3(256), 0, st, dphi, S,
                           w, n_elem, n_int, n_q, fe);
    SSX_CHECK_LAUNCH("ssx_force");
    return SSX_OK;
}

}  // namespace

// Dispatch over the supported (dim, n_p) pairs: 3D P1/P2/P4 tetrahedra, 2D P1/P2/P4 triangles.
#define SSX_DISPATCH(dim, n_p, CALL)                                                           \
    do {                                                                                       \
        if (dim == 3 && n_p == 4) return CALL(3, 4);                                           \
        if (dim == 3 && n_p == 10) return CALL(3, 10);                                         \
        if (dim == 3 && n_p == 35) return CALL(3, 35);                                         \
        if (dim == 2 && n_p == 3) return CALL(2, 3);                                           \
        if (dim == 2 && n_p == 6) return CALL(2, 6);                                           \
        if (dim == 2 && n_p == 15) return CALL(2, 15);                                         \
        ssx::set_error("unsupported element: dim=%d n_p=%d", dim, n_p);                        \
        return SSX_EINVAL;                                                                     \
    } while (0)

extern "C" {

int ssx_strain(int dim, int n_p, const double *dphi, const int32_t *edofs, const double *u, int64_t n_elem, int n_q,
               double *E, void *stream) {
    SSX_REQUIRE(n_elem >= 0 && n_q > 0, "ssx_strain: bad sizes");
    SSX_REQUIRE(dphi && edofs && u && E, "ssx_strain: null pointer");
    if (n_elem == 0) return SSX_OK;
    const int64_t n_int = n_elem * n_q;
    hipStream_t st = ssx::as_stream(stream);
#define CALL(D, P) launch_strain<D, P>(dphi, edofs, u, n_int, n_q, E, st)
    SSX_DISPATCH(dim, n_p, CALL);
#undef CALL
}

int ssx_force_elem(int dim, int n_p, const double *dphi, const double *S, const double *w, int64_t n_elem, int n_q,
                   double *fe, void *stream) {
    SSX_REQUIRE(n_elem >= 0 && n_q > 0, "ssx_force_elem: bad sizes");
    SSX_REQUIRE(dphi && S && w && fe, "ssx_force_elem: null pointer");
    if (n_elem == 0) return SSX_OK;
    const int64_t n_int = n_elem * n_q;
    hipStream_t st = ssx::as_stream(stream);
#define CALL(D, P) launch_force_elem<D, P>(dphi, S, w, n_elem, n_int, n_q, fe, st)
    SSX_DISPATCH(dim, n_p, CALL);
#undef CALL
}

int ssx_force_gather(const double *fe, const int32_t *row_slot_ptr, const int32_t *slot_elem, const uint8_t *slot_lrow,
                     int64_t n_rows, int nld, double *F, void *stream) {
    SSX_REQUIRE(n_rows >= 0 && nld > 0, "ssx_force_gather: bad sizes");
    if (n_rows == 0) return SSX_OK;
    SSX_REQUIRE(fe && row_slot_ptr && F, "ssx_force_gather: null pointer");
    hipLaunchKernelGGL(k_force_gather, dim3(ssx::grid_for(n_rows, 256)), dim3(256), 0, ssx::as_stream(stream), fe,
                       row_slot_ptr, slot_elem, slot_lrow, n_rows, nld, F);
    SSX_CHECK_LAUNCH("ssx_force_gather");
    return SSX_OK;
}

int ssx_force_gather_nodes(int dim, int nld, const double *fe, const int32_t *ns_ptr, const int32_t *ns_elem,
                           const uint8_t *ns_lnode, const uint8_t *row_free, int64_t n_nodes, double *F, void *stream) {
    SSX_REQUIRE((dim == 2 || dim == 3) && nld > 0 && n_nodes >= 0, "ssx_force_gather_nodes: bad sizes");
    if (n_nodes == 0) return SSX_OK;
    SSX_REQUIRE(fe && ns_ptr && ns_elem && ns_lnode && row_free && F, "ssx_force_gather_nodes: null pointer");
    hipLaunchKernelGGL(k_force_gather_nodes, dim3(ssx::grid_for(n_nodes * dim, 256)), dim3(256), 0,
                       ssx::as_stream(stream), fe, ns_ptr, ns_elem, ns_lnode, row_free, n_nodes, dim, nld, F);
    SSX_CHECK_LAUNCH("ssx_force_gather_nodes");
    return SSX_OK;
}

int ssx_tangent_rows(int dim, int n_p, const double *dphi, const double *DS, const double *w,
                     const int32_t *row_slot_ptr, const int32_t *slot_elem, const uint8_t *slot_lrow,
                     const int32_t *slot_pos, const int32_t *indptr, const uint8_t *row_free, int64_t n_rows,
                     int64_t n_elem, int n_q, double r, const double *elastic, double *out, float *out32,
                     int64_t max_block_nnz, void *stream) {
    SSX_REQUIRE(n_rows >= 0 && n_elem >= 0 && n_q > 0, "ssx_tangent_rows: bad sizes");
    if (n_rows == 0) return SSX_OK;
    SSX_REQUIRE(dphi && DS && w && row_slot_ptr && indptr && row_free && out, "ssx_tangent_rows: null pointer");
    const int64_t n_int = n_elem * n_q;
    hipStream_t st = ssx::as_stream(stream);
#define CALL(D, P)                                                                                                 \
    launch_tangent<D, P>(dphi, DS, w, row_slot_ptr, slot_elem, slot_lrow, slot_pos, indptr, row_free, n_rows, n_int, \
                         n_q, r, elastic, out, out32, max_block_nnz, st)
    SSX_DISPATCH(dim, n_p, CALL);
#undef CALL
}

}  // extern "C"
