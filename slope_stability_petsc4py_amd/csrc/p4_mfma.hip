// P4 tetrahedra: element tangent matrices on the FP64 matrix cores, then a deterministic row
// gather into the fixed CSR pattern.
//
// ke = sum_q B_q^T (w_q D_q) B_q is one GEMM per element, ke (105 x 105) = A (105 x 144) * Bm (144 x 105)
// with K = 24 points x 6 strains, A[i][k] = B_q[s][i] and Bm[k][j] = sum_t w_q D_q[s][t] B_q[t][j]
// (k = 6 q + s).  One wavefront per element; the 105 x 105 result is padded to 7 x 7 tiles of
// 16 x 16 and each tile accumulates 36 v_mfma_f64_16x16x4 steps.  Both operands are formed per lane
// from the element's dphi and w*DS staged in LDS (B is sparse: every column of B_q has 3 nonzeros,
// so an operand value costs <= 3 FMAs).  DS is symmetric for associated flow, so only the upper
// tile triangle (28 of 49 tiles) is computed and mirrored on write.
//
// Stage 2 walks every row's slots in element order (the reference's `rows` order,
// cython/assemble_tangent_values_3d.c:111-214) and adds ke[e][a][:] at the slot's CSR positions,
// then applies the constrained diagonal, K_r blend and FP32 copy.
//
// Reference behaviour: cython/assemble_tangent_values_3d.c:13-109 (element matrices + scatter),
// :111-214 (rows); flop count 2*105*105*144 per element.
#include "ssx_common.h"

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int NP4 = 35;
constexpr int NLD4 = 105;
constexpr int NQ4 = 24;
constexpr int KDIM = NQ4 * 6;  // 144
constexpr int TILES = 7;        // ceil(105 / 16)

// Column of B for local dof i = 3*node + d (gradients (gx, gy, gz) of that node), as the six
// strain-row coefficients, branch-free (0/1 weights): no divergent branch selects an address or
// an LDS offset anywhere in this file (hipcc for gfx950 has misplaced such selections under the
// wrong execution mask).
__device__ __forceinline__ void bcol(int d, double gx, double gy, double gz, double c[6]) {
    const double c0 = d == 0 ? 1.0 : 0.0, c1 = d == 1 ? 1.0 : 0.0, c2 = d == 2 ? 1.0 : 0.0;
    c[0] = c0 * gx;
    c[1] = c1 * gy;
    c[2] = c2 * gz;
    c[3] = c0 * gy + c1 * gx;
    c[4] = c1 * gz + c2 * gy;
    c[5] = c0 * gz + c2 * gx;
}

// Entry s of a six-vector, branch-free.
__device__ __forceinline__ double pick6(int s, const double c[6]) {
    double v = 0.0;
#pragma unroll
    for (int r = 0; r < 6; ++r) v += (s == r ? 1.0 : 0.0) * c[r];
    return v;
}

__global__ void __launch_bounds__(64) k_ke_p4_mfma(const double *__restrict__ dphi, const double *__restrict__ DS,
                                                   const double *__restrict__ w, int64_t n_elem, int64_t n_int,
                                                   double r, const double *__restrict__ shear,
                                                   const double *__restrict__ lame, double *__restrict__ ke) {
    __shared__ double sg[3][NQ4][NP4];  // gradients of the element's basis functions
    __shared__ double sd[NQ4][36];      // w_q * DS_q (column-major 6x6)
    const int64_t e = blockIdx.x;
    if (e >= n_elem) return;
    const int lane = threadIdx.x;
    const int64_t g0 = e * NQ4;
    for (int t = lane; t < 3 * NQ4 * NP4; t += 64) {
        const int c = t / (NQ4 * NP4), rem = t % (NQ4 * NP4);
        sg[c][rem / NP4][rem % NP4] = dphi[((int64_t)c * n_int + g0) * NP4 + rem];
    }
    // w_q * D_q, or w_q * ((1 - r) DS_q + r D_el) with the elastic tangent
    // D_el = lame * vol vol^T + 2 G * ident (ident = diag(1, 1, 1, 1/2, 1/2, 1/2)) when the
    // regularisation is folded in (shear != nullptr)
    const bool fold = shear != nullptr;
    for (int t = lane; t < NQ4 * 36; t += 64) {
        const int q = t / 36, k = t % 36;
        const int i = k % 6, j = k / 6;
        double d = DS[(g0 + q) * 36 + k];
        if (fold) {
            const double G = shear[g0 + q], L = lame[g0 + q];
            const double del = (i < 3 && j < 3 ? L : 0.0) + (i == j ? (i < 3 ? 2.0 * G : G) : 0.0);
            d = (1.0 - r) * d + r * del;
        }
        sd[q][k] = w[g0 + q] * d;
    }
    __syncthreads();

    const int col = lane & 15;   // i for A, j for B and D
    const int kq = lane >> 4;    // k offset within the 4-wide step
    double *out = ke + e * (int64_t)(NLD4 * NLD4);
    for (int ti = 0; ti < TILES; ++ti) {
        d4 acc[TILES];
#pragma unroll
        for (int tj = 0; tj < TILES; ++tj) acc[tj] = d4{0.0, 0.0, 0.0, 0.0};
        const int ia = 16 * ti + col;
        // padded rows/columns (>= 105) use a valid node and a zero weight
        const double va = ia < NLD4 ? 1.0 : 0.0;
        const int iac = ia < NLD4 ? ia : NLD4 - 1;
        const int na = iac / 3, da = iac % 3;
        for (int ks = 0; ks < KDIM / 4; ++ks) {
            const int k = 4 * ks + kq;
            const int q = k / 6, s = k % 6;
            double ca[6];
            bcol(da, sg[0][q][na], sg[1][q][na], sg[2][q][na], ca);
            const double a = va * pick6(s, ca);
            const double *D = sd[q];
            // row s of wD: D column-major, D[s + 6 t]
            double drow[6];
#pragma unroll
            for (int t = 0; t < 6; ++t) drow[t] = D[s + 6 * t];
#pragma unroll
            for (int tj = 0; tj < TILES; ++tj) {
                if (tj < ti) continue;
                const int jb = 16 * tj + col;
                const double vb = jb < NLD4 ? 1.0 : 0.0;
                const int jbc = jb < NLD4 ? jb : NLD4 - 1;
                const int nb = jbc / 3, db = jbc % 3;
                double cb[6];
                bcol(db, sg[0][q][nb], sg[1][q][nb], sg[2][q][nb], cb);
                // row s of (wD) B: sum_t wD[s][t] B[t][jb]
                double b = drow[0] * cb[0];
#pragma unroll
                for (int t = 1; t < 6; ++t) b += drow[t] * cb[t];
                acc[tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, vb * b, acc[tj], 0, 0, 0);
            }
        }
        // accumulator layout of v_mfma_f64_16x16x4f64: register r of lane l holds row
        // (l / 16) + 4 r of column l % 16
#pragma unroll
        for (int tj = 0; tj < TILES; ++tj) {
            if (tj < ti) continue;
            const int j = 16 * tj + col;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = 16 * ti + kq + 4 * r;
                if (i < NLD4 && j < NLD4) {
                    out[(int64_t)i * NLD4 + j] = acc[tj][r];
                    if (tj > ti) out[(int64_t)j * NLD4 + i] = acc[tj][r];
                }
            }
        }
    }
}

// Row gather of element matrices: one lane per row, slots in element order; LDS row-block tile
// as in the P2 tangent kernel (USE_LDS) or read-modify-write on the lane's own row in HBM.
// NLD > 0: compile-time element size, columns handled in register chunks of CH (positions and
// values loaded as vectors, then branch-free LDS updates with constrained columns sent to the
// dummy word buf[L]); NLD == 0: any element size.
constexpr int kRowsKe = 64;

template <int NLD, bool USE_LDS>
__global__ void __launch_bounds__(kRowsKe) k_scatter_ke_rows(const double *__restrict__ ke, int nld,
                                                             const int32_t *__restrict__ row_slot_ptr,
                                                             const int32_t *__restrict__ slot_elem,
                                                             const uint8_t *__restrict__ slot_lrow,
                                                             const int32_t *__restrict__ slot_pos,
                                                             const int32_t *__restrict__ indptr,
                                                             const uint8_t *__restrict__ row_free, int64_t n_rows,
                                                             double r, const double *__restrict__ elastic,
                                                             double *__restrict__ out, float *__restrict__ out32) {
    extern __shared__ double lds[];
    const int64_t r0 = (int64_t)blockIdx.x * kRowsKe;
    const int64_t r1 = r0 + kRowsKe < n_rows ? r0 + kRowsKe : n_rows;
    const int64_t v0 = indptr[r0];
    const int64_t L = (int64_t)indptr[r1] - v0;
    double *buf = USE_LDS ? lds : out + v0;
    for (int64_t i = threadIdx.x; i < L; i += kRowsKe) buf[i] = 0.0;
    __syncthreads();
    const int64_t row = r0 + threadIdx.x;
    if (row < r1) {
        if (!row_free[row]) {
            buf[indptr[row] - v0] = 1.0;
        } else {
            const int32_t s1 = row_slot_ptr[row + 1];
            for (int32_t s = row_slot_ptr[row]; s < s1; ++s) {
                if constexpr (NLD > 0 && USE_LDS) {
                    constexpr int CH = NLD % 15 == 0 ? 15 : NLD;
                    const double *kr = ke + ((int64_t)slot_elem[s] * NLD + slot_lrow[s]) * NLD;
                    const int32_t *pos = slot_pos + (int64_t)s * NLD;
                    for (int b0 = 0; b0 < NLD; b0 += CH) {
                        int32_t pv[CH];
                        double kv[CH];
#pragma unroll
                        for (int b = 0; b < CH; ++b) {
                            pv[b] = pos[b0 + b];
                            kv[b] = kr[b0 + b];
                        }
#pragma unroll
                        for (int b = 0; b < CH; ++b) {
                            const int64_t idx = pv[b] >= 0 ? (int64_t)pv[b] - v0 : L;
                            buf[idx] += kv[b];
                        }
                    }
                } else {
                    const double *kr = ke + ((int64_t)slot_elem[s] * nld + slot_lrow[s]) * nld;
                    const int32_t *pos = slot_pos + (int64_t)s * nld;
                    for (int b = 0; b < nld; ++b) {
                        const int32_t p = pos[b];
                        if (p >= 0) buf[p - v0] += kr[b];
                    }
                }
            }
        }
    }
    __syncthreads();
    const double om = 1.0 - r;
    for (int64_t i = threadIdx.x; i < L; i += kRowsKe) {
        double v = buf[i];
        if (elastic) v = om * v + r * elastic[v0 + i];
        if (USE_LDS) out[v0 + i] = v;
        else buf[i] = v;
        if (out32) out32[v0 + i] = (float)v;
    }
}

template <int NLD>
int launch_scatter(const double *ke, int nld, const int32_t *row_slot_ptr, const int32_t *slot_elem,
                   const uint8_t *slot_lrow, const int32_t *slot_pos, const int32_t *indptr, const uint8_t *row_free,
                   int64_t n_rows, double r, const double *elastic, double *out, float *out32, int64_t max_block_nnz,
                   hipStream_t st) {
    const unsigned grid = ssx::grid_for(n_rows, kRowsKe);
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
        lds_ok = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_scatter_ke_rows<NLD, true>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes) == hipSuccess;
    if (lds_ok)
        hipLaunchKernelGGL((k_scatter_ke_rows<NLD, true>), dim3(grid), dim3(kRowsKe), lds_bytes, st, ke, nld,
                           row_slot_ptr, slot_elem, slot_lrow, slot_pos, indptr, row_free, n_rows, r, elastic, out,
                           out32);
    else
        hipLaunchKernelGGL((k_scatter_ke_rows<NLD, false>), dim3(grid), dim3(kRowsKe), 0, st, ke, nld, row_slot_ptr,
                           slot_elem, slot_lrow, slot_pos, indptr, row_free, n_rows, r, elastic, out, out32);
    SSX_CHECK_LAUNCH("ssx_scatter_element_rows");
    return SSX_OK;
}

}  // namespace

extern "C" {

int ssx_element_matrices_p4(const double *dphi, const double *DS, const double *w, int64_t n_elem, double r,
                            const double *shear, const double *lame, double *ke, void *stream) {
    SSX_REQUIRE(n_elem >= 0, "ssx_element_matrices_p4: n_elem < 0");
    if (n_elem == 0) return SSX_OK;
    SSX_REQUIRE(dphi && DS && w && ke, "ssx_element_matrices_p4: null pointer");
    SSX_REQUIRE((shear == nullptr) == (lame == nullptr), "ssx_element_matrices_p4: shear and lame go together");
    hipLaunchKernelGGL(k_ke_p4_mfma, dim3((unsigned)n_elem), dim3(64), 0, ssx::as_stream(stream), dphi, DS, w, n_elem,
                       n_elem * NQ4, r, shear, lame, ke);
    SSX_CHECK_LAUNCH("ssx_element_matrices_p4");
    return SSX_OK;
}

int ssx_scatter_element_rows(const double *ke, int nld, const int32_t *row_slot_ptr, const int32_t *slot_elem,
                             const uint8_t *slot_lrow, const int32_t *slot_pos, const int32_t *indptr,
                             const uint8_t *row_free, int64_t n_rows, double r, const double *elastic, double *out,
                             float *out32, int64_t max_block_nnz, void *stream) {
    SSX_REQUIRE(n_rows >= 0 && nld > 0, "ssx_scatter_element_rows: bad sizes");
    if (n_rows == 0) return SSX_OK;
    SSX_REQUIRE(ke && row_slot_ptr && indptr && row_free && out, "ssx_scatter_element_rows: null pointer");
    hipStream_t st = ssx::as_stream(stream);
#define CALL(N) launch_scatter<N>(ke, nld, row_slot_ptr, slot_elem, slot_lrow, slot_pos, indptr, row_free, n_rows, r, \
                                  elastic, out, out32, max_block_nnz, st)
    if (nld == 12) return CALL(12);
    if (nld == 30) return CALL(30);
    if (nld == 105) return CALL(105);
    return CALL(0);
#undef CALL
}

}  // extern "C"
