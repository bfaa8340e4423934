// P4 tetrahedra: element tangent matrices on the FP64 matrix cores.
//
// ke = sum_q B_q^T (w_q D_q) B_q is one GEMM per element, ke (105 x 105) = A (105 x 144) * Bm (144 x 105)
// with K = 24 points x 6 strains, A[i][k] = B_q[s][i] and Bm[k][j] = sum_t w_q D_q[s][t] B_q[t][j]
// (k = 6 q + s).  One wavefront per element; the 105 x 105 result is padded to 7 x 7 tiles of
// 16 x 16 and each tile accumulates 36 v_mfma_f64_16x16x4 steps.  Both operands are formed per lane
// from the element's dphi and w*DS staged in LDS (B is sparse: every column of B_q has 3 nonzeros,
// so an operand value costs <= 3 FMAs).  DS is symmetric for associated flow, so only the upper
// tile triangle (28 of 49 tiles) is computed and mirrored on write.
//
// The element matrices are then gathered per mesh node into the fixed CSR pattern
// (ssx_gather_node_rows in node_assembly.hip), in the reference's `rows` summation order
// (cython/assemble_tangent_values_3d.c:111-214).
//
// Reference behaviour: cython/assemble_tangent_values_3d.c:13-109 (element matrices); flop count
// 2*105*105*144 per element.
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

}  // extern "C"
