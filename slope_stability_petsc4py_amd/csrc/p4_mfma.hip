// P4 tetrahedra: element tangent matrices on the FP64 matrix cores.
//
// ke = sum_q B_q^T (w_q D_q) B_q is one GEMM per element, ke (105 x 105) = A (105 x 144) * Bm (144 x 105)
// with K = 24 points x 6 strains, A[i][k] = B_q[s][i] and Bm[k][j] = sum_t w_q D_q[s][t] B_q[t][j]
// (k = 6 q + s).  The 105 x 105 result is padded to 7 x 7 tiles of 16 x 16 and each tile
// accumulates 36 v_mfma_f64_16x16x4 steps.  Both operands are formed from the element's dphi and
// w*DS staged in LDS (B is sparse: every column of B_q has 3 nonzeros, so an operand value costs
// <= 3 FMAs).  DS is symmetric for associated flow, so only the upper tile triangle (28 of 49
// tiles) is computed and mirrored on write.
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

// One workgroup of four wavefronts per element.  The Bm operand (w D_r B, 144 x 112 padded) is
// formed once per element into LDS, in three chunks of KC = 48 k values (8 points), fragment-major:
// sb[tj][ks][lane] is the value lane `lane` feeds v_mfma_f64_16x16x4 at k-step ks of tile column
// tj (k = 4 ks + lane / 16, j = 16 tj + lane % 16), so each MFMA reads one contiguous 512-byte
// fragment.  The A operand (B^T) costs a few VALU operations per k-step and tile row, and is formed
// in registers.  The 28 upper tiles are split 7 per wavefront: wavefront w takes tile row w (tiles
// w..6) and tile row 7 - w (tiles 7 - w..6), so its slot t is (w, w + t) for t < 7 - w, else
// (7 - w, t).  LDS: 20,160 B gradients + 6,912 B w D + 43,008 B fragments = 70,080 B, so two
// elements (eight wavefronts) share a CU and one element's operand formation overlaps the other's
// MFMAs.
constexpr int NPAD = 16 * TILES;   // 112
constexpr int KC = 48;             // k values per LDS chunk (8 points x 6 strains)
constexpr int KSC = KC / 4;        // MFMA k-steps per chunk
constexpr int NCHUNK = KDIM / KC;  // 3

__global__ void __launch_bounds__(256, 2) k_ke_p4_mfma(const double *__restrict__ dphi, const double *__restrict__ DS,
                                                      const double *__restrict__ w, int64_t n_elem, int64_t n_int,
                                                      double r, const double *__restrict__ shear,
                                                      const double *__restrict__ lame, double *__restrict__ ke) {
    __shared__ double sg[3][NQ4][NP4];        // gradients of the element's basis functions
    __shared__ double sd[NQ4][36];            // w_q * D_q (column-major 6x6)
    __shared__ double sb[TILES][KSC][64];     // Bm fragments of one chunk
    const int64_t e = blockIdx.x;
    if (e >= n_elem) return;
    const int tid = threadIdx.x;
    const int wv = tid >> 6, lane = tid & 63;
    const int64_t g0 = e * NQ4;
    for (int t = tid; t < 3 * NQ4 * NP4; t += 256) {
        const int c = t / (NQ4 * NP4), rem = t % (NQ4 * NP4);
        sg[c][rem / NP4][rem % NP4] = dphi[((int64_t)c * n_int + g0) * NP4 + rem];
    }
    // w_q * D_q, or w_q * ((1 - r) DS_q + r D_el) with the elastic tangent
    // D_el = lame * vol vol^T + 2 G * ident (ident = diag(1, 1, 1, 1/2, 1/2, 1/2)) when the
    // regularisation is folded in (shear != nullptr)
    const bool fold = shear != nullptr;
    for (int t = tid; t < NQ4 * 36; t += 256) {
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

    const int col = lane & 15;   // i for A, j for B
    const int kq = lane >> 4;    // k offset within the 4-wide step
    const int tiA = wv, tiB = TILES - wv;          // tile rows of this wavefront (tiB = 7: none)
    const int nA = TILES - wv;                     // slots of tile row tiA
    // A-operand rows of the two tile rows (padded rows >= 105 use a valid node and a zero weight)
    const int iaA = 16 * tiA + col, iaB = 16 * (tiB < TILES ? tiB : 0) + col;
    const double vaA = iaA < NLD4 ? 1.0 : 0.0, vaB = (tiB < TILES && iaB < NLD4) ? 1.0 : 0.0;
    const int iacA = iaA < NLD4 ? iaA : NLD4 - 1, iacB = iaB < NLD4 ? iaB : NLD4 - 1;
    const int naA = iacA / 3, daA = iacA % 3, naB = iacB / 3, daB = iacB % 3;

    d4 acc[TILES];
#pragma unroll
    for (int t = 0; t < TILES; ++t) acc[t] = d4{0.0, 0.0, 0.0, 0.0};
    double *sbf = &sb[0][0][0];
    for (int ch = 0; ch < NCHUNK; ++ch) {
        __syncthreads();   // staging done (first chunk) / the previous chunk's fragments are read
        // Bm fragments of this chunk: element f = (tj * KSC + ksl) * 64 + fl
        for (int f = tid; f < TILES * KSC * 64; f += 256) {
            const int fl = f & 63, tk = f >> 6;
            const int tj = tk / KSC, ksl = tk - (tk / KSC) * KSC;
            const int k = ch * KC + 4 * ksl + (fl >> 4);
            const int q = k / 6, s = k - (k / 6) * 6;
            const int jb = 16 * tj + (fl & 15);
            const double vb = jb < NLD4 ? 1.0 : 0.0;
            const int jbc = jb < NLD4 ? jb : NLD4 - 1;
            const int nb = jbc / 3, db = jbc % 3;
            double cb[6];
            bcol(db, sg[0][q][nb], sg[1][q][nb], sg[2][q][nb], cb);
            // row s of (wD) B: sum_t wD[s][t] B[t][jb], D column-major (D[s + 6 t])
            const double *D = sd[q];
            double b = D[s] * cb[0];
#pragma unroll
            for (int t = 1; t < 6; ++t) b += D[s + 6 * t] * cb[t];
            sbf[f] = vb * b;
        }
        __syncthreads();
        for (int ksl = 0; ksl < KSC; ++ksl) {
            const int k = ch * KC + 4 * ksl + kq;
            const int q = k / 6, s = k - (k / 6) * 6;
            double ca[6];
            bcol(daA, sg[0][q][naA], sg[1][q][naA], sg[2][q][naA], ca);
            const double aA = vaA * pick6(s, ca);
            bcol(daB, sg[0][q][naB], sg[1][q][naB], sg[2][q][naB], ca);
            const double aB = vaB * pick6(s, ca);
#pragma unroll
            for (int t = 0; t < TILES; ++t) {
                const bool first = t < nA;                // wavefront-uniform
                const int tj = first ? tiA + t : t;
                const double b = sbf[(tj * KSC + ksl) * 64 + lane];
                acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(first ? aA : aB, b, acc[t], 0, 0, 0);
            }
        }
    }
    // accumulator layout of v_mfma_f64_16x16x4f64: register r of lane l holds row
    // (l / 16) + 4 r of column l % 16
    double *out = ke + e * (int64_t)(NLD4 * NLD4);
#pragma unroll
    for (int t = 0; t < TILES; ++t) {
        const bool first = t < nA;
        const int ti = first ? tiA : tiB, tj = first ? tiA + t : t;
        const int j = 16 * tj + col;
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
            const int i = 16 * ti + kq + 4 * rr;
            if (i < NLD4 && j < NLD4) {
                out[(int64_t)i * NLD4 + j] = acc[t][rr];
                if (tj > ti) out[(int64_t)j * NLD4 + i] = acc[t][rr];
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
    hipLaunchKernelGGL(k_ke_p4_mfma, dim3((unsigned)n_elem), dim3(256), 0, ssx::as_stream(stream), dphi, DS, w, n_elem,
                       n_elem * NQ4, r, shear, lame, ke);
    SSX_CHECK_LAUNCH("ssx_element_matrices_p4");
    return SSX_OK;
}

}  // extern "C"
