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

// No divergent branch selects an address or an LDS offset anywhere in this file (hipcc for gfx950
// has misplaced such selections under the wrong execution mask): the sparsity of B is read from
// bit tables into plain integer arithmetic, and zero entries are 0/1 weights on a valid load.

// One workgroup of four wavefronts per element.  The Bm operand (w D_r B, 144 x 112 padded) is
// formed once per element into LDS, in three chunks of KC = 48 k values (8 points), fragment-major:
// sb[tj][ks][lane] is the value lane `lane` feeds v_mfma_f64_16x16x4 at k-step ks of tile column
// tj (k = 4 ks + lane / 16, j = 16 tj + lane % 16), so each MFMA reads one contiguous 512-byte
// fragment.  A column of B has three nonzeros, so a Bm value costs three FMAs; an A value (an
// entry of B) is one gradient read or zero, formed in registers per k-step and tile row.  The 28 upper tiles are split 7 per wavefront: wavefront w takes tile row w (tiles
// w..6) and tile row 7 - w (tiles 7 - w..6), so its slot t is (w, w + t) for t < 7 - w, else
// (7 - w, t).  LDS: 20,160 B gradients + 6,912 B w D + 43,008 B fragments = 70,080 B, so two
// elements (eight wavefronts) share a CU and one element's operand formation overlaps the other's
// MFMAs.
constexpr int NPAD = 16 * TILES;   // 112

// Sparsity of B (strain rows e11 e22 e33 g12 g23 g13, displacement component d of a node):
// kStrainC, 2 bits at 3 s + d: the gradient component (0 x, 1 y, 2 z) in row s, 3 = zero.
constexpr uint64_t strain_c_table() {
    const int c[6][3] = {{0, 3, 3}, {3, 1, 3}, {3, 3, 2}, {1, 0, 3}, {3, 2, 1}, {2, 3, 0}};
    uint64_t v = 0;
    for (int s = 0; s < 6; ++s)
        for (int d = 0; d < 3; ++d) v |= (uint64_t)c[s][d] << (2 * (3 * s + d));
    return v;
}
constexpr uint64_t kStrainC = strain_c_table();
// Per component d (12 bits at 12 d): the three nonzero rows t1 < t2 < t3 of a column of B
// (kRowT, 4 bits each) and the gradient components c1, c2, c3 in them (kGradC).
//   d = 0: rows 0, 3, 5 <- x, y, z;  d = 1: rows 1, 3, 4 <- y, x, z;  d = 2: rows 2, 4, 5 <- z, y, x
constexpr uint64_t kRowT = (0x0 | 0x3 << 4 | 0x5 << 8) | (uint64_t)(0x1 | 0x3 << 4 | 0x4 << 8) << 12 |
                           (uint64_t)(0x2 | 0x4 << 4 | 0x5 << 8) << 24;
constexpr uint64_t kGradC = (0x0 | 0x1 << 4 | 0x2 << 8) | (uint64_t)(0x1 | 0x0 << 4 | 0x2 << 8) << 12 |
                            (uint64_t)(0x2 | 0x1 << 4 | 0x0 << 8) << 24;
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
    const double *sgf = &sg[0][0][0];
    const double *sdf = &sd[0][0];
    double *sbf = &sb[0][0][0];
    // Bm formation: thread f < 2 * NPAD forms column j = f % NPAD for the 24 k values (4 points x 6
    // strains) of half h = f / NPAD of each chunk; the last 32 threads do not form.  Column j (node
    // fn, component fd) of B_q has three nonzeros: rows t1 < t2 < t3 holding gradient components
    // c1, c2, c3 (kRowT / kGradC), so per point three gradient reads serve the six strains, and
    // Bm[k][j] = sum_i wD_q[s][t_i] g_{c_i}.
    constexpr int GS = NQ4 * NP4;                 // stride of a gradient component in sg
    const bool former = tid < 2 * NPAD;
    const int fj = tid % NPAD, fh = tid / NPAD;
    const double fv = fj < NLD4 ? 1.0 : 0.0;
    const int fjc = fj < NLD4 ? fj : NLD4 - 1;
    const int fn = fjc / 3, fd = fjc - 3 * (fjc / 3);
    const int ftd = (int)((kRowT >> (12 * fd)) & 0xfff), ftc = (int)((kGradC >> (12 * fd)) & 0xfff);
    const int o1 = 6 * (ftd & 15), o2 = 6 * ((ftd >> 4) & 15), o3 = 6 * (ftd >> 8);
    const int g1 = (ftc & 15) * GS + fn, g2 = ((ftc >> 4) & 15) * GS + fn, g3 = (ftc >> 8) * GS + fn;
    double *fout = sbf + (fj >> 4) * (KSC * 64) + (fj & 15);   // + ksl * 64 + kq * 16
    for (int ch = 0; ch < NCHUNK; ++ch) {
        __syncthreads();   // staging done (first chunk) / the previous chunk's fragments are read
        if (former) {
#pragma unroll
            for (int qi = 0; qi < 4; ++qi) {
                const int q = ch * (KC / 6) + fh * 4 + qi;
                const double *gq = sgf + q * NP4;
                const double G1 = gq[g1], G2 = gq[g2], G3 = gq[g3];
                const double *D = sdf + q * 36;
#pragma unroll
                for (int s6 = 0; s6 < 6; ++s6) {
                    const double b = D[s6 + o1] * G1 + D[s6 + o2] * G2 + D[s6 + o3] * G3;
                    const int kl = 6 * qi + s6;              // k within the half chunk (24 = 6 k-steps)
                    fout[(fh * 6 + (kl >> 2)) * 64 + (kl & 3) * 16] = fv * b;
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int ksl = 0; ksl < KSC; ++ksl) {
            const int k = ch * KC + 4 * ksl + kq;
            const int q = k / 6, s = k - 6 * (k / 6);
            // A value B_q[s][i]: gradient component kStrainC(s, d) of node n, or zero
            const int cA = (int)((kStrainC >> (2 * (3 * s + daA))) & 3);
            const int cB = (int)((kStrainC >> (2 * (3 * s + daB))) & 3);
            const double aA = (cA < 3 ? vaA : 0.0) * sgf[(cA < 3 ? cA : 0) * (NQ4 * NP4) + q * NP4 + naA];
            const double aB = (cB < 3 ? vaB : 0.0) * sgf[(cB < 3 ? cB : 0) * (NQ4 * NP4) + q * NP4 + naB];
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
