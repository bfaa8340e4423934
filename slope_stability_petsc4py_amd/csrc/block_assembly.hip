// Tangent assembly of P2 tetrahedra through symmetric node-pair blocks, on gfx950:
//
//   phase 1 (k_ke_blocks_p2)  one wavefront per element at a time; lane p < 55 owns the node pair
//                             (a, b), a <= b, and computes the 3x3 block K_e[3a + i, 3b + j] over
//                             the 11 points.  The point tangent D is the same for every lane of the
//                             wavefront, so it is read with scalar loads into SGPRs and every FMA
//                             takes its D operand from there (the gradients come from LDS); the
//                             blocks below the diagonal are not computed.
//   phase 2 (k_gather_blocks) one wavefront per mesh node sums the node's three rows over its
//                             elements (element order, an LDS copy of the rows) and writes them to
//                             the fixed CSR pattern in one coalesced pass (FP64 values and the FP32
//                             copy).  Row 3a + i of K_e at column 3b + j is block (a, b)[i][j] when
//                             b >= a, else block (b, a)[j][i].
//
// The blocks take 55 * 9 = 495 doubles per element against 900 for the full element matrix, so
// the intermediate written and read back is 0.55x the full matrices.  The lower blocks are the
// transposes of the upper ones: the Mohr-Coulomb tangent of the return mapping is symmetric
// (associated flow with the reduced sin(phi)) and so is the elastic tangent.
//
// Block (a, b) = sum_q w_q B_a^T D_q B_b with B_a the 6x3 strain-displacement block of node a:
// B_a[s, i] = g_a[k] for s = V[i][k] (Voigt index of the pair (i, k), engineering shears).  The
// lane forms t[i][s'] = sum_k g_a[k] D[V[i][k]][s'] (18 values) and adds sum_l t[i][V[j][l]] g_b[l].
// With the regularisation K_r = (1-r) K + r K_el folded in, t is scaled by w (1-r) and the elastic
// part w r B_a^T D_el B_b = w r (lame g_a[i] g_b[j] + G (g_a[j] g_b[i] + delta_ij g_a.g_b)) enters
// through t (D_el has the closed form lame m m^T + 2G on the normal and G on the shear diagonal).
//
// Summation order of a CSR entry: element blocks over the node's elements in element order (the
// order of the reference's rows kernel), each block summed over the points in order (with branch
// ids: the points the return mapping left elastic in order, then the others in order); bitwise
// reproducible.
//
// Reference behaviour: cython/assemble_tangent_values_3d.c:13-109 (element matrices), :111-214
// (rows); fem/distributed_tangent.py:1559-1612 (constrained diagonal = 1); constitutive/problem.py:
// 1595-1644 (K_r on the same pattern).
#include "ssx_common.h"

namespace {

constexpr int kNP = 10;                       // P2 tetrahedron
constexpr int kNQ = 11;                       // 11-point rule
constexpr int kNPair = kNP * (kNP + 1) / 2;   // 55 node pairs a <= b
constexpr int kBlk = 9;                       // 3x3 block

// Voigt index of the index pair (i, k): [e11 e22 e33 g12 g23 g13]
__host__ __device__ constexpr int voigt(int i, int k) {
    return i == k ? i : ((i + k == 1) ? 3 : ((i + k == 3) ? 4 : 5));
}

// Start of row a of the upper triangle (pairs (a, a), (a, a+1), ... in row-major order).
__host__ __device__ constexpr int pair_row_start(int a) { return a * kNP - a * (a - 1) / 2; }

// Word of D[s][m] in a point's tangent: s + 6 m (all 36), or the packed upper entry of (min, max).
template <bool SYM>
__host__ __device__ constexpr int ds_at(int s, int m) {
    return !SYM ? s + 6 * m
                : (s <= m ? s * 6 - s * (s - 1) / 2 + (m - s) : m * 6 - m * (m - 1) / 2 + (s - m));
}

// Value of lane k's v (k uniform), as two 32-bit reads.
__device__ __forceinline__ double lane_value(double v, int k) {
    const long long x = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)x, k), hi = __builtin_amdgcn_readlane((int)(x >> 32), k);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Lane -> node pair (a, b), a <= b: lanes 0..54 in row-major order over the upper triangle; idle
// lanes repeat the last pair (their stores are masked).
__device__ __forceinline__ void lane_pair(int lane, int &p, int &a, int &b) {
    p = lane < kNPair ? lane : kNPair - 1;
    a = 0;
#pragma unroll
    for (int k = 1; k < kNP; ++k) a += p >= pair_row_start(k) ? 1 : 0;
    b = a + (p - pair_row_start(a));
}

// Point data of one element staged in LDS: gradients [c][q][node] (330), then w, shear, lame and the
// elastic flag per point (11 each).
constexpr int kStG = 3 * kNQ * kNP;
constexpr int kStRegs = (kStG + 63) / 64;   // gradient values per lane in flight

// FOLD: shear/lame given, the regularisation r folded in (K_r); EL: per-point branch ids given, and a
// point the return mapping left elastic (branch 0) has D = D_el exactly, so its block is the closed
// form w (lame g_a g_b^T + G (g_b g_a^T + (g_a.g_b) I)) (and (1-r) D_el + r D_el = D_el).
//
// Persistent wavefronts walk elements e, e + grid, ...; the next element's point data is loaded
// into registers while the current element is computed from LDS and staged into the other LDS
// buffer afterwards; one load per lane touches every 128-byte line of the next element's D, so
// the scalar loads of D find it in the L2.
//
// GEO: the element has a constant Jacobian (affine P2 tetrahedron).  In place of its 330 gradient
// values and 11 weights the next element's 10 factors gm[e] = (J^-1 row-major, |det J|) are
// fetched (lanes 0..9), and staging forms the gradients from them and the rule's reference
// gradients h_k(q, n), held in LDS for the kernel's life: g[c][q][n] = sum_k gm[e][3 c + k]
// h_k(q, n) (six rounds of three FMAs per lane, one component c per round), w_q = gm[e][9] wf[q].
// The point loop is the same in both forms.
// SYM: DS holds the 21 entries D[s][m], s <= m, of each point (packed row by row) instead of 36; the
// lower entries are read from their mirror (the tangent is symmetric), so the point loop is the
// same with other constant offsets.
template <bool FOLD, bool EL, bool GEO, bool SYM = false>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4, 8)))
k_ke_blocks_p2(const double *__restrict__ dphi, const double *__restrict__ DS, const double *__restrict__ w,
               int64_t n_elem, double r, const double *__restrict__ shear, const double *__restrict__ lame,
               const int8_t *__restrict__ branch, const double *__restrict__ gm, const double *__restrict__ ref,
               double *__restrict__ kb) {
    // per-element staging: gradients [c][q][n], then w, shear, lame and the elastic flag per point
    constexpr int DW = SYM ? 21 : 36;              // words of a point's D
    constexpr int OFF = kStG;
    constexpr int NREG = GEO ? 0 : kStRegs;
    constexpr int NQN = kNQ * kNP;                 // 110 values per component
    __shared__ double sg[2][OFF + 4 * kNQ + 1];   // + 1: the element's material is uniform (EL)
    __shared__ double sh[GEO ? kStG + kNQ : 1];   // GEO: h_k(q, n) at k NQN + q kNP + n, then wf
    const int lane = threadIdx.x;
    int p, a, b;
    lane_pair(lane, p, a, b);
    const int64_t n_int = n_elem * kNQ;
    const double om = FOLD ? 1.0 - r : 1.0;
    int64_t e = blockIdx.x;
    if (e >= n_elem) return;   // whole wavefront
    if constexpr (GEO) {
        for (int i = lane; i < kStG + kNQ; i += 64) sh[i] = ref[i];
        __syncthreads();
    }

    double rg[NREG > 0 ? NREG : 1], rw = 0.0, rs = 0.0, rl = 0.0, rgm = 0.0, touch;
    int rb = 1;
    // global -> registers for element ee (every index in bounds; lanes past the data read index 0
    // or the element's last word)
    auto fetch = [&](int64_t ee) {
#pragma unroll
        for (int k = 0; k < NREG; ++k) {
            const int t = lane + 64 * k;
            const int tc = t < kStG ? t : 0;
            const int c = tc / (kNQ * kNP), rem = tc - c * (kNQ * kNP);
            rg[k] = dphi[((int64_t)c * n_int + ee * kNQ) * kNP + rem];
        }
        const int lq = lane < kNQ ? lane : 0;
        if constexpr (!GEO) rw = w[ee * kNQ + lq];
        if constexpr (FOLD || EL) {
            rs = shear[ee * kNQ + lq];
            rl = lame[ee * kNQ + lq];
        }
        if constexpr (EL) rb = branch[ee * kNQ + lq];
        if constexpr (GEO) rgm = gm[ee * 10 + (lane < 10 ? lane : 9)];
        touch = DS[ee * (kNQ * DW) + (lane * 16 < kNQ * DW - 1 ? lane * 16 : kNQ * DW - 1)];
    };
    // registers -> LDS buffer bb
    auto stage = [&](int bb) {
        asm volatile("" ::"v"(touch));   // the line touch is complete (and kept) before the buffer turns
#pragma unroll
        for (int k = 0; k < NREG; ++k) {
            const int t = lane + 64 * k;
            if (t < kStG) sg[bb][t] = rg[k];
        }
        if constexpr (GEO) {
            double m[10];
#pragma unroll
            for (int k = 0; k < 10; ++k) m[k] = lane_value(rgm, k);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int rem = lane + 64 * h;
                    if (rem < NQN)
                        sg[bb][c * NQN + rem] = fma(m[3 * c + 2], sh[2 * NQN + rem],
                                                    fma(m[3 * c + 1], sh[NQN + rem], m[3 * c] * sh[rem]));
                }
            }
            rw = m[9] * sh[kStG + (lane < kNQ ? lane : 0)];
        }
        if (lane < kNQ) {
            sg[bb][OFF + lane] = rw;
            sg[bb][OFF + kNQ + lane] = rs;
            sg[bb][OFF + 2 * kNQ + lane] = rl;
            sg[bb][OFF + 3 * kNQ + lane] = rb == 0 ? 1.0 : 0.0;
        }
        if constexpr (EL) {
            // lanes past the 11 points hold point 0's values
            const double s0 = lane_value(rs, 0), l0 = lane_value(rl, 0);
            const bool uni = __all(rs == s0 && rl == l0);
            if (lane == 0) sg[bb][OFF + 4 * kNQ] = uni ? 1.0 : 0.0;
        }
    };
    fetch(e);
    stage(0);
    __syncthreads();
    int buf = 0;
    for (; e < n_elem; e += gridDim.x) {
        const int64_t en = e + gridDim.x;
        if (en < n_elem) fetch(en);
        const double *L = sg[buf];
        double acc[kBlk];
#pragma unroll
        for (int k = 0; k < kBlk; ++k) acc[k] = 0.0;
        auto grads = [&](int q, double ha[3], double gb[3]) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                ha[d] = L[d * NQN + q * kNP + a];
                gb[d] = L[d * NQN + q * kNP + b];
            }
        };
        // Two point loops, each with one arithmetic path and a uniform skip: the points the return
        // mapping left elastic (closed form), then the others.  (One loop with both paths makes the
        // compiler keep the two paths' accumulators in different registers and copy all nine
        // between them at the join and on the back edge, 18 moves per point.)
        if constexpr (EL) {
            if (L[OFF + 4 * kNQ] != 0.0) {
                // One material over the element (uniform): M = sum over the elastic points of
                // w g_a g_b^T (9 FMAs per point), then the closed form once,
                // lame M + G (M^T + tr(M) I).
                double M[kBlk];
#pragma unroll
                for (int k = 0; k < kBlk; ++k) M[k] = 0.0;
#pragma unroll 1
                for (int q = 0; q < kNQ; ++q) {
                    if (L[OFF + 3 * kNQ + q] == 0.0) continue;   // uniform: one point for the whole wavefront
                    const double wq = L[OFF + q];
                    double ha[3], gb[3];
                    grads(q, ha, gb);
#pragma unroll
                    for (int i = 0; i < 3; ++i) {
                        const double wa = wq * ha[i];
#pragma unroll
                        for (int j = 0; j < 3; ++j) M[3 * i + j] = fma(wa, gb[j], M[3 * i + j]);
                    }
                }
                const double lm = L[OFF + 2 * kNQ], gs = L[OFF + kNQ];
                const double tr = M[0] + M[4] + M[8];
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j)
                        acc[3 * i + j] = fma(lm, M[3 * i + j], gs * (M[3 * j + i] + (i == j ? tr : 0.0)));
            } else {
#pragma unroll 1
                for (int q = 0; q < kNQ; ++q) {
                    if (L[OFF + 3 * kNQ + q] == 0.0) continue;   // uniform: one point for the whole wavefront
                    const double wq = L[OFF + q];
                    double ha[3], gb[3];
                    grads(q, ha, gb);
                    const double wl = wq * L[OFF + 2 * kNQ + q], wg = wq * L[OFF + kNQ + q];
                    const double dg = wg * fma(ha[2], gb[2], fma(ha[1], gb[1], ha[0] * gb[0]));
#pragma unroll
                    for (int i = 0; i < 3; ++i) {
                        const double li = wl * ha[i], gi = wg * gb[i];
#pragma unroll
                        for (int j = 0; j < 3; ++j)
                            acc[3 * i + j] = fma(gi, ha[j], fma(li, gb[j], acc[3 * i + j]));
                        acc[4 * i] += dg;
                    }
                }
            }
        }
#pragma unroll 1
        for (int q = 0; q < kNQ; ++q) {
            if constexpr (EL) {
                if (L[OFF + 3 * kNQ + q] != 0.0) continue;   // uniform: taken in the first loop
            }
            const double wq = L[OFF + q];
            double ha[3], gb[3];
            grads(q, ha, gb);
            const double *D = DS + (e * kNQ + q) * DW;   // D[s][m] = D[ds_at(s, m)]; uniform: scalar loads
            const double c1 = wq * om;
            const double ga[3] = {c1 * ha[0], c1 * ha[1], c1 * ha[2]};
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                // t[s'] = sum_k ga[k] D[V[i][k]][s'] (row i of w (1-r) B_a^T D)
                double t[6];
#pragma unroll
                for (int m = 0; m < 6; ++m) {
                    double v = ga[0] * D[ds_at<SYM>(voigt(i, 0), m)];
                    v = fma(ga[1], D[ds_at<SYM>(voigt(i, 1), m)], v);
                    v = fma(ga[2], D[ds_at<SYM>(voigt(i, 2), m)], v);
                    t[m] = v;
                }
                if constexpr (FOLD) {
                    // + w r (B_a^T D_el) row i: lame g_a[i] on the normal strains, 2G g_a[i] more on
                    // strain i, G g_a[k] on the shear strain of (i, k), k != i
                    const double c2 = wq * r;
                    const double cl = c2 * L[OFF + 2 * kNQ + q], cg = c2 * L[OFF + kNQ + q];
                    const double li = cl * ha[i];
#pragma unroll
                    for (int m = 0; m < 3; ++m) t[m] += li;
                    t[i] = fma(2.0 * cg, ha[i], t[i]);
#pragma unroll
                    for (int k = 0; k < 3; ++k)
                        if (k != i) t[voigt(i, k)] = fma(cg, ha[k], t[voigt(i, k)]);
                }
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    double v = acc[3 * i + j];
                    v = fma(t[voigt(j, 0)], gb[0], v);
                    v = fma(t[voigt(j, 1)], gb[1], v);
                    v = fma(t[voigt(j, 2)], gb[2], v);
                    acc[3 * i + j] = v;
                }
            }
        }
        // The next element is staged before this element's blocks are stored: its loads were
        // issued before the point loops, and the wait for them (in issue order) then does not also
        // wait for the stores.
        if (en < n_elem) stage(buf ^ 1);
        if (lane < kNPair) {
            double *o = kb + (e * kNPair + p) * kBlk;
#pragma unroll
            for (int k = 0; k < kBlk; ++k) o[k] = acc[k];
        }
        __syncthreads();
        buf ^= 1;
    }
}

// ------------------------------------------------------------------------------------------
// Phase 2: node gather from the blocks.  LDS holds the node's three rows, each of length L + 1
// (free rows share one length L; word L of a row absorbs its constrained columns, whose offset is
// L), plus one spare word.  Lane v (and v + 64) handles row c = v / 30 and element column 3b + j =
// v % 30.  For the element's local node a, K_e[3a + c, 3b + j] is block (a, b)[c][j] when b >= a,
// else block (b, a)[j][c]; its offset in the element's 495 words splits into a part that depends
// only on a (the same for the wavefront) and a part that depends only on the lane:
//   b >= a:  9 (p0(a) - a) + (9b + 3c + j),      b < a:  9a + (9 (p0(b) - b) + 3j + c),
// with p0(x) the first pair of row x of the upper triangle.
//
// The node's slot metadata (element, local node) for up to 64 slots comes in one vector load and
// each slot's pair is read from its lane (readlane).  Slots go two at a time: both slots' loads are
// issued, then both are added in slot order.  Every add is unconditional -- a slot past the node's
// last one and the lanes past the 90 values add into the spare word -- so the loads feed no branch
// and stay in flight together (a conditional add lets the compiler sink each slot's loads into its
// branch, which serialises them).  LDS operations of one wavefront complete in order, so slot
// order needs no barrier beyond the wavefront's own.
// ------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(64) k_gather_blocks(const double *__restrict__ kb, const int32_t *__restrict__ ns_ptr,
                                                      const int32_t *__restrict__ ns_elem,
                                                      const uint8_t *__restrict__ ns_lnode,
                                                      const int16_t *__restrict__ ns_off,
                                                      const int32_t *__restrict__ indptr,
                                                      const uint8_t *__restrict__ row_free,
                                                      const int32_t *__restrict__ nodes, double r,
                                                      const double *__restrict__ elastic, double *__restrict__ out,
                                                      float *__restrict__ out32) {
    constexpr int DIM = 3, NLD = 3 * kNP, NV = DIM * NLD, NPASS = (NV + 63) / 64, DEPTH = 2;
    extern __shared__ double acc[];
    const int64_t node = nodes ? (int64_t)nodes[blockIdx.x] : (int64_t)blockIdx.x;
    const int lane = threadIdx.x;
    const int64_t r0 = DIM * node;
    int fr[DIM];
    int64_t st[DIM];
    int L = 0;
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
        fr[c] = row_free[r0 + c];
        st[c] = indptr[r0 + c];
        const int len = indptr[r0 + c + 1] - indptr[r0 + c];
        L = fr[c] ? len : L;
    }
    const int32_t s0 = ns_ptr[node], s1 = ns_ptr[node + 1];
    const int L1 = L + 1;
    const int spare = DIM * L1;
    for (int i = lane; i <= spare; i += 64) acc[i] = 0.0;

    int rb[NPASS], col[NPASS], bn[NPASS], pup[NPASS], plo[NPASS];
#pragma unroll
    for (int k = 0; k < NPASS; ++k) {
        const int v = lane + 64 * k;
        const int vc = v < NV ? v : NV - 1;   // lanes past the values read the last one (added to the spare word)
        const int c = vc / NLD;
        col[k] = vc - c * NLD;
        const int b = col[k] / 3, j = col[k] - 3 * b;
        bn[k] = b;
        pup[k] = kBlk * b + 3 * c + j;
        plo[k] = kBlk * (pair_row_start(b) - b) + 3 * j + c;
        rb[k] = v < NV ? c * L1 : -1;
    }
    __syncthreads();
    for (int32_t cs = s0; cs < s1; cs += 64) {
        const int nc = (s1 - cs) < 64 ? (s1 - cs) : 64;
        const int ml = lane < nc ? lane : nc - 1;
        const int my_e = ns_elem[cs + ml];
        const int my_a = ns_lnode[cs + ml];
        for (int base = 0; base < nc; base += DEPTH) {
            double val[DEPTH][NPASS];
            int o[DEPTH][NPASS];
#pragma unroll
            for (int d = 0; d < DEPTH; ++d) {
                const int k = base + d < nc ? base + d : nc - 1;
                const int64_t e = __builtin_amdgcn_readlane(my_e, k);
                const int a = __builtin_amdgcn_readlane(my_a, k);
                const double *bp = kb + e * (kNPair * kBlk);
                const int up0 = kBlk * (pair_row_start(a) - a), lo0 = kBlk * a;
                const int16_t *off = ns_off + (int64_t)(cs + k) * NLD;
#pragma unroll
                for (int kk = 0; kk < NPASS; ++kk) {
                    val[d][kk] = bp[bn[kk] >= a ? up0 + pup[kk] : lo0 + plo[kk]];
                    o[d][kk] = off[col[kk]];
                }
            }
#pragma unroll
            for (int d = 0; d < DEPTH; ++d) {
                const bool slot_ok = base + d < nc;   // uniform
#pragma unroll
                for (int kk = 0; kk < NPASS; ++kk) {
                    const int idx = (slot_ok && rb[kk] >= 0) ? rb[kk] + o[d][kk] : spare;
                    acc[idx] += val[d][kk];
                }
                __syncthreads();
            }
        }
    }
    const double om = 1.0 - r;
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
        if (fr[c]) {
            for (int i = lane; i < L; i += 64) {
                double v = acc[c * L1 + i];
                if (elastic) v = om * v + r * elastic[st[c] + i];
                out[st[c] + i] = v;
                if (out32) out32[st[c] + i] = (float)v;
            }
        } else if (lane == 0) {
            out[st[c]] = 1.0;  // constrained row: its diagonal only
            if (out32) out32[st[c]] = 1.0f;
        }
    }
}

// Persistent grid of the block kernel: 64 wavefronts per CU, about 3.5 times the resident count
// (LDS holds 18 per CU).  Measured on MI355X on the 117k-element bench mesh: 300 us at 32, 270-275
// us at 64-128, 285 us with one element per wavefront; later wavefronts start as earlier ones end,
// which keeps the CUs' memory and arithmetic phases from running in step.
constexpr int kBlkWavesPerCU = 64;

unsigned blocks_grid(int64_t n_elem) {
    static int n_cu = 0;
    if (n_cu == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n_cu <= 0)
            n_cu = 256;
    }
    const int64_t want = (int64_t)n_cu * kBlkWavesPerCU;
    return (unsigned)(n_elem < want ? n_elem : want);
}

}  // namespace

extern "C" {

int ssx_element_blocks_3d(int n_p, const double *dphi, const double *DS, const double *w, int64_t n_elem, int n_q,
                          double r, const double *shear, const double *lame, const int8_t *branch, double *kb,
                          void *stream) {
    SSX_REQUIRE(n_p == kNP && n_q == kNQ, "ssx_element_blocks_3d: only P2 tetrahedra with the 11-point rule "
                "(n_p 10, n_q 11), got n_p %d n_q %d", n_p, n_q);
    SSX_REQUIRE(n_elem >= 0 && n_elem <= 0x7fffffff, "ssx_element_blocks_3d: bad element count");
    if (n_elem == 0) return SSX_OK;
    SSX_REQUIRE(dphi && DS && w && kb, "ssx_element_blocks_3d: null pointer");
    SSX_REQUIRE((shear == nullptr) == (lame == nullptr), "ssx_element_blocks_3d: shear and lame go together");
    SSX_REQUIRE(branch == nullptr || shear != nullptr, "ssx_element_blocks_3d: branch ids need shear and lame");
    hipStream_t st = ssx::as_stream(stream);
    const unsigned grid = blocks_grid(n_elem);
    // FOLD (shear/lame given) applies r; with r = 0 and branch ids the fold terms vanish
#define KB(F_, E_)                                                                                          \
    hipLaunchKernelGGL((k_ke_blocks_p2<F_, E_, false>), dim3(grid), dim3(64), 0, st, dphi, DS, w, n_elem, r, shear, \
                       lame, branch, nullptr, nullptr, kb)
    if (branch) KB(true, true);
    else if (shear) KB(true, false);
    else KB(false, false);
#undef KB
    SSX_CHECK_LAUNCH("ssx_element_blocks_3d");
    return SSX_OK;
}

int ssx_element_blocks_affine_3d(int n_p, const double *gm, const double *ref, const double *DS, int ds_sym,
                                 int64_t n_elem, int n_q, double r, const double *shear, const double *lame,
                                 const int8_t *branch, double *kb, void *stream) {
    SSX_REQUIRE(n_p == kNP && n_q == kNQ, "ssx_element_blocks_affine_3d: only P2 tetrahedra with the 11-point rule "
                "(n_p 10, n_q 11), got n_p %d n_q %d", n_p, n_q);
    SSX_REQUIRE(n_elem >= 0 && n_elem <= 0x7fffffff, "ssx_element_blocks_affine_3d: bad element count");
    if (n_elem == 0) return SSX_OK;
    SSX_REQUIRE(gm && ref && DS && kb, "ssx_element_blocks_affine_3d: null pointer");
    SSX_REQUIRE((shear == nullptr) == (lame == nullptr), "ssx_element_blocks_affine_3d: shear and lame go together");
    SSX_REQUIRE(branch == nullptr || shear != nullptr, "ssx_element_blocks_affine_3d: branch ids need shear and lame");
    hipStream_t st = ssx::as_stream(stream);
    const unsigned grid = blocks_grid(n_elem);
#define KB(F_, E_, S_)                                                                                   \
    hipLaunchKernelGGL((k_ke_blocks_p2<F_, E_, true, S_>), dim3(grid), dim3(64), 0, st, nullptr, DS, nullptr, \
                       n_elem, r, shear, lame, branch, gm, ref, kb)
    if (ds_sym) {
        if (branch) KB(true, true, true);
        else if (shear) KB(true, false, true);
        else KB(false, false, true);
    } else {
        if (branch) KB(true, true, false);
        else if (shear) KB(true, false, false);
        else KB(false, false, false);
    }
#undef KB
    SSX_CHECK_LAUNCH("ssx_element_blocks_affine_3d");
    return SSX_OK;
}

int ssx_gather_node_blocks(int n_p, const double *kb, const int32_t *ns_ptr, const int32_t *ns_elem,
                           const uint8_t *ns_lnode, const int16_t *ns_off, const int32_t *indptr,
                           const uint8_t *row_free, const int32_t *nodes, int64_t n_nodes, int max_len, double r,
                           const double *elastic, double *out, float *out32, void *stream) {
    SSX_REQUIRE(n_p == kNP, "ssx_gather_node_blocks: only P2 tetrahedra (n_p 10), got %d", n_p);
    SSX_REQUIRE(n_nodes >= 0 && n_nodes <= 0x7fffffff && max_len >= 0 && max_len < 32768,
                "ssx_gather_node_blocks: bad sizes");
    if (n_nodes == 0) return SSX_OK;
    SSX_REQUIRE(kb && ns_ptr && ns_elem && ns_lnode && ns_off && indptr && row_free && out,
                "ssx_gather_node_blocks: null pointer");
    const size_t lds = ((size_t)3 * (max_len + 1) + 1) * sizeof(double);
    SSX_REQUIRE(lds <= 65536, "ssx_gather_node_blocks: rows of %d entries exceed the LDS", max_len);
    hipStream_t st = ssx::as_stream(stream);
    hipLaunchKernelGGL(k_gather_blocks, dim3((unsigned)n_nodes), dim3(64), lds, st, kb, ns_ptr, ns_elem, ns_lnode,
                       ns_off, indptr, row_free, nodes, r, elastic, out, out32);
    SSX_CHECK_LAUNCH("ssx_gather_node_blocks");
    return SSX_OK;
}

}  // extern "C"
