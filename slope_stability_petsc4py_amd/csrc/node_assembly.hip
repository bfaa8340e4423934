// Two-phase tangent assembly for tetrahedra on gfx950:
//
//   phase 1 (k_ke_3d)        one wavefront per element computes the element matrix
//                            K_e = sum_q w_q B_q^T D_q B_q (dim*n_p square, row-major) once;
//   phase 2 (k_gather_nodes) one wavefront per mesh node gathers the node's dim rows from the
//                            element matrices of the elements around it, in element order, into an
//                            LDS copy of those rows, then writes them to the fixed CSR pattern with
//                            one coalesced pass (FP64 values and the FP32 copy).
//
// The regularised operator K_r = (1-r) K_t + r K_el is linear in D, so phase 1 can fold it into
// the point tangent: D_r = (1-r) DS + r D_el with D_el = lame * vol vol^T + 2 G * ident (the elastic
// tangent of the return mapping).  That replaces the read of the assembled elastic values in the
// blend; the blend form (phase 2 reading K_el) is kept for element matrices made elsewhere.
//
// Free rows of one node share their column list (the pattern rule: sorted unique free columns of
// every element touching the node), so the CSR positions of an element's columns are stored once
// per (node, element) pair as int16 offsets from the row start: n_p*dim*2 bytes instead of
// dim*(n_p*dim*4) bytes of per-row positions.  Constrained rows hold only their diagonal (= 1).
//
// Summation order of a CSR entry: element matrices are summed over the node's elements in element
// order, each element matrix being the sum over its points -- the order of the reference's `rows`
// kernel (cython/assemble_tangent_values_3d.c:111-214).  Every LDS update of one element completes
// (barrier) before the next element's updates start, so results are bitwise reproducible.
//
// Reference behaviour: cython/assemble_tangent_values_3d.c:13-109 (element matrices), :111-214
// (rows); fem/distributed_tangent.py:1559-1612 (constrained diagonal = 1); constitutive/problem.py:
// 1595-1644 (K_r on the same pattern).
#include "ssx_common.h"

namespace {

// Row a = 3*node + c of B^T for a node with gradient (x, y, z), as coefficients on the strain
// rows [e11 e22 e33 g12 g23 g13]; branch-free (0/1 weights): the same code for every c.
__device__ __forceinline__ void brow(int c, double x, double y, double z, double b[6]) {
    const double c0 = c == 0 ? 1.0 : 0.0, c1 = c == 1 ? 1.0 : 0.0, c2 = c == 2 ? 1.0 : 0.0;
    b[0] = c0 * x;
    b[1] = c1 * y;
    b[2] = c2 * z;
    b[3] = c0 * y + c1 * x;
    b[4] = c1 * z + c2 * y;
    b[5] = c0 * z + c2 * x;
}

// ------------------------------------------------------------------------------------------
// Phase 1: element matrices of P1/P2 tetrahedra.  Lane (h, a): a = lane % 32 is the element row
// (a < 3*NP), h = lane / 32 selects the column half (nodes NH*h .. NH*h + NH - 1).  Per point the
// lane forms t = w (D_r b_a) (six values) and adds t . B for its 3*NH columns.  The element's point
// data (D, w, material) is the same for every lane of the wavefront.
// ------------------------------------------------------------------------------------------
template <int NP>
__global__ void __launch_bounds__(64) k_ke_3d(const double *__restrict__ dphi, const double *__restrict__ DS,
                                              const double *__restrict__ w, int64_t n_elem, int n_q, double r,
                                              const double *__restrict__ shear, const double *__restrict__ lame,
                                              double *__restrict__ ke) {
    constexpr int NLD = 3 * NP;
    constexpr int NH = NP / 2;
    static_assert(NP % 2 == 0 && NLD <= 32, "two column halves of an element with at most 32 rows");
    const int64_t e = blockIdx.x;
    const int lane = threadIdx.x;
    const int h = lane >> 5;
    const int a_raw = lane & 31;
    const bool valid = a_raw < NLD;
    const int a = valid ? a_raw : NLD - 1;  // clamped row: in-bounds reads, store masked
    const int ni = a / 3, ci = a - 3 * ni;
    const int64_t n_int = n_elem * n_q;
    const bool fold = shear != nullptr;
    const double om = fold ? 1.0 - r : 1.0;
    const double rr = fold ? r : 0.0;
    double acc[3 * NH];
#pragma unroll
    for (int k = 0; k < 3 * NH; ++k) acc[k] = 0.0;
    for (int q = 0; q < n_q; ++q) {
        const int64_t g = e * n_q + q;
        const double *gx = dphi + g * NP;
        const double *gy = dphi + (n_int + g) * NP;
        const double *gz = dphi + (2 * n_int + g) * NP;
        const double *D = DS + g * 36;
        const double wg = w[g];
        double b[6];
        brow(ci, gx[ni], gy[ni], gz[ni], b);
        // D b (D column-major: D[s + 6 m]); D_el b = lame tr(b) vol + 2 G ident b
        double t[6];
#pragma unroll
        for (int m = 0; m < 6; ++m) {
            double v = b[0] * D[6 * m];
#pragma unroll
            for (int s = 1; s < 6; ++s) v += b[s] * D[s + 6 * m];
            t[m] = v;
        }
        if (fold) {
            const double G = shear[g], L = lame[g];
            const double tr = L * (b[0] + b[1] + b[2]);
#pragma unroll
            for (int m = 0; m < 6; ++m) {
                const double del = (m < 3 ? tr + 2.0 * G * b[m] : G * b[m]);
                t[m] = om * t[m] + rr * del;
            }
        }
#pragma unroll
        for (int m = 0; m < 6; ++m) t[m] *= wg;
#pragma unroll
        for (int k = 0; k < NH; ++k) {
            const int j = NH * h + k;
            const double xj = gx[j], yj = gy[j], zj = gz[j];
            acc[3 * k + 0] += t[0] * xj + t[3] * yj + t[5] * zj;
            acc[3 * k + 1] += t[3] * xj + t[1] * yj + t[4] * zj;
            acc[3 * k + 2] += t[5] * xj + t[4] * yj + t[2] * zj;
        }
    }
    if (valid) {
        double *o = ke + ((int64_t)e * NLD + a) * NLD + 3 * NH * h;
#pragma unroll
        for (int k = 0; k < 3 * NH; ++k) o[k] = acc[k];
    }
}

// ------------------------------------------------------------------------------------------
// Phase 1, pipelined form for P1 (NP = 4, NQ = 1; P2 meshes with the 11-point rule take the
// node-pair blocks of block_assembly.hip instead): persistent
// wavefronts walk elements e, e + grid, ...; while one element is computed from LDS, the next
// element's point data (gradients, D, w and the material) is already in flight into registers
// and is written to the other LDS buffer afterwards, so the global-memory latency is paid once
// per element and overlapped with the previous element's arithmetic.
//
// LDS per buffer: gradients sg[q][c][h][6] (the two column halves of NH <= 5 nodes each padded to
// 6, so a lane reads its columns as aligned pairs), D transposed to rows sd[q][s][m] = D[s + 6 m]
// (a lane reads the three rows of D its B row touches), and w, shear, lame per point.
// ------------------------------------------------------------------------------------------
template <int NP, int NQ>
struct KePipe {
    static constexpr int NLD = 3 * NP;
    static constexpr int NH = NP / 2;
    static constexpr int HP = 6;                   // padded half width
    static constexpr int G = NQ * 3 * 2 * HP;      // gradient words
    static constexpr int D = NQ * 36;              // D words
    static constexpr int M = NQ * 3;               // w, shear, lame
    static constexpr int WORDS = G + D + M;
    static constexpr int NGR = (NQ * 3 * NP + 63) / 64;   // gradient values per lane in flight
    static constexpr int NDR = (D + 63) / 64;
    static constexpr int NMR = (M + 63) / 64;
};

template <int NP, int NQ, bool FOLD>
__device__ __forceinline__ void ke_pipe_body(const double *__restrict__ dphi, const double *__restrict__ DS,
                                             const double *__restrict__ w, int64_t n_elem, double r,
                                             const double *__restrict__ shear, const double *__restrict__ lame,
                                             double *__restrict__ ke, double (*lds)[KePipe<NP, NQ>::WORDS]) {
    using P = KePipe<NP, NQ>;
    constexpr int NLD = P::NLD, NH = P::NH, HP = P::HP;
    const int lane = threadIdx.x;
    const int h = lane >> 5;
    const int a_raw = lane & 31;
    const bool valid = a_raw < NLD;
    const int a = valid ? a_raw : NLD - 1;
    const int ni = a / 3, ci = a - 3 * ni;
    const int64_t n_int = n_elem * NQ;
    // rows of D touched by row a of B^T (component ci): s0 = ci, s1 = 3 + (ci == 2), s2 = 5 - (ci == 1)
    const int s0 = ci, s1 = 3 + (ci == 2 ? 1 : 0), s2 = 5 - (ci == 1 ? 1 : 0);
    const int hi = ni / NH, ki = ni - NH * (ni / NH);   // node ni's half and slot

    double gr[P::NGR], dr[P::NDR], mr[P::NMR][3];
    // global -> registers for element e (all loads in bounds; lanes past the data read index 0)
    auto fetch = [&](int64_t e) {
#pragma unroll
        for (int k = 0; k < P::NGR; ++k) {
            const int t = lane + 64 * k;                       // t = (c * NQ + q) * NP + node
            const int tc = t < NQ * 3 * NP ? t : 0;
            const int c = tc / (NQ * NP), rem = tc - c * (NQ * NP);
            gr[k] = dphi[((int64_t)c * n_int + e * NQ) * NP + rem];
        }
#pragma unroll
        for (int k = 0; k < P::NDR; ++k) {
            const int t = lane + 64 * k;
            dr[k] = DS[e * P::D + (t < P::D ? t : 0)];
        }
#pragma unroll
        for (int k = 0; k < P::NMR; ++k) {
            // t = kind * NQ + q: w, shear and lame are each read at an in-bounds index (no per-lane
            // choice between base pointers); the value is selected when staged, so nothing here
            // waits for the loads
            const int t = lane + 64 * k;
            const int tc = t < P::M ? t : 0;
            const int kind = tc / NQ, q = tc - kind * NQ;
            mr[k][0] = w[e * NQ + q];
            if constexpr (FOLD) {
                mr[k][1] = shear[e * NQ + q];
                mr[k][2] = lame[e * NQ + q];
            }
        }
    };
    // registers -> LDS buffer b
    auto stage = [&](int b) {
        double *L = lds[b];
#pragma unroll
        for (int k = 0; k < P::NGR; ++k) {
            const int t = lane + 64 * k;
            if (t < NQ * 3 * NP) {
                const int c = t / (NQ * NP), rem = t - c * (NQ * NP);
                const int q = rem / NP, node = rem - q * NP;
                const int hh = node / NH, kk = node - NH * (node / NH);
                L[((q * 3 + c) * 2 + hh) * HP + kk] = gr[k];
            }
        }
#pragma unroll
        for (int k = 0; k < P::NDR; ++k) {
            const int t = lane + 64 * k;                       // t = q*36 + s + 6 m  -> row-major (s, m)
            if (t < P::D) {
                const int q = t / 36, rem = t - 36 * q;
                const int m = rem / 6, s = rem - 6 * m;
                L[P::G + q * 36 + s * 6 + m] = dr[k];
            }
        }
#pragma unroll
        for (int k = 0; k < P::NMR; ++k) {
            const int t = lane + 64 * k;
            const int kind = (t < P::M ? t : 0) / NQ;
            double v = kind == 0 ? mr[k][0] : 0.0;
            if constexpr (FOLD) v += (kind == 1 ? mr[k][1] : 0.0) + (kind == 2 ? mr[k][2] : 0.0);
            if (t < P::M) L[P::G + P::D + t] = v;
        }
    };

    int64_t e = blockIdx.x;
    if (e >= n_elem) return;   // whole wavefront (one wavefront per workgroup)
    fetch(e);
    stage(0);
    __syncthreads();
    const double om = FOLD ? 1.0 - r : 1.0;
    int buf = 0;
    for (; e < n_elem; e += gridDim.x) {
        const int64_t en = e + gridDim.x;
        if (en < n_elem) fetch(en);
        const double *L = lds[buf];
        double acc[3 * NH];
#pragma unroll
        for (int k = 0; k < 3 * NH; ++k) acc[k] = 0.0;
#pragma unroll 1
        for (int q = 0; q < NQ; ++q) {
            const double *gq = L + q * 3 * 2 * HP;
            const double xi = gq[(0 * 2 + hi) * HP + ki], yi = gq[(1 * 2 + hi) * HP + ki], zi = gq[(2 * 2 + hi) * HP + ki];
            const double wg = L[P::G + P::D + q];
            // B row a has coefficients (c0, c1, c2) on strain rows (s0, s1, s2); they are scaled by
            // w (1 - r) once, so t = w (1 - r) D b_a costs 18 FMAs
            const double c0 = ci == 0 ? xi : (ci == 1 ? yi : zi);
            const double c1 = ci == 1 ? xi : yi;
            const double c2 = ci == 2 ? xi : zi;
            const double wo = wg * om;
            const double wc0 = wo * c0, wc1 = wo * c1, wc2 = wo * c2;
            double t[6];
            const double *D0 = L + P::G + q * 36 + s0 * 6;
            const double *D1 = L + P::G + q * 36 + s1 * 6;
            const double *D2 = L + P::G + q * 36 + s2 * 6;
#pragma unroll
            for (int m = 0; m < 6; ++m) t[m] = fma(wc2, D2[m], fma(wc1, D1[m], wc0 * D0[m]));
            if constexpr (FOLD) {
                // + w r D_el b_a with D_el b = lame (b0 + b1 + b2) vol + 2 G ident b; here
                // b0 + b1 + b2 = c0 and b is (c0, c1, c2) on rows (s0, s1, s2)
                const double Gq = L[P::G + P::D + NQ + q], Lq = L[P::G + P::D + 2 * NQ + q];
                const double rw = r * wg;
                const double tr = Lq * c0, g2 = 2.0 * Gq;
                const double b3 = ci != 2 ? c1 : 0.0;
                const double b4 = ci == 2 ? c1 : (ci == 1 ? c2 : 0.0);
                const double b5 = ci != 1 ? c2 : 0.0;
                t[0] = fma(rw, fma(g2, ci == 0 ? c0 : 0.0, tr), t[0]);
                t[1] = fma(rw, fma(g2, ci == 1 ? c0 : 0.0, tr), t[1]);
                t[2] = fma(rw, fma(g2, ci == 2 ? c0 : 0.0, tr), t[2]);
                const double rg = rw * Gq;
                t[3] = fma(rg, b3, t[3]);
                t[4] = fma(rg, b4, t[4]);
                t[5] = fma(rg, b5, t[5]);
            }
            const double *gx = gq + (0 * 2 + h) * HP, *gy = gq + (1 * 2 + h) * HP, *gz = gq + (2 * 2 + h) * HP;
#pragma unroll
            for (int k = 0; k < NH; ++k) {
                const double xj = gx[k], yj = gy[k], zj = gz[k];
                acc[3 * k + 0] = fma(t[5], zj, fma(t[3], yj, fma(t[0], xj, acc[3 * k + 0])));
                acc[3 * k + 1] = fma(t[4], zj, fma(t[1], yj, fma(t[3], xj, acc[3 * k + 1])));
                acc[3 * k + 2] = fma(t[2], zj, fma(t[4], yj, fma(t[5], xj, acc[3 * k + 2])));
            }
        }
        if (valid) {
            double *o = ke + (e * NLD + a) * NLD + 3 * NH * h;
#pragma unroll
            for (int k = 0; k < 3 * NH; ++k) o[k] = acc[k];
        }
        if (en < n_elem) stage(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
}

// Double-buffered LDS (the next element is staged while nothing waits for it).
template <int NP, int NQ, bool FOLD>
__global__ void __launch_bounds__(64) k_ke_3d_pipe(const double *__restrict__ dphi, const double *__restrict__ DS,
                                                   const double *__restrict__ w, int64_t n_elem, double r,
                                                   const double *__restrict__ shear, const double *__restrict__ lame,
                                                   double *__restrict__ ke) {
    __shared__ double lds[2][KePipe<NP, NQ>::WORDS];
    ke_pipe_body<NP, NQ, FOLD>(dphi, DS, w, n_elem, r, shear, lame, ke, lds);
}

// ------------------------------------------------------------------------------------------
// Phase 2: node gather.  One wavefront per node; LDS holds the node's DIM rows, each of length
// L + 1 (free rows share one length L; word L of a row absorbs its constrained columns, whose
// offset is L), plus a spare word that absorbs the updates of constrained rows.  Each
// element contributes its DIM contiguous rows DIM*li .. DIM*li + DIM-1 (DIM*NLD values, read
// coalesced); lane p handles value p (and p + 64, ...), i.e. row c = p / NLD, column b = p % NLD.
// Within one element the targets are distinct (distinct columns); elements are applied one after
// the other with a barrier in between, and the next element's values are loaded before the
// barrier (software pipelining).
// ------------------------------------------------------------------------------------------
template <int DIM, int NLD>
__global__ void __launch_bounds__(64) k_gather_nodes(const double *__restrict__ ke, const int32_t *__restrict__ ns_ptr,
                                                     const int32_t *__restrict__ ns_elem,
                                                     const uint8_t *__restrict__ ns_lnode,
                                                     const int16_t *__restrict__ ns_off,
                                                     const int32_t *__restrict__ indptr,
                                                     const uint8_t *__restrict__ row_free,
                                                     const int32_t *__restrict__ nodes, double r,
                                                     const double *__restrict__ elastic, double *__restrict__ out,
                                                     float *__restrict__ out32) {
    constexpr int NV = DIM * NLD;
    constexpr int NPASS = (NV + 63) / 64;
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
    const int L1 = L + 1;
    const int spare = DIM * L1;
    for (int i = lane; i <= spare; i += 64) acc[i] = 0.0;

    // per-pass constant part of the LDS target: c*(L+1) for a free row, -1 (the spare word) otherwise
    int rowbase[NPASS], col[NPASS];
    bool live[NPASS];
#pragma unroll
    for (int k = 0; k < NPASS; ++k) {
        const int p = lane + 64 * k;
        live[k] = p < NV;
        const int pc = live[k] ? p : 0;
        const int c = pc / NLD;
        col[k] = pc - c * NLD;
        int frc = 0;
#pragma unroll
        for (int cc = 0; cc < DIM; ++cc) frc += (c == cc) ? fr[cc] : 0;
        rowbase[k] = (live[k] && frc) ? c * L1 : -1;
    }

    const int32_t s0 = ns_ptr[node], s1 = ns_ptr[node + 1];
    double v[NPASS];
    int o[NPASS];
    auto load = [&](int32_t s, double *vv, int *oo) {
        const int64_t e = ns_elem[s];
        const int li = ns_lnode[s];
        const double *kb = ke + (e * NLD + (int64_t)DIM * li) * NLD;
        const int16_t *off = ns_off + (int64_t)s * NLD;
#pragma unroll
        for (int k = 0; k < NPASS; ++k) {
            const int p = live[k] ? lane + 64 * k : 0;
            vv[k] = kb[p];
            oo[k] = off[col[k]];
        }
    };
    if (s0 < s1) load(s0, v, o);
    __syncthreads();
    for (int32_t s = s0; s < s1; ++s) {
        double vn[NPASS];
        int on[NPASS];
        load(s + 1 < s1 ? s + 1 : s, vn, on);
#pragma unroll
        for (int k = 0; k < NPASS; ++k) {
            const int idx = rowbase[k] >= 0 ? rowbase[k] + o[k] : spare;
            acc[idx] += v[k];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < NPASS; ++k) {
            v[k] = vn[k];
            o[k] = on[k];
        }
    }
    const double om = 1.0 - r;
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
        if (fr[c]) {
            for (int i = lane; i < L; i += 64) {
                double val = acc[c * L1 + i];
                if (elastic) val = om * val + r * elastic[st[c] + i];
                out[st[c] + i] = val;
                if (out32) out32[st[c] + i] = (float)val;
            }
        } else if (lane == 0) {
            out[st[c]] = 1.0;  // constrained row: its diagonal only
            if (out32) out32[st[c]] = 1.0f;
        }
    }
}

template <int DIM, int NLD>
int launch_gather(const double *ke, const int32_t *ns_ptr, const int32_t *ns_elem, const uint8_t *ns_lnode,
                  const int16_t *ns_off, const int32_t *indptr, const uint8_t *row_free, const int32_t *nodes,
                  int64_t n_nodes, int max_len, double r, const double *elastic, double *out, float *out32,
                  hipStream_t st) {
    const size_t lds = ((size_t)DIM * (max_len + 1) + 1) * sizeof(double);
    if (lds > 65536) {
        static int max_lds = -1;
        if (max_lds < 0) {
            int dev = 0, v = 65536;
            if (hipGetDevice(&dev) != hipSuccess ||
                hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
                v = 65536;
            max_lds = v;
        }
        SSX_REQUIRE(lds <= (size_t)max_lds, "ssx_gather_node_rows: rows of %d entries exceed the LDS", max_len);
        if (hipFuncSetAttribute(reinterpret_cast<const void *>(&k_gather_nodes<DIM, NLD>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess) {
            ssx::set_error("ssx_gather_node_rows: cannot raise the LDS limit to %zu bytes", lds);
            return SSX_ERUNTIME;
        }
    }
    hipLaunchKernelGGL((k_gather_nodes<DIM, NLD>), dim3((unsigned)n_nodes), dim3(64), lds, st, ke, ns_ptr, ns_elem,
                       ns_lnode, ns_off, indptr, row_free, nodes, r, elastic, out, out32);
    SSX_CHECK_LAUNCH("ssx_gather_node_rows");
    return SSX_OK;
}

}  // namespace

// Persistent grid of the pipelined element-matrix kernel: 24 wavefronts per CU (measured best
// of 8-32 on MI355X; the kernel holds 8 resident per CU).
constexpr int kKeWavesPerCU = 24;

extern "C" {

int ssx_element_matrices_3d(int n_p, const double *dphi, const double *DS, const double *w, int64_t n_elem, int n_q,
                            double r, const double *shear, const double *lame, double *ke, void *stream) {
    SSX_REQUIRE(n_elem >= 0 && n_q > 0, "ssx_element_matrices_3d: bad sizes");
    if (n_elem == 0) return SSX_OK;
    SSX_REQUIRE(dphi && DS && w && ke, "ssx_element_matrices_3d: null pointer");
    SSX_REQUIRE((shear == nullptr) == (lame == nullptr), "ssx_element_matrices_3d: shear and lame go together");
    SSX_REQUIRE(n_elem <= 0x7fffffff, "ssx_element_matrices_3d: too many elements for one launch");
    hipStream_t st = ssx::as_stream(stream);
    if (n_p == 4 && n_q == 1) {
        static int n_cu = 0;
        if (n_cu == 0) {
            int dev = 0;
            if (hipGetDevice(&dev) != hipSuccess ||
                hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n_cu <= 0)
                n_cu = 256;
        }
        const int64_t want = (int64_t)n_cu * kKeWavesPerCU;
        const unsigned grid = (unsigned)(n_elem < want ? n_elem : want);
        const bool fold = shear != nullptr;
#define PIPE(NP_, NQ_, F_) \
    hipLaunchKernelGGL((k_ke_3d_pipe<NP_, NQ_, F_>), dim3(grid), dim3(64), 0, st, dphi, DS, w, n_elem, r, shear, lame, ke)
        if (fold) PIPE(4, 1, true); else PIPE(4, 1, false);
#undef PIPE
        SSX_CHECK_LAUNCH("ssx_element_matrices_3d");
        return SSX_OK;
    }
    if (n_p == 10)
        hipLaunchKernelGGL((k_ke_3d<10>), dim3((unsigned)n_elem), dim3(64), 0, st, dphi, DS, w, n_elem, n_q, r, shear,
                           lame, ke);
    else if (n_p == 4)
        hipLaunchKernelGGL((k_ke_3d<4>), dim3((unsigned)n_elem), dim3(64), 0, st, dphi, DS, w, n_elem, n_q, r, shear,
                           lame, ke);
    else {
        ssx::set_error("ssx_element_matrices_3d: n_p must be 4 or 10");
        return SSX_EINVAL;
    }
    SSX_CHECK_LAUNCH("ssx_element_matrices_3d");
    return SSX_OK;
}

int ssx_gather_node_rows(int dim, int nld, const double *ke, const int32_t *ns_ptr, const int32_t *ns_elem,
                         const uint8_t *ns_lnode, const int16_t *ns_off, const int32_t *indptr,
                         const uint8_t *row_free, const int32_t *nodes, int64_t n_nodes, int max_len, double r,
                         const double *elastic, double *out, float *out32, void *stream) {
    SSX_REQUIRE(n_nodes >= 0 && max_len >= 0 && max_len < 32768, "ssx_gather_node_rows: bad sizes");
    if (n_nodes == 0) return SSX_OK;
    SSX_REQUIRE(ke && ns_ptr && ns_elem && ns_lnode && ns_off && indptr && row_free && out,
                "ssx_gather_node_rows: null pointer");
    SSX_REQUIRE(n_nodes <= 0x7fffffff, "ssx_gather_node_rows: too many nodes for one launch");
    hipStream_t st = ssx::as_stream(stream);
#define CALL(D, N) \
    launch_gather<D, N>(ke, ns_ptr, ns_elem, ns_lnode, ns_off, indptr, row_free, nodes, n_nodes, max_len, r, elastic, \
                        out, out32, st)
    if (dim == 3 && nld == 12) return CALL(3, 12);
    if (dim == 3 && nld == 30) return CALL(3, 30);
    if (dim == 3 && nld == 105) return CALL(3, 105);
#undef CALL
    ssx::set_error("ssx_gather_node_rows: unsupported element (dim=%d, nld=%d)", dim, nld);
    return SSX_EINVAL;
}

}  // extern "C"
