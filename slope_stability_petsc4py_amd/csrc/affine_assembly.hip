// Strain gather and internal force of P2 tetrahedra with a constant Jacobian per element, on gfx950.
//
// A tetrahedron with straight edges and its edge nodes at the midpoints is an affine image of the
// reference element, so J is the same at all of its points and the physical gradients are
//   dphi_d(e, q, n) = sum_k G_e[3 d + k] h_k(q, n),   w(e, q) = G_e[9] wf[q],
// with G_e = (J^-1 row-major, |det J|) (10 doubles per element) and h_k(q, n) the rule's reference
// gradients (330 doubles for P2 at 11 points, the same for every element).  The kernels form the
// gradients in registers from G_e and an LDS copy of h instead of reading the per-point dphi (240
// bytes per point) and w (8 bytes) from HBM: the strain gather then reads u, the DOF table and 10
// doubles per element, the force kernel S and 10 doubles per element.
//
// The arithmetic is that of the per-point kernels (assembly.hip) on gradients that agree with the
// per-point dphi to rounding (the same cofactor formula, taken at the rule's first point).
//
// Reference behaviour: cython/assemble_tangent_values_3d.c:216-273 (strain), :275-327 (force);
// fem/assembly.py:139-245 (gradients and weights).
#include "ssx_common.h"

namespace {

constexpr int kNP = 10;                   // P2 tetrahedron
constexpr int kNQ = 11;                   // 11-point rule
constexpr int kNS = 6;                    // [e11 e22 e33 g12 g23 g13]
constexpr int kTab = 3 * kNQ * kNP;       // h_k(q, n) at (k kNQ + q) kNP + n
constexpr int kGW = 10;                   // words of G_e
constexpr int kBlock = 256;

// Strain: one lane per point g = e kNQ + q; the block's strains leave through LDS in one coalesced
// pass over E[g0 kNS .. (g0 + np) kNS).  The factors of the block's elements (at most kEpbS) and
// their 30 displacement values are staged in LDS first, each gathered once per element (read
// directly, every lane issued 69 loads, 30 of them dependent on a dof index, of values its
// element's other 10 lanes read too).
constexpr int kEpbS = kBlock / kNQ + 2;   // elements a block's points can touch

__global__ void __launch_bounds__(kBlock) k_strain_affine(const double *__restrict__ gm, const double *__restrict__ ref,
                                                          const int32_t *__restrict__ edofs,
                                                          const double *__restrict__ u, int64_t n_int,
                                                          double *__restrict__ E) {
    __shared__ double sh[kTab];
    __shared__ double so[kBlock * (kNS + 1)];
    __shared__ double sgm[kEpbS * kGW];
    __shared__ double su[kEpbS * 3 * kNP];
    const int t = threadIdx.x;
    for (int i = t; i < kTab; i += kBlock) sh[i] = ref[i];
    const int64_t g0 = (int64_t)blockIdx.x * kBlock;
    const int np = (int)(n_int - g0 < kBlock ? n_int - g0 : kBlock);
    const int64_t e0 = g0 / kNQ;
    const int ne = (int)((g0 + np - 1) / kNQ + 1 - e0);
    for (int i = t; i < ne * kGW; i += kBlock) sgm[i] = gm[e0 * kGW + i];
    for (int i = t; i < ne * 3 * kNP; i += kBlock) su[i] = u[edofs[e0 * (3 * kNP) + i]];
    __syncthreads();
    if (t < np) {
        const int64_t g = g0 + t;
        const int64_t e = g / kNQ;
        const int q = (int)(g - e * kNQ);
        const int le = (int)(e - e0);
        const double *G = sgm + le * kGW;
        double m[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) m[k] = G[k];
        const double *ue = su + le * (3 * kNP);
        const double *h0 = sh + q * kNP, *h1 = sh + (kNQ + q) * kNP, *h2 = sh + (2 * kNQ + q) * kNP;
        double a[kNS];
#pragma unroll
        for (int k = 0; k < kNS; ++k) a[k] = 0.0;
#pragma unroll
        for (int j = 0; j < kNP; ++j) {
            const double ux = ue[3 * j], uy = ue[3 * j + 1], uz = ue[3 * j + 2];
            const double r0 = h0[j], r1 = h1[j], r2 = h2[j];
            const double x = fma(m[2], r2, fma(m[1], r1, m[0] * r0));
            const double y = fma(m[5], r2, fma(m[4], r1, m[3] * r0));
            const double z = fma(m[8], r2, fma(m[7], r1, m[6] * r0));
            a[0] += x * ux;
            a[1] += y * uy;
            a[2] += z * uz;
            a[3] += y * ux + x * uy;
            a[4] += z * uy + y * uz;
            a[5] += z * ux + x * uz;
        }
        double *o = so + t * (kNS + 1);
#pragma unroll
        for (int k = 0; k < kNS; ++k) o[k] = a[k];
    }
    __syncthreads();
    for (int i = t; i < np * kNS; i += kBlock) E[g0 * kNS + i] = so[(i / kNS) * (kNS + 1) + i % kNS];
}

// Force, stage 1: element-local vector f_e[a] = sum_q w_q B_a(q)^T S_q, one lane per (element, node)
// (the same sums, in the same order, as the per-point kernel); stage 2 is the node gather.  The
// stresses and factors of the block's elements (at most kEpb) are staged in LDS by coalesced loads
// first: read directly, every lane issued 76 loads of values its element's other lanes read too.
constexpr int kEpb = kBlock / kNP + 2;   // elements a block's lanes can touch

__global__ void __launch_bounds__(kBlock) k_force_affine(const double *__restrict__ gm, const double *__restrict__ ref,
                                                         const double *__restrict__ S, int64_t n_elem,
                                                         double *__restrict__ fe) {
    __shared__ double sh[kTab + kNQ];
    __shared__ double ss[kEpb * kNQ * kNS];
    __shared__ double sgm[kEpb * kGW];
    for (int i = threadIdx.x; i < kTab + kNQ; i += kBlock) sh[i] = ref[i];
    const int64_t t0 = (int64_t)blockIdx.x * kBlock;
    const int64_t t_end = t0 + kBlock < n_elem * kNP ? t0 + kBlock : n_elem * kNP;
    const int64_t e0 = t0 / kNP;
    const int ne = (int)((t_end - 1) / kNP + 1 - e0);
    for (int i = threadIdx.x; i < ne * kNQ * kNS; i += kBlock) ss[i] = S[e0 * (kNQ * kNS) + i];
    for (int i = threadIdx.x; i < ne * kGW; i += kBlock) sgm[i] = gm[e0 * kGW + i];
    __syncthreads();
    const int64_t t = t0 + threadIdx.x;
    if (t >= n_elem * kNP) return;
    const int64_t e = t / kNP;
    const int node = (int)(t - e * kNP);
    const int le = (int)(e - e0);
    const double *G = sgm + le * kGW;
    double m[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) m[k] = G[k];
    const double det = G[9];
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;
#pragma unroll
    for (int q = 0; q < kNQ; ++q) {
        const double r0 = sh[q * kNP + node], r1 = sh[(kNQ + q) * kNP + node], r2 = sh[(2 * kNQ + q) * kNP + node];
        const double gx = fma(m[2], r2, fma(m[1], r1, m[0] * r0));
        const double gy = fma(m[5], r2, fma(m[4], r1, m[3] * r0));
        const double gz = fma(m[8], r2, fma(m[7], r1, m[6] * r0));
        const double wg = det * sh[kTab + q];
        const double *s = ss + (le * kNQ + q) * kNS;
        const double s0 = s[0], s1 = s[1], s2 = s[2], s3 = s[3], s4 = s[4], s5 = s[5];
        f0 += wg * (gx * s0 + gy * s3 + gz * s5);
        f1 += wg * (gx * s3 + gy * s1 + gz * s4);
        f2 += wg * (gx * s5 + gy * s4 + gz * s2);
    }
    double *o = fe + e * (3 * kNP) + 3 * node;
    o[0] = f0;
    o[1] = f1;
    o[2] = f2;
}

}  // namespace

extern "C" {

int ssx_strain_affine_3d(int n_p, const double *gm, const double *ref, const int32_t *edofs, const double *u,
                         int64_t n_elem, int n_q, double *E, void *stream) {
    SSX_REQUIRE(n_p == kNP && n_q == kNQ, "ssx_strain_affine_3d: only P2 tetrahedra with the 11-point rule "
                "(n_p 10, n_q 11), got n_p %d n_q %d", n_p, n_q);
    SSX_REQUIRE(n_elem >= 0 && n_elem <= 0x7fffffff, "ssx_strain_affine_3d: bad element count");
    if (n_elem == 0) return SSX_OK;
    SSX_REQUIRE(gm && ref && edofs && u && E, "ssx_strain_affine_3d: null pointer");
    const int64_t n_int = n_elem * kNQ;
    hipLaunchKernelGGL(k_strain_affine, dim3(ssx::grid_for(n_int, kBlock)), dim3(kBlock), 0, ssx::as_stream(stream), gm,
                       ref, edofs, u, n_int, E);
    SSX_CHECK_LAUNCH("ssx_strain_affine_3d");
    return SSX_OK;
}

int ssx_force_elem_affine_3d(int n_p, const double *gm, const double *ref, const double *S, int64_t n_elem, int n_q,
                             double *fe, void *stream) {
    SSX_REQUIRE(n_p == kNP && n_q == kNQ, "ssx_force_elem_affine_3d: only P2 tetrahedra with the 11-point rule "
                "(n_p 10, n_q 11), got n_p %d n_q %d", n_p, n_q);
    SSX_REQUIRE(n_elem >= 0 && n_elem <= 0x7fffffff, "ssx_force_elem_affine_3d: bad element count");
    if (n_elem == 0) return SSX_OK;
    SSX_REQUIRE(gm && ref && S && fe, "ssx_force_elem_affine_3d: null pointer");
    hipLaunchKernelGGL(k_force_affine, dim3(ssx::grid_for(n_elem * kNP, kBlock)), dim3(kBlock), 0,
                       ssx::as_stream(stream), gm, ref, S, n_elem, fe);
    SSX_CHECK_LAUNCH("ssx_force_elem_affine_3d");
    return SSX_OK;
}

}  // extern "C"
