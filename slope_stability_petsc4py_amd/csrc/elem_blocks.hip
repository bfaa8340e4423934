// Element tangent matrices of P1/P2 tetrahedra by node blocks on the vector ALUs: one lane per
// node pair (i <= j) of an element computes the 3x3 block K_e[(i,c),(j,d)] = sum_q w_q
// b_ic^T D_q b_jd, with b_ic the (3-nonzero) strain-displacement column of dof (i,c); the mirrored
// block is written for i < j (DS is symmetric for associated flow).  A workgroup is one element,
// so every lane reads the same D_q / w_q (uniform loads) and its own two gradient triples.
// The matrices then go through ssx_scatter_element_rows (deterministic row gather, slot order).
//
// Reference behaviour: the element stage of cython/assemble_tangent_values_3d.c:45-94.
#include "ssx_common.h"

namespace {

template <int NP>
__global__ void __launch_bounds__(64) k_ke_blocks(const double *__restrict__ dphi, const double *__restrict__ DS,
                                                  const double *__restrict__ w, int64_t n_elem, int n_q,
                                                  double *__restrict__ ke) {
    constexpr int NLD = 3 * NP;
    constexpr int NB = NP * (NP + 1) / 2;
    const int64_t e = blockIdx.x;
    const int t = threadIdx.x;
    if (t >= NB || e >= n_elem) return;
    // (i, j) with i <= j from the linear index t
    int i = 0, rem = t;
    while (rem >= NP - i) {
        rem -= NP - i;
        ++i;
    }
    const int j = i + rem;
    const int64_t n_int = n_elem * n_q;
    double acc[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) acc[k] = 0.0;
    for (int q = 0; q < n_q; ++q) {
        const int64_t g = e * n_q + q;
        const double xi = dphi[g * NP + i], yi = dphi[(n_int + g) * NP + i], zi = dphi[(2 * n_int + g) * NP + i];
        const double xj = dphi[g * NP + j], yj = dphi[(n_int + g) * NP + j], zj = dphi[(2 * n_int + g) * NP + j];
        const double *D = DS + g * 36;
        const double wg = w[g];
        // t_c[m] = w * sum_s b_ic[s] D[s + 6 m]
        double tc[3][6];
#pragma unroll
        for (int m = 0; m < 6; ++m) {
            const double *Dm = D + 6 * m;
            tc[0][m] = wg * (xi * Dm[0] + yi * Dm[3] + zi * Dm[5]);
            tc[1][m] = wg * (yi * Dm[1] + xi * Dm[3] + zi * Dm[4]);
            tc[2][m] = wg * (zi * Dm[2] + yi * Dm[4] + xi * Dm[5]);
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            acc[3 * c + 0] += tc[c][0] * xj + tc[c][3] * yj + tc[c][5] * zj;
            acc[3 * c + 1] += tc[c][3] * xj + tc[c][1] * yj + tc[c][4] * zj;
            acc[3 * c + 2] += tc[c][5] * xj + tc[c][4] * yj + tc[c][2] * zj;
        }
    }
    double *K = ke + e * (int64_t)(NLD * NLD);
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            K[(3 * i + c) * NLD + 3 * j + d] = acc[3 * c + d];
            if (i != j) K[(3 * j + d) * NLD + 3 * i + c] = acc[3 * c + d];
        }
}

}  // namespace

extern "C" int ssx_element_matrices_blocks(int n_p, const double *dphi, const double *DS, const double *w,
                                           int64_t n_elem, int n_q, double *ke, void *stream) {
    SSX_REQUIRE(n_elem >= 0 && n_q > 0, "ssx_element_matrices_blocks: bad sizes");
    if (n_elem == 0) return SSX_OK;
    SSX_REQUIRE(dphi && DS && w && ke, "ssx_element_matrices_blocks: null pointer");
    hipStream_t st = ssx::as_stream(stream);
    if (n_p == 10)
        hipLaunchKernelGGL((k_ke_blocks<10>), dim3((unsigned)n_elem), dim3(64), 0, st, dphi, DS, w, n_elem, n_q, ke);
    else if (n_p == 4)
        hipLaunchKernelGGL((k_ke_blocks<4>), dim3((unsigned)n_elem), dim3(64), 0, st, dphi, DS, w, n_elem, n_q, ke);
    else {
        ssx::set_error("ssx_element_matrices_blocks: n_p must be 4 or 10 (3D)");
        return SSX_EINVAL;
    }
    SSX_CHECK_LAUNCH("ssx_element_matrices_blocks");
    return SSX_OK;
}
