// Mohr-Coulomb constitutive operator on gfx950: strength reduction (Davis A/B/C) and the
// return mapping with its consistent tangent, one lane per quadrature point.
//
// Reference behaviour: constitutive/reduction.py:14-78 (Davis), cython/constitutive_3D_kernel.h:68-351
// (3D return mapping), constitutive/problem.py:143-475 (2D plane strain).
//
// HBM layout (point-major, the reference C kernels' layout): E[p*6 + c], S[p*6 + c], DS[p*36 + k]
// with k = i + 6*j holding dS_i/dE_j (column-major 6x6, the reference's DS convention).  A point's
// record is contiguous (48 B / 288 B), so the assembly kernels that read all of a point's tangent
// touch 3 cache lines per point, and a wavefront's stores fill whole lines in L2 before write-back.
// Material data comes either as per-point arrays or as a per-element material id plus a small
// per-material table, which keeps the per-point read at 48 B of strain instead of 80 B.
#include "ssx_common.h"
#include <math.h>

namespace {

constexpr double kTwoPiThird = 2.0943951023931956;  // 2*pi/3

__device__ __forceinline__ double vol(int i) { return i < 3 ? 1.0 : 0.0; }
__device__ __forceinline__ double idn(int i, int j) { return i != j ? 0.0 : (i < 3 ? 1.0 : 0.5); }

// d(eps^2)/dE entry (i = Voigt row of eps^2, j = engineering strain column).
__device__ __forceinline__ double dsq(int i, int j, const double et[6]) {
    // Closed form of the derivative of the squared strain tensor with respect to the engineering
    // strain vector (tensor shear = half the engineering shear).
    switch (i + 6 * j) {
        case 0: return 2.0 * et[0];
        case 3: return et[3];
        case 5: return et[5];
        case 7: return 2.0 * et[1];
        case 9: return et[3];
        case 10: return et[4];
        case 14: return 2.0 * et[2];
        case 16: return et[4];
        case 17: return et[5];
        case 18: return et[3];
        case 19: return et[3];
        case 21: return 0.5 * (et[0] + et[1]);
        case 22: return 0.5 * et[5];
        case 23: return 0.5 * et[4];
        case 25: return et[4];
        case 26: return et[4];
        case 27: return 0.5 * et[5];
        case 28: return 0.5 * (et[1] + et[2]);
        case 29: return 0.5 * et[3];
        case 30: return et[5];
        case 32: return et[5];
        case 33: return 0.5 * et[4];
        case 34: return 0.5 * et[3];
        case 35: return 0.5 * (et[0] + et[2]);
        default: return 0.0;
    }
}

__device__ __forceinline__ void eig_proj(const double esq[6], const double et[6], double apb, double atb,
                                         double inv, double P[6]) {
#pragma unroll
    for (int i = 0; i < 6; ++i) P[i] = inv * (esq[i] - apb * et[i] + (i < 3 ? atb : 0.0));
}

// Packed symmetric layout of a 6x6 tangent: the 21 entries DS[s + 6 m], s <= m, row by row
// (s = 0: m = 0..5, s = 1: m = 1..5, ...); row s starts at packed index 6 s - s (s - 1) / 2.

// Tangent outputs of a return-mapping launch: point-major DS (n x 36, or n x 21 packed for the
// packed kernel) and its optional FP32 copy in the same layout; both may be null (stresses only).
struct DSWriter {
    double *ds;
    float *ds32;
    __device__ __forceinline__ bool on() const { return ds != nullptr || ds32 != nullptr; }
};

// Return mapping at one point. Returns the branch id: 0 elastic, 1 smooth, 2 left edge,
// 3 right edge, 4 apex.  Follows cython/constitutive_3D_kernel.h:68-351 formula by formula
// (same thresholds: Q < 1e-30 -> theta = 0, clamped acos argument, exact-tie edge choice).
template <class Writer>
__device__ int mc3d_point(const double E[6], double cb, double sp, double G, double lam, double S[6],
                          const Writer &w) {
    const double sq = sp;  // associated flow (sin psi = sin phi)
    const double et[6] = {E[0], E[1], E[2], 0.5 * E[3], 0.5 * E[4], 0.5 * E[5]};
    double esq[6];
    esq[0] = et[0] * et[0] + et[3] * et[3] + et[5] * et[5];
    esq[1] = et[1] * et[1] + et[3] * et[3] + et[4] * et[4];
    esq[2] = et[2] * et[2] + et[4] * et[4] + et[5] * et[5];
    esq[3] = et[0] * et[3] + et[1] * et[3] + et[4] * et[5];
    esq[4] = et[3] * et[5] + et[1] * et[4] + et[2] * et[4];
    esq[5] = et[0] * et[5] + et[3] * et[4] + et[2] * et[5];

    const double I1 = et[0] + et[1] + et[2];
    const double I2 = et[0] * et[1] + et[0] * et[2] + et[1] * et[2] - et[3] * et[3] - et[4] * et[4] - et[5] * et[5];
    const double I3 = et[0] * et[1] * et[2] - et[2] * et[3] * et[3] - et[1] * et[5] * et[5] - et[0] * et[4] * et[4] +
                      2.0 * et[3] * et[4] * et[5];
    double Q = (I1 * I1 - 3.0 * I2) / 9.0;
    Q = Q < 0.0 ? 0.0 : Q;
    const double R = (-2.0 * I1 * I1 * I1 + 9.0 * I1 * I2 - 27.0 * I3) / 54.0;
    const double rQ = sqrt(Q);
    double th = 0.0;
    if (!(Q < 1e-30)) {
        double c = R / (rQ * Q);
        c = fmin(fmax(c, -1.0), 1.0);
        th = acos(c) / 3.0;
    }
    const double third = I1 / 3.0;
    const double e1 = -2.0 * rQ * cos(th + kTwoPiThird) + third;
    const double e2 = -2.0 * rQ * cos(th - kTwoPiThird) + third;
    const double e3 = -2.0 * rQ * cos(th) + third;

    const double f = 2.0 * G * ((1.0 + sp) * e1 - (1.0 - sp) * e3) + 2.0 * lam * sp * I1 - cb;

    if (f <= 0.0) {
#pragma unroll
        for (int i = 0; i < 6; ++i) S[i] = (i < 3 ? lam * I1 : 0.0) + 2.0 * G * et[i];
        if (w.on()) {
#pragma unroll
            for (int j = 0; j < 6; ++j)
#pragma unroll
                for (int i = 0; i < 6; ++i) w.put(i + 6 * j, lam * vol(i) * vol(j) + 2.0 * G * idn(i, j));
        }
        return 0;
    }

    const double g_sl = (e1 - e2) / (1.0 + sq);
    const double g_sr = (e2 - e3) / (1.0 - sq);
    const double den_s = 4.0 * lam * sp * sq + 4.0 * G * (1.0 + sp * sq);
    const double l_s = f / den_s;

    if (l_s <= g_sl && l_s <= g_sr) {
        const double d1 = 1.0 / ((e1 - e2) * (e1 - e3));
        const double d2 = 1.0 / ((e2 - e1) * (e2 - e3));
        const double d3 = 1.0 / ((e3 - e1) * (e3 - e2));
        double P1[6], P2[6], P3[6];
        eig_proj(esq, et, e2 + e3, e2 * e3, d1, P1);
        eig_proj(esq, et, e1 + e3, e1 * e3, d2, P2);
        eig_proj(esq, et, e1 + e2, e1 * e2, d3, P3);
        const double s1 = lam * I1 + 2.0 * G * e1 - l_s * (2.0 * lam * sq + 2.0 * G * (1.0 + sq));
        const double s2 = lam * I1 + 2.0 * G * e2 - l_s * 2.0 * lam * sq;
        const double s3 = lam * I1 + 2.0 * G * e3 - l_s * (2.0 * lam * sq - 2.0 * G * (1.0 - sq));
#pragma unroll
        for (int i = 0; i < 6; ++i) S[i] = s1 * P1[i] + s2 * P2[i] + s3 * P3[i];
        if (w.on()) {
            double Dp[6], Dq[6];
#pragma unroll
            for (int i = 0; i < 6; ++i) {
                Dp[i] = 2.0 * G * ((1.0 + sp) * P1[i] - (1.0 - sp) * P3[i]) + 2.0 * lam * sp * vol(i);
                Dq[i] = 2.0 * G * ((1.0 + sq) * P1[i] - (1.0 - sq) * P3[i]) + 2.0 * lam * sq * vol(i);
            }
            const double inv = 1.0 / den_s;
            const double a1 = 2.0 * e1 - e2 - e3, b1 = e2 - e3;
            const double a2 = 2.0 * e2 - e1 - e3, b2 = e1 - e3;
            const double a3 = 2.0 * e3 - e1 - e2, b3 = e1 - e2;
#pragma unroll
            for (int j = 0; j < 6; ++j) {
#pragma unroll
                for (int i = 0; i < 6; ++i) {
                    const double dd = dsq(i, j, et), id = idn(i, j);
                    const double o1 = P1[i] * P1[j], o2 = P2[i] * P2[j], o3 = P3[i] * P3[j];
                    const double dP1 = d1 * (dd - id * (e2 + e3) - a1 * o1 - b1 * (o2 - o3));
                    const double dP2 = d2 * (dd - id * (e1 + e3) - a2 * o2 - b2 * (o1 - o3));
                    const double dP3 = d3 * (dd - id * (e1 + e2) - a3 * o3 - b3 * (o1 - o2));
                    w.put(i + 6 * j, s1 * dP1 + s2 * dP2 + s3 * dP3 + lam * vol(i) * vol(j) +
                                         2.0 * G * (o1 + o2 + o3) - Dq[i] * Dp[j] * inv);
                }
            }
        }
        return 1;
    }

    if (g_sl < g_sr) {
        const double den_l = 4.0 * lam * sp * sq + G * (1.0 + sp) * (1.0 + sq) + 2.0 * G * (1.0 - sp) * (1.0 - sq);
        const double l_l = (G * ((1.0 + sp) * (e1 + e2) - 2.0 * (1.0 - sp) * e3) + 2.0 * lam * sp * I1 - cb) / den_l;
        const double g_la = (e1 + e2 - 2.0 * e3) / (3.0 - sq);
        if (l_l >= g_sl && l_l <= g_la) {
            const double d3 = 1.0 / ((e3 - e1) * (e3 - e2));
            double P3[6], P12[6];
            eig_proj(esq, et, e1 + e2, e1 * e2, d3, P3);
#pragma unroll
            for (int i = 0; i < 6; ++i) P12[i] = vol(i) - P3[i];
            const double s1 = lam * I1 + G * (e1 + e2) - l_l * (2.0 * lam * sq + G * (1.0 + sq));
            const double s3 = lam * I1 + 2.0 * G * e3 - l_l * (2.0 * lam * sq - 2.0 * G * (1.0 - sq));
#pragma unroll
            for (int i = 0; i < 6; ++i) S[i] = s1 * P12[i] + s3 * P3[i];
            if (w.on()) {
                double Dp[6], Dq[6];
#pragma unroll
                for (int i = 0; i < 6; ++i) {
                    Dp[i] = G * ((1.0 + sp) * P12[i] - 2.0 * (1.0 - sp) * P3[i]) + 2.0 * lam * sp * vol(i);
                    Dq[i] = G * ((1.0 + sq) * P12[i] - 2.0 * (1.0 - sq) * P3[i]) + 2.0 * lam * sq * vol(i);
                }
                const double inv = 1.0 / den_l, m12 = e1 + e2;
#pragma unroll
                for (int j = 0; j < 6; ++j) {
#pragma unroll
                    for (int i = 0; i < 6; ++i) {
                        const double dd = dsq(i, j, et), id = idn(i, j);
                        const double o33 = P3[i] * P3[j], oAA = P12[i] * P12[j];
                        const double dP3 = d3 * (dd - id * m12 - (P12[i] * et[j] + et[i] * P12[j]) + m12 * oAA +
                                                 (m12 - 2.0 * e3) * o33 + e3 * (P12[i] * P3[j] + P3[i] * P12[j]));
                        w.put(i + 6 * j, (s3 - s1) * dP3 + lam * vol(i) * vol(j) + G * (oAA + 2.0 * o33) -
                                             Dq[i] * Dp[j] * inv);
                    }
                }
            }
            return 2;
        }
    }

    if (g_sl > g_sr) {
        const double den_r = 4.0 * lam * sp * sq + 2.0 * G * (1.0 + sp) * (1.0 + sq) + G * (1.0 - sp) * (1.0 - sq);
        const double l_r = (G * (2.0 * (1.0 + sp) * e1 - (1.0 - sp) * (e2 + e3)) + 2.0 * lam * sp * I1 - cb) / den_r;
        const double g_ra = (2.0 * e1 - e2 - e3) / (3.0 + sq);
        if (l_r >= g_sr && l_r <= g_ra) {
            const double d1 = 1.0 / ((e1 - e2) * (e1 - e3));
            double P1[6], P23[6];
            eig_proj(esq, et, e2 + e3, e2 * e3, d1, P1);
#pragma unroll
            for (int i = 0; i < 6; ++i) P23[i] = vol(i) - P1[i];
            const double s1 = lam * I1 + 2.0 * G * e1 - l_r * (2.0 * lam * sq + 2.0 * G * (1.0 + sq));
            const double s3 = lam * I1 + G * (e2 + e3) - l_r * (2.0 * lam * sq - G * (1.0 - sq));
#pragma unroll
            for (int i = 0; i < 6; ++i) S[i] = s1 * P1[i] + s3 * P23[i];
            if (w.on()) {
                double Dp[6], Dq[6];
#pragma unroll
                for (int i = 0; i < 6; ++i) {
                    Dp[i] = G * (2.0 * (1.0 + sp) * P1[i] - (1.0 - sp) * P23[i]) + 2.0 * lam * sp * vol(i);
                    Dq[i] = G * (2.0 * (1.0 + sq) * P1[i] - (1.0 - sq) * P23[i]) + 2.0 * lam * sq * vol(i);
                }
                const double inv = 1.0 / den_r, m23 = e2 + e3;
#pragma unroll
                for (int j = 0; j < 6; ++j) {
#pragma unroll
                    for (int i = 0; i < 6; ++i) {
                        const double dd = dsq(i, j, et), id = idn(i, j);
                        const double o11 = P1[i] * P1[j], oBB = P23[i] * P23[j];
                        const double dP1 = d1 * (dd - id * m23 - (P23[i] * et[j] + et[i] * P23[j]) + m23 * oBB +
                                                 (m23 - 2.0 * e1) * o11 + e1 * (P23[i] * P1[j] + P1[i] * P23[j]));
                        w.put(i + 6 * j, (s1 - s3) * dP1 + lam * vol(i) * vol(j) + G * (2.0 * o11 + oBB) -
                                             Dq[i] * Dp[j] * inv);
                    }
                }
            }
            return 3;
        }
    }

    const double sa = cb / (2.0 * sp);
#pragma unroll
    for (int i = 0; i < 6; ++i) S[i] = i < 3 ? sa : 0.0;
    if (w.on()) {
#pragma unroll
        for (int k = 0; k < 36; ++k) w.put(k, 0.0);
    }
    return 4;
}

// Material data at one point: per-point arrays, or a per-element id into a table of
// (c_bar, sin_phi, shear, lame) rows.
struct MatSource {
    const double *c_bar, *sin_phi, *shear, *lame;  // per point (table == nullptr)
    const int32_t *mat_id;                           // per element (table != nullptr)
    const double *table;                             // [n_mat][4]
    int n_q;
    __device__ __forceinline__ void get(int64_t p, double &cb, double &sp, double &G, double &lam) const {
        if (table) {
            const double *row = table + 4 * (int64_t)mat_id[p / n_q];
            cb = row[0]; sp = row[1]; G = row[2]; lam = row[3];
        } else {
            cb = c_bar[p]; sp = sin_phi[p]; G = shear[p]; lam = lame[p];
        }
    }
};

// Tangent rows of a workgroup's points staged in LDS (row stride 37 doubles), written out to the
// point-major DS array afterwards by consecutive lanes at consecutive addresses.
struct LdsDSWriter {
    double *row;
    bool enabled;
    __device__ __forceinline__ void put(int k, double v) const { row[k] = v; }
    __device__ __forceinline__ bool on() const { return enabled; }
};

// The packed form of the same rows: only the 21 upper entries D[s][m], s <= m, row by row (the
// packed layout above), so a workgroup's rows are its points' packed tangents back to back.
// k is a compile-time constant at every call: the 15 lower entries are neither stored nor formed
// (137 us against 166 us for the return mapping of 1.29M stratified points).  A product that the
// full form shares between entries (i, j) and (j, i) may fuse into an FMA here, so a packed entry
// can differ from the full kernel's in its last bit.
struct LdsSymDSWriter {
    double *row;
    bool enabled;
    __device__ __forceinline__ void put(int k, double v) const {
        const int s = k % 6, m = k / 6;
        if (s <= m) row[s * 6 - s * (s - 1) / 2 + (m - s)] = v;
    }
    __device__ __forceinline__ bool on() const { return enabled; }
};

constexpr int kCBlock = 128;   // points per workgroup
constexpr int kDStride = 37;   // LDS row stride of a point's tangent (36 + 1 against bank conflicts)

// One lane per point.  The block's strains are read through LDS (one coalesced pass), the return
// mapping writes the point's tangent into LDS rows, and the block's DS (and the FP32 copy of DS)
// and then S leave in coalesced passes: the block owns the contiguous ranges S[p0*6 .. (p0+np)*6)
// and DS[p0*36 .. (p0+np)*36).  One LDS area serves all three stagings in turn (E, then DS, then
// S, with barriers between), which keeps a block at 38 KB of LDS: 4 blocks (8 wavefronts) per CU.
// Persistent blocks walk the point blocks blk, blk + grid, ...; the next block's strains and
// material are loaded into registers while the current block's points are mapped, so their
// latency, and the drain of the current block's stores, overlap the arithmetic.
//
// SYM: the tangent is written packed (21 values per point); the LDS rows are then packed too (stride
// 21), the lower entries are never formed, and the block's rows leave as one contiguous range.
template <bool SYM>
__global__ void __launch_bounds__(kCBlock) k_constitutive_3d(const double *__restrict__ E, MatSource mat, int64_t n,
                                                             double *__restrict__ S, DSWriter w,
                                                             int8_t *__restrict__ branch) {
    constexpr int RS = SYM ? 21 : kDStride;
    __shared__ double sds[kCBlock * RS];
    const int t = threadIdx.x;
    const int64_t nb = (n + kCBlock - 1) / kCBlock;
    int64_t blk = blockIdx.x;
    if (blk >= nb) return;   // whole workgroup
    double pe[6], pcb = 0.0, psp = 0.0, pG = 0.0, plam = 0.0;
    // global -> registers for point block bk (lanes past the data read index 0 of the block)
    auto fetch = [&](int64_t bk) {
        const int64_t q0 = bk * kCBlock;
        const int qn = (int)(n - q0 < kCBlock ? n - q0 : kCBlock);
#pragma unroll
        for (int m = 0; m < 6; ++m) {
            const int i = t + kCBlock * m;
            pe[m] = E[q0 * 6 + (i < 6 * qn ? i : 0)];
        }
        mat.get(q0 + (t < qn ? t : 0), pcb, psp, pG, plam);
    };
    fetch(blk);
    const bool want = w.on();
    for (; blk < nb; blk += gridDim.x) {
        const int64_t p0 = blk * kCBlock;
        const int np = (int)(n - p0 < kCBlock ? n - p0 : kCBlock);
#pragma unroll
        for (int m = 0; m < 6; ++m) {
            const int i = t + kCBlock * m;
            if (i < 6 * np) sds[(i / 6) * 7 + i % 6] = pe[m];
        }
        const double cb = pcb, sp = psp, G = pG, lam = plam;
        __syncthreads();
        double e[6], s[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            e[c] = t < np ? sds[t * 7 + c] : 0.0;
            s[c] = 0.0;
        }
        __syncthreads();   // every lane holds its strain: the area now takes the tangent rows
        if (blk + gridDim.x < nb) fetch(blk + gridDim.x);
        if (t < np) {
            int b;
            if constexpr (SYM) {
                b = mc3d_point(e, cb, sp, G, lam, s, LdsSymDSWriter{sds + t * RS, want});
            } else {
                b = mc3d_point(e, cb, sp, G, lam, s, LdsDSWriter{sds + t * RS, want});
            }
            if (branch) branch[p0 + t] = (int8_t)b;
        }
        __syncthreads();
        if (want && !SYM) {
            for (int i = t; i < 36 * np; i += kCBlock) {
                const double v = sds[(i / 36) * RS + i % 36];
                if (w.ds) w.ds[p0 * 36 + i] = v;
                if (w.ds32) w.ds32[p0 * 36 + i] = (float)v;
            }
        } else if (want) {
            for (int i = t; i < 21 * np; i += kCBlock) {
                const double v = sds[i];
                if (w.ds) w.ds[p0 * 21 + i] = v;
                if (w.ds32) w.ds32[p0 * 21 + i] = (float)v;
            }
        }
        if (want) __syncthreads();   // tangent rows read out: the area now takes the stresses
        if (t < np) {
#pragma unroll
            for (int c = 0; c < 6; ++c) sds[t * 7 + c] = s[c];
        }
        __syncthreads();
        for (int i = t; i < 6 * np; i += kCBlock) S[p0 * 6 + i] = sds[(i / 6) * 7 + i % 6];
        __syncthreads();   // stresses read out: the area takes the next block's strains
    }
}

// Persistent grid of the 3D return mapping: 32 workgroups per CU (4 resident).  Measured on MI355X
// on 1.29M points: 190 us with one workgroup per point block and no prefetch, 165 / 160 / 153 us
// at 8 / 16 / 24-64 workgroups per CU.
constexpr int kCBlocksPerCU = 32;

unsigned const3d_grid(int64_t n) {
    static int n_cu = 0;
    if (n_cu == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n_cu <= 0)
            n_cu = 256;
    }
    const int64_t nb = (n + kCBlock - 1) / kCBlock;
    const int64_t want = (int64_t)n_cu * kCBlocksPerCU;
    return (unsigned)(nb < want ? nb : want);
}

// ---------------------------------------------------------------------------------------------
// Plane strain (2D).  Strain rows [e11, e22, g12]; stress rows [s11, s22, s12, s33]; DS is 3x3
// column-major (9 rows).  Follows constitutive/problem.py:143-475: in-plane eigenpairs from the
// 2x2 block, the out-of-plane eigenvalue 0, the reference's sort with its tie rules, and the same
// five branches with the 2D denominators.
// ---------------------------------------------------------------------------------------------
__device__ int mc2d_point(const double E[3], double cb, double sp, double G, double K, double lam, double S[4],
                          double DS[9], bool want_ds) {
    const double ex = E[0], ey = E[1], gxy = E[2], ez = 0.0;
    const double I1p = ex + ey;
    const double I2p = sqrt((ex - ey) * (ex - ey) + gxy * gxy);
    const double a1 = 0.5 * (I1p + I2p), a2 = 0.5 * (I1p - I2p), a3 = ez;
    // in-plane projections in strain representation (rows 11, 22, 12, 33)
    double P0a[4], P0b[4];
    double D0a[9];
    if (I2p == 0.0) {
        P0a[0] = 1.0; P0a[1] = 1.0; P0a[2] = 0.0; P0a[3] = 0.0;
#pragma unroll
        for (int k = 0; k < 9; ++k) D0a[k] = 0.0;
    } else {
        P0a[0] = (ex - a2) / I2p; P0a[1] = (ey - a2) / I2p; P0a[2] = (gxy / 2.0) / I2p; P0a[3] = 0.0;
    }
    P0b[0] = 1.0 - P0a[0]; P0b[1] = 1.0 - P0a[1]; P0b[2] = -P0a[2]; P0b[3] = 0.0;
    if (I2p != 0.0) {
        const double x1 = P0a[0], x2 = P0a[1], x3 = P0a[2];
        const double y1 = P0b[0], y2 = P0b[1], y3 = P0b[2];
        D0a[0] = (1.0 - x1 * x1 - y1 * y1) / I2p;
        D0a[1] = -(x2 * x1 + y2 * y1) / I2p;
        D0a[2] = -(x3 * x1 + y3 * y1) / I2p;
        D0a[3] = -(x1 * x2 + y1 * y2) / I2p;
        D0a[4] = (1.0 - x2 * x2 - y2 * y2) / I2p;
        D0a[5] = -(x3 * x2 + y3 * y2) / I2p;
        D0a[6] = -(x1 * x3 + y1 * y3) / I2p;
        D0a[7] = -(x2 * x3 + y2 * y3) / I2p;
        D0a[8] = (0.5 - x3 * x3 - y3 * y3) / I2p;
    }
    const double P0c[4] = {0.0, 0.0, 0.0, 1.0};
    // sort (problem.py:224-241): test2 -> (a1, a3, a2); test3 -> (a3, a1, a2)
    const bool t2 = (a1 >= a3) && (a3 > a2);
    const bool t3 = a3 > a1;
    // t2 and t3 exclude each other (a1 >= a3 vs a3 > a1); the ordered projections are selected by
    // value (no per-lane choice between arrays, which would put them in scratch memory)
    const double e1 = t3 ? a3 : a1;
    const double e2 = t3 ? a1 : (t2 ? a3 : a2);
    const double e3 = (t2 || t3) ? a2 : a3;
    double P1[4], P2[4], P3[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        P1[i] = t3 ? P0c[i] : P0a[i];
        P2[i] = t3 ? P0a[i] : (t2 ? P0c[i] : P0b[i]);
        P3[i] = (t2 || t3) ? P0b[i] : P0c[i];
    }
    // 0 -> +D0a, 1 -> -D0a, 2 -> zero
    const int D1 = t3 ? 2 : 0;
    const int D2 = t3 ? 0 : (t2 ? 2 : 1);
    const int D3 = (t2 || t3) ? 1 : 2;
    auto Dm = [&](int which, int k) -> double { return which == 0 ? D0a[k] : (which == 1 ? -D0a[k] : 0.0); };

    const double tr = e1 + e2 + e3;
    const double f = 2.0 * G * ((1.0 + sp) * e1 - (1.0 - sp) * e3) + 2.0 * lam * sp * tr - cb;
    const double g_sl = (e1 - e2) / (1.0 + sp);
    const double g_sr = (e2 - e3) / (1.0 - sp);
    const double g_la = (e1 + e2 - 2.0 * e3) / (3.0 - sp);
    const double g_ra = (2.0 * e1 - e2 - e3) / (3.0 + sp);
    const double den_s = 4.0 * lam * sp * sp + 2.0 * G * (1.0 + sp) * (1.0 + sp) + 2.0 * G * (1.0 - sp) * (1.0 - sp);
    const double den_l = 4.0 * lam * sp * sp + G * (1.0 + sp) * (1.0 + sp) + 2.0 * G * (1.0 - sp) * (1.0 - sp);
    const double den_r = 4.0 * lam * sp * sp + 2.0 * G * (1.0 + sp) * (1.0 + sp) + G * (1.0 - sp) * (1.0 - sp);
    (void)K;
    const double l_s = f / den_s;
    const double l_l = (G * ((1.0 + sp) * (e1 + e2) - 2.0 * (1.0 - sp) * e3) + 2.0 * lam * sp * tr - cb) / den_l;
    const double l_r = (G * (2.0 * (1.0 + sp) * e1 - (1.0 - sp) * (e2 + e3)) + 2.0 * lam * sp * tr - cb) / den_r;
    const double v3[3] = {1.0, 1.0, 0.0};

    if (f <= 0.0) {
        // S = lame * vol4 E_tr + 2G ident4 E_tr with E_tr = [ex, ey, gxy, 0]
        S[0] = lam * (ex + ey) + 2.0 * G * ex;
        S[1] = lam * (ex + ey) + 2.0 * G * ey;
        S[2] = 2.0 * G * 0.5 * gxy;
        S[3] = lam * (ex + ey);
        if (want_ds) {
            for (int j = 0; j < 3; ++j)
                for (int i = 0; i < 3; ++i)
                    DS[i + 3 * j] = v3[i] * v3[j] * lam + 2.0 * (i == j ? (i < 2 ? 1.0 : 0.5) : 0.0) * G;
        }
        return 0;
    }
    if (l_s <= fmin(g_sl, g_sr)) {
        const double s1 = lam * tr + 2.0 * G * e1 - l_s * (2.0 * lam * sp + 2.0 * G * (1.0 + sp));
        const double s2 = lam * tr + 2.0 * G * e2 - l_s * (2.0 * lam * sp);
        const double s3 = lam * tr + 2.0 * G * e3 - l_s * (2.0 * lam * sp - 2.0 * G * (1.0 - sp));
        for (int i = 0; i < 4; ++i) S[i] = s1 * P1[i] + s2 * P2[i] + s3 * P3[i];
        if (want_ds) {
            double q6[3];
            for (int i = 0; i < 3; ++i)
                q6[i] = 2.0 * G * (1.0 + sp) * P1[i] - 2.0 * G * (1.0 - sp) * P3[i] + v3[i] * (2.0 * lam * sp);
            for (int j = 0; j < 3; ++j)
                for (int i = 0; i < 3; ++i) {
                    const int k = i + 3 * j;
                    const double m1 = s1 * Dm(D1, k) + s2 * Dm(D2, k) + s3 * Dm(D3, k);
                    const double m2 = v3[i] * v3[j] * lam;
                    const double m3 = 2.0 * G * P1[i] * P1[j];
                    const double m4 = 2.0 * G * P2[i] * P2[j];
                    const double m5 = 2.0 * G * P3[i] * P3[j];
                    const double m6 = q6[i] * q6[j] / den_s;
                    DS[k] = m1 + m2 + m3 + m4 + m5 - m6;
                }
        }
        return 1;
    }
    if ((g_sl < g_sr) && (l_l >= g_sl) && (l_l <= g_la)) {
        const double s1 = lam * tr + G * (e1 + e2) - l_l * (2.0 * lam * sp + G * (1.0 + sp));
        const double s3 = lam * tr + 2.0 * G * e3 - l_l * (2.0 * lam * sp - 2.0 * G * (1.0 - sp));
        for (int i = 0; i < 4; ++i) S[i] = s1 * (P1[i] + P2[i]) + s3 * P3[i];
        if (want_ds) {
            double P12[3], q6[3];
            for (int i = 0; i < 3; ++i) {
                P12[i] = P1[i] + P2[i];
                q6[i] = G * (1.0 + sp) * P12[i] - 2.0 * G * (1.0 - sp) * P3[i] + v3[i] * (2.0 * lam * sp);
            }
            for (int j = 0; j < 3; ++j)
                for (int i = 0; i < 3; ++i) {
                    const int k = i + 3 * j;
                    const double m1 = s1 * (Dm(D1, k) + Dm(D2, k)) + s3 * Dm(D3, k);
                    DS[k] = m1 + v3[i] * v3[j] * lam + G * P12[i] * P12[j] + 2.0 * G * P3[i] * P3[j] -
                            q6[i] * q6[j] / den_l;
                }
        }
        return 2;
    }
    if ((g_sl > g_sr) && (l_r >= g_sr) && (l_r <= g_ra)) {
        const double s1 = lam * tr + 2.0 * G * e1 - l_r * (2.0 * lam * sp + 2.0 * G * (1.0 + sp));
        const double s3 = lam * tr + G * (e2 + e3) - l_r * (2.0 * lam * sp - G * (1.0 - sp));
        for (int i = 0; i < 4; ++i) S[i] = s1 * P1[i] + s3 * (P2[i] + P3[i]);
        if (want_ds) {
            double P23[3], q6[3];
            for (int i = 0; i < 3; ++i) {
                P23[i] = P2[i] + P3[i];
                q6[i] = 2.0 * G * (1.0 + sp) * P1[i] - G * (1.0 - sp) * P23[i] + v3[i] * (2.0 * lam * sp);
            }
            for (int j = 0; j < 3; ++j)
                for (int i = 0; i < 3; ++i) {
                    const int k = i + 3 * j;
                    const double m1 = s1 * Dm(D1, k) + s3 * (Dm(D2, k) + Dm(D3, k));
                    DS[k] = m1 + v3[i] * v3[j] * lam + 2.0 * G * P1[i] * P1[j] + G * P23[i] * P23[j] -
                            q6[i] * q6[j] / den_r;
                }
        }
        return 3;
    }
    const double sa = cb / (2.0 * sp);
    S[0] = sa; S[1] = sa; S[2] = 0.0; S[3] = sa;
    if (want_ds)
        for (int k = 0; k < 9; ++k) DS[k] = 0.0;
    return 4;
}

__global__ void __launch_bounds__(256) k_constitutive_2d(const double *__restrict__ E, MatSource mat,
                                                         const double *__restrict__ bulk, int64_t n,
                                                         double *__restrict__ S, double *__restrict__ DS,
                                                         float *__restrict__ DS32, int8_t *__restrict__ branch) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    double e[3] = {E[p * 3], E[p * 3 + 1], E[p * 3 + 2]};
    double cb, sp, G, lam;
    mat.get(p, cb, sp, G, lam);
    const double K = bulk ? bulk[p] : 0.0;
    double s[4], ds[9];
    const bool want = DS != nullptr || DS32 != nullptr;
    const int b = mc2d_point(e, cb, sp, G, K, lam, s, ds, want);
#pragma unroll
    for (int c = 0; c < 4; ++c) S[p * 4 + c] = s[c];
    if (DS) {
#pragma unroll
        for (int k = 0; k < 9; ++k) DS[p * 9 + k] = ds[k];
    }
    if (DS32) {
#pragma unroll
        for (int k = 0; k < 9; ++k) DS32[p * 9 + k] = (float)ds[k];
    }
    if (branch) branch[p] = (int8_t)b;
}

// Davis strength reduction per point (constitutive/reduction.py:14-78).  type 0/1/2 = A/B/C.
__device__ void davis(int type, double c0, double phi, double psi, double lam, double &cb, double &sp) {
    if (type == 0) {
        const double beta = cos(phi) * cos(psi) / (1.0 - sin(phi) * sin(psi));
        const double c0l = beta * c0 / lam;
        const double phil = atan(beta * tan(phi) / lam);
        cb = 2.0 * c0l * cos(phil);
        sp = sin(phil);
    } else if (type == 1) {
        const double c01 = c0 / lam;
        const double phi1 = atan(tan(phi) / lam);
        const double psi1 = atan(tan(psi) / lam);
        const double beta = cos(phi1) * cos(psi1) / (1.0 - sin(phi1) * sin(psi1));
        const double phil = atan(beta * tan(phi1));
        cb = 2.0 * (beta * c01) * cos(phil);
        sp = sin(phil);
    } else {
        const double c01 = c0 / lam;
        const double phi1 = atan(tan(phi) / lam);
        const double beta = phi1 > psi ? cos(phi1) * cos(psi) / (1.0 - sin(phi1) * sin(psi)) : 1.0;
        const double phil = atan(beta * tan(phi1));
        cb = 2.0 * (beta * c01) * cos(phil);
        sp = sin(phil);
    }
}

// Full point-major DS (n x 36) from the packed symmetric one (n x 21): entry s + 6 m takes packed
// entry (min(s, m), max(s, m)).
__global__ void __launch_bounds__(256) k_ds_unpack_sym(int64_t n, const double *__restrict__ DSp,
                                                       double *__restrict__ DS) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * 36) return;
    const int64_t p = i / 36;
    const int k = (int)(i - p * 36);
    const int r = k % 6, c = k / 6;
    const int s = r < c ? r : c, m = r < c ? c : r;
    DS[i] = DSp[p * 21 + s * 6 - s * (s - 1) / 2 + (m - s)];
}

__global__ void k_davis(const double *__restrict__ c0, const double *__restrict__ phi,
                        const double *__restrict__ psi, int64_t n, double lam, int type, double *__restrict__ cb,
                        double *__restrict__ sp) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    double a, b;
    davis(type, c0[p], phi[p], psi[p], lam, a, b);
    cb[p] = a;
    sp[p] = b;
}

}  // namespace

extern "C" {

int ssx_reduction(const double *c0, const double *phi, const double *psi, int64_t n, double lam, int davis_type,
                  double *c_bar, double *sin_phi, void *stream) {
    SSX_REQUIRE(n >= 0, "ssx_reduction: n < 0");
    SSX_REQUIRE(lam > 0.0, "ssx_reduction: Reduction parameter lambda must be positive");
    SSX_REQUIRE(davis_type >= 0 && davis_type <= 2, "ssx_reduction: Incorrect choice of the Davis approach");
    if (n == 0) return SSX_OK;
    hipLaunchKernelGGL(k_davis, dim3(ssx::grid_for(n, 256)), dim3(256), 0, ssx::as_stream(stream), c0, phi, psi, n,
                       lam, davis_type, c_bar, sin_phi);
    SSX_CHECK_LAUNCH("ssx_reduction");
    return SSX_OK;
}

int ssx_constitutive_3d(const double *E, const double *c_bar, const double *sin_phi, const double *shear,
                        const double *lame, int64_t n, double *S, double *DS, float *DS32, int8_t *branch,
                        void *stream) {
    SSX_REQUIRE(n >= 0, "ssx_constitutive_3d: n < 0");
    SSX_REQUIRE(E && S && c_bar && sin_phi && shear && lame, "ssx_constitutive_3d: null pointer");
    if (n == 0) return SSX_OK;
    MatSource m{c_bar, sin_phi, shear, lame, nullptr, nullptr, 1};
    DSWriter w{DS, DS32};
    hipLaunchKernelGGL(k_constitutive_3d<false>, dim3(const3d_grid(n)), dim3(kCBlock), 0, ssx::as_stream(stream), E, m, n,
                       S, w, branch);
    SSX_CHECK_LAUNCH("ssx_constitutive_3d");
    return SSX_OK;
}

int ssx_constitutive_3d_sym(const double *E, const double *c_bar, const double *sin_phi, const double *shear,
                            const double *lame, int64_t n, double *S, double *DSp, int8_t *branch, void *stream) {
    SSX_REQUIRE(n >= 0, "ssx_constitutive_3d_sym: n < 0");
    SSX_REQUIRE(E && S && c_bar && sin_phi && shear && lame, "ssx_constitutive_3d_sym: null pointer");
    if (n == 0) return SSX_OK;
    MatSource m{c_bar, sin_phi, shear, lame, nullptr, nullptr, 1};
    DSWriter w{DSp, nullptr};
    hipLaunchKernelGGL(k_constitutive_3d<true>, dim3(const3d_grid(n)), dim3(kCBlock), 0, ssx::as_stream(stream), E, m, n,
                       S, w, branch);
    SSX_CHECK_LAUNCH("ssx_constitutive_3d_sym");
    return SSX_OK;
}

int ssx_ds_unpack_sym(int64_t n, const double *DSp, double *DS, void *stream) {
    SSX_REQUIRE(n >= 0, "ssx_ds_unpack_sym: n < 0");
    if (n == 0) return SSX_OK;
    SSX_REQUIRE(DSp && DS, "ssx_ds_unpack_sym: null pointer");
    hipLaunchKernelGGL(k_ds_unpack_sym, dim3(ssx::grid_for(n * 36, 256)), dim3(256), 0, ssx::as_stream(stream), n,
                       DSp, DS);
    SSX_CHECK_LAUNCH("ssx_ds_unpack_sym");
    return SSX_OK;
}

int ssx_constitutive_3d_mat(const double *E, const int32_t *mat_id, const double *table, int n_mat, int n_q,
                            int64_t n, double *S, double *DS, float *DS32, int8_t *branch, void *stream) {
    SSX_REQUIRE(n >= 0 && n_q > 0 && n % n_q == 0, "ssx_constitutive_3d_mat: n must be a multiple of n_q");
    SSX_REQUIRE(E && S && mat_id && table && n_mat > 0, "ssx_constitutive_3d_mat: null pointer");
    if (n == 0) return SSX_OK;
    MatSource m{nullptr, nullptr, nullptr, nullptr, mat_id, table, n_q};
    DSWriter w{DS, DS32};
    hipLaunchKernelGGL(k_constitutive_3d<false>, dim3(const3d_grid(n)), dim3(kCBlock), 0, ssx::as_stream(stream), E, m, n,
                       S, w, branch);
    SSX_CHECK_LAUNCH("ssx_constitutive_3d_mat");
    return SSX_OK;
}

int ssx_constitutive_2d(const double *E, const double *c_bar, const double *sin_phi, const double *shear,
                        const double *bulk, const double *lame, int64_t n, double *S, double *DS, float *DS32,
                        int8_t *branch, void *stream) {
    SSX_REQUIRE(n >= 0, "ssx_constitutive_2d: n < 0");
    SSX_REQUIRE(E && S && c_bar && sin_phi && shear && lame, "ssx_constitutive_2d: null pointer");
    if (n == 0) return SSX_OK;
    MatSource m{c_bar, sin_phi, shear, lame, nullptr, nullptr, 1};
    hipLaunchKernelGGL(k_constitutive_2d, dim3(ssx::grid_for(n, 256)), dim3(256), 0, ssx::as_stream(stream), E, m,
                       bulk, n, S, DS, DS32, branch);
    SSX_CHECK_LAUNCH("ssx_constitutive_2d");
    return SSX_OK;
}

}  // extern "C"
