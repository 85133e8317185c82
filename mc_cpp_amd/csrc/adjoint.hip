// adjoint.hip -- hand-written adjoint of the switched nonbonded energy
// (the derivative of computeNonBonded_switch, src/FField_MDBAS.cpp:336-424).
//
// E = C sum S q_i q_j / r + 4 sum S e_i e_j (u^12 - u^6),  u = (s_i+s_j)/r,
// C = 332.0716*418.4, S the switching polynomial.  One wavefront per atom i
// walks i's FULL list row (both directions), so every per-atom sum --
// dE/dx_i, dE/de_i, dE/ds_i, dE/dq_i -- is owned by one wavefront: no
// atomics, deterministic FP64 results.  LJ pair arithmetic FP32; Coulomb,
// S, dS/dr^2 and all sums FP64.  A batched launch covers many systems
// (blockIdx.y = system), and a second kernel reduces the per-atom parameter
// gradients to per-type sums in FP64.
#include "terms.h"

namespace {

constexpr int kGW = 4;

__device__ __forceinline__ double wsum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Adds one pair's gradient (switched when S < 1 or dS != 0) to the lane's
// sums: Coulomb C q_i q_j / r in FP64, LJ 4 e_i e_j (u^12 - u^6) with
// u = (s_i + s_j) / r in FP32 (list and all-pair terms) or FP64 (1-4 terms,
// whose own parameters give the 1-4 components).
struct GradAcc {
    double fx = 0, fy = 0, fz = 0, ge = 0, gs = 0, gq = 0, ge14 = 0, gs14 = 0;
};

__device__ __forceinline__ void switch_of(const DevSys &s, double r2, bool sw, double &S, double &dS)
{
    S = 1.0; dS = 0.0;
    if (sw && r2 > s.cton2) {
        const double a = s.ctoff2 - r2, b = s.ctoff2 + 2.0 * r2 - 3.0 * s.cton2;
        S = a * a * b * s.switch2;
        dS = (-2.0 * a * b + 2.0 * a * a) * s.switch2;
    }
}

__device__ __forceinline__ void pair32(const DevSys &s, const double d[3], double r2, float qi, float qj, float2 ei,
                                       float2 ej, bool sw, GradAcc &A)
{
    const double C = MCX_CHGCHARMM * MCX_KCALTOIU;
    float ri32;
    const double ri = mcx_rsqrt(r2, ri32);
    const float ri2 = ri32 * ri32;
    const float sij = ei.y + ej.y;
    const float u = sij * ri32;
    const float t = u * u;
    const float t3 = t * t * t;
    const float t6 = t3 * t3;
    const float lj = t6 - t3;                        // u^12 - u^6
    const float ee = ei.x * ej.x;
    const double Ee = (double)qi * (double)qj * ri;  // q_i q_j / r   (times C)
    const float Ev = ee * lj;                        // e_i e_j (u^12 - u^6)  (times 4)
    double S, dS;
    switch_of(s, r2, sw, S, dS);
    const double dEe = -0.5 * Ee * (ri * ri);                            // d(q q / r)/dr2
    const double dEv = (double)(ee * fmaf(-6.0f, t6, 3.0f * t3) * ri2);  // d(ee(t^6-t^3))/dr2
    const double gpair = S * (C * dEe + 4.0 * dEv) + dS * (C * Ee + 4.0 * (double)Ev);
    // r2 = |x_j - x_i|^2 -> dr2/dx_i = -2 d
    A.fx -= 2.0 * gpair * d[0];
    A.fy -= 2.0 * gpair * d[1];
    A.fz -= 2.0 * gpair * d[2];
    A.ge += S * 4.0 * (double)(ej.x * lj);
    if (sij != 0.f) A.gs += S * 4.0 * (double)(ee * fmaf(12.0f, t6, -6.0f * t3) / sij);
    A.gq += S * C * ((double)qj * ri);
}

__device__ __forceinline__ void pair14(const DevSys &s, int i, int j, bool sw, GradAcc &A)
{
    const double C = MCX_CHGCHARMM * MCX_KCALTOIU;
    double d[3];
    mcx_disp_A(s, i, j, d);
    const double r2 = fma(d[2], d[2], fma(d[1], d[1], d[0] * d[0]));
    if (sw && !(r2 <= s.ctoff2)) return;
    const double ri = 1.0 / sqrt(r2);
    const double ei = s.eps1464[i], ej = s.eps1464[j], sij = s.sig1464[i] + s.sig1464[j];
    const double qi = s.q64[i], qj = s.q64[j];
    const double u = sij * ri, u2 = u * u, t3 = u2 * u2 * u2, t6 = t3 * t3;
    const double ee = ei * ej, Ee = qi * qj * ri, Ev = ee * (t6 - t3);
    double S, dS;
    switch_of(s, r2, sw, S, dS);
    const double gpair = S * (C * (-0.5 * Ee * ri * ri) + 4.0 * ee * (-6.0 * t6 + 3.0 * t3) * ri * ri) +
                         dS * (C * Ee + 4.0 * Ev);
    A.fx -= 2.0 * gpair * d[0];
    A.fy -= 2.0 * gpair * d[1];
    A.fz -= 2.0 * gpair * d[2];
    A.ge14 += S * 4.0 * ej * (t6 - t3);
    if (sij != 0.0) A.gs14 += S * 4.0 * ee * (12.0 * t6 - 6.0 * t3) / sij;
    A.gq += S * C * qj * ri;
}

// Gradient row of atom i: SWITCH mode walks i's full list row (both
// directions), FULL mode every other atom except i's excluded partners
// (pair (a, b), a < b, excluded iff b is in a's visible exclusion row; the
// sorted partner list is merged against the lane's ascending j); then i's
// 1-4 partners.  Lane 0 writes the eight components into g (8 n values:
// dE/dx, dE/dy, dE/dz, dE/deps, dE/dsig, dE/dq, dE/deps14, dE/dsig14).
template <bool FULL>
__device__ __forceinline__ void grad_row(const DevSys &s, int i, int lane, double *__restrict__ g)
{
    const float qi = __int_as_float(s.ipos[i].w);
    const float2 ei = s.es[i];
    GradAcc A;
    if (FULL) {
        const int *xr = s.xfull_idx + s.xfull_off[i];
        const int nx = s.xfull_off[i + 1] - s.xfull_off[i];
        int p = 0;
        for (int j = lane; j < s.n; j += 64) {
            while (p < nx && xr[p] < j) p++;
            if (j == i || (p < nx && xr[p] == j)) continue;
            double d[3];
            mcx_disp_A(s, i, j, d);
            const double r2 = fma(d[2], d[2], fma(d[1], d[1], d[0] * d[0]));
            pair32(s, d, r2, qi, __int_as_float(s.ipos[j].w), ei, s.es[j], false, A);
        }
    } else {
        const int *row = s.nbr + (size_t)i * s.cap;
        const int k1 = s.cnt[i];
        for (int k = lane; k < k1; k += 64) {
            const int j = row[k];
            double d[3];
            mcx_disp_A(s, i, j, d);
            const double r2 = fma(d[2], d[2], fma(d[1], d[1], d[0] * d[0]));
            if (!(r2 <= s.ctoff2)) continue;
            pair32(s, d, r2, qi, __int_as_float(s.ipos[j].w), ei, s.es[j], true, A);
        }
    }
    if (s.p14_off)
        for (int k = s.p14_off[i] + lane; k < s.p14_off[i + 1]; k += 64) pair14(s, i, s.p14_idx[k], !FULL, A);
    const double v[8] = {wsum(A.fx), wsum(A.fy), wsum(A.fz), wsum(A.ge), wsum(A.gs), wsum(A.gq), wsum(A.ge14),
                         wsum(A.gs14)};
    if (lane == 0) {
        const size_t n = (size_t)s.n;
#pragma unroll
        for (int c = 0; c < 8; c++) g[c * n + i] = v[c];
    }
}

template <bool FULL>
__global__ __launch_bounds__(64 * kGW) void k_grad(DevSys s, double *__restrict__ g)
{
    const int i = blockIdx.x * kGW + (threadIdx.x >> 6);
    if (i < s.n) grad_row<FULL>(s, i, threadIdx.x & 63, g);
}

// One launch for many systems (SWITCH mode): blockIdx.y selects the system.
__global__ __launch_bounds__(64 * kGW) void k_grad_batch(const DevSys *__restrict__ sys, double *const *__restrict__ g)
{
    const DevSys &s = sys[blockIdx.y];
    const int i = blockIdx.x * kGW + (threadIdx.x >> 6);
    if (i < s.n) grad_row<false>(s, i, threadIdx.x & 63, g[blockIdx.y]);
}

// Per-type sums of the per-atom parameter gradients (components 3..7 of
// g8: eps, sig, q, eps14, sig14), one block per type, fixed summation order.
__global__ __launch_bounds__(256) void k_type_sums(const double *__restrict__ g8, int n, const int *__restrict__ type,
                                                   double *__restrict__ out)
{
    __shared__ double red[5][4];
    const int t = blockIdx.x;
    double a[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < n; i += 256)
        if (type[i] == t)
            for (int c = 0; c < 5; c++) a[c] += g8[(size_t)(3 + c) * n + i];
    for (int c = 0; c < 5; c++) {
        const double v = wsum(a[c]);
        if ((threadIdx.x & 63) == 0) red[c][threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x < 5) {
        const int c = threadIdx.x;
        out[5 * t + c] = red[c][0] + red[c][1] + red[c][2] + red[c][3];
    }
}

}  // namespace

void launch_nonbonded_grad(const DevSys &s, double *g8, hipStream_t st)
{
    if (s.cutmode == 0)     // all pairs
        hipLaunchKernelGGL(k_grad<true>, dim3((s.n + kGW - 1) / kGW), dim3(64 * kGW), 0, st, s, g8);
    else
        hipLaunchKernelGGL(k_grad<false>, dim3((s.n + kGW - 1) / kGW), dim3(64 * kGW), 0, st, s, g8);
}

void launch_nonbonded_grad_batch(const DevSys *sys, double *const *g8, int nsys, int nmax, hipStream_t st)
{
    hipLaunchKernelGGL(k_grad_batch, dim3((nmax + kGW - 1) / kGW, nsys), dim3(64 * kGW), 0, st, sys, g8);
}

void launch_type_sums(const double *g8, int n, const int *type, int ntypes, double *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_type_sums, dim3(ntypes), dim3(256), 0, st, g8, n, type, out);
}
