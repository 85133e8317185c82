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

// Gradient row of atom i; lane 0 writes the six components into g.
__device__ __forceinline__ void grad_row(const DevSys &s, int i, int lane, double *__restrict__ g)
{
    const double C = MCX_CHGCHARMM * MCX_KCALTOIU;
    const float qi = __int_as_float(s.ipos[i].w);
    const float2 ei = s.es[i];
    const int *row = s.nbr + (size_t)i * s.cap;
    const int k1 = s.cnt[i];
    double fx = 0, fy = 0, fz = 0, ge = 0, gs = 0, gq = 0;
    for (int k = lane; k < k1; k += 64) {
        const int j = row[k];
        const float qj = __int_as_float(s.ipos[j].w);
        const float2 ej = s.es[j];
        double d[3];
        mcx_disp_A(s, i, j, d);
        const double r2 = fma(d[2], d[2], fma(d[1], d[1], d[0] * d[0]));
        if (!(r2 <= s.ctoff2)) continue;
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
        double S = 1.0, dS = 0.0;
        if (r2 > s.cton2) {
            const double a = s.ctoff2 - r2, b = s.ctoff2 + 2.0 * r2 - 3.0 * s.cton2;
            S = a * a * b * s.switch2;
            dS = (-2.0 * a * b + 2.0 * a * a) * s.switch2;
        }
        // dE/dr2 for this pair
        const double dEe = -0.5 * Ee * (ri * ri);                            // d(q q / r)/dr2
        const double dEv = (double)(ee * fmaf(-6.0f, t6, 3.0f * t3) * ri2);  // d(ee(t^6-t^3))/dr2
        const double gpair = S * (C * dEe + 4.0 * dEv) + dS * (C * Ee + 4.0 * (double)Ev);
        // r2 = |x_j - x_i|^2 -> dr2/dx_i = -2 d
        fx -= 2.0 * gpair * d[0];
        fy -= 2.0 * gpair * d[1];
        fz -= 2.0 * gpair * d[2];
        ge += S * 4.0 * (double)(ej.x * lj);
        if (sij != 0.f) gs += S * 4.0 * (double)(ee * fmaf(12.0f, t6, -6.0f * t3) / sij);
        gq += S * C * ((double)qj * ri);
    }
    fx = wsum(fx); fy = wsum(fy); fz = wsum(fz);
    ge = wsum(ge); gs = wsum(gs); gq = wsum(gq);
    if (lane == 0) {
        const size_t n = (size_t)s.n;
        g[i] = fx; g[n + i] = fy; g[2 * n + i] = fz;
        g[3 * n + i] = ge; g[4 * n + i] = gs; g[5 * n + i] = gq;
    }
}

__global__ __launch_bounds__(64 * kGW) void k_grad(DevSys s, double *__restrict__ g)
{
    const int i = blockIdx.x * kGW + (threadIdx.x >> 6);
    if (i < s.n) grad_row(s, i, threadIdx.x & 63, g);
}

// One launch for many systems: blockIdx.y selects the system.
__global__ __launch_bounds__(64 * kGW) void k_grad_batch(const DevSys *__restrict__ sys, double *const *__restrict__ g)
{
    const DevSys &s = sys[blockIdx.y];
    const int i = blockIdx.x * kGW + (threadIdx.x >> 6);
    if (i < s.n) grad_row(s, i, threadIdx.x & 63, g[blockIdx.y]);
}

// Per-type sums of the per-atom parameter gradients (components 3..5 of
// g6), one block per type, fixed summation order.
__global__ __launch_bounds__(256) void k_type_sums(const double *__restrict__ g6, int n, const int *__restrict__ type,
                                                   double *__restrict__ out)
{
    __shared__ double red[3][4];
    const int t = blockIdx.x;
    double a[3] = {0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < n; i += 256)
        if (type[i] == t)
            for (int c = 0; c < 3; c++) a[c] += g6[(size_t)(3 + c) * n + i];
    for (int c = 0; c < 3; c++) {
        const double v = wsum(a[c]);
        if ((threadIdx.x & 63) == 0) red[c][threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int c = threadIdx.x;
        out[3 * t + c] = red[c][0] + red[c][1] + red[c][2] + red[c][3];
    }
}

}  // namespace

void launch_nonbonded_grad(const DevSys &s, double *g6, hipStream_t st)
{
    hipLaunchKernelGGL(k_grad, dim3((s.n + kGW - 1) / kGW), dim3(64 * kGW), 0, st, s, g6);
}

void launch_nonbonded_grad_batch(const DevSys *sys, double *const *g6, int nsys, int nmax, hipStream_t st)
{
    hipLaunchKernelGGL(k_grad_batch, dim3((nmax + kGW - 1) / kGW, nsys), dim3(64 * kGW), 0, st, sys, g6);
}

void launch_type_sums(const double *g6, int n, const int *type, int ntypes, double *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_type_sums, dim3(ntypes), dim3(256), 0, st, g6, n, type, out);
}
