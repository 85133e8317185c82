// terms.h -- device functions shared by the fused bonded kernel and the
// per-move dE kernel: one bonded term or one listed pair, evaluated for
// coordinates fetched through a caller-supplied accessor (so the dE kernel
// can substitute trial coordinates of moved atoms).
#pragma once

#include "mcx_internal.h"

#pragma clang fp contract(off)

__device__ __forceinline__ double mcx_ref_rint(double v)
{
    return (double)(v >= 0. ? (int)(v + 0.5) : (int)(v - 0.5));
}

// PerConditions::applyPBC on a displacement (src/PerConditions.cpp:244-265)
__device__ __forceinline__ void mcx_pbc(const DevSys &s, double d[3])
{
    if (s.pbtype != 0) {
        d[0] -= s.L[0] * mcx_ref_rint(s.rL[0] * d[0]);
        d[1] -= s.L[1] * mcx_ref_rint(s.rL[1] * d[1]);
        d[2] -= s.L[2] * mcx_ref_rint(s.rL[2] * d[2]);
    }
}

// Tools::distance2 with delta = a2 - a1 (include/Tools.hpp:173-188)
__device__ __forceinline__ double mcx_dist2(const DevSys &s, const double a1[3], const double a2[3], double d[3])
{
    d[0] = a2[0] - a1[0];
    d[1] = a2[1] - a1[1];
    d[2] = a2[2] - a1[2];
    mcx_pbc(s, d);
    return d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
}

// Fractional coordinate -> 32-bit fixed point (wraps modulo the box).
__device__ __forceinline__ int mcx_to_fixed(double x, double rL)
{
    double v = x * rL;
    v = v - floor(v);
    return (int)(unsigned)(long long)(v * 4294967296.0);
}

__device__ __forceinline__ int4 mcx_fixed4(const DevSys &s, double x, double y, double z, float q)
{
    return make_int4(mcx_to_fixed(x, s.rL[0]), mcx_to_fixed(y, s.rL[1]), mcx_to_fixed(z, s.rL[2]),
                     __float_as_int(q));
}

// wrapping int32 difference b - a (minimum image in units of L/2^32)
__device__ __forceinline__ int mcx_wdiff(int b, int a) { return (int)((unsigned)b - (unsigned)a); }

// Minimum-image displacement x_j - x_i in A, FP64, from the fixed-point
// mirror (or from the FP64 master copy without periodic boundaries).
__device__ __forceinline__ void mcx_disp_A(const DevSys &s, int i, int j, double d[3])
{
    if (s.pbtype != 0) {
        const int4 a = s.ipos[i], b = s.ipos[j];
        d[0] = (double)mcx_wdiff(b.x, a.x) * s.unit[0];
        d[1] = (double)mcx_wdiff(b.y, a.y) * s.unit[1];
        d[2] = (double)mcx_wdiff(b.z, a.z) * s.unit[2];
    } else {
        d[0] = s.x[j] - s.x[i]; d[1] = s.y[j] - s.y[i]; d[2] = s.z[j] - s.z[i];
    }
}

// 1/sqrt(r2): the FP32 hardware estimate refined by one FP64 Newton step
// (relative error ~1e-14); the FP32 estimate is returned for the LJ part.
__device__ __forceinline__ double mcx_rsqrt(double r2, float &ri32)
{
    ri32 = __builtin_amdgcn_rsqf((float)r2);
    const double ri = (double)ri32;
    const double e = fma(-r2 * ri, ri, 1.0);
    return fma(0.5 * ri, e, ri);
}

// One listed pair from an FP64 displacement (already minimum-imaged, A):
// FP64 q_j/r, FP32 e_j*(t^6 - t^3) with t = ((s_i+s_j)/r)^2.
__device__ __forceinline__ void mcx_pair(double dx, double dy, double dz, float qj, float ej, float sij,
                                         double &r2, double &c, float &v)
{
    r2 = fma(dz, dz, fma(dy, dy, dx * dx));
    float ri32;
    const double ri = mcx_rsqrt(r2, ri32);
    const float u = sij * ri32;
    const float t = u * u;
    const float t3 = t * t * t;
    c = (double)qj * ri;
    v = ej * fmaf(t3, t3, -t3);
}

// Displacement x_j - x_i in the all-pair kernels' length unit:
// PB 1 cubic (L/2^32, exact int difference), 2 orthorhombic (A), 0 none (A).
template <int PB>
__device__ __forceinline__ void mcx_disp_k(const DevSys &s, const int4 &pi, const int4 &pj, double xi, double yi,
                                           double zi, int j, double &dx, double &dy, double &dz)
{
    if (PB == 1) {
        dx = (double)mcx_wdiff(pj.x, pi.x);
        dy = (double)mcx_wdiff(pj.y, pi.y);
        dz = (double)mcx_wdiff(pj.z, pi.z);
    } else if (PB == 2) {
        dx = (double)mcx_wdiff(pj.x, pi.x) * s.unit[0];
        dy = (double)mcx_wdiff(pj.y, pi.y) * s.unit[1];
        dz = (double)mcx_wdiff(pj.z, pi.z) * s.unit[2];
    } else {
        dx = s.x[j] - xi;
        dy = s.y[j] - yi;
        dz = s.z[j] - zi;
    }
}

// The per-pair values of the all-pair kernel, in two parts so that the
// Lennard-Jones part is skipped for atoms without it: 1/r in FP64 (kernel
// length units) from the FP64 hardware estimate (relative error <= 2^-23)
// and one Newton step (~2e-14), returning the estimate r0 too; and
// t^6 - t^3 in FP32 with t = (sij * r0)^2.  The excluded-pair correction
// calls the same two, so a pair added by the tile kernel is removed with
// identical values.
template <int PB>
__device__ __forceinline__ double mcx_tile_r2(const DevSys &s, const int4 &pi, const int4 &pj, double xi, double yi,
                                              double zi, int j)
{
    double dx, dy, dz;
    mcx_disp_k<PB>(s, pi, pj, xi, yi, zi, j, dx, dy, dz);
    return fma(dz, dz, fma(dy, dy, dx * dx));
}

__device__ __forceinline__ double mcx_tile_newton(double r2, double r0)
{
    const double e = fma(-r2 * r0, r0, 1.0);
    return fma(0.5 * r0, e, r0);
}

template <int PB>
__device__ __forceinline__ double mcx_tile_ri(const DevSys &s, const int4 &pi, const int4 &pj, double xi, double yi,
                                              double zi, int j, double &r0)
{
    const double r2 = mcx_tile_r2<PB>(s, pi, pj, xi, yi, zi, j);
    r0 = __builtin_amdgcn_rsq(r2);
    return mcx_tile_newton(r2, r0);
}

__device__ __forceinline__ float mcx_tile_lj(float sij, double r0)
{
    const float u = sij * (float)r0;
    const float t = u * u;
    const float t3 = t * t * t;
    return fmaf(t3, t3, -t3);
}

template <int PB>
__device__ __forceinline__ void mcx_tile_pair(const DevSys &s, const int4 &pi, const int4 &pj, double xi, double yi,
                                              double zi, int j, float sij, double &ri, float &lj)
{
    double r0;
    ri = mcx_tile_ri<PB>(s, pi, pj, xi, yi, zi, j, r0);
    lj = mcx_tile_lj(sij, r0);
}

__device__ __forceinline__ void mcx_mimg64(const DevSys &s, double d[3])
{
    if (s.pbtype != 0) {
        d[0] -= s.L[0] * rint(d[0] * s.rL[0]);
        d[1] -= s.L[1] * rint(d[1] * s.rL[1]);
        d[2] -= s.L[2] * rint(d[2] * s.rL[2]);
    }
}

// Switching function S(r2) in FP64 (src/FField_MDBAS.cpp:393-399).
__device__ __forceinline__ double mcx_switch(const DevSys &s, double r2)
{
    if (r2 <= s.cton2) return 1.0;
    const double a = s.ctoff2 - r2;
    const double b = fma(2.0, r2, s.ctoff2 - 3.0 * s.cton2);    // ctoff2 + 2 r2 - 3 cton2
    return (a * a) * (b * s.switch2);
}

struct CoordFetch {
    const double *x, *y, *z;
    // trial coordinates of atoms [mfirst, mlast] (LDS) when non-null
    const double *tx, *ty, *tz;
    int mfirst, mlast;
    __device__ __forceinline__ void get(int a, double d[3]) const
    {
        if (tx && a >= mfirst && a <= mlast) {
            d[0] = tx[a - mfirst]; d[1] = ty[a - mfirst]; d[2] = tz[a - mfirst];
        } else {
            d[0] = x[a]; d[1] = y[a]; d[2] = z[a];
        }
    }
};

// Energy of one term of kind `kind` (iu for bonded kinds; for the pair
// kinds the raw FP64 values elec_raw = q_i q_j / r * S and
// vdw_raw = e_i e_j (t^6 - t^3) * S, to be scaled by 332.0716*418.4 and 4).
// Bonded formulas restate src/FField_MDBAS.cpp:490-786 operation by
// operation in FP64.
__device__ inline void mcx_term_energy_rec(const DevSys &s, int kind, const int4 id, int type, const double4 p,
                                           const CoordFetch &f, double &e, double &eel, double &evdw)
{
    e = 0.0; eel = 0.0; evdw = 0.0;
    if (kind == kBond || kind == kUB) {
        double di[3], dj[3], d[3];
        f.get(id.x, di); f.get(id.y, dj);
        const double r = sqrt(mcx_dist2(s, di, dj, d));
        if (kind == kBond && type == 1) {
            const double ma = exp(-p.z * (r - p.y));
            const double mb = ma * ma;
            e = p.x * (mb - 2. * ma) + p.x;
        } else {
            const double x = r - p.y;
            e = 0.5 * p.x * (x * x);
        }
    } else if (kind == kAngle) {
        double di[3], dj[3], dk[3], dab[3], dbc[3];
        f.get(id.x, di); f.get(id.y, dj); f.get(id.z, dk);
        const double rab = sqrt(mcx_dist2(s, di, dj, dab));
        const double rbc = sqrt(mcx_dist2(s, dk, dj, dbc));
        const double cost = (dab[0] * dbc[0] + dab[1] * dbc[1] + dab[2] * dbc[2]) / (rab * rbc);
        const double th = acos(cost);
        const double x = th - p.y;
        e = 0.5 * p.x * (x * x);
    } else if (kind == kDihe || kind == kImpr) {
        double di[3], dj[3], dk[3], dl[3], dab[3], dbc[3], dcd[3], pb[3], pc[3];
        f.get(id.x, di); f.get(id.y, dj); f.get(id.z, dk); f.get(id.w, dl);
        dab[0] = di[0] - dj[0]; dab[1] = di[1] - dj[1]; dab[2] = di[2] - dj[2];
        mcx_pbc(s, dab);
        const double rbc = sqrt(mcx_dist2(s, dk, dj, dbc));
        dcd[0] = dk[0] - dl[0]; dcd[1] = dk[1] - dl[1]; dcd[2] = dk[2] - dl[2];
        mcx_pbc(s, dcd);
        pb[0] = dab[1] * dbc[2] - dab[2] * dbc[1];
        pb[1] = dab[2] * dbc[0] - dab[0] * dbc[2];
        pb[2] = dab[0] * dbc[1] - dab[1] * dbc[0];
        const double rpb = sqrt(pb[0] * pb[0] + pb[1] * pb[1] + pb[2] * pb[2]);
        pc[0] = dbc[1] * dcd[2] - dbc[2] * dcd[1];
        pc[1] = dbc[2] * dcd[0] - dbc[0] * dcd[2];
        pc[2] = dbc[0] * dcd[1] - dbc[1] * dcd[0];
        const double rpc = sqrt(pc[0] * pc[0] + pc[1] * pc[1] + pc[2] * pc[2]);
        const double pbpc = pb[0] * pc[0] + pb[1] * pc[1] + pb[2] * pc[2];
        const double cosp = pbpc / (rpb * rpc);
        const double sinp = (dbc[0] * (pc[1] * pb[2] - pc[2] * pb[1]) + dbc[1] * (pb[0] * pc[2] - pb[2] * pc[0]) +
                             dbc[2] * (pc[0] * pb[1] - pc[1] * pb[0])) / (rpb * rpc * rbc);
        double phi = atan2(sinp, cosp);
        if (type == 2) {
            phi = phi - p.y;
            phi = phi - mcx_ref_rint(phi / MCX_TWOPI) * MCX_TWOPI;
            e = 0.5 * p.x * (phi * phi);
        } else {
            e = p.x * (1. + cos(p.z * phi - p.y));
        }
    } else {   // kPair14 (14 parameters, switched in SWITCH mode) or kExclPair
        double di[3], dj[3];
        f.get(id.x, di); f.get(id.y, dj);
        const bool p14 = kind == kPair14;
        if (!p14) {
            // excluded pair: the all-pair kernel's own values, subtracted
            const int4 pi = s.ipos[id.x], pj = s.ipos[id.y];
            const float2 ei = s.esk[id.x], ej = s.esk[id.y];
            double ri;
            float lj;
            const double xi = s.x[id.x], yi = s.y[id.x], zi = s.z[id.x];
            if (s.pbtype == 1) mcx_tile_pair<1>(s, pi, pj, xi, yi, zi, id.y, ei.y + ej.y, ri, lj);
            else if (s.pbtype == 2) mcx_tile_pair<2>(s, pi, pj, xi, yi, zi, id.y, ei.y + ej.y, ri, lj);
            else mcx_tile_pair<0>(s, pi, pj, xi, yi, zi, id.y, ei.y + ej.y, ri, lj);
            eel = -((double)__int_as_float(pi.w) * ((double)__int_as_float(pj.w) * ri)) * s.iunit;
            evdw = lj == 0.f || ej.x == 0.f ? 0.0 : -((double)ei.x * ((double)ej.x * (double)lj));
            return;
        }
        // FP64 throughout: the per-move dE differences this value
        double d[3] = {dj[0] - di[0], dj[1] - di[1], dj[2] - di[2]};
        mcx_mimg64(s, d);
        const double r2 = fma(d[2], d[2], fma(d[1], d[1], d[0] * d[0]));
        double sw = 1.0;
        if (s.cutmode == 1) {
            if (!(r2 <= s.ctoff2)) return;
            sw = mcx_switch(s, r2);
        }
        float ri32;
        const double ri = mcx_rsqrt(r2, ri32);
        const double sij = s.sig1464[id.x] + s.sig1464[id.y];
        const double t = sij * sij * (ri * ri);
        const double t3 = t * t * t;
        eel = sw * (s.q64[id.x] * (s.q64[id.y] * ri));
        evdw = sw * (s.eps1464[id.x] * (s.eps1464[id.y] * fma(t3, t3, -t3)));
    }
}

__device__ inline void mcx_term_energy(const DevSys &s, const DevTerms &t, int kind, int k, const CoordFetch &f,
                                       double &e, double &eel, double &evdw)
{
    mcx_term_energy_rec(s, kind, t.idx[kind][k], t.type[kind][k], t.par[kind][k], f, e, eel, evdw);
}
