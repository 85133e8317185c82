// listcore.h -- device helpers shared by the full list build (list.hip) and
// the incremental update (list_patch.hip): the reference's pair test and
// minimum image, cell indexing and a wavefront prefix scan.
#pragma once

#include "mcx_internal.h"

#pragma clang fp contract(off)

namespace {

__device__ __forceinline__ double ref_rint(double v)
{
    return (double)(v >= 0. ? (int)(v + 0.5) : (int)(v - 0.5));
}

// wrapping int32 difference b - a: the minimum image in units of L/2^32
__device__ __forceinline__ int wdiff(int b, int a) { return (int)((unsigned)b - (unsigned)a); }

__device__ __forceinline__ int cell_index(double x, double rL, int nc)
{
    double u = x * rL;
    u = u - floor(u);
    int c = (int)(u * (double)nc);
    return c < 0 ? 0 : (c >= nc ? nc - 1 : c);
}

// Tools::distance2 + the r2 <= cutnb2 test, operation by operation
template <bool PBC>
__device__ __forceinline__ bool within(const DevSys &s, double4 pj, double xi, double yi, double zi)
{
    double dx = pj.x - xi;
    double dy = pj.y - yi;
    double dz = pj.z - zi;
    if (PBC) {
        dx -= s.L[0] * ref_rint(s.rL[0] * dx);
        dy -= s.L[1] * ref_rint(s.rL[1] * dy);
        dz -= s.L[2] * ref_rint(s.rL[2] * dz);
    }
    const double r2 = dx * dx + dy * dy + dz * dz;
    return r2 <= s.cutnb2;
}

// Inclusive prefix sum over the 64 lanes of a wavefront: row_shr 1, 2, 4, 8
// within rows of 16 lanes, then row_bcast 15 and 31 carry the row totals.
__device__ __forceinline__ int wave_incl_scan(int v)
{
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);
    return v;
}

}  // namespace
