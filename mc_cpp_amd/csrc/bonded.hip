// bonded.hip -- one fused launch for every non-list term of getE
// (src/FField_MDBAS.cpp:108-192): bonds, Urey-Bradley, angles, dihedrals,
// impropers (:490-786), the 1-4 pairs (:267-318 / :426-488) and, in FULL
// mode, the excluded pairs that the all-pairs tile kernel included and that
// are subtracted here.  A thread evaluates one term of a term-type-tagged
// work list; per-component FP64 block sums go to `partial` (7 per block:
// elec_raw, vdw_raw, bond, ang, ub, dihe, impr).
#include "terms.h"

namespace {

constexpr int kBT = 256;

__device__ __forceinline__ double wsum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__global__ __launch_bounds__(kBT) void k_bonded(DevSys s, DevTerms t, double *__restrict__ partial)
{
    __shared__ double sm[7][kBT / 64];
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    const CoordFetch f{s.x, s.y, s.z, nullptr, nullptr, nullptr, 0, -1};
    const int total = t.start[kNumKinds];
    for (int g = blockIdx.x * kBT + threadIdx.x; g < total; g += gridDim.x * kBT) {
        int kind = 0;
        while (g >= t.start[kind + 1]) kind++;
        const int k = g - t.start[kind];
        double e, eel, ev;
        mcx_term_energy(s, t, kind, k, f, e, eel, ev);
        acc[0] += eel;
        acc[1] += ev;
        // component slots: bond 2, ang 3, ub 4, dihe 5, impr 6
        const int slot = kind == kBond ? 2 : kind == kAngle ? 3 : kind == kUB ? 4 : kind == kDihe ? 5 : 6;
        if (kind <= kImpr) acc[slot] += e;
    }
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
    for (int c = 0; c < 7; c++) {
        const double v = wsum(acc[c]);
        if (l == 0) sm[c][w] = v;
    }
    __syncthreads();
    if (threadIdx.x < 7) {
        double v = 0.0;
        for (int k = 0; k < kBT / 64; k++) v += sm[threadIdx.x][k];
        partial[blockIdx.x * 7 + threadIdx.x] = v;
    }
}

}  // namespace

int bonded_blocks(const DevTerms &t)
{
    int b = (t.start[kNumKinds] + kBT - 1) / kBT;
    if (b < 1) b = 1;
    if (b > 1024) b = 1024;
    return b;
}

void launch_bonded(const DevSys &s, const DevTerms &t, double *partial, int nblocks, hipStream_t st)
{
    hipLaunchKernelGGL(k_bonded, dim3(nblocks), dim3(kBT), 0, st, s, t, partial);
}
