// nonbonded.hip -- LJ + Coulomb pair energies on gfx950.
//
//  * k_full_energy: all pairs i<j with minimum image and no cutoff
//    (computeNonBonded_full, src/FField_MDBAS.cpp:194-265).  Triangular
//    64x64 tiles, one wavefront each (18,528 of them for 12,288 atoms, which
//    balances the 1,024 SIMDs far better than 256-atom tiles: 61 us against
//    78 us); one i atom per lane, the j atom uniform across the
//    wavefront so that its data arrives through scalar loads and LJ is
//    skipped by a scalar branch when eps_j = 0.  Excluded pairs are included
//    here and subtracted by the fused bonded kernel, which keeps this loop
//    free of per-pair branches.
//  * k_switch_energy: the switched sum over the Verlet list
//    (computeNonBonded_switch, :336-424).  A group of 16, 32 or 64 lanes
//    per atom row of the sorted full list (consecutive rows share a
//    wavefront), reading the j > i suffix in steps whose loads are all in
//    flight together, the next step's indices loaded during this step's
//    arithmetic.
//
// Pair arithmetic: the displacement is the wrapping int32 difference of the
// fixed-point fractional coordinates (the exact minimum image, no rint),
// converted to FP64; r^2 in FP64; 1/r from a hardware estimate (FP32 v_rsq
// in the list kernel, FP64 v_rsq in the all-pair tiles) refined by one FP64
// Newton step; the Coulomb sum in FP64.  The LJ part uses the
// FP32 estimate of 1/r and FP32 arithmetic, folded into FP64 every 8 pairs.
// FP64 FMAs issue at the FP32 rate on this GPU, so the Coulomb refinement
// costs a few issue slots and buys the 1e-6 relative total that water's
// 700-fold Coulomb cancellation demands.
#include "terms.h"

#include <cstdlib>

namespace {

constexpr int kTile = kFullTile;

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Block-wide FP64 sum of two values; result valid in thread 0.
template <int NT>
__device__ __forceinline__ void block_sum2(double &a, double &b)
{
    __shared__ double sa[NT / 64], sb[NT / 64];
    a = wave_sum(a);
    b = wave_sum(b);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) { sa[w] = a; sb[w] = b; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double ta = 0.0, tb = 0.0;
        for (int k = 0; k < NT / 64; k++) { ta += sa[k]; tb += sb[k]; }
        a = ta; b = tb;
    }
}

template <int PB>
__device__ __forceinline__ void disp(const DevSys &s, const int4 &pi, const int4 &pj, double xi, double yi,
                                     double zi, int j, double &dx, double &dy, double &dz)
{
    mcx_disp_k<PB>(s, pi, pj, xi, yi, zi, j, dx, dy, dz);
}

// Is the wave-uniform value v non-zero in any bit?  A scalar compare: the
// compiler would otherwise test a uniform float with a vector instruction.
__device__ __forceinline__ bool scalar_nonzero(float v)
{
    int r;
    asm("s_cmp_lg_u32 %1, 0\n\ts_cselect_b32 %0, 1, 0" : "=s"(r) : "s"(v) : "scc");
    return r != 0;
}

// One j atom of a tile against this lane's i atom.  j is uniform across the
// wavefront, so its position and parameters arrive as scalar operands, and
// the test for eps_j == 0 (the water hydrogens) is a scalar bit test: the
// Lennard-Jones part is a scalar branch.
template <int PB, bool DIAG>
__device__ __forceinline__ void tile_pair_acc(const DevSys &s, int i, int j, const int4 pi, const int4 pj,
                                              const TileAtom &qe, double xi, double yi, double zi, float si,
                                              double &accC, double &accV)
{
    double r0;
    const double ri = mcx_tile_ri<PB>(s, pi, pj, xi, yi, zi, j, r0);
    if (DIAG && !(j > i)) return;
    accC = fma(qe.q, ri, accC);
    if (scalar_nonzero(qe.eps)) accV = fma((double)qe.eps, (double)mcx_tile_lj(si + qe.sig, r0), accV);
}

#ifndef MCX_TILE_ILP
#define MCX_TILE_ILP 4
#endif

// The j atoms of a tile in order.  Off the diagonal, K j atoms go together:
// their scalar loads are issued at once, their distance chains and 1/r
// estimates are independent, then each Lennard-Jones branch, then the
// Newton steps, into two accumulator pairs, so that the chains overlap and
// a wavefront waits on the scalar loads once per K atoms.
template <int PB, bool DIAG>
__device__ __forceinline__ void full_tile(const DevSys &s, int i, int j0, int j1, const int4 pi, double xi, double yi,
                                          double zi, float si, double &accC, double &accV)
{
    constexpr int K = MCX_TILE_ILP;
    const int4 *__restrict__ ipos = s.ipos;
    const TileAtom *__restrict__ tq = s.tq;
    int j = j0;
    if (!DIAG && K > 1) {
        double accC2 = 0.0, accV2 = 0.0;
        for (; j + K - 1 < j1; j += K) {
            int4 p[K];
            TileAtom q[K];
#pragma unroll
            for (int k = 0; k < K; k++) { p[k] = ipos[j + k]; q[k] = tq[j + k]; }
            double r2[K], r0[K];
#pragma unroll
            for (int k = 0; k < K; k++) {
                r2[k] = mcx_tile_r2<PB>(s, pi, p[k], xi, yi, zi, j + k);
                r0[k] = __builtin_amdgcn_rsq(r2[k]);
            }
#pragma unroll
            for (int k = 0; k < K; k++)
                if (scalar_nonzero(q[k].eps)) {
                    double &av = (k & 1) ? accV2 : accV;
                    av = fma((double)q[k].eps, (double)mcx_tile_lj(si + q[k].sig, r0[k]), av);
                }
#pragma unroll
            for (int k = 0; k < K; k++) {
                double &ac = (k & 1) ? accC2 : accC;
                ac = fma(q[k].q, mcx_tile_newton(r2[k], r0[k]), ac);
            }
        }
        accC += accC2;
        accV += accV2;
    }
    for (; j < j1; j++)
        tile_pair_acc<PB, DIAG>(s, i, j, pi, ipos[j], tq[j], xi, yi, zi, si, accC, accV);
}

// Tile b of the triangular enumeration that the tile list holds: row I
// has the tiles (I, I), (I, I+1), ..., (I, nt-1) and starts at
// S(I) = I*nt - I*(I-1)/2.
__device__ __forceinline__ int2 tri_tile(int b, int nt)
{
    auto S = [nt](long long I) { return I * nt - I * (I - 1) / 2; };
    const double a = 2.0 * nt + 1.0;
    int I = (int)(0.5 * (a - sqrt(fmax(a * a - 8.0 * b, 0.0))));
    I = I < 0 ? 0 : (I >= nt ? nt - 1 : I);
    while (I > 0 && S(I) > b) I--;
    while (I + 1 < nt && S(I + 1) <= b) I++;
    return make_int2(I, I + (int)(b - S(I)));
}

// LIST: tile b read from `tiles`; otherwise computed (tri_tile over nt
// tile rows), for callers that keep no tile list.
template <int PB, bool LIST = true>
__global__ __launch_bounds__(kTile) void k_full_energy(DevSys s, const int2 *__restrict__ tiles,
                                                        double *__restrict__ partial, int nt = 0)
{
    const int2 tl = LIST ? tiles[blockIdx.x] : tri_tile(blockIdx.x, nt);
    const int i = tl.x * kTile + threadIdx.x;
    const bool valid = i < s.n;
    const int ii = valid ? i : 0;
    const int4 pi = s.ipos[ii];
    const float2 ei = s.esk[ii];
    const double xi = PB == 0 ? s.x[ii] : 0.0, yi = PB == 0 ? s.y[ii] : 0.0, zi = PB == 0 ? s.z[ii] : 0.0;
    const int j0 = tl.y * kTile;
    const int j1 = min(j0 + kTile, s.n);
    double accC = 0.0, accV = 0.0;
    if (tl.x == tl.y) full_tile<PB, true>(s, i, j0, j1, pi, xi, yi, zi, ei.y, accC, accV);
    else full_tile<PB, false>(s, i, j0, j1, pi, xi, yi, zi, ei.y, accC, accV);
    double ec = valid ? (double)__int_as_float(pi.w) * accC * s.iunit : 0.0;
    double ev = valid ? (double)ei.x * accV : 0.0;
    block_sum2<kTile>(ec, ev);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = ec;
        partial[2 * blockIdx.x + 1] = ev;
    }
}

// ---------------------------------------------------------------------------

constexpr int kSwWaves = 4;
constexpr int kSwUnroll = 4;    // entries per lane whose loads are in flight together

// LPR lanes per row, 64/LPR consecutive rows per wavefront.  A row's j > i
// suffix averages ~290 entries on water, so a wavefront per row leaves
// ~40 % of its lanes idle in the last chunk; a group of LPR lanes per row
// keeps them busy (consecutive rows -- the atoms of one or two molecules --
// have suffixes of similar length).  Each lane walks its row with stride
// LPR, kSwUnroll entries at a time: the j atoms' positions and parameters
// are gathered for the step's indices, the next step's indices are loaded,
// then the arithmetic runs while those are in flight.  A lane keeps its own sums
// (q_i and eps_i folded in at the end of the row); one wavefront sum at the
// end.
template <int PB, int LPR>
__global__ __launch_bounds__(64 * kSwWaves) void k_switch_energy(DevSys s, double *__restrict__ partial)
{
    constexpr int RPW = 64 / LPR;                // rows per wavefront
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int sub = lane % LPR;
    double ec = 0.0, ev = 0.0;
    // kernel length unit -> A^2
    const double u2 = PB == 1 ? s.unit[0] * s.unit[0] : 1.0;
    for (int i0 = (blockIdx.x * kSwWaves + w) * RPW; i0 < s.n; i0 += gridDim.x * kSwWaves * RPW) {
        const int i = i0 + lane / LPR;
        const bool live = i < s.n;
        const int ii = live ? i : i0;
        const int4 pi = s.ipos[ii];
        const float2 ei = s.esk[ii];
        const double xi = PB == 0 ? s.x[ii] : 0.0, yi = PB == 0 ? s.y[ii] : 0.0, zi = PB == 0 ? s.z[ii] : 0.0;
        const int *__restrict__ row = s.nbr + (size_t)ii * s.cap;
        const int k1 = live ? s.cnt[ii] : 0;
        double c = 0.0, v = 0.0;
        int kb = (live ? s.half[ii] : 0) + sub;
        int j[kSwUnroll];
#pragma unroll
        for (int u = 0; u < kSwUnroll; u++) {
            const int k = kb + LPR * u;
            j[u] = k < k1 ? row[k] : -1;
        }
        while (__ballot(kb < k1) != 0ull) {
            int4 pj[kSwUnroll];
            float2 ej[kSwUnroll];
#pragma unroll
            for (int u = 0; u < kSwUnroll; u++) {
                const int jj = j[u] < 0 ? ii : j[u];   // a valid address; the pair is not summed
                pj[u] = s.ipos[jj];
                ej[u] = s.esk[jj];
            }
            // the next step's indices, in flight during this step's arithmetic
            int jn[kSwUnroll];
            kb += LPR * kSwUnroll;
#pragma unroll
            for (int u = 0; u < kSwUnroll; u++) {
                const int k = kb + LPR * u;
                jn[u] = k < k1 ? row[k] : -1;
            }
#pragma unroll
            for (int u = 0; u < kSwUnroll; u++) {
                if (j[u] < 0) continue;
                double dx, dy, dz;
                disp<PB>(s, pi, pj[u], xi, yi, zi, j[u], dx, dy, dz);
                const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
                const double r2A = r2 * u2;
                if (r2A <= s.ctoff2) {
                    float ri32;
                    const double ri = mcx_rsqrt(r2, ri32);
                    const float uu = (ei.y + ej[u].y) * ri32;
                    const float t = uu * uu;
                    const float t3 = t * t * t;
                    const double vj = (double)(ej[u].x * fmaf(t3, t3, -t3));
                    double sw = 1.0;
                    if (r2A > s.cton2) {
                        const double a = s.ctoff2 - r2A;
                        sw = a * a * (s.ctoff2 + 2.0 * r2A - 3.0 * s.cton2) * s.switch2;
                    }
                    c = fma(sw * (double)__int_as_float(pj[u].w), ri, c);
                    v = fma(sw, vj, v);
                }
            }
#pragma unroll
            for (int u = 0; u < kSwUnroll; u++) j[u] = jn[u];
        }
        ec = fma((double)__int_as_float(pi.w), c, ec);
        ev = fma((double)ei.x, v, ev);
    }
    ec = wave_sum(ec) * s.iunit;
    ev = wave_sum(ev);
    // lanes hold identical wave sums; keep lane 0's copy only
    if (lane != 0) { ec = 0.0; ev = 0.0; }
    block_sum2<64 * kSwWaves>(ec, ev);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = ec;
        partial[2 * blockIdx.x + 1] = ev;
    }
}

// thread t sums blocks t, t+256, ... of component c in order, then a
// fixed-shape tree: the result does not depend on scheduling.
__device__ __forceinline__ void reduce_partials(const double *__restrict__ partial, int nblocks, int ncomp,
                                                double *__restrict__ out)
{
    __shared__ double sm[256];
    for (int c = 0; c < ncomp; c++) {
        double v = 0.0;
        for (int b = threadIdx.x; b < nblocks; b += 256) v += partial[(size_t)b * ncomp + c];
        sm[threadIdx.x] = v;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if (threadIdx.x < o) sm[threadIdx.x] += sm[threadIdx.x + o];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[c] = sm[0];
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_reduce(const double *__restrict__ partial, int nblocks, int ncomp,
                                                 double *__restrict__ out)
{
    reduce_partials(partial, nblocks, ncomp, out);
}

// two independent sums in one launch, one block each
__global__ __launch_bounds__(256) void k_reduce2(const double *__restrict__ pa, int na, int ca, double *__restrict__ oa,
                                                  const double *__restrict__ pb, int nb, int cb, double *__restrict__ ob)
{
    if (blockIdx.x == 0) reduce_partials(pa, na, ca, oa);
    else reduce_partials(pb, nb, cb, ob);
}

}  // namespace

void launch_full_energy(const DevSys &s, const int2 *tiles, int ntiles, int tile, double *partial, hipStream_t st)
{
    (void)tile;
    const int pb = s.pbtype == 0 ? 0 : (s.pbtype == 1 ? 1 : 2);
    if (pb == 1) hipLaunchKernelGGL(k_full_energy<1>, dim3(ntiles), dim3(kTile), 0, st, s, tiles, partial);
    else if (pb == 2) hipLaunchKernelGGL(k_full_energy<2>, dim3(ntiles), dim3(kTile), 0, st, s, tiles, partial);
    else hipLaunchKernelGGL(k_full_energy<0>, dim3(ntiles), dim3(kTile), 0, st, s, tiles, partial);
}

void launch_full_energy_listless(const DevSys &s, double *partial, hipStream_t st)
{
    const int nt = (s.n + kTile - 1) / kTile;
    const int ntiles = nt * (nt + 1) / 2;
    const int pb = s.pbtype == 0 ? 0 : (s.pbtype == 1 ? 1 : 2);
    if (pb == 1)
        hipLaunchKernelGGL((k_full_energy<1, false>), dim3(ntiles), dim3(kTile), 0, st, s, nullptr, partial, nt);
    else if (pb == 2)
        hipLaunchKernelGGL((k_full_energy<2, false>), dim3(ntiles), dim3(kTile), 0, st, s, nullptr, partial, nt);
    else
        hipLaunchKernelGGL((k_full_energy<0, false>), dim3(ntiles), dim3(kTile), 0, st, s, nullptr, partial, nt);
}

// Lanes per row of the switched kernel: the fewest of 16, 32, 64 that still
// give at least 24 wavefronts per CU (6,144), so that enough of them are in
// flight to cover the load latency (MCX_SW_LPR = 8, 16, 32 or 64 forces one).
static int sw_lpr(int n)
{
    static const int forced = [] {
        const char *e = std::getenv("MCX_SW_LPR");
        const int k = e ? std::atoi(e) : 0;
        return (k == 8 || k == 16 || k == 32 || k == 64) ? k : 0;
    }();
    if (forced) return forced;
    for (int k = 16; k < 64; k *= 2)
        if ((long long)n * k / 64 >= 6144) return k;
    return 64;
}

int switch_energy_blocks(int n)
{
    const int rows = kSwWaves * (64 / sw_lpr(n));
    int b = (n + rows - 1) / rows;
    return b < 1 ? 1 : b;
}

template <int LPR>
static void launch_switch_lpr(const DevSys &s, double *partial, int nblocks, hipStream_t st)
{
    const int pb = s.pbtype == 0 ? 0 : (s.pbtype == 1 ? 1 : 2);
    const dim3 grid(nblocks), block(64 * kSwWaves);
    if (pb == 1) hipLaunchKernelGGL((k_switch_energy<1, LPR>), grid, block, 0, st, s, partial);
    else if (pb == 2) hipLaunchKernelGGL((k_switch_energy<2, LPR>), grid, block, 0, st, s, partial);
    else hipLaunchKernelGGL((k_switch_energy<0, LPR>), grid, block, 0, st, s, partial);
}

void launch_switch_energy(const DevSys &s, double *partial, int nblocks, hipStream_t st)
{
    switch (sw_lpr(s.n)) {
    case 8: launch_switch_lpr<8>(s, partial, nblocks, st); break;
    case 32: launch_switch_lpr<32>(s, partial, nblocks, st); break;
    case 64: launch_switch_lpr<64>(s, partial, nblocks, st); break;
    default: launch_switch_lpr<16>(s, partial, nblocks, st); break;
    }
}

void launch_reduce(const double *partial, int nblocks, int ncomp, double *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(256), 0, st, partial, nblocks, ncomp, out);
}

void launch_reduce2(const double *pa, int na, int ca, double *oa, const double *pb, int nb, int cb, double *ob,
                    hipStream_t st)
{
    hipLaunchKernelGGL(k_reduce2, dim3(2), dim3(256), 0, st, pa, na, ca, oa, pb, nb, cb, ob);
}
