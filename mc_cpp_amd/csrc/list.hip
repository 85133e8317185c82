// list.hip -- Verlet pair-list build on gfx950
// (List_nonBonded::update_verlet_list, src/List_nonBonded.cpp:836-915).
//
// Pair set: {(i,j): r2(i,j) <= (ctoff+dcut)^2, not (frozen_i and frozen_j),
// max(i,j) not in the visible exclusion row of min(i,j)}, with r2 formed
// exactly as Tools::distance2 does (include/Tools.hpp:151-170): FP64
// difference, PerConditions minimum image with the truncating rint
// (src/PerConditions.cpp:244-265, include/Tools.hpp:221-224), squares
// summed left to right, no contraction.  The result is the reference's set
// bit for bit.
//
// Layout: atoms are binned into a periodic cell grid (FP64, cells at least
// (ctoff+dcut)/reach wide); then one wavefront per atom i tests the atoms of
// the surrounding cells and sets bit j of an LDS bitmap for every hit.
// Scanning the bitmap in word order is the stream compaction: it writes the
// row's j in ascending order into an ELL row of capacity `cap` (full list,
// both directions; the j > i suffix is the reference's half row).
#include "mcx_internal.h"

#pragma clang fp contract(off)

namespace {

__device__ __forceinline__ double ref_rint(double v)
{
    return (double)(v >= 0. ? (int)(v + 0.5) : (int)(v - 0.5));
}

__global__ void k_zero_int(int *p, int n)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) p[k] = 0;
}

__device__ __forceinline__ int cell_index(double x, double rL, int nc)
{
    double u = x * rL;
    u = u - floor(u);
    int c = (int)(u * (double)nc);
    return c < 0 ? 0 : (c >= nc ? nc - 1 : c);
}

__global__ void k_cell_count(DevSys s, CellGrid g)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= s.n) return;
    int c = 0;
    if (g.ncell > 1) {
        const int cx = cell_index(s.x[a], s.rL[0], g.nc[0]);
        const int cy = cell_index(s.y[a], s.rL[1], g.nc[1]);
        const int cz = cell_index(s.z[a], s.rL[2], g.nc[2]);
        c = (cz * g.nc[1] + cy) * g.nc[0] + cx;
    }
    g.cell_of[a] = c;
    atomicAdd(&g.cell_count[c], 1);
}

// exclusive scan of cell counts, one block
__global__ __launch_bounds__(1024) void k_cell_scan(CellGrid g)
{
    __shared__ int sm[1024];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < g.ncell; base += 1024) {
        const int c = base + threadIdx.x;
        const int v = c < g.ncell ? g.cell_count[c] : 0;
        sm[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {
            int t = threadIdx.x >= o ? sm[threadIdx.x - o] : 0;
            __syncthreads();
            sm[threadIdx.x] += t;
            __syncthreads();
        }
        if (c < g.ncell) g.cell_start[c] = carry + sm[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry += sm[1023];
        __syncthreads();
    }
}

__global__ void k_cell_fill(DevSys s, CellGrid g)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= s.n) return;
    const int c = g.cell_of[a];
    const int p = g.cell_start[c] + atomicAdd(&g.cursor[c], 1);
    g.cell_atoms[p] = a;
}

template <bool PBC>
__device__ __forceinline__ bool pair_in_list(const DevSys &s, int i, int j, double xi, double yi, double zi)
{
    double dx = s.x[j] - xi;
    double dy = s.y[j] - yi;
    double dz = s.z[j] - zi;
    if (PBC) {
        dx -= s.L[0] * ref_rint(s.rL[0] * dx);
        dy -= s.L[1] * ref_rint(s.rL[1] * dy);
        dz -= s.L[2] * ref_rint(s.rL[2] * dz);
    }
    const double r2 = dx * dx + dy * dy + dz * dz;
    if (!(r2 <= s.cutnb2)) return false;
    if (s.frozen[i] && s.frozen[j]) return false;
    const int lo = i < j ? i : j, hi = i < j ? j : i;
    for (int k = s.excl_off[lo]; k < s.excl_off[lo + 1]; k++)
        if (s.excl_idx[k] == hi) return false;
    return true;
}

template <bool PBC>
__global__ __launch_bounds__(256) void k_build_rows(DevSys s, CellGrid g, int words, int wpb)
{
    extern __shared__ uint32_t bitmap[];
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int i = blockIdx.x * wpb + w;
    uint32_t *B = bitmap + (size_t)w * words;
    for (int k = lane; k < words; k += 64) B[k] = 0u;
    __syncthreads();
    if (i < s.n) {

    const double xi = s.x[i], yi = s.y[i], zi = s.z[i];
    const int ci = g.cell_of[i];
    const int cx = ci % g.nc[0], cy = (ci / g.nc[0]) % g.nc[1], cz = ci / (g.nc[0] * g.nc[1]);
    // neighbour cell ranges; a dimension with fewer than 2*reach+1 cells is
    // searched whole so that no cell is visited twice
    int lo[3], hi[3];
    const int cc[3] = {cx, cy, cz};
    for (int d = 0; d < 3; d++) {
        if (2 * g.reach[d] + 1 >= g.nc[d]) { lo[d] = 0; hi[d] = g.nc[d] - 1; }
        else { lo[d] = cc[d] - g.reach[d]; hi[d] = cc[d] + g.reach[d]; }
    }
    for (int zc = lo[2]; zc <= hi[2]; zc++) {
        const int wz = (zc + g.nc[2]) % g.nc[2];
        for (int yc = lo[1]; yc <= hi[1]; yc++) {
            const int wy = (yc + g.nc[1]) % g.nc[1];
            for (int xc = lo[0]; xc <= hi[0]; xc++) {
                const int wx = (xc + g.nc[0]) % g.nc[0];
                const int c = (wz * g.nc[1] + wy) * g.nc[0] + wx;
                const int a0 = g.cell_start[c], a1 = a0 + g.cell_count[c];
                for (int a = a0 + lane; a < a1; a += 64) {
                    const int j = g.cell_atoms[a];
                    if (j != i && pair_in_list<PBC>(s, i, j, xi, yi, zi))
                        atomicOr(&B[j >> 5], 1u << (j & 31));
                }
            }
        }
    }
    }
    __syncthreads();
    if (i >= s.n) return;

    // ordered compaction of the bitmap
    int *row = s.nbr + (size_t)i * s.cap;
    int total = 0, nlow = 0;
    const int iw = i >> 5;
    for (int base = 0; base < words; base += 64) {
        const int k = base + lane;
        uint32_t wd = k < words ? B[k] : 0u;
        const int pc = __popc(wd);
        // inclusive wave prefix sum
        int incl = pc;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            int t = __shfl_up(incl, o, 64);
            if (lane >= o) incl += t;
        }
        const int tot = __shfl(incl, 63, 64);
        int pos = total + incl - pc;
        // atoms below i: whole words below i's word, plus the low bits of it
        int lowc = 0;
        if (k < iw) lowc = pc;
        else if (k == iw) lowc = __popc(wd & ((1u << (i & 31)) - 1u));
        while (wd) {
            const int b = __ffs(wd) - 1;
            wd &= wd - 1u;
            if (pos < s.cap) row[pos] = k * 32 + b;
            pos++;
        }
        for (int o = 32; o > 0; o >>= 1) lowc += __shfl_xor(lowc, o, 64);
        nlow += lowc;
        total += tot;
    }
    if (lane == 0) {
        s.cnt[i] = total < s.cap ? total : s.cap;
        s.half[i] = nlow < s.cap ? nlow : s.cap;
        if (total > s.cap) atomicMax(s.overflow, total);
    }
}

}  // namespace

size_t list_build_lds_bytes(int n)
{
    const int words = (n + 31) / 32;
    return (size_t)words * 4;
}

void launch_list_build(const DevSys &s, const CellGrid &g, hipStream_t st)
{
    const int nb = (s.n + 255) / 256;
    hipLaunchKernelGGL(k_zero_int, dim3(64), dim3(256), 0, st, g.cell_count, g.ncell);
    hipLaunchKernelGGL(k_zero_int, dim3(64), dim3(256), 0, st, g.cursor, g.ncell);
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, st, s.overflow, 1);
    hipLaunchKernelGGL(k_cell_count, dim3(nb), dim3(256), 0, st, s, g);
    hipLaunchKernelGGL(k_cell_scan, dim3(1), dim3(1024), 0, st, g);
    hipLaunchKernelGGL(k_cell_fill, dim3(nb), dim3(256), 0, st, s, g);
    const int words = (s.n + 31) / 32;
    const size_t per_wave = (size_t)words * 4;
    int wpb = 4;
    while (wpb > 1 && per_wave * wpb > 64 * 1024) wpb >>= 1;
    const size_t lds = per_wave * wpb;
    const int blocks = (s.n + wpb - 1) / wpb;
    if (s.pbtype == 0)
        hipLaunchKernelGGL(k_build_rows<false>, dim3(blocks), dim3(64 * wpb), lds, st, s, g, words, wpb);
    else
        hipLaunchKernelGGL(k_build_rows<true>, dim3(blocks), dim3(64 * wpb), lds, st, s, g, words, wpb);
}
