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
    g.cpos[p] = make_double4(s.x[a], s.y[a], s.z[a], (double)(2 * a + (s.frozen[a] ? 1 : 0)));
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

constexpr int kMaxRuns = 64;   // 2 runs for each of up to 25 (y, z) cell columns, rounded up
constexpr int kUnroll = 4;

template <bool PBC>
__global__ __launch_bounds__(256) void k_build_rows(DevSys s, CellGrid g, int words, int wpb)
{
    extern __shared__ uint32_t bitmap[];
    __shared__ int run_start[4][kMaxRuns], run_len[4][kMaxRuns];
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int i = blockIdx.x * wpb + w;
    const bool active = i < s.n;
    uint32_t *B = bitmap + (size_t)w * words;
    for (int k = lane; k < words; k += 64) B[k] = 0u;

    // phase 1: the candidate runs of row i.  One lane per (y, z) column of
    // neighbour cells; the x cells of a column are contiguous in cell
    // order, except where the range wraps (then two runs).
    double xi = 0.0, yi = 0.0, zi = 0.0;
    bool fi = false;
    int ncol = 0;
    if (active) {
        xi = s.x[i]; yi = s.y[i]; zi = s.z[i];
        fi = s.frozen[i] != 0;
        const int ci = g.cell_of[i];
        const int cc[3] = {ci % g.nc[0], (ci / g.nc[0]) % g.nc[1], ci / (g.nc[0] * g.nc[1])};
        int lo[3], cnt[3];
        for (int d = 0; d < 3; d++) {
            if (2 * g.reach[d] + 1 >= g.nc[d]) { lo[d] = 0; cnt[d] = g.nc[d]; }
            else { lo[d] = cc[d] - g.reach[d]; cnt[d] = 2 * g.reach[d] + 1; }
        }
        ncol = cnt[1] * cnt[2];
        for (int col = lane; col < kMaxRuns / 2; col += 64) {
            int s0 = 0, l0 = 0, s1 = 0, l1 = 0;
            if (col < ncol) {
                const int wy = (lo[1] + col % cnt[1] + g.nc[1]) % g.nc[1];
                const int wz = (lo[2] + col / cnt[1] + g.nc[2]) % g.nc[2];
                const int base = (wz * g.nc[1] + wy) * g.nc[0];
                int x0 = lo[0], x1 = lo[0] + cnt[0] - 1, a0 = 0, b0 = -1;
                if (x0 < 0) { a0 = x0 + g.nc[0]; b0 = g.nc[0] - 1; x0 = 0; }
                else if (x1 >= g.nc[0]) { a0 = 0; b0 = x1 - g.nc[0]; x1 = g.nc[0] - 1; }
                s0 = g.cell_start[base + x0];
                l0 = g.cell_start[base + x1] + g.cell_count[base + x1] - s0;
                if (b0 >= a0) {
                    s1 = g.cell_start[base + a0];
                    l1 = g.cell_start[base + b0] + g.cell_count[base + b0] - s1;
                }
            }
            run_start[w][2 * col] = s0; run_len[w][2 * col] = l0;
            run_start[w][2 * col + 1] = s1; run_len[w][2 * col + 1] = l1;
        }
    }
    __syncthreads();

    // phase 2: stream the runs, kUnroll loads in flight per lane; every hit
    // sets bit j of the row bitmap
    if (active) {
        for (int r = 0; r < 2 * ncol; r++) {
            const int rs = run_start[w][r], rl = run_len[w][r];
            for (int off = 0; off < rl; off += 64 * kUnroll) {
                double4 pj[kUnroll];
#pragma unroll
                for (int u = 0; u < kUnroll; u++) {
                    const int t = off + u * 64 + lane;
                    pj[u] = t < rl ? g.cpos[rs + t] : make_double4(0.0, 0.0, 0.0, -1.0);
                }
#pragma unroll
                for (int u = 0; u < kUnroll; u++) {
                    const int code = (int)pj[u].w;
                    const int j = code >> 1;
                    if (code >= 0 && j != i && !(fi && (code & 1)) && within<PBC>(s, pj[u], xi, yi, zi))
                        atomicOr(&B[j >> 5], 1u << (j & 31));
                }
            }
        }
    }
    __syncthreads();
    // phase 3: exclusions only remove pairs -- clear the bits of i's
    // excluded partners
    if (active) {
        for (int k = s.xfull_off[i] + lane; k < s.xfull_off[i + 1]; k += 64) {
            const int e = s.xfull_idx[k];
            atomicAnd(&B[e >> 5], ~(1u << (e & 31)));
        }
    }
    __syncthreads();
    if (!active) return;

    // phase 4: ordered compaction of the bitmap (no barriers below)
    int *row = s.nbr + (size_t)i * s.cap;
    int total = 0, nlow = 0;
    const int iw = i >> 5;
    for (int base = 0; base < words; base += 64) {
        const int k = base + lane;
        uint32_t wd = k < words ? B[k] : 0u;
        const int pc = __popc(wd);
        int incl = pc;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            int t = __shfl_up(incl, o, 64);
            if (lane >= o) incl += t;
        }
        const int tot = __shfl(incl, 63, 64);
        int pos = total + incl - pc;
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
    while (wpb > 1 && per_wave * wpb > 60 * 1024) wpb >>= 1;
    const size_t lds = per_wave * wpb;
    const int blocks = (s.n + wpb - 1) / wpb;
    if (s.pbtype == 0)
        hipLaunchKernelGGL(k_build_rows<false>, dim3(blocks), dim3(64 * wpb), lds, st, s, g, words, wpb);
    else
        hipLaunchKernelGGL(k_build_rows<true>, dim3(blocks), dim3(64 * wpb), lds, st, s, g, words, wpb);
}
