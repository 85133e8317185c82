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
// Binning: atoms go into a periodic cell grid (cells at least
// (ctoff+dcut)/reach wide): blocks of 1,024 atoms count into LDS histograms
// added to global accumulators, the last block turns the totals into cell
// and row-group starts and a table of row groups (grids of at most
// kBinCells cells; larger ones count with global atomics over more
// launches); the atoms are then placed, and each cell's atoms put in
// ascending order, which makes the list rows (and every sum over them) the
// same from build to build.
//
// Rows: a workgroup of four wavefronts takes a group of up to G atoms of
// one cell and streams the surrounding cells' atoms once for all G rows
// (cell-ordered records with 2*atom + frozen).  In a periodic box each
// candidate is first tested with an FP32 r^2 from the fixed-point mirror,
// taken about the cell centre when the box is large enough for that to be
// every row atom's minimum image; a candidate with any FP32 r^2 within
// 1e-5 (relative) of (ctoff+dcut)^2 -- the FP32 chain's own error is below
// 1e-6 -- is decided by the FP64 test on the FP64 coordinates.  Without
// periodic boundaries every candidate takes the FP64 test.  Every hit sets
// bit j of the row's LDS bitmap; scanning the bitmap in word order is the
// stream compaction: it writes the row's j in ascending order into an ELL
// row of capacity `cap` (full list, both directions; the j > i suffix from
// half[i] on is the reference's half row).  Each wavefront compacts every
// fourth row of the group.
#include "listcore.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>

#pragma clang fp contract(off)

namespace {

constexpr int kBinCells = 8192;    // LDS counters of the one-launch binning
constexpr int kMaxRuns = 64;       // 2 runs for each of up to 25 (y, z) cell columns, rounded up
constexpr int kUnroll = 4;
constexpr int kRowWaves = 4;       // wavefronts per row-group workgroup
constexpr double kBand = 1e-5;     // relative band around (ctoff+dcut)^2 decided in FP64

__global__ void k_zero_int(int *p, int n)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) p[k] = 0;
}

__device__ __forceinline__ int cell_of_atom(const DevSys &s, const CellGrid &g, int a)
{
    if (g.ncell == 1) return 0;
    const int cx = cell_index(s.x[a], s.rL[0], g.nc[0]);
    const int cy = cell_index(s.y[a], s.rL[1], g.nc[1]);
    const int cz = cell_index(s.z[a], s.rL[2], g.nc[2]);
    return (cz * g.nc[1] + cy) * g.nc[0] + cx;
}

// the cell-ordered record of atom a at position p
__device__ __forceinline__ void store_record(const DevSys &s, const CellGrid &g, int a, int p)
{
    g.cell_atoms[p] = a;
    const int code = 2 * a + (s.frozen[a] ? 1 : 0);
    if (s.pbtype != 0) {
        const int4 q = s.ipos[a];
        g.cfix[p] = make_int4(q.x, q.y, q.z, code);
    } else {
        g.cpos[p] = make_double4(s.x[a], s.y[a], s.z[a], (double)code);
    }
}

// Block-wide exclusive scan of f(in[c]) for c in [0, n) into out (and into
// out2 when given; out2 may alias `in`: chunk c is read before it is
// written), NT threads; returns the total.  sm: NT ints of LDS.
template <int NT, class F>
__device__ int block_scan(const int *in, int *out, int n, int *sm, int *carry, F f, int *out2 = nullptr)
{
    if (threadIdx.x == 0) *carry = 0;
    __syncthreads();
    for (int base = 0; base < n; base += NT) {
        const int c = base + threadIdx.x;
        const int v = c < n ? f(in[c]) : 0;
        sm[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < NT; o <<= 1) {
            const int t = threadIdx.x >= o ? sm[threadIdx.x - o] : 0;
            __syncthreads();
            sm[threadIdx.x] += t;
            __syncthreads();
        }
        if (c < n) {
            out[c] = *carry + sm[threadIdx.x] - v;
            if (out2) out2[c] = *carry + sm[threadIdx.x] - v;
        }
        __syncthreads();
        if (threadIdx.x == NT - 1) *carry += sm[NT - 1];
        __syncthreads();
    }
    const int total = *carry;
    __syncthreads();      // every thread has read the total before the next reset
    return total;
}

struct Identity {
    __device__ int operator()(int v) const { return v; }
};
struct Groups {
    int g;
    __device__ int operator()(int v) const { return (v + g - 1) / g; }
};

// The row-group table: group k of cell c (ceil(count / group) per cell)
// is (c, first record, atoms, 0), so that a row-group workgroup finds its
// atoms with one load.  Threads of a block share out the cells.
__device__ void write_groups(const CellGrid &g, int group, int nthreads)
{
    for (int c = threadIdx.x; c < g.ncell; c += nthreads) {
        const int cnt = g.cell_count[c], gs = g.group_start[c], cs = g.cell_start[c];
        for (int k = 0; k * group < cnt; k++)
            g.group_info[gs + k] = make_int4(c, cs + k * group, min(group, cnt - k * group), 0);
    }
}

// Binning (ncell <= kBinCells), pass 1: each block bins kBinAtoms atoms
// (kBinAtoms / kBinT per thread, their loads in flight together) into an
// LDS histogram and adds it to the global accumulators g.cell_acc, which
// are zero between builds.  The last block to finish takes the totals
// (resetting the accumulators), writes cell counts, cell starts, row-group
// starts and the group table, clears the fill cursors and the overflow
// flag, and re-arms the ticket.
constexpr int kBinT = 256;
constexpr int kBinPer = 4;
constexpr int kBinAtoms = kBinT * kBinPer;

__global__ __launch_bounds__(kBinT) void k_bin_count(DevSys s, CellGrid g, int group)
{
    __shared__ int hist[kBinCells];
    __shared__ int sm[kBinT];
    __shared__ int carry;
    __shared__ bool last;
    const int nc = g.ncell;
    for (int c = threadIdx.x; c < nc; c += kBinT) hist[c] = 0;
    __syncthreads();
    const int a0 = blockIdx.x * kBinAtoms + threadIdx.x;
    int cell[kBinPer];
#pragma unroll
    for (int u = 0; u < kBinPer; u++) {
        const int a = a0 + u * kBinT;
        cell[u] = a < s.n ? cell_of_atom(s, g, a) : -1;
    }
#pragma unroll
    for (int u = 0; u < kBinPer; u++) {
        const int a = a0 + u * kBinT;
        if (cell[u] < 0) continue;
        g.cell_of[a] = cell[u];
        g.ipos_ref[a] = s.ipos[a];
        atomicAdd(&hist[cell[u]], 1);
    }
    __syncthreads();
    for (int c = threadIdx.x; c < nc; c += kBinT)
        if (hist[c]) atomicAdd(&g.cell_acc[c], hist[c]);
    // every wavefront's additions have completed at the barrier; one
    // device-scope fence per block orders them before its ticket
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        last = atomicAdd(g.ticket, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    // the totals, read and reset by atomic exchange (device scope: every
    // block's additions have landed before its ticket)
    for (int c = threadIdx.x; c < nc; c += kBinT) {
        const int v = atomicExch(&g.cell_acc[c], 0);
        g.cell_count[c] = v;
        hist[c] = v;
        g.cursor[c] = 0;
    }
    if (threadIdx.x == 0) s.overflow[0] = 0;
    __syncthreads();
    const int ng = block_scan<kBinT>(hist, g.group_start, nc, sm, &carry, Groups{group});
    if (threadIdx.x == 0) g.group_start[nc] = ng;
    block_scan<kBinT>(hist, g.cell_start, nc, sm, &carry, Identity{});
    __syncthreads();
    write_groups(g, group, kBinT);
    if (threadIdx.x == 0) *g.ticket = 0u;
}

// Pass 2: atoms into their cells' ranges (unordered; k_cell_sort orders
// them).  One cell: the records in atom order, no sort.
__global__ __launch_bounds__(kBinT) void k_bin_fill(DevSys s, CellGrid g)
{
    const int a0 = blockIdx.x * kBinAtoms + threadIdx.x;
    if (g.ncell == 1) {
        for (int u = 0; u < kBinPer; u++)
            if (a0 + u * kBinT < s.n) store_record(s, g, a0 + u * kBinT, a0 + u * kBinT);
        return;
    }
    int cell[kBinPer];
#pragma unroll
    for (int u = 0; u < kBinPer; u++) {
        const int a = a0 + u * kBinT;
        cell[u] = a < s.n ? g.cell_of[a] : -1;
    }
#pragma unroll
    for (int u = 0; u < kBinPer; u++) {
        if (cell[u] < 0) continue;
        const int c = cell[u];
        g.cell_tmp[g.cell_start[c] + atomicAdd(&g.cursor[c], 1)] = a0 + u * kBinT;
    }
}

// Binning for larger grids: global counters over several launches.
__global__ void k_cell_count(DevSys s, CellGrid g)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= s.n) return;
    const int c = cell_of_atom(s, g, a);
    g.cell_of[a] = c;
    g.ipos_ref[a] = s.ipos[a];
    atomicAdd(&g.cell_count[c], 1);
}

__global__ __launch_bounds__(1024) void k_cell_scan(CellGrid g, int group)
{
    __shared__ int sm[1024];
    __shared__ int carry;
    block_scan<1024>(g.cell_count, g.cell_start, g.ncell, sm, &carry, Identity{});
    const int ng = block_scan<1024>(g.cell_count, g.group_start, g.ncell, sm, &carry, Groups{group});
    if (threadIdx.x == 0) g.group_start[g.ncell] = ng;
    __syncthreads();
    write_groups(g, group, 1024);
}

__global__ void k_cell_fill(DevSys s, CellGrid g)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= s.n) return;
    const int c = g.cell_of[a];
    g.cell_tmp[g.cell_start[c] + atomicAdd(&g.cursor[c], 1)] = a;
}

// Ascending atom order within each cell (one block per cell, rank sort:
// atom indices are distinct), so that the cell records, the row groups
// and the order of every list row do not depend on the order in which the
// binning atomics landed.
__global__ __launch_bounds__(256) void k_cell_sort(DevSys s, CellGrid g)
{
    const int c = blockIdx.x;
    const int st = g.cell_start[c], cnt = g.cell_count[c];
    const int *src = g.cell_tmp + st;
    for (int e = threadIdx.x; e < cnt; e += 256) {
        const int a = src[e];
        int rank = 0;
        for (int k = 0; k < cnt; k++) rank += src[k] < a;
        store_record(s, g, a, st + rank);
    }
}

// Candidate runs of the cells around cell ci: thread t < 32 takes (y, z)
// column t of neighbour cells; the x cells of a column are contiguous in
// cell order, except where the range wraps (then two runs).  Returns the
// number of columns (2 runs each); the runs are visible after the caller's
// barrier.
__device__ __forceinline__ int cell_runs(const CellGrid &g, int ci, int t, int *run_start, int *run_len)
{
    const int cc[3] = {ci % g.nc[0], (ci / g.nc[0]) % g.nc[1], ci / (g.nc[0] * g.nc[1])};
    int lo[3], cnt[3];
    for (int d = 0; d < 3; d++) {
        if (2 * g.reach[d] + 1 >= g.nc[d]) { lo[d] = 0; cnt[d] = g.nc[d]; }
        else { lo[d] = cc[d] - g.reach[d]; cnt[d] = 2 * g.reach[d] + 1; }
    }
    const int ncol = cnt[1] * cnt[2];
    if (t < kMaxRuns / 2) {
        const int col = t;
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
        run_start[2 * col] = s0; run_len[2 * col] = l0;
        run_start[2 * col + 1] = s1; run_len[2 * col + 1] = l1;
    }
    return ncol;
}

// Ordered compaction of row i's bitmap B (LDS) into its ELL row; sets cnt,
// half (entries below i) and the overflow high-water mark.  One wavefront,
// no barriers.
constexpr int kSparseBits = 8;     // words with more set bits are written cooperatively

__device__ void compact_row(const DevSys &s, const uint32_t *B, int words, int i, int lane)
{
    int *row = s.nbr + (size_t)i * s.cap;
    int total = 0, nlow = 0;
    const int iw = i >> 5;
    for (int base = 0; base < words; base += 64) {
        const int k = base + lane;
        uint32_t wd = k < words ? B[k] : 0u;
        if (__ballot(wd != 0u) == 0ull) {          // an empty chunk (most of them in a large box)
            if (iw >= base && iw < base + 64) nlow = total;
            continue;
        }
        const int pc = __popc(wd);
        const int incl = wave_incl_scan(pc);
        int pos = total + incl - pc;
        if (iw >= base && iw < base + 64) {
            const int below = pos + __popc(wd & ((1u << (i & 31)) - 1u));
            nlow = __builtin_amdgcn_readlane(below, iw - base);
        }
        // Sparse words (at most kSparseBits set): each lane writes its own.
        // Denser words -- a neighbour row sets runs of consecutive indices --
        // go two per step with 32 lanes each, one bit per lane, so the loop
        // count no longer follows the densest word of the chunk.
        const bool dense = pc > kSparseBits;
        if (!dense) {
            while (wd) {
                const int b = __ffs(wd) - 1;
                wd &= wd - 1u;
                if (pos < s.cap) row[pos] = k * 32 + b;
                pos++;
            }
        }
        uint64_t dm = __ballot(dense);
        while (dm) {
            const int l0 = __builtin_ctzll(dm);
            dm &= dm - 1;
            const int l1 = dm ? __builtin_ctzll(dm) : -1;
            if (dm) dm &= dm - 1;
            const int src = lane < 32 ? l0 : l1;
            const int from = src < 0 ? l0 : src;
            const uint32_t w = (uint32_t)__shfl((int)wd, from, 64);
            const int p0 = __shfl(pos, from, 64);
            const int b = lane & 31;
            if (src >= 0 && ((w >> b) & 1u)) {
                const int p = p0 + __popc(w & ((1u << b) - 1u));
                if (p < s.cap) row[p] = (base + src) * 32 + b;
            }
        }
        total += __builtin_amdgcn_readlane(incl, 63);
    }
    if (lane == 0) {
        s.cnt[i] = total < s.cap ? total : s.cap;
        s.half[i] = nlow < s.cap ? nlow : s.cap;
        if (total > s.cap) atomicMax(s.overflow, total);
    }
}

// Rows of a group of up to G atoms of one cell (group `blockIdx.x` belongs
// to the last cell whose group_start is not above it), four wavefronts.
// Per candidate j the G row tests set bits of a hit mask; a candidate with
// some r^2 in the band around the list radius redoes its tests in FP64.
// Self pairs (r^2 = 0) are cleared with the exclusions.
template <int G, bool PBC>
__global__ __launch_bounds__(64 * kRowWaves) void k_build_groups(DevSys s, CellGrid g, int words, float cut2,
                                                                  float band, int center_ok)
{
    constexpr int NT = 64 * kRowWaves;
    extern __shared__ uint32_t bitmap[];     // G rows of `words` words
    __shared__ int run_start[kMaxRuns], run_len[kMaxRuns], run_pref[kMaxRuns + 1];
    __shared__ int s_a[G];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int grp = blockIdx.x;
    const int ngrp = g.group_start[g.ncell];
    const int4 gi = g.group_info[grp];                // (cell, first record, atoms); valid when grp < ngrp
    if (grp >= ngrp) return;                          // uniform over the block
    for (int k = tid; k < G * words; k += NT) bitmap[k] = 0u;
    const int na = gi.z, ci = gi.x;
    if (tid < na) s_a[tid] = g.cell_atoms[gi.y + tid];
    const int nrun = 2 * cell_runs(g, ci, tid, run_start, run_len);
    __syncthreads();
    // the runs as one index space: prefix of their lengths (wavefront 0)
    if (tid < 64) {
        const int incl = wave_incl_scan(tid < nrun ? run_len[tid] : 0);
        run_pref[tid + 1] = incl;
        if (tid == 0) run_pref[0] = 0;
    }
    __syncthreads();
    const int total = run_pref[nrun];

    // row atoms: FP32 offsets from the cell centre (periodic), FP64 copies
    // for the band test; absent rows sit far away
    int4 pc = make_int4(0, 0, 0, 0);
    if (PBC) {
        const int cc[3] = {ci % g.nc[0], (ci / g.nc[0]) % g.nc[1], ci / (g.nc[0] * g.nc[1])};
        pc.x = (int)(unsigned)(long long)((cc[0] + 0.5) / g.nc[0] * 4294967296.0);
        pc.y = (int)(unsigned)(long long)((cc[1] + 0.5) / g.nc[1] * 4294967296.0);
        pc.z = (int)(unsigned)(long long)((cc[2] + 0.5) / g.nc[2] * 4294967296.0);
    }
    const float ux = (float)s.unit[0], uy = (float)s.unit[1], uz = (float)s.unit[2];
    int4 pi[G];
    float ox[G], oy[G], oz[G];
    unsigned fmask = 0u;
#pragma unroll
    for (int q = 0; q < G; q++) {
        if (PBC) {
            // the row atoms' own cell records: fixed-point position, 2*atom + frozen
            pi[q] = g.cfix[gi.y + (q < na ? q : 0)];
            if (q < na && (pi[q].w & 1)) fmask |= 1u << q;
            ox[q] = q < na ? (float)wdiff(pi[q].x, pc.x) * ux : 1e18f;
            oy[q] = (float)wdiff(pi[q].y, pc.y) * uy;
            oz[q] = (float)wdiff(pi[q].z, pc.z) * uz;
        } else {
            const int aa = q < na ? s_a[q] : s_a[0];
            if (q < na && s.frozen[aa]) fmask |= 1u << q;
        }
    }
    const unsigned live = na >= 32 ? ~0u : ((1u << na) - 1u);
    int r = 0;                                   // run of this thread's next candidate
    {
        for (int base = 0; base < total; base += NT * kUnroll) {
            int code[kUnroll];
            int4 pj[kUnroll];
            double4 pd[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; u++) {
                const int f = base + u * NT + tid;
                code[u] = -1;
                if (f < total) {
                    while (run_pref[r + 1] <= f) r++;
                    const int t = run_start[r] + f - run_pref[r];
                    if (PBC) { pj[u] = g.cfix[t]; code[u] = pj[u].w; }
                    else { pd[u] = g.cpos[t]; code[u] = (int)pd[u].w; }
                }
            }
#pragma unroll
            for (int u = 0; u < kUnroll; u++) {
                if (code[u] < 0) continue;
                const int j = code[u] >> 1;
                unsigned hits = 0u;
                bool exact = !PBC;
                if (PBC) {
                    float dmin = 3.0e38f;
                    if (center_ok) {
                        const float cx = (float)wdiff(pj[u].x, pc.x) * ux;
                        const float cy = (float)wdiff(pj[u].y, pc.y) * uy;
                        const float cz = (float)wdiff(pj[u].z, pc.z) * uz;
#pragma unroll
                        for (int q = 0; q < G; q++) {
                            const float dx = cx - ox[q], dy = cy - oy[q], dz = cz - oz[q];
                            const float d = fmaf(dz, dz, fmaf(dy, dy, dx * dx)) - cut2;
                            dmin = fminf(dmin, fabsf(d));
                            hits |= d < 0.f ? (1u << q) : 0u;
                        }
                    } else {
#pragma unroll
                        for (int q = 0; q < G; q++) {
                            const float dx = (float)wdiff(pj[u].x, pi[q].x) * ux;
                            const float dy = (float)wdiff(pj[u].y, pi[q].y) * uy;
                            const float dz = (float)wdiff(pj[u].z, pi[q].z) * uz;
                            const float d = fmaf(dz, dz, fmaf(dy, dy, dx * dx)) - cut2;
                            dmin = fminf(dmin, fabsf(d));
                            hits |= d < 0.f ? (1u << q) : 0u;
                        }
                    }
                    exact = dmin <= band;
                }
                if (exact) {
                    // FP64 test of every row (the reference's r2 and test)
                    const double4 pjd = PBC ? make_double4(s.x[j], s.y[j], s.z[j], 0.0) : pd[u];
                    hits = 0u;
                    for (int q = 0; q < na; q++) {
                        const int aq = s_a[q];
                        if (within<PBC>(s, pjd, s.x[aq], s.y[aq], s.z[aq])) hits |= 1u << q;
                    }
                }
                hits &= live;
                if (code[u] & 1) hits &= ~fmask;
                const uint32_t jb = 1u << (j & 31);
                const int jw = j >> 5;
#pragma unroll
                for (int q = 0; q < G; q++)
                    if (hits & (1u << q)) atomicOr(&bitmap[q * words + jw], jb);
            }
        }
    }
    __syncthreads();
    // exclusions and the self pair only remove pairs
    for (int q = 0; q < na; q++) {
        const int aq = s_a[q];
        if (tid == 0) atomicAnd(&bitmap[q * words + (aq >> 5)], ~(1u << (aq & 31)));
        for (int k = s.xfull_off[aq] + tid; k < s.xfull_off[aq + 1]; k += NT) {
            const int e = s.xfull_idx[k];
            atomicAnd(&bitmap[q * words + (e >> 5)], ~(1u << (e & 31)));
        }
    }
    __syncthreads();
    for (int q = wv; q < na; q += kRowWaves) compact_row(s, bitmap + q * words, words, s_a[q], lane);
}

template <int G>
void launch_groups(const DevSys &s, const CellGrid &g, int words, float cut2, float band, int center_ok,
                   hipStream_t st)
{
    // groups: at most ceil(n / G) + ncell; surplus blocks return at once
    const dim3 grid((s.n + G - 1) / G + g.ncell), block(64 * kRowWaves);
    const size_t lds = (size_t)words * 4 * G;
    if (s.pbtype != 0)
        hipLaunchKernelGGL((k_build_groups<G, true>), grid, block, lds, st, s, g, words, cut2, band, center_ok);
    else
        hipLaunchKernelGGL((k_build_groups<G, false>), grid, block, lds, st, s, g, words, cut2, band, center_ok);
}

// ---------------------------------------------------------------------------
// Rows of a group of up to kMaxGroup atoms of one cell, from the group's
// candidates in ascending atom order (four wavefronts).
//
//  1. Every candidate record of the cells around the group's cell sets its
//     atom's bit in one LDS bitmap over all atoms (one bitmap per group, not
//     per row).
//  2. The bitmap, scanned in word order, gives the candidates in ascending
//     atom order, kCand at a time.
//  3. Each candidate is tested against every row atom (the FP32 prefilter
//     about the cell centre, the reference's FP64 test inside the band), a
//     32-bit hit mask per candidate.
//  4. Exclusions and the self pair clear their row's bit (a binary search in
//     the sorted candidates).
//  5. Each row takes its candidates in order: a ballot of the row's bit over
//     64 candidates gives each hit its place, so the row is written in
//     ascending order with coalesced stores; rows carry their counts from one
//     batch of candidates to the next.
constexpr int kMaxGroup = 32;      // rows of a group (hit masks are 32 bits)
constexpr int kCand = 4096;        // sorted candidates handled per batch
constexpr int kSRT = 256;          // threads of the sorted-candidate kernel
constexpr int kSRW = kSRT / 64;

size_t sorted_lds_bytes(int words) { return (size_t)words * 4; }

template <int G, bool PBC>
__global__ __launch_bounds__(kSRT) void k_build_sorted(DevSys s, CellGrid g, int words, float cut2, float band,
                                                       int center_ok)
{
    constexpr int RW = G / kSRW;            // rows per wavefront in the ordered output
    extern __shared__ uint32_t bits[];      // [words] the group's candidate atoms
    __shared__ int cand[kCand];
    __shared__ uint32_t hit[kCand];
    __shared__ int run_start[kMaxRuns], run_len[kMaxRuns], run_pref[kMaxRuns + 1];
    __shared__ int s_a[G], s_xpref[kMaxGroup + 1];
    __shared__ int4 s_pi[G];                // row atoms' cell records (fixed-point position, 2*atom + frozen)
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int grp = blockIdx.x;
    const int ngrp = g.group_start[g.ncell];
    if (grp >= ngrp) return;                          // uniform over the block
    const int4 gi = g.group_info[grp];                // (cell, first record, atoms)
    const int na = min(gi.z, G), ci = gi.x;
    for (int k = tid; k < words; k += kSRT) bits[k] = 0u;
    if (tid < na) {
        s_a[tid] = g.cell_atoms[gi.y + tid];
        if (PBC) s_pi[tid] = g.cfix[gi.y + tid];
    }
    const int nrun = 2 * cell_runs(g, ci, tid, run_start, run_len);
    __syncthreads();
    if (tid < 64) {
        const int incl = wave_incl_scan(tid < nrun ? run_len[tid] : 0);
        run_pref[tid + 1] = incl;
        if (tid == 0) run_pref[0] = 0;
        // exclusion partners of each row (plus the self pair), as a prefix
        const int nx = tid < na ? s.xfull_off[s_a[tid] + 1] - s.xfull_off[s_a[tid]] + 1 : 0;
        const int xi = wave_incl_scan(nx);
        if (tid < kMaxGroup) s_xpref[tid + 1] = xi;
        if (tid == 0) s_xpref[0] = 0;
    }
    // row atoms in registers: FP32 offsets from the cell centre (periodic),
    // fixed-point positions, frozen flags
    int4 pc = make_int4(0, 0, 0, 0);
    if (PBC) {
        const int cc[3] = {ci % g.nc[0], (ci / g.nc[0]) % g.nc[1], ci / (g.nc[0] * g.nc[1])};
        pc.x = (int)(unsigned)(long long)((cc[0] + 0.5) / g.nc[0] * 4294967296.0);
        pc.y = (int)(unsigned)(long long)((cc[1] + 0.5) / g.nc[1] * 4294967296.0);
        pc.z = (int)(unsigned)(long long)((cc[2] + 0.5) / g.nc[2] * 4294967296.0);
    }
    const float ux = (float)s.unit[0], uy = (float)s.unit[1], uz = (float)s.unit[2];
    __syncthreads();
    float ox[G], oy[G], oz[G];
    unsigned fmask = 0u;
#pragma unroll
    for (int q = 0; q < G; q++) {
        const bool in = q < na;
        if (PBC) {
            const int4 p = s_pi[in ? q : 0];
            if (in && (p.w & 1)) fmask |= 1u << q;
            ox[q] = in ? (float)wdiff(p.x, pc.x) * ux : 1e18f;
            oy[q] = (float)wdiff(p.y, pc.y) * uy;
            oz[q] = (float)wdiff(p.z, pc.z) * uz;
        } else if (in && s.frozen[s_a[q]]) {
            fmask |= 1u << q;
        }
    }
    // 1. the candidate atoms' bits
    {
        constexpr int U = 8;                          // record loads in flight per thread
        const int total = run_pref[nrun];
        int r = 0;
        for (int base = 0; base < total; base += kSRT * U) {
            int j[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int f = base + u * kSRT + tid;
                j[u] = -1;
                if (f < total) {
                    while (run_pref[r + 1] <= f) r++;
                    j[u] = g.cell_atoms[run_start[r] + f - run_pref[r]];
                }
            }
#pragma unroll
            for (int u = 0; u < U; u++)
                if (j[u] >= 0) atomicOr(&bits[j[u] >> 5], 1u << (j[u] & 31));
        }
    }
    __syncthreads();
    const unsigned live = na >= 32 ? ~0u : ((1u << na) - 1u);
    const int nxt = s_xpref[na];
    // this wavefront's rows of the ordered output: wv, wv + 4, ...
    int r_cnt[RW], r_low[RW], r_a[RW];
#pragma unroll
    for (int t = 0; t < RW; t++) {
        const int q = wv + kSRW * t;
        r_cnt[t] = 0;
        r_low[t] = 0;
        r_a[t] = q < na ? s_a[q] : 0;
    }

    // 3-5 over the batch cand[0, nc)
    auto process = [&](int nc) {
        // 3. hit masks; the next candidate's position is loaded during this
        // one's tests
        int4 qn = make_int4(0, 0, 0, 0);
        if (PBC && tid < nc) qn = g.ipos_ref[cand[tid]];
        for (int k = tid; k < nc; k += kSRT) {
            const int j = cand[k];
            unsigned hits = 0u;
            bool exact = !PBC;
            if (PBC) {
                const int4 q = qn;
                if (k + kSRT < nc) qn = g.ipos_ref[cand[k + kSRT]];
                float dmin = 3.0e38f;
                if (center_ok) {
                    const float cx = (float)wdiff(q.x, pc.x) * ux;
                    const float cy = (float)wdiff(q.y, pc.y) * uy;
                    const float cz = (float)wdiff(q.z, pc.z) * uz;
#pragma unroll
                    for (int r = 0; r < G; r++) {
                        if (r >= na) break;
                        const float dx = cx - ox[r], dy = cy - oy[r], dz = cz - oz[r];
                        const float d = fmaf(dz, dz, fmaf(dy, dy, dx * dx)) - cut2;
                        dmin = fminf(dmin, fabsf(d));
                        hits |= d < 0.f ? (1u << r) : 0u;
                    }
                } else {
                    // small boxes: the row atoms' own minimum images (LDS)
                    for (int r = 0; r < na; r++) {
                        const int4 p = s_pi[r];
                        const float dx = (float)wdiff(q.x, p.x) * ux;
                        const float dy = (float)wdiff(q.y, p.y) * uy;
                        const float dz = (float)wdiff(q.z, p.z) * uz;
                        const float d = fmaf(dz, dz, fmaf(dy, dy, dx * dx)) - cut2;
                        dmin = fminf(dmin, fabsf(d));
                        hits |= d < 0.f ? (1u << r) : 0u;
                    }
                }
                exact = dmin <= band;
            }
            if (exact) {
                // FP64 test of every row (the reference's r2 and test)
                const double4 pjd = make_double4(s.x[j], s.y[j], s.z[j], 0.0);
                hits = 0u;
                for (int r = 0; r < na; r++) {
                    const int aq = s_a[r];
                    if (within<PBC>(s, pjd, s.x[aq], s.y[aq], s.z[aq])) hits |= 1u << r;
                }
            }
            hits &= live;
            if (fmask && s.frozen[j]) hits &= ~fmask;
            hit[k] = hits;
        }
        __syncthreads();
        // 4. exclusions and the self pair: clear the row's bit where the
        // partner is among this batch's candidates
        for (int t = tid; t < nxt; t += kSRT) {
            int r = 0;
            while (s_xpref[r + 1] <= t) r++;
            const int aq = s_a[r];
            const int x = t - s_xpref[r];
            const int e = x == 0 ? aq : s.xfull_idx[s.xfull_off[aq] + x - 1];
            int lo = 0, hi = nc;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (cand[mid] < e) lo = mid + 1; else hi = mid;
            }
            if (lo < nc && cand[lo] == e) atomicAnd(&hit[lo], ~(1u << r));
        }
        __syncthreads();
        // 5. this wavefront's rows, in ascending order
        for (int base = 0; base < nc; base += 64) {
            const int k = base + lane;
            const uint32_t m = k < nc ? hit[k] : 0u;
            const int j = k < nc ? cand[k] : 0x7fffffff;
#pragma unroll
            for (int t = 0; t < RW; t++) {
                const int q = wv + kSRW * t;
                if (q >= na) break;
                const bool h = (m >> q) & 1u;
                const uint64_t b = __ballot(h);
                if (h) {
                    const int pos = r_cnt[t] + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32),
                                                                              __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0u));
                    if (pos < s.cap) s.nbr[(size_t)r_a[t] * s.cap + pos] = j;
                }
                r_low[t] += __popcll(b & __ballot(j < r_a[t]));
                r_cnt[t] += __popcll(b);
            }
        }
        __syncthreads();
    };

    // 2. candidates in ascending atom order, kCand at a time (a chunk of 64
    // words holds at most 2,048 atoms, so it always fits after a flush);
    // every wavefront scans the same chunk, lanes of wavefront w write the
    // words k = w (mod 4)
    int fill = 0;
    for (int base = 0; base < words; base += 64) {
        const int k = base + lane;
        const uint32_t wd = k < words ? bits[k] : 0u;
        const int pc_ = __popc(wd);
        const int incl = wave_incl_scan(pc_);
        const int tot = __builtin_amdgcn_readlane(incl, 63);
        if (tot == 0) continue;                      // uniform: every wavefront reads the same words
        if (fill + tot > kCand) {
            __syncthreads();
            process(fill);
            fill = 0;
        }
        if ((lane & (kSRW - 1)) == wv) {
            uint32_t w = wd;
            int pos = fill + incl - pc_;
            while (w) {
                const int b = __ffs(w) - 1;
                w &= w - 1u;
                cand[pos++] = k * 32 + b;
            }
        }
        fill += tot;
    }
    __syncthreads();
    if (fill > 0) process(fill);
    if (lane == 0) {
#pragma unroll
        for (int t = 0; t < RW; t++) {
            const int q = wv + kSRW * t;
            if (q >= na) break;
            const int c = r_cnt[t], lo = r_low[t];
            s.cnt[r_a[t]] = c < s.cap ? c : s.cap;
            s.half[r_a[t]] = lo < s.cap ? lo : s.cap;
            if (c > s.cap) atomicMax(s.overflow, c);
        }
    }
}

// Group size of the sorted-candidate kernel: 16 rows (the instance with 16
// row registers; 8, 12, 24 and 32 measured slower on the 92,384-atom box),
// or MCX_LIST_GROUP (1-32).
int sorted_group_size()
{
    const char *e = std::getenv("MCX_LIST_GROUP");
    const int v = e ? std::max(0, std::min(kMaxGroup, std::atoi(e))) : 0;
    return v ? v : 16;
}

// the kernel instance (row registers) for a group size
int sorted_template(int group) { return group <= 8 ? 8 : group <= 16 ? 16 : group <= 24 ? 24 : 32; }

template <int G>
void launch_sorted(const DevSys &s, const CellGrid &g, int group, int words, float cut2, float band, int center_ok,
                   hipStream_t st)
{
    // groups: at most ceil(n / group) + ncell; surplus blocks return at once
    const dim3 grid((s.n + group - 1) / group + g.ncell), block(kSRT);
    const size_t lds = sorted_lds_bytes(words);
    if (s.pbtype != 0)
        hipLaunchKernelGGL((k_build_sorted<G, true>), grid, block, lds, st, s, g, words, cut2, band, center_ok);
    else
        hipLaunchKernelGGL((k_build_sorted<G, false>), grid, block, lds, st, s, g, words, cut2, band, center_ok);
}

}  // namespace

size_t list_build_lds_bytes(int n)
{
    const int words = (n + 31) / 32;
    return (size_t)words * 4;
}

void launch_list_build(const DevSys &s, const CellGrid &g, hipStream_t st)
{
    const int words = (s.n + 31) / 32;
    const size_t per_row = (size_t)words * 4;
    // Rows from sorted candidates from 32,768 atoms up, while the group's
    // bitmap fits beside the candidate batch (up to ~229k atoms): there the
    // per-row bitmaps of k_build_groups cost O(N) per row (92,384 atoms:
    // 1.18 -> 0.80 ms).  Below, k_build_groups is faster (12,288 atoms:
    // 126 us against 135 us at best), its rows' bitmaps being short.
    // MCX_LIST_SORTED=1 (0) takes (never takes) the sorted kernel at any size.
    const char *sv = std::getenv("MCX_LIST_SORTED");
    const bool fits = sorted_lds_bytes(words) <= 28 * 1024;
    const bool sorted = fits && (sv ? sv[0] == '1' : s.n >= 32768);
    int G = 8;
    if (sorted) G = sorted_group_size();
    else
        while (G > 1 && per_row * G > 60 * 1024) G >>= 1;
    if (g.ncell <= kBinCells) {
        const int nb = (s.n + kBinAtoms - 1) / kBinAtoms;
        hipLaunchKernelGGL(k_bin_count, dim3(nb), dim3(kBinT), 0, st, s, g, G);
        hipLaunchKernelGGL(k_bin_fill, dim3(nb), dim3(kBinT), 0, st, s, g);
    } else {
        const int nb = (s.n + 255) / 256;
        hipLaunchKernelGGL(k_zero_int, dim3(64), dim3(256), 0, st, g.cell_count, g.ncell);
        hipLaunchKernelGGL(k_zero_int, dim3(64), dim3(256), 0, st, g.cursor, g.ncell);
        hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, st, s.overflow, 1);
        hipLaunchKernelGGL(k_cell_count, dim3(nb), dim3(256), 0, st, s, g);
        hipLaunchKernelGGL(k_cell_scan, dim3(1), dim3(1024), 0, st, g, G);
        hipLaunchKernelGGL(k_cell_fill, dim3(nb), dim3(256), 0, st, s, g);
    }
    if (g.ncell > 1) hipLaunchKernelGGL(k_cell_sort, dim3(g.ncell), dim3(256), 0, st, s, g);
    // the FP32 test relative to the cell centre needs the minimum image
    // about the centre to be the one about every atom of the cell for all
    // pairs up to the list radius: L/2 > (ctoff + dcut) + half a cell
    // diagonal (+ 1e-3 A) on every axis
    int center_ok = s.pbtype != 0;
    double diag2 = 0.0;
    for (int d = 0; d < 3; d++) diag2 += (s.L[d] / g.nc[d]) * (s.L[d] / g.nc[d]);
    const double rcut = std::sqrt(s.cutnb2);
    for (int d = 0; d < 3; d++)
        if (!(0.5 * s.L[d] > rcut + 0.5 * std::sqrt(diag2) + 1e-3)) center_ok = 0;
    const float cut2 = (float)s.cutnb2, band = (float)(s.cutnb2 * kBand);
    if (sorted) {
        switch (sorted_template(G)) {
        case 8: launch_sorted<8>(s, g, G, words, cut2, band, center_ok, st); break;
        case 16: launch_sorted<16>(s, g, G, words, cut2, band, center_ok, st); break;
        case 24: launch_sorted<24>(s, g, G, words, cut2, band, center_ok, st); break;
        default: launch_sorted<32>(s, g, G, words, cut2, band, center_ok, st); break;
        }
        return;
    }
    switch (G) {
    case 8: launch_groups<8>(s, g, words, cut2, band, center_ok, st); break;
    case 4: launch_groups<4>(s, g, words, cut2, band, center_ok, st); break;
    case 2: launch_groups<2>(s, g, words, cut2, band, center_ok, st); break;
    default: launch_groups<1>(s, g, words, cut2, band, center_ok, st); break;
    }
}
