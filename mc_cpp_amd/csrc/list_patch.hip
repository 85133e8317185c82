// list_patch.hip -- the Verlet list brought up to date after Metropolis moves
// (List_nonBonded::update_verlet_list, src/List_nonBonded.cpp:836-915, as
// MC_metropolis::run asks for it every 50 steps, src/MC_metropolis.cpp:151-154)
// by recomputing only what the moves since the last update can change.
//
// A pair's membership depends only on the two atoms' coordinates.  Pairs of
// atoms that did not move keep their membership, so:
//   * every atom a moved since the last update gets its full row rebuilt:
//     candidates from the cells around its current position (cells of the
//     last full build; atoms moved since that build are stale in them and are
//     tested from an explicit list instead), the FP64 test of the full build
//     (Tools::distance2, truncating rint, no contraction), exclusions and the
//     self pair cleared, written in ascending order;
//   * every other row j only gains or loses moved atoms: the difference of
//     a's new and old rows gives the edits "insert a into row j" and "remove
//     a from row j", applied as a sorted merge.
// Rows stay in ascending order, so the list equals a full rebuild's bit for
// bit (the tests compare it with the FP64 oracle's after runs).
//
// Limits handled by a full build instead (s.overflow is set; the next
// Metropolis chunk then changes nothing and the host rebuilds in full):
// a row outgrowing the capacity, more than kPatchEdits edits on one row,
// more than kPatchRows moved atoms in one update, more than kPatchStale
// atoms moved since the last full build.
#include "listcore.h"

namespace {

constexpr int kPT = 256;               // threads per block of the apply kernel (four wavefronts)
constexpr int kPatchBlocks = 256;

__device__ __forceinline__ void add_edit(const ListPatch &p, const DevSys &s, int j, int v)
{
    const int e = atomicAdd(&p.edit_cnt[j], 1);
    if (e == 0) p.touched[atomicAdd(p.touched_cnt, 1)] = j;
    if (e < kPatchEdits) p.edit_val[(size_t)j * kPatchEdits + e] = v;
    else atomicMax(s.overflow, 1);          // too many edits for one row: full build
}

// Is the update beyond what the patch takes (then both kernels do nothing
// and a full build follows)?
__device__ __forceinline__ bool patch_abort(const ListPatch &p, int nm, int nst)
{
    return nm > kPatchRows || nst > kPatchStale;
}

// Rows of the moved atoms, rebuilt into p.newrow, and the edits of every
// other row.  One block per moved atom at a time; LDS: the new and the old
// row and the atoms moved since the full build, as bitmaps over all atoms.
// Candidates are the cell-ordered records of the full build (fixed-point
// position and atom code in one 16-byte load), in the flattened runs of the
// (y, z) cell columns around the atom, kRowsU per thread in flight; a
// candidate's exact minimum-image displacement from the fixed-point
// positions decides unless its r^2 lies within kPatchBand (relative) of the
// list radius, where the reference's FP64 test on the master coordinates
// decides.  Atoms moved since the full build are tested from their list.
constexpr int kPRT = 1024;
constexpr int kRowsU = 4;
constexpr double kPatchBand = 1e-6;

template <bool PBC>
__global__ __launch_bounds__(kPRT) void k_patch_rows(DevSys s, CellGrid g, ListPatch p, int words)
{
    extern __shared__ uint32_t bm[];
    uint32_t *Bn = bm, *Bo = bm + words, *Bs = bm + 2 * words;
    __shared__ int run_start[64], run_len[64], run_pref[65];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int nm = *p.moved_cnt, nst = *p.stale_cnt;
    if (patch_abort(p, nm, nst)) {
        if (blockIdx.x == 0 && tid == 0) atomicMax(s.overflow, 1);
        return;
    }
    if ((int)blockIdx.x >= nm) return;
    // atoms moved since the full build: their cell records are stale
    for (int k = tid; k < words; k += kPRT) Bs[k] = 0u;
    __syncthreads();
    for (int k = tid; k < nst; k += kPRT) {
        const int j = p.stale_list[k];
        atomicOr(&Bs[j >> 5], 1u << (j & 31));
    }
    const double band = kPatchBand * s.cutnb2;
    for (int b = blockIdx.x; b < nm; b += gridDim.x) {
        const int a = p.moved_list[b];
        const double xa = s.x[a], ya = s.y[a], za = s.z[a];
        const int4 pa = s.ipos[a];
        const bool fa = s.frozen[a] != 0;
        for (int k = tid; k < 2 * words; k += kPRT) bm[k] = 0u;
        // runs of cells: the x cells of each (y, z) column around a's cell
        // are contiguous in cell order, except where the range wraps
        if (wv == 0) {
            int cc[3], lo[3], cnt[3];
            cc[0] = PBC ? cell_index(xa, s.rL[0], g.nc[0]) : 0;
            cc[1] = PBC ? cell_index(ya, s.rL[1], g.nc[1]) : 0;
            cc[2] = PBC ? cell_index(za, s.rL[2], g.nc[2]) : 0;
            for (int d = 0; d < 3; d++) {
                if (g.ncell == 1 || 2 * g.reach[d] + 1 >= g.nc[d]) { lo[d] = 0; cnt[d] = g.nc[d]; }
                else { lo[d] = cc[d] - g.reach[d]; cnt[d] = 2 * g.reach[d] + 1; }
            }
            const int ncol = cnt[1] * cnt[2];     // at most 25
            int rs = 0, rl = 0;
            const int col = lane >> 1;
            if (col < ncol) {
                const int wy = (lo[1] + col % cnt[1] + g.nc[1]) % g.nc[1];
                const int wz = (lo[2] + col / cnt[1] + g.nc[2]) % g.nc[2];
                const int base = (wz * g.nc[1] + wy) * g.nc[0];
                int x0 = lo[0], x1 = lo[0] + cnt[0] - 1;
                int a0 = 0, b0 = -1;                 // the wrapped part, if any
                if (x0 < 0) { a0 = x0 + g.nc[0]; b0 = g.nc[0] - 1; x0 = 0; }
                else if (x1 >= g.nc[0]) { a0 = 0; b0 = x1 - g.nc[0]; x1 = g.nc[0] - 1; }
                if ((lane & 1) == 0) {
                    rs = g.cell_start[base + x0];
                    rl = g.cell_start[base + x1] + g.cell_count[base + x1] - rs;
                } else if (b0 >= a0) {
                    rs = g.cell_start[base + a0];
                    rl = g.cell_start[base + b0] + g.cell_count[base + b0] - rs;
                }
            }
            const int incl = wave_incl_scan(rl);
            run_start[lane] = rs;
            run_len[lane] = rl;
            run_pref[lane + 1] = incl;
            if (lane == 0) run_pref[0] = 0;
        }
        __syncthreads();
        // the old row
        const int c_old = s.cnt[a];
        for (int k = tid; k < c_old; k += kPRT) {
            const int j = s.nbr[(size_t)a * s.cap + k];
            atomicOr(&Bo[j >> 5], 1u << (j & 31));
        }
        const int total = run_pref[64];
        int r = 0;
        for (int base = 0; base < total; base += kPRT * kRowsU) {
            int4 rec[kRowsU];
            double4 recd[kRowsU];
            bool have[kRowsU];
#pragma unroll
            for (int u = 0; u < kRowsU; u++) {
                const int f = base + u * kPRT + tid;
                have[u] = f < total;
                if (have[u]) {
                    while (run_pref[r + 1] <= f) r++;
                    const int t = run_start[r] + f - run_pref[r];
                    if (PBC) rec[u] = g.cfix[t];
                    else recd[u] = g.cpos[t];
                }
            }
#pragma unroll
            for (int u = 0; u < kRowsU; u++) {
                if (!have[u]) continue;
                const int code = PBC ? rec[u].w : (int)recd[u].w;
                const int j = code >> 1;
                if (j == a || ((Bs[j >> 5] >> (j & 31)) & 1u) || (fa && (code & 1))) continue;
                bool in;
                if (PBC) {
                    const double dx = (double)wdiff(rec[u].x, pa.x) * s.unit[0];
                    const double dy = (double)wdiff(rec[u].y, pa.y) * s.unit[1];
                    const double dz = (double)wdiff(rec[u].z, pa.z) * s.unit[2];
                    const double r2 = dx * dx + dy * dy + dz * dz;
                    in = r2 <= s.cutnb2;
                    if (fabs(r2 - s.cutnb2) <= band)
                        in = within<true>(s, make_double4(s.x[j], s.y[j], s.z[j], 0.0), xa, ya, za);
                } else {
                    in = within<false>(s, recd[u], xa, ya, za);
                }
                if (in) atomicOr(&Bn[j >> 5], 1u << (j & 31));
            }
        }
        for (int k = tid; k < nst; k += kPRT) {
            const int j = p.stale_list[k];
            if (j == a || (fa && s.frozen[j])) continue;
            if (within<PBC>(s, make_double4(s.x[j], s.y[j], s.z[j], 0.0), xa, ya, za))
                atomicOr(&Bn[j >> 5], 1u << (j & 31));
        }
        __syncthreads();
        // exclusions only remove pairs
        for (int k = s.xfull_off[a] + tid; k < s.xfull_off[a + 1]; k += kPRT) {
            const int e = s.xfull_idx[k];
            atomicAnd(&Bn[e >> 5], ~(1u << (e & 31)));
        }
        __syncthreads();
        // the new row in ascending order (wavefront 0), and the edits of the
        // rows of atoms that did not move (all wavefronts)
        if (wv == 0) {
            int *row = p.newrow + (size_t)b * s.cap;
            int total_n = 0, nlow = 0;
            for (int base = 0; base < words; base += 64) {
                const int k = base + lane;
                uint32_t wd = k < words ? Bn[k] : 0u;
                const int pc = __popc(wd);
                const int incl = wave_incl_scan(pc);
                int pos = total_n + incl - pc;
                const int below = __popc(k < (a >> 5) ? wd : (k == (a >> 5) ? (wd & ((1u << (a & 31)) - 1u)) : 0u));
                nlow += __builtin_amdgcn_readlane(wave_incl_scan(below), 63);
                while (wd) {
                    const int bit = __ffs(wd) - 1;
                    wd &= wd - 1u;
                    if (pos < s.cap) row[pos] = k * 32 + bit;
                    pos++;
                }
                total_n += __builtin_amdgcn_readlane(incl, 63);
            }
            if (lane == 0) {
                p.newcnt[b] = total_n < s.cap ? total_n : s.cap;
                p.newhalf[b] = nlow < s.cap ? nlow : s.cap;
                if (total_n > s.cap) atomicMax(s.overflow, total_n);
            }
        }
        for (int k = tid; k < words; k += kPRT) {
            const uint32_t wn = Bn[k], wo = Bo[k];
            uint32_t ins = wn & ~wo, del = wo & ~wn;
            while (ins) {
                const int j = k * 32 + __ffs(ins) - 1;
                ins &= ins - 1u;
                if (!p.moved_flag[j]) add_edit(p, s, j, a + 1);
            }
            while (del) {
                const int j = k * 32 + __ffs(del) - 1;
                del &= del - 1u;
                if (!p.moved_flag[j]) add_edit(p, s, j, -(a + 1));
            }
        }
        __syncthreads();
    }
}

// Applies the edits (one wavefront per edited row: a sorted merge through
// an LDS copy of the row) and copies the rebuilt rows of the moved atoms into
// place; the last block clears the update's counters.
__global__ __launch_bounds__(kPT) void k_patch_apply(DevSys s, ListPatch p)
{
    extern __shared__ int rowbuf[];           // [4][cap]
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int nm = *p.moved_cnt, nst = *p.stale_cnt, nt = *p.touched_cnt;
    const bool abort_ = patch_abort(p, nm, nst);
    int *buf = rowbuf + (size_t)wv * s.cap;
    for (int it = blockIdx.x * 4 + wv; !abort_ && it < nt + nm; it += gridDim.x * 4) {
        if (it < nt) {
            const int j = p.touched[it];
            const int E = min(p.edit_cnt[j], kPatchEdits);
            const int ev = lane < E ? p.edit_val[(size_t)j * kPatchEdits + lane] : 0;
            const int val = ev > 0 ? ev - 1 : -ev - 1;   // the atom inserted or removed
            const bool ins = ev > 0;
            // edits below each edit's atom: inserts and removals
            int ni_v = 0, nd_v = 0;
            for (int e = 0; e < E; e++) {
                const int ve = __builtin_amdgcn_readlane(ev, e);
                const int a_e = ve > 0 ? ve - 1 : -ve - 1;
                if (a_e < val) { if (ve > 0) ni_v++; else nd_v++; }
            }
            int *row = s.nbr + (size_t)j * s.cap;
            const int c = s.cnt[j];
            // every slot the merge should fill starts out invalid, so that a
            // slot left unfilled is caught below instead of copied into the list
            for (int k = lane; k < min(c + E, s.cap); k += 64) buf[k] = -1;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            int less_v = 0;                          // old entries below an inserted atom
            constexpr int U = 8;                     // row chunks loaded together
            for (int base = 0; base < c; base += 64 * U) {
                int xs[U];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const int k = base + u * 64 + lane;
                    xs[u] = k < c ? row[k] : 0x7fffffff;
                }
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const int k = base + u * 64 + lane;
                    const int x = xs[u];
                    int ni = 0, nd = 0;
                    bool gone = false;
                    for (int e = 0; e < E; e++) {
                        const int ve = __builtin_amdgcn_readlane(ev, e);
                        const int a_e = ve > 0 ? ve - 1 : -ve - 1;
                        if (a_e < x) { if (ve > 0) ni++; else nd++; }
                        if (ve < 0 && a_e == x) gone = true;
                        const int below = __popcll(__ballot(k < c && x < a_e));
                        if (lane == e) less_v += below;
                    }
                    const int dst = k - nd + ni;
                    if (k < c && !gone && dst < s.cap) buf[dst] = x;
                }
            }
            const int newc = c + __popcll(__ballot(lane < E && ins)) - __popcll(__ballot(lane < E && !ins));
            if (lane < E && ins) {
                const int dst = less_v - nd_v + ni_v;
                if (dst < s.cap) buf[dst] = val;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            bool bad = false;
            for (int k = lane; k < newc && k < s.cap; k += 64) {
                const int x = buf[k];
                bad = bad || x < 0 || x >= s.n || (k > 0 && buf[k - 1] >= x);
            }
            if (newc > s.cap || newc < 0 || __ballot(bad) != 0ull) {
                // more than the capacity: the host grows the list; an
                // inconsistent merge (never expected): a full build
                if (lane == 0) atomicMax(s.overflow, newc > 1 ? newc : 1);
            } else {
                int nlow = 0;
                for (int k = lane; k < newc; k += 64) {
                    const int x = buf[k];
                    row[k] = x;
                    nlow += x < j;
                }
                for (int o = 32; o > 0; o >>= 1) nlow += __shfl_xor(nlow, o, 64);
                if (lane == 0) { s.cnt[j] = newc; s.half[j] = nlow; }
            }
            if (lane == 0) p.edit_cnt[j] = 0;
        } else {
            const int b = it - nt;
            const int a = p.moved_list[b];
            const int c = p.newcnt[b];
            const int *src = p.newrow + (size_t)b * s.cap;
            int *row = s.nbr + (size_t)a * s.cap;
#pragma unroll 8
            for (int k = lane; k < c; k += 64) row[k] = src[k];
            if (lane == 0) {
                s.cnt[a] = c;
                s.half[a] = p.newhalf[b];
                p.moved_flag[a] = 0;
            }
        }
    }
    __syncthreads();
    if (tid == 0) {
        __threadfence();
        if (atomicAdd(p.ticket, 1u) == gridDim.x - 1) {
            if (!abort_) { *p.moved_cnt = 0; *p.touched_cnt = 0; }
            *p.ticket = 0u;
        }
    }
}

__global__ void k_patch_reset(int n, ListPatch p)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        p.moved_flag[k] = 0;
        p.stale_flag[k] = 0;
        p.edit_cnt[k] = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *p.moved_cnt = 0;
        *p.stale_cnt = 0;
        *p.touched_cnt = 0;
        *p.ticket = 0u;
    }
}

}  // namespace

bool list_patch_supported(const DevSys &s)
{
    const size_t words = ((size_t)s.n + 31) / 32;
    return 3 * words * 4 <= 64 * 1024 && (size_t)4 * s.cap * 4 <= 64 * 1024;
}

void launch_list_patch(const DevSys &s, const CellGrid &g, const ListPatch &p, hipStream_t st)
{
    const int words = (s.n + 31) / 32;
    const size_t lds = (size_t)3 * words * 4;
    if (s.pbtype != 0)
        hipLaunchKernelGGL((k_patch_rows<true>), dim3(kPatchBlocks), dim3(kPRT), lds, st, s, g, p, words);
    else
        hipLaunchKernelGGL((k_patch_rows<false>), dim3(kPatchBlocks), dim3(kPRT), lds, st, s, g, p, words);
    hipLaunchKernelGGL(k_patch_apply, dim3(kPatchBlocks), dim3(kPT), (size_t)4 * s.cap * 4, st, s, p);
}

void launch_list_patch_reset(const DevSys &s, const ListPatch &p, hipStream_t st)
{
    hipLaunchKernelGGL(k_patch_reset, dim3(64), dim3(256), 0, st, s.n, p);
}
