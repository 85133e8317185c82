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
// more than kEditBuf edits from one moved atom, more than kPatchRows moved
// atoms in one update, more than kPatchStale atoms moved since the last
// full build.
#include "listcore.h"

namespace {

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
// row, the atoms moved since the full build and those moved since the last
// update, as bitmaps over all atoms.  Everything that does not depend on the
// moved atom (the counts, the moved and the stale atoms, the stale atoms'
// coordinates) is read at the start, the moved atom's own data in one round
// after it.  Candidates are the cell-ordered records of the full build
// (fixed-point position and atom code in one 16-byte load), in the flattened
// runs of the (y, z) cell columns around the atom, found by a binary search
// over the runs, kRowsU per thread in flight; a candidate's exact
// minimum-image displacement from the fixed-point positions decides unless
// its r^2 lies within kPatchBand (relative) of the list radius, where the
// reference's FP64 test on the master coordinates decides.  Atoms moved since
// the full build are tested from their list.  Every thread writes the atoms
// of its bitmap word into the new row, at the block-wide prefix of the
// words' populations.
constexpr int kPRT = 1024;
constexpr int kPatchBlocks = 128;          // moved atoms beyond it: a block takes several in turn
constexpr int kRowsU = 8;
constexpr int kStaleU = kPatchStale / kPRT;   // stale atoms per thread
constexpr int kMovedU = kPatchRows / kPRT;    // moved atoms per thread
constexpr int kOldU = 2;                      // old-row entries per thread read ahead
constexpr int kEditBuf = 2048;                // edits one moved atom's block gathers (more: full build)
constexpr double kPatchBand = 1e-6;
static_assert(kPatchStale % kPRT == 0 && kPatchRows % kPRT == 0, "read-ahead sized to the limits");

template <bool PBC>
__global__ __launch_bounds__(kPRT) void k_patch_rows(DevSys s, CellGrid g, ListPatch p, int words)
{
    extern __shared__ uint32_t bm[];
    uint32_t *Bn = bm, *Bo = bm + words, *Bs = bm + 2 * words, *Bm = bm + 3 * words;
    __shared__ int run_start[64], run_pref[65];
    __shared__ int wsum[2][kPRT / 64], wlow[2][kPRT / 64], wedit[2][kPRT / 64];
    __shared__ int s_edit[kEditBuf];          // the block's edits: (row << 1) | removal
    const int tid = threadIdx.x, lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    // read before the counts are known; an entry is used only below its count
    const int nm = *p.moved_cnt, nst = *p.stale_cnt;
    int mv[kMovedU], st[kStaleU];
#pragma unroll
    for (int q = 0; q < kMovedU; q++) {
        const int k = q * kPRT + tid;
        mv[q] = k < s.n ? p.moved_list[k] : 0;
    }
#pragma unroll
    for (int q = 0; q < kStaleU; q++) {
        const int k = q * kPRT + tid;
        st[q] = k < s.n ? p.stale_list[k] : 0;
    }
    const int a_first = (int)blockIdx.x < s.n ? p.moved_list[blockIdx.x] : 0;
    // atoms moved since the full build (their cell records are stale) and
    // since the last update (their rows are rebuilt, they take no edits)
    for (int k = tid; k < 2 * words; k += kPRT) Bs[k] = 0u;
    __syncthreads();
    if (patch_abort(p, nm, nst)) {
        if (blockIdx.x == 0 && tid == 0) atomicMax(s.overflow, 1);
        return;
    }
    if ((int)blockIdx.x >= nm) return;
    // the stale atoms' master coordinates, for their explicit test
    double sx[kStaleU], sy[kStaleU], sz[kStaleU];
    bool sfr[kStaleU];
#pragma unroll
    for (int q = 0; q < kStaleU; q++) {
        sx[q] = sy[q] = sz[q] = 0.0;
        sfr[q] = false;
        if (q * kPRT + tid < nst) {
            const int j = st[q];
            sx[q] = s.x[j];
            sy[q] = s.y[j];
            sz[q] = s.z[j];
            sfr[q] = s.frozen[j] != 0;
        }
    }
#pragma unroll
    for (int q = 0; q < kStaleU; q++)
        if (q * kPRT + tid < nst) atomicOr(&Bs[st[q] >> 5], 1u << (st[q] & 31));
#pragma unroll
    for (int q = 0; q < kMovedU; q++)
        if (q * kPRT + tid < nm) atomicOr(&Bm[mv[q] >> 5], 1u << (mv[q] & 31));
    const double band = kPatchBand * s.cutnb2;
    for (int b = blockIdx.x; b < nm; b += gridDim.x) {
        const int a = b == (int)blockIdx.x ? a_first : p.moved_list[b];
        const double xa = s.x[a], ya = s.y[a], za = s.z[a];
        const int4 pa = s.ipos[a];
        const bool fa = s.frozen[a] != 0;
        const int c_old = s.cnt[a];
        const int x0 = s.xfull_off[a], x1 = s.xfull_off[a + 1];
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
                int x0c = lo[0], x1c = lo[0] + cnt[0] - 1;
                int a0 = 0, b0 = -1;                 // the wrapped part, if any
                if (x0c < 0) { a0 = x0c + g.nc[0]; b0 = g.nc[0] - 1; x0c = 0; }
                else if (x1c >= g.nc[0]) { a0 = 0; b0 = x1c - g.nc[0]; x1c = g.nc[0] - 1; }
                if ((lane & 1) == 0) {
                    rs = g.cell_start[base + x0c];
                    rl = g.cell_start[base + x1c] + g.cell_count[base + x1c] - rs;
                } else if (b0 >= a0) {
                    rs = g.cell_start[base + a0];
                    rl = g.cell_start[base + b0] + g.cell_count[base + b0] - rs;
                }
            }
            const int incl = wave_incl_scan(rl);
            run_start[lane] = rs;
            run_pref[lane + 1] = incl;
            if (lane == 0) run_pref[0] = 0;
        }
        // the old row and the excluded partners, read ahead
        int ov[kOldU];
#pragma unroll
        for (int q = 0; q < kOldU; q++) {
            const int k = q * kPRT + tid;
            ov[q] = k < c_old ? s.nbr[(size_t)a * s.cap + k] : -1;
        }
        const int xe = x0 + tid < x1 ? s.xfull_idx[x0 + tid] : -1;
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kOldU; q++)
            if (ov[q] >= 0) atomicOr(&Bo[ov[q] >> 5], 1u << (ov[q] & 31));
        for (int k = kOldU * kPRT + tid; k < c_old; k += kPRT) {
            const int j = s.nbr[(size_t)a * s.cap + k];
            atomicOr(&Bo[j >> 5], 1u << (j & 31));
        }
        const int total = run_pref[64];
        for (int base = 0; base < total; base += kPRT * kRowsU) {
            int4 rec[kRowsU];
            double4 recd[kRowsU];
            bool have[kRowsU];
#pragma unroll
            for (int u = 0; u < kRowsU; u++) {
                const int f = base + u * kPRT + tid;
                have[u] = f < total;
                if (have[u]) {
                    // the last run starting at or before f (run_pref ascends, run_pref[0] = 0)
                    int r = 0;
#pragma unroll
                    for (int stp = 32; stp > 0; stp >>= 1)
                        if (run_pref[r + stp] <= f) r += stp;
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
#pragma unroll
        for (int q = 0; q < kStaleU; q++) {
            if (q * kPRT + tid >= nst) continue;
            const int j = st[q];
            if (j == a || (fa && sfr[q])) continue;
            if (within<PBC>(s, make_double4(sx[q], sy[q], sz[q], 0.0), xa, ya, za))
                atomicOr(&Bn[j >> 5], 1u << (j & 31));
        }
        __syncthreads();
        // exclusions only remove pairs
        if (xe >= 0) atomicAnd(&Bn[xe >> 5], ~(1u << (xe & 31)));
        for (int k = x0 + kPRT + tid; k < x1; k += kPRT) {
            const int e = s.xfull_idx[k];
            atomicAnd(&Bn[e >> 5], ~(1u << (e & 31)));
        }
        __syncthreads();
        // the new row in ascending order: the atoms of word k go to the
        // block-wide prefix of the populations of the words before it (two
        // sets of wavefront totals, so that one barrier per pass suffices)
        // The edits of the rows of atoms that did not move (a inserted where
        // the new row has an atom the old one lacks, removed where the old
        // row has one the new one lacks) are gathered the same way into
        // s_edit and then applied one per thread.
        int *row = p.newrow + (size_t)b * s.cap;
        const int aw = a >> 5;
        int done_n = 0, done_low = 0, done_e = 0;
        for (int wb = 0, pass = 0; wb < words; wb += kPRT, pass ^= 1) {
            const int k = wb + tid;
            uint32_t wd = 0u, ins = 0u, del = 0u;
            if (k < words) {
                const uint32_t wo = Bo[k], keep = ~Bm[k];
                wd = Bn[k];
                ins = wd & ~wo & keep;
                del = wo & ~wd & keep;
            }
            const int pc = __popc(wd), ec = __popc(ins) + __popc(del);
            const int below = k < aw ? pc : (k == aw ? __popc(wd & ((1u << (a & 31)) - 1u)) : 0);
            const int ip = wave_incl_scan(pc);
            const int il = wave_incl_scan(below);
            const int ie = wave_incl_scan(ec);
            if (lane == 63) { wsum[pass][wv] = ip; wlow[pass][wv] = il; wedit[pass][wv] = ie; }
            __syncthreads();
            const int sp = wave_incl_scan(lane < kPRT / 64 ? wsum[pass][lane] : 0);
            const int sl = wave_incl_scan(lane < kPRT / 64 ? wlow[pass][lane] : 0);
            const int se = wave_incl_scan(lane < kPRT / 64 ? wedit[pass][lane] : 0);
            int pos = done_n + ip - pc, epos = done_e + ie - ec;
            if (wv > 0) {
                pos += __builtin_amdgcn_readlane(sp, wv - 1);
                epos += __builtin_amdgcn_readlane(se, wv - 1);
            }
            while (wd) {
                const int bit = __ffs(wd) - 1;
                wd &= wd - 1u;
                if (pos < s.cap) row[pos] = k * 32 + bit;
                pos++;
            }
            while (ins) {
                const int bit = __ffs(ins) - 1;
                ins &= ins - 1u;
                if (epos < kEditBuf) s_edit[epos] = (k * 32 + bit) << 1;
                epos++;
            }
            while (del) {
                const int bit = __ffs(del) - 1;
                del &= del - 1u;
                if (epos < kEditBuf) s_edit[epos] = ((k * 32 + bit) << 1) | 1;
                epos++;
            }
            done_n += __builtin_amdgcn_readlane(sp, kPRT / 64 - 1);
            done_low += __builtin_amdgcn_readlane(sl, kPRT / 64 - 1);
            done_e += __builtin_amdgcn_readlane(se, kPRT / 64 - 1);
        }
        if (tid == 0) {
            p.newcnt[b] = done_n < s.cap ? done_n : s.cap;
            p.newhalf[b] = done_low < s.cap ? done_low : s.cap;
            if (done_n > s.cap) atomicMax(s.overflow, done_n);
            if (done_e > kEditBuf) atomicMax(s.overflow, 1);   // a row changed beyond a patch: full build
        }
        __syncthreads();
        for (int i = tid; i < done_e && i < kEditBuf; i += kPRT) {
            const int e = s_edit[i];
            add_edit(p, s, e >> 1, (e & 1) ? -(a + 1) : a + 1);
        }
        __syncthreads();
    }
}

// Applies the edits (one workgroup per edited row: a sorted merge through an
// LDS copy of the row, every old entry moved by the number of edits below
// it) and copies the rebuilt rows of the moved atoms into place; the last
// block clears the update's counters.
constexpr int kAT = 256;
constexpr int kAU = 8;                 // row entries per thread in flight
constexpr int kApplyBlocks = 512;

__global__ __launch_bounds__(kAT) void k_patch_apply(DevSys s, ListPatch p)
{
    extern __shared__ int buf[];              // [cap]: the merged row
    __shared__ int s_less[kPatchEdits], s_bad, s_red[kAT / 64];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int nm = *p.moved_cnt, nst = *p.stale_cnt, nt = *p.touched_cnt;
    // this block's first edited row, read before nt is known (used below it)
    const int j_first = (int)blockIdx.x < s.n ? p.touched[blockIdx.x] : 0;
    const bool abort_ = patch_abort(p, nm, nst);
    for (int it = blockIdx.x; !abort_ && it < nt + nm; it += gridDim.x) {
        if (it < nt) {
            const int j = it == (int)blockIdx.x ? j_first : p.touched[it];
            const int E = min(p.edit_cnt[j], kPatchEdits);
            const int evr = lane < kPatchEdits ? p.edit_val[(size_t)j * kPatchEdits + lane] : 0;
            const int c = s.cnt[j];
            int *row = s.nbr + (size_t)j * s.cap;
            const int ev = lane < E ? evr : 0;       // lane e: edit e (every wavefront holds them)
            const int val = ev > 0 ? ev - 1 : -ev - 1;
            const bool ins = ev > 0;
            // edits below each edit's atom: inserts and removals
            int ni_v = 0, nd_v = 0;
            for (int e = 0; e < E; e++) {
                const int ve = __builtin_amdgcn_readlane(ev, e);
                const int a_e = ve > 0 ? ve - 1 : -ve - 1;
                if (a_e < val) { if (ve > 0) ni_v++; else nd_v++; }
            }
            // every slot the merge should fill starts out invalid, so that a
            // slot left unfilled is caught below instead of copied into the list
            for (int k = tid; k < min(c + E, s.cap); k += kAT) buf[k] = -1;
            if (tid < kPatchEdits) s_less[tid] = 0;
            if (tid == 0) s_bad = 0;
            __syncthreads();
            int less_v = 0;                          // this wavefront's old entries below edit e's atom
            for (int base = 0; base < c; base += kAT * kAU) {
                int xs[kAU];
#pragma unroll
                for (int u = 0; u < kAU; u++) {
                    const int k = base + u * kAT + tid;
                    xs[u] = k < c ? row[k] : 0x7fffffff;
                }
#pragma unroll
                for (int u = 0; u < kAU; u++) {
                    const int k = base + u * kAT + tid;
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
                    if (k < c && !gone && dst >= 0 && dst < s.cap) buf[dst] = x;
                }
            }
            if (lane < E && ins && less_v) atomicAdd(&s_less[lane], less_v);
            const int newc = c + __popcll(__ballot(lane < E && ins)) - __popcll(__ballot(lane < E && !ins));
            __syncthreads();
            if (wv == 0 && lane < E && ins) {
                const int dst = s_less[lane] - nd_v + ni_v;
                if (dst >= 0 && dst < s.cap) buf[dst] = val;
            }
            __syncthreads();
            bool bad = false;
            for (int k = tid; k < newc && k < s.cap; k += kAT) {
                const int x = buf[k];
                bad = bad || x < 0 || x >= s.n || (k > 0 && buf[k - 1] >= x);
            }
            if (__ballot(bad) != 0ull && lane == 0) s_bad = 1;
            __syncthreads();
            const bool ok = newc <= s.cap && newc >= 0 && s_bad == 0;
            int nlow = 0;
            if (ok) {
                for (int k = tid; k < newc; k += kAT) {
                    const int x = buf[k];
                    row[k] = x;
                    nlow += x < j;
                }
            }
            for (int o = 32; o > 0; o >>= 1) nlow += __shfl_xor(nlow, o, 64);
            if (lane == 0) s_red[wv] = nlow;
            __syncthreads();
            if (tid == 0) {
                if (ok) {
                    int h = 0;
                    for (int w = 0; w < kAT / 64; w++) h += s_red[w];
                    s.cnt[j] = newc;
                    s.half[j] = h;
                } else {
                    // more than the capacity: the host grows the list; an
                    // inconsistent merge (never expected): a full build
                    atomicMax(s.overflow, newc > 1 ? newc : 1);
                }
                p.edit_cnt[j] = 0;
            }
        } else {
            const int b = it - nt;
            const int a = p.moved_list[b];
            const int c = p.newcnt[b];
            const int hf = p.newhalf[b];
            const int *src = p.newrow + (size_t)b * s.cap;
            int *row = s.nbr + (size_t)a * s.cap;
            for (int base = 0; base < c; base += kAT * kAU) {
                int v[kAU];
#pragma unroll
                for (int u = 0; u < kAU; u++) {
                    const int k = base + u * kAT + tid;
                    v[u] = k < c ? src[k] : 0;
                }
#pragma unroll
                for (int u = 0; u < kAU; u++) {
                    const int k = base + u * kAT + tid;
                    if (k < c) row[k] = v[u];
                }
            }
            if (tid == 0) {
                s.cnt[a] = c;
                s.half[a] = hf;
                p.moved_flag[a] = 0;
            }
        }
    }
    __syncthreads();
    // every block read the counts before it takes its ticket, so the last
    // one can clear them; no device-scope fence is needed (nothing here reads
    // another block's writes, and the next kernel sees them all)
    if (tid == 0) {
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
    return 4 * words * 4 + kEditBuf * 4 + 1024 <= 64 * 1024 && (size_t)s.cap * 4 <= 60 * 1024;
}

void launch_list_patch(const DevSys &s, const CellGrid &g, const ListPatch &p, hipStream_t st)
{
    const int words = (s.n + 31) / 32;
    const size_t lds = (size_t)4 * words * 4;
    if (s.pbtype != 0)
        hipLaunchKernelGGL((k_patch_rows<true>), dim3(kPatchBlocks), dim3(kPRT), lds, st, s, g, p, words);
    else
        hipLaunchKernelGGL((k_patch_rows<false>), dim3(kPatchBlocks), dim3(kPRT), lds, st, s, g, p, words);
    hipLaunchKernelGGL(k_patch_apply, dim3(kApplyBlocks), dim3(kAT), (size_t)s.cap * 4, st, s, p);
}

void launch_list_patch_reset(const DevSys &s, const ListPatch &p, hipStream_t st)
{
    hipLaunchKernelGGL(k_patch_reset, dim3(64), dim3(256), 0, st, s.n, p);
}
