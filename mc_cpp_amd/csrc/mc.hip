// mc.hip -- per-move energy change and the Metropolis loop on the device.
//
// MC_metropolis::run (src/MC_metropolis.cpp:45-173) forms dE as the
// difference of two full-system energies (:94,122-123).  Here one
// workgroup evaluates only the terms that touch the moved atoms:
//   * nonbonded pairs from the moved atoms' full list rows (SWITCH) or all
//     atoms (FULL): each pair's old and new energy from the exact
//     fixed-point displacement and its change, in FP64 (mcx_pair_change);
//   * bonded and 1-4 terms through per-move term records, FP64.
// Force-bias translations (not in the reference) add two force passes over
// the same pairs: the net force on the moved set before the move (for the
// drift) and after it (for the reverse proposal density).
// The move itself (src/MC_metropolis.cpp:84-119; include/Move.hpp:42-183),
// the draws (src/MC.cpp:127-156: std::mt19937 + generate_canonical, or a
// Philox4x32-10 stream), the test of apply_criterion (:175-204) and the
// dmax adaptation (src/MC.cpp:165-173) all run in the same workgroup, so a
// chunk of moves is one launch with no host round trip.
#include "../../include/mcx.h"
#include "mcmove.h"

#include <cfloat>
#include <climits>

namespace {

// 512 threads: eight wavefronts of up to 256 VGPRs each, which the
// speculative kernel's proposal, walk and resolve state needs without
// spilling to scratch (at 1,024 threads the 128-VGPR limit spilled ~600
// bytes per lane, and every round's release fence wrote the spills back)
#ifndef MCX_MC_THREADS
#define MCX_MC_THREADS 512
#endif
#ifndef MCX_MC_UNROLL
#define MCX_MC_UNROLL 4
#endif
constexpr int kMCT = MCX_MC_THREADS;
constexpr int kPairU = MCX_MC_UNROLL;   // list entries in flight per thread
constexpr int kMCW = kMCT / 64;

// Shared state of one move.
struct MoveLDS {
    int first, last, nm, total, acc, type, inst, pivot;
    int scan;                    // this trial leaves the list's validity: every atom is a candidate
    int rigid;                   // the set moves rigidly: terms among its own atoms are unchanged
    int stop;                    // an accepted move left it: end the chunk for a rebuild
    double r[3], R[3][3], piv[3];
    double tx[kMaxMoveAtoms], ty[kMaxMoveAtoms], tz[kMaxMoveAtoms];
    double dx[kMaxMoveAtoms], dy[kMaxMoveAtoms], dz[kMaxMoveAtoms];   // unwrapped displacement of each moved atom
    int pref[kMaxMoveAtoms + 1];
    int4 ip[kMaxMoveAtoms];      // fixed-point mirror of the moved atoms (old positions)
    double4 mq[kMaxMoveAtoms];   // (charge, eps, sig) of the moved atoms
    int mc[kMaxMoveAtoms];       // list-row length (or n) of the moved atoms
    double ox[kMaxMoveAtoms], oy[kMaxMoveAtoms], oz[kMaxMoveAtoms];   // old coordinates
    double ms[kMaxMoveAtoms];    // masses (centre-of-mass pivots)
    double xpiv[3];              // pivot coordinates when the pivot is not moved
    int ra0, ra1;                // term-record range of the move (ra1 < 0: use the per-atom adjacency)
    double red[6][kMCW];
    double de;
    double xi[3];                // force bias: the Gaussian draw
    double fold[3], fnew[3];     // force bias: net force on the set before / after (iu/A)
};

__device__ __forceinline__ bool excluded(const DevSys &s, int i, int j)
{
    const int lo = i < j ? i : j, hi = i < j ? j : i;
    for (int k = s.excl_off[lo]; k < s.excl_off[lo + 1]; k++)
        if (s.excl_idx[k] == hi) return true;
    return false;
}

// dE contribution of pair (i, j) with i moved (index a) and j either moved
// (index b >= 0, new coordinates tx/ty/tz[b]) or fixed (b < 0).
// Adds q_i-scaled elec_raw and e_i-scaled vdw_raw; with `force` set and j
// fixed, also the force on i at its new position to fnew.
__device__ __forceinline__ void pair_delta(const DevSys &s, const MoveLDS &m, int i, int a, int j, int b,
                                           const int4 pj, const double4 pq, double &dec, double &dev, bool force,
                                           double (&fnew)[3])
{
    const int4 pi = m.ip[a];
    double o[3];                                  // old displacement x_j - x_i
    if (s.pbtype != 0) {
        o[0] = (double)mcx_wdiff(pj.x, pi.x) * s.unit[0];
        o[1] = (double)mcx_wdiff(pj.y, pi.y) * s.unit[1];
        o[2] = (double)mcx_wdiff(pj.z, pi.z) * s.unit[2];
    } else {
        o[0] = s.x[j] - s.x[i]; o[1] = s.y[j] - s.y[i]; o[2] = s.z[j] - s.z[i];
    }
    double g[3] = {-m.dx[a], -m.dy[a], -m.dz[a]};  // change of the displacement
    if (b >= 0) { g[0] += m.dx[b]; g[1] += m.dy[b]; g[2] += m.dz[b]; }
    double c, v;
    const double4 qa = m.mq[a];
    // no Lennard-Jones change when every moved atom of this wavefront's
    // entries has eps = 0 (the hydrogens of a water)
    const bool lj = __ballot(qa.y != 0.0) != 0ull;
    if (mcx_pair_change(s, o[0], o[1], o[2], g[0], g[1], g[2], qa.x, qa.y, qa.z, pq.x, pq.y, pq.z, true, c, v, lj)) {
        dec += c;
        dev += v;
    }
    if (force && b < 0) {
        double dn[3] = {o[0] + g[0], o[1] + g[1], o[2] + g[2]};
        if (s.pbtype != 0 && !(fabs(dn[0]) < 0.5 * s.L[0] && fabs(dn[1]) < 0.5 * s.L[1] && fabs(dn[2]) < 0.5 * s.L[2]))
            mcx_mimg64(s, dn);
        mcx_pair_force(s, dn, qa.x, qa.y, qa.z, pq.x, pq.y, pq.z, fnew);
    }
}

// Net nonbonded force on the moved set at its old position (force-bias
// drift): pairs of a moved atom with an atom outside the set, over the same
// list rows as move_delta; pairs inside the set cancel.  Result in m.fold
// after the final barrier.
__device__ __forceinline__ void force_old(const DevSys &s, MoveLDS &m)
{
    const int tid = threadIdx.x;
    double f[3] = {0.0, 0.0, 0.0};
    const int total = m.total, nm = m.nm;
    for (int e = tid; e < total; e += kMCT) {
        int lo = 0, hi = nm - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (m.pref[mid] <= e) lo = mid; else hi = mid - 1;
        }
        const int k = e - m.pref[lo];
        const int i = m.first + lo;
        const int j = s.cutmode == 0 ? k : s.nbr[(size_t)i * s.cap + k];
        if (j >= m.first && j <= m.last) continue;
        if (s.cutmode == 0 && excluded(s, i, j)) continue;
        const int4 pi = m.ip[lo], pj = s.ipos[j];
        double d[3];
        if (s.pbtype != 0) {
            d[0] = (double)mcx_wdiff(pj.x, pi.x) * s.unit[0];
            d[1] = (double)mcx_wdiff(pj.y, pi.y) * s.unit[1];
            d[2] = (double)mcx_wdiff(pj.z, pi.z) * s.unit[2];
        } else {
            d[0] = s.x[j] - s.x[i]; d[1] = s.y[j] - s.y[i]; d[2] = s.z[j] - s.z[i];
        }
        const double4 qa = m.mq[lo], pq = s.qes[j];
        mcx_pair_force(s, d, qa.x, qa.y, qa.z, pq.x, pq.y, pq.z, f);
    }
    for (int c = 0; c < 3; c++) f[c] = wsum_dpp(f[c]);
    const int w = tid >> 6;
    if ((tid & 63) == 0) { m.red[3][w] = f[0]; m.red[4][w] = f[1]; m.red[5][w] = f[2]; }
    __syncthreads();
    if (tid < 3) {
        double v = 0.0;
        for (int k = 0; k < kMCW; k++) v += m.red[3 + tid][k];
        m.fold[tid] = v;
    }
    __syncthreads();
}


// Does the moved set hold every atom of term record id?  (Unused slots are
// negative.)
__device__ __forceinline__ bool set_holds(const MoveLDS &m, const int4 id)
{
    return id.x >= m.first && id.x <= m.last && id.y >= m.first && id.y <= m.last &&
           (id.z < 0 || (id.z >= m.first && id.z <= m.last)) && (id.w < 0 || (id.w >= m.first && id.w <= m.last));
}

// dE of the move described in m (trial coordinates, displacements and the
// cached moved-atom data set by pick_load and trial_w0), in three parts: the nonbonded
// pairs of the moved atoms' list rows (threads 64.., each issuing the loads
// of kPairU list entries together), the bonded and 1-4 terms (wavefront 0,
// less those among the atoms of a rigidly moved set), and the workgroup
// reduction into m.de.

// Nonbonded part (threads 64..): raw Coulomb and LJ sums, and with `fb` the
// net force on the set at its new position.
// pt: the pair thread (0 .. kMCT-65) whose entries this thread evaluates;
// the one-workgroup kernels pass threadIdx.x - 64, the wide kernel the
// thread of the wavefront it stands in for when a proposal's pairs are
// split over several workgroups.
// A trial that leaves the list's validity (an atom beyond half the skin from
// where the list was built), with a cutoff: every atom is a candidate.  The
// atoms stream past each pair thread kScanU at a time; an atom farther from
// moved atom 0's old position than the cutoff plus the set's old extent, and
// farther from its new position than the cutoff plus the set's new extent
// (FP32 minimum-image distances, a 0.01 A margin), has no pair within the
// cutoff of any moved atom before or after and is skipped; the others go
// through the exact pair change as the list entries do.
constexpr int kScanU = 4;

__device__ __forceinline__ void delta_pairs_scan(const DevSys &s, const MoveLDS &m, bool fb, double &dec,
                                                 double &dev, double (&fnew)[3], int pt)
{
    constexpr int NP = kMCT - 64;
    const int nm = m.nm;
    const int4 c0 = m.ip[0];
    const float ux = (float)s.unit[0], uy = (float)s.unit[1], uz = (float)s.unit[2];
    const float Lx = (float)s.L[0], Ly = (float)s.L[1], Lz = (float)s.L[2];
    const float rLx = (float)s.rL[0], rLy = (float)s.rL[1], rLz = (float)s.rL[2];
    const float d0x = (float)m.dx[0], d0y = (float)m.dy[0], d0z = (float)m.dz[0];
    // the set's extent about atom 0, old and new
    float eo2 = 0.f, en2 = 0.f;
    for (int a = 1; a < nm; a++) {
        const float ox = (float)mcx_wdiff(m.ip[a].x, c0.x) * ux, oy = (float)mcx_wdiff(m.ip[a].y, c0.y) * uy,
                    oz = (float)mcx_wdiff(m.ip[a].z, c0.z) * uz;
        eo2 = fmaxf(eo2, fmaf(oz, oz, fmaf(oy, oy, ox * ox)));
        float nx = ox + ((float)m.dx[a] - d0x), ny = oy + ((float)m.dy[a] - d0y), nz = oz + ((float)m.dz[a] - d0z);
        nx -= Lx * rintf(nx * rLx);
        ny -= Ly * rintf(ny * rLy);
        nz -= Lz * rintf(nz * rLz);
        en2 = fmaxf(en2, fmaf(nz, nz, fmaf(ny, ny, nx * nx)));
    }
    const float rc = sqrtf((float)s.ctoff2);
    const float ro = rc + sqrtf(eo2) + 0.01f, rn = rc + sqrtf(en2) + 0.01f;
    const float ro2 = ro * ro, rn2 = rn * rn;
    for (int j0 = 0; j0 < s.n; j0 += NP * kScanU) {
        int4 pj[kScanU];
#pragma unroll
        for (int u = 0; u < kScanU; u++) {
            const int j = j0 + u * NP + pt;
            pj[u] = s.ipos[j < s.n ? j : 0];
        }
#pragma unroll
        for (int u = 0; u < kScanU; u++) {
            const int j = j0 + u * NP + pt;
            if (j >= s.n) continue;
            const float ox = (float)mcx_wdiff(pj[u].x, c0.x) * ux, oy = (float)mcx_wdiff(pj[u].y, c0.y) * uy,
                        oz = (float)mcx_wdiff(pj[u].z, c0.z) * uz;
            float nx = ox - d0x, ny = oy - d0y, nz = oz - d0z;
            nx -= Lx * rintf(nx * rLx);
            ny -= Ly * rintf(ny * rLy);
            nz -= Lz * rintf(nz * rLz);
            if (fmaf(oz, oz, fmaf(oy, oy, ox * ox)) > ro2 && fmaf(nz, nz, fmaf(ny, ny, nx * nx)) > rn2) continue;
            const double4 pq = s.qes[j];
            for (int a = 0; a < nm; a++) {
                const int i = m.first + a;
                if (j == i || excluded(s, i, j)) continue;
                if (s.frozen[i] && s.frozen[j]) continue;   // not a list pair
                int b = -1;
                if (j >= m.first && j <= m.last) {
                    if (j < i) continue;     // pair of two moved atoms: counted once
                    b = j - m.first;
                }
                pair_delta(s, m, i, a, j, b, pj[u], pq, dec, dev, fb, fnew);
            }
        }
    }
}

__device__ __forceinline__ void delta_pairs(const DevSys &s, const MoveLDS &m, bool fb, double &dec, double &dev,
                                            double (&fnew)[3], int pt = threadIdx.x - 64)
{
    constexpr int NP = kMCT - 64;
    if (m.scan && s.cutmode != 0) {
        delta_pairs_scan(s, m, fb, dec, dev, fnew, pt);
        return;
    }
    const int total = m.total, nm = m.nm;
    const bool all = s.cutmode == 0 || m.scan;
    // Row starts of a set of up to four atoms, read once: the moved atom of
    // an entry is then found without a chain of dependent LDS reads.
    const bool few = nm <= 4;
    const int p1 = nm > 1 ? m.pref[1] : INT_MAX, p2 = nm > 2 ? m.pref[2] : INT_MAX;
    const int p3 = nm > 3 ? m.pref[3] : INT_MAX;
    for (int base = 0; base < total; base += kPairU * NP) {
        // entries past the end are clamped to the last one (and dropped
        // below), so that every load is unconditional and all of them are
        // in flight together
        int av[kPairU], jv[kPairU];
#pragma unroll
        for (int u = 0; u < kPairU; u++) {
            const int e0 = base + u * NP + pt;
            const int e = e0 < total ? e0 : total - 1;
            int lo, start;
            if (few) {
                lo = (e >= p1) + (e >= p2) + (e >= p3);
                start = lo == 0 ? 0 : lo == 1 ? p1 : lo == 2 ? p2 : p3;
            } else {
                lo = 0;                        // last a with pref[a] <= e
                int hi = nm - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (m.pref[mid] <= e) lo = mid; else hi = mid - 1;
                }
                start = m.pref[lo];
            }
            const int k = e - start;
            av[u] = e0 < total ? lo : -1;
            jv[u] = all ? k : s.nbr[(size_t)(m.first + lo) * s.cap + k];
        }
        int4 pj[kPairU];
        double4 pq[kPairU];
#pragma unroll
        for (int u = 0; u < kPairU; u++) { pj[u] = s.ipos[jv[u]]; pq[u] = s.qes[jv[u]]; }
#pragma unroll
        for (int u = 0; u < kPairU; u++) {
            const int a = av[u], j = jv[u];
            if (a < 0) continue;
            const int i = m.first + a;
            if (all && (j == i || excluded(s, i, j))) continue;
            if (m.scan && s.frozen[i] && s.frozen[j]) continue;   // not a list pair
            int b = -1;
            if (j >= m.first && j <= m.last) {
                if (j < i) continue;     // pair of two moved atoms: counted once
                b = j - m.first;
            }
            pair_delta(s, m, i, a, j, b, pj[u], pq[u], dec, dev, fb, fnew);
        }
    }
}

// Bonded and 1-4 part (wavefront 0).
__device__ __forceinline__ void delta_terms(const DevSys &s, const DevMoves &mv, const MoveLDS &m, double &dec,
                                            double &dev, double &deb)
{
    const int tid = threadIdx.x;
    const CoordFetch fo{s.x, s.y, s.z, nullptr, nullptr, nullptr, 0, -1};
    const CoordFetch fn{s.x, s.y, s.z, m.tx, m.ty, m.tz, m.first, m.last};
    if (m.ra1 >= 0) {
        // self-contained records of the move instance (their loads overlap
        // the other wavefronts' pair loop)
        for (int e = m.ra0 + tid; e < m.ra1; e += 64) {
            const int4 id = mv.rec_ids[e];
            if (m.rigid && set_holds(m, id)) continue;
            const int kk = mv.rec_kk[e];
            const int ty = mv.rec_type[e];
            const double4 par = mv.rec_par[e];
            double e0, c0, v0, e1, c1, v1;
            mcx_term_energy_rec(s, kk >> 28, id, ty, par, fo, e0, c0, v0);
            mcx_term_energy_rec(s, kk >> 28, id, ty, par, fn, e1, c1, v1);
            deb += e1 - e0;
            dec += c1 - c0;
            dev += v1 - v0;
        }
    } else {
        // terms touching the moved atoms through the per-atom adjacency,
        // each counted at its lowest moved atom
        const int a0 = mv.adj_off[m.first], a1 = mv.adj_off[m.last + 1];
        for (int e = a0 + tid; e < a1; e += 64) {
            const int4 id = mv.adj_ids[e];
            const int kk = mv.adj_kk[e];
            int owner = m.first;
            while (mv.adj_off[owner + 1] <= e) owner++;
            int lowest = 0x7fffffff;
            const int ids[4] = {id.x, id.y, id.z, id.w};
            for (int q = 0; q < 4; q++)
                if (ids[q] >= m.first && ids[q] <= m.last && ids[q] < lowest) lowest = ids[q];
            if (lowest != owner || (m.rigid && set_holds(m, id))) continue;
            double e0, c0, v0, e1, c1, v1;
            mcx_term_energy_rec(s, kk >> 28, id, mv.adj_type[e], mv.adj_par[e], fo, e0, c0, v0);
            mcx_term_energy_rec(s, kk >> 28, id, mv.adj_type[e], mv.adj_par[e], fn, e1, c1, v1);
            deb += e1 - e0;
            dec += c1 - c0;
            dev += v1 - v0;
        }
    }
}

// Workgroup reduction of the parts: m.de (and m.fnew with fb) valid after
// the final barrier.  With wave0_only, m.de is needed by wavefront 0 alone:
// the final workgroup barrier becomes a wavefront barrier (the other
// wavefronts write partials again only after the next step's first
// barrier, when wavefront 0 has read these).
__device__ __forceinline__ void delta_finish(MoveLDS &m, bool fb, double dec, double dev, double deb,
                                             double (&fnew)[3], bool wave0_only = false)
{
    const int tid = threadIdx.x;
    dec = wsum_dpp(dec);
    dev = wsum_dpp(dev);
    deb = wsum_dpp(deb);
    const int w = tid >> 6;
    if ((tid & 63) == 0) { m.red[0][w] = dec; m.red[1][w] = dev; m.red[2][w] = deb; }
    if (fb) {
        for (int c = 0; c < 3; c++) fnew[c] = wsum_dpp(fnew[c]);
        if ((tid & 63) == 0) { m.red[3][w] = fnew[0]; m.red[4][w] = fnew[1]; m.red[5][w] = fnew[2]; }
    }
    __syncthreads();
    if (tid == 0) {
        double c = 0.0, v = 0.0, bb = 0.0;
        for (int k = 0; k < kMCW; k++) { c += m.red[0][k]; v += m.red[1][k]; bb += m.red[2][k]; }
        m.de = MCX_CHGCHARMM * MCX_KCALTOIU * c + 4.0 * v + bb;
    }
    if (fb && tid >= 1 && tid <= 3) {
        double v = 0.0;
        for (int k = 0; k < kMCW; k++) v += m.red[2 + tid][k];
        m.fnew[tid - 1] = v;
    }
    if (wave0_only) wave_sync();
    else __syncthreads();
}

// All three parts (the one-shot dE kernel).
__device__ __forceinline__ void move_delta(const DevSys &s, const DevTerms &t, const DevMoves &mv, MoveLDS &m)
{
    double dec = 0.0, dev = 0.0, deb = 0.0;
    double fnew[3] = {0.0, 0.0, 0.0};
    const bool fb = m.type == MCX_MV_FBIAS_DEV;
    if (__builtin_amdgcn_readfirstlane(threadIdx.x) >= 64) delta_pairs(s, m, fb, dec, dev, fnew);
    else delta_terms(s, mv, m, dec, dev, deb);
    delta_finish(m, fb, dec, dev, deb, fnew);
}

// Every datum of the moved atoms and the pivot, one thread each (the
// one-shot dE kernel).
__device__ __forceinline__ void load_moved(const DevSys &s, const DevMoves &mv, MoveLDS &m)
{
    const int a = threadIdx.x;
    if (a < m.nm) {
        const int i = m.first + a;
        m.ox[a] = s.x[i]; m.oy[a] = s.y[i]; m.oz[a] = s.z[i];
        m.ip[a] = s.ipos[i];
        m.mq[a] = s.qes[i];
        m.mc[a] = s.cutmode == 0 ? s.n : s.cnt[i];
        m.ms[a] = mv.mass ? mv.mass[i] : 1.0;
    } else if (a == kMaxMoveAtoms) {
        m.scan = 0;
        m.rigid = 1;
        const int pv = m.pivot;
        if (pv >= 0 && (pv < m.first || pv > m.last)) {
            // wavefront 0 commits positions with vector stores; a plain load
            // at this uniform address may be compiled as a scalar-cache load,
            // which those stores do not update: read through atomic loads
            m.xpiv[0] = __hip_atomic_load(s.x + pv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            m.xpiv[1] = __hip_atomic_load(s.y + pv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            m.xpiv[2] = __hip_atomic_load(s.z + pv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    }
    __syncthreads();
}

// Records atom i as moved for the incremental list update: in the moves
// since the last update and in those since the last full build.
__device__ __forceinline__ void record_move(const MCRunArgs &a, int i)
{
    const ListPatch &p = a.patch;
    if (atomicExch(&p.moved_flag[i], 1) == 0) p.moved_list[atomicAdd(p.moved_cnt, 1)] = i;
    if (atomicExch(&p.stale_flag[i], 1) == 0) p.stale_list[atomicAdd(p.stale_cnt, 1)] = i;
}

// Is fixed-point position p farther than half the list skin from its
// position r at the last list build?
__device__ __forceinline__ bool beyond_half_skin(const DevSys &s, int4 p, int4 r, double half_skin2)
{
    const double dx = (double)mcx_wdiff(p.x, r.x) * s.unit[0];
    const double dy = (double)mcx_wdiff(p.y, r.y) * s.unit[1];
    const double dz = (double)mcx_wdiff(p.z, r.z) * s.unit[2];
    return dx * dx + dy * dy + dz * dz > half_skin2;
}

// Wavefront 0, lane a < nm: trial coordinates (MOVE_TRN::tr_a_b /
// MOVE_ROT::apply_rotation, include/Move.hpp:42-50,120-143, each atom
// re-wrapped) and displacements of moved atom a; m.scan and m.rigid from
// ballots.
__device__ __forceinline__ void trial_w0(const DevSys &s, MoveLDS &m, int lane, const int4 *ipos_ref,
                                         double half_skin2)
{
    const int a = lane;
    bool far = false, beyond = false;
    if (a < m.nm) {
        const double x = m.ox[a], y = m.oy[a], z = m.oz[a];
        double ux, uy, uz;
        if (m.type != MCX_MV_ROT_DEV) {   // TRN or force-bias translation
            ux = x + m.r[0]; uy = y + m.r[1]; uz = z + m.r[2];
        } else {
            double piv[3];
            const int pv = m.pivot;
            if (pv >= m.first && pv <= m.last) {
                piv[0] = m.ox[pv - m.first]; piv[1] = m.oy[pv - m.first]; piv[2] = m.oz[pv - m.first];
            } else if (pv >= 0) {
                piv[0] = m.xpiv[0]; piv[1] = m.xpiv[1]; piv[2] = m.xpiv[2];
            } else {   // Tools::getCentreOfMass, include/Tools.hpp:95-113
                double mt_ = 0.0, c0 = 0.0, c1 = 0.0, c2 = 0.0;
                for (int b = 0; b < m.nm; b++) {
                    const double ms = m.ms[b];
                    mt_ += ms; c0 += ms * m.ox[b]; c1 += ms * m.oy[b]; c2 += ms * m.oz[b];
                }
                piv[0] = c0 / mt_; piv[1] = c1 / mt_; piv[2] = c2 / mt_;
            }
            const double px = x - piv[0], py = y - piv[1], pz = z - piv[2];
            // A rotation whose atoms all lie within a quarter box of the
            // pivot moves the set rigidly, like a translation: distances,
            // angles and torsions inside it change by rounding only
            // (~1e-15 relative) and the terms among its own atoms are
            // skipped.  The pivot is applied to wrapped coordinates
            // (include/Move.hpp:120-143), so a set the box edge splits is
            // not rigid and keeps them.
            const double lq = 0.25 * fmin(s.L[0], fmin(s.L[1], s.L[2]));
            far = s.pbtype != 0 && px * px + py * py + pz * pz >= lq * lq;
            ux = piv[0] + m.R[0][0] * px + m.R[1][0] * py + m.R[2][0] * pz;
            uy = piv[1] + m.R[0][1] * px + m.R[1][1] * py + m.R[2][1] * pz;
            uz = piv[2] + m.R[0][2] * px + m.R[1][2] * py + m.R[2][2] * pz;
        }
        m.dx[a] = ux - x; m.dy[a] = uy - y; m.dz[a] = uz - z;
        m.tx[a] = wrap1(s, ux, 0); m.ty[a] = wrap1(s, uy, 1); m.tz[a] = wrap1(s, uz, 2);
        beyond = ipos_ref && beyond_half_skin(s, mcx_fixed4(s, m.tx[a], m.ty[a], m.tz[a], 0.f),
                                              ipos_ref[m.first + a], half_skin2);
    }
    const bool scan = __ballot(beyond) != 0ull, rigid = __ballot(far) == 0ull;
    if (a == 0) { m.scan = scan; m.rigid = rigid; }
    wave_sync();
}

// Wavefront 0: prefix of the moved atoms' list-row lengths (every atom when
// the trial leaves the list's validity), the flattened pair index of
// delta_pairs and force_old.
__device__ __forceinline__ void prefix_w0(const DevSys &s, MoveLDS &m, int lane)
{
    const int nm = m.nm;
    int c = 0;
    if (lane < nm) {
        c = m.scan ? s.n : m.mc[lane];
        m.mc[lane] = c;
    }
    int v = c;
    for (int off = 1; off < 64; off <<= 1) {
        const int u = __shfl_up(v, off, 64);
        if (lane >= off) v += u;
    }
    if (lane < nm) m.pref[lane] = v - c;
    if (lane == nm - 1) { m.pref[nm] = v; m.total = v; }
    wave_sync();
}

// Phase A of the chain loop (wavefront 0): draw the move, read the moved
// atoms and the pivot (their loads issued before the displacement draws, so
// they are in flight meanwhile), publish everything in m; then, except for
// a force-bias move (whose trial waits for the force), the trial and the
// row prefix.  Everything the other wavefronts need is in m at the next
// workgroup barrier.
template <class D>
__device__ __forceinline__ void pick_load(const DevSys &s, const DevMoves &mv, D &dr, const double *dmax,
                                          MoveLDS &m, int lane, const int4 *ipos_ref, double half_skin2)
{
    int tp, in;
    int4 rec;
    int2 rr;
    pick_instance(mv, dr, lane, tp, in, rec, rr);
    const int type = mv.type[tp];
    const int first = rec.x, last = rec.y, nm = last - first + 1, pv = rec.z;
    const bool mine = lane < nm;
    const int i = first + (mine ? lane : 0);
    double ox = 0.0, oy = 0.0, oz = 0.0, ms = 1.0;
    int4 ip = make_int4(0, 0, 0, 0);
    double4 mq = make_double4(0.0, 0.0, 0.0, 0.0);
    int mc = 0;
    if (mine) {
        ox = s.x[i]; oy = s.y[i]; oz = s.z[i];
        ip = s.ipos[i];
        mq = s.qes[i];
        mc = s.cutmode == 0 ? s.n : s.cnt[i];
        if (mv.mass) ms = mv.mass[i];
    }
    const bool pout = pv >= 0 && (pv < first || pv > last);
    double p0 = 0.0, p1 = 0.0, p2 = 0.0;
    if (pout && lane == 0) {
        // a uniform address this wavefront wrote with vector stores: atomic
        // loads, which are never compiled as scalar-cache loads (the scalar
        // cache does not see those stores)
        p0 = __hip_atomic_load(s.x + pv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        p1 = __hip_atomic_load(s.y + pv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        p2 = __hip_atomic_load(s.z + pv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    double r[3];
    double R[3][3];
    pick_displacement(dr, dmax, tp, type, lane, r, R);
    if (lane == 0) {
        m.type = type; m.inst = in; m.acc = tp;
        m.first = first; m.last = last; m.pivot = pv;
        m.ra0 = rr.x; m.ra1 = rr.y;
        m.nm = nm;
        m.r[0] = r[0]; m.r[1] = r[1]; m.r[2] = r[2];
        if (type == MCX_MV_ROT_DEV)
            for (int a = 0; a < 3; a++) for (int b = 0; b < 3; b++) m.R[a][b] = R[a][b];
        if (pout) { m.xpiv[0] = p0; m.xpiv[1] = p1; m.xpiv[2] = p2; }
        m.scan = 0;
        m.rigid = 1;
    }
    if (mine) {
        m.ox[lane] = ox; m.oy[lane] = oy; m.oz[lane] = oz;
        m.ip[lane] = ip;
        m.mq[lane] = mq;
        m.mc[lane] = mc;
        m.ms[lane] = ms;
    }
    wave_sync();
    if (type != MCX_MV_FBIAS_DEV) trial_w0(s, m, lane, ipos_ref, half_skin2);
    prefix_w0(s, m, lane);
}

template <int PB, int CUT>
__global__ __launch_bounds__(kMCT) void k_mc_chunk(DevSys s0, DevTerms t, DevMoves mv, MCRunArgs arg)
{
    const DevSys s = specialise<PB, CUT>(s0);
    if (arg.bail && ((s.overflow && s.overflow[0] != 0) || arg.bail[0] != INT_MAX)) {
        // the list this chunk would use is incomplete: leave every state
        // untouched; the host grows the list and runs again from here
        if (threadIdx.x == 0) atomicMin(arg.bail, arg.step0);
        return;
    }
    __shared__ MoveLDS m;
    __shared__ uint32_t mt[kMTN];
    __shared__ double dmax[kMaxMoveTypes];
    __shared__ int acc_tmp[kMaxMoveTypes];
    __shared__ long long n_trial[kMaxMoveTypes], n_acc[kMaxMoveTypes];
    __shared__ double etot;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const bool w0 = __builtin_amdgcn_readfirstlane(tid) < 64;   // wave-uniform
    WaveDraws dr;
    if (w0) {
        dr.replay = arg.rng_mode == 1;
        dr.mt = mt;
        dr.key = make_uint2((uint32_t)arg.seed, (uint32_t)(arg.seed >> 32));
        dr.chain = arg.stream;
        dr.load(arg.mt_state, lane);
        if (lane < kMaxMoveTypes && lane < mv.ntypes) {
            dmax[lane] = arg.dmax[lane]; acc_tmp[lane] = arg.acc_tmp[lane];
            n_trial[lane] = arg.trials[lane]; n_acc[lane] = arg.accepted[lane];
        }
        if (lane == 0) etot = arg.etot[0];
    }
    if (arg.rng_mode == 1)
        for (int k = tid; k < kMTN; k += kMCT) mt[k] = arg.mt_state[k];
    if (tid == 0) m.stop = 0;
    __syncthreads();

    // The step loop, in two copies with the same sequence of workgroup
    // barriers: wavefront 0 runs the chain (pick, moved-atom loads, trial,
    // bonded terms, Metropolis test, commit) and the other wavefronts the
    // nonbonded pairs.  Separate loops keep each role's values out of the
    // other's register allocation.
    int step = 0;
    if (w0) {
        for (; step < arg.nsteps; step++) {
            const int st = arg.step0 + step + 1;
            // ---- phase A (lanes in step): pick and draw the move
            // (src/MC_metropolis.cpp:85-116; src/MC.cpp:137-156), read the
            // moved atoms, form the trial
            if (!m.stop) pick_load(s, mv, dr, dmax, m, lane, arg.ipos_ref, arg.half_skin2);
            __syncthreads();
            if (m.stop) break;
            const bool fb = m.type == MCX_MV_FBIAS_DEV;
            if (fb) {
                // force-bias proposal: dx = beta A F + sqrt(2A) xi with F the
                // net nonbonded force on the set in kcal/mol/A (iu/A /
                // 418.4), A = dmax (A^2) and beta = 1/(R T) in mol/kcal
                force_old(s, m);
                if (tid == 0) {
                    const double A = dmax[m.acc], bA = arg.beta * A * mv.fb_scale[m.acc], sq = sqrt(2.0 * A);
                    for (int c = 0; c < 3; c++) {
                        m.xi[c] = m.r[c];
                        m.r[c] = bA * m.fold[c] + sq * m.xi[c];
                    }
                }
                wave_sync();
                trial_w0(s, m, lane, arg.ipos_ref, arg.half_skin2);
                if (m.scan) prefix_w0(s, m, lane);
                __syncthreads();
            }
            double dec = 0.0, dev = 0.0, deb = 0.0;
            double fnew[3] = {0.0, 0.0, 0.0};
            delta_terms(s, mv, m, dec, dev, deb);
            delta_finish(m, fb, dec, dev, deb, fnew);
            // ---- phase E+F: Metropolis test (apply_criterion,
            // src/MC_metropolis.cpp:175-204), bookkeeping, commit of the
            // moved atoms; the next phase-A barrier publishes the commit
            const int tp = m.acc;
            const double de = m.de;
            int acc;
            double lnq = 0.0;
            if (m.type == MCX_MV_FBIAS_DEV) {
                // min(1, exp(-beta dE) q(new->old)/q(old->new)), q the
                // Gaussian of the proposal: log ratio =
                // (|dx - bA F_old|^2 - |-dx - bA F_new|^2) / (4A); beta dE
                // as the reference's test forms it (dE in iu), so force-bias
                // and TRN/ROT moves leave the same density invariant
                // whatever the drift's units
                const double A = dmax[tp], bA = arg.beta * A * mv.fb_scale[tp];
                double fw = 0.0, rv = 0.0;
                for (int c = 0; c < 3; c++) {
                    const double a = m.r[c] - bA * m.fold[c], b = -m.r[c] - bA * m.fnew[c];
                    fw += a * a;
                    rv += b * b;
                }
                lnq = (fw - rv) / (4.0 * A);
                const double x = lnq - arg.beta * de;
                acc = x >= 0.0 ? 1 : (dr.next(lane) < exp(x));
            } else {
                acc = metropolis(dr, arg.beta, de, lane);
            }
            if (lane == 0) {
                n_trial[tp]++;
                if (acc) { etot += de; n_acc[tp]++; }
                const int o = st - 1;
                if (arg.acc_out) arg.acc_out[o] = (unsigned char)acc;
                if (arg.de_out) arg.de_out[o] = de;
                if (arg.mvtyp_out) arg.mvtyp_out[o] = tp;
                if (arg.mvinst_out) arg.mvinst_out[o] = m.inst;
                if (arg.lnq_out) arg.lnq_out[o] = lnq;
                if (arg.disp_out) {
                    const bool tr = m.type != MCX_MV_ROT_DEV;
                    for (int c = 0; c < 3; c++) arg.disp_out[3 * o + c] = tr ? m.r[c] : 0.0;
                }
                adapt_dmax(mv, st, tp, acc, dmax, acc_tmp);
            }
            bool far = false;
            if (acc && lane < m.nm) {
                const int i = m.first + lane;
                s.x[i] = m.tx[lane]; s.y[i] = m.ty[lane]; s.z[i] = m.tz[lane];
                const int4 np = mcx_fixed4(s, m.tx[lane], m.ty[lane], m.tz[lane], __int_as_float(m.ip[lane].w));
                s.ipos[i] = np;
                far = arg.ipos_ref && beyond_half_skin(s, np, arg.ipos_ref[i], arg.half_skin2);
                if (arg.record_moves) record_move(arg, i);
            }
            if (__ballot(far) != 0ull && lane == 0) m.stop = 1;
        }
    } else {
        for (; step < arg.nsteps; step++) {
            __syncthreads();                  // phase A
            if (m.stop) break;
            const bool fb = m.type == MCX_MV_FBIAS_DEV;
            if (fb) {
                force_old(s, m);
                __syncthreads();              // the force-bias trial (wavefront 0)
            }
            double dec = 0.0, dev = 0.0;
            double fnew[3] = {0.0, 0.0, 0.0};
            delta_pairs(s, m, fb, dec, dev, fnew);
            delta_finish(m, fb, dec, dev, 0.0, fnew);
        }
    }
    if (arg.stop_at && tid == 0) {
        arg.stop_at[0] = step;
        arg.stop_at[1] = m.stop;
    }
    __syncthreads();
    if (w0) {
        if (lane < mv.ntypes) {
            arg.dmax[lane] = dmax[lane]; arg.acc_tmp[lane] = acc_tmp[lane];
            arg.trials[lane] = n_trial[lane]; arg.accepted[lane] = n_acc[lane];
        }
        if (lane == 0) arg.etot[0] = etot;
        dr.save(arg.mt_state, lane);
    }
    if (arg.rng_mode == 1)
        for (int k = tid; k < kMTN; k += kMCT) arg.mt_state[k] = mt[k];
}

// ---------------------------------------------------------------------------
// Read-ahead proposals.  MC_metropolis::apply_criterion draws its uniform
// only when dE exceeds DBL_EPSILON (src/MC_metropolis.cpp:180-187), so a
// rejected move always consumed exactly its proposal draws and one
// acceptance draw, and left the state unchanged: the next move's proposal
// then starts at a draw position known in advance and sees the same
// coordinates.  k_mc_spec uses this to take phase A off the critical path:
// while the other wavefronts evaluate move k's pairs, wavefront 0 draws and
// loads move k+1 into a second buffer on the assumption that move k is
// rejected (about four moves in five).  If it is, the next step starts
// from that buffer at once.  If move k is accepted, the buffer is kept
// only when the acceptance took its draw (dE above DBL_EPSILON) and moved
// none of the buffered move's atoms or its pivot; otherwise it is dropped
// (nothing was consumed from the stream) and move k+1 is drawn again after
// the commit.  The chain is the one k_mc_chunk produces.  No read-ahead is
// done across a step after which dmax is adapted.

// Is dmax adapted after (1-based) step st?  (src/MC_metropolis.cpp:137-148)
__device__ __forceinline__ bool adapts_at(const DevMoves &mv, long long st)
{
    for (int k = 0; k < mv.ntypes; k++)
        if (mv.each[k] != 0 && (st % mv.each[k]) == 0) return true;
    return false;
}

struct PipeLDS {
    MoveLDS m[2];          // the move under evaluation and the read-ahead one
    int n;                 // 1: evaluate m[cur]; 0: the chunk is over
    int cur;
    int stop;
};

// The Metropolis loop with read-ahead proposals, for move sets without
// force-bias moves; same arguments and the same chain as k_mc_chunk, and
// the same two workgroup barriers per step in both role loops.
template <int PB, int CUT>
__global__ __launch_bounds__(kMCT) void k_mc_spec(DevSys s0, DevTerms t, DevMoves mv, MCRunArgs arg)
{
    const DevSys s = specialise<PB, CUT>(s0);
    if (arg.bail && ((s.overflow && s.overflow[0] != 0) || arg.bail[0] != INT_MAX)) {
        // the list this chunk would use is incomplete: leave every state
        // untouched; the host grows the list and runs again from here
        if (threadIdx.x == 0) atomicMin(arg.bail, arg.step0);
        return;
    }
    __shared__ PipeLDS g;
    __shared__ uint32_t mt[kMTN];
    __shared__ double dmax[kMaxMoveTypes];
    __shared__ int acc_tmp[kMaxMoveTypes];
    __shared__ long long n_trial[kMaxMoveTypes], n_acc[kMaxMoveTypes];
    __shared__ double etot;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const bool w0 = __builtin_amdgcn_readfirstlane(tid) < 64;   // wave-uniform
    if (arg.rng_mode == 1)
        for (int k = tid; k < kMTN; k += kMCT) mt[k] = arg.mt_state[k];
    if (tid == 0) g.stop = 0;
    __syncthreads();
    LookDraws ld;
    if (w0) {
        ld.d.replay = arg.rng_mode == 1;
        ld.d.mt = mt;
        ld.d.key = make_uint2((uint32_t)arg.seed, (uint32_t)(arg.seed >> 32));
        ld.d.chain = arg.stream;
        ld.load(arg.mt_state, lane);
        if (lane < kMaxMoveTypes && lane < mv.ntypes) {
            dmax[lane] = arg.dmax[lane]; acc_tmp[lane] = arg.acc_tmp[lane];
            n_trial[lane] = arg.trials[lane]; n_acc[lane] = arg.accepted[lane];
        }
        if (lane == 0) etot = arg.etot[0];
        wave_sync();
    }

    int step = 0;
    if (w0) {
        bool stop = false, ready = false;
        int cur = 0;
        for (;;) {
            int n = 0;
            if (step < arg.nsteps && !stop) {
                if (!ready) pick_load(s, mv, ld, dmax, g.m[cur], lane, arg.ipos_ref, arg.half_skin2);
                n = 1;
            }
            if (lane == 0) { g.n = n; g.cur = cur; }
            __syncthreads();                  // the move is published
            if (n == 0) break;
            MoveLDS &m = g.m[cur];
            double dec = 0.0, dev = 0.0, deb = 0.0;
            double fnew[3] = {0.0, 0.0, 0.0};
            delta_terms(s, mv, m, dec, dev, deb);
            // read ahead: the next proposal, valid if this move is rejected
            // (then its acceptance draw was taken)
            int nd = -1;
            if (step + 1 < arg.nsteps && !adapts_at(mv, (long long)arg.step0 + step + 1)) {
                PeekDraws pk{ld.d.v, ld.v1, ld.d.cur + 1, true};
                const int i0 = pk.idx;
                pick_load(s, mv, pk, dmax, g.m[cur ^ 1], lane, arg.ipos_ref, arg.half_skin2);
                if (pk.ok) nd = pk.idx - i0;
            }
            delta_finish(m, false, dec, dev, deb, fnew, true);
            // Metropolis test (apply_criterion, src/MC_metropolis.cpp:175-204),
            // bookkeeping, commit of the moved atoms
            const int st = arg.step0 + step + 1;
            const int tp = m.acc;
            const double de = m.de;
            const int acc = metropolis(ld, arg.beta, de, lane);
            if (lane == 0) {
                n_trial[tp]++;
                if (acc) { etot += de; n_acc[tp]++; }
                const int o = st - 1;
                if (arg.acc_out) arg.acc_out[o] = (unsigned char)acc;
                if (arg.de_out) arg.de_out[o] = de;
                if (arg.mvtyp_out) arg.mvtyp_out[o] = tp;
                if (arg.mvinst_out) arg.mvinst_out[o] = m.inst;
                if (arg.lnq_out) arg.lnq_out[o] = 0.0;
                if (arg.disp_out) {
                    const bool tr = m.type != MCX_MV_ROT_DEV;
                    for (int c = 0; c < 3; c++) arg.disp_out[3 * o + c] = tr ? m.r[c] : 0.0;
                }
                adapt_dmax(mv, st, tp, acc, dmax, acc_tmp);
            }
            step++;
            if (acc) {
                bool far = false;
                if (lane < m.nm) {
                    const int i = m.first + lane;
                    s.x[i] = m.tx[lane]; s.y[i] = m.ty[lane]; s.z[i] = m.tz[lane];
                    const int4 np = mcx_fixed4(s, m.tx[lane], m.ty[lane], m.tz[lane], __int_as_float(m.ip[lane].w));
                    s.ipos[i] = np;
                    far = arg.ipos_ref && beyond_half_skin(s, np, arg.ipos_ref[i], arg.half_skin2);
                    if (arg.record_moves) record_move(arg, i);
                }
                if (__ballot(far) != 0ull) stop = true;
                // An acceptance that took its acceptance draw (dE above
                // DBL_EPSILON) leaves the stream where the read-ahead
                // assumed; the proposal stays valid if none of its atoms or
                // its pivot moved (and the chunk goes on: a move that left
                // the list's validity ends it, and the next chunk draws anew).
                const MoveLDS &mn = g.m[cur ^ 1];
                const bool apart = mn.last < m.first || mn.first > m.last;
                const bool pivot_ok = mn.pivot < 0 || mn.pivot < m.first || mn.pivot > m.last;
                ready = !stop && nd >= 0 && de > DBL_EPSILON && apart && pivot_ok;
            } else {
                ready = nd >= 0;
            }
            if (ready) {
                // the read-ahead proposal is the next move: take its draws
                for (int q = 0; q < nd; q++) (void)ld.next(lane);
                cur ^= 1;
            }
            wave_sync();
        }
        if (lane == 0) g.stop = stop;
    } else {
        for (;;) {
            __syncthreads();                  // the move is published
            if (g.n == 0) break;
            MoveLDS &m = g.m[g.cur];
            double dec = 0.0, dev = 0.0;
            double fnew[3] = {0.0, 0.0, 0.0};
            delta_pairs(s, m, false, dec, dev, fnew);
            delta_finish(m, false, dec, dev, 0.0, fnew, true);
        }
    }
    if (arg.stop_at && tid == 0) {
        arg.stop_at[0] = step;
        arg.stop_at[1] = g.stop;
    }
    __syncthreads();
    if (w0) {
        if (lane < mv.ntypes) {
            arg.dmax[lane] = dmax[lane]; arg.acc_tmp[lane] = acc_tmp[lane];
            arg.trials[lane] = n_trial[lane]; arg.accepted[lane] = n_acc[lane];
        }
        if (lane == 0) arg.etot[0] = etot;
        ld.save(arg.mt_state, lane);
    }
    if (arg.rng_mode == 1)
        for (int k = tid; k < kMTN; k += kMCT) arg.mt_state[k] = mt[k];
}

// ---------------------------------------------------------------------------
// Speculative rounds over several workgroups.  The read-ahead argument above
// extends to any depth: as long as the moves before it are rejected, the
// proposal g moves ahead starts at a draw position known in advance (each
// rejected move used its proposal draws and one acceptance draw) and sees
// the coordinates as they are now.  k_mc_wide runs nwg workgroups (one per
// CU, all resident: a cooperative launch).  In a round every workgroup
// parses the same stretch of the draw stream into up to nwg proposals;
// workgroup g draws and loads proposal g and evaluates its dE with the pair
// and bonded code of k_mc_chunk; the dE values meet on a board in global
// memory behind a grid-wide barrier; then every workgroup resolves the round
// identically in chain order up to and including the first accepted move,
// and writes that move's new coordinates itself (the same values from every
// workgroup), so that its own view is current without a second barrier.
// The next round starts after the last resolved move: its acceptance draw
// consumed, unless it was accepted with dE <= DBL_EPSILON (no draw).  A
// round never extends past a dmax adaptation step.  The chain is k_mc_chunk's
// bit for bit: the same proposals, the same dE (same code, same reduction
// order), the same decisions.
//
// Draws: each workgroup generates the stream (MT19937 state in LDS, or
// Philox) into a ring of kWideRing draws ahead of the chain's position.

struct RingDraws {
    const double *ring;
    long long idx, end;
    bool ok;
    __device__ __forceinline__ double next(int)
    {
        double u = 0.5;                   // past the generated draws: ends a rejection loop; ok cleared
        if (idx < end) u = ring[idx & (kWideRing - 1)];
        else ok = false;
        idx++;
        return u;
    }
};

// The draws of one proposal without its loads: exactly what pick_instance
// and pick_displacement consume for a TRN, ROT or force-bias move.
template <class D>
__device__ __forceinline__ void skip_proposal(const DevMoves &mv, D &dr, int lane)
{
    const int tp = (int)((double)mv.ntypes * dr.next(lane));
    (void)dr.next(lane);
    const int type = mv.type[tp < mv.ntypes ? tp : mv.ntypes - 1];
    if (type == MCX_MV_FBIAS_DEV) {
        // force bias: four draws of the Box-Muller normals
        for (int k = 0; k < 4; k++) (void)dr.next(lane);
        return;
    }
    double r0, r1, r2;
    do {
        r0 = dr.next(lane) * 2.0 + (-1.0);
        r1 = dr.next(lane) * 2.0 + (-1.0);
        r2 = dr.next(lane) * 2.0 + (-1.0);
    } while ((r0 * r0 + r1 * r1 + r2 * r2) > 1.0);
    if (type == MCX_MV_ROT_DEV) (void)dr.next(lane);
}

constexpr int kWideRec = 256;         // accepted moves a launch records in LDS (more: recorded at once)

struct WideLDS {
    MoveLDS m;                         // this workgroup's proposal
    double ring[kWideRing];
    int endq[kWideRing];               // position of the acceptance draw of the proposal starting at
                                       // each ring position, relative to it (-1: not yet known)
    long long start[kMaxWide];         // draw position of each proposal of the round
    long long accpos[kMaxWide];        // position of its acceptance draw
    double accu[kMaxWide];             // its acceptance draw
    int nprop;                         // proposals of the round (0: the chunk is over)
    int fail;
    double lnq;                        // force bias: log proposal-density ratio of this workgroup's proposal
};

// Grid-wide barrier of a round: every workgroup's writes before it are
// visible to every workgroup after it (agent-scope release before the
// arrival, acquire after the wait).  A wait gives up after spin_ticks of
// the constant-rate clock, or as soon as another workgroup gave up, and
// then reports failure: the kernel ends and the host reports the error.
__device__ __forceinline__ bool round_barrier(const WideArgs &w, unsigned target, int &s_fail)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_fetch_add(w.sync, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const long long t0 = (long long)wall_clock64();
        int fail = 0;
        for (;;) {
            // the count and the give-up flag read together: one round trip per poll
            const unsigned v = __hip_atomic_load(w.sync, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const int e = __hip_atomic_load(w.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (v >= target) break;
            if (e != 0 || (long long)wall_clock64() - t0 > w.spin_ticks) {
                fail = 1;
                __hip_atomic_store(w.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        s_fail = fail;
    }
    __syncthreads();
    return s_fail == 0;
}

// End of a k_mc_wide workgroup: the last one to finish clears the launch's
// counters for the next launch (every workgroup has passed its last
// rendezvous or barrier by then).
__device__ __forceinline__ void wide_exit(const WideArgs &w)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        if (__hip_atomic_fetch_add(w.exit_ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == (unsigned)w.nwg - 1) {
            __hip_atomic_store(w.sync, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(w.arrive, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(w.go, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(w.exit_ctr, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// First step >= st0 (1-based) after which dmax is adapted, or INT_MAX.
__device__ __forceinline__ long long next_adapt(const DevMoves &mv, long long st0)
{
    long long best = LLONG_MAX;
    for (int k = 0; k < mv.ntypes; k++) {
        const long long e = mv.each[k];
        if (e <= 0) continue;
        const long long s = ((st0 + e - 1) / e) * e;
        if (s < best) best = s;
    }
    return best;
}

// FB: the move set has force-bias types (their proposals carry a force pass
// and the force-bias acceptance); without, the instance has no such code.
template <int PB, int CUT, bool FB = false>
__global__ __launch_bounds__(kMCT) void k_mc_wide(DevSys s0, DevTerms t, DevMoves mv, MCRunArgs arg, WideArgs w)
{
    const DevSys s = specialise<PB, CUT>(s0);
    if (arg.bail && ((s.overflow && s.overflow[0] != 0) || arg.bail[0] != INT_MAX)) {
        // the list this chunk would use is incomplete: leave every state
        // untouched; the host grows the list and runs again from here
        if (threadIdx.x == 0 && blockIdx.x == 0) atomicMin(arg.bail, arg.step0);
        return;
    }
    __shared__ WideLDS g;
    __shared__ uint32_t mt[kMTN];
    __shared__ double dmax[kMaxMoveTypes];
    __shared__ int acc_tmp[kMaxMoveTypes];
    __shared__ long long n_trial[kMaxMoveTypes], n_acc[kMaxMoveTypes];
    __shared__ double etot;
    __shared__ int s_stop, s_done, s_go;
    __shared__ int2 rec[kWideRec];      // accepted moves (first atom, atoms) to record at the end
    __shared__ int rec_pref[kWideRec + 1];
    __shared__ int s_nrec;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int me = blockIdx.x;
    const int prop = me / w.split, part = me % w.split;   // this workgroup's proposal and share of its pairs
    const bool w0 = __builtin_amdgcn_readfirstlane(tid) < 64;   // wave-uniform
    MoveLDS &m = g.m;
    // Rendezvous: the rounds need every workgroup resident at once.  The
    // last to arrive opens it; one that waits longer than rdv_ticks closes
    // it instead; the first of the two decisions holds for all.  A closed
    // rendezvous changes nothing: the host runs the chunk again.
    if (tid == 0) {
        const unsigned t = __hip_atomic_fetch_add(w.arrive, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int go = 0;
        if (t == (unsigned)w.nwg - 1) {
            int expect = 0;
            go = __hip_atomic_compare_exchange_strong(w.go, &expect, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_AGENT) ? 1 : expect;
        } else {
            const long long t0 = (long long)wall_clock64();
            while ((go = __hip_atomic_load(w.go, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) == 0) {
                if ((long long)wall_clock64() - t0 > w.rdv_ticks) {
                    int expect = 0;
                    go = __hip_atomic_compare_exchange_strong(w.go, &expect, 2, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                              __HIP_MEMORY_SCOPE_AGENT) ? 2 : expect;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        s_go = go;
    }
    __syncthreads();
    if (s_go != 1) {
        if (tid == 0 && me == 0) {
            __hip_atomic_store(w.err, 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (arg.bail) atomicMin(arg.bail, arg.step0);
            if (arg.stop_at) { arg.stop_at[0] = 0; arg.stop_at[1] = 0; }
        }
        wide_exit(w);
        return;
    }
    if (arg.rng_mode == 1)
        for (int k = tid; k < kMTN; k += kMCT) mt[k] = arg.mt_state[k];
    if (tid == 0) { s_stop = 0; s_done = 0; g.fail = 0; s_nrec = 0; }
    __syncthreads();
    WaveDraws gen;
    long long pos = 0, filled = 0;
    long long eq_hi = 0;           // endq holds the ends of proposals starting in [pos, eq_hi)
    long long g_start_l = 0;       // lane q: start of proposal q of the round
    // wavefront 0: 32 more draws into the ring
    auto gen32 = [&]() {
        gen.refill(lane);
        if (lane < 32) g.ring[(filled + lane) & (kWideRing - 1)] = gen.v;
        filled += 32;
        wave_sync();
    };
    // generate up to `upto` while the ring has room
    auto top_up = [&](long long upto) {
        while (filled < upto && filled - pos + 32 <= kWideRing) gen32();
    };
    // where the proposals starting at q0 .. q0+63 end (their acceptance
    // draw, relative to the start; -1: past the generated draws)
    auto parse_chunk = [&](long long q0) {
        const long long q = q0 + lane;
        RingDraws d{g.ring, q, filled, true};
        skip_proposal(mv, d, lane);
        g.endq[q & (kWideRing - 1)] = (d.ok && d.idx < filled) ? (int)(d.idx - q) : -1;
        wave_sync();
    };
    // during a round's dE: fill the ring, parse up to four chunks ahead
    auto parse_ahead = [&]() {
        top_up(pos + kWideRing);
        if (eq_hi < pos) eq_hi = pos;
        for (int c = 0; c < 4 && eq_hi + 64 - pos <= kWideRing; c++) {
            parse_chunk(eq_hi);
            eq_hi += 64;
        }
    };
    if (w0) {
        gen.replay = arg.rng_mode == 1;
        gen.mt = mt;
        gen.key = make_uint2((uint32_t)arg.seed, (uint32_t)(arg.seed >> 32));
        gen.chain = arg.stream;
        gen.load(arg.mt_state, lane);
        const uint32_t *st = arg.mt_state;
        const uint32_t flag = st[kRngWords - 1];
        if (flag == 2u) {
            // drawn-but-unused draws of an earlier wide chunk
            const int cnt = (int)st[kRngWords];
            for (int k = lane; k < cnt; k += 64)
                g.ring[k] = __hiloint2double((int)st[kRngWords + 2 + 2 * k], (int)st[kRngWords + 1 + 2 * k]);
            filled = cnt;
        } else {
            // the rest of the current batch, then the batch read ahead
            const int c = gen.cur;
            if (lane >= c && lane < 32) g.ring[lane - c] = gen.v;
            filled = 32 - c;
            if (flag == 1u) {
                if (lane < 32)
                    g.ring[filled + lane] = __hiloint2double((int)st[kRngAhead + 2 * lane + 1], (int)st[kRngAhead + 2 * lane]);
                filled += 32;
            }
        }
        if (lane < kMaxMoveTypes && lane < mv.ntypes) {
            dmax[lane] = arg.dmax[lane]; acc_tmp[lane] = arg.acc_tmp[lane];
            n_trial[lane] = arg.trials[lane]; n_acc[lane] = arg.accepted[lane];
        }
        if (lane == 0) etot = arg.etot[0];
        wave_sync();
    }
    __syncthreads();

    int done = 0;
    unsigned round = 0;
    bool ok = true;
    for (;;) {
        // ---- the round's proposals (wavefront 0 of every workgroup, the
        // same in all): draw positions and acceptance draws, by following
        // the chain of proposals through endq (where each proposal ends;
        // mostly parsed ahead during the previous round's dE).  Where a
        // round stops depends on the draw stream alone, never on how much
        // a workgroup generated or parsed ahead.
        if (w0) {
            int np = 0;
            const int left = arg.nsteps - done;
            if (eq_hi < pos) eq_hi = pos;
            if (left > 0 && !s_stop) {
                long long wmax = min(left, w.nwg / w.split);
                const long long na = next_adapt(mv, (long long)arg.step0 + done + 1);
                if (na != LLONG_MAX) wmax = min(wmax, na - ((long long)arg.step0 + done));
                long long p = pos;
                while (np < wmax && p + 64 - pos <= kWideRing) {
                    // proposals already parsed: a compact loop (the parse and
                    // generation below are rare, and keeping them out of this
                    // loop keeps its code short), in 32-bit offsets from pos
                    {
                        const int qhi = (int)(eq_hi - pos), wm = (int)wmax;
                        const int rb = (int)(pos & (kWideRing - 1));
                        int q = (int)(p - pos);
                        while (q < qhi) {
                            const int e = g.endq[(rb + q) & (kWideRing - 1)];
                            if (e < 0) break;
                            if (lane == np) g_start_l = pos + q;
                            q += e + 1;
                            np++;
                            if (!(np < wm && q <= kWideRing - 64)) break;
                        }
                        p = pos + q;
                    }
                    if (!(np < wmax && p + 64 - pos <= kWideRing)) break;
                    if (p >= eq_hi) {
                        top_up(eq_hi + 192);
                        parse_chunk(eq_hi);
                        eq_hi += 64;
                        continue;
                    }
                    int e = g.endq[p & (kWideRing - 1)];
                    if (e < 0) {
                        // the proposal at p ran past the draws generated when
                        // it was parsed: generate more (while the ring has
                        // room) and parse again; with a full ring the result
                        // is final
                        const bool full = filled - pos + 32 > kWideRing;
                        if (!full) gen32();
                        parse_chunk(p);
                        if (p + 64 > eq_hi) eq_hi = p + 64;
                        if (full && g.endq[p & (kWideRing - 1)] < 0) break;
                        continue;
                    }
                    if (lane == np) g_start_l = p;
                    p += e + 1;
                    np++;
                }
            }
            // a chunk that still has moves always gets a proposal, unless one
            // proposal needs more draws than the ring holds (then stop and
            // report: the host raises an error)
            if (np == 0 && left > 0 && !s_stop && lane == 0)
                __hip_atomic_store(w.err, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            // lane q < np: proposal q's start, acceptance-draw position and draw
            if (lane < np) {
                const long long ap = g_start_l + g.endq[g_start_l & (kWideRing - 1)];
                g.start[lane] = g_start_l;
                g.accpos[lane] = ap;
                g.accu[lane] = g.ring[ap & (kWideRing - 1)];
            }
            if (lane == 0) g.nprop = np;
            wave_sync();
            // this workgroup's proposal: draw and load it, form the trial
            if (prop < np) {
                RingDraws d{g.ring, g.start[prop], filled, true};
                pick_load(s, mv, d, dmax, m, lane, arg.ipos_ref, arg.half_skin2);
            }
        }
        __syncthreads();
        const int np = g.nprop;
        if (np == 0) break;
        // ---- dE of this workgroup's proposal (its share of the pairs when
        // a proposal is split over several workgroups): each wavefront's
        // sums go to the board, where the resolution adds them in the
        // order of delta_finish; meanwhile wavefront 0 fills the draw ring
        // and parses ahead for the next round
        if (FB && prop < np && m.type == MCX_MV_FBIAS_DEV) {
            // force-bias proposal (one workgroup per proposal): the net force
            // on the set at its old position, the drifted trial, then dE and
            // the force at the new position, reduced as the one-proposal
            // kernel reduces them; dE goes to the board as the bonded part of
            // wavefront 0 (the other sums zero, so the resolution's sum is
            // dE exactly) and the log proposal-density ratio beside it
            force_old(s, m);
            if (w0) {
                if (lane == 0) {
                    const double A = dmax[m.acc], bA = arg.beta * A * mv.fb_scale[m.acc], sq = sqrt(2.0 * A);
                    for (int c = 0; c < 3; c++) {
                        m.xi[c] = m.r[c];
                        m.r[c] = bA * m.fold[c] + sq * m.xi[c];
                    }
                }
                wave_sync();
                trial_w0(s, m, lane, arg.ipos_ref, arg.half_skin2);
                if (m.scan) prefix_w0(s, m, lane);
            }
            __syncthreads();
            double dec = 0.0, dev = 0.0, deb = 0.0;
            double fnew[3] = {0.0, 0.0, 0.0};
            if (w0) delta_terms(s, mv, m, dec, dev, deb);
            else delta_pairs(s, m, true, dec, dev, fnew);
            delta_finish(m, true, dec, dev, deb, fnew);
            if (w0) {
                if (lane == 0) {
                    const int tp = m.acc;
                    const double A = dmax[tp], bA = arg.beta * A * mv.fb_scale[tp];
                    double fw = 0.0, rv = 0.0;
                    for (int c = 0; c < 3; c++) {
                        const double a = m.r[c] - bA * m.fold[c], b = -m.r[c] - bA * m.fnew[c];
                        fw += a * a;
                        rv += b * b;
                    }
                    g.lnq = (fw - rv) / (4.0 * A);
                    w.blnq[prop] = g.lnq;
                }
                if (lane < kMCW) {
                    double *r = w.bred + ((size_t)prop * kMCW + lane) * 3;
                    r[0] = 0.0; r[1] = 0.0; r[2] = lane == 0 ? m.de : 0.0;
                }
                if (lane == 0) w.bmeta[prop] = make_int4(m.acc, m.first, m.nm, m.inst);
                for (int a = lane; a < m.nm; a += 64) {
                    double *b = w.bxyz + ((size_t)prop * kMaxMoveAtoms + a) * 3;
                    b[0] = m.tx[a]; b[1] = m.ty[a]; b[2] = m.tz[a];
                    w.bip[(size_t)prop * kMaxMoveAtoms + a] =
                        mcx_fixed4(s, m.tx[a], m.ty[a], m.tz[a], __int_as_float(m.ip[a].w));
                }
                parse_ahead();
            }
        } else if (prop < np) {
            double dec = 0.0, dev = 0.0, deb = 0.0;
            double fnew[3] = {0.0, 0.0, 0.0};
            const int v0 = 1 + part * (kMCW - 1) / w.split, v1 = 1 + (part + 1) * (kMCW - 1) / w.split;
            const int wr = tid >> 6;                  // this wavefront; wavefront 0 keeps the bonded terms
            const int vw = v0 + wr - 1;               // the wavefront of the one-workgroup kernel it stands in for
            if (w0) {
                if (part == 0) delta_terms(s, mv, m, dec, dev, deb);
                parse_ahead();
            } else if (vw < v1) {
                delta_pairs(s, m, false, dec, dev, fnew, (vw - 1) * 64 + lane);
            }
            if (w0 ? part == 0 : vw < v1) {
                dec = wsum_dpp(dec);
                dev = wsum_dpp(dev);
                deb = wsum_dpp(deb);
                if (lane == 0) {
                    double *r = w.bred + ((size_t)prop * kMCW + (w0 ? 0 : vw)) * 3;
                    r[0] = dec; r[1] = dev; r[2] = deb;
                }
            }
            if (w0 && part == 0) {
                if (lane == 0) w.bmeta[prop] = make_int4(m.acc, m.first, m.nm, m.inst);
                for (int a = lane; a < m.nm; a += 64) {
                    double *b = w.bxyz + ((size_t)prop * kMaxMoveAtoms + a) * 3;
                    b[0] = m.tx[a]; b[1] = m.ty[a]; b[2] = m.tz[a];
                    w.bip[(size_t)prop * kMaxMoveAtoms + a] =
                        mcx_fixed4(s, m.tx[a], m.ty[a], m.tz[a], __int_as_float(m.ip[a].w));
                }
            }
        } else if (w0) {
            parse_ahead();
        }
        round++;
        if (!round_barrier(w, round * (unsigned)w.nwg, g.fail)) { ok = false; break; }
        // ---- resolve in chain order (wavefront 0, the same in every
        // workgroup).  Lane q takes proposal q's dE and acceptance test and
        // also reads its first four trial positions, so that the commit of
        // a move of up to four atoms needs no further round trip.
        if (w0) {
            // lane q < np: proposal q's dE, its wavefront sums added in the
            // order of delta_finish
            double de_l = 0.0;
            if (lane < np) {
                const double *r = w.bred + (size_t)lane * kMCW * 3;
                double c = 0.0, v = 0.0, bb = 0.0;
#pragma unroll
                for (int k = 0; k < kMCW; k++) { c += r[3 * k]; v += r[3 * k + 1]; bb += r[3 * k + 2]; }
                de_l = MCX_CHGCHARMM * MCX_KCALTOIU * c + 4.0 * v + bb;
            }
            const int4 meta_l = lane < np ? w.bmeta[lane] : make_int4(0, 0, 0, 0);
            const int pq = lane >> 2, pa = lane & 3;
            double cx = 0.0, cy = 0.0, cz = 0.0;
            int4 cip = make_int4(0, 0, 0, 0);
            if (pq < np) {
                const double *b = w.bxyz + ((size_t)pq * kMaxMoveAtoms + pa) * 3;
                cx = b[0]; cy = b[1]; cz = b[2];
                cip = w.bip[(size_t)pq * kMaxMoveAtoms + pa];
            }
            // MC_metropolis::apply_criterion (src/MC_metropolis.cpp:175-204)
            // force bias: min(1, exp(lnq - beta dE)), the acceptance draw taken
            // only below 1 (as the one-proposal kernel tests it)
            const bool fb_l = FB && lane < np && mv.type[meta_l.x] == MCX_MV_FBIAS_DEV;
            const double xfb = fb_l ? w.blnq[lane] - arg.beta * de_l : 0.0;
            const bool free_acc = fb_l ? xfb >= 0.0 : de_l <= DBL_EPSILON;
            const bool acc_l = lane < np && (free_acc || g.accu[lane < kMaxWide ? lane : 0] <
                                                              (fb_l ? exp(xfb) : exp(-arg.beta * de_l)));
            const uint64_t accm = __ballot(acc_l);
            const int acc_at = accm ? __builtin_ctzll(accm) : -1;
            const int R = acc_at >= 0 ? acc_at + 1 : np;
            // bookkeeping of the R resolved moves: trials per type, the
            // accepted move, then the dmax adaptation of MC::adjust_dmax
            // (src/MC.cpp:165-173), which a round's last move alone can meet
            // (rounds end at adaptation steps)
            int ntr = 0;
            for (int k = 0; k < mv.ntypes; k++) {
                const int c = __popcll(__ballot(lane < R && meta_l.x == k));
                if (lane == k) ntr = c;
            }
            if (lane < mv.ntypes) n_trial[lane] += ntr;
            wave_sync();
            if (lane == 0) {
                if (acc_at >= 0) {
                    const int tp = __builtin_amdgcn_readlane(meta_l.x, acc_at);
                    etot += __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(de_l), acc_at),
                                             __builtin_amdgcn_readlane(__double2loint(de_l), acc_at));
                    n_acc[tp]++;
                    acc_tmp[tp]++;
                }
                adapt_dmax(mv, (long long)arg.step0 + done + R, __builtin_amdgcn_readlane(meta_l.x, R - 1), 0, dmax,
                           acc_tmp);
            }
            const double de_own = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(de_l), prop < np ? prop : 0),
                                                   __builtin_amdgcn_readlane(__double2loint(de_l), prop < np ? prop : 0));
            if (prop < R && part == 0 && lane == 0) {
                // this workgroup's own move: its per-step records
                const int o = arg.step0 + done + prop;
                if (arg.acc_out) arg.acc_out[o] = (unsigned char)(prop == acc_at);
                if (arg.de_out) arg.de_out[o] = de_own;
                if (arg.mvtyp_out) arg.mvtyp_out[o] = m.acc;
                if (arg.mvinst_out) arg.mvinst_out[o] = m.inst;
                if (arg.lnq_out) arg.lnq_out[o] = FB && m.type == MCX_MV_FBIAS_DEV ? g.lnq : 0.0;
                if (arg.disp_out) {
                    const bool tr = m.type != MCX_MV_ROT_DEV;
                    for (int c = 0; c < 3; c++) arg.disp_out[3 * o + c] = tr ? m.r[c] : 0.0;
                }
            }
            // the chain's next draw: after the last resolved proposal's
            // acceptance draw, or at it if that move was accepted with
            // dE <= DBL_EPSILON (apply_criterion then draws nothing)
            const bool last_free = __builtin_amdgcn_readlane((int)free_acc, R - 1) != 0;
            pos = g.accpos[R - 1] + ((acc_at >= 0 && last_free) ? 0 : 1);
            // the accepted move's new coordinates, written by every
            // workgroup (the same values)
            bool far = false;
            if (acc_at >= 0) {
                const int first = __builtin_amdgcn_readlane(meta_l.y, acc_at);
                const int nm = __builtin_amdgcn_readlane(meta_l.z, acc_at);
                // positions read above: the first four atoms of proposals 0-15
                const int a0 = acc_at < 16 ? 4 : 0;
                if (acc_at < 16 && pq == acc_at && pa < nm) {
                    const int i = first + pa;
                    s.x[i] = cx; s.y[i] = cy; s.z[i] = cz;
                    s.ipos[i] = cip;
                    far = arg.ipos_ref && beyond_half_skin(s, cip, arg.ipos_ref[i], arg.half_skin2);
                }
                for (int a = a0 + lane; a < nm; a += 64) {
                    const int i = first + a;
                    const double *b = w.bxyz + ((size_t)acc_at * kMaxMoveAtoms + a) * 3;
                    const int4 np4 = w.bip[(size_t)acc_at * kMaxMoveAtoms + a];
                    s.x[i] = b[0]; s.y[i] = b[1]; s.z[i] = b[2];
                    s.ipos[i] = np4;
                    far = far || (arg.ipos_ref && beyond_half_skin(s, np4, arg.ipos_ref[i], arg.half_skin2));
                }
                // the moved atoms, recorded for the list update at the end of
                // the launch (off this round's critical path)
                if (arg.record_moves && me == 0) {
                    if (s_nrec < kWideRec) {
                        if (lane == 0) rec[s_nrec] = make_int2(first, nm);
                        wave_sync();
                        if (lane == 0) s_nrec++;
                    } else {
                        for (int a = lane; a < nm; a += 64) record_move(arg, first + a);
                    }
                }
            }
            done += R;
            if (__ballot(far) != 0ull && lane == 0) s_stop = 1;
            if (lane == 0) s_done = done;
            wave_sync();
        }
        __syncthreads();
        done = s_done;
    }
    if (!ok) {
        wide_exit(w);
        return;
    }
    if (arg.stop_at && tid == 0 && me == 0) {
        arg.stop_at[0] = done;
        arg.stop_at[1] = s_stop;
    }
    wide_exit(w);
    if (me != 0) return;
    if (arg.record_moves) {
        // the launch's moved atoms, one thread each
        if (tid == 0) {
            int acc = 0;
            for (int r = 0; r < s_nrec; r++) { rec_pref[r] = acc; acc += rec[r].y; }
            rec_pref[s_nrec] = acc;
        }
        __syncthreads();
        const int nr = s_nrec, total = rec_pref[nr];
        for (int t = tid; t < total; t += kMCT) {
            int lo = 0, hi = nr - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (rec_pref[mid] <= t) lo = mid; else hi = mid - 1;
            }
            record_move(arg, rec[lo].x + (t - rec_pref[lo]));
        }
    }
    if (w0) {
        if (lane < mv.ntypes) {
            arg.dmax[lane] = dmax[lane]; arg.acc_tmp[lane] = acc_tmp[lane];
            arg.trials[lane] = n_trial[lane]; arg.accepted[lane] = n_acc[lane];
        }
        if (lane == 0) arg.etot[0] = etot;
        // generator state, then the drawn-but-unused draws [pos, filled)
        gen.save(arg.mt_state, lane);
        uint32_t *st = arg.mt_state;
        const int cnt = (int)(filled - pos);
        for (int k = lane; k < cnt; k += 64) {
            const double u = g.ring[(pos + k) & (kWideRing - 1)];
            st[kRngWords + 1 + 2 * k] = (uint32_t)__double2loint(u);
            st[kRngWords + 2 + 2 * k] = (uint32_t)__double2hiint(u);
        }
        if (lane == 0) {
            st[kRngWords] = (uint32_t)cnt;
            st[kRngWords - 1] = 2u;
        }
    }
    if (arg.rng_mode == 1)
        for (int k = tid; k < kMTN; k += kMCT) arg.mt_state[k] = mt[k];
}

// one-shot dE for caller-supplied trial coordinates
__global__ __launch_bounds__(kMCT) void k_delta(DevSys s, DevTerms t, DevMoves mv, int first, int last,
                                                const double *__restrict__ nxyz, double *out)
{
    __shared__ MoveLDS m;
    const int nm = last - first + 1;
    if (threadIdx.x == 0) {
        m.first = first; m.last = last; m.nm = nm; m.pivot = -1; m.type = MCX_MV_TRN_DEV; m.scan = 0;
        m.ra0 = 0; m.ra1 = -1;
    }
    __syncthreads();
    load_moved(s, mv, m);
    if (threadIdx.x < nm) {
        const int a = threadIdx.x;
        const double nx = nxyz[3 * a], ny = nxyz[3 * a + 1], nz = nxyz[3 * a + 2];
        double d[3] = {nx - m.ox[a], ny - m.oy[a], nz - m.oz[a]};
        mcx_pbc(s, d);
        m.dx[a] = d[0]; m.dy[a] = d[1]; m.dz[a] = d[2];
        m.tx[a] = nx; m.ty[a] = ny; m.tz[a] = nz;
    }
    if (threadIdx.x == 64) {
        int acc = 0;
        for (int b = 0; b < nm; b++) { m.pref[b] = acc; acc += m.mc[b]; }
        m.pref[nm] = acc;
        m.total = acc;
        m.rigid = 0;      // arbitrary trial coordinates
    }
    __syncthreads();
    move_delta(s, t, mv, m);
    if (threadIdx.x == 0) out[0] = m.de;
}

}  // namespace

void launch_delta_energy(const DevSys &s, const DevTerms &t, const DevMoves &mv, int first, int last,
                         const double *newxyz, double *de, hipStream_t st)
{
    hipLaunchKernelGGL(k_delta, dim3(1), dim3(kMCT), 0, st, s, t, mv, first, last, newxyz, de);
}

// The fixed-point mirror of every atom from the FP64 master coordinates
// (after a caller wrote them in place), charge bits kept.
__global__ void k_mirror(DevSys s)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < s.n; i += gridDim.x * blockDim.x) {
        const int qb = s.ipos[i].w;
        s.ipos[i] = mcx_fixed4(s, s.x[i], s.y[i], s.z[i], __int_as_float(qb));
    }
}

// The same from a packed upload (x[n], y[n], z[n]), which also becomes the
// FP64 master copy.
__global__ void k_mirror_packed(DevSys s, const double *__restrict__ xyz)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < s.n; i += gridDim.x * blockDim.x) {
        const double x = xyz[i], y = xyz[s.n + i], z = xyz[2 * s.n + i];
        s.x[i] = x;
        s.y[i] = y;
        s.z[i] = z;
        const int qb = s.ipos[i].w;
        s.ipos[i] = mcx_fixed4(s, x, y, z, __int_as_float(qb));
    }
}

hipError_t launch_mc_wide(const DevSys &s, const DevTerms &t, const DevMoves &mv, const MCRunArgs &a,
                          const WideArgs &w, hipStream_t st)
{
    DevSys s_ = s;
    DevTerms t_ = t;
    DevMoves mv_ = mv;
    MCRunArgs a_ = a;
    WideArgs w_ = w;
    bool fb = false;
    for (int k = 0; k < mv.ntypes; k++) fb = fb || mv.type[k] == MCX_MV_FBIAS_DEV;
    if (fb && s.pbtype == MCX_PBC_CUBIC && s.cutmode == MCX_CUT_SWITCH)
        hipLaunchKernelGGL((k_mc_wide<MCX_PBC_CUBIC, MCX_CUT_SWITCH, true>), dim3(w.nwg), dim3(kMCT), 0, st, s_, t_, mv_,
                           a_, w_);
    else if (fb)
        hipLaunchKernelGGL((k_mc_wide<-1, -1, true>), dim3(w.nwg), dim3(kMCT), 0, st, s_, t_, mv_, a_, w_);
    else if (s.pbtype == MCX_PBC_CUBIC && s.cutmode == MCX_CUT_SWITCH)
        hipLaunchKernelGGL((k_mc_wide<MCX_PBC_CUBIC, MCX_CUT_SWITCH>), dim3(w.nwg), dim3(kMCT), 0, st, s_, t_, mv_, a_, w_);
    else if (s.pbtype == MCX_PBC_CUBIC && s.cutmode == MCX_CUT_FULL)
        hipLaunchKernelGGL((k_mc_wide<MCX_PBC_CUBIC, MCX_CUT_FULL>), dim3(w.nwg), dim3(kMCT), 0, st, s_, t_, mv_, a_, w_);
    else
        hipLaunchKernelGGL((k_mc_wide<-1, -1>), dim3(w.nwg), dim3(kMCT), 0, st, s_, t_, mv_, a_, w_);
    return hipGetLastError();
}

int mc_wide_max_workgroups(int device)
{
    // one workgroup of kMCT threads per CU at most (LDS and registers), so
    // the co-resident limit is the occupancy times the CU count
    int ncu = 0, per = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, reinterpret_cast<const void *>(&k_mc_wide<-1, -1>),
                                                     kMCT, 0) != hipSuccess)
        return 0;
    return ncu * per;
}

void launch_mc_chunk(const DevSys &s, const DevTerms &t, const DevMoves &mv, const MCRunArgs &a, hipStream_t st)
{
    // the water and argon boxes (cubic, switched) and cubic all-pair runs get
    // their own instances; other systems run the generic one
    if (a.spec) {
        if (s.pbtype == MCX_PBC_CUBIC && s.cutmode == MCX_CUT_SWITCH)
            hipLaunchKernelGGL((k_mc_spec<MCX_PBC_CUBIC, MCX_CUT_SWITCH>), dim3(1), dim3(kMCT), 0, st, s, t, mv, a);
        else if (s.pbtype == MCX_PBC_CUBIC && s.cutmode == MCX_CUT_FULL)
            hipLaunchKernelGGL((k_mc_spec<MCX_PBC_CUBIC, MCX_CUT_FULL>), dim3(1), dim3(kMCT), 0, st, s, t, mv, a);
        else
            hipLaunchKernelGGL((k_mc_spec<-1, -1>), dim3(1), dim3(kMCT), 0, st, s, t, mv, a);
    } else if (s.pbtype == MCX_PBC_CUBIC && s.cutmode == MCX_CUT_SWITCH)
        hipLaunchKernelGGL((k_mc_chunk<MCX_PBC_CUBIC, MCX_CUT_SWITCH>), dim3(1), dim3(kMCT), 0, st, s, t, mv, a);
    else if (s.pbtype == MCX_PBC_CUBIC && s.cutmode == MCX_CUT_FULL)
        hipLaunchKernelGGL((k_mc_chunk<MCX_PBC_CUBIC, MCX_CUT_FULL>), dim3(1), dim3(kMCT), 0, st, s, t, mv, a);
    else
        hipLaunchKernelGGL((k_mc_chunk<-1, -1>), dim3(1), dim3(kMCT), 0, st, s, t, mv, a);
}

void launch_mirror_from_coords(const DevSys &s, hipStream_t st)
{
    hipLaunchKernelGGL(k_mirror, dim3((s.n + 255) / 256), dim3(256), 0, st, s);
}

void launch_mirror_from_packed(const DevSys &s, const double *xyz, hipStream_t st)
{
    hipLaunchKernelGGL(k_mirror_packed, dim3((s.n + 255) / 256), dim3(256), 0, st, s, xyz);
}
