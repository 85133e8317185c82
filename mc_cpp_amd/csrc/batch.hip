// batch.hip -- many independent Metropolis chains (replicas) in one launch,
// and the replica-exchange step of a temperature ladder.
//
// One workgroup runs one replica's chain for a whole sweep: the launch is
// persistent over the replica's moves, with the same move, draw and
// acceptance code as the single-chain kernel (MC_metropolis::run,
// src/MC_metropolis.cpp:45-204).  At the start the replica's fixed-point
// image (x, y, z as 32-bit fractions of the box plus an atom-type word) is
// copied into LDS -- up to ~9,200 atoms fit in gfx950's 160 KB -- so the
// per-move scan over all atoms reads LDS only.  Candidates are filtered in
// FP32 against the moved set's bounding sphere, compacted per wavefront
// into an LDS queue, and evaluated densely with the exact per-pair change
// of mcx_pair_change (FP64 distance chain and Coulomb, FP32 LJ, FP64 switch
// and sums).
//
// No pair list is kept: a pair counts when its minimum-image distance is
// inside the cutoff.  That equals the reference's list-based energy
// (src/FField_MDBAS.cpp:336-424 over the list of
// src/List_nonBonded.cpp:836-915) whenever no pair has closed by more than
// the list skin since the last rebuild.  Excluded pairs are taken out by
// subtracting the identical per-pair values, and pairs of two moved atoms
// are evaluated once by wavefront 0 together with the bonded terms.
#include "../../include/mcx.h"
#include "mcmove.h"

#include <algorithm>
#include <type_traits>

namespace {

constexpr int kQ = 128;     // entries of a wavefront's candidate queue
constexpr int kBRec = 32;   // term records of a move staged in LDS

// The move as wavefront 0 publishes it for the scan and commits it: two
// copies, the move under evaluation and the one read ahead.
struct BMove {
    int first, last, nm, check;
    float thr;                 // squared radius of the candidate filter (A^2)
    int4 ip[kBatchAtoms];      // old fixed-point positions, type word (index | frozen << 16) in .w
    double gx[kBatchAtoms], gy[kBatchAtoms], gz[kBatchAtoms];   // unwrapped displacements
    double tx[kBatchAtoms], ty[kBatchAtoms], tz[kBatchAtoms];   // trial coordinates (wrapped)
    double4 qe[kBatchAtoms];   // (charge, eps, sig, 0)
};

// What only wavefront 0 reads before it reads the next move ahead (one
// copy), and the per-wavefront partial sums.
struct BStage {
    double ox[kBatchAtoms], oy[kBatchAtoms], oz[kBatchAtoms];   // old coordinates
    double ms[kBatchAtoms];
    double red[16];
    int4 rid[kBRec];           // the move's first kBRec term records
    int rkk[kBRec], rty[kBRec];
    double4 rpar[kBRec];
    int2 pent[kBRec];          // and pair corrections
};

// Wavefront 0's registers of a move.
struct BRegs {
    int tp, first, last, nm, ra0, ra1, pa0, pa1;
    bool rigid;
};

// Per-type move parameters of a replica (the DevMoves fields indexed by
// the move type), held in LDS for the whole launch.
struct MoveTab {
    int ntypes;
    int type[kMaxMoveTypes], ninst[kMaxMoveTypes], off[kMaxMoveTypes], each[kMaxMoveTypes];
    double target[kMaxMoveTypes];
};

template <bool IMG>
struct ImgFetch {
    const int4 *img;           // LDS image
    const int4 *ipos;          // global mirror (image off)
    const int *atype;
    __device__ __forceinline__ int4 operator()(int j) const
    {
        if (IMG) return img[j];
        const int4 p = ipos[j];
        return make_int4(p.x, p.y, p.z, atype[j]);
    }
};

// Cross-lane steps on FP64 values through DPP (both halves moved alike).
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double lane63(double v)
{
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), 63),
                            __builtin_amdgcn_readlane(__double2loint(v), 63));
}

// Sum over the 64 lanes of a wavefront (all lanes active): pair and quad
// swaps, row rotations by 4 and 8 (every lane then holds its row's sum),
// then the row broadcasts 15 and 31 carry rows 0-2 into row 3; the total is
// read from lane 63.
__device__ __forceinline__ double wave_total(double v)
{
    v += dpp_move<0xb1>(v);    // quad_perm [1,0,3,2]
    v += dpp_move<0x4e>(v);    // quad_perm [2,3,0,1]
    v += dpp_move<0x124>(v);   // row_ror:4
    v += dpp_move<0x128>(v);   // row_ror:8
    v += dpp_move<0x142>(v);   // row_bcast:15
    v += dpp_move<0x143>(v);   // row_bcast:31
    return lane63(v);
}

// Maximum of non-negative values over the wavefront, same steps.
__device__ __forceinline__ double wave_max0(double v)
{
    v = fmax(v, dpp_move<0xb1>(v));
    v = fmax(v, dpp_move<0x4e>(v));
    v = fmax(v, dpp_move<0x124>(v));
    v = fmax(v, dpp_move<0x128>(v));
    v = fmax(v, dpp_move<0x142>(v));
    v = fmax(v, dpp_move<0x143>(v));
    return lane63(v);
}

// Pair change of moved atom a (old image p, displacement g, parameters qa)
// against atom j, which does not move.
__device__ __forceinline__ bool pair_moved_fixed(const DevSys &s, const int4 &p, double gx, double gy, double gz,
                                                 const double4 &qa, const int4 &pj, const double4 &P, bool check,
                                                 double &dec, double &dev)
{
    const double ox = (double)mcx_wdiff(pj.x, p.x) * s.unit[0];
    const double oy = (double)mcx_wdiff(pj.y, p.y) * s.unit[1];
    const double oz = (double)mcx_wdiff(pj.z, p.z) * s.unit[2];
    double c, v;
    const bool lj = __ballot(qa.y != 0.0) != 0ull;   // some moved atom here has eps != 0
    if (mcx_pair_change(s, ox, oy, oz, -gx, -gy, -gz, qa.x, qa.y, qa.z, P.x, P.y, P.z, check, c, v, lj)) {
        dec += c;
        dev += v;
        return true;
    }
    return false;
}

// Nonbonded part of dE from wavefronts 1..NT/64-1: every atom outside the
// moved range, U tiles of 64 candidates per wavefront in flight, filtered
// against the moved set's bounding sphere and queued per wavefront; queued
// candidates are evaluated 64 at a time against the NM moved atoms (NM = 0:
// the move's own count, read from LDS per atom).  Returns this lane's
// share of dE in iu.
template <int NT, int U, bool IMG, int NM, typename QT>
__device__ __forceinline__ double scan_pairs(const DevSys &s, const BMove &m, const ImgFetch<IMG> &f,
                                             const double4 *ptab, QT *queue, bool frozen_any, unsigned &npair)
{
    constexpr int NSW = NT / 64 - 1;
    constexpr int NR = NM > 0 ? NM : 1;
    const int lane = threadIdx.x & 63;
    const int w = (threadIdx.x >> 6) - 1;
    QT *q = queue + w * kQ;
    const int nm = NM > 0 ? NM : m.nm;
    const int first = m.first, nsc = s.n - nm;
    const int4 c0 = m.ip[0];
    const float thr = m.thr;
    const bool check = m.check != 0;
    const float ux = (float)s.unit[0], uy = (float)s.unit[1], uz = (float)s.unit[2];
    int4 pa[NR];
    double ga[NR][3];
    double4 qa[NR];
    if (NM > 0) {
#pragma unroll
        for (int a = 0; a < NR; a++) {
            pa[a] = m.ip[a];
            ga[a][0] = m.gx[a]; ga[a][1] = m.gy[a]; ga[a][2] = m.gz[a];
            qa[a] = m.qe[a];
        }
    }
    double dec = 0.0, dev = 0.0;
    auto process = [&](int j) {
        const int4 pj = f(j);
        const double4 P = ptab[pj.w & 0xffff];
        const bool fj = (pj.w >> 16) != 0;
        auto one = [&](const int4 &p, double gx, double gy, double gz, const double4 &qq) {
            if (frozen_any && fj && (p.w >> 16)) return;
            npair += pair_moved_fixed(s, p, gx, gy, gz, qq, pj, P, check, dec, dev);
        };
        if (NM > 0) {
#pragma unroll
            for (int a = 0; a < NR; a++) one(pa[a], ga[a][0], ga[a][1], ga[a][2], qa[a]);
        } else {
            for (int a = 0; a < nm; a++) one(m.ip[a], m.gx[a], m.gy[a], m.gz[a], m.qe[a]);
        }
    };
    int qn = 0;   // queued entries (wave-uniform)
    for (int base = w * 64 * U; base < nsc; base += NSW * 64 * U) {
        int jv[U];
        int4 pj[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int jj = base + u * 64 + lane;
            jv[u] = jj < nsc ? (jj < first ? jj : jj + nm) : -1;
            pj[u] = f(jv[u] >= 0 ? jv[u] : 0);
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const float dx = (float)mcx_wdiff(pj[u].x, c0.x) * ux;
            const float dy = (float)mcx_wdiff(pj[u].y, c0.y) * uy;
            const float dz = (float)mcx_wdiff(pj[u].z, c0.z) * uz;
            const bool pass = jv[u] >= 0 && fmaf(dz, dz, fmaf(dy, dy, dx * dx)) <= thr;
            const uint64_t bal = __ballot(pass);
            if (pass) {
                const int pos = qn + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                                                    __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
                q[pos] = (QT)jv[u];
            }
            qn += __popcll(bal);
            if (qn >= 64) {
                wave_sync();
                const int j = (int)q[qn - 64 + lane];
                wave_sync();
                qn -= 64;
                process(j);
            }
        }
    }
    wave_sync();
    if (lane < qn) process((int)q[lane]);
    return MCX_CHGCHARMM * MCX_KCALTOIU * dec + 4.0 * dev;
}

// Phase A of a move (wavefront 0, lanes in step): the move
// (src/MC_metropolis.cpp:85-116; src/MC.cpp:137-156), the moved atoms and
// the trial coordinates (MOVE_TRN::tr_a_b / MOVE_ROT::apply_rotation,
// include/Move.hpp:42-50,120-143), with no barrier in between, into m and
// sg; the draws come from D (the stream, or a read-ahead cursor).
template <bool IMG, class D>
__device__ __forceinline__ BRegs batch_pick(const DevSys &s, const DevMoves &mv, const MoveTab &tab,
                                            const double *dmax, const double4 *ptab, const ImgFetch<IMG> &f,
                                            D &dr, BMove &m, BStage &sg, int lane)
{
    int tp = 0, first = 0, last = 0, nm = 0, ra0 = 0, ra1 = 0, pa0 = 0, pa1 = 0, inst = 0;
    bool rigid = false;
    tp = (int)((double)tab.ntypes * dr.next(lane));
    inst = (int)((double)tab.ninst[tp] * dr.next(lane));
    const int k = tab.off[tp] + inst;
    const int4 rec = mv.inst[k];
    const int2 rr = mv.inst_adj[k];
    const int2 rp = mv.inst_pairs[k];
    double rv[3];
    do {
        rv[0] = dr.next(lane) * 2.0 + (-1.0);
        rv[1] = dr.next(lane) * 2.0 + (-1.0);
        rv[2] = dr.next(lane) * 2.0 + (-1.0);
    } while ((rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2]) > 1.0);
    const int type = tab.type[tp];
    double Rm[3][3];
    if (type == MCX_MV_TRN_DEV) {
        rv[0] = rv[0] * dmax[tp]; rv[1] = rv[1] * dmax[tp]; rv[2] = rv[2] * dmax[tp];
    } else {
        const double rang = dmax[tp] * (dr.next(lane) * 2.0 + (-1.0));
        rot_mat(Rm, rv, rang);
    }
    first = rec.x; last = rec.y;
    const int pivot = rec.z;
    nm = last - first + 1;
    ra0 = rr.x; ra1 = rr.y;
    pa0 = rp.x; pa1 = rp.y;
    double x = 0.0, y = 0.0, z = 0.0;
    int4 pl = make_int4(0, 0, 0, 0);
    if (lane < nm) {
        const int i = first + lane;
        x = s.x[i]; y = s.y[i]; z = s.z[i];
        pl = f(i);
        sg.ox[lane] = x; sg.oy[lane] = y; sg.oz[lane] = z;
        sg.ms[lane] = (pivot < 0 && mv.mass) ? mv.mass[i] : 1.0;
    }
    if (lane < kBRec && lane < ra1 - ra0) {
        const int e = ra0 + lane;
        sg.rid[lane] = mv.rec_ids[e]; sg.rkk[lane] = mv.rec_kk[e]; sg.rty[lane] = mv.rec_type[e];
        sg.rpar[lane] = mv.rec_par[e];
    }
    if (lane < kBRec && lane < pa1 - pa0) sg.pent[lane] = mv.pair_ent[pa0 + lane];
    double piv[3] = {0.0, 0.0, 0.0};
    if (type != MCX_MV_TRN_DEV) {
        if (pivot >= 0 && (pivot < first || pivot > last)) {
            // a uniform address this wavefront writes with vector
            // stores: atomic loads are never compiled as scalar-cache
            // loads, which would not see those stores
            piv[0] = __hip_atomic_load(s.x + pivot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            piv[1] = __hip_atomic_load(s.y + pivot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            piv[2] = __hip_atomic_load(s.z + pivot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        } else {
            wave_sync();
            if (pivot >= 0) {
                piv[0] = sg.ox[pivot - first]; piv[1] = sg.oy[pivot - first]; piv[2] = sg.oz[pivot - first];
            } else {   // Tools::getCentreOfMass, include/Tools.hpp:95-113
                double mt_ = 0.0, c0 = 0.0, c1 = 0.0, c2 = 0.0;
                for (int b = 0; b < nm; b++) {
                    const double ms = sg.ms[b];
                    mt_ += ms; c0 += ms * sg.ox[b]; c1 += ms * sg.oy[b]; c2 += ms * sg.oz[b];
                }
                piv[0] = c0 / mt_; piv[1] = c1 / mt_; piv[2] = c2 / mt_;
            }
        }
    }
    const int p0x = __builtin_amdgcn_readlane(pl.x, 0), p0y = __builtin_amdgcn_readlane(pl.y, 0);
    const int p0z = __builtin_amdgcn_readlane(pl.z, 0);
    double g2 = 0.0, rad2 = 0.0;
    bool far = false;
    if (lane < nm) {
        double ux, uy, uz;
        if (type == MCX_MV_TRN_DEV) {
            ux = x + rv[0]; uy = y + rv[1]; uz = z + rv[2];
        } else {
            const double px = x - piv[0], py = y - piv[1], pz = z - piv[2];
            const double lq = 0.25 * fmin(s.L[0], fmin(s.L[1], s.L[2]));
            far = s.pbtype != 0 && px * px + py * py + pz * pz >= lq * lq;
            ux = piv[0] + Rm[0][0] * px + Rm[1][0] * py + Rm[2][0] * pz;
            uy = piv[1] + Rm[0][1] * px + Rm[1][1] * py + Rm[2][1] * pz;
            uz = piv[2] + Rm[0][2] * px + Rm[1][2] * py + Rm[2][2] * pz;
        }
        const double ddx = ux - x, ddy = uy - y, ddz = uz - z;
        m.gx[lane] = ddx; m.gy[lane] = ddy; m.gz[lane] = ddz;
        m.tx[lane] = wrap1(s, ux, 0); m.ty[lane] = wrap1(s, uy, 1); m.tz[lane] = wrap1(s, uz, 2);
        m.ip[lane] = pl;
        m.qe[lane] = ptab[pl.w & 0xffff];
        g2 = ddx * ddx + ddy * ddy + ddz * ddz;
        const double rx = (double)mcx_wdiff(pl.x, p0x) * s.unit[0];
        const double ry = (double)mcx_wdiff(pl.y, p0y) * s.unit[1];
        const double rz = (double)mcx_wdiff(pl.z, p0z) * s.unit[2];
        rad2 = rx * rx + ry * ry + rz * rz;
    }
    const double gm = sqrt(wave_max0(g2)), rm = sqrt(wave_max0(rad2));
    // A translation, or a rotation whose atoms all lie within a
    // quarter box of the pivot, moves the set rigidly: distances,
    // angles and torsions inside it change by rounding only
    // (~1e-15 relative), so the terms among its own atoms are
    // skipped.  A rotation of a set the box edge splits (the pivot
    // is applied to wrapped coordinates, include/Move.hpp:120-143)
    // is not rigid and keeps them.
    rigid = __ballot(far) == 0;
    if (lane == 0) {
        // |x_j - x_a| >= |x_j - x_0| - rm and a moves by at most gm, so
        // a candidate farther than rc + rm + gm from atom 0's old
        // position is outside the cutoff before and after the move;
        // inside that sphere, o + g stays within the half box when
        // rc + rm + 2 gm < L/2 on every axis
        const double rc = sqrt(s.ctoff2) + rm + gm;
        const double reach = rc + gm;
        m.first = first; m.last = last; m.nm = nm;
        m.thr = s.cutmode == 0 ? __builtin_huge_valf() : (float)(rc * rc * 1.00001 + 1e-4);
        m.check = s.cutmode == 0 ||
                  !(reach * 1.00001 < 0.5 * s.L[0] && reach * 1.00001 < 0.5 * s.L[1] &&
                    reach * 1.00001 < 0.5 * s.L[2]);
    }
    return BRegs{tp, first, last, nm, ra0, ra1, pa0, pa1, rigid};
}

// Is dmax adapted after (1-based) step st?  (src/MC.cpp:165-173)
__device__ __forceinline__ bool tab_adapts(const MoveTab &tab, long long st)
{
    for (int k = 0; k < tab.ntypes; k++)
        if (tab.each[k] != 0 && (st % tab.each[k]) == 0) return true;
    return false;
}

template <int NT, int U, bool IMG, bool CS>
__global__ __launch_bounds__(NT) void k_batch_sweep(BatchArgs a)
{
    using QT = typename std::conditional<IMG, unsigned short, int>::type;
    constexpr int NW = NT / 64;
    extern __shared__ int4 dyn[];
    __shared__ BMove mb[2];
    __shared__ BStage sg;
    __shared__ int mcur;
    __shared__ MoveTab tab;
    __shared__ uint32_t mt[kMTN];
    __shared__ double4 ptab[kBatchTypes];
    __shared__ double dmax[kMaxMoveTypes];
    __shared__ int acc_tmp[kMaxMoveTypes];
    __shared__ long long n_trial[kMaxMoveTypes], n_acc[kMaxMoveTypes];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const bool w0 = __builtin_amdgcn_readfirstlane(tid) < 64;   // wave-uniform
    const int r = a.order[blockIdx.x];
    const BatchRep &R = a.reps[r];
    // fields used in the pair loops: kept in registers; with CS the box type
    // and cutoff mode are compile-time constants
    const DevSys s = CS ? specialise<MCX_PBC_CUBIC, MCX_CUT_SWITCH>(R.s) : R.s;
    const DevMoves &mv = R.mv;
    int4 *img = dyn;
    QT *queue = reinterpret_cast<QT *>(dyn + (IMG ? a.img_cap : 0));
    const ImgFetch<IMG> f{img, s.ipos, R.atype};
    const bool frozen_any = R.frozen_any != 0;

    // ---- the replica's state: image, tables, draws, counters
    if (IMG)
        for (int k = tid; k < s.n; k += NT) {
            const int4 p = s.ipos[k];
            img[k] = make_int4(p.x, p.y, p.z, R.atype[k]);
        }
    for (int k = tid; k < R.ntab; k += NT) ptab[k] = R.ptab[k];
    const double beta = 1.0 / ((MCX_RBOLTZUI / MCX_KCALTOIU) * a.temps[a.slot_of_rep[r]]);
    LookDraws ld;
    double etot = 0.0;
    long long first_step = 0;
    if (w0) {
        ld.d.replay = R.rng_mode == 1;
        ld.d.mt = mt;
        ld.d.key = make_uint2((uint32_t)R.key, (uint32_t)(R.key >> 32));
        ld.d.chain = R.stream;
        const int nt = mv.ntypes;
        if (lane < nt) {
            dmax[lane] = R.dmax[lane]; acc_tmp[lane] = R.acc_tmp[lane];
            n_trial[lane] = R.trials[lane]; n_acc[lane] = R.accepted[lane];
            tab.type[lane] = mv.type[lane]; tab.ninst[lane] = mv.ninst[lane]; tab.off[lane] = mv.off[lane];
            tab.each[lane] = mv.each[lane]; tab.target[lane] = mv.target[lane];
        }
        if (lane == 0) tab.ntypes = nt;
        etot = R.etot[0];
        first_step = R.step[0];
    }
    if (R.rng_mode == 1)
        for (int k = tid; k < kMTN; k += NT) mt[k] = R.rng[k];
    __syncthreads();
    if (w0) ld.load(R.rng, lane);      // may generate the read-ahead batch from mt
    const int nmoves = a.nmoves > 0 ? a.nmoves : R.sweep_len;
    unsigned npair = 0;     // this lane's in-cutoff pair evaluations (old and new) of the scan

    // The move loop in two copies with the same barriers: wavefront 0 runs
    // the chain (move, trial, bonded terms, Metropolis test, commit) and
    // the other wavefronts the pair scan, so that each role's values stay
    // out of the other's register allocation.
    if (w0) {
        int c = 0;                 // buffer of the move under evaluation
        bool ready = false;        // mb[c] already holds it (read ahead)
        BRegs q{};
        for (int mi = 0; mi < nmoves; mi++) {
            // ---- wavefront 0: phase A, unless the move was read ahead
            if (!ready) q = batch_pick<IMG>(s, mv, tab, dmax, ptab, f, ld, mb[c], sg, lane);
            if (lane == 0) mcur = c;
            __syncthreads();
            BMove &m = mb[c];
            const int tp = q.tp, first = q.first, last = q.last, nm = q.nm;
            const int ra0 = q.ra0, ra1 = q.ra1, pa0 = q.pa0, pa1 = q.pa1;
            const bool rigid = q.rigid;
            // ---- dE: wavefronts 1.. scan the nonbonded pairs; wavefront 0
            // takes the bonded and 1-4 terms, removes excluded pairs the scan
            // counted, and evaluates pairs of two moved atoms
            double part = 0.0;
            double dec = 0.0, dev = 0.0, deb = 0.0;
            // moved atoms' old and trial coordinates from LDS, others from memory
            const CoordFetch fo{s.x, s.y, s.z, sg.ox, sg.oy, sg.oz, first, last};
            const CoordFetch fn{s.x, s.y, s.z, m.tx, m.ty, m.tz, first, last};
            auto term = [&](const int4 &id, int kk, int ty_, const double4 &par) {
                double e0, c0, v0, e1, c1, v1;
                mcx_term_energy_rec(s, kk >> 28, id, ty_, par, fo, e0, c0, v0);
                mcx_term_energy_rec(s, kk >> 28, id, ty_, par, fn, e1, c1, v1);
                deb += e1 - e0;
                dec += c1 - c0;
                dev += v1 - v0;
            };
            auto inside = [&](const int4 &id) {
                return id.x >= first && id.x <= last && id.y >= first && id.y <= last &&
                       (id.z < 0 || (id.z >= first && id.z <= last)) && (id.w < 0 || (id.w >= first && id.w <= last));
            };
            for (int e = ra0 + lane; e < ra1; e += 64) {
                const int k = e - ra0;
                const int4 id = k < kBRec ? sg.rid[k] : mv.rec_ids[e];
                if (rigid && inside(id)) continue;
                if (k < kBRec) term(id, sg.rkk[k], sg.rty[k], sg.rpar[k]);
                else term(id, mv.rec_kk[e], mv.rec_type[e], mv.rec_par[e]);
            }
            // pair corrections: excluded partners outside the moved range,
            // removed with the scan's own values, and pairs of two moved atoms
            for (int e = pa0 + lane; e < pa1; e += 64) {
                const int2 pe = e - pa0 < kBRec ? sg.pent[e - pa0] : mv.pair_ent[e];
                const int ia = pe.x & 0xffff;
                const int4 p = m.ip[ia];
                double c = 0.0, v = 0.0;
                if ((pe.x >> 16) == 0) {
                    const int4 pj = f(pe.y);
                    pair_moved_fixed(s, p, m.gx[ia], m.gy[ia], m.gz[ia], m.qe[ia], pj, ptab[pj.w & 0xffff], true,
                                     c, v);
                    dec -= c;
                    dev -= v;
                } else if (!rigid) {
                    const int ib = pe.y;
                    const int4 pj = m.ip[ib];
                    const double ox = (double)mcx_wdiff(pj.x, p.x) * s.unit[0];
                    const double oy = (double)mcx_wdiff(pj.y, p.y) * s.unit[1];
                    const double oz = (double)mcx_wdiff(pj.z, p.z) * s.unit[2];
                    if (mcx_pair_change(s, ox, oy, oz, m.gx[ib] - m.gx[ia], m.gy[ib] - m.gy[ia], m.gz[ib] - m.gz[ia],
                                        m.qe[ia].x, m.qe[ia].y, m.qe[ia].z, m.qe[ib].x, m.qe[ib].y,
                                        m.qe[ib].z, true, c, v)) {
                        dec += c;
                        dev += v;
                    }
                }
            }
            part = MCX_CHGCHARMM * MCX_KCALTOIU * dec + 4.0 * dev + deb;
            part = wave_total(part);
            // ---- read ahead: the next move into the other buffer, valid if
            // this one is rejected (then its acceptance draw was taken); the
            // staged old coordinates and records of this move are no longer
            // read
            int nd = -1;
            BRegs qn{};
            if (mi + 1 < nmoves && !tab_adapts(tab, first_step + mi + 1)) {
                PeekDraws pk{ld.d.v, ld.v1, ld.d.cur + 1, true};
                const int i0 = pk.idx;
                qn = batch_pick<IMG>(s, mv, tab, dmax, ptab, f, pk, mb[c ^ 1], sg, lane);
                if (pk.ok) nd = pk.idx - i0;
            }
            if (lane == 0) sg.red[tid >> 6] = part;
            __syncthreads();
            // ---- wavefront 0: Metropolis test (apply_criterion,
            // src/MC_metropolis.cpp:175-204), bookkeeping and commit; the next
            // move's barrier publishes the commit
            const double de = wave_total(lane < NW ? sg.red[lane] : 0.0);
            const int acc = metropolis(ld, beta, de, lane);
            if (lane == 0) {
                n_trial[tp]++;
                if (acc) { etot += de; n_acc[tp]++; }
                if (mi < a.rec_len) {
                    const size_t at = (size_t)r * a.rec_len + mi;
                    if (a.acc_out) a.acc_out[at] = (unsigned char)acc;
                    if (a.de_out) a.de_out[at] = de;
                }
                // dmax adaptation of MC::adjust_dmax (src/MC.cpp:165-173)
                const long long st = first_step + mi + 1;
                if (acc) acc_tmp[tp]++;
                for (int k = 0; k < tab.ntypes; k++) {
                    if (tab.each[k] != 0 && (st % tab.each[k]) == 0) {
                        const double ratio = (double)acc_tmp[k] / (double)tab.each[k];
                        if (ratio > tab.target[k] / 100) dmax[k] *= 1.10;
                        else dmax[k] *= 0.9;
                        acc_tmp[k] = 0;
                    }
                }
            }
            if (acc && lane < nm) {
                const int i = first + lane;
                const double tx = m.tx[lane], ty = m.ty[lane], tz = m.tz[lane];
                const int tw = m.ip[lane].w;
                s.x[i] = tx; s.y[i] = ty; s.z[i] = tz;
                const int4 p = mcx_fixed4(s, tx, ty, tz, (float)m.qe[lane].x);
                s.ipos[i] = p;
                if (IMG) img[i] = make_int4(p.x, p.y, p.z, tw);
            }
            if (acc) {
                ready = false;
            } else if (nd >= 0) {
                // the move read ahead is the next one: take its draws
                for (int k = 0; k < nd; k++) (void)ld.next(lane);
                c ^= 1;
                q = qn;
                ready = true;
            } else {
                ready = false;
            }
        }
    } else {
        for (int mi = 0; mi < nmoves; mi++) {
            __syncthreads();              // the move is published
            double part = 0.0;
            const BMove &m = mb[mcur];
            const int n_m = m.nm;
            if (n_m == 3) part = scan_pairs<NT, U, IMG, 3, QT>(s, m, f, ptab, queue, frozen_any, npair);
            else if (n_m == 1) part = scan_pairs<NT, U, IMG, 1, QT>(s, m, f, ptab, queue, frozen_any, npair);
            else part = scan_pairs<NT, U, IMG, 0, QT>(s, m, f, ptab, queue, frozen_any, npair);
            part = wave_total(part);
            if (lane == 0) sg.red[tid >> 6] = part;
            __syncthreads();
        }
    }
    // pair evaluations of the launch: one atomic per wavefront
    {
        const double tw = wave_total((double)npair);
        if (lane == 0 && tw > 0.0) atomicAdd(R.npair, (unsigned long long)tw);
    }
    __syncthreads();
    if (w0) {
        if (lane < tab.ntypes) {
            R.dmax[lane] = dmax[lane]; R.acc_tmp[lane] = acc_tmp[lane];
            R.trials[lane] = n_trial[lane]; R.accepted[lane] = n_acc[lane];
        }
        if (lane == 0) { R.etot[0] = etot; R.step[0] = first_step + nmoves; }
        ld.save(R.rng, lane);
    }
    if (R.rng_mode == 1)
        for (int k = tid; k < kMTN; k += NT) R.rng[k] = mt[k];
}

// Replica exchange between neighbouring ladder slots (k, k+1), k of the
// round's parity: accepted with min(1, exp((b_k - b_k+1)(E_k - E_k+1))),
// b = 1/((0.83144621/418.4) T) as the Metropolis test applies it to E in iu.
// The uniform of pair k in round t is Philox4x32-10 block
// (k, t_lo, t_hi, 0x52455843) under the batch key.
__global__ __launch_bounds__(256) void k_exchange(ExchArgs x)
{
    const unsigned long long round = x.counter[0];
    const int parity = (int)(round & 1ull);
    const uint2 key = make_uint2((uint32_t)x.key, (uint32_t)(x.key >> 32));
    for (int k = parity + 2 * (int)threadIdx.x; k + 1 < x.nrep; k += 2 * (int)blockDim.x) {
        const int ra = x.rep_of_slot[k], rb = x.rep_of_slot[k + 1];
        const double ba = 1.0 / ((MCX_RBOLTZUI / MCX_KCALTOIU) * x.temps[k]);
        const double bb = 1.0 / ((MCX_RBOLTZUI / MCX_KCALTOIU) * x.temps[k + 1]);
        const double delta = (ba - bb) * (x.etot[ra] - x.etot[rb]);
        const uint4 w = philox(make_uint4((uint32_t)k, (uint32_t)round, (uint32_t)(round >> 32), 0x52455843u), key);
        const double u = philox_u53(w.x, w.y);
        const bool acc = delta >= 0.0 || u < exp(delta);
        x.att[k]++;
        if (acc) {
            x.acc[k]++;
            x.rep_of_slot[k] = rb;
            x.rep_of_slot[k + 1] = ra;
            x.slot_of_rep[ra] = k + 1;
            x.slot_of_rep[rb] = k;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) x.counter[0] = round + 1;
}

using BatchKernel = void (*)(BatchArgs);

template <int NT, int U, bool IMG, bool CS>
constexpr BatchKernel kvar()
{
    return &k_batch_sweep<NT, U, IMG, CS>;
}

BatchKernel variant(int threads, int unroll, bool img, bool cs = false)
{
#define MCX_BV(NT, U)                                                                           \
    if (threads == NT && unroll == U)                                                           \
        return img ? (cs ? kvar<NT, U, true, true>() : kvar<NT, U, true, false>())              \
                   : (cs ? kvar<NT, U, false, true>() : kvar<NT, U, false, false>());
    MCX_BV(256, 1) MCX_BV(256, 2) MCX_BV(256, 4)
    MCX_BV(512, 1) MCX_BV(512, 2) MCX_BV(512, 4)
    MCX_BV(768, 1) MCX_BV(768, 2) MCX_BV(768, 4)
    MCX_BV(1024, 1) MCX_BV(1024, 2) MCX_BV(1024, 4)
#undef MCX_BV
    return nullptr;
}

}  // namespace

bool batch_variant_ok(int threads, int unroll, bool img) { return variant(threads, unroll, img) != nullptr; }

size_t batch_static_lds(int threads, int unroll, bool img)
{
    size_t most = 0;
    for (int cs = 0; cs < 2; cs++) {
        BatchKernel k = variant(threads, unroll, img, cs != 0);
        if (!k) throw mcx_error(MCX_ERR_ARG, "no batched kernel for this thread count / unroll");
        hipFuncAttributes at{};
        HIPCHK(hipFuncGetAttributes(&at, reinterpret_cast<const void *>(k)));
        most = std::max(most, (size_t)at.sharedSizeBytes);
    }
    return most;
}

size_t batch_dynamic_lds(int threads, bool img, int img_cap)
{
    return (img ? (size_t)img_cap * sizeof(int4) : 0) +
           (size_t)(threads / 64 - 1) * kQ * (img ? sizeof(unsigned short) : sizeof(int));
}

void batch_prepare(int threads, int unroll, bool img, int img_cap)
{
    for (int cs = 0; cs < 2; cs++) {
        BatchKernel k = variant(threads, unroll, img, cs != 0);
        if (!k) throw mcx_error(MCX_ERR_ARG, "no batched kernel for this thread count / unroll");
        HIPCHK(hipFuncSetAttribute(reinterpret_cast<const void *>(k), hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)batch_dynamic_lds(threads, img, img_cap)));
    }
}

void launch_batch_sweep(const BatchArgs &a, int nrep, int threads, int unroll, bool img, hipStream_t st)
{
    BatchKernel k = variant(threads, unroll, img, a.cs != 0);
    if (!k) throw mcx_error(MCX_ERR_ARG, "no batched kernel for this thread count / unroll");
    hipLaunchKernelGGL(k, dim3(nrep), dim3(threads), batch_dynamic_lds(threads, img, a.img_cap), st, a);
}

void launch_exchange(const ExchArgs &x, hipStream_t st)
{
    hipLaunchKernelGGL(k_exchange, dim3(1), dim3(256), 0, st, x);
}
