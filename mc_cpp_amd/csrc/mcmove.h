// mcmove.h -- device code shared by the single-chain Metropolis kernel and
// the batched replica kernel: the random draws (std::mt19937 replay or
// Philox4x32-10), the rigid-body move of include/Move.hpp and the change of
// one pair's nonbonded energy under a move.
#pragma once

#include "terms.h"

#include <cfloat>

namespace {

// ---------------------------------------------------------------------------
// Random draws.  Wavefront 0 of a workgroup runs the serial part of every
// move with all 64 lanes in step; it keeps a batch of 32 draws in registers
// (lane l holds draw l) and reads the next one with a wave-uniform shuffle.
// A batch is refilled in parallel: 64 MT19937 outputs (twist done by the
// wavefront in four dependency-free phases) or 16 Philox4x32-10 blocks.

constexpr int kMTN = 624;

__device__ __forceinline__ uint32_t mt_temper(uint32_t y)
{
    y ^= (y >> 11);
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= (y >> 18);
    return y;
}

__device__ __forceinline__ uint32_t mt_new(uint32_t cur, uint32_t nxt, uint32_t far)
{
    const uint32_t y = (cur & 0x80000000u) | (nxt & 0x7fffffffu);
    uint32_t v = far ^ (y >> 1);
    if (y & 1u) v ^= 0x9908b0dfu;
    return v;
}

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// MT19937 state regeneration by one wavefront.  New mt[i] depends on old
// mt[i+1] and on mt[i+397] (old for i < 227, new for i >= 227), so the
// ranges [0,227), [227,454), [454,623) and {623} are each free of
// intra-phase dependencies; every phase reads all its inputs before writing.
__device__ __forceinline__ void mt_twist_wave(uint32_t *mt, int lane)
{
    const int lo[4] = {0, 227, 454, 623};
    const int hi[4] = {227, 454, 623, 624};
    for (int ph = 0; ph < 4; ph++) {
        uint32_t v[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int i = lo[ph] + r * 64 + lane;
            if (i < hi[ph]) v[r] = mt_new(mt[i], mt[(i + 1) % kMTN], mt[(i + 397) % kMTN]);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int i = lo[ph] + r * 64 + lane;
            if (i < hi[ph]) mt[i] = v[r];
        }
        wave_sync();
    }
}

__device__ __forceinline__ double canonical2(uint32_t lo, uint32_t hi)
{
    // libstdc++ generate_canonical<double,53> over two 32-bit outputs
    double sum = (double)lo;
    sum += (double)hi * 4294967296.0;
    double r = sum / 18446744073709551616.0;
    if (r >= 1.0) r = 0.99999999999999989;
    return r;
}

__device__ __forceinline__ uint4 philox(uint4 c, uint2 k)
{
    for (int r = 0; r < 10; r++) {
        const uint32_t hi0 = __umulhi(0xD2511F53u, c.x), lo0 = 0xD2511F53u * c.x;
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, c.z), lo1 = 0xCD9E8D57u * c.z;
        c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
        k.x += 0x9E3779B9u;
        k.y += 0xBB67AE85u;
    }
    return c;
}

// 53-bit uniform in [0,1) from two Philox words
__device__ __forceinline__ double philox_u53(uint32_t a, uint32_t b)
{
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) / 9007199254740992.0;
}

// The saved draw state has the kRngWords layout of mcx_internal.h.
struct WaveDraws {
    double v;          // this lane's draw of the current batch
    int cur;           // next draw (uniform)
    int mt_idx;        // MT19937 output index (uniform)
    uint64_t ctr;      // Philox block counter (uniform)
    bool replay;
    uint2 key;
    uint32_t chain;    // Philox stream
    uint32_t *mt;      // LDS state

    __device__ __forceinline__ void load(const uint32_t *st, int lane)
    {
        mt_idx = (int)st[624];
        ctr = ((uint64_t)st[625] << 32) | st[626];
        cur = (int)st[627];
        v = lane < 32 ? __hiloint2double((int)st[628 + 2 * lane + 1], (int)st[628 + 2 * lane]) : 0.0;
    }
    __device__ __forceinline__ void save(uint32_t *st, int lane) const
    {
        if (lane == 0) {
            st[624] = (uint32_t)mt_idx;
            st[625] = (uint32_t)(ctr >> 32);
            st[626] = (uint32_t)ctr;
            st[627] = (uint32_t)cur;
        }
        if (lane < 32) {
            st[628 + 2 * lane] = (uint32_t)__double2loint(v);
            st[628 + 2 * lane + 1] = (uint32_t)__double2hiint(v);
        }
    }
    __device__ __forceinline__ void refill(int lane)
    {
        if (replay) {
            const int p = mt_idx + lane;
            uint32_t y = p < kMTN ? mt[p] : 0u;
            if (mt_idx + 64 > kMTN) {
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                mt_twist_wave(mt, lane);
                if (p >= kMTN) y = mt[p - kMTN];
            }
            // a batch ending exactly at 624 leaves the index at 624 (twist due
            // before the next output); only a batch that crossed it wraps
            mt_idx = mt_idx + 64 > kMTN ? mt_idx + 64 - kMTN : mt_idx + 64;
            const uint32_t out = mt_temper(y);
            const uint32_t lo = __shfl(out, (2 * lane) & 63, 64), hi = __shfl(out, (2 * lane + 1) & 63, 64);
            v = canonical2(lo, hi);
        } else {
            uint4 w = make_uint4(0, 0, 0, 0);
            if (lane < 16) w = philox(make_uint4((uint32_t)(ctr + lane), (uint32_t)((ctr + lane) >> 32), chain, 0x6d63u), key);
            ctr += 16;
            const int src = (lane >> 1) & 15;
            const uint32_t x0 = __shfl(w.x, src, 64), x1 = __shfl(w.y, src, 64);
            const uint32_t x2 = __shfl(w.z, src, 64), x3 = __shfl(w.w, src, 64);
            const uint32_t a = (lane & 1) ? x2 : x0, b = (lane & 1) ? x3 : x1;
            v = philox_u53(a, b);
        }
        cur = 0;
    }
    __device__ __forceinline__ double next(int lane)
    {
        if (cur == 32) refill(lane);
        const double u = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), cur),
                                          __builtin_amdgcn_readlane(__double2loint(v), cur));
        cur++;
        return u;
    }
};

// Read-ahead proposals.  MC_metropolis::apply_criterion draws its uniform
// only when dE exceeds DBL_EPSILON (src/MC_metropolis.cpp:180-187), so a
// rejected move always used its proposal draws and one acceptance draw and
// left the state unchanged: the next move's proposal then starts at a draw
// position known in advance.  The chain kernels draw and load it while the
// current move is evaluated, from the two batches below.

// The draw stream with the batch after the current one already generated,
// so that up to 32 draws beyond the current batch can be read ahead.
struct LookDraws {
    WaveDraws d;   // current batch (d.v, d.cur) and the generator
    double v1;     // the next batch

    __device__ __forceinline__ void gen(int lane)
    {
        const double keep = d.v;
        const int c = d.cur;
        d.refill(lane);
        v1 = d.v;
        d.v = keep;
        d.cur = c;
    }
    __device__ __forceinline__ void load(const uint32_t *st, int lane)
    {
        d.load(st, lane);
        if (st[kRngWords - 1])
            v1 = lane < 32 ? __hiloint2double((int)st[kRngAhead + 2 * lane + 1], (int)st[kRngAhead + 2 * lane]) : 0.0;
        else
            gen(lane);
    }
    __device__ __forceinline__ void save(uint32_t *st, int lane) const
    {
        d.save(st, lane);
        if (lane < 32) {
            st[kRngAhead + 2 * lane] = (uint32_t)__double2loint(v1);
            st[kRngAhead + 2 * lane + 1] = (uint32_t)__double2hiint(v1);
        }
        if (lane == 0) st[kRngWords - 1] = 1u;
    }
    __device__ __forceinline__ double next(int lane)
    {
        if (d.cur == 32) {
            d.v = v1;
            d.cur = 0;
            gen(lane);
        }
        const double u = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(d.v), d.cur),
                                          __builtin_amdgcn_readlane(__double2loint(d.v), d.cur));
        d.cur++;
        return u;
    }
};

// Read-ahead cursor over the two batches of a LookDraws (nothing consumed).
// Past their end it returns 0.5 (which ends the unit-ball rejection loop)
// and clears ok: that proposal is dropped.
struct PeekDraws {
    double v0, v1;
    int idx;       // draw index from the start of the current batch
    bool ok;
    __device__ __forceinline__ double next(int)
    {
        double u = 0.5;
        if (idx < 64) {
            const double src = idx < 32 ? v0 : v1;
            const int l = idx & 31;
            u = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(src), l),
                                 __builtin_amdgcn_readlane(__double2loint(src), l));
        } else {
            ok = false;
        }
        idx++;
        return u;
    }
};

// One DPP move of an FP64 value (both halves; lanes the pattern or the
// row mask leaves out receive 0).
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp64(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWS, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWS, 0xf, false);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double wsum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// The same sum by DPP register moves instead of LDS round trips, for code
// that every lane of the wavefront reaches: quad swaps, half-row and row
// mirrors (every lane of a row then holds the row's sum), row_bcast15 and
// row_bcast31 carry the row sums into lane 63, which is broadcast.
__device__ __forceinline__ double wsum_dpp(double v)
{
    v += dpp64<0xB1, 0xf>(v);    // quad_perm [1,0,3,2]
    v += dpp64<0x4E, 0xf>(v);    // quad_perm [2,3,0,1]
    v += dpp64<0x141, 0xf>(v);   // row_half_mirror
    v += dpp64<0x140, 0xf>(v);   // row_mirror
    v += dpp64<0x142, 0xa>(v);   // row_bcast15 into rows 1 and 3
    v += dpp64<0x143, 0xc>(v);   // row_bcast31 into rows 2 and 3
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), 63),
                            __builtin_amdgcn_readlane(__double2loint(v), 63));
}

__device__ __forceinline__ double wmax(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// find_rot_mat, include/Move.hpp:77-118
__device__ __forceinline__ void rot_mat(double mat[3][3], double vec[3], double angle)
{
    const double angleRad = angle * MCX_PI / 180.0;
    const double rn = 1.0 / sqrt(vec[0] * vec[0] + vec[1] * vec[1] + vec[2] * vec[2]);
    vec[0] *= rn; vec[1] *= rn; vec[2] *= rn;
    double T[3][3], T2[3][3];
    T[0][0] = T[1][1] = T[2][2] = 0.0;
    T[2][1] = vec[0]; T[1][2] = -vec[0];
    T[0][2] = vec[1]; T[2][0] = -vec[1];
    T[1][0] = vec[2]; T[0][1] = -vec[2];
    const double v1 = sin(angleRad), v2 = 1.0 - cos(angleRad);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) { mat[i][j] = (i == j ? 1.0 : 0.0); mat[i][j] += T[i][j] * v1; T2[i][j] = 0.0; }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) T2[i][j] += T[i][k] * T[k][j];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) mat[i][j] += T2[i][j] * v2;
}

__device__ __forceinline__ double wrap1(const DevSys &s, double v, int k)
{
    return s.pbtype != 0 ? v - s.L[k] * mcx_ref_rint(s.rL[k] * v) : v;
}

// The system with the box type and cutoff mode made compile-time
// constants (PB, CUT >= 0): the branches on them fold away, and a cubic
// box's three edges become one value.  PB or CUT < 0 keeps the field.
template <int PB, int CUT>
__device__ __forceinline__ DevSys specialise(const DevSys &s0)
{
    DevSys s = s0;
    if (PB >= 0) s.pbtype = PB;
    if (CUT >= 0) s.cutmode = CUT;
    if (PB == MCX_PBC_CUBIC) {
        for (int d = 1; d < 3; d++) {
            s.L[d] = s0.L[0]; s.rL[d] = s0.rL[0]; s.unit[d] = s0.unit[0];
            s.Lf[d] = s0.Lf[0]; s.rLf[d] = s0.rLf[0];
        }
    }
    return s;
}

// Phase A of a move (src/MC_metropolis.cpp:85-116; src/MC.cpp:137-156), run
// by wavefront 0 with its lanes in step, in two parts.  pick_instance draws
// the move type and instance and reads the instance record (atom range,
// pivot, term-record range); pick_displacement draws the rndSphere
// direction (rejection in the unit ball) and, for rotations, the angle and
// the rotation matrix.  A force-bias move draws a Gaussian vector instead
// (Box-Muller; left in r for the drift to be added).
// The draw source D is anything with a wave-uniform next(lane).
template <class D>
__device__ __forceinline__ void pick_instance(const DevMoves &mv, D &dr, int lane, int &tp, int &in,
                                              int4 &rec, int2 &rr)
{
    tp = (int)((double)mv.ntypes * dr.next(lane));
    in = (int)((double)mv.ninst[tp] * dr.next(lane));
    rec = mv.inst[mv.off[tp] + in];
    rr = mv.inst_adj[mv.off[tp] + in];
}

template <class D>
__device__ __forceinline__ void pick_displacement(D &dr, const double *dmax, int tp, int type, int lane,
                                                  double (&r)[3], double (&R)[3][3])
{
    if (type == MCX_MV_FBIAS_DEV) {
        // force bias: three standard normals by Box-Muller from four draws
        const double u1 = dr.next(lane), u2 = dr.next(lane), u3 = dr.next(lane), u4 = dr.next(lane);
        const double a = sqrt(-2.0 * log(1.0 - u1)), b = sqrt(-2.0 * log(1.0 - u3));
        r[0] = a * cos(MCX_TWOPI * u2);
        r[1] = a * sin(MCX_TWOPI * u2);
        r[2] = b * cos(MCX_TWOPI * u4);
    } else {
        do {
            r[0] = dr.next(lane) * 2.0 + (-1.0);
            r[1] = dr.next(lane) * 2.0 + (-1.0);
            r[2] = dr.next(lane) * 2.0 + (-1.0);
        } while ((r[0] * r[0] + r[1] * r[1] + r[2] * r[2]) > 1.0);
    }
    if (type == MCX_MV_TRN_DEV) {
        r[0] = r[0] * dmax[tp]; r[1] = r[1] * dmax[tp]; r[2] = r[2] * dmax[tp];
    } else if (type == MCX_MV_ROT_DEV) {
        const double rang = dmax[tp] * (dr.next(lane) * 2.0 + (-1.0));
        rot_mat(R, r, rang);
    }
}

// Both parts; lane 0 publishes the move in m.
template <class M>
__device__ __forceinline__ void pick_move(const DevMoves &mv, WaveDraws &dr, const double *dmax, M &m, int lane)
{
    int tp, in;
    int4 rec;
    int2 rr;
    pick_instance(mv, dr, lane, tp, in, rec, rr);
    const int type = mv.type[tp];
    double r[3];
    double R[3][3];
    pick_displacement(dr, dmax, tp, type, lane, r, R);
    if (lane == 0) {
        m.type = type; m.inst = in; m.acc = tp;
        m.first = rec.x; m.last = rec.y; m.pivot = rec.z;
        m.ra0 = rr.x; m.ra1 = rr.y;
        m.nm = rec.y - rec.x + 1;
        m.r[0] = r[0]; m.r[1] = r[1]; m.r[2] = r[2];
        if (type == MCX_MV_ROT_DEV)
            for (int a = 0; a < 3; a++) for (int b = 0; b < 3; b++) m.R[a][b] = R[a][b];
    }
}

// MC_metropolis::apply_criterion (src/MC_metropolis.cpp:175-204): a draw is
// consumed only when dE exceeds DBL_EPSILON.
template <class D>
__device__ __forceinline__ int metropolis(D &dr, double beta, double de, int lane)
{
    if (de <= DBL_EPSILON) return 1;
    const double alpha = dr.next(lane);
    return alpha < exp(-beta * de);
}

// Acceptance bookkeeping and the dmax update of MC::adjust_dmax
// (src/MC.cpp:165-173) after step st (1-based).
__device__ __forceinline__ void adapt_dmax(const DevMoves &mv, long long st, int tp, int acc, double *dmax,
                                           int *acc_tmp)
{
    if (acc) acc_tmp[tp]++;
    for (int k = 0; k < mv.ntypes; k++) {
        if (mv.each[k] != 0 && (st % mv.each[k]) == 0) {
            const double ratio = (double)acc_tmp[k] / (double)mv.each[k];
            if (ratio > mv.target[k] / 100) dmax[k] *= 1.10;
            else dmax[k] *= 0.9;
            acc_tmp[k] = 0;
        }
    }
}

// Force on atom i from atom j, FP64: d = x_j - x_i (minimum image), the
// pair energy of mcx_pair_change; adds -dE_pair/dx_i = 2 (dE/dr^2) d to f
// (iu/A).  Nothing is added beyond the cutoff.
__device__ __forceinline__ void mcx_pair_force(const DevSys &s, const double d[3], double qi, double ei, double si,
                                               double qj, double ej, double sj, double f[3])
{
    const double r2 = fma(d[2], d[2], fma(d[1], d[1], d[0] * d[0]));
    const bool full = s.cutmode == 0;
    if (!full && !(r2 <= s.ctoff2)) return;
    float f32;
    const double ri = mcx_rsqrt(r2, f32);
    const double ri2 = ri * ri;
    const double sij = si + sj;
    const double t = sij * sij * ri2;
    const double t3 = t * t * t;
    const double C = MCX_CHGCHARMM * MCX_KCALTOIU;
    const double qq = C * qi * qj, ee = 4.0 * ei * ej;
    const double E0 = qq * ri + ee * fma(t3, t3, -t3);
    const double dE0 = -0.5 * qq * ri * ri2 + ee * fma(-6.0, t3 * t3, 3.0 * t3) * ri2;   // dE0/dr^2
    double S = 1.0, dS = 0.0;
    if (!full && r2 > s.cton2) {
        const double a = s.ctoff2 - r2, b = s.ctoff2 + 2.0 * r2 - 3.0 * s.cton2;
        S = a * a * b * s.switch2;
        dS = (-2.0 * a * b + 2.0 * a * a) * s.switch2;
    }
    const double g = S * dE0 + dS * E0;
    f[0] += 2.0 * g * d[0];
    f[1] += 2.0 * g * d[1];
    f[2] += 2.0 * g * d[2];
}

// Change of one pair's nonbonded sums when atom i and possibly j move.
// (ox, oy, oz) is the old minimum-image displacement x_j - x_i and g its
// change.  The new displacement is o + g, re-imaged when it leaves the half
// box on some axis (tested only when `check` is set: a caller that knows
// |o + g| < L/2 on every axis skips it), so old and new refer to the same
// stored positions.  The pair energy (charges, eps, sig as given, FP64) is
// evaluated at both displacements in FP64 and differenced: a move's dE is a
// small difference of pair terms up to 1e4 times larger, so the FP32
// rounding of a parameter alone (~6e-8) would exceed the 1e-5 bar on dE
// (on gfx950, FP64 FMA issues at the non-packed FP32 rate).  1/r comes
// from the FP32 estimate and one FP64 Newton step (~1e-14).  Returns false
// (and adds nothing) when neither distance is inside the cutoff; otherwise
// c = q_i*dC and v = eps_i*dV, the raw sums before the 332.0716*418.4 and
// 4 factors of src/FField_MDBAS.cpp:420-421.  With lj false (a
// wavefront-uniform flag, set only when ei is zero on every lane of the
// call) the Lennard-Jones part is skipped: v = ei * 0 adds the same zero.
__device__ __forceinline__ bool mcx_pair_change(const DevSys &s, double ox, double oy, double oz, double gx, double gy,
                                                double gz, double qi, double ei, double si, double qj, double ej,
                                                double sj, bool check, double &c, double &v, bool lj = true)
{
    double dn[3] = {ox + gx, oy + gy, oz + gz};
    const bool full = s.cutmode == 0;
    if (check && s.pbtype != 0 &&
        !(fabs(dn[0]) < 0.5 * s.L[0] && fabs(dn[1]) < 0.5 * s.L[1] && fabs(dn[2]) < 0.5 * s.L[2]))
        mcx_mimg64(s, dn);
    const double r2o = fma(oz, oz, fma(oy, oy, ox * ox));
    const double r2n = fma(dn[2], dn[2], fma(dn[1], dn[1], dn[0] * dn[0]));
    const bool ino = full || r2o <= s.ctoff2, inn = full || r2n <= s.ctoff2;
    if (!ino && !inn) return false;
    float f32;
    const double rio = mcx_rsqrt(r2o, f32), rin = mcx_rsqrt(r2n, f32);
    const double So = ino ? (full ? 1.0 : mcx_switch(s, r2o)) : 0.0;
    const double Sn = inn ? (full ? 1.0 : mcx_switch(s, r2n)) : 0.0;
    const double dc = qj * fma(Sn, rin, -(So * rio));
    double dv = 0.0;
    if (lj) {
        const double sij = si + sj;
        const double s2 = sij * sij;
        const double to = s2 * (rio * rio), tn = s2 * (rin * rin);
        const double t3o = to * to * to, t3n = tn * tn * tn;
        dv = ej * fma(Sn, fma(t3n, t3n, -t3n), -(So * fma(t3o, t3o, -t3o)));
    }
    c = qi * dc;
    v = ei * dv;
    return true;
}

}  // namespace
