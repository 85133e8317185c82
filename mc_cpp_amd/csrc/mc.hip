// mc.hip -- per-move energy change and the Metropolis loop on the device.
//
// MC_metropolis::run (src/MC_metropolis.cpp:45-173) forms dE as the
// difference of two full-system energies (:94,122-123).  Here one
// workgroup evaluates only the terms that touch the moved atoms:
//   * nonbonded pairs from the moved atoms' full list rows (SWITCH) or all
//     atoms (FULL), with the pair arithmetic in FP32 written in difference
//     form -- r_new^2 - r_old^2 = delta.(2 d_old + delta), and 1/r, t^3,
//     t^6-t^3 differences factored through it -- so that a small dE is not
//     the cancellation of two large FP32 energies; the switching polynomial
//     and all sums are FP64;
//   * bonded and 1-4 terms through a per-atom adjacency list, FP64.
// The move itself (src/MC_metropolis.cpp:84-119; include/Move.hpp:42-183),
// the draws (src/MC.cpp:127-156: std::mt19937 + generate_canonical, or a
// Philox4x32-10 stream), the test of apply_criterion (:175-204) and the
// dmax adaptation (src/MC.cpp:165-173) all run in the same workgroup, so a
// chunk of moves is one launch with no host round trip.
#include "terms.h"

#include <cfloat>

namespace {

constexpr int kMCT = 512;
constexpr int kMCW = kMCT / 64;

struct MT {
    uint32_t mt[624];
    int idx;
};

__device__ uint32_t mt_next(MT &g)
{
    if (g.idx >= 624) {
        for (int i = 0; i < 624; i++) {
            const uint32_t y = (g.mt[i] & 0x80000000u) | (g.mt[(i + 1) % 624] & 0x7fffffffu);
            uint32_t v = g.mt[(i + 397) % 624] ^ (y >> 1);
            if (y & 1u) v ^= 0x9908b0dfu;
            g.mt[i] = v;
        }
        g.idx = 0;
    }
    uint32_t y = g.mt[g.idx++];
    y ^= (y >> 11);
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= (y >> 18);
    return y;
}

// libstdc++ generate_canonical<double,53> over two 32-bit outputs
__device__ double mt_canonical(MT &g)
{
    double sum = (double)mt_next(g);
    sum += (double)mt_next(g) * 4294967296.0;
    double r = sum / 18446744073709551616.0;
    if (r >= 1.0) r = 0.99999999999999989;
    return r;
}

__device__ uint4 philox(uint4 c, uint2 k)
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

struct Rng {
    MT *mt;           // replay stream (LDS) or null
    uint2 key;        // Philox key
    uint32_t chain;   // Philox stream id
    uint64_t ctr;     // Philox block counter
    uint4 buf;
    int nbuf;
    __device__ double u01()
    {
        if (mt) return mt_canonical(*mt);
        if (nbuf == 0) {
            buf = philox(make_uint4((uint32_t)ctr, (uint32_t)(ctr >> 32), chain, 0x6d63u), key);
            ctr++;
            nbuf = 2;
        }
        uint32_t a, b;
        if (nbuf == 2) { a = buf.x; b = buf.y; } else { a = buf.z; b = buf.w; }
        nbuf--;
        return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) / 9007199254740992.0;
    }
};

__device__ __forceinline__ double wsum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Shared state of one move.
struct MoveLDS {
    int first, last, nm, total, acc, type, inst;
    double r[3], R[3][3], piv[3];
    double tx[kMaxMoveAtoms], ty[kMaxMoveAtoms], tz[kMaxMoveAtoms];
    double dx[kMaxMoveAtoms], dy[kMaxMoveAtoms], dz[kMaxMoveAtoms];   // unwrapped displacement of each moved atom
    int pref[kMaxMoveAtoms + 1];
    double red[3][kMCW];
    double de;
};

// dE contribution of pair (i, j) with i moved (index a) and j either moved
// (index b >= 0, new coordinates tx/ty/tz[b]) or fixed (b < 0).
// Adds q_i-scaled elec_raw and e_i-scaled vdw_raw.
__device__ void pair_delta(const DevSys &s, const MoveLDS &m, int i, int a, int j, int b, double &dec, double &dev)
{
    const float qi = __int_as_float(s.ipos[i].w), qj = __int_as_float(s.ipos[j].w);
    const float2 ei = s.es[i];
    const float2 ej = s.es[j];
    double o[3];
    mcx_disp_A(s, i, j, o);                       // old displacement x_j - x_i
    double g[3] = {-m.dx[a], -m.dy[a], -m.dz[a]};  // change of the displacement
    if (b >= 0) { g[0] += m.dx[b]; g[1] += m.dy[b]; g[2] += m.dz[b]; }
    const double nx = o[0] + g[0], ny = o[1] + g[1], nz = o[2] + g[2];
    const bool full = s.cutmode == 0;
    const float sij = ei.y + ej.y;
    bool consistent = true;
    if (s.pbtype != 0)
        consistent = fabs(nx) < 0.5 * s.L[0] && fabs(ny) < 0.5 * s.L[1] && fabs(nz) < 0.5 * s.L[2];
    double dc, dv;
    if (consistent) {
        const double r2o = fma(o[2], o[2], fma(o[1], o[1], o[0] * o[0]));
        // r_new^2 - r_old^2 = g.(2 o + g), free of cancellation
        const double D = fma(g[2], fma(2.0, o[2], g[2]), fma(g[1], fma(2.0, o[1], g[1]), g[0] * fma(2.0, o[0], g[0])));
        const double r2n = r2o + D;
        const bool ino = full || r2o <= s.ctoff2, inn = full || r2n <= s.ctoff2;
        if (!ino && !inn) return;
        float rio32, rin32;
        const double rio = mcx_rsqrt(r2o, rio32), rin = mcx_rsqrt(r2n, rin32);
        const float s2 = sij * sij;
        const float to = s2 * rio32 * rio32, tn = s2 * rin32 * rin32;
        const float t3o = to * to * to, t3n = tn * tn * tn;
        const double co = (double)qj * rio, cn = (double)qj * rin;
        const float vo = ej.x * fmaf(t3o, t3o, -t3o), vn = ej.x * fmaf(t3n, t3n, -t3n);
        const double So = full ? 1.0 : mcx_switch(s, r2o), Sn = full ? 1.0 : mcx_switch(s, r2n);
        if (ino && inn) {
            // 1/r_n - 1/r_o = -D / (r_o r_n (r_o + r_n)); t^3 and t^6-t^3
            // differences factored through t_n - t_o = -s^2 D / (r_o^2 r_n^2)
            const double dri = -D * rio * rin / fma(r2o, rio, r2n * rin);
            const float Df = (float)D;
            const float dt = -s2 * Df * (rio32 * rio32) * (rin32 * rin32);
            const float dt3 = dt * fmaf(tn, tn, fmaf(tn, to, to * to));
            const float dL = dt3 * (t3n + t3o - 1.0f);
            const double dS = Sn - So;
            dc = Sn * ((double)qj * dri) + co * dS;
            dv = Sn * (double)(ej.x * dL) + (double)vo * dS;
        } else if (inn) {
            dc = Sn * cn;
            dv = Sn * (double)vn;
        } else {
            dc = -So * co;
            dv = -So * (double)vo;
        }
    } else {
        // image changed between the old and new displacement: evaluate both
        double xjn, yjn, zjn;
        if (b >= 0) { xjn = m.tx[b]; yjn = m.ty[b]; zjn = m.tz[b]; }
        else { xjn = s.x[j]; yjn = s.y[j]; zjn = s.z[j]; }
        double dn[3] = {xjn - m.tx[a], yjn - m.ty[a], zjn - m.tz[a]};
        mcx_mimg64(s, dn);
        double r2o, r2n, co, cn;
        float vo, vn;
        mcx_pair(o[0], o[1], o[2], qj, ej.x, sij, r2o, co, vo);
        mcx_pair(dn[0], dn[1], dn[2], qj, ej.x, sij, r2n, cn, vn);
        double So = 1.0, Sn = 1.0;
        if (!full) {
            So = r2o <= s.ctoff2 ? mcx_switch(s, r2o) : 0.0;
            Sn = r2n <= s.ctoff2 ? mcx_switch(s, r2n) : 0.0;
        }
        dc = Sn * cn - So * co;
        dv = Sn * (double)vn - So * (double)vo;
    }
    dec += (double)qi * dc;
    dev += (double)ei.x * dv;
}

__device__ __forceinline__ bool excluded(const DevSys &s, int i, int j)
{
    const int lo = i < j ? i : j, hi = i < j ? j : i;
    for (int k = s.excl_off[lo]; k < s.excl_off[lo + 1]; k++)
        if (s.excl_idx[k] == hi) return true;
    return false;
}

// dE of the move described in m (trial coordinates and displacements set).
// Returns the total in m.de (valid after the final barrier).
__device__ void move_delta(const DevSys &s, const DevTerms &t, const DevMoves &mv, MoveLDS &m)
{
    const int tid = threadIdx.x;
    if (tid == 0) {
        int acc = 0;
        for (int a = 0; a < m.nm; a++) {
            m.pref[a] = acc;
            acc += s.cutmode == 0 ? s.n : s.cnt[m.first + a];
        }
        m.pref[m.nm] = acc;
        m.total = acc;
    }
    __syncthreads();
    double dec = 0.0, dev = 0.0, deb = 0.0;
    for (int e = tid; e < m.total; e += kMCT) {
        int a = 0;
        while (m.pref[a + 1] <= e) a++;
        const int k = e - m.pref[a];
        const int i = m.first + a;
        int j;
        if (s.cutmode == 0) {
            j = k;
            if (j == i) continue;
            if (excluded(s, i, j)) continue;
        } else {
            j = s.nbr[(size_t)i * s.cap + k];
        }
        int b = -1;
        if (j >= m.first && j <= m.last) {
            if (j < i) continue;     // pair of two moved atoms: counted once
            b = j - m.first;
        }
        pair_delta(s, m, i, a, j, b, dec, dev);
    }
    // bonded and 1-4 terms touching the moved atoms, each counted at its
    // lowest moved atom
    const CoordFetch fo{s.x, s.y, s.z, nullptr, nullptr, nullptr, 0, -1};
    const CoordFetch fn{s.x, s.y, s.z, m.tx, m.ty, m.tz, m.first, m.last};
    const int a0 = mv.adj_off[m.first], a1 = mv.adj_off[m.last + 1];
    for (int e = a0 + tid; e < a1; e += kMCT) {
        const int g = mv.adj_term[e];
        int owner = m.first;
        while (mv.adj_off[owner + 1] <= e) owner++;
        int kind = 0;
        while (g >= t.start[kind + 1]) kind++;
        const int k = g - t.start[kind];
        const int4 id = t.idx[kind][k];
        int lowest = 0x7fffffff;
        const int ids[4] = {id.x, id.y, id.z, id.w};
        for (int q = 0; q < 4; q++)
            if (ids[q] >= m.first && ids[q] <= m.last && ids[q] < lowest) lowest = ids[q];
        if (lowest != owner) continue;
        double e0, c0, v0, e1, c1, v1;
        mcx_term_energy(s, t, kind, k, fo, e0, c0, v0);
        mcx_term_energy(s, t, kind, k, fn, e1, c1, v1);
        deb += e1 - e0;
        dec += c1 - c0;
        dev += v1 - v0;
    }
    dec = wsum(dec);
    dev = wsum(dev);
    deb = wsum(deb);
    const int w = tid >> 6;
    if ((tid & 63) == 0) { m.red[0][w] = dec; m.red[1][w] = dev; m.red[2][w] = deb; }
    __syncthreads();
    if (tid == 0) {
        double c = 0.0, v = 0.0, bb = 0.0;
        for (int k = 0; k < kMCW; k++) { c += m.red[0][k]; v += m.red[1][k]; bb += m.red[2][k]; }
        m.de = MCX_CHGCHARMM * MCX_KCALTOIU * c + 4.0 * v + bb;
    }
    __syncthreads();
}

// find_rot_mat, include/Move.hpp:77-118
__device__ void rot_mat(double mat[3][3], double vec[3], double angle)
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

// Trial coordinates of the moved atoms (threads 0..nm-1).
__device__ void make_trial(const DevSys &s, MoveLDS &m)
{
    const int a = threadIdx.x;
    if (a < m.nm) {
        const int i = m.first + a;
        const double x = s.x[i], y = s.y[i], z = s.z[i];
        double ux, uy, uz;
        if (m.type == MCX_MV_TRN_DEV) {
            ux = x + m.r[0]; uy = y + m.r[1]; uz = z + m.r[2];
        } else {
            const double px = x - m.piv[0], py = y - m.piv[1], pz = z - m.piv[2];
            ux = m.piv[0] + m.R[0][0] * px + m.R[1][0] * py + m.R[2][0] * pz;
            uy = m.piv[1] + m.R[0][1] * px + m.R[1][1] * py + m.R[2][1] * pz;
            uz = m.piv[2] + m.R[0][2] * px + m.R[1][2] * py + m.R[2][2] * pz;
        }
        m.dx[a] = ux - x; m.dy[a] = uy - y; m.dz[a] = uz - z;
        m.tx[a] = wrap1(s, ux, 0); m.ty[a] = wrap1(s, uy, 1); m.tz[a] = wrap1(s, uz, 2);
    }
    __syncthreads();
}

__global__ __launch_bounds__(kMCT) void k_mc_chunk(DevSys s, DevTerms t, DevMoves mv, MCRunArgs arg)
{
    __shared__ MoveLDS m;
    __shared__ MT mt;
    __shared__ double dmax[kMaxMoveTypes];
    __shared__ int acc_tmp[kMaxMoveTypes];
    __shared__ Rng rng;
    const int tid = threadIdx.x;
    if (tid == 0) {
        for (int k = 0; k < mv.ntypes; k++) { dmax[k] = arg.dmax[k]; acc_tmp[k] = arg.acc_tmp[k]; }
        rng.mt = nullptr;
        rng.key = make_uint2((uint32_t)arg.seed, (uint32_t)(arg.seed >> 32));
        rng.chain = 0;
        rng.nbuf = 0;
        rng.ctr = ((uint64_t)arg.mt_state[624] << 32) | arg.mt_state[625];
        if (arg.rng_mode == 1) {
            rng.mt = &mt;
            mt.idx = (int)arg.mt_state[624];
        }
    }
    if (arg.rng_mode == 1)
        for (int k = tid; k < 624; k += kMCT) mt.mt[k] = arg.mt_state[k];
    __syncthreads();

    for (int step = 0; step < arg.nsteps; step++) {
        const int st = arg.step0 + step + 1;
        if (tid == 0) {
            const int tp = (int)((double)mv.ntypes * rng.u01());
            const int in = (int)((double)mv.ninst[tp] * rng.u01());
            const int g = mv.off[tp] + in;
            m.type = mv.type[tp];
            m.inst = in;
            m.first = mv.first[g];
            m.last = mv.last[g];
            m.nm = m.last - m.first + 1;
            double r[3];
            do {
                r[0] = rng.u01() * 2.0 + (-1.0);
                r[1] = rng.u01() * 2.0 + (-1.0);
                r[2] = rng.u01() * 2.0 + (-1.0);
            } while ((r[0] * r[0] + r[1] * r[1] + r[2] * r[2]) > 1.0);
            if (m.type == MCX_MV_TRN_DEV) {
                m.r[0] = r[0] * dmax[tp]; m.r[1] = r[1] * dmax[tp]; m.r[2] = r[2] * dmax[tp];
            } else {
                const double rang = dmax[tp] * (rng.u01() * 2.0 + (-1.0));
                rot_mat(m.R, r, rang);
                const int pv = mv.pivot[g];
                if (pv >= 0) { m.piv[0] = s.x[pv]; m.piv[1] = s.y[pv]; m.piv[2] = s.z[pv]; }
                else {
                    double mt_ = 0.0, c0 = 0.0, c1 = 0.0, c2 = 0.0;
                    for (int i = m.first; i <= m.last; i++) {
                        const double ms = mv.mass[i];
                        mt_ += ms; c0 += ms * s.x[i]; c1 += ms * s.y[i]; c2 += ms * s.z[i];
                    }
                    m.piv[0] = c0 / mt_; m.piv[1] = c1 / mt_; m.piv[2] = c2 / mt_;
                }
            }
            arg.trials[tp]++;
            m.acc = tp;   // remember the type until the test
        }
        __syncthreads();
        make_trial(s, m);
        move_delta(s, t, mv, m);
        if (tid == 0) {
            const int tp = m.acc;
            const double de = m.de;
            int acc;
            if (de <= DBL_EPSILON) acc = 1;
            else {
                const double alpha = rng.u01();
                acc = alpha < exp(-arg.beta * de);
            }
            if (acc) { arg.etot[0] += de; arg.accepted[tp]++; acc_tmp[tp]++; }
            const int o = st - 1;
            if (arg.acc_out) arg.acc_out[o] = (unsigned char)acc;
            if (arg.de_out) arg.de_out[o] = de;
            if (arg.mvtyp_out) arg.mvtyp_out[o] = tp;
            if (arg.mvinst_out) arg.mvinst_out[o] = m.inst;
            for (int k = 0; k < mv.ntypes; k++) {
                if (mv.each[k] != 0 && (st % mv.each[k]) == 0) {
                    const double ratio = (double)acc_tmp[k] / (double)mv.each[k];
                    if (ratio > mv.target[k] / 100) dmax[k] *= 1.10;
                    else dmax[k] *= 0.9;
                    acc_tmp[k] = 0;
                }
            }
            m.acc = acc;
        }
        __syncthreads();
        if (m.acc && tid < m.nm) {
            const int i = m.first + tid;
            s.x[i] = m.tx[tid]; s.y[i] = m.ty[tid]; s.z[i] = m.tz[tid];
            const float q = __int_as_float(s.ipos[i].w);
            s.ipos[i] = mcx_fixed4(s, m.tx[tid], m.ty[tid], m.tz[tid], q);
        }
        __syncthreads();
    }
    if (tid == 0) {
        for (int k = 0; k < mv.ntypes; k++) { arg.dmax[k] = dmax[k]; arg.acc_tmp[k] = acc_tmp[k]; }
        if (arg.rng_mode == 1) arg.mt_state[624] = (uint32_t)mt.idx;
        else { arg.mt_state[624] = (uint32_t)(rng.ctr >> 32); arg.mt_state[625] = (uint32_t)rng.ctr; }
    }
    if (arg.rng_mode == 1)
        for (int k = tid; k < 624; k += kMCT) arg.mt_state[k] = mt.mt[k];
}

// one-shot dE for caller-supplied trial coordinates
__global__ __launch_bounds__(kMCT) void k_delta(DevSys s, DevTerms t, DevMoves mv, int first, int last,
                                                const double *__restrict__ nxyz, double *out)
{
    __shared__ MoveLDS m;
    const int nm = last - first + 1;
    if (threadIdx.x == 0) { m.first = first; m.last = last; m.nm = nm; }
    if (threadIdx.x < nm) {
        const int a = threadIdx.x, i = first + a;
        const double nx = nxyz[3 * a], ny = nxyz[3 * a + 1], nz = nxyz[3 * a + 2];
        double d[3] = {nx - s.x[i], ny - s.y[i], nz - s.z[i]};
        mcx_pbc(s, d);
        m.dx[a] = d[0]; m.dy[a] = d[1]; m.dz[a] = d[2];
        m.tx[a] = nx; m.ty[a] = ny; m.tz[a] = nz;
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

void launch_mc_chunk(const DevSys &s, const DevTerms &t, const DevMoves &mv, const MCRunArgs &a, hipStream_t st)
{
    hipLaunchKernelGGL(k_mc_chunk, dim3(1), dim3(kMCT), 0, st, s, t, mv, a);
}
