// mcx_internal.h -- shared definitions of the gfx950 energy backend.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

// Internal units and constants (include/Constants.hpp:29-42 of mc_cpp).
#define MCX_KCALTOIU 418.4
#define MCX_RBOLTZUI 0.83144621
#define MCX_CHGCHARMM 332.0716
#define MCX_PI 3.141592653589793
#define MCX_TWOPI 6.283185307179586

// Algorithmic work per unit, frozen here so that the benchmark's roofline
// figures use one set of constants (FMA = 2 FLOP, rsqrt = 1, compares and
// rounding = 0):
//   switched list pair : 17 FLOP per list entry + 23 per in-cutoff entry
//   full O(N^2) pair   : 32 FLOP per pair
//   list build bytes   : 24*N + 4*(N+1) + 4*P
#define MCX_FLOP_LIST_ENTRY 17.0
#define MCX_FLOP_LIST_INCUT 23.0
#define MCX_FLOP_FULL_PAIR 32.0

#define MCX_MV_TRN_DEV 1   // move type codes as in include/Move.hpp:33-38
#define MCX_MV_ROT_DEV 2
#define MCX_MV_FBIAS_DEV 3 // force-bias translation (not in the reference)

constexpr int kMaxMoveAtoms = 64;           // atoms per move instance on the device path
#ifndef MCX_FULL_TILE
#define MCX_FULL_TILE 64
#endif
constexpr int kFullTile = MCX_FULL_TILE;   // i and j atoms per tile of the O(N^2) kernel
constexpr int kMaxMoveTypes = 8;
// Saved draw state of a chain (uint32 words): [0,624) MT19937 state, 624 MT
// index, 625/626 Philox block counter (hi, lo), 627 cursor into the draw
// batch, [628,692) the batch (32 doubles), [692,756) the batch after it,
// drawn ahead for read-ahead proposals, and 756 (1: that batch is present).
constexpr int kRngAhead = 692;
constexpr int kRngWords = 757;

// Everything a kernel needs about one system, passed by value.
// Per-atom parameters of the all-pair kernel's j atom, read as one 16-byte
// scalar load: q (the FP32 value of ipos.w, widened to FP64 so that it is
// an FP64 operand as it stands) and eps, sig (esk).
struct alignas(16) TileAtom {
    double q;
    float eps, sig;
};

struct DevSys {
    int n;
    int pbtype;
    double L[3], rL[3];
    float Lf[3], rLf[3];
    double ctoff2, cton2, switch2, cutnb2;
    int cutmode;
    // coordinates: FP64 master copy + a 32-bit fixed-point mirror of the
    // fractional coordinates (x/L_k mod 1 scaled by 2^32) with the charge's
    // FP32 bits in .w.  A wrapping int32 difference of two mirrors is the
    // minimum-image displacement in units of L_k/2^32, exactly.
    double *x, *y, *z;
    int4 *ipos;
    double unit[3];      // L_k / 2^32 (1 when pbtype == NONE)
    double iunit;        // 2^32 / L for cubic boxes (kernel length unit), else 1
    float2 *es;          // (eps, sig) FP32, sig in A
    float2 *esk;         // (eps, sig) FP32, sig in kernel length units
    TileAtom *tq;        // q of ipos.w widened to FP64, and esk: the all-pair kernel's j data
    float2 *es14;        // (eps14, sig14)
    const double *q64, *eps64, *sig64, *eps1464, *sig1464;
    const double4 *qes;  // (charge, eps, sig in A, 0) FP64 per atom: the per-move dE path
    const int *frozen;
    // exclusion rows (visible entries, CSR)
    const int *excl_off, *excl_idx;
    // excluded partners of each atom in both directions (pair (a,b), a<b,
    // is excluded iff b is in the visible row of a), CSR
    const int *xfull_off, *xfull_idx;
    // 1-4 partners of each atom in both directions (one entry per 1-4 pair
    // record), CSR
    const int *p14_off, *p14_idx;
    // neighbour list: full rows (all j != i, ascending) in ELL layout
    int cap;
    int *nbr;            // [n*cap]
    int *cnt;            // [n]
    int *half;           // [n] index of the first j > i in the row
    int *overflow;       // [1]
};

// Bonded terms in a tagged work list for the fused bonded kernel.
enum TermKind : int {
    kBond = 0, kUB = 1, kAngle = 2, kDihe = 3, kImpr = 4, kPair14 = 5, kExclPair = 6, kNumKinds = 7
};

struct DevTerms {
    int nterm[kNumKinds];
    int start[kNumKinds + 1];   // prefix of nterm: global work index ranges
    const int4 *idx[kNumKinds]; // atom indices (unused lanes = -1)
    const int *type[kNumKinds];
    const double4 *par[kNumKinds]; // k, r0/theta0/phi0, beta/mult, unused
};

struct DevMoves {
    int ntypes;
    int type[kMaxMoveTypes];
    int ninst[kMaxMoveTypes];
    int off[kMaxMoveTypes];
    double dmax0[kMaxMoveTypes];
    double target[kMaxMoveTypes];
    int each[kMaxMoveTypes];
    // force-bias drift beta A F per type: F in iu/A times 1/418.4 (the drift
    // in kcal/mol/A, MCX_MV_FBIAS) or times 1 (the Metropolis test's own
    // beta applied to the iu force, MCX_MV_FBIAS_ACC)
    double fb_scale[kMaxMoveTypes];
    const int4 *inst;                  // flat instance table: (first, last, pivot, 0)
    const int2 *inst_adj;              // range of the instance's term records
    const int2 *inst_pairs;            // range of the instance's pair corrections in pair_ent
    // pair corrections of a list-free dE: (a, j) = moved atom a (offset in
    // the range) and an excluded partner j outside it, to subtract; or
    // (a | 1 << 16, b) = a non-excluded pair of moved atoms a < b, to add
    const int2 *pair_ent;
    // term records: every bonded term / 1-4 pair touching an instance, once
    const int4 *rec_ids;
    const int *rec_kk;                 // kind << 28 | index within kind
    const int *rec_type;
    const double4 *rec_par;
    const double *mass;                // for centre-of-mass pivots
    // per-atom adjacency of bonded terms and 1-4 pairs (CSR over atoms):
    // the term's atoms and (kind << 28 | index within kind)
    const int *adj_off;
    const int4 *adj_ids;
    const int *adj_kk;                 // (kind << 28 | index), bit 27.. of index unused
    const int *adj_type;
    const double4 *adj_par;
};

struct HostTerms {
    std::vector<int4> idx[kNumKinds];
    std::vector<int> type[kNumKinds];
    std::vector<double4> par[kNumKinds];
};

#define HIPCHK(expr)                                                                   \
    do {                                                                               \
        hipError_t _e = (expr);                                                        \
        if (_e != hipSuccess) throw mcx_error(MCX_ERR_HIP, std::string(#expr) + ": " + \
                                              hipGetErrorString(_e));                  \
    } while (0)

struct mcx_error {
    int code;
    std::string msg;
    mcx_error(int c, std::string m) : code(c), msg(std::move(m)) {}
};

// ---- launchers implemented in the .hip files ----
// nonbonded.hip
void launch_full_energy(const DevSys &s, const int2 *tiles, int ntiles, int tile, double *partial,
                        hipStream_t st);
void launch_switch_energy(const DevSys &s, double *partial, int nblocks, hipStream_t st);
// k_full_energy with the tile of each block computed instead of read from a
// tile list: ceil(n/64)*(ceil(n/64)+1)/2 blocks, two partials each.
void launch_full_energy_listless(const DevSys &s, double *partial, hipStream_t st);
void launch_reduce(const double *partial, int nblocks, int ncomp, double *out, hipStream_t st);
// the two sums of launch_reduce (pa -> oa, pb -> ob) in one launch
void launch_reduce2(const double *pa, int na, int ca, double *oa, const double *pb, int nb, int cb, double *ob,
                    hipStream_t st);
int switch_energy_blocks(int n);
// list.hip
struct CellGrid {
    int nc[3];
    int ncell;
    int reach[3];            // cells searched in each direction
    int *cell_of, *cell_start, *cell_count, *cell_atoms, *cursor;
    double4 *cpos;           // cell-ordered (x, y, z, 2*atom + frozen)
    int4 *cfix;              // cell-ordered (fixed-point x, y, z, 2*atom + frozen)
    int4 *ipos_ref;          // fixed-point positions at this build (displacement test)
    int *cell_tmp;           // binned atoms before the in-cell sort
    int *group_start;        // [ncell + 1] first row group of each cell (periodic build)
    int4 *group_info;        // [n + ncell] row group -> (cell, first record, atoms, 0)
    int *cell_acc;           // [ncell] binning accumulators, zero between builds
    unsigned *ticket;        // [1] blocks of the binning pass done, zero between builds
};
void launch_list_build(const DevSys &s, const CellGrid &g, hipStream_t st);
size_t list_build_lds_bytes(int n);
// list_patch.hip: the list brought up to date after Metropolis moves by
// recomputing only what they can change.  Atoms moved since the last
// update get their rows rebuilt; every other row only gains or loses moved
// atoms, applied as sorted edits.  The result equals a full rebuild bit for
// bit.  Atoms moved since the last full build (binning) are scanned
// explicitly, as their cells are stale.
constexpr int kPatchEdits = 32;      // edits one row can take in one update
constexpr int kPatchRows = 4096;     // moved atoms one update can take
constexpr int kPatchStale = 3072;    // atoms moved since the binning before a full build is asked for
struct ListPatch {
    int *moved_flag, *moved_list, *moved_cnt;   // moved since the last update (flag, list, count)
    int *stale_flag, *stale_list, *stale_cnt;   // moved since the last full build
    int *edit_cnt, *edit_val;                    // [n], [n * kPatchEdits]: +(a+1) insert a, -(a+1) delete a
    int *touched, *touched_cnt;                  // rows with edits
    int *newrow, *newcnt, *newhalf;              // [kPatchRows * cap], [kPatchRows] rebuilt rows of moved atoms
    unsigned *ticket;                            // blocks of the apply kernel finished
};
// Queues the two patch kernels; a row outgrowing the capacity, more edits
// or moved atoms than one update takes, or kPatchStale atoms moved since the
// binning set s.overflow (the next Metropolis chunk then bails out and the
// host runs a full build).  Returns false when this system cannot be patched
// (too many atoms for the LDS bitmaps, rows too long for the apply kernel).
bool list_patch_supported(const DevSys &s);
void launch_list_patch(const DevSys &s, const CellGrid &g, const ListPatch &p, hipStream_t st);
// zeroes the patch state (after a full build)
void launch_list_patch_reset(const DevSys &s, const ListPatch &p, hipStream_t st);
// bonded.hip
void launch_bonded(const DevSys &s, const DevTerms &t, double *partial, int nblocks, hipStream_t st);
int bonded_blocks(const DevTerms &t);
// mc.hip
struct DevMCState;
void launch_delta_energy(const DevSys &s, const DevTerms &t, const DevMoves &mv, int first, int last,
                         const double *newxyz, double *de, hipStream_t st);
struct MCRunArgs {
    int nsteps, step0, rng_mode;
    uint64_t seed;
    uint32_t stream;       // Philox stream id
    double beta;
    double *etot;          // device scalar
    uint32_t *mt_state;    // [692] device: MT19937 state, index, Philox counter, draw batch
    double *dmax;          // [ntypes] device
    long long *trials, *accepted;   // [ntypes] device
    int *acc_tmp;          // [ntypes]
    unsigned char *acc_out; double *de_out; int *mvtyp_out; int *mvinst_out;  // device or null
    double *disp_out;      // [3*nsteps] translation of TRN and force-bias moves, or null
    double *lnq_out;       // [nsteps] log proposal-density ratio of force-bias moves, or null
    // displacement-triggered list (auto mode): positions at the last build;
    // a trial farther than half the skin from them is evaluated over every
    // atom, and an accepted one ends the chunk (moves done -> *stop_at)
    const int4 *ipos_ref;  // null: fixed rebuild schedule
    double half_skin2;     // (dcut/2)^2, A^2
    int *stop_at;          // device scalar
    // fixed rebuild schedule, list rebuilds queued without a host wait: a
    // chunk that finds the preceding rebuild overflowed (s.overflow) or an
    // earlier chunk bailed out changes nothing and records its first step
    int *bail;             // device scalar (INT_MAX: none), or null
    // evaluate several proposals per round (no force-bias move type)
    int spec;
    // incremental list updates: every committed atom is recorded (null: no
    // record; the list is rebuilt in full)
    ListPatch patch;
    int record_moves;
};
void launch_mc_chunk(const DevSys &s, const DevTerms &t, const DevMoves &mv, const MCRunArgs &a,
                     hipStream_t st);
// Speculative rounds over several workgroups (k_mc_wide): workgroup g of a
// round evaluates the proposal g moves ahead on the assumption that the
// moves before it are rejected; the rounds resolve in chain order.
constexpr int kMaxWide = 32;        // workgroups of one launch
constexpr int kDefaultWide = 12;    // proposals per round unless MCX_WIDE says otherwise
constexpr int kDefaultSplit = 2;    // workgroups per proposal unless MCX_WIDE_SPLIT says otherwise
constexpr int kWideRing = 512;      // draws generated ahead (power of two)
// the draw state of a wide run: the common layout (kRngWords words, last
// word = 2) followed by the count of drawn-but-unused draws and the draws
constexpr int kRngWordsWide = kRngWords + 1 + 2 * kWideRing;
struct WideArgs {
    int nwg;                 // workgroups of the launch (<= kMaxWide)
    unsigned *sync;          // [1] round-barrier arrivals, zero at launch (the last workgroup to finish clears it)
    unsigned *exit_ctr;      // [1] workgroups finished, zero at launch (likewise)
    unsigned *arrive;        // [1] workgroups arrived at the start-of-launch rendezvous (likewise)
    int *go;                 // [1] rendezvous decision: 0 open, 1 all resident, 2 abandoned (likewise)
    long long rdv_ticks;     // give-up time of the rendezvous (s_memrealtime ticks)
    int *err;                // [1] set when a round barrier gave up waiting
    long long spin_ticks;    // give-up time of a barrier wait (s_memrealtime ticks)
    int split;               // workgroups per proposal (its pair wavefronts shared out), nwg a multiple
    double *bred;            // [kMaxWide * 16 * 3] wavefront sums (elec, vdw, bonded) of each proposal
    int4 *bmeta;             // [kMaxWide] (type index, first atom, moved atoms, instance)
    double *bxyz;            // [kMaxWide * kMaxMoveAtoms * 3] trial coordinates
    int4 *bip;               // [kMaxWide * kMaxMoveAtoms] trial fixed-point mirror
    double *blnq;            // [kMaxWide] force bias: log proposal-density ratio of each proposal
};
// Launch of nwg workgroups.  The kernel starts with a rendezvous: unless
// every workgroup is resident within rdv_ticks, it changes nothing, sets
// *err = 3 and the bail-out step; the host then runs the chunk again.
hipError_t launch_mc_wide(const DevSys &s, const DevTerms &t, const DevMoves &mv, const MCRunArgs &a,
                          const WideArgs &w, hipStream_t st);
// workgroups of k_mc_wide that can be resident at once on the device
int mc_wide_max_workgroups(int device);
// adjoint.hip
// adjoint.hip: g8 = 8 n values per system (dE/dx, dE/dy, dE/dz, dE/deps,
// dE/dsig, dE/dq, dE/deps14, dE/dsig14); type sums: 5 per type (eps, sig,
// q, eps14, sig14)
void launch_nonbonded_grad(const DevSys &s, double *g8, hipStream_t st);
void launch_nonbonded_grad_batch(const DevSys *sys, double *const *g8, int nsys, int nmax, hipStream_t st);
void launch_type_sums(const double *g8, int n, const int *type, int ntypes, double *out, hipStream_t st);
// mc.hip: the fixed-point mirror recomputed from the device coordinates
// (charge bits kept)
void launch_mirror_from_coords(const DevSys &s, hipStream_t st);
// the FP64 coordinates from a packed device copy (x[n], y[n], z[n]) into
// s.x, s.y, s.z, and the mirror from them
void launch_mirror_from_packed(const DevSys &s, const double *xyz, hipStream_t st);
// batch.hip: one Metropolis chain per replica, one workgroup each
constexpr int kBatchTypes = 64;     // distinct (q, eps, sig) atom types of a replica
constexpr int kBatchAtoms = 32;     // atoms per move instance
struct BatchRep {
    DevSys s;
    DevMoves mv;
    const int *atype;        // [n] type index | frozen << 16
    const double4 *ptab;     // [ntab] (charge, eps, sig in A, 0)
    int ntab;
    int sweep_len;           // moves per sweep
    int frozen_any;
    int rng_mode;            // 1: std::mt19937 replay, 0: Philox
    uint64_t key;            // Philox key
    uint32_t stream;         // Philox stream
    uint32_t *rng;           // [692] draw state
    double *etot;            // running total energy (iu)
    double *dmax;            // [kMaxMoveTypes]
    long long *trials, *accepted;   // [kMaxMoveTypes]
    int *acc_tmp;            // [kMaxMoveTypes]
    long long *step;         // moves done since the chain was (re)started
    unsigned long long *npair;   // in-cutoff pair changes evaluated (each: old and new energy)
};
struct BatchArgs {
    const BatchRep *reps;
    const int *order;            // workgroup -> replica
    const double *temps;         // [nrep] temperature of each ladder slot
    const int *slot_of_rep;      // [nrep]
    int img_cap;                 // atoms held by the LDS image
    int nmoves;                  // moves per replica in this launch (<= 0: the replica's sweep_len)
    int rec_len;                 // moves recorded per replica (the launch's first rec_len)
    unsigned char *acc_out;      // [nrep * rec_len] or null
    double *de_out;              // [nrep * rec_len] or null
    int cs;                      // every replica cubic with a switched cutoff: the specialised kernel
};
struct ExchArgs {
    int nrep;
    uint64_t key;
    const double *temps;         // [nrep] slot temperatures
    const double *etot;          // [nrep] replica energies
    int *slot_of_rep, *rep_of_slot;
    unsigned long long *counter; // exchange rounds so far
    long long *att, *acc;        // [nrep - 1] per neighbouring slot pair
};
bool batch_variant_ok(int threads, int unroll, bool img);
size_t batch_static_lds(int threads, int unroll, bool img);
size_t batch_dynamic_lds(int threads, bool img, int img_cap);
void batch_prepare(int threads, int unroll, bool img, int img_cap);   // sets the kernel's LDS limit
void launch_batch_sweep(const BatchArgs &a, int nrep, int threads, int unroll, bool img, hipStream_t st);
void launch_exchange(const ExchArgs &x, hipStream_t st);
