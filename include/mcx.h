/*
 * mcx.h -- C ABI of the MI355X (gfx950) energy backend for mc_cpp.
 *
 * One opaque handle per simulated system (or per batch of replicas).  All
 * arrays are plain host pointers that the library copies; the caller keeps
 * ownership.  Every call returns an int status (MCX_OK on success); the
 * message of the last failure on a handle is read with mcx_last_error().
 * Calls on one handle are serialised on that handle's HIP stream; distinct
 * handles are independent.
 *
 * Units are the reference's internal units (src/IO_MDBAS.cpp:130-505,
 * include/Constants.hpp:29): energies in 10 J/mol (1 kcal/mol = 418.4 iu),
 * eps = sqrt(eps_kcal*418.4), sig = sigma/2, bonded k already *418.4,
 * angles in radians.
 *
 * Reference interfaces replaced (file:line in the mc_cpp tree):
 *   FField ctor / cut mode        include/FField.hpp:58-59, src/FField.cpp:27-51
 *   PerConditions                 src/PerConditions.cpp:103-134
 *   IO_MDBAS -> AtomList / FField::set*List   src/IO_MDBAS.cpp:130-505, src/FField.cpp:127-180
 *   List_nonBonded                include/List_nonBonded.hpp:50-59, src/List_nonBonded.cpp:143-915
 *   FField::getE / getE(double[10]) include/FField.hpp:95-96, src/FField_MDBAS.cpp:53-106
 *   FField::askListUpdate         include/FField.hpp:98, src/FField.cpp:222-235
 *   MC_metropolis::run            src/MC_metropolis.cpp:45-204
 *   FORTRAN C ABI (kcal/mol)      src/FORTRAN_interface.cpp:21-106
 */
#ifndef MCX_H
#define MCX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MCX_OK                 0
#define MCX_ERR_ARG            1  /* bad argument or size */
#define MCX_ERR_HIP            2  /* HIP runtime failure */
#define MCX_ERR_LIST_OVERFLOW  3  /* pair-list capacity exceeded (the library regrows and rebuilds) */
#define MCX_ERR_BOX_TOO_SMALL  4  /* cutoff + skin exceeds half the box in fast (non-compat) paths */
#define MCX_ERR_STATE          5  /* call order violated (e.g. energy before coordinates) */
#define MCX_ERR_UNSUPPORTED    6  /* configuration outside what the device path implements */

#define MCX_PBC_NONE   0
#define MCX_PBC_CUBIC  1
#define MCX_PBC_ORBIC  2

#define MCX_CUT_FULL   0
#define MCX_CUT_SWITCH 1

#define MCX_MV_TRN 1
#define MCX_MV_ROT 2
#define MCX_MV_FBIAS 3   /* force-bias (smart MC) translation; not in the reference */
#define MCX_MV_FBIAS_ACC 4 /* the same with the drift taken at the Metropolis test's own beta; not in the reference */
#define MCX_MODE_RESIDUE 1
#define MCX_MODE_ALL     2
#define MCX_MODE_ATOM    4

typedef struct mcx_ctx mcx_ctx;

/* Creates a handle for natom atoms on HIP device `device`.
 * Replaces the FField/AtomList/Ensemble construction of Parser_XML
 * (src/Parser.cpp:77-215). */
int mcx_create(int natom, int device, mcx_ctx **out);
void mcx_destroy(mcx_ctx *h);
const char *mcx_last_error(const mcx_ctx *h);

/* PerConditions(type, a, b, c) -- src/PerConditions.cpp:31-61,103-134.
 * For MCX_PBC_CUBIC only `a` is used. */
int mcx_set_box(mcx_ctx *h, int pbtype, double a, double b, double c);

/* FField(..., cutMode, ctoff, cuton, dcut) -- include/FField.hpp:58-59;
 * temperature of Ens_NVT (src/Ens_NVT.cpp:23-31) used by the Metropolis test. */
int mcx_set_cut(mcx_ctx *h, int cutmode, double ctoff, double cton, double dcut, double temp);

/* AtomList per-atom arrays as IO_MDBAS::read_ff fills them
 * (src/IO_MDBAS.cpp:183-186,460-465; src/IO_MDBAS.cpp:121 for resid).
 * frozen: 0/1.  resid: residue_id_global (1-based, consecutive). */
int mcx_set_atoms(mcx_ctx *h, const double *charge, const double *eps, const double *sig,
                  const double *eps14, const double *sig14, const double *mass,
                  const int *frozen, const int *resid);

/* FField::setBndList/setUbList/setAngList/setDiheList/setImprList
 * (src/FField.cpp:157-180) with the records of include/{Bond,Bond_UB,Angle,
 * Dihedral,Dihedral_improper}.hpp.  Atom indices 0-based. */
int mcx_set_bonds(mcx_ctx *h, int n, const int *a, const int *b, const int *type,
                  const double *k, const double *r0, const double *beta);
int mcx_set_ub(mcx_ctx *h, int n, const int *a, const int *b, const double *k, const double *r0);
int mcx_set_angles(mcx_ctx *h, int n, const int *a, const int *b, const int *c,
                   const double *k, const double *theta0);
int mcx_set_dihedrals(mcx_ctx *h, int n, const int *a, const int *b, const int *c, const int *d,
                      const int *type, const double *k, const double *phi0, const double *mult);
int mcx_set_impropers(mcx_ctx *h, int n, const int *a, const int *b, const int *c, const int *d,
                      const int *type, const double *k, const double *phi0, const double *mult);

/* List_nonBonded::build_exclude_list (src/List_nonBonded.cpp:143-775):
 * builds the exclusion rows and the 1-4 pair list on the host from the
 * bonded lists and uploads them.  Must follow the mcx_set_{bonds..} calls. */
int mcx_build_exclusions(mcx_ctx *h);

/* getExclPair/getExclList/getNPair14/getNeighList14
 * (src/List_nonBonded.cpp:1343-1361).  Rows are written back to back into
 * `idx` (visible entries only); returns sizes through *nidx, *npair14. */
int mcx_get_exclusions(mcx_ctx *h, int *exclPair, int *idx, long long cap, long long *nidx,
                       int *pairs14, int cap14, int *npair14);

/* AtomList::setCoords / getCoords for all atoms (src/FORTRAN_interface.cpp:77-82).
 * mcx_set_coords marks the pair list stale (the next energy rebuilds it). */
int mcx_set_coords(mcx_ctx *h, const double *x, const double *y, const double *z);

/* Coordinates changed by Monte Carlo moves, as MOVE_TRN / MOVE_ROT change
 * AtomList in place (include/Move.hpp:42-143): the pair list stays the one
 * built at the last update, as the reference's Verlet list does until
 * askListUpdate (src/FField.cpp:222-235), so a following mcx_energy is the
 * reference's getE over that list.  The arrays go up in one copy from a
 * pinned staging buffer of the handle (the call returns once they are
 * staged; the next call waits for that copy) and the fixed-point mirror is
 * recomputed on the device.  The first upload of a handle behaves as
 * mcx_set_coords. */
int mcx_set_coords_moved(mcx_ctx *h, const double *x, const double *y, const double *z);
int mcx_get_coords(mcx_ctx *h, double *x, double *y, double *z);

/* List_nonBonded::update_verlet_list (src/List_nonBonded.cpp:836-915) on the
 * GPU: cell binning + in-kernel PerConditions minimum image + stream
 * compaction.  The pair set equals the reference's bit for bit.  Reports
 * the number of (i<j) pairs.  FField::askListUpdate (src/FField.cpp:222-235)
 * maps to this call; it is a no-op in MCX_CUT_FULL mode. */
int mcx_update_list(mcx_ctx *h, long long *npairs);

/* getNeighPair/getNeighList (src/List_nonBonded.cpp:1363-1376): the half
 * list, row i holding ascending j > i, written back to back into `flat`. */
int mcx_get_list(mcx_ctx *h, int *neighPair, int *flat, long long cap);

/* Row capacity of the pair list (entries per atom), the sizeList of
 * List_nonBonded::init_verlet_list (src/List_nonBonded.cpp:827-830, sized
 * there as max_i neighPair * 1.04 + 1 at the first build).  set > 0 makes
 * it exactly `set` and marks the list stale; set == 0 leaves it.  *current
 * (if non-null) receives the capacity in use.  Unlike the reference, a build
 * that needs more -- also one inside mcx_mc_run -- grows it and rebuilds. */
int mcx_list_capacity(mcx_ctx *h, int set, int *current);

/* FField::getE(double ener[10]) (src/FField_MDBAS.cpp:77-106):
 * {tot, pot, kin, elec, vdw, bond, ang, ub, dihe, impr} in iu. */
int mcx_energy(mcx_ctx *h, double ener[10]);

/* mcx_energy with no device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Partial sums and reduced components live in one static 4 MiB device
 * array per GPU, held until the process ends; all-pair mode supports at
 * most 46,272 atoms, switched mode 1,048,544 (MCX_ERR_ARG beyond).
 * Switched mode needs a current pair list (MCX_ERR_STATE otherwise: call
 * mcx_update_list, which may allocate, beforehand).  Same result bits as
 * mcx_energy.  Calls on one GPU share the array: calls that overlap (other
 * handles, streams or threads on that GPU) race and must be serialised. */
int mcx_energy_noalloc(mcx_ctx *h, double ener[10]);

/* Energy change of moving atoms [first,last] to the given coordinates
 * (arrays of last-first+1 entries), against the current list; equal to
 * getE(new) - getE(old) as MC_metropolis::run forms it
 * (src/MC_metropolis.cpp:94,122-123).  Coordinates are not committed. */
int mcx_delta_energy(mcx_ctx *h, int first, int last, const double *nx, const double *ny,
                     const double *nz, double *de);

/* One Monte Carlo move type, as List_Moves::addNewMoveType registers it
 * (src/List_Moves.cpp:170-211,257-457).  sele: 0/1 per atom or NULL = all. */
typedef struct mcx_move_type {
    int type;          /* MCX_MV_TRN / MCX_MV_ROT / MCX_MV_FBIAS / MCX_MV_FBIAS_ACC */
    int mode;          /* MCX_MODE_RESIDUE / MCX_MODE_ATOM / MCX_MODE_ALL */
    double dmax;       /* A for TRN, degrees for ROT, the step A (A^2) for FBIAS */
    double target;     /* acceptance target in % */
    int each;          /* dmax update period in steps, 0 = never */
    const int *sele;
} mcx_move_type;

typedef struct mcx_mc_stats {
    double etot;           /* running total after the last step (iu) */
    long long trials[8];
    long long accepted[8];
    double dmax[8];        /* final dmax per move type */
    long long list_updates;
    double seconds;        /* wall time of the device loop */
    /* With the environment variable MCX_KERNEL_EVENTS=k (k >= 1): the
     * device time of every k-th Metropolis kernel launch of the run and of
     * every k-th in-run list update, each bracketed by HIP events on the
     * handle's stream, and how many were timed (0 else). */
    double chain_kernel_ms;
    long long chain_launches;
    double list_kernel_ms;
    long long list_launches;
} mcx_mc_stats;

#define MCX_LIST_AUTO  (-1) /* list_every: rebuild when an accepted move carries an atom beyond half the skin */
#define MCX_MC_REPLAY   1  /* draw from std::mt19937(seed) + generate_canonical exactly as src/MC.cpp:119-156 */
#define MCX_MC_PHILOX   0  /* counter-based Philox4x32-10 proposals (no replay) */

/* MC_metropolis::run (src/MC_metropolis.cpp:45-173) with the per-move dE
 * evaluated on the device and the Metropolis test of apply_criterion
 * (:175-204), including its beta = 1/((0.83144621/418.4)*T) applied to dE
 * in iu.  The list is rebuilt every `list_every` steps (the reference uses
 * 50, :151-154).  list_every = MCX_LIST_AUTO (periodic boxes; not in the
 * reference) rebuilds instead after any accepted move that carries an atom
 * farther than dcut/2 from its position at the last build, and evaluates a
 * trial that does so over every atom: every dE then sums exactly the pairs
 * inside the cutoff.  Optional per-step outputs (nsteps entries each):
 * accepted flag, dE, move type, move instance. */
int mcx_mc_run(mcx_ctx *h, const mcx_move_type *types, int ntypes, int nsteps, uint64_t seed,
               int rng_mode, int list_every, mcx_mc_stats *stats, unsigned char *acc_out,
               double *de_out, int *mvtyp_out, int *mvinst_out);

/* Per-step records of mcx_mc_run_traced; every pointer may be NULL. */
typedef struct mcx_mc_trace {
    unsigned char *acc;    /* [nsteps] accepted flag */
    double *de;            /* [nsteps] dE (iu) */
    int *mvtyp;            /* [nsteps] move type index */
    int *mvinst;           /* [nsteps] move instance */
    double *disp;          /* [3*nsteps] translation (A) of TRN and FBIAS moves, 0 for ROT */
    double *lnq;           /* [nsteps] ln q(new->old)/q(old->new) of FBIAS moves, 0 otherwise */
    /* Frames, as MC_metropolis::run writes tr.xyz / ener.dat every save_each
     * steps (src/MC_metropolis.cpp:157-163): when save_each > 0 and on_frame
     * is set, on_frame(user, step, etot, x, y, z) is called after every step
     * that is a multiple of save_each with the running total (iu) and the
     * coordinates (host arrays, valid during the call). */
    int save_each;
    void (*on_frame)(void *user, int step, double etot, const double *x, const double *y, const double *z);
    void *user;
} mcx_mc_trace;

/* mcx_mc_run with the trace records above; mcx_mc_run is this call with
 * disp = lnq = NULL.  Not in the reference (it adds the force-bias records).
 *
 * MCX_MV_FBIAS moves (force-bias / smart Monte Carlo, Rossky, Doll and
 * Friedman, J. Chem. Phys. 69, 4628 (1978); not in the reference) translate
 * the move's atom set by
 *     dx = beta A F + sqrt(2A) xi,   xi ~ N(0, I),
 * F the net nonbonded force on the set from atoms outside it in kcal/mol/A
 * (the list pairs of the set's atoms; bonded and 1-4 terms are not part of
 * F), A = dmax (A^2, adapted like a TRN step), beta = 1/(R T) in mol/kcal,
 * and accept with min(1, exp(-beta dE) q(new->old) / q(old->new)), beta dE
 * formed as for TRN/ROT moves (dE in iu, as the reference's test does), ln q
 * ratio = (|dx - beta A F_old|^2 - |-dx - beta A F_new|^2) / (4A).  Draws per move: move type, instance, four uniforms for three
 * normals (Box-Muller on 1 - u), and the acceptance uniform only when the
 * exponent is negative.  MCX_MV_FBIAS_ACC moves are the same with F in
 * iu/A (418.4 times the kcal/mol/A force): the drift then belongs to the
 * density the test samples, exp(-beta E) with E in iu, 418.4 times colder
 * than 1/(RT) in kcal/mol, so A must be ~418 times smaller for the same
 * drift.  Batches (mcx_batch_create) refuse force-bias types. */
int mcx_mc_run_traced(mcx_ctx *h, const mcx_move_type *types, int ntypes, int nsteps, uint64_t seed,
                      int rng_mode, int list_every, mcx_mc_stats *stats, const mcx_mc_trace *trace);

/* Philox stream id used by mcx_mc_run in MCX_MC_PHILOX mode (default 0).
 * Replica r of a batch draws from stream r, so a handle set to stream r
 * replays that replica's proposals.  Not in the reference. */
int mcx_set_rng_stream(mcx_ctx *h, uint32_t stream);

/* Adjoint of the nonbonded energy elec + vdw: SWITCH mode the list pairs
 * inside the cutoff (computeNonBonded_switch, src/FField_MDBAS.cpp:336-424),
 * FULL mode every non-excluded pair (computeNonBonded_full, :194-265), and in
 * both the 1-4 pairs (:267-318, :426-488).  dE/dx,dy,dz per atom and dE/deps,
 * dE/dsig, dE/dq per atom (iu, with eps/sig in the AtomList convention; the
 * 1-4 pairs' own eps14/sig14 derivatives come from
 * mcx_nonbonded_grad_device).  Not in the reference. */
int mcx_nonbonded_grad(mcx_ctx *h, double *gx, double *gy, double *gz,
                       double *geps, double *gsig, double *gq);

/* The same gradients with the parameter part summed per atom type (FP64,
 * fixed order): type[i] in [0, ntypes) for every atom; gx/gy/gz per atom
 * (n entries), geps/gsig/gq per type (ntypes entries).  Not in the reference. */
int mcx_nonbonded_grad_types(mcx_ctx *h, int ntypes, const int *type, double *gx, double *gy, double *gz,
                             double *geps, double *gsig, double *gq);

/* Batched adjoint: the gradients of nsys systems (SWITCH mode, 1-4 pairs
 * included, one device) in one launch.  g6[k] receives system k's 6*n values: gx, gy, gz, geps,
 * gsig, gq back to back.  Errors are reported on hs[0].  Not in the
 * reference. */
int mcx_nonbonded_grad_batch(mcx_ctx *const *hs, int nsys, double *const *g6);

/* Device pointer access for zero-copy interop (PyTorch tensors, HIP code):
 * which = 0 x, 1 y, 2 z (double[N] master coordinates). */
int mcx_device_coords(mcx_ctx *h, int which, void **ptr);

/* The caller wrote the master coordinates in place on the device (through
 * mcx_device_coords): the fixed-point mirror is recomputed on the device,
 * the pair list marked stale and the host copy refreshed when next needed.
 * Work the caller queued on its own streams must have completed.  Not in
 * the reference. */
int mcx_coords_updated(mcx_ctx *h);

/* The adjoint of mcx_nonbonded_grad with its outputs in device memory (no
 * host round trip): gpos[3 n] = dE/dx, dE/dy, dE/dz per atom; when ntypes
 * > 0, gtype[5 ntypes] = dE/deps, dE/dsig, dE/dq, dE/deps14, dE/dsig14
 * summed per atom type in FP64 in a fixed order (type: n host indices in
 * [0, ntypes); pass type[i] = i, ntypes = n for per-atom values).  Runs on
 * the handle's stream and returns after it.  Not in the reference. */
int mcx_nonbonded_grad_device(mcx_ctx *h, double *gpos, int ntypes, const int *type, double *gtype);

/* Timing probe used by the benchmark: runs `reps` back-to-back launches of
 * kernel `which` (0 full O(N^2) energy, 1 switched list energy, 2 list
 * build, 3 fused bonded, 4 nonbonded adjoint) on the handle's stream and
 * returns the mean
 * device time per launch from HIP events. */
int mcx_time_kernel(mcx_ctx *h, int which, int reps, double *ms_per_launch);

/* ---------------------------------------------------------------------------
 * Batched replicas.  Not in the reference, whose MC_metropolis::run
 * (src/MC_metropolis.cpp:45-204) drives one chain: a batch runs one such
 * chain per member handle, all replicas in one persistent kernel launch per
 * sweep, with the same moves, draws and acceptance test.  Each replica's dE
 * counts every pair whose minimum-image distance is inside the cutoff (no
 * pair list), which equals the reference's list-based dE whenever no pair
 * closes by more than the list skin between rebuilds.  Replica exchange
 * swaps temperatures of neighbouring ladder slots on the device.
 *
 * Member handles must be fully set up (atoms, bonded terms, exclusions,
 * coordinates, periodic box) on one device, with at most 32 atoms per move
 * instance and at most 64 distinct (charge, eps, sig) atom types; they must
 * outlive the batch.  The batch updates their device coordinates in place.
 * After changing a member's set-up, call mcx_batch_reset.
 */
typedef struct mcx_batch mcx_batch;

typedef struct mcx_batch_stats {
    long long moves;               /* moves run by the call, all replicas */
    long long accepted;            /* of which accepted */
    long long exchange_attempts;   /* replica-exchange attempts of the call */
    long long exchange_accepted;
    long long sweeps;              /* sweeps since the last reset */
    double seconds;                /* wall time of the call's device work */
    long long pair_evals;          /* in-cutoff pair changes evaluated (old and new energy each) */
} mcx_batch_stats;

#define MCX_MAX_MOVE_TYPES 8

/* Creates a batch over nrep handles.  Replica r has ntypes[r] move types
 * (1..MCX_MAX_MOVE_TYPES) as mcx_mc_run takes them; `types` holds them
 * replica after replica (selections are per replica).
 * Replica r starts in ladder slot r at its handle's temperature; the chains
 * start as mcx_batch_reset(b, 0, MCX_MC_PHILOX, NULL) leaves them.  On
 * failure the message is read with mcx_last_error(reps[0]). */
int mcx_batch_create(mcx_ctx *const *reps, int nrep, const mcx_move_type *types, const int *ntypes,
                     mcx_batch **out);
void mcx_batch_destroy(mcx_batch *b);
const char *mcx_batch_last_error(const mcx_batch *b);

/* Restarts every chain from its handle's current coordinates: energies are
 * recomputed, dmax and counters reset, replicas return to slot r, and the
 * draws restart -- MCX_MC_REPLAY: std::mt19937(seeds[r]) (seeds NULL:
 * seed + r) as mcx_mc_run draws; MCX_MC_PHILOX: key `seed`, stream r. */
int mcx_batch_reset(mcx_batch *b, uint64_t seed, int rng_mode, const uint64_t *seeds);

/* Temperatures (K) of the nrep ladder slots. */
int mcx_batch_set_ladder(mcx_batch *b, const double *temps);

/* Moves per sweep of each replica (default: its number of residues). */
int mcx_batch_set_sweep(mcx_batch *b, const int *moves_per_sweep);

/* Runs nsweeps sweeps of every replica; when exchange_every > 0, a
 * replica-exchange round follows every exchange_every-th sweep (counted
 * since the reset; rounds alternate between even and odd slot pairs).
 * use_graph != 0 replays the sweeps and exchanges as a captured hipGraph. */
int mcx_batch_run(mcx_batch *b, int nsweeps, int exchange_every, int use_graph, mcx_batch_stats *stats);

/* Runs nmoves moves of every replica in one launch, with no exchange, and
 * writes each replica's accepted flags and dE ([nrep][nmoves], either may
 * be NULL). */
int mcx_batch_moves(mcx_batch *b, int nmoves, unsigned char *acc_out, double *de_out);

/* Per-replica state: running energy (iu), ladder slot, and per move type
 * (MCX_MAX_MOVE_TYPES entries per replica, zero past its own types) trials,
 * accepted moves and dmax.  Any pointer may be NULL. */
int mcx_batch_state(mcx_batch *b, double *etot, int *slot_of_rep, long long *trials, long long *accepted,
                    double *dmax);

/* Replica-exchange attempts and acceptances of each neighbouring slot pair
 * (k, k+1), nrep-1 entries each, since the last reset. */
int mcx_batch_exchange_stats(mcx_batch *b, long long *attempts, long long *accepted);

/* Launch shape: threads per replica workgroup (256, 512, 768 or 1024) and the
 * number of 64-candidate tiles each wavefront keeps in flight (1, 2, 4). */
int mcx_batch_set_config(mcx_batch *b, int threads, int unroll);

/* Autotuner: times one sweep of each launch shape on the batch as it is,
 * restores the state after each trial, and keeps the fastest shape. */
int mcx_batch_tune(mcx_batch *b, int *threads, int *unroll, double *ms_per_sweep);

#ifdef __cplusplus
}
#endif
#endif
