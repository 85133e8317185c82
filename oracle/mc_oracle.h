/*
 * mc_oracle.h -- FP64 CPU restatement of the mc_cpp energy hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * backend in mc_cpp_amd/: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product never links it.
 *
 * Every function restates a reference routine (cited per function in
 * mc_oracle.c) with the same operation order, the same truncating rint
 * and no floating-point contraction (built with -ffp-contract=off).
 *
 * Units follow the reference: energies in internal units (iu,
 * 1 kcal/mol = 418.4 iu), per-atom eps = sqrt(eps_kcal*418.4),
 * sig = sigma/2, bonded force constants already multiplied by 418.4,
 * angles in radians.
 */
#ifndef MC_ORACLE_H
#define MC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_PBC_NONE = 0, ORC_PBC_CUBIC = 1, ORC_PBC_ORBIC = 2 };
enum { ORC_CUT_FULL = 0, ORC_CUT_SWITCH = 1 };
enum { ORC_MV_TRN = 1, ORC_MV_ROT = 2 };
enum { ORC_MODE_RESIDUE = 1, ORC_MODE_ALL = 2, ORC_MODE_ATOM = 4 };

/* All arrays are caller-owned (numpy) except the ones the oracle fills
 * through orc_build_exclusions / orc_update_verlet_list, which it allocates
 * and frees in orc_free_lists. */
typedef struct orc_system {
    int natom;
    int pbtype;
    double pbx, pby, pbz, rpbx, rpby, rpbz;
    int cutmode;
    double ctoff, cton, dcut;
    double temp;

    double *x, *y, *z;
    double *charge, *eps, *sig, *eps14, *sig14, *mass;
    int *frozen;
    int *resid;          /* residue_id_global, 1-based consecutive */

    int nbond;  int *bond_a, *bond_b, *bond_type; double *bond_k, *bond_r0, *bond_beta;
    int nub;    int *ub_a, *ub_b; double *ub_k, *ub_r0;
    int nangle; int *ang_a, *ang_b, *ang_c; double *ang_k, *ang_t0;
    int ndihe;  int *dih_a, *dih_b, *dih_c, *dih_d, *dih_type; double *dih_k, *dih_phi0, *dih_mult;
    int nimpr;  int *imp_a, *imp_b, *imp_c, *imp_d, *imp_type; double *imp_k, *imp_phi0, *imp_mult;

    /* exclusion lists (filled by orc_build_exclusions) */
    int *exclPair;       /* [natom] visible length of each row */
    int *exclOff;        /* [natom+1] row offsets into exclIdx (row capacity = tmpPair+1) */
    int *exclIdx;
    int nPair14;
    int *neighList14;    /* [2*nPair14] */
    int pair14_overflow; /* 1 if more 1-4 pairs than the reference's 5*nDihedral capacity */

    /* Verlet list (filled by orc_update_verlet_list): row i holds neighPair[i]
     * ascending j > i at neighList[i*cap .. ]; cap grows as needed. */
    int *neighPair;
    int *neighList;
    int cap;
    int sizeList;        /* what the reference's init_verlet_list would allocate */

    /* energy components of the last evaluation (iu) */
    double tot, pot, kin, elec, vdw, bond, ang, ub, dihe, impr;
} orc_system;

void orc_set_box(orc_system *s, int pbtype, double a, double b, double c);
void orc_apply_pbc(const orc_system *s, double d[3]);
double orc_rint(double x);

int  orc_build_exclusions(orc_system *s);
void orc_free_lists(orc_system *s);
long long orc_update_verlet_list(orc_system *s, int nthreads);

double orc_getE(orc_system *s, double ener[10]);
void orc_nonbonded_full(orc_system *s);
void orc_nonbonded14(orc_system *s);
void orc_nonbonded_switch(orc_system *s);
/* threads of the switched pair loops; 1 (default) = serial reference order */
void orc_set_energy_threads(int n);
void orc_nonbonded14_switch(orc_system *s);
void orc_ebond(orc_system *s);
void orc_eang(orc_system *s);
void orc_eub(orc_system *s);
void orc_edihe(orc_system *s);
void orc_eimpr(orc_system *s);

/* energy of every nonbonded (list or all-pair), 1-4 and bonded term that
 * touches at least one atom of [first,last]; used for dE = E_set(new)-E_set(old) */
double orc_moving_set_energy(orc_system *s, int first, int last);
/* the same with every pair the list would hold at the current coordinates */
double orc_moving_set_energy_exact(orc_system *s, int first, int last);

/* MT19937 (C++ [rand.eng.mers]) + libstdc++ generate_canonical<double,53> */
typedef struct orc_mt { uint32_t mt[624]; int idx; } orc_mt;
void orc_mt_seed(orc_mt *g, uint64_t seed);
uint32_t orc_mt_next(orc_mt *g);
double orc_mt_canonical(orc_mt *g);

/* Monte Carlo move types, restating List_Moves for TRN/ROT with
 * RESIDUE/ATOM/ALL modes and an atom-mask selection. */
typedef struct orc_movetype {
    int type;            /* ORC_MV_TRN / ORC_MV_ROT */
    int mode;            /* ORC_MODE_* */
    const int *sele;     /* [natom] 0/1, NULL = all */
    double dmax, target; int each;
} orc_movetype;

typedef struct orc_mc_record {
    int *mvtyp, *mvinst;   /* [nsteps] */
    double *de;            /* [nsteps] */
    unsigned char *acc;    /* [nsteps] */
} orc_mc_record;

/* Runs the Metropolis loop of MC_metropolis::run.  mode 0 = reference
 * compat (two full getE per move), 1 = moving-set dE, 2 = moving-set dE
 * over every pair the list would hold at the current coordinates (the
 * batched path's definition; equals mode 1 until an atom outruns the list
 * skin).  Returns final etot; final dmax per type written to dmax_out. */
double orc_mc_run(orc_system *s, const orc_movetype *types, int ntypes,
                  int nsteps, uint64_t seed, int mode, int list_every,
                  orc_mc_record *rec, double *dmax_out, int nthreads);

/* FP64 reference gradients of the nonbonded switched energy:
 * dE/dx,dy,dz [natom] and dE/deps, dE/dsig, dE/dq per atom (iu). */
/* dE/dx, dE/dy, dE/dz, dE/deps, dE/dsig, dE/dq, dE/deps14, dE/dsig14 (8 n
   values) of elec + vdw, list and 1-4 pairs, SWITCH or FULL mode */
void orc_nonbonded_grad(orc_system *s, double *g8);
void orc_nonbonded_switch_grad(orc_system *s, double *gx, double *gy, double *gz,
                               double *geps, double *gsig, double *gq);

#ifdef __cplusplus
}
#endif
#endif
