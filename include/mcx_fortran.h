/* mcx_fortran.h -- the reference's Fortran/C energy interface, served by the
 * gfx950 library.
 *
 * Same names, argument lists, units and call order as
 * src/FORTRAN_interface.cpp:21-106 (one static system per process, strings
 * NUL-terminated by the caller, energies in kcal/mol), so a Fortran program
 * written against the reference's `charmmene` library (test/fortran_test/
 * test.f95) links against libmcx.so unchanged.  Unlike the reference these
 * entry points never exit(): a failure is reported by mcx_f_last_error()
 * returning a non-empty string (and get_energy_ filling NaN).
 *
 * One behavioural difference, by design: after set_coords_ the reference
 * keeps evaluating switched energies over the Verlet list built in
 * set_ff_and_cor_, which no longer holds every pair inside the cutoff once
 * atoms move; here the list is rebuilt for the new coordinates. */
#ifndef MCX_FORTRAN_H
#define MCX_FORTRAN_H

#ifdef __cplusplus
extern "C" {
#endif

/* set_pbc_: PerConditions(type, a, b, c, alpha, beta, gamma)
 * (src/FORTRAN_interface.cpp:23-29); type "none", "cubic" or "orbic". */
void set_pbc_(char *type, double *a, double *b, double *c, double *alpha, double *beta, double *gamma);

/* set_ens_: Ens_NVT(natom, volume, T) (src/FORTRAN_interface.cpp:32-35). */
void set_ens_(int *natom, double *Temp);

/* set_ff_and_cor_: FField_MDBAS + IO_MDBAS + List_nonBonded
 * (src/FORTRAN_interface.cpp:37-55): reads the MDBAS forcefield and COR
 * files, builds the exclusions and, in "switch" mode, the Verlet list. */
void set_ff_and_cor_(char *cutoffmode, double *ctoff, double *cton, double *dcut, char *FF_file, char *COR_file);

/* get_energy_: getE(ener) / 418.4 -- kcal/mol, order {tot, pot, kin,
 * elec, vdw, bond, ang, ub, dihe, impr} (src/FORTRAN_interface.cpp:57-75). */
void get_energy_(double ener[10]);

/* set_coords_, get_mass_, clean_ (src/FORTRAN_interface.cpp:77-105). */
void set_coords_(double x[], double y[], double z[]);
void get_mass_(double mass[]);
void clean_(void);

/* Not in the reference: the last failure of the calls above ("" if none). */
const char *mcx_f_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* MCX_FORTRAN_H */
