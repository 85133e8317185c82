// ref_driver.cpp -- a C entry-point harness around the reference's own
// classes, compiled together with the reference sources into
// oracle/_ref/libmcref.so by oracle/Makefile (target "ref").
//
// TEST INFRASTRUCTURE ONLY: used by tests/ to check the FP64 restatement
// (oracle/mc_oracle.c) against the real reference code, and optionally as
// the CPU baseline.  It constructs the objects the way Parser_XML does
// (src/Parser.cpp:60-279) and exposes energies, lists and the Metropolis
// loop with plain pointers.  Console output of the reference is captured
// into a buffer readable through refd_output().
#include <cstdint>
#include <cstring>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "AtomList.hpp"
#include "Ens_NVT.hpp"
#include "FField_MDBAS.hpp"
#include "IO_MDBAS.hpp"
#include "List_Moves.hpp"
#include "List_nonBonded.hpp"
#include "MC_metropolis.hpp"
#include "PerConditions.hpp"

namespace {

struct RefSys {
    AtomList at;
    PerConditions *pbc = nullptr;
    Ens_NVT *ens = nullptr;
    FField_MDBAS *ff = nullptr;
    IO_MDBAS *io = nullptr;
    List_nonBonded *lst = nullptr;
    int natom = 0;
};

std::string g_out;

struct Capture {
    std::stringstream ss;
    std::streambuf *old;
    Capture() : old(std::cout.rdbuf(ss.rdbuf())) {}
    ~Capture() { std::cout.rdbuf(old); g_out = ss.str(); }
};

}  // namespace

extern "C" {

const char *refd_output() { return g_out.c_str(); }

void *refd_create(const char *cor, const char *fff, const char *pbtype, double a, double b, double c,
                  int natom, double temp, const char *cutmode, double ctoff, double cton, double dcut)
{
    Capture cap;
    RefSys *s = new RefSys();
    s->natom = natom;
    s->pbc = new PerConditions(std::string(pbtype), a, b, c, 90.0, 90.0, 90.0);
    s->ens = new Ens_NVT(natom, s->pbc->computeVol(), temp);
    s->ff = new FField_MDBAS(s->at, *s->pbc, *s->ens, std::string(cutmode), ctoff, cton, dcut);
    s->io = new IO_MDBAS(std::string(cor), std::string(fff), *s->ff, s->at, *s->pbc, *s->ens);
    s->lst = new List_nonBonded(s->at, *s->ff, *s->pbc, *s->ens);
    s->ff->setExcl(*s->lst);
    return s;
}

void refd_destroy(void *h)
{
    RefSys *s = static_cast<RefSys *>(h);
    delete s->lst; delete s->io; delete s->ff; delete s->ens; delete s->pbc;
    delete s;
}

void refd_energy(void *h, double ener[10])
{
    Capture cap;
    static_cast<RefSys *>(h)->ff->getE(ener);
}

void refd_update_list(void *h)
{
    Capture cap;
    static_cast<RefSys *>(h)->lst->update_verlet_list();
}

void refd_get_coords(void *h, double *x, double *y, double *z)
{
    RefSys *s = static_cast<RefSys *>(h);
    for (int i = 0; i < s->natom; i++) { x[i] = s->at.getX(i); y[i] = s->at.getY(i); z[i] = s->at.getZ(i); }
}

void refd_set_coords(void *h, const double *x, const double *y, const double *z)
{
    RefSys *s = static_cast<RefSys *>(h);
    for (int i = 0; i < s->natom; i++) s->at.setCoords(i, x[i], y[i], z[i]);
}

// Verlet list: row counts, then rows packed back to back (ascending j).
long long refd_get_list(void *h, int *neighPair, int *flat, long long cap)
{
    RefSys *s = static_cast<RefSys *>(h);
    const std::vector<int> &np = s->lst->getNeighPair();
    const std::vector<std::vector<int>> &nl = s->lst->getNeighList();
    long long o = 0;
    for (int i = 0; i < s->natom; i++) {
        neighPair[i] = np[i];
        for (int k = 0; k < np[i]; k++) { if (o < cap) flat[o] = nl[i][k]; o++; }
    }
    return o;
}

// Exclusion rows as the reference exposes them (visible exclPair[i] entries).
long long refd_get_excl(void *h, int *exclPair, int *flat, long long cap, int *npair14, int *list14, int cap14)
{
    RefSys *s = static_cast<RefSys *>(h);
    const std::vector<int> &ep = s->lst->getExclPair();
    const std::vector<std::vector<int>> &el = s->lst->getExclList();
    long long o = 0;
    for (int i = 0; i < s->natom; i++) {
        exclPair[i] = ep[i];
        for (int k = 0; k < ep[i]; k++) { if (o < cap) flat[o] = el[i][k]; o++; }
    }
    *npair14 = s->lst->getNPair14();
    const std::vector<int> &l14 = s->lst->getNeighList14();
    for (int k = 0; k < 2 * *npair14 && k < cap14; k++) list14[k] = l14[k];
    return o;
}

// Runs MC_metropolis::run() for nsteps with the given move types
// (type 1 TRN / 2 ROT; mode "residue"/"atom"/"all"; selection "all").
// Final dmax values are written to dmax_out; console output via refd_output().
void refd_mc(void *h, int ntypes, const int *type, const char *const *mode, const double *dmax,
             const double *target, const int *each, int nsteps, uint64_t seed, double *dmax_out)
{
    RefSys *s = static_cast<RefSys *>(h);
    Capture cap;
    List_Moves mv(s->at, *s->ff, s->natom);
    for (int t = 0; t < ntypes; t++)
        mv.addNewMoveType(type[t] == 1 ? "trn" : "rot", mode[t], "all", "", dmax[t], target[t], each[t]);
    s->ff->setMcmvlst(mv);
    {
        MC_metropolis mc(s->at, *s->pbc, *s->ens, *s->ff, mv, nsteps, 0, seed);
        mc.run();
    }
    for (int t = 0; t < ntypes; t++) dmax_out[t] = mv.getMoveLimitsList()[t];
}

}  // extern "C"
