// ref_mcx_driver.cpp -- the reference's own Metropolis loop
// (MC_metropolis::run, src/MC_metropolis.cpp:45-173) driving either the
// reference's FField_MDBAS or the drop-in FField_MCX of
// include/mcx_ffield_adapter.hpp, compiled with the reference sources into
// oracle/_ref/libmcref_mcx.so by oracle/Makefile (target "ref-mcx").
//
// TEST INFRASTRUCTURE ONLY: it shows that the adapter works where the
// reference would use it.  Objects are built the way Parser_XML builds them
// (src/Parser.cpp:60-279); the only change between the two modes is the
// FField class.  Every getE() the loop makes is recorded, so a test can
// compare the two runs' energy sequences call for call (the initial total,
// then eold and enew of each step, :62,94,122).
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
#include "mcx_ffield_adapter.hpp"

namespace {

// FField with every getE() value appended to `log`
template <class FF>
struct Recorded : FF {
    std::vector<double> *log;
    template <class... A>
    Recorded(std::vector<double> *l, A &&...a) : FF(std::forward<A>(a)...), log(l) {}
    double getE() override
    {
        const double e = FF::getE();
        log->push_back(e);
        return e;
    }
};

struct RefMcx {
    AtomList at;
    PerConditions *pbc = nullptr;
    Ens_NVT *ens = nullptr;
    FField *ff = nullptr;
    FField_MCX *gpu = nullptr;      // == ff when the drop-in is used
    IO_MDBAS *io = nullptr;
    List_nonBonded *lst = nullptr;
    std::vector<double> energies;
    int natom = 0;
};

std::string g_out, g_err;

struct Capture {
    std::stringstream ss;
    std::streambuf *old;
    Capture() : old(std::cout.rdbuf(ss.rdbuf())) {}
    ~Capture() { std::cout.rdbuf(old); g_out = ss.str(); }
};

}  // namespace

extern "C" {

const char *refx_output() { return g_out.c_str(); }
const char *refx_error() { return g_err.c_str(); }

// use_gpu = 0: FField_MDBAS (the reference); 1: FField_MCX on `device`.
// NULL on failure (message in refx_error()).
void *refx_create(const char *cor, const char *fff, const char *pbtype, double a, double b, double c, int natom,
                  double temp, const char *cutmode, double ctoff, double cton, double dcut, int use_gpu, int device)
{
    Capture cap;
    RefMcx *s = new RefMcx();
    try {
        s->natom = natom;
        s->pbc = new PerConditions(std::string(pbtype), a, b, c, 90.0, 90.0, 90.0);
        s->ens = new Ens_NVT(natom, s->pbc->computeVol(), temp);
        if (use_gpu) {
            auto *g = new Recorded<FField_MCX>(&s->energies, s->at, *s->pbc, *s->ens, std::string(cutmode), ctoff,
                                               cton, dcut, device);
            s->ff = g;
            s->gpu = g;
        } else {
            s->ff = new Recorded<FField_MDBAS>(&s->energies, s->at, *s->pbc, *s->ens, std::string(cutmode), ctoff,
                                               cton, dcut);
        }
        s->io = new IO_MDBAS(std::string(cor), std::string(fff), *s->ff, s->at, *s->pbc, *s->ens);
        s->lst = new List_nonBonded(s->at, *s->ff, *s->pbc, *s->ens);
        s->ff->setExcl(*s->lst);
        if (s->gpu) s->gpu->uploadTopology();
    } catch (const std::exception &e) {
        g_err = e.what();
        delete s->lst; delete s->io; delete s->ff; delete s->ens; delete s->pbc;
        delete s;
        return nullptr;
    }
    return s;
}

void refx_destroy(void *h)
{
    RefMcx *s = static_cast<RefMcx *>(h);
    delete s->lst; delete s->io; delete s->ff; delete s->ens; delete s->pbc;
    delete s;
}

void refx_energy(void *h, double ener[10])
{
    Capture cap;
    static_cast<RefMcx *>(h)->ff->getE(ener);
}

void refx_get_coords(void *h, double *x, double *y, double *z)
{
    RefMcx *s = static_cast<RefMcx *>(h);
    for (int i = 0; i < s->natom; i++) { x[i] = s->at.getX(i); y[i] = s->at.getY(i); z[i] = s->at.getZ(i); }
}

// MC_metropolis::run for nsteps (type 1 TRN / 2 ROT; mode "residue" /
// "atom" / "all"; selection "all"); final dmax per type into dmax_out.
// Returns the number of getE() values recorded since refx_create (read
// them with refx_energies), or -1 on an exception (refx_error()).
long long refx_mc(void *h, int ntypes, const int *type, const char *const *mode, const double *dmax,
                  const double *target, const int *each, int nsteps, uint64_t seed, double *dmax_out)
{
    RefMcx *s = static_cast<RefMcx *>(h);
    Capture cap;
    try {
        List_Moves mv(s->at, *s->ff, s->natom);
        for (int t = 0; t < ntypes; t++)
            mv.addNewMoveType(type[t] == 1 ? "trn" : "rot", mode[t], "all", "", dmax[t], target[t], each[t]);
        s->ff->setMcmvlst(mv);
        {
            MC_metropolis mc(s->at, *s->pbc, *s->ens, *s->ff, mv, nsteps, 0, seed);
            mc.run();
        }
        for (int t = 0; t < ntypes; t++) dmax_out[t] = mv.getMoveLimitsList()[t];
    } catch (const std::exception &e) {
        g_err = e.what();
        return -1;
    }
    return (long long)s->energies.size();
}

void refx_energies(void *h, double *out, long long n)
{
    RefMcx *s = static_cast<RefMcx *>(h);
    for (long long k = 0; k < n && k < (long long)s->energies.size(); k++) out[k] = s->energies[k];
}

}  // extern "C"
