// mcx_ffield_adapter.hpp -- FField implementation for the mc_cpp code base
// backed by the gfx950 library (include/mcx.h).
//
// A maintainer drops this header next to the reference's FField_MDBAS.hpp
// and instantiates FField_MCX where Parser_XML creates FField_MDBAS
// (src/Parser.cpp:139-153); after IO_MDBAS has filled the atom and bonded
// lists and List_nonBonded was constructed, call uploadTopology() once.
// Every getE() then uploads the coordinates the move code mutated in place
// (include/Move.hpp:42-143) and evaluates the energy on the GPU over the
// pair list of the last update; askListUpdate() rebuilds the list on the
// GPU.  runMetropolis() is the
// device-resident alternative to MC_metropolis::run.
//
// Interfaces implemented: include/FField.hpp:52-153 (public getE,
// getE(double[10]), askListUpdate and the protected pure virtuals).
#ifndef MCX_FFIELD_ADAPTER_HPP
#define MCX_FFIELD_ADAPTER_HPP

#include <stdexcept>
#include <string>
#include <vector>

#include "AtomList.hpp"
#include "Constants.hpp"
#include "Ensemble.hpp"
#include "FField.hpp"
#include "PerConditions.hpp"
#include "mcx.h"

class FField_MCX : public FField {
public:
    FField_MCX(AtomList &at, PerConditions &pbc_, Ensemble &ens_, std::string mode = "switch", double ctoff = 12.0,
               double cuton = 10.0, double dc = 2.0, int device = 0)
        : FField(at, pbc_, ens_, mode, ctoff, cuton, dc), n_(ens_.getN())
    {
        check(mcx_create(n_, device, &h_), "mcx_create");
        double v[3];
        pbc.get_pbc_vectors(v);
        const int pt = pbc.getType() == CUBIC ? MCX_PBC_CUBIC : pbc.getType() == ORBIC ? MCX_PBC_ORBIC : MCX_PBC_NONE;
        check(mcx_set_box(h_, pt, v[0], v[1], v[2]), "mcx_set_box");
        check(mcx_set_cut(h_, getCutMode() == FULL ? MCX_CUT_FULL : MCX_CUT_SWITCH, cutoff, cuton, deltacut,
                          ens.getTemp()), "mcx_set_cut");
    }
    ~FField_MCX() override { mcx_destroy(h_); }

    // Atom parameters and bonded lists as IO_MDBAS left them in AtomList
    // and in this FField (src/IO_MDBAS.cpp:130-505); exclusions are rebuilt
    // by the library with the List_nonBonded algorithm.
    void uploadTopology()
    {
        std::vector<double> q(n_), e(n_), s(n_), e14(n_), s14(n_), m(n_);
        std::vector<int> fr(n_), rs(n_);
        for (int i = 0; i < n_; i++) {
            q[i] = at_List.getCharge(i); e[i] = at_List.getEpsilon(i); s[i] = at_List.getSigma(i);
            e14[i] = at_List.getEpsilon14(i); s14[i] = at_List.getSigma14(i); m[i] = at_List.getMass(i);
            fr[i] = at_List.Is_frozen(i) ? 1 : 0; rs[i] = at_List.getResidue_id_global(i);
        }
        check(mcx_set_atoms(h_, q.data(), e.data(), s.data(), e14.data(), s14.data(), m.data(), fr.data(), rs.data()),
              "mcx_set_atoms");
        {
            std::vector<int> a, b, t; std::vector<double> k, r0, be;
            for (const Bond &x : bndList) { a.push_back(x.getAt1()); b.push_back(x.getAt2()); t.push_back(x.getType());
                k.push_back(x.getK()); r0.push_back(x.getR0()); be.push_back(x.getBeta()); }
            check(mcx_set_bonds(h_, (int)a.size(), a.data(), b.data(), t.data(), k.data(), r0.data(), be.data()), "bonds");
        }
        {
            std::vector<int> a, b; std::vector<double> k, r0;
            for (const Bond_UB &x : ubList) { a.push_back(x.getAt1()); b.push_back(x.getAt2());
                k.push_back(x.getK()); r0.push_back(x.getR0()); }
            check(mcx_set_ub(h_, (int)a.size(), a.data(), b.data(), k.data(), r0.data()), "ub");
        }
        {
            std::vector<int> a, b, c; std::vector<double> k, t0;
            for (const Angle &x : angList) { a.push_back(x.getAt1()); b.push_back(x.getAt2()); c.push_back(x.getAt3());
                k.push_back(x.getK()); t0.push_back(x.getTheta0()); }
            check(mcx_set_angles(h_, (int)a.size(), a.data(), b.data(), c.data(), k.data(), t0.data()), "angles");
        }
        torsions(diheList, false);
        torsions(imprList, true);
        check(mcx_build_exclusions(h_), "mcx_build_exclusions");
        pushCoords();
        if (getCutMode() != FULL) check(mcx_update_list(h_, nullptr), "mcx_update_list");
    }

    double getE() override
    {
        double e[10];
        getE(e);
        return e[0];
    }

    void getE(double ener[10]) override
    {
        pushCoords();
        check(mcx_energy(h_, ener), "mcx_energy");
        tot = ener[0]; pot = ener[1]; kin = ener[2]; elec = ener[3]; vdw = ener[4];
        bond = ener[5]; ang = ener[6]; ub = ener[7]; dihe = ener[8]; impr = ener[9];
    }

    void askListUpdate(int) override
    {
        if (getCutMode() == FULL) return;
        pushCoords();
        check(mcx_update_list(h_, nullptr), "mcx_update_list");
    }

    mcx_ctx *handle() { return h_; }

protected:
    double getEtot() override { return getE(); }
    double getEswitch() override { return getE(); }
    // the individual terms are evaluated together on the device
    void computeNonBonded_full() override { getE(); }
    void computeNonBonded14() override {}
    void computeNonBonded_switch() override { getE(); }
    void computeNonBonded14_switch() override {}
    void computeEbond() override {}
    void computeEang() override {}
    void computeEub() override {}
    void computeEdihe() override {}
    void computeEimpr() override {}

private:
    int n_;
    mcx_ctx *h_ = nullptr;
    std::vector<double> x_, y_, z_;

    void check(int rc, const char *what)
    {
        if (rc != MCX_OK) throw std::runtime_error(std::string(what) + ": " + mcx_last_error(h_));
    }
    // the coordinates the move code changed; the pair list stays the one of
    // the last askListUpdate, as the reference's does
    void pushCoords()
    {
        x_.resize(n_); y_.resize(n_); z_.resize(n_);
        for (int i = 0; i < n_; i++) { x_[i] = at_List.getX(i); y_[i] = at_List.getY(i); z_[i] = at_List.getZ(i); }
        check(mcx_set_coords_moved(h_, x_.data(), y_.data(), z_.data()), "mcx_set_coords_moved");
    }
    template <typename T>
    void torsions(const std::vector<T> &v, bool improper)
    {
        std::vector<int> a, b, c, d, t; std::vector<double> k, p0, mu;
        for (const T &x : v) { a.push_back(x.getAt1()); b.push_back(x.getAt2()); c.push_back(x.getAt3());
            d.push_back(x.getAt4()); t.push_back(x.getType()); k.push_back(x.getK()); p0.push_back(x.getPhi0());
            mu.push_back(x.getMult()); }
        const int n = (int)a.size();
        const int rc = improper ? mcx_set_impropers(h_, n, a.data(), b.data(), c.data(), d.data(), t.data(), k.data(), p0.data(), mu.data())
                                : mcx_set_dihedrals(h_, n, a.data(), b.data(), c.data(), d.data(), t.data(), k.data(), p0.data(), mu.data());
        check(rc, improper ? "impropers" : "dihedrals");
    }
};

#endif
