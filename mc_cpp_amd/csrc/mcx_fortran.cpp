// mcx_fortran.cpp -- the reference's Fortran/C energy interface
// (src/FORTRAN_interface.cpp:21-106) on top of the mcx_* C ABI, with a reader
// of the MDBAS COR and forfield files (IO_MDBAS::read_coord / read_ff,
// src/IO_MDBAS.cpp:77-505) so that the calls take file names as the
// reference's do.
#include "../../include/mcx.h"
#include "../../include/mcx_fortran.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

constexpr double kKcalToIu = 418.4;              // include/Constants.hpp:29
constexpr double kPi = 3.141592653589793;         // include/Constants.hpp:41

struct Terms {
    std::vector<int> a, b, c, d, type;
    std::vector<double> k, r0, beta, t0, phi0, mult;
};

// One process-wide system, as the reference keeps its static singletons.
struct FortranState {
    int pbtype = MCX_PBC_NONE;
    double box[3] = {0.0, 0.0, 0.0};
    int natom = 0;
    double temp = 300.0;
    int cutmode = MCX_CUT_FULL;
    std::vector<double> mass;
    mcx_ctx *h = nullptr;
    std::string error;
};
FortranState g;

std::string squeeze_lower(const char *s, bool lower)
{
    std::string out;
    for (const char *p = s; p && *p; p++) {
        const unsigned char c = (unsigned char)*p;
        if (std::isspace(c)) continue;                // Tools::str_rm_blank_spaces
        out.push_back(lower ? (char)std::tolower(c) : (char)c);
    }
    return out;
}

void check(int rc, const char *what)
{
    if (rc != MCX_OK) {
        std::string msg = std::string(what) + ": ";
        msg += g.h ? mcx_last_error(g.h) : "no handle";
        throw std::runtime_error(msg);
    }
}

// IO_MDBAS::read_coord (src/IO_MDBAS.cpp:77-127): header lines start with
// '*'; then the atom count; then 10 fields per atom.
void read_cor(const std::string &path, int natom, std::vector<double> &x, std::vector<double> &y,
              std::vector<double> &z, std::vector<int> &resid)
{
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open coordinates file " + path);
    std::string line;
    while (std::getline(f, line))
        if (line.empty() || line[0] != '*') break;
    const int n = std::atoi(line.c_str());
    if (n != natom)
        throw std::runtime_error("number of atoms at the top of coordinates file differs from set_ens_");
    x.resize(n); y.resize(n); z.resize(n); resid.resize(n);
    for (int i = 0; i < n; i++) {
        int id, ire, res;
        std::string ren, atl, sen;
        double wei;
        if (!(f >> id >> ire >> ren >> atl >> x[i] >> y[i] >> z[i] >> sen >> res >> wei))
            throw std::runtime_error("IO error on coordinates file " + path);
        resid[i] = ire;
    }
}

// IO_MDBAS::read_ff (src/IO_MDBAS.cpp:130-505) with its unit conversions:
// force constants x 418.4, angles in radians, eps = sqrt(eps_kcal * 418.4).
struct ForceField {
    std::vector<double> charge, mass, eps, sig, eps14, sig14;
    std::vector<int> frozen;
    Terms bonds, ub, angles, dihe, impr;
};

ForceField read_ff(const std::string &path, int n)
{
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open forcefield file " + path);
    ForceField F;
    F.charge.assign(n, 0.0); F.mass.assign(n, 0.0); F.eps.assign(n, 0.0); F.sig.assign(n, 0.0);
    F.eps14.assign(n, 0.0); F.sig14.assign(n, 0.0); F.frozen.assign(n, 0);
    std::string line;
    auto records = [&](int count) {
        std::vector<std::vector<std::string>> out;
        while ((int)out.size() < count) {
            if (!std::getline(f, line)) throw std::runtime_error("truncated forcefield file " + path);
            if (!line.empty() && line[0] == '#') continue;
            std::istringstream ss(line);
            std::vector<std::string> t;
            for (std::string w; ss >> w;) t.push_back(w);
            out.push_back(t);
        }
        return out;
    };
    auto need = [&](const std::vector<std::string> &t, size_t k) {
        if (t.size() < k) throw std::runtime_error("short record in forcefield file " + path);
    };
    auto atom = [&](const std::string &s) {
        const int i = std::stoi(s) - 1;
        if (i < 0 || i >= n) throw std::runtime_error("atom index out of range in " + path);
        return i;
    };
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream ss(line);
        std::string kw;
        int count = 0;
        ss >> kw;
        if (kw.empty()) continue;
        std::transform(kw.begin(), kw.end(), kw.begin(), [](unsigned char c) { return (char)std::tolower(c); });
        if (kw == "end") break;
        ss >> count;
        if (kw == "atoms") {
            for (auto &t : records(count)) {
                need(t, 6);
                const int i = atom(t[0]);
                F.charge[i] = std::stod(t[3]); F.mass[i] = std::stod(t[4]); F.frozen[i] = std::stoi(t[5]) != 0;
            }
        } else if (kw == "bonds") {
            for (auto &t : records(count)) {
                need(t, 6);
                F.bonds.a.push_back(atom(t[0])); F.bonds.b.push_back(atom(t[1])); F.bonds.type.push_back(std::stoi(t[2]));
                F.bonds.k.push_back(std::stod(t[3]) * kKcalToIu); F.bonds.r0.push_back(std::stod(t[4]));
                F.bonds.beta.push_back(std::stod(t[5]));
            }
        } else if (kw == "constraints") {
            if (count != 0) throw std::runtime_error("a non-empty Constraints section is not readable by the reference");
        } else if (kw == "urey-bradley") {
            for (auto &t : records(count)) {
                need(t, 4);
                F.ub.a.push_back(atom(t[0])); F.ub.b.push_back(atom(t[1]));
                F.ub.k.push_back(std::stod(t[2]) * kKcalToIu); F.ub.r0.push_back(std::stod(t[3]));
            }
        } else if (kw == "angles") {
            for (auto &t : records(count)) {
                need(t, 6);
                F.angles.a.push_back(atom(t[0])); F.angles.b.push_back(atom(t[1])); F.angles.c.push_back(atom(t[2]));
                F.angles.type.push_back(std::stoi(t[3]));
                F.angles.k.push_back(std::stod(t[4]) * kKcalToIu);
                F.angles.t0.push_back(std::stod(t[5]) * kPi / 180.);
            }
        } else if (kw == "dihedrals" || kw == "impropers") {
            Terms &T = kw == "dihedrals" ? F.dihe : F.impr;
            for (auto &t : records(count)) {
                need(t, 9);
                T.a.push_back(atom(t[0])); T.b.push_back(atom(t[1])); T.c.push_back(atom(t[2])); T.d.push_back(atom(t[3]));
                T.type.push_back(std::stoi(t[4]));
                T.k.push_back(std::stod(t[6]) * kKcalToIu);
                T.phi0.push_back(std::stod(t[7]) * kPi / 180.);
                T.mult.push_back(std::stod(t[8]));
            }
        } else if (kw == "vdw") {
            for (auto &t : records(count)) {
                need(t, 7);
                const int i = atom(t[0]);
                F.eps[i] = std::sqrt(std::stod(t[2]) * kKcalToIu); F.sig[i] = std::stod(t[3]);
                F.eps14[i] = std::sqrt(std::stod(t[5]) * kKcalToIu); F.sig14[i] = std::stod(t[6]);
            }
        } else {
            throw std::runtime_error("unknown forcefield keyword '" + kw + "' in " + path);
        }
    }
    return F;
}

template <class F>
void guarded(F &&body)
{
    try {
        body();
    } catch (const std::exception &e) {
        g.error = e.what();
    }
}

}  // namespace

extern "C" {

void set_pbc_(char *type, double *a, double *b, double *c, double *alpha, double *beta, double *gamma)
{
    (void)alpha; (void)beta; (void)gamma;
    guarded([&] {
        g.error.clear();
        const std::string t = squeeze_lower(type, true);
        if (t == "cubic") { g.pbtype = MCX_PBC_CUBIC; g.box[0] = g.box[1] = g.box[2] = *a; }
        else if (t == "orbic") { g.pbtype = MCX_PBC_ORBIC; g.box[0] = *a; g.box[1] = *b; g.box[2] = *c; }
        else { g.pbtype = MCX_PBC_NONE; g.box[0] = *a; g.box[1] = *b; g.box[2] = *c; }   // "none" and unknown types
    });
}

void set_ens_(int *natom, double *Temp)
{
    guarded([&] {
        g.error.clear();
        g.natom = *natom;
        g.temp = *Temp;
    });
}

void set_ff_and_cor_(char *cutoffmode, double *ctoff, double *cton, double *dcut, char *FF_file, char *COR_file)
{
    guarded([&] {
        g.error.clear();
        if (g.h) { mcx_destroy(g.h); g.h = nullptr; }
        const std::string cut = squeeze_lower(cutoffmode, true);
        g.cutmode = cut == "switch" ? MCX_CUT_SWITCH : MCX_CUT_FULL;
        if (cut != "switch" && cut != "full") throw std::runtime_error("cutoff mode must be 'full' or 'switch'");
        const int n = g.natom;
        std::vector<double> x, y, z;
        std::vector<int> resid;
        read_cor(squeeze_lower(COR_file, false), n, x, y, z, resid);
        ForceField F = read_ff(squeeze_lower(FF_file, false), n);
        check(mcx_create(n, 0, &g.h), "mcx_create");
        check(mcx_set_box(g.h, g.pbtype, g.box[0], g.box[1], g.box[2]), "mcx_set_box");
        check(mcx_set_cut(g.h, g.cutmode, *ctoff, *cton, *dcut, g.temp), "mcx_set_cut");
        check(mcx_set_atoms(g.h, F.charge.data(), F.eps.data(), F.sig.data(), F.eps14.data(), F.sig14.data(),
                            F.mass.data(), F.frozen.data(), resid.data()), "mcx_set_atoms");
        check(mcx_set_bonds(g.h, (int)F.bonds.a.size(), F.bonds.a.data(), F.bonds.b.data(), F.bonds.type.data(),
                            F.bonds.k.data(), F.bonds.r0.data(), F.bonds.beta.data()), "mcx_set_bonds");
        check(mcx_set_ub(g.h, (int)F.ub.a.size(), F.ub.a.data(), F.ub.b.data(), F.ub.k.data(), F.ub.r0.data()),
              "mcx_set_ub");
        check(mcx_set_angles(g.h, (int)F.angles.a.size(), F.angles.a.data(), F.angles.b.data(), F.angles.c.data(),
                             F.angles.k.data(), F.angles.t0.data()), "mcx_set_angles");
        check(mcx_set_dihedrals(g.h, (int)F.dihe.a.size(), F.dihe.a.data(), F.dihe.b.data(), F.dihe.c.data(),
                                F.dihe.d.data(), F.dihe.type.data(), F.dihe.k.data(), F.dihe.phi0.data(),
                                F.dihe.mult.data()), "mcx_set_dihedrals");
        check(mcx_set_impropers(g.h, (int)F.impr.a.size(), F.impr.a.data(), F.impr.b.data(), F.impr.c.data(),
                                F.impr.d.data(), F.impr.type.data(), F.impr.k.data(), F.impr.phi0.data(),
                                F.impr.mult.data()), "mcx_set_impropers");
        check(mcx_build_exclusions(g.h), "mcx_build_exclusions");
        check(mcx_set_coords(g.h, x.data(), y.data(), z.data()), "mcx_set_coords");
        if (g.cutmode == MCX_CUT_SWITCH) {
            long long np = 0;
            check(mcx_update_list(g.h, &np), "mcx_update_list");   // List_nonBonded builds the list
        }
        g.mass = F.mass;
    });
}

void get_energy_(double ener[10])
{
    guarded([&] {
        for (int k = 0; k < 10; k++) ener[k] = std::numeric_limits<double>::quiet_NaN();
        if (!g.h) throw std::runtime_error("get_energy_ before set_ff_and_cor_");
        check(mcx_energy(g.h, ener), "mcx_energy");
        for (int k = 0; k < 10; k++) ener[k] /= kKcalToIu;
    });
}

void set_coords_(double x[], double y[], double z[])
{
    guarded([&] {
        if (!g.h) throw std::runtime_error("set_coords_ before set_ff_and_cor_");
        // AtomList::setCoords only: the list built at set-up stays, as the
        // reference's does (src/FORTRAN_interface.cpp:77-82)
        check(mcx_set_coords_moved(g.h, x, y, z), "mcx_set_coords_moved");
    });
}

void get_mass_(double mass[])
{
    guarded([&] {
        if (!g.h) throw std::runtime_error("get_mass_ before set_ff_and_cor_");
        std::copy(g.mass.begin(), g.mass.end(), mass);
    });
}

void clean_(void)
{
    if (g.h) mcx_destroy(g.h);
    g = FortranState();
}

const char *mcx_f_last_error(void)
{
    return g.error.c_str();
}

}  // extern "C"
