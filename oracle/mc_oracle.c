/*
 * mc_oracle.c -- FP64 CPU restatement of the mc_cpp energy hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the
 * product in mc_cpp_amd/.
 *
 * Build: oracle/Makefile (-O2 -ffp-contract=off -fopenmp).  Contraction is
 * off so every expression rounds exactly as the reference's source text
 * states; the reference's own CMake build (-O3 -march=native) would let GCC
 * contract, which makes its low bits build-dependent.
 *
 * Reference citations are relative to the mc_cpp source tree.
 */
#include "mc_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static const double KCALTOIU = 418.4;        /* include/Constants.hpp:29 */
static const double RBOLTZUI = 0.83144621;   /* include/Constants.hpp:35 */
static const double CHGCHARMM = 332.0716;    /* include/Constants.hpp:37 */
static const double PI_ = 3.141592653589793; /* include/Constants.hpp:41 */
static const double TWOPI = 6.283185307179586; /* include/Constants.hpp:42 */

/* ------------------------------------------------------------------ */
/* L1 geometry                                                          */
/* ------------------------------------------------------------------ */

/* Tools::rint, include/Tools.hpp:221-224: truncation after an FP64 +-0.5 */
double orc_rint(double x)
{
    return (double)(x >= 0. ? (int)(x + 0.5) : (int)(x - 0.5));
}

/* PerConditions::set_pbc_vectors, src/PerConditions.cpp:103-134 */
void orc_set_box(orc_system *s, int pbtype, double a, double b, double c)
{
    s->pbtype = pbtype;
    if (pbtype == ORC_PBC_CUBIC) { b = a; c = a; }
    s->pbx = a; s->pby = b; s->pbz = c;
    if (pbtype == ORC_PBC_CUBIC || pbtype == ORC_PBC_ORBIC) {
        s->rpbx = 1.0 / s->pbx; s->rpby = 1.0 / s->pby; s->rpbz = 1.0 / s->pbz;
    } else {
        s->rpbx = s->rpby = s->rpbz = 0.0;
    }
}

/* PerConditions::applyPBC(double[3]), src/PerConditions.cpp:244-265 */
void orc_apply_pbc(const orc_system *s, double d[3])
{
    if (s->pbtype == ORC_PBC_CUBIC || s->pbtype == ORC_PBC_ORBIC) {
        d[0] -= s->pbx * orc_rint(s->rpbx * d[0]);
        d[1] -= s->pby * orc_rint(s->rpby * d[1]);
        d[2] -= s->pbz * orc_rint(s->rpbz * d[2]);
    }
}

static inline void getc3(const orc_system *s, int i, double d[3])
{
    d[0] = s->x[i]; d[1] = s->y[i]; d[2] = s->z[i];
}

/* Tools::distance2, include/Tools.hpp:151-188: delta = a2 - a1, PBC, r2 */
static inline double distance2v(const orc_system *s, const double a1[3], const double a2[3], double delta[3])
{
    delta[0] = a2[0] - a1[0];
    delta[1] = a2[1] - a1[1];
    delta[2] = a2[2] - a1[2];
    orc_apply_pbc(s, delta);
    return delta[0] * delta[0] + delta[1] * delta[1] + delta[2] * delta[2];
}

static inline double distance2(const orc_system *s, const double a1[3], const double a2[3])
{
    double d[3];
    return distance2v(s, a1, a2, d);
}

static inline double X2(double x) { return x * x; }
static inline double X3(double x) { return x * x * x; }
static inline double X6(double x) { return x * x * x * x * x * x; }
static inline double X12(double x) { return x * x * x * x * x * x * x * x * x * x * x * x; }

/* ------------------------------------------------------------------ */
/* R9: exclusion lists, src/List_nonBonded.cpp:143-775                  */
/* ------------------------------------------------------------------ */

typedef struct ivec { int *v; int n, cap; } ivec;
static void iv_push(ivec *a, int x)
{
    if (a->n == a->cap) { a->cap = a->cap ? 2 * a->cap : 16; a->v = (int *)realloc(a->v, sizeof(int) * a->cap); }
    a->v[a->n++] = x;
}
static int iv_has(const ivec *a, int x)
{
    for (int k = 0; k < a->n; k++) if (a->v[k] == x) return 1;
    return 0;
}

/* the "if not already in tempAtom[ii], add to both" block repeated through
 * excl_angles/dihedrals/impropers/connectivity (e.g. :235-251) */
static int excl_add(ivec *tmp, int ii, int jj)
{
    if (iv_has(&tmp[ii], jj)) return 0;
    iv_push(&tmp[ii], jj);
    iv_push(&tmp[jj], ii);
    return 1;
}

typedef struct pvec { int *v; int n, cap; } pvec;
static void pv_push(pvec *a, int x, int y)
{
    if (a->n == a->cap) { a->cap = a->cap ? 2 * a->cap : 16; a->v = (int *)realloc(a->v, sizeof(int) * 2 * a->cap); }
    a->v[2 * a->n] = x; a->v[2 * a->n + 1] = y; a->n++;
}

void orc_free_lists(orc_system *s)
{
    free(s->exclPair); free(s->exclOff); free(s->exclIdx); free(s->neighList14);
    free(s->neighPair); free(s->neighList);
    s->exclPair = s->exclOff = s->exclIdx = s->neighList14 = NULL;
    s->neighPair = s->neighList = NULL;
    s->cap = 0;
}

int orc_build_exclusions(orc_system *s)
{
    const int N = s->natom;
    ivec *tmp = (ivec *)calloc(N, sizeof(ivec));
    ivec *con = (ivec *)calloc(N, sizeof(ivec));
    pvec v14 = {0, 0, 0};

    /* excl_bonds :184-213 (unconditional append, also connectivity) */
    for (int i = 0; i < s->nbond; i++) {
        int ii = s->bond_a[i], jj = s->bond_b[i];
        iv_push(&con[ii], jj); iv_push(&con[jj], ii);
        iv_push(&tmp[ii], jj); iv_push(&tmp[jj], ii);
    }
    /* excl_angles :215-299 */
    for (int i = 0; i < s->nangle; i++) {
        int a = s->ang_a[i], b = s->ang_b[i], c = s->ang_c[i];
        excl_add(tmp, a, b); excl_add(tmp, a, c); excl_add(tmp, b, c);
    }
    /* excl_dihedrals :301-462 (the a-d pair feeds the 1-4 list) */
    for (int i = 0; i < s->ndihe; i++) {
        int a = s->dih_a[i], b = s->dih_b[i], c = s->dih_c[i], d = s->dih_d[i];
        excl_add(tmp, a, b); excl_add(tmp, a, c);
        if (excl_add(tmp, a, d)) pv_push(&v14, a, d);
        excl_add(tmp, b, c); excl_add(tmp, b, d); excl_add(tmp, c, d);
    }
    /* excl_impropers :464-619 */
    for (int i = 0; i < s->nimpr; i++) {
        int a = s->imp_a[i], b = s->imp_b[i], c = s->imp_c[i], d = s->imp_d[i];
        excl_add(tmp, a, b); excl_add(tmp, a, c); excl_add(tmp, a, d);
        excl_add(tmp, b, c); excl_add(tmp, b, d); excl_add(tmp, c, d);
    }
    /* excl_connectivity :621-703 */
    for (int ia = 0; ia < N; ia++) {
        for (int j = 0; j < con[ia].n; j++) {
            int ib = con[ia].v[j];
            for (int k = 0; k < con[ib].n; k++) {
                int ic = con[ib].v[k];
                if (ic == ia) continue;
                excl_add(tmp, ia, ic);
                for (int l = 0; l < con[ic].n; l++) {
                    int id = con[ic].v[l];
                    if (id == ia || id == ib) continue;
                    if (excl_add(tmp, ia, id)) pv_push(&v14, ia, id);
                }
            }
        }
    }

    free(s->exclPair); free(s->exclOff); free(s->exclIdx); free(s->neighList14);
    s->exclPair = (int *)calloc(N, sizeof(int));
    s->exclOff = (int *)calloc(N + 1, sizeof(int));
    for (int i = 0; i < N; i++) s->exclOff[i + 1] = s->exclOff[i] + tmp[i].n + 1;
    s->exclIdx = (int *)calloc(s->exclOff[N] > 0 ? s->exclOff[N] : 1, sizeof(int));

    /* excl_final_Lists :705-766: circular half rule + sorted insertion */
    const int hn = N / 2, hm1 = (N - 1) / 2;
    for (int i = 0; i < N; i++) {
        int *L = s->exclIdx + s->exclOff[i];
        int jj = 0;
        for (int j = 0; j < tmp[i].n; j++) {
            int e = tmp[i].v[j];
            if (((e > i) && ((e - i) < hn)) || ((e < i) && ((e - i + N) < hm1))) {
                L[jj] = e;
                for (int k = jj; k > 0; k--) {
                    if (L[k] < L[k - 1]) { int t = L[k]; L[k] = L[k - 1]; L[k - 1] = t; }
                }
                jj++;
            }
        }
        s->exclPair[i] = jj;
    }
    /* :748-766 -- entries < i are shifted out; the saved entry is written one
     * past the visible end (index exclPair[i]), so the visible row loses it
     * and repeats its last element.  Row capacity tmpPair+1 holds that write. */
    for (int i = 0; i < N; i++) {
        int *L = s->exclIdx + s->exclOff[i];
        const int np = s->exclPair[i];
        for (int jj = 0; jj < np; jj++) {
            if (L[0] < i) {
                int e = L[0];
                for (int kk = 0; kk < np - 1; kk++) L[kk] = L[kk + 1];
                L[np] = e;
            }
        }
    }
    /* 1-4 list :768-773 */
    s->nPair14 = v14.n;
    s->pair14_overflow = v14.n > 5 * s->ndihe;
    s->neighList14 = (int *)malloc(sizeof(int) * (2 * v14.n + 1));
    if (v14.n) memcpy(s->neighList14, v14.v, sizeof(int) * 2 * v14.n);

    for (int i = 0; i < N; i++) { free(tmp[i].v); free(con[i].v); }
    free(tmp); free(con); free(v14.v);
    return 0;
}

static inline int excl_member(const orc_system *s, int i, int j)
{
    const int *L = s->exclIdx + s->exclOff[i];
    for (int k = 0; k < s->exclPair[i]; k++) if (L[k] == j) return 1;
    return 0;
}

/* ------------------------------------------------------------------ */
/* R10/R11: Verlet list, src/List_nonBonded.cpp:777-915                 */
/* ------------------------------------------------------------------ */

long long orc_update_verlet_list(orc_system *s, int nthreads)
{
    const int N = s->natom;
    double cutnb2 = s->ctoff + s->dcut;
    cutnb2 *= cutnb2;
    free(s->neighPair);
    s->neighPair = (int *)calloc(N, sizeof(int));

    /* pass 1 counts (init_verlet_list :789-825), pass 2 fills (:852-911) */
    int maxc = 0;
#ifdef _OPENMP
    #pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1) reduction(max:maxc)
#endif
    for (int i = 0; i < N; i++) {
        double di[3], dj[3];
        getc3(s, i, di);
        int c = 0;
        for (int j = i + 1; j < N; j++) {
            getc3(s, j, dj);
            double r2 = distance2(s, di, dj);
            if (r2 <= cutnb2) {
                int ex = 0;
                if (s->frozen[i] && s->frozen[j]) ex = 1;
                else ex = excl_member(s, i, j);
                if (!ex) c++;
            }
        }
        s->neighPair[i] = c;
        if (c > maxc) maxc = c;
    }
    s->sizeList = (int)(maxc * (1. + 2. * 0.02)) + 1;
    if (maxc + 1 > s->cap || s->neighList == NULL) {
        free(s->neighList);
        s->cap = maxc + 1;
        s->neighList = (int *)malloc(sizeof(int) * (size_t)N * s->cap);
    }
    long long total = 0;
#ifdef _OPENMP
    #pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1) reduction(+:total)
#endif
    for (int i = 0; i < N; i++) {
        double di[3], dj[3];
        getc3(s, i, di);
        int *row = s->neighList + (size_t)i * s->cap;
        int c = 0;
        for (int j = i + 1; j < N; j++) {
            getc3(s, j, dj);
            double r2 = distance2(s, di, dj);
            if (r2 <= cutnb2) {
                int ex = 0;
                if (s->frozen[i] && s->frozen[j]) ex = 1;
                else ex = excl_member(s, i, j);
                if (!ex) row[c++] = j;
            }
        }
        total += c;
    }
    return total;
}

/* ------------------------------------------------------------------ */
/* R3-R6 nonbonded, src/FField_MDBAS.cpp:194-488                        */
/* ------------------------------------------------------------------ */

static inline double e_elec(double qi, double qj, double rt) { return qi * qj * rt; }   /* :321-325 */
static inline double e_vdw(double ei, double ej, double si, double sj, double rt)     /* :328-334 */
{
    return ei * ej * (X12((si + sj) * rt) - X6((si + sj) * rt));
}

/* computeNonBonded_full :194-265, including the sorted-pointer exclusion walk */
void orc_nonbonded_full(orc_system *s)
{
    double lelec = 0., lvdw = 0.;
    const int N = s->natom;
    for (int i = 0; i < N - 1; i++) {
        double di[3], dj[3];
        getc3(s, i, di);
        const double qi = s->charge[i], epsi = s->eps[i], sigi = s->sig[i];
        const int *L = s->exclIdx + s->exclOff[i];
        const int np = s->exclPair[i];
        int k = 0;
        for (int j = i + 1; j < N; j++) {
            int exclude = 0;
            if ((np > 0) && (L[k] == j)) {
                exclude = 1;
                k++;
                if (k >= np) k = np - 1;
            }
            if (!exclude) {
                getc3(s, j, dj);
                double rt = distance2(s, di, dj);
                rt = sqrt(rt);
                rt = 1. / rt;
                lelec += e_elec(qi, s->charge[j], rt);
                lvdw += e_vdw(epsi, s->eps[j], sigi, s->sig[j], rt);
            }
        }
    }
    s->elec = CHGCHARMM * KCALTOIU * lelec;
    s->vdw = 4.0 * lvdw;
}

/* computeNonBonded14 :267-318 */
void orc_nonbonded14(orc_system *s)
{
    double lelec = 0., levdw = 0.;
    for (int k = 0; k < s->nPair14; k++) {
        int i = s->neighList14[2 * k], j = s->neighList14[2 * k + 1];
        double di[3], dj[3];
        getc3(s, i, di); getc3(s, j, dj);
        double r2 = distance2(s, di, dj);
        double r = sqrt(r2);
        double rt = 1. / r;
        lelec += e_elec(s->charge[i], s->charge[j], rt);
        levdw += e_vdw(s->eps14[i], s->eps14[j], s->sig14[i], s->sig14[j], rt);
    }
    s->elec += CHGCHARMM * KCALTOIU * lelec;
    s->vdw += 4.0 * levdw;
}

/* one switched pair, :383-405 */
static inline void switched_pair(const orc_system *s, double r2, double qi, double qj,
                                 double ei, double ej, double si, double sj,
                                 double ctoff2, double cton2, double switch2,
                                 double *pe, double *pv)
{
    double r = sqrt(r2);
    double rt = 1. / r;
    double pelec = e_elec(qi, qj, rt);
    double pvdw = e_vdw(ei, ej, si, sj, rt);
    double sw = 1.0;
    if (r2 > cton2) {
        double switch1 = ctoff2 - r2;
        sw = X2(switch1) * (ctoff2 + 2. * r2 - 3. * cton2) * switch2;
    }
    *pe = pelec * sw;
    *pv = pvdw * sw;
}

/* Threads of the switched pair loops (orc_nonbonded_switch and the SWITCH
 * branch of orc_moving_set_energy).  1, the default, keeps the reference's
 * serial summation order bit for bit.  More deal the rows out dynamically
 * with per-thread sums, the OpenMP form the reference gives its list build
 * (src/List_nonBonded.cpp:847-851) and leaves commented out on its energy
 * loops (src/FField_MDBAS.cpp:211): the totals then differ by rounding
 * only.  Used to time the CPU on all cores. */
static int g_energy_threads = 1;

void orc_set_energy_threads(int n) { g_energy_threads = n > 1 ? n : 1; }

/* one row of computeNonBonded_switch :336-424, entries k0..k1-1, summed
 * into *le / *lv pair by pair */
static inline void switch_row(const orc_system *s, int i, int k0, int k1, double ctoff2, double cton2,
                              double switch2, double *le, double *lv)
{
    double di[3], dj[3];
    getc3(s, i, di);
    const double qi = s->charge[i], epsi = s->eps[i], sigi = s->sig[i];
    const int *row = s->neighList + (size_t)i * s->cap;
    for (int k = k0; k < k1; k++) {
        int j = row[k];
        getc3(s, j, dj);
        double r2 = distance2(s, di, dj);
        if (r2 <= ctoff2) {
            double pe, pv;
            switched_pair(s, r2, qi, s->charge[j], epsi, s->eps[j], sigi, s->sig[j],
                          ctoff2, cton2, switch2, &pe, &pv);
            *le += pe;
            *lv += pv;
        }
    }
}

/* computeNonBonded_switch :336-424 */
void orc_nonbonded_switch(orc_system *s)
{
    double lelec = 0., levdw = 0.;
    const double ctoff2 = s->ctoff * s->ctoff;
    const double cton2 = s->cton * s->cton;
    const double switch2 = 1. / (X3(ctoff2 - cton2));
    const int nt = g_energy_threads;
    if (nt > 1) {
        #pragma omp parallel for schedule(dynamic, 16) num_threads(nt) reduction(+:lelec, levdw)
        for (int i = 0; i < s->natom; i++)
            switch_row(s, i, 0, s->neighPair[i], ctoff2, cton2, switch2, &lelec, &levdw);
    } else {
        for (int i = 0; i < s->natom; i++)
            switch_row(s, i, 0, s->neighPair[i], ctoff2, cton2, switch2, &lelec, &levdw);
    }
    s->elec = CHGCHARMM * KCALTOIU * lelec;
    s->vdw = 4.0 * levdw;
}

/* computeNonBonded14_switch :426-488 */
void orc_nonbonded14_switch(orc_system *s)
{
    double lelec = 0., levdw = 0.;
    const double ctoff2 = s->ctoff * s->ctoff;
    const double cton2 = s->cton * s->cton;
    const double switch2 = 1. / (X3(ctoff2 - cton2));
    for (int k = 0; k < s->nPair14; k++) {
        int i = s->neighList14[2 * k], j = s->neighList14[2 * k + 1];
        double di[3], dj[3];
        getc3(s, i, di); getc3(s, j, dj);
        double r2 = distance2(s, di, dj);
        if (r2 <= ctoff2) {
            double pe, pv;
            switched_pair(s, r2, s->charge[i], s->charge[j], s->eps14[i], s->eps14[j],
                          s->sig14[i], s->sig14[j], ctoff2, cton2, switch2, &pe, &pv);
            lelec += pe;
            levdw += pv;
        }
    }
    s->elec += CHGCHARMM * KCALTOIU * lelec;
    s->vdw += 4.0 * levdw;
}

/* ------------------------------------------------------------------ */
/* R7 bonded, src/FField_MDBAS.cpp:490-786                              */
/* ------------------------------------------------------------------ */

static double bond_term(const orc_system *s, int ll)
{
    double di[3], dj[3];
    getc3(s, s->bond_a[ll], di); getc3(s, s->bond_b[ll], dj);
    double d = sqrt(distance2(s, di, dj));
    double r0 = s->bond_r0[ll], k = s->bond_k[ll];
    if (s->bond_type[ll] == 1) {               /* BMORSE :519-527 */
        double beta = s->bond_beta[ll];
        double ma = exp(-beta * (d - r0));
        double mb = X2(ma);
        return k * (mb - 2. * ma) + k;
    }
    return 0.5 * k * X2(d - r0);                /* BHARM and default :515-531 */
}

static double angle_term(const orc_system *s, int ll)            /* :547-572 */
{
    double di[3], dj[3], dk[3], dab[3], dbc[3];
    getc3(s, s->ang_a[ll], di); getc3(s, s->ang_b[ll], dj); getc3(s, s->ang_c[ll], dk);
    double rab = sqrt(distance2v(s, di, dj, dab));
    double rbc = sqrt(distance2v(s, dk, dj, dbc));
    double cost = (dab[0] * dbc[0] + dab[1] * dbc[1] + dab[2] * dbc[2]) / (rab * rbc);
    double theta = acos(cost);
    return 0.5 * s->ang_k[ll] * X2(theta - s->ang_t0[ll]);
}

static double ub_term(const orc_system *s, int ll)               /* :584-598 */
{
    double di[3], dj[3];
    getc3(s, s->ub_a[ll], di); getc3(s, s->ub_b[ll], dj);
    double d = sqrt(distance2(s, di, dj));
    return 0.5 * s->ub_k[ll] * X2(d - s->ub_r0[ll]);
}

/* shared body of computeEdihe :617-691 and computeEimpr :710-784 */
static double torsion_term(const orc_system *s, int i, int j, int k, int l,
                           int type, double kst, double phi0, double mult)
{
    double di[3], dj[3], dk[3], dl[3], dab[3], dbc[3], dcd[3], pb[3], pc[3];
    getc3(s, i, di); getc3(s, j, dj); getc3(s, k, dk); getc3(s, l, dl);
    dab[0] = di[0] - dj[0]; dab[1] = di[1] - dj[1]; dab[2] = di[2] - dj[2];   /* vec_substract(dj,di) */
    orc_apply_pbc(s, dab);
    double rbc = sqrt(distance2v(s, dk, dj, dbc));
    dcd[0] = dk[0] - dl[0]; dcd[1] = dk[1] - dl[1]; dcd[2] = dk[2] - dl[2];   /* vec_substract(dl,dk) */
    orc_apply_pbc(s, dcd);
    pb[0] = dab[1] * dbc[2] - dab[2] * dbc[1];
    pb[1] = dab[2] * dbc[0] - dab[0] * dbc[2];
    pb[2] = dab[0] * dbc[1] - dab[1] * dbc[0];
    double r2pb = X2(pb[0]) + X2(pb[1]) + X2(pb[2]);
    double rpb = sqrt(r2pb);
    pc[0] = dbc[1] * dcd[2] - dbc[2] * dcd[1];
    pc[1] = dbc[2] * dcd[0] - dbc[0] * dcd[2];
    pc[2] = dbc[0] * dcd[1] - dbc[1] * dcd[0];
    double r2pc = X2(pc[0]) + X2(pc[1]) + X2(pc[2]);
    double rpc = sqrt(r2pc);
    double pbpc = pb[0] * pc[0] + pb[1] * pc[1] + pb[2] * pc[2];
    double cosp = pbpc / (rpb * rpc);
    double sinp = (dbc[0] * (pc[1] * pb[2] - pc[2] * pb[1]) + dbc[1] * (pb[0] * pc[2] - pb[2] * pc[0]) +
                   dbc[2] * (pc[0] * pb[1] - pc[1] * pb[0])) / (rpb * rpc * rbc);
    double phi = atan2(sinp, cosp);
    if (type == 2) {                             /* DHARM */
        phi = phi - phi0;
        phi = phi - orc_rint(phi / TWOPI) * TWOPI;
        return 0.5 * kst * (phi * phi);
    }
    return kst * (1. + cos(mult * phi - phi0)); /* DCOS and default */
}

void orc_ebond(orc_system *s) { double e = 0.; for (int l = 0; l < s->nbond; l++) e += bond_term(s, l); s->bond = e; }
void orc_eang(orc_system *s) { double e = 0.; for (int l = 0; l < s->nangle; l++) e += angle_term(s, l); s->ang = e; }
void orc_eub(orc_system *s) { double e = 0.; for (int l = 0; l < s->nub; l++) e += ub_term(s, l); s->ub = e; }
void orc_edihe(orc_system *s)
{
    double e = 0.;
    for (int l = 0; l < s->ndihe; l++)
        e += torsion_term(s, s->dih_a[l], s->dih_b[l], s->dih_c[l], s->dih_d[l], s->dih_type[l],
                          s->dih_k[l], s->dih_phi0[l], s->dih_mult[l]);
    s->dihe = e;
}
void orc_eimpr(orc_system *s)
{
    double e = 0.;
    for (int l = 0; l < s->nimpr; l++)
        e += torsion_term(s, s->imp_a[l], s->imp_b[l], s->imp_c[l], s->imp_d[l], s->imp_type[l],
                          s->imp_k[l], s->imp_phi0[l], s->imp_mult[l]);
    s->impr = e;
}

/* getE(double[10]) / getEtot / getEswitch, src/FField_MDBAS.cpp:77-192 */
double orc_getE(orc_system *s, double ener[10])
{
    /* the reference builds the list in the List_nonBonded constructor
     * (src/List_nonBonded.cpp:62-100) before any getE */
    if (s->cutmode == ORC_CUT_SWITCH && s->neighPair == NULL) orc_update_verlet_list(s, 1);
    if (s->cutmode == ORC_CUT_FULL) { orc_nonbonded_full(s); orc_nonbonded14(s); }
    else { orc_nonbonded_switch(s); orc_nonbonded14_switch(s); }
    if (s->nbond > 0) orc_ebond(s);
    if (s->nangle > 0) orc_eang(s);
    if (s->nub > 0) orc_eub(s);
    if (s->ndihe > 0) orc_edihe(s);
    if (s->nimpr > 0) orc_eimpr(s);
    s->pot = s->elec + s->vdw + s->bond + s->ang + s->ub + s->dihe + s->impr;
    s->tot = s->pot + s->kin;
    if (ener) {
        ener[0] = s->tot; ener[1] = s->pot; ener[2] = s->kin; ener[3] = s->elec; ener[4] = s->vdw;
        ener[5] = s->bond; ener[6] = s->ang; ener[7] = s->ub; ener[8] = s->dihe; ener[9] = s->impr;
    }
    return s->tot;
}

/* ------------------------------------------------------------------ */
/* Moving-set energy: the terms of getE that touch atoms [first,last].  */
/* The reference has this only as commented-out code                    */
/* (src/MC_metropolis.cpp:93,121; src/FField_MDBAS.cpp:788-902); here  */
/* E_set(new) - E_set(old) equals getE(new) - getE(old) term by term.   */
/* ------------------------------------------------------------------ */

static int lower_bound_row(const int *row, int n, int v)
{
    int lo = 0, hi = n;
    while (lo < hi) { int m = (lo + hi) / 2; if (row[m] < v) lo = m + 1; else hi = m; }
    return lo;
}

/* 1-4 pairs and bonded terms touching [first,last] (used by both forms of
 * the moving-set energy below); returns elec_raw/vdw_raw of the 1-4 pairs
 * through lelec and levdw and the bonded sum as the result. */
static double moving_set_rest(orc_system *s, int first, int last, double *lelec, double *levdw)
{
#define INSET(a) ((a) >= first && (a) <= last)
    if (s->cutmode == ORC_CUT_SWITCH) {
        const double ctoff2 = s->ctoff * s->ctoff, cton2 = s->cton * s->cton;
        const double switch2 = 1. / (X3(ctoff2 - cton2));
        for (int k = 0; k < s->nPair14; k++) {
            int i = s->neighList14[2 * k], j = s->neighList14[2 * k + 1];
            if (!INSET(i) && !INSET(j)) continue;
            double di[3], dj[3];
            getc3(s, i, di); getc3(s, j, dj);
            double r2 = distance2(s, di, dj);
            if (r2 <= ctoff2) {
                double pe, pv;
                switched_pair(s, r2, s->charge[i], s->charge[j], s->eps14[i], s->eps14[j],
                              s->sig14[i], s->sig14[j], ctoff2, cton2, switch2, &pe, &pv);
                *lelec += pe; *levdw += pv;
            }
        }
    } else {
        for (int k = 0; k < s->nPair14; k++) {
            int i = s->neighList14[2 * k], j = s->neighList14[2 * k + 1];
            if (!INSET(i) && !INSET(j)) continue;
            double di[3], dj[3];
            getc3(s, i, di); getc3(s, j, dj);
            double rt = 1. / sqrt(distance2(s, di, dj));
            *lelec += e_elec(s->charge[i], s->charge[j], rt);
            *levdw += e_vdw(s->eps14[i], s->eps14[j], s->sig14[i], s->sig14[j], rt);
        }
    }
    double e = 0.;
    for (int l = 0; l < s->nbond; l++) if (INSET(s->bond_a[l]) || INSET(s->bond_b[l])) e += bond_term(s, l);
    for (int l = 0; l < s->nangle; l++)
        if (INSET(s->ang_a[l]) || INSET(s->ang_b[l]) || INSET(s->ang_c[l])) e += angle_term(s, l);
    for (int l = 0; l < s->nub; l++) if (INSET(s->ub_a[l]) || INSET(s->ub_b[l])) e += ub_term(s, l);
    for (int l = 0; l < s->ndihe; l++)
        if (INSET(s->dih_a[l]) || INSET(s->dih_b[l]) || INSET(s->dih_c[l]) || INSET(s->dih_d[l]))
            e += torsion_term(s, s->dih_a[l], s->dih_b[l], s->dih_c[l], s->dih_d[l], s->dih_type[l],
                              s->dih_k[l], s->dih_phi0[l], s->dih_mult[l]);
    for (int l = 0; l < s->nimpr; l++)
        if (INSET(s->imp_a[l]) || INSET(s->imp_b[l]) || INSET(s->imp_c[l]) || INSET(s->imp_d[l]))
            e += torsion_term(s, s->imp_a[l], s->imp_b[l], s->imp_c[l], s->imp_d[l], s->imp_type[l],
                              s->imp_k[l], s->imp_phi0[l], s->imp_mult[l]);
#undef INSET
    return e;
}

double orc_moving_set_energy(orc_system *s, int first, int last)
{
    const int N = s->natom;
    double lelec = 0., levdw = 0.;
#define INSET(a) ((a) >= first && (a) <= last)
    if (s->cutmode == ORC_CUT_SWITCH) {
        const double ctoff2 = s->ctoff * s->ctoff, cton2 = s->cton * s->cton;
        const double switch2 = 1. / (X3(ctoff2 - cton2));
        /* rows of moved atoms in full; of the others, the entries that are
           moved atoms (rows hold ascending j > i) */
#define MS_ROW(i_)                                                                 \
        do {                                                                       \
            const int *row = s->neighList + (size_t)(i_) * s->cap;                 \
            const int np = s->neighPair[i_];                                       \
            int k0 = 0, k1 = np;                                                   \
            if (!INSET(i_)) {                                                      \
                if ((i_) >= last) break;                                           \
                k0 = lower_bound_row(row, np, first);                              \
                k1 = lower_bound_row(row, np, last + 1);                           \
            }                                                                      \
            switch_row(s, (i_), k0, k1, ctoff2, cton2, switch2, &lelec, &levdw);   \
        } while (0)
        const int nt = g_energy_threads;
        if (nt > 1) {
            #pragma omp parallel for schedule(dynamic, 256) num_threads(nt) reduction(+:lelec, levdw)
            for (int i = 0; i < N; i++) MS_ROW(i);
        } else {
            for (int i = 0; i < N; i++) MS_ROW(i);
        }
#undef MS_ROW
    } else {
        for (int i = 0; i < N; i++) {
            for (int j = i + 1; j < N; j++) {
                if (!INSET(i) && !INSET(j)) continue;
                if (excl_member(s, i, j)) continue;
                double di[3], dj[3];
                getc3(s, i, di); getc3(s, j, dj);
                double rt = 1. / sqrt(distance2(s, di, dj));
                lelec += e_elec(s->charge[i], s->charge[j], rt);
                levdw += e_vdw(s->eps[i], s->eps[j], s->sig[i], s->sig[j], rt);
            }
        }
    }
#undef INSET
    const double eb = moving_set_rest(s, first, last, &lelec, &levdw);
    return CHGCHARMM * KCALTOIU * lelec + 4.0 * levdw + eb;
}

/* The same energy with the pair list replaced by its definition at the
 * current coordinates: every pair (i in [first,last], any j) that
 * update_verlet_list (src/List_nonBonded.cpp:836-915) would list -- not
 * excluded, not both frozen -- and whose distance is inside the cutoff.
 * Equal to orc_moving_set_energy right after a list update; it defines the
 * dE of the batched replica path, which keeps no list.  Not in the
 * reference.  FULL mode has no list: same as orc_moving_set_energy. */
double orc_moving_set_energy_exact(orc_system *s, int first, int last)
{
    if (s->cutmode != ORC_CUT_SWITCH) return orc_moving_set_energy(s, first, last);
    const int N = s->natom;
    const double ctoff2 = s->ctoff * s->ctoff, cton2 = s->cton * s->cton;
    const double switch2 = 1. / (X3(ctoff2 - cton2));
    double lelec = 0., levdw = 0.;
    for (int i = first; i <= last; i++) {
        double di[3], dj[3];
        getc3(s, i, di);
        for (int j = 0; j < N; j++) {
            if (j == i || (j >= first && j <= last && j < i)) continue;
            const int lo = i < j ? i : j, hi = i < j ? j : i;
            if ((s->frozen[lo] && s->frozen[hi]) || excl_member(s, lo, hi)) continue;
            getc3(s, j, dj);
            double r2 = distance2(s, di, dj);
            if (r2 <= ctoff2) {
                double pe, pv;
                switched_pair(s, r2, s->charge[i], s->charge[j], s->eps[i], s->eps[j], s->sig[i], s->sig[j], ctoff2,
                              cton2, switch2, &pe, &pv);
                lelec += pe; levdw += pv;
            }
        }
    }
    const double eb = moving_set_rest(s, first, last, &lelec, &levdw);
    return CHGCHARMM * KCALTOIU * lelec + 4.0 * levdw + eb;
}

/* ------------------------------------------------------------------ */
/* R14 RNG: std::mt19937 ([rand.eng.mers], [rand.predef]) and           */
/* libstdc++ generate_canonical<double,53> (bits/random.tcc:3348-3380), */
/* uniform_real_distribution (bits/random.h:1866-1871); call sites      */
/* src/MC.cpp:119-140.                                                  */
/* ------------------------------------------------------------------ */

void orc_mt_seed(orc_mt *g, uint64_t seed)
{
    g->mt[0] = (uint32_t)(seed & 0xffffffffu);   /* seed taken mod 2^32 */
    for (int i = 1; i < 624; i++)
        g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (uint32_t)i;
    g->idx = 624;
}

uint32_t orc_mt_next(orc_mt *g)
{
    if (g->idx >= 624) {
        for (int i = 0; i < 624; i++) {
            uint32_t y = (g->mt[i] & 0x80000000u) | (g->mt[(i + 1) % 624] & 0x7fffffffu);
            uint32_t v = g->mt[(i + 397) % 624] ^ (y >> 1);
            if (y & 1u) v ^= 0x9908b0dfu;
            g->mt[i] = v;
        }
        g->idx = 0;
    }
    uint32_t y = g->mt[g->idx++];
    y ^= (y >> 11);
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= (y >> 18);
    return y;
}

double orc_mt_canonical(orc_mt *g)
{
    double sum = 0.0, tmp = 1.0;
    for (int k = 0; k < 2; k++) {
        sum += (double)orc_mt_next(g) * tmp;
        tmp *= 4294967296.0;
    }
    double ret = sum / tmp;
    if (ret >= 1.0) ret = nextafter(1.0, 0.0);
    return ret;
}

static inline double u_alpha(orc_mt *g) { return orc_mt_canonical(g) * (1.0 - 0.0) + 0.0; }
static inline double u_move(orc_mt *g) { return orc_mt_canonical(g) * (1.0 - (-1.0)) + (-1.0); }

/* ------------------------------------------------------------------ */
/* R15 moves: include/Move.hpp:42-183; R16 backup src/AtomList.cpp     */
/* ------------------------------------------------------------------ */

static void wrap_atom(const orc_system *s, int i)
{
    /* PerConditions::applyPBC(double&,double&,double&) :268-285 */
    if (s->pbtype == ORC_PBC_CUBIC || s->pbtype == ORC_PBC_ORBIC) {
        s->x[i] -= s->pbx * orc_rint(s->rpbx * s->x[i]);
        s->y[i] -= s->pby * orc_rint(s->rpby * s->y[i]);
        s->z[i] -= s->pbz * orc_rint(s->rpbz * s->z[i]);
    }
}

static void translate_range(orc_system *s, int a, int b, double rx, double ry, double rz)
{
    for (int i = a; i <= b; i++) {   /* tr_a_b :42-50 */
        s->x[i] += rx; s->y[i] += ry; s->z[i] += rz;
        wrap_atom(s, i);
    }
}

/* find_rot_mat :77-118 (vec is normalised in place, as the reference does) */
static void find_rot_mat(double mat[3][3], double vec[3], double angle)
{
    double angleRad = angle * PI_ / 180.0;
    double rn = 1.0 / sqrt(vec[0] * vec[0] + vec[1] * vec[1] + vec[2] * vec[2]);
    vec[0] *= rn; vec[1] *= rn; vec[2] *= rn;
    double T[3][3], T2[3][3] = {{0., 0., 0.}, {0., 0., 0.}, {0., 0., 0.}};
    T[0][0] = T[1][1] = T[2][2] = 0.0;
    T[2][1] = vec[0]; T[1][2] = -vec[0];
    T[0][2] = vec[1]; T[2][0] = -vec[1];
    T[1][0] = vec[2]; T[0][1] = -vec[2];
    double v1 = sin(angleRad), v2 = 1.0 - cos(angleRad);
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) mat[i][j] += T[i][j] * v1;
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) for (int k = 0; k < 3; k++) T2[i][j] += T[i][k] * T[k][j];
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) mat[i][j] += T2[i][j] * v2;
}

/* apply_rotation :120-143 */
static void rotate_range(orc_system *s, const double piv[3], double R[3][3], int a, int b)
{
    for (int i = a; i <= b; i++) {
        double x = s->x[i] - piv[0], y = s->y[i] - piv[1], z = s->z[i] - piv[2];
        double nx = piv[0] + R[0][0] * x + R[1][0] * y + R[2][0] * z;
        double ny = piv[1] + R[0][1] * x + R[1][1] * y + R[2][1] * z;
        double nz = piv[2] + R[0][2] * x + R[1][2] * y + R[2][2] * z;
        s->x[i] = nx; s->y[i] = ny; s->z[i] = nz;
        wrap_atom(s, i);
    }
}

/* Tools::getCentreOfMass(at_List, first, last), include/Tools.hpp:95-113 */
static void centre_of_mass(const orc_system *s, int a, int b, double c[3])
{
    double mt = 0.0; c[0] = c[1] = c[2] = 0.0;
    for (int i = a; i <= b; i++) {
        double m = s->mass[i];
        mt += m; c[0] += m * s->x[i]; c[1] += m * s->y[i]; c[2] += m * s->z[i];
    }
    c[0] /= mt; c[1] /= mt; c[2] /= mt;
}

/* move instance tables, src/List_Moves.cpp:257-457 and gtrsfl :630-697.
 * An instance is one atom range [first,last] plus a pivot (-1 = COM). */
typedef struct orc_inst { int first, last, pivot; } orc_inst;

/* gtrsfl :630-697.  resLast[k] = last atom of residue k (resLast[0] = -1);
 * the bounds test is inclusive on both sides, so for the last atom of a
 * residue the search can return the following residue. */
static int *residue_last_table(const orc_system *s, int *nres)
{
    const int N = s->natom;
    int iresl = 0;
    for (int i = 0; i < N; i++) if (s->resid[i] > iresl) iresl = s->resid[i];
    int *resLast = (int *)malloc(sizeof(int) * (iresl + 1));
    resLast[0] = -1;
    int last = 1, idx = 0;
    for (int it = 0; it < N; it++) if (s->resid[it] > last) { last++; resLast[++idx] = it - 1; }
    resLast[iresl] = N - 1;
    *nres = iresl;
    return resLast;
}

static int residue_range(const int *resLast, int nres, int atomidx, int *a1, int *a2)
{
    int iresf = 1, iresl = nres;
    while (iresf <= iresl) {
        int mid = (iresl + iresf) / 2;
        if (resLast[mid - 1] <= atomidx && resLast[mid] >= atomidx) {
            *a1 = resLast[mid - 1] + 1; *a2 = resLast[mid]; return 1;
        } else if (resLast[mid] <= atomidx) iresf = mid + 1;
        else iresl = mid - 1;
    }
    return 0;
}

static orc_inst *build_instances(const orc_system *s, const orc_movetype *t, int *n)
{
    const int N = s->natom;
    orc_inst *v = (orc_inst *)malloc(sizeof(orc_inst) * (N + 1));
    int c = 0, nres = 0;
    int *resLast = residue_last_table(s, &nres);
    if (t->mode == ORC_MODE_ALL) {
        /* makeMoveList gives the selected ranges; only a contiguous "all"
         * selection is supported here */
        v[0].first = 0; v[0].last = N - 1; v[0].pivot = -1; c = 1;
    } else {
        for (int i = 0; i < N; i++) {
            if (t->sele && !t->sele[i]) continue;
            if (t->mode == ORC_MODE_RESIDUE) {
                int a1 = i, a2 = i;
                residue_range(resLast, nres, i, &a1, &a2);
                v[c].first = a1; v[c].last = a2;
            } else {
                v[c].first = i; v[c].last = i;
            }
            v[c].pivot = (t->type == ORC_MV_ROT) ? i : -1;
            c++;
        }
    }
    free(resLast);
    *n = c;
    return v;
}

/* MC::adjust_dmax, src/MC.cpp:165-173 */
static void adjust_dmax(double *dmax, double target, int each, int acc)
{
    double ratio = (double)acc / (double)each;
    if (ratio > target / 100) *dmax *= 1.10;
    else *dmax *= 0.9;
}

/* MC_metropolis::run :45-173 and apply_criterion :175-204 */
double orc_mc_run(orc_system *s, const orc_movetype *types, int ntypes,
                  int nsteps, uint64_t seed, int mode, int list_every,
                  orc_mc_record *rec, double *dmax_out, int nthreads)
{
    const int N = s->natom;
    orc_mt g;
    orc_mt_seed(&g, seed);
    orc_inst **inst = (orc_inst **)malloc(sizeof(orc_inst *) * ntypes);
    int *ninst = (int *)malloc(sizeof(int) * ntypes);
    double *dmax = (double *)malloc(sizeof(double) * ntypes);
    int *accTmp = (int *)calloc(ntypes, sizeof(int));
    for (int t = 0; t < ntypes; t++) { inst[t] = build_instances(s, &types[t], &ninst[t]); dmax[t] = types[t].dmax; }
    double *bx = (double *)malloc(sizeof(double) * N), *by = (double *)malloc(sizeof(double) * N),
           *bz = (double *)malloc(sizeof(double) * N);
    const double beta = 1.0 / ((RBOLTZUI / KCALTOIU) * s->temp);

    double etot = orc_getE(s, NULL);
    for (int st = 1; st <= nsteps; st++) {
        int imvtyp = (int)((double)ntypes * u_alpha(&g));
        int imvatm = (int)((double)ninst[imvtyp] * u_alpha(&g));
        const orc_inst in = inst[imvtyp][imvatm];
        double eold = mode == 0 ? orc_getE(s, NULL)
                      : mode == 2 ? orc_moving_set_energy_exact(s, in.first, in.last)
                                  : orc_moving_set_energy(s, in.first, in.last);
        for (int i = in.first; i <= in.last; i++) { bx[i] = s->x[i]; by[i] = s->y[i]; bz[i] = s->z[i]; }
        double r[3];
        do { r[0] = u_move(&g); r[1] = u_move(&g); r[2] = u_move(&g); }   /* rndSphere :142-156 */
        while ((r[0] * r[0] + r[1] * r[1] + r[2] * r[2]) > 1.0);
        if (types[imvtyp].type == ORC_MV_TRN) {
            r[0] *= dmax[imvtyp]; r[1] *= dmax[imvtyp]; r[2] *= dmax[imvtyp];
            translate_range(s, in.first, in.last, r[0], r[1], r[2]);
        } else {
            double rang = dmax[imvtyp] * u_move(&g);
            double R[3][3] = {{1., 0., 0.}, {0., 1., 0.}, {0., 0., 1.}};
            find_rot_mat(R, r, rang);
            double piv[3];
            if (in.pivot >= 0) { piv[0] = s->x[in.pivot]; piv[1] = s->y[in.pivot]; piv[2] = s->z[in.pivot]; }
            else centre_of_mass(s, in.first, in.last, piv);
            rotate_range(s, piv, R, in.first, in.last);
        }
        double enew = mode == 0 ? orc_getE(s, NULL)
                      : mode == 2 ? orc_moving_set_energy_exact(s, in.first, in.last)
                                  : orc_moving_set_energy(s, in.first, in.last);
        double de = enew - eold;
        int acc;
        if (de <= DBL_EPSILON) acc = 1;
        else {
            double alpha = u_alpha(&g);
            double accf = exp(-beta * de);
            acc = alpha < accf;
        }
        if (acc) { etot += de; accTmp[imvtyp]++; }
        else for (int i = in.first; i <= in.last; i++) { s->x[i] = bx[i]; s->y[i] = by[i]; s->z[i] = bz[i]; }
        if (rec) { rec->mvtyp[st - 1] = imvtyp; rec->mvinst[st - 1] = imvatm; rec->de[st - 1] = de; rec->acc[st - 1] = (unsigned char)acc; }
        for (int t = 0; t < ntypes; t++) {
            if (types[t].each != 0 && (st % types[t].each) == 0) {
                adjust_dmax(&dmax[t], types[t].target, types[t].each, accTmp[t]);
                accTmp[t] = 0;
            }
        }
        if (list_every > 0 && st % list_every == 0 && s->cutmode != ORC_CUT_FULL)
            orc_update_verlet_list(s, nthreads);
    }
    if (dmax_out) for (int t = 0; t < ntypes; t++) dmax_out[t] = dmax[t];
    for (int t = 0; t < ntypes; t++) free(inst[t]);
    free(inst); free(ninst); free(dmax); free(accTmp); free(bx); free(by); free(bz);
    return etot;
}

/* ------------------------------------------------------------------ */
/* Gradient of the switched nonbonded energy (list-based), FP64.  Not   */
/* in the reference; the derivative of computeNonBonded_switch :336-424 */
/* term by term, written out here as the checker for the adjoint path.  */
/* E = C*sum S qi qj/r + 4*sum S ei ej (u^12 - u^6), u = (si+sj)/r,      */
/* C = 332.0716*418.4.                                                   */
/* ------------------------------------------------------------------ */
void orc_nonbonded_switch_grad(orc_system *s, double *gx, double *gy, double *gz,
                               double *geps, double *gsig, double *gq)
{
    const int N = s->natom;
    const double C = CHGCHARMM * KCALTOIU;
    const double ctoff2 = s->ctoff * s->ctoff, cton2 = s->cton * s->cton;
    const double switch2 = 1. / (X3(ctoff2 - cton2));
    for (int i = 0; i < N; i++) { gx[i] = gy[i] = gz[i] = geps[i] = gsig[i] = gq[i] = 0.0; }
    for (int i = 0; i < N; i++) {
        double di[3], dj[3], d[3];
        getc3(s, i, di);
        const int *row = s->neighList + (size_t)i * s->cap;
        for (int k = 0; k < s->neighPair[i]; k++) {
            int j = row[k];
            getc3(s, j, dj);
            double r2 = distance2v(s, di, dj, d);      /* d = xj - xi (min image) */
            if (r2 > ctoff2) continue;
            double r = sqrt(r2), rinv = 1.0 / r;
            double sij = s->sig[i] + s->sig[j];
            double u = sij * rinv, u6 = X6(u), u12 = u6 * u6;   /* u = 0 when sij = 0 */
            double eij = s->eps[i] * s->eps[j];
            double qq = s->charge[i] * s->charge[j];
            double Ee = C * qq * rinv;
            double Ev = 4.0 * eij * (u12 - u6);
            double S = 1.0, dSdr2 = 0.0;
            if (r2 > cton2) {
                double a = ctoff2 - r2, b = ctoff2 + 2. * r2 - 3. * cton2;
                S = a * a * b * switch2;
                dSdr2 = (-2.0 * a * b + 2.0 * a * a) * switch2;
            }
            /* dE/dr2 of S*(Ee+Ev) */
            double dEe_dr2 = -0.5 * Ee / r2;
            double dEv_dr2 = 4.0 * eij * (-6.0 * u12 + 3.0 * u6) / r2;
            double g = S * (dEe_dr2 + dEv_dr2) + dSdr2 * (Ee + Ev);
            /* r2 = |xj - xi|^2 -> dr2/dxj = 2d, dr2/dxi = -2d */
            gx[j] += 2.0 * g * d[0]; gy[j] += 2.0 * g * d[1]; gz[j] += 2.0 * g * d[2];
            gx[i] -= 2.0 * g * d[0]; gy[i] -= 2.0 * g * d[1]; gz[i] -= 2.0 * g * d[2];
            /* parameters */
            double dv_deij = S * 4.0 * (u12 - u6);
            geps[i] += dv_deij * s->eps[j]; geps[j] += dv_deij * s->eps[i];
            if (sij != 0.0) {
                double dv_dsij = S * 4.0 * eij * (12.0 * u12 - 6.0 * u6) / sij;
                gsig[i] += dv_dsij; gsig[j] += dv_dsij;
            }
            double de_dqq = S * C * rinv;
            gq[i] += de_dqq * s->charge[j]; gq[j] += de_dqq * s->charge[i];
        }
    }
}

/* ------------------------------------------------------------------ */
/* Gradient of the whole nonbonded energy elec + vdw, FP64.  Not in the  */
/* reference; the derivative of computeNonBonded_switch (:336-424) with  */
/* computeNonBonded14_switch (:426-488) in SWITCH mode, and of           */
/* computeNonBonded_full (:194-265) with computeNonBonded14 (:267-318)   */
/* in FULL mode.  g8 = 8 n values: dE/dx, dE/dy, dE/dz, dE/deps, dE/dsig, */
/* dE/dq, dE/deps14, dE/dsig14 per atom (parameters in the AtomList      */
/* convention).                                                          */
/* ------------------------------------------------------------------ */

/* one pair's C q_i q_j / r + 4 e_i e_j (u^12 - u^6), u = (s_i + s_j) / r,
   times the switch when `sw`; its gradient added to both atoms, the LJ
   parameter parts to components ce (eps) and cs (sig) */
static void pair_grad(const orc_system *s, int i, int j, const double d[3], double r2, double ei, double ej,
                      double si, double sj, int sw, double ctoff2, double cton2, double switch2,
                      double *g8, int ce, int cs)
{
    const int N = s->natom;
    const double C = CHGCHARMM * KCALTOIU;
    double r = sqrt(r2), rinv = 1.0 / r;
    double sij = si + sj;
    double u = sij * rinv, u6 = X6(u), u12 = u6 * u6;
    double eij = ei * ej;
    double qq = s->charge[i] * s->charge[j];
    double Ee = C * qq * rinv;
    double Ev = 4.0 * eij * (u12 - u6);
    double S = 1.0, dSdr2 = 0.0;
    if (sw && r2 > cton2) {
        double a = ctoff2 - r2, b = ctoff2 + 2. * r2 - 3. * cton2;
        S = a * a * b * switch2;
        dSdr2 = (-2.0 * a * b + 2.0 * a * a) * switch2;
    }
    double g = S * (-0.5 * Ee / r2 + 4.0 * eij * (-6.0 * u12 + 3.0 * u6) / r2) + dSdr2 * (Ee + Ev);
    for (int c = 0; c < 3; c++) {          /* d = x_j - x_i: dr2/dx_j = 2d, dr2/dx_i = -2d */
        g8[(size_t)c * N + j] += 2.0 * g * d[c];
        g8[(size_t)c * N + i] -= 2.0 * g * d[c];
    }
    double dv_deij = S * 4.0 * (u12 - u6);
    g8[(size_t)ce * N + i] += dv_deij * ej;
    g8[(size_t)ce * N + j] += dv_deij * ei;
    if (sij != 0.0) {
        double dv_dsij = S * 4.0 * eij * (12.0 * u12 - 6.0 * u6) / sij;
        g8[(size_t)cs * N + i] += dv_dsij;
        g8[(size_t)cs * N + j] += dv_dsij;
    }
    double de_dqq = S * C * rinv;
    g8[5 * (size_t)N + i] += de_dqq * s->charge[j];
    g8[5 * (size_t)N + j] += de_dqq * s->charge[i];
}

void orc_nonbonded_grad(orc_system *s, double *g8)
{
    const int N = s->natom;
    const double ctoff2 = s->ctoff * s->ctoff, cton2 = s->cton * s->cton;
    const double switch2 = 1. / (X3(ctoff2 - cton2));
    const int sw = s->cutmode == ORC_CUT_SWITCH;
    for (size_t k = 0; k < 8 * (size_t)N; k++) g8[k] = 0.0;
    double di[3], dj[3], d[3];
    if (sw) {
        for (int i = 0; i < N; i++) {
            getc3(s, i, di);
            const int *row = s->neighList + (size_t)i * s->cap;
            for (int k = 0; k < s->neighPair[i]; k++) {
                int j = row[k];
                getc3(s, j, dj);
                double r2 = distance2v(s, di, dj, d);
                if (r2 > ctoff2) continue;
                pair_grad(s, i, j, d, r2, s->eps[i], s->eps[j], s->sig[i], s->sig[j], 1, ctoff2, cton2, switch2,
                          g8, 3, 4);
            }
        }
    } else {
        for (int i = 0; i < N - 1; i++) {
            getc3(s, i, di);
            for (int j = i + 1; j < N; j++) {
                if (excl_member(s, i, j)) continue;
                getc3(s, j, dj);
                double r2 = distance2v(s, di, dj, d);
                pair_grad(s, i, j, d, r2, s->eps[i], s->eps[j], s->sig[i], s->sig[j], 0, ctoff2, cton2, switch2,
                          g8, 3, 4);
            }
        }
    }
    for (int k = 0; k < s->nPair14; k++) {
        int i = s->neighList14[2 * k], j = s->neighList14[2 * k + 1];
        getc3(s, i, di); getc3(s, j, dj);
        double r2 = distance2v(s, di, dj, d);
        if (sw && r2 > ctoff2) continue;
        pair_grad(s, i, j, d, r2, s->eps14[i], s->eps14[j], s->sig14[i], s->sig14[j], sw, ctoff2, cton2, switch2,
                  g8, 6, 7);
    }
}
