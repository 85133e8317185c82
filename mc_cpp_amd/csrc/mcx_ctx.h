// mcx_ctx.h -- the handle behind the C ABI (one simulated system on one
// GPU) and the host helpers shared by the single-system and batched entry
// points.
#pragma once

#include "../../include/mcx.h"
#include "mcx_internal.h"

#include <cstring>
#include <string>
#include <vector>

template <typename T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    void alloc(size_t count)
    {
        if (count <= n && p) return;
        if (p) HIPCHK(hipFree(p));
        p = nullptr;
        n = 0;
        if (count == 0) count = 1;
        HIPCHK(hipMalloc((void **)&p, count * sizeof(T)));
        n = count;
    }
    void upload(const T *h, size_t count, hipStream_t st)
    {
        alloc(count);
        if (count) HIPCHK(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, st));
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

// Device buffers of a move set (instance table and term records).
struct MoveBufs {
    DBuf<int4> rid;
    DBuf<int> rkk, rty;
    DBuf<double4> rpar;
    DBuf<int2> inst_adj, inst_pairs, pair_ent;
    DBuf<int4> inst;
    void release()
    {
        rid.release(); rkk.release(); rty.release(); rpar.release(); inst_adj.release(); inst.release();
        inst_pairs.release(); pair_ent.release();
    }
};

struct mcx_ctx {
    int n = 0, device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // parameters
    int pbtype = MCX_PBC_CUBIC;
    double box[3] = {0, 0, 0};
    int cutmode = MCX_CUT_SWITCH;
    double ctoff = 12.0, cton = 10.0, dcut = 2.0, temp = 300.0;
    bool have_atoms = false, have_coords = false, have_excl = false, list_valid = false;
    std::vector<double> q, eps, sig, eps14, sig14, mass;
    std::vector<int> frozen, resid;
    HostTerms terms;    // kinds 0..4 from the caller, 5/6 from the exclusion builder
    std::vector<int> excl_pair, excl_off, excl_idx;   // visible rows
    std::vector<int> pairs14;
    // device buffers
    DBuf<double> x, y, z, q64, eps64, sig64, eps1464, sig1464, mass64;
    DBuf<int4> ipos;
    DBuf<float2> es, esk, es14;
    DBuf<TileAtom> tq;
    DBuf<double4> qes;
    std::vector<double> hx, hy, hz;   // host copy of the last uploaded coordinates
    // pinned host staging: mcx_set_coords_moved packs x, y, z into stage_h
    // for one upload (stage_ev marks its end), mcx_energy's nine sums land
    // in res_h
    double *stage_h = nullptr, *res_h = nullptr;
    size_t stage_cap = 0;
    hipEvent_t stage_ev = nullptr;
    DBuf<double> xyz_in;
    bool host_stale = false;          // the device coordinates were changed in place since
    DBuf<int> frozen_d, excl_off_d, excl_idx_d, xfull_off_d, xfull_idx_d, p14_off_d, p14_idx_d;
    DBuf<double4> cpos;
    DBuf<int> nbr, cnt, half, overflow;
    int cap = 0;
    DBuf<int> cell_of, cell_start, cell_count, cell_atoms, cursor, group_start, cell_tmp, cell_acc;
    DBuf<int4> group_info;
    DBuf<unsigned> bin_ticket;
    DBuf<int4> cfix, ipos_ref;
    CellGrid grid{};
    DBuf<int2> tiles;
    int ntiles = 0;
    DBuf<double> partial, result;
    DBuf<int4> t_idx[kNumKinds];
    DBuf<int> t_type[kNumKinds];
    DBuf<double4> t_par[kNumKinds];
    DBuf<int> adj_off, adj_kk, adj_type;
    DBuf<int4> adj_ids;
    DBuf<double4> adj_par;
    std::vector<int> adj_off_h, adj_kk_h;
    std::vector<int> xfull_off_h, xfull_idx_h;   // excluded partners, both directions (CSR)
    MoveBufs mvb;
    uint32_t rng_stream = 0;   // Philox stream of mcx_mc_run
    DBuf<double> newxyz, scalar;
    long long last_npairs = 0;
    // incremental list updates (list_patch.hip); valid while every move since
    // the last full build was recorded
    DBuf<int> p_moved_flag, p_moved_list, p_moved_cnt, p_stale_flag, p_stale_list, p_stale_cnt;
    DBuf<int> p_edit_cnt, p_edit_val, p_touched, p_touched_cnt, p_newrow, p_newcnt, p_newhalf;
    DBuf<unsigned> p_ticket;
    bool patch_alloc = false, patch_valid = false;

    ListPatch patch() const
    {
        ListPatch p{};
        p.moved_flag = p_moved_flag.p; p.moved_list = p_moved_list.p; p.moved_cnt = p_moved_cnt.p;
        p.stale_flag = p_stale_flag.p; p.stale_list = p_stale_list.p; p.stale_cnt = p_stale_cnt.p;
        p.edit_cnt = p_edit_cnt.p; p.edit_val = p_edit_val.p;
        p.touched = p_touched.p; p.touched_cnt = p_touched_cnt.p;
        p.newrow = p_newrow.p; p.newcnt = p_newcnt.p; p.newhalf = p_newhalf.p;
        p.ticket = p_ticket.p;
        return p;
    }

    DevSys sys() const
    {
        DevSys s{};
        s.n = n;
        s.pbtype = pbtype;
        for (int k = 0; k < 3; k++) {
            s.L[k] = box[k];
            s.rL[k] = pbtype == MCX_PBC_NONE ? 0.0 : 1.0 / box[k];
            s.Lf[k] = (float)s.L[k];
            s.rLf[k] = (float)s.rL[k];
        }
        s.ctoff2 = ctoff * ctoff;
        s.cton2 = cton * cton;
        const double dd = s.ctoff2 - s.cton2;
        s.switch2 = 1. / (dd * dd * dd);
        double cn = ctoff + dcut;
        s.cutnb2 = cn * cn;
        s.cutmode = cutmode;
        s.x = x.p; s.y = y.p; s.z = z.p;
        s.ipos = ipos.p;
        s.es = es.p; s.esk = esk.p; s.tq = tq.p; s.es14 = es14.p;
        for (int k = 0; k < 3; k++) s.unit[k] = pbtype == MCX_PBC_NONE ? 1.0 : box[k] / 4294967296.0;
        s.iunit = pbtype == MCX_PBC_CUBIC ? 4294967296.0 / box[0] : 1.0;
        s.qes = qes.p;
        s.q64 = q64.p; s.eps64 = eps64.p; s.sig64 = sig64.p; s.eps1464 = eps1464.p; s.sig1464 = sig1464.p;
        s.frozen = frozen_d.p;
        s.excl_off = excl_off_d.p; s.excl_idx = excl_idx_d.p;
        s.xfull_off = xfull_off_d.p; s.xfull_idx = xfull_idx_d.p;
        s.p14_off = p14_off_d.p; s.p14_idx = p14_idx_d.p;
        s.cap = cap;
        s.nbr = nbr.p; s.cnt = cnt.p; s.half = half.p; s.overflow = overflow.p;
        return s;
    }

    DevTerms dterms() const
    {
        DevTerms t{};
        t.start[0] = 0;
        for (int k = 0; k < kNumKinds; k++) {
            t.nterm[k] = (int)terms.idx[k].size();
            t.start[k + 1] = t.start[k] + t.nterm[k];
            t.idx[k] = t_idx[k].p;
            t.type[k] = t_type[k].p;
            t.par[k] = t_par[k].p;
        }
        // the excluded-pair correction applies to FULL mode only
        if (cutmode != MCX_CUT_FULL) {
            t.nterm[kExclPair] = 0;
            t.start[kExclPair + 1] = t.start[kExclPair];
        }
        return t;
    }

    // timing events of mcx_mc_run's kernel launches (MCX_KERNEL_EVENTS=1),
    // kept for the next run
    std::vector<hipEvent_t> ev_pool;

    ~mcx_ctx()
    {
        DBuf<double> *d[] = {&x, &y, &z, &q64, &eps64, &sig64, &eps1464, &sig1464, &mass64, &partial, &result,
                             &newxyz, &scalar, &xyz_in};
        for (auto *b : d) b->release();
        if (stage_ev) (void)hipEventSynchronize(stage_ev);
        if (stage_h) (void)hipHostFree(stage_h);
        if (res_h) (void)hipHostFree(res_h);
        if (stage_ev) (void)hipEventDestroy(stage_ev);
        ipos.release(); es.release(); esk.release(); tq.release(); es14.release(); qes.release();
        cpos.release();
        cfix.release();
        ipos_ref.release();
        adj_ids.release();
        adj_par.release();
        adj_type.release();
        mvb.release();
        DBuf<int> *ib[] = {&frozen_d, &excl_off_d, &excl_idx_d, &xfull_off_d, &xfull_idx_d, &p14_off_d, &p14_idx_d, &nbr, &cnt, &half, &overflow, &cell_of,
                           &cell_start, &cell_count, &cell_atoms, &cursor, &group_start, &cell_tmp, &adj_off, &adj_kk,
                           &cell_acc};
        for (auto *b : ib) b->release();
        tiles.release();
        group_info.release();
        bin_ticket.release();
        DBuf<int> *pb[] = {&p_moved_flag, &p_moved_list, &p_moved_cnt, &p_stale_flag, &p_stale_list, &p_stale_cnt,
                           &p_edit_cnt, &p_edit_val, &p_touched, &p_touched_cnt, &p_newrow, &p_newcnt, &p_newhalf};
        for (auto *b : pb) b->release();
        p_ticket.release();
        for (int k = 0; k < kNumKinds; k++) { t_idx[k].release(); t_type[k].release(); t_par[k].release(); }
        for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};


// noalloc.hip: looks up the device's copy of the static scratch array of
// mcx_energy_noalloc once, with that device current (called by mcx_create).
int noalloc_init_device(int device);

// host helpers (mcx_api.cpp)
namespace mcxi {
long long build_list(mcx_ctx *h);
void energy_device(mcx_ctx *h, double out9[9]);
void assemble(const double r[9], double ener[10]);
void sync_host_coords(mcx_ctx *h);
DevMoves build_moves(mcx_ctx *h, const mcx_move_type *types, int ntypes, MoveBufs &buf, int max_atoms);
void mt_seed(uint32_t *st, uint64_t seed);
}  // namespace mcxi
