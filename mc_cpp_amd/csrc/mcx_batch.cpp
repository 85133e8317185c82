// mcx_batch.cpp -- host side of the batched replicas: the mcx_batch_* entry
// points of include/mcx.h, the per-replica descriptors and run-state pools,
// hipGraph capture of sweeps + exchange rounds, and the launch-shape tuner.
#include "mcx_ctx.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>
#include <memory>
#include <tuple>
#include <type_traits>

using namespace mcxi;

struct mcx_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int nrep = 0;
    std::vector<mcx_ctx *> reps;
    std::vector<std::vector<mcx_move_type>> types;   // per replica
    std::vector<MoveBufs> mbufs;
    std::vector<DevMoves> moves;
    std::vector<int> sweep_len, ntab, frozen_any;
    std::vector<double> temps;
    std::vector<BatchRep> hrep;
    DBuf<BatchRep> drep;
    DBuf<int> order, slot_of_rep, rep_of_slot, atype, acc_tmp;
    DBuf<double4> ptab;
    DBuf<double> dtemps, etot, dmax;
    DBuf<uint32_t> rng;
    DBuf<long long> trials, accepted, step, xatt, xacc;
    DBuf<unsigned long long> xcount, npair;
    std::vector<size_t> atype_off;   // offset of replica r's atom types in the pool
    int max_n = 0;
    size_t lds_limit = 0;
    int threads = 1024, unroll = 2;
    uint64_t seed = 0;
    long long sweeps = 0;
    // captured sweep(s) + exchange, keyed by (exchange_every, threads, unroll)
    hipGraphExec_t gexec = nullptr;
    std::tuple<int, int, int> gkey{-1, -1, -1};

    bool img_fits(int nt, int u) const
    {
        return batch_static_lds(nt, u, true) + batch_dynamic_lds(nt, true, max_n) <= lds_limit && max_n < 65536;
    }
    BatchArgs args(int nmoves, int rec_len, unsigned char *acc_out, double *de_out, bool img) const
    {
        BatchArgs a{};
        a.reps = drep.p;
        a.order = order.p;
        a.temps = dtemps.p;
        a.slot_of_rep = slot_of_rep.p;
        a.img_cap = img ? max_n : 0;
        a.nmoves = nmoves;
        a.rec_len = rec_len;
        a.acc_out = acc_out;
        a.de_out = de_out;
        a.cs = 1;
        for (const mcx_ctx *h : reps)
            if (h->pbtype != MCX_PBC_CUBIC || h->cutmode != MCX_CUT_SWITCH) a.cs = 0;
        return a;
    }
    ExchArgs xargs() const
    {
        ExchArgs x{};
        x.nrep = nrep;
        x.key = seed;
        x.temps = dtemps.p;
        x.etot = etot.p;
        x.slot_of_rep = slot_of_rep.p;
        x.rep_of_slot = rep_of_slot.p;
        x.counter = xcount.p;
        x.att = xatt.p;
        x.acc = xacc.p;
        return x;
    }
    void launch_sweep(int nmoves, int rec_len, unsigned char *acc_out, double *de_out)
    {
        const bool img = img_fits(threads, unroll);
        launch_batch_sweep(args(nmoves, rec_len, acc_out, de_out, img), nrep, threads, unroll, img, stream);
        HIPCHK(hipGetLastError());
    }
    void drop_graph()
    {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        gexec = nullptr;
        gkey = std::make_tuple(-1, -1, -1);
    }
    void prepare()
    {
        const bool img = img_fits(threads, unroll);
        batch_prepare(threads, unroll, img, img ? max_n : 0);
    }
    ~mcx_batch()
    {
        drop_graph();
        for (auto &m : mbufs) m.release();
        drep.release(); order.release(); slot_of_rep.release(); rep_of_slot.release(); atype.release();
        acc_tmp.release(); ptab.release(); dtemps.release(); etot.release(); dmax.release(); rng.release();
        trials.release(); accepted.release(); step.release(); xatt.release(); xacc.release(); xcount.release();
        npair.release();
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

template <typename F>
int bguard(mcx_batch *b, F &&f)
{
    if (!b) return MCX_ERR_ARG;
    try {
        HIPCHK(hipSetDevice(b->device));
        f();
        return MCX_OK;
    } catch (const mcx_error &e) {
        b->err = e.msg;
        return e.code;
    } catch (const std::exception &e) {
        b->err = e.what();
        return MCX_ERR_ARG;
    }
}

// Distinct (charge, eps, sig) triples of a replica and each atom's index
// into them, with the frozen flag in bit 16.
void type_table(const mcx_ctx *h, std::vector<double4> &tab, std::vector<int> &tw)
{
    std::map<std::tuple<double, double, double>, int> seen;
    tw.resize(h->n);
    for (int i = 0; i < h->n; i++) {
        const auto key = std::make_tuple(h->q[i], h->eps[i], h->sig[i]);
        auto it = seen.find(key);
        int t;
        if (it == seen.end()) {
            t = (int)tab.size();
            if (t >= kBatchTypes)
                throw mcx_error(MCX_ERR_UNSUPPORTED, "batched replicas support at most 64 distinct atom types");
            seen[key] = t;
            tab.push_back(make_double4(std::get<0>(key), std::get<1>(key), std::get<2>(key), 0.0));
        } else {
            t = it->second;
        }
        tw[i] = t | (h->frozen[i] ? (1 << 16) : 0);
    }
}

int residue_count(const mcx_ctx *h)
{
    int c = 0, last = -1;
    for (int i = 0; i < h->n; i++)
        if (h->resid[i] != last) { c++; last = h->resid[i]; }
    return c;
}

// Descriptors from the handles' current set-up, and the replica order
// (largest work first, so the longest chains start first).
void build_descriptors(mcx_batch *b)
{
    std::vector<double4> tab;
    std::vector<int> tw_all;
    b->hrep.assign(b->nrep, BatchRep{});
    b->ntab.assign(b->nrep, 0);
    b->frozen_any.assign(b->nrep, 0);
    b->atype_off.assign(b->nrep, 0);
    std::vector<size_t> tab_off(b->nrep);
    for (int r = 0; r < b->nrep; r++) {
        mcx_ctx *h = b->reps[r];
        std::vector<double4> t;
        std::vector<int> tw;
        type_table(h, t, tw);
        tab_off[r] = tab.size();
        b->atype_off[r] = tw_all.size();
        b->ntab[r] = (int)t.size();
        for (int v : tw) b->frozen_any[r] |= (v >> 16) & 1;
        tab.insert(tab.end(), t.begin(), t.end());
        tw_all.insert(tw_all.end(), tw.begin(), tw.end());
    }
    b->atype.upload(tw_all.data(), tw_all.size(), b->stream);
    b->ptab.upload(tab.data(), tab.size(), b->stream);
    for (int r = 0; r < b->nrep; r++) {
        BatchRep &R = b->hrep[r];
        R.s = b->reps[r]->sys();
        R.mv = b->moves[r];
        R.atype = b->atype.p + b->atype_off[r];
        R.ptab = b->ptab.p + tab_off[r];
        R.ntab = b->ntab[r];
        R.sweep_len = b->sweep_len[r];
        R.frozen_any = b->frozen_any[r];
        R.rng = b->rng.p + (size_t)r * kRngWords;
        R.etot = b->etot.p + r;
        R.dmax = b->dmax.p + (size_t)r * kMaxMoveTypes;
        R.trials = b->trials.p + (size_t)r * kMaxMoveTypes;
        R.accepted = b->accepted.p + (size_t)r * kMaxMoveTypes;
        R.acc_tmp = b->acc_tmp.p + (size_t)r * kMaxMoveTypes;
        R.step = b->step.p + r;
        R.npair = b->npair.p + r;
    }
    std::vector<int> ord(b->nrep);
    for (int r = 0; r < b->nrep; r++) ord[r] = r;
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) {
        return (long long)b->reps[x]->n * b->sweep_len[x] > (long long)b->reps[y]->n * b->sweep_len[y];
    });
    b->order.upload(ord.data(), ord.size(), b->stream);
}

void upload_descriptors(mcx_batch *b, uint64_t seed, int rng_mode)
{
    for (int r = 0; r < b->nrep; r++) {
        b->hrep[r].rng_mode = rng_mode == MCX_MC_REPLAY ? 1 : 0;
        b->hrep[r].key = seed;
        b->hrep[r].stream = (uint32_t)r;
    }
    b->drep.upload(b->hrep.data(), b->hrep.size(), b->stream);
}

void reset(mcx_batch *b, uint64_t seed, int rng_mode, const uint64_t *seeds)
{
    if (rng_mode != MCX_MC_REPLAY && rng_mode != MCX_MC_PHILOX) throw mcx_error(MCX_ERR_ARG, "unknown rng_mode");
    const int R = b->nrep;
    b->drop_graph();
    b->seed = seed;
    build_descriptors(b);
    upload_descriptors(b, seed, rng_mode);
    // starting energies from each handle's current coordinates
    std::vector<double> e0(R);
    for (int r = 0; r < R; r++) {
        mcx_ctx *h = b->reps[r];
        HIPCHK(hipSetDevice(h->device));
        double r9[9], ener[10];
        energy_device(h, r9);
        assemble(r9, ener);
        e0[r] = ener[0];
    }
    HIPCHK(hipSetDevice(b->device));
    std::vector<uint32_t> st((size_t)R * kRngWords, 0);
    for (int r = 0; r < R; r++) {
        uint32_t *w = st.data() + (size_t)r * kRngWords;
        if (rng_mode == MCX_MC_REPLAY) mt_seed(w, seeds ? seeds[r] : seed + (uint64_t)r);
        w[627] = 32;   // empty draw batch: the first draw refills it
    }
    std::vector<double> dm((size_t)R * kMaxMoveTypes, 0.0);
    for (int r = 0; r < R; r++)
        for (size_t t = 0; t < b->types[r].size(); t++) dm[(size_t)r * kMaxMoveTypes + t] = b->types[r][t].dmax;
    std::vector<long long> z8((size_t)R * kMaxMoveTypes, 0), zr(R, 0), zx(std::max(R - 1, 1), 0);
    std::vector<int> zi((size_t)R * kMaxMoveTypes, 0), ident(R);
    for (int r = 0; r < R; r++) ident[r] = r;
    const unsigned long long zc = 0;
    b->rng.upload(st.data(), st.size(), b->stream);
    b->etot.upload(e0.data(), R, b->stream);
    b->dmax.upload(dm.data(), dm.size(), b->stream);
    b->trials.upload(z8.data(), z8.size(), b->stream);
    b->accepted.upload(z8.data(), z8.size(), b->stream);
    b->acc_tmp.upload(zi.data(), zi.size(), b->stream);
    b->step.upload(zr.data(), R, b->stream);
    b->slot_of_rep.upload(ident.data(), R, b->stream);
    b->rep_of_slot.upload(ident.data(), R, b->stream);
    b->xatt.upload(zx.data(), zx.size(), b->stream);
    b->xacc.upload(zx.data(), zx.size(), b->stream);
    b->xcount.upload(&zc, 1, b->stream);
    std::vector<unsigned long long> zp(R, 0);
    b->npair.upload(zp.data(), R, b->stream);
    HIPCHK(hipStreamSynchronize(b->stream));
    b->sweeps = 0;
}

// after device work: the handles' host copies follow the device coordinates
// and their pair lists are stale
void finish(mcx_batch *b)
{
    HIPCHK(hipStreamSynchronize(b->stream));
    for (mcx_ctx *h : b->reps) {
        sync_host_coords(h);
        h->list_valid = false;
    }
}

unsigned long long sum_pairs(mcx_batch *b)
{
    std::vector<unsigned long long> a(b->nrep);
    HIPCHK(hipMemcpy(a.data(), b->npair.p, a.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    unsigned long long s = 0;
    for (unsigned long long v : a) s += v;
    return s;
}

long long sum_accepted(mcx_batch *b)
{
    std::vector<long long> a((size_t)b->nrep * kMaxMoveTypes);
    HIPCHK(hipMemcpy(a.data(), b->accepted.p, a.size() * sizeof(long long), hipMemcpyDeviceToHost));
    long long s = 0;
    for (long long v : a) s += v;
    return s;
}

void exchange_totals(mcx_batch *b, long long &att, long long &acc)
{
    std::vector<long long> x(std::max(b->nrep - 1, 1)), y(x.size());
    HIPCHK(hipMemcpy(x.data(), b->xatt.p, x.size() * sizeof(long long), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(y.data(), b->xacc.p, y.size() * sizeof(long long), hipMemcpyDeviceToHost));
    att = acc = 0;
    for (size_t k = 0; k + 1 < (size_t)b->nrep; k++) { att += x[k]; acc += y[k]; }
}

// per-replica coordinates and run state, for the tuner's trials
struct Snapshot {
    std::vector<DBuf<double>> xyz;
    std::vector<DBuf<int4>> ip;
    DBuf<uint32_t> rng;
    DBuf<double> etot, dmax;
    DBuf<long long> trials, accepted, step;
    DBuf<int> acc_tmp;
    template <typename T>
    static void copy(T *dst, const T *src, size_t n, hipStream_t st)
    {
        HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToDevice, st));
    }
    void take(mcx_batch *b)
    {
        const int R = b->nrep;
        xyz.resize(3 * R);
        ip.resize(R);
        for (int r = 0; r < R; r++) {
            mcx_ctx *h = b->reps[r];
            DBuf<double> *src[3] = {&h->x, &h->y, &h->z};
            for (int k = 0; k < 3; k++) {
                xyz[3 * r + k].alloc(h->n);
                copy(xyz[3 * r + k].p, src[k]->p, h->n, b->stream);
            }
            ip[r].alloc(h->n);
            copy(ip[r].p, h->ipos.p, h->n, b->stream);
        }
        const size_t m = (size_t)R * kMaxMoveTypes;
        rng.alloc((size_t)R * kRngWords); copy(rng.p, b->rng.p, (size_t)R * kRngWords, b->stream);
        etot.alloc(R); copy(etot.p, b->etot.p, R, b->stream);
        step.alloc(R); copy(step.p, b->step.p, R, b->stream);
        dmax.alloc(m); copy(dmax.p, b->dmax.p, m, b->stream);
        trials.alloc(m); copy(trials.p, b->trials.p, m, b->stream);
        accepted.alloc(m); copy(accepted.p, b->accepted.p, m, b->stream);
        acc_tmp.alloc(m); copy(acc_tmp.p, b->acc_tmp.p, m, b->stream);
    }
    void restore(mcx_batch *b)
    {
        const int R = b->nrep;
        for (int r = 0; r < R; r++) {
            mcx_ctx *h = b->reps[r];
            DBuf<double> *dst[3] = {&h->x, &h->y, &h->z};
            for (int k = 0; k < 3; k++) copy(dst[k]->p, xyz[3 * r + k].p, h->n, b->stream);
            copy(h->ipos.p, ip[r].p, h->n, b->stream);
        }
        const size_t m = (size_t)R * kMaxMoveTypes;
        copy(b->rng.p, rng.p, (size_t)R * kRngWords, b->stream);
        copy(b->etot.p, etot.p, R, b->stream);
        copy(b->step.p, step.p, R, b->stream);
        copy(b->dmax.p, dmax.p, m, b->stream);
        copy(b->trials.p, trials.p, m, b->stream);
        copy(b->accepted.p, accepted.p, m, b->stream);
        copy(b->acc_tmp.p, acc_tmp.p, m, b->stream);
    }
    void release()
    {
        for (auto &d : xyz) d.release();
        for (auto &d : ip) d.release();
        rng.release(); etot.release(); dmax.release(); trials.release(); accepted.release(); step.release();
        acc_tmp.release();
    }
};

}  // namespace

extern "C" {

int mcx_batch_create(mcx_ctx *const *reps, int nrep, const mcx_move_type *types, const int *ntypes,
                     mcx_batch **out)
{
    if (!out) return MCX_ERR_ARG;
    *out = nullptr;
    if (!reps || nrep < 1 || !types || !ntypes) return MCX_ERR_ARG;
    for (int r = 0; r < nrep; r++)
        if (!reps[r] || ntypes[r] < 1 || ntypes[r] > kMaxMoveTypes) return MCX_ERR_ARG;
    auto b = std::unique_ptr<mcx_batch>(new mcx_batch());
    b->device = reps[0]->device;
    int rc = bguard(b.get(), [&] {
        for (int r = 0; r < nrep; r++) {
            const mcx_ctx *h = reps[r];
            if (h->device != b->device) throw mcx_error(MCX_ERR_ARG, "replicas on different devices");
            if (!h->have_atoms || !h->have_coords || !h->have_excl)
                throw mcx_error(MCX_ERR_STATE, "replica " + std::to_string(r) + " is not set up");
            if (h->pbtype == MCX_PBC_NONE)
                throw mcx_error(MCX_ERR_UNSUPPORTED, "batched replicas need a periodic box");
            for (int q = 0; q < r; q++)
                if (reps[q] == h) throw mcx_error(MCX_ERR_ARG, "a handle appears twice in the batch");
        }
        HIPCHK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        int lim = 0;
        HIPCHK(hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, b->device));
        b->lds_limit = (size_t)lim;
        b->nrep = nrep;
        b->reps.assign(reps, reps + nrep);
        b->types.resize(nrep);
        size_t at = 0;
        for (int r = 0; r < nrep; r++) {
            b->types[r].assign(types + at, types + at + ntypes[r]);
            at += ntypes[r];
            for (const auto &t : b->types[r])
                if (t.type == MCX_MV_FBIAS || t.type == MCX_MV_FBIAS_ACC)
                    throw mcx_error(MCX_ERR_UNSUPPORTED, "force-bias moves run in mcx_mc_run only");
        }
        b->mbufs.resize(nrep);
        b->moves.resize(nrep);
        b->sweep_len.resize(nrep);
        b->temps.resize(nrep);
        for (int r = 0; r < nrep; r++) {
            mcx_ctx *h = b->reps[r];
            b->moves[r] = build_moves(h, b->types[r].data(), ntypes[r], b->mbufs[r], kBatchAtoms);
            b->sweep_len[r] = residue_count(h);
            b->temps[r] = h->temp;
            b->max_n = std::max(b->max_n, h->n);
        }
        for (auto &tr : b->types)
            for (auto &t : tr) t.sele = nullptr;   // selections were consumed by build_moves
        const size_t R = nrep, m = R * kMaxMoveTypes;
        b->rng.alloc(R * kRngWords); b->etot.alloc(R); b->dmax.alloc(m); b->trials.alloc(m);
        b->accepted.alloc(m); b->acc_tmp.alloc(m); b->step.alloc(R);
        b->slot_of_rep.alloc(R); b->rep_of_slot.alloc(R);
        b->xatt.alloc(std::max(nrep - 1, 1)); b->xacc.alloc(std::max(nrep - 1, 1)); b->xcount.alloc(1);
        b->npair.alloc(R);
        b->dtemps.upload(b->temps.data(), R, b->stream);
        reset(b.get(), 0, MCX_MC_PHILOX, nullptr);
        b->prepare();
    });
    if (rc != MCX_OK) {
        // the error text goes with the first replica handle, which outlives the batch
        reps[0]->err = b->err;
        return rc;
    }
    *out = b.release();
    return MCX_OK;
}

void mcx_batch_destroy(mcx_batch *b)
{
    if (!b) return;
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->stream);
    delete b;
}

const char *mcx_batch_last_error(const mcx_batch *b) { return b ? b->err.c_str() : "null batch"; }

int mcx_batch_reset(mcx_batch *b, uint64_t seed, int rng_mode, const uint64_t *seeds)
{
    return bguard(b, [&] { reset(b, seed, rng_mode, seeds); });
}

int mcx_batch_set_ladder(mcx_batch *b, const double *temps)
{
    return bguard(b, [&] {
        if (!temps) throw mcx_error(MCX_ERR_ARG, "null temperatures");
        for (int k = 0; k < b->nrep; k++)
            if (!(temps[k] > 0.0)) throw mcx_error(MCX_ERR_ARG, "temperatures must be positive");
        b->temps.assign(temps, temps + b->nrep);
        b->dtemps.upload(b->temps.data(), b->nrep, b->stream);
        HIPCHK(hipStreamSynchronize(b->stream));
    });
}

int mcx_batch_set_sweep(mcx_batch *b, const int *moves_per_sweep)
{
    return bguard(b, [&] {
        if (!moves_per_sweep) throw mcx_error(MCX_ERR_ARG, "null sweep lengths");
        for (int r = 0; r < b->nrep; r++) {
            if (moves_per_sweep[r] < 0) throw mcx_error(MCX_ERR_ARG, "negative sweep length");
            b->sweep_len[r] = moves_per_sweep[r];
            b->hrep[r].sweep_len = moves_per_sweep[r];
        }
        b->drep.upload(b->hrep.data(), b->hrep.size(), b->stream);
        std::vector<int> ord(b->nrep);
        for (int r = 0; r < b->nrep; r++) ord[r] = r;
        std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) {
            return (long long)b->reps[x]->n * b->sweep_len[x] > (long long)b->reps[y]->n * b->sweep_len[y];
        });
        b->order.upload(ord.data(), ord.size(), b->stream);
        HIPCHK(hipStreamSynchronize(b->stream));
    });
}

int mcx_batch_run(mcx_batch *b, int nsweeps, int exchange_every, int use_graph, mcx_batch_stats *stats)
{
    return bguard(b, [&] {
        if (nsweeps < 0 || exchange_every < 0) throw mcx_error(MCX_ERR_ARG, "negative sweep count");
        const long long acc0 = stats ? sum_accepted(b) : 0;
        const unsigned long long np0 = stats ? sum_pairs(b) : 0;
        long long xa0 = 0, xc0 = 0;
        if (stats) exchange_totals(b, xa0, xc0);
        HIPCHK(hipStreamSynchronize(b->stream));
        const auto t0 = std::chrono::steady_clock::now();
        const bool exch = exchange_every > 0 && b->nrep > 1;
        auto one = [&] {
            b->launch_sweep(0, 0, nullptr, nullptr);
            b->sweeps++;
            if (exch && b->sweeps % exchange_every == 0) {
                launch_exchange(b->xargs(), b->stream);
                HIPCHK(hipGetLastError());
            }
        };
        int left = nsweeps;
        if (use_graph && exch) {
            // align to an exchange boundary, then replay [sweep x every, exchange]
            while (left > 0 && b->sweeps % exchange_every != 0) { one(); left--; }
            if (left >= exchange_every) {
                const auto key = std::make_tuple(exchange_every, b->threads, b->unroll);
                if (!b->gexec || b->gkey != key) {
                    b->drop_graph();
                    hipGraph_t g = nullptr;
                    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
                    const bool img = b->img_fits(b->threads, b->unroll);
                    for (int k = 0; k < exchange_every; k++)
                        launch_batch_sweep(b->args(0, 0, nullptr, nullptr, img), b->nrep, b->threads, b->unroll, img,
                                           b->stream);
                    launch_exchange(b->xargs(), b->stream);
                    HIPCHK(hipStreamEndCapture(b->stream, &g));
                    const hipError_t ie = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
                    (void)hipGraphDestroy(g);
                    HIPCHK(ie);
                    b->gkey = key;
                }
                while (left >= exchange_every) {
                    HIPCHK(hipGraphLaunch(b->gexec, b->stream));
                    b->sweeps += exchange_every;
                    left -= exchange_every;
                }
            }
        } else if (use_graph && left > 0) {
            const auto key = std::make_tuple(0, b->threads, b->unroll);
            if (!b->gexec || b->gkey != key) {
                b->drop_graph();
                hipGraph_t g = nullptr;
                HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
                const bool img = b->img_fits(b->threads, b->unroll);
                launch_batch_sweep(b->args(0, 0, nullptr, nullptr, img), b->nrep, b->threads, b->unroll, img,
                                   b->stream);
                HIPCHK(hipStreamEndCapture(b->stream, &g));
                const hipError_t ie = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                HIPCHK(ie);
                b->gkey = key;
            }
            for (; left > 0; left--) {
                HIPCHK(hipGraphLaunch(b->gexec, b->stream));
                b->sweeps++;
            }
        }
        for (; left > 0; left--) one();
        HIPCHK(hipStreamSynchronize(b->stream));
        const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        finish(b);
        if (stats) {
            stats->moves = 0;
            for (int r = 0; r < b->nrep; r++) stats->moves += (long long)b->sweep_len[r] * nsweeps;
            stats->accepted = sum_accepted(b) - acc0;
            long long xa = 0, xc = 0;
            exchange_totals(b, xa, xc);
            stats->exchange_attempts = xa - xa0;
            stats->exchange_accepted = xc - xc0;
            stats->sweeps = b->sweeps;
            stats->seconds = secs;
            stats->pair_evals = (long long)(sum_pairs(b) - np0);
        }
    });
}

int mcx_batch_moves(mcx_batch *b, int nmoves, unsigned char *acc_out, double *de_out)
{
    return bguard(b, [&] {
        if (nmoves < 1) throw mcx_error(MCX_ERR_ARG, "nmoves must be positive");
        DBuf<unsigned char> acc;
        DBuf<double> de;
        struct Release {
            DBuf<unsigned char> &a;
            DBuf<double> &d;
            ~Release() { a.release(); d.release(); }
        } rel{acc, de};
        const size_t n = (size_t)b->nrep * nmoves;
        if (acc_out) acc.alloc(n);
        if (de_out) de.alloc(n);
        b->launch_sweep(nmoves, (acc_out || de_out) ? nmoves : 0, acc_out ? acc.p : nullptr, de_out ? de.p : nullptr);
        HIPCHK(hipStreamSynchronize(b->stream));
        if (acc_out) HIPCHK(hipMemcpy(acc_out, acc.p, n, hipMemcpyDeviceToHost));
        if (de_out) HIPCHK(hipMemcpy(de_out, de.p, n * sizeof(double), hipMemcpyDeviceToHost));
        finish(b);
    });
}

int mcx_batch_state(mcx_batch *b, double *etot, int *slot_of_rep, long long *trials, long long *accepted,
                    double *dmax)
{
    return bguard(b, [&] {
        HIPCHK(hipStreamSynchronize(b->stream));
        const int R = b->nrep;
        const size_t m = (size_t)R * kMaxMoveTypes;
        if (etot) HIPCHK(hipMemcpy(etot, b->etot.p, R * sizeof(double), hipMemcpyDeviceToHost));
        if (slot_of_rep) HIPCHK(hipMemcpy(slot_of_rep, b->slot_of_rep.p, R * sizeof(int), hipMemcpyDeviceToHost));
        if (trials) HIPCHK(hipMemcpy(trials, b->trials.p, m * sizeof(long long), hipMemcpyDeviceToHost));
        if (accepted) HIPCHK(hipMemcpy(accepted, b->accepted.p, m * sizeof(long long), hipMemcpyDeviceToHost));
        if (dmax) HIPCHK(hipMemcpy(dmax, b->dmax.p, m * sizeof(double), hipMemcpyDeviceToHost));
    });
}

int mcx_batch_exchange_stats(mcx_batch *b, long long *attempts, long long *accepted)
{
    return bguard(b, [&] {
        HIPCHK(hipStreamSynchronize(b->stream));
        const size_t n = (size_t)std::max(b->nrep - 1, 0);
        if (n == 0) return;
        if (attempts) HIPCHK(hipMemcpy(attempts, b->xatt.p, n * sizeof(long long), hipMemcpyDeviceToHost));
        if (accepted) HIPCHK(hipMemcpy(accepted, b->xacc.p, n * sizeof(long long), hipMemcpyDeviceToHost));
    });
}

int mcx_batch_set_config(mcx_batch *b, int threads, int unroll)
{
    return bguard(b, [&] {
        if (!batch_variant_ok(threads, unroll, true))
            throw mcx_error(MCX_ERR_ARG, "threads must be 256, 512, 768 or 1024 and unroll 1, 2 or 4");
        b->threads = threads;
        b->unroll = unroll;
        b->prepare();
    });
}

int mcx_batch_tune(mcx_batch *b, int *threads, int *unroll, double *ms_per_sweep)
{
    return bguard(b, [&] {
        Snapshot snap;
        struct Release {
            Snapshot &s;
            ~Release() { s.release(); }
        } rel{snap};
        snap.take(b);
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        double best = 1e300;
        int bt = b->threads, bu = b->unroll;
        const int T[4] = {256, 512, 768, 1024}, U[3] = {1, 2, 4};
        try {
            for (int nt : T)
                for (int u : U) {
                    b->threads = nt;
                    b->unroll = u;
                    b->prepare();
                    b->launch_sweep(0, 0, nullptr, nullptr);      // warm-up
                    snap.restore(b);
                    HIPCHK(hipEventRecord(e0, b->stream));
                    b->launch_sweep(0, 0, nullptr, nullptr);
                    HIPCHK(hipEventRecord(e1, b->stream));
                    HIPCHK(hipEventSynchronize(e1));
                    float ms = 0.f;
                    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
                    snap.restore(b);
                    if (ms < best) { best = ms; bt = nt; bu = u; }
                }
        } catch (...) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        HIPCHK(hipEventDestroy(e0));
        HIPCHK(hipEventDestroy(e1));
        HIPCHK(hipStreamSynchronize(b->stream));
        b->threads = bt;
        b->unroll = bu;
        b->prepare();
        b->drop_graph();
        if (threads) *threads = bt;
        if (unroll) *unroll = bu;
        if (ms_per_sweep) *ms_per_sweep = best;
    });
}

}  // extern "C"
