// mcx_api.cpp -- host side of the gfx950 energy backend: the C ABI of
// include/mcx.h, device buffer management, the exclusion-list builder and
// the move-instance tables (both host work in the reference too).
#include "mcx_ctx.h"

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>

namespace mcxi {

int fail(mcx_ctx *h, int code, const std::string &msg)
{
    if (h) h->err = msg;
    return code;
}

template <typename F>
int guarded(mcx_ctx *h, F &&f)
{
    if (!h) return MCX_ERR_ARG;
    try {
        HIPCHK(hipSetDevice(h->device));
        f();
        return MCX_OK;
    } catch (const mcx_error &e) {
        return fail(h, e.code, e.msg);
    } catch (const std::exception &e) {
        return fail(h, MCX_ERR_ARG, e.what());
    }
}

// ---- exclusion lists: List_nonBonded::build_exclude_list restated
// (src/List_nonBonded.cpp:143-775) ----
void build_exclusions_host(mcx_ctx *h)
{
    const int N = h->n;
    std::vector<std::vector<int>> tmp(N), con(N);
    std::vector<int> v14;
    auto add = [&](int a, int b) -> bool {
        for (int e : tmp[a]) if (e == b) return false;
        tmp[a].push_back(b);
        tmp[b].push_back(a);
        return true;
    };
    const HostTerms &T = h->terms;
    for (const int4 &b : T.idx[kBond]) {     // excl_bonds :184-213
        con[b.x].push_back(b.y); con[b.y].push_back(b.x);
        tmp[b.x].push_back(b.y); tmp[b.y].push_back(b.x);
    }
    for (const int4 &a : T.idx[kAngle]) { add(a.x, a.y); add(a.x, a.z); add(a.y, a.z); }   // :215-299
    for (const int4 &d : T.idx[kDihe]) {                                                   // :301-462
        add(d.x, d.y); add(d.x, d.z);
        if (add(d.x, d.w)) { v14.push_back(d.x); v14.push_back(d.w); }
        add(d.y, d.z); add(d.y, d.w); add(d.z, d.w);
    }
    for (const int4 &d : T.idx[kImpr]) {                                                   // :464-619
        add(d.x, d.y); add(d.x, d.z); add(d.x, d.w); add(d.y, d.z); add(d.y, d.w); add(d.z, d.w);
    }
    for (int ia = 0; ia < N; ia++)                                                          // :621-703
        for (int ib : con[ia])
            for (int ic : con[ib]) {
                if (ic == ia) continue;
                add(ia, ic);
                for (int id : con[ic]) {
                    if (id == ia || id == ib) continue;
                    if (add(ia, id)) { v14.push_back(ia); v14.push_back(id); }
                }
            }
    // excl_final_Lists :705-766 -- circular half rule, sorted insertion, and
    // the rotation of entries < i, which drops them from the visible row
    // and repeats the row's last entry
    const int hn = N / 2, hm1 = (N - 1) / 2;
    std::vector<std::vector<int>> rows(N);
    h->excl_pair.assign(N, 0);
    for (int i = 0; i < N; i++) {
        std::vector<int> &L = rows[i];
        L.assign(tmp[i].size() + 1, 0);
        int jj = 0;
        for (int e : tmp[i]) {
            if (((e > i) && ((e - i) < hn)) || ((e < i) && ((e - i + N) < hm1))) {
                L[jj] = e;
                for (int k = jj; k > 0; k--)
                    if (L[k] < L[k - 1]) std::swap(L[k], L[k - 1]);
                jj++;
            }
        }
        h->excl_pair[i] = jj;
        for (int r = 0; r < jj; r++) {
            if (L[0] < i) {
                const int e = L[0];
                for (int kk = 0; kk < jj - 1; kk++) L[kk] = L[kk + 1];
                L[jj] = e;
            }
        }
    }
    h->excl_off.assign(N + 1, 0);
    h->excl_idx.clear();
    for (int i = 0; i < N; i++) {
        h->excl_off[i + 1] = h->excl_off[i] + h->excl_pair[i];
        h->excl_idx.insert(h->excl_idx.end(), rows[i].begin(), rows[i].begin() + h->excl_pair[i]);
    }
    h->pairs14 = v14;
    // pair terms for the fused kernel: 1-4 pairs, and the distinct excluded
    // pairs (i < j) whose all-pair contribution FULL mode must remove
    HostTerms &W = h->terms;
    W.idx[kPair14].clear(); W.type[kPair14].clear(); W.par[kPair14].clear();
    for (size_t k = 0; k + 1 < v14.size(); k += 2) {
        W.idx[kPair14].push_back(make_int4(v14[k], v14[k + 1], -1, -1));
        W.type[kPair14].push_back(0);
        W.par[kPair14].push_back(make_double4(0, 0, 0, 0));
    }
    W.idx[kExclPair].clear(); W.type[kExclPair].clear(); W.par[kExclPair].clear();
    for (int i = 0; i < N; i++) {
        int prev = -1;
        for (int k = h->excl_off[i]; k < h->excl_off[i + 1]; k++) {
            const int e = h->excl_idx[k];
            if (e <= i || e == prev) continue;
            prev = e;
            W.idx[kExclPair].push_back(make_int4(i, e, -1, -1));
            W.type[kExclPair].push_back(0);
            W.par[kExclPair].push_back(make_double4(0, 0, 0, 0));
        }
    }
}

void upload_terms(mcx_ctx *h)
{
    for (int k = 0; k < kNumKinds; k++) {
        h->t_idx[k].upload(h->terms.idx[k].data(), h->terms.idx[k].size(), h->stream);
        h->t_type[k].upload(h->terms.type[k].data(), h->terms.type[k].size(), h->stream);
        h->t_par[k].upload(h->terms.par[k].data(), h->terms.par[k].size(), h->stream);
    }
    // per-atom adjacency of bonded terms and 1-4 pairs: atom ids inline and
    // (kind << 28 | index) so that the dE kernel needs no indirection
    const int N = h->n;
    std::vector<std::vector<std::pair<int4, int>>> adj(N);
    for (int k = 0; k <= kPair14; k++) {
        const auto &v = h->terms.idx[k];
        for (size_t t = 0; t < v.size(); t++) {
            const int ids[4] = {v[t].x, v[t].y, v[t].z, v[t].w};
            for (int q = 0; q < 4; q++) {
                const int a = ids[q];
                if (a < 0) continue;
                bool dup = false;
                for (int p = 0; p < q; p++) dup |= ids[p] == a;
                if (!dup) adj[a].push_back(std::make_pair(v[t], (k << 28) | (int)t));
            }
        }
    }
    std::vector<int> off(N + 1, 0), kk, ty;
    std::vector<int4> ids;
    std::vector<double4> par;
    for (int i = 0; i < N; i++) {
        off[i + 1] = off[i] + (int)adj[i].size();
        for (const auto &p : adj[i]) {
            ids.push_back(p.first);
            kk.push_back(p.second);
            const int kind = p.second >> 28, k = p.second & 0x0fffffff;
            ty.push_back(h->terms.type[kind][k]);
            par.push_back(h->terms.par[kind][k]);
        }
    }
    h->adj_off_h = off;
    h->adj_kk_h = kk;
    h->adj_off.upload(off.data(), off.size(), h->stream);
    h->adj_ids.upload(ids.data(), ids.size(), h->stream);
    h->adj_kk.upload(kk.data(), kk.size(), h->stream);
    h->adj_type.upload(ty.data(), ty.size(), h->stream);
    h->adj_par.upload(par.data(), par.size(), h->stream);
    h->excl_off_d.upload(h->excl_off.data(), h->excl_off.size(), h->stream);
    h->excl_idx_d.upload(h->excl_idx.data(), h->excl_idx.size(), h->stream);
    // both-direction excluded partners (pair (a,b), a<b, excluded iff b is
    // in the visible row of a), used by the list build to clear bits
    std::vector<std::vector<int>> xf(N);
    for (int a = 0; a < N; a++)
        for (int k = h->excl_off[a]; k < h->excl_off[a + 1]; k++) {
            const int b = h->excl_idx[k];
            if (b <= a) continue;
            xf[a].push_back(b);
            xf[b].push_back(a);
        }
    std::vector<int> xo(N + 1, 0), xi;
    for (int a = 0; a < N; a++) {
        std::sort(xf[a].begin(), xf[a].end());
        xf[a].erase(std::unique(xf[a].begin(), xf[a].end()), xf[a].end());
        xo[a + 1] = xo[a] + (int)xf[a].size();
        xi.insert(xi.end(), xf[a].begin(), xf[a].end());
    }
    h->xfull_off_d.upload(xo.data(), xo.size(), h->stream);
    h->xfull_idx_d.upload(xi.data(), xi.size(), h->stream);
    h->xfull_off_h = xo;
    h->xfull_idx_h = xi;
    // both-direction 1-4 partners, one entry per 1-4 pair record (the
    // adjoint kernel's per-atom rows)
    std::vector<int> po(N + 1, 0), pi;
    const int n14 = (int)h->pairs14.size() / 2;
    for (int k = 0; k < n14; k++) { po[h->pairs14[2 * k] + 1]++; po[h->pairs14[2 * k + 1] + 1]++; }
    for (int a = 0; a < N; a++) po[a + 1] += po[a];
    pi.resize(po[N]);
    {
        std::vector<int> cur(po.begin(), po.end() - 1);
        for (int k = 0; k < n14; k++) {
            const int a = h->pairs14[2 * k], b = h->pairs14[2 * k + 1];
            pi[cur[a]++] = b;
            pi[cur[b]++] = a;
        }
    }
    h->p14_off_d.upload(po.data(), po.size(), h->stream);
    h->p14_idx_d.upload(pi.data(), pi.size(), h->stream);
}

void setup_grid(mcx_ctx *h)
{
    CellGrid &g = h->grid;
    const double cutnb = h->ctoff + h->dcut;
    const int reach = 2;
    for (int d = 0; d < 3; d++) {
        g.reach[d] = reach;
        if (h->pbtype == MCX_PBC_NONE) { g.nc[d] = 1; continue; }
        const double wmin = cutnb / reach * (1.0 + 1e-9) + 1e-6;
        int nc = (int)std::floor(h->box[d] / wmin);
        g.nc[d] = std::max(1, std::min(nc, 1024));
    }
    g.ncell = g.nc[0] * g.nc[1] * g.nc[2];
    h->cell_of.alloc(h->n);
    h->cell_atoms.alloc(h->n);
    h->cell_start.alloc(g.ncell);
    h->cell_count.alloc(g.ncell);
    h->cursor.alloc(g.ncell);
    h->cpos.alloc(h->n);
    h->cfix.alloc(h->n);
    h->ipos_ref.alloc(h->n);
    h->cell_tmp.alloc(h->n);
    h->group_start.alloc(g.ncell + 1);
    h->group_info.alloc((size_t)h->n + g.ncell + 1);
    h->cell_acc.alloc(g.ncell);
    h->bin_ticket.alloc(1);
    // the binning's accumulators and ticket are zero between builds (its
    // last block resets them); set here, stream-ordered after earlier builds
    HIPCHK(hipMemsetAsync(h->cell_acc.p, 0, sizeof(int) * g.ncell, h->stream));
    HIPCHK(hipMemsetAsync(h->bin_ticket.p, 0, sizeof(unsigned), h->stream));
    g.group_info = h->group_info.p;
    g.cell_acc = h->cell_acc.p;
    g.ticket = h->bin_ticket.p;
    g.cpos = h->cpos.p;
    g.cfix = h->cfix.p;
    g.ipos_ref = h->ipos_ref.p;
    g.cell_tmp = h->cell_tmp.p;
    g.group_start = h->group_start.p;
    g.cell_of = h->cell_of.p;
    g.cell_atoms = h->cell_atoms.p;
    g.cell_start = h->cell_start.p;
    g.cell_count = h->cell_count.p;
    g.cursor = h->cursor.p;
}

void ensure_list_capacity(mcx_ctx *h, int cap)
{
    cap = (cap + 31) & ~31;
    if (cap <= h->cap && h->nbr.p) return;
    h->cap = cap;
    h->nbr.alloc((size_t)h->n * cap);
    h->cnt.alloc(h->n);
    h->half.alloc(h->n);
    h->overflow.alloc(1);
}

long long build_list(mcx_ctx *h)
{
    if (!h->have_coords || !h->have_excl) throw mcx_error(MCX_ERR_STATE, "coordinates and exclusions must be set before the list");
    if ((size_t)((h->n + 31) / 32) * 4 > 60 * 1024) throw mcx_error(MCX_ERR_UNSUPPORTED, "list build supports up to 491,520 atoms");
    setup_grid(h);
    if (h->cap == 0) {
        double vol = h->box[0] * h->box[1] * h->box[2];
        double rho = h->pbtype == MCX_PBC_NONE || vol <= 0 ? 0.1 : h->n / vol;
        double rc = h->ctoff + h->dcut;
        int est = (int)(rho * 4.18879 * rc * rc * rc * 1.3) + 64;
        ensure_list_capacity(h, std::min(est, h->n));
    }
    for (int attempt = 0; attempt < 4; attempt++) {
        DevSys s = h->sys();
        launch_list_build(s, h->grid, h->stream);
        HIPCHK(hipGetLastError());
        int ov = 0;
        HIPCHK(hipMemcpyAsync(&ov, h->overflow.p, sizeof(int), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        if (ov == 0) {
            h->list_valid = true;
            if (h->patch_alloc) {
                // a fresh binning: nothing has moved since this build
                launch_list_patch_reset(h->sys(), h->patch(), h->stream);
                HIPCHK(hipGetLastError());
                h->patch_valid = true;
            }
            // count the half pairs
            std::vector<int> c(h->n), hf(h->n);
            HIPCHK(hipMemcpy(c.data(), h->cnt.p, sizeof(int) * h->n, hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(hf.data(), h->half.p, sizeof(int) * h->n, hipMemcpyDeviceToHost));
            long long np = 0;
            for (int i = 0; i < h->n; i++) np += c[i] - hf[i];
            h->last_npairs = np;
            return np;
        }
        ensure_list_capacity(h, (int)(ov * 1.25) + 32);
    }
    throw mcx_error(MCX_ERR_LIST_OVERFLOW, "pair list capacity could not be grown");
}

// list rebuild without the host read-back of counts (used inside MC runs)
void rebuild_list_fast(mcx_ctx *h)
{
    DevSys s = h->sys();
    launch_list_build(s, h->grid, h->stream);
    int ov = 0;
    HIPCHK(hipMemcpyAsync(&ov, h->overflow.p, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    if (ov) {
        ensure_list_capacity(h, (int)(ov * 1.25) + 32);
        build_list(h);
    } else if (h->patch_alloc) {
        launch_list_patch_reset(h->sys(), h->patch(), h->stream);
        HIPCHK(hipGetLastError());
        h->patch_valid = true;
    }
}

// Buffers of the incremental list update, allocated once per handle (the
// rebuilt-row buffer grows with the list capacity); new state starts empty.
void ensure_patch(mcx_ctx *h)
{
    const size_t n = (size_t)h->n;
    const size_t rows = std::min(n, (size_t)kPatchRows);
    h->p_newrow.alloc(rows * (size_t)std::max(h->cap, 1));
    if (h->patch_alloc) return;
    h->p_moved_flag.alloc(n); h->p_moved_list.alloc(n); h->p_moved_cnt.alloc(1);
    h->p_stale_flag.alloc(n); h->p_stale_list.alloc(n); h->p_stale_cnt.alloc(1);
    h->p_edit_cnt.alloc(n); h->p_edit_val.alloc(n * kPatchEdits);
    h->p_touched.alloc(n); h->p_touched_cnt.alloc(1);
    h->p_newcnt.alloc(rows); h->p_newhalf.alloc(rows);
    h->p_ticket.alloc(1);
    h->patch_alloc = true;
    h->patch_valid = false;
}

void setup_tiles(mcx_ctx *h)
{
    const int nt = (h->n + kFullTile - 1) / kFullTile;
    std::vector<int2> t;
    for (int I = 0; I < nt; I++)
        for (int J = I; J < nt; J++) t.push_back(make_int2(I, J));
    h->ntiles = (int)t.size();
    h->tiles.upload(t.data(), t.size(), h->stream);
}

void energy_device(mcx_ctx *h, double out9[9])
{
    if (!h->have_coords || !h->have_atoms || !h->have_excl)
        throw mcx_error(MCX_ERR_STATE, "atoms, coordinates and exclusions must be set before getE");
    DevSys s = h->sys();
    DevTerms t = h->dterms();
    int nb_nb;
    h->result.alloc(16);
    if (h->cutmode == MCX_CUT_FULL) {
        if (h->ntiles == 0) setup_tiles(h);
        nb_nb = h->ntiles;
        h->partial.alloc((size_t)nb_nb * 2 + (size_t)bonded_blocks(t) * 7);
        launch_full_energy(s, h->tiles.p, h->ntiles, kFullTile, h->partial.p, h->stream);
    } else {
        if (!h->list_valid) build_list(h);
        s = h->sys();
        nb_nb = switch_energy_blocks(h->n);
        h->partial.alloc((size_t)nb_nb * 2 + (size_t)bonded_blocks(t) * 7);
        launch_switch_energy(s, h->partial.p, nb_nb, h->stream);
    }
    // the bonded partials after the nonbonded ones, both sums in one launch
    const int nbb = bonded_blocks(t);
    double *pb = h->partial.p + (size_t)nb_nb * 2;
    launch_bonded(s, t, pb, nbb, h->stream);
    launch_reduce2(h->partial.p, nb_nb, 2, h->result.p, pb, nbb, 7, h->result.p + 2, h->stream);
    HIPCHK(hipGetLastError());
    if (!h->res_h) HIPCHK(hipHostMalloc((void **)&h->res_h, 16 * sizeof(double), hipHostMallocDefault));
    HIPCHK(hipMemcpyAsync(h->res_h, h->result.p, 9 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    std::memcpy(out9, h->res_h, 9 * sizeof(double));
}

void assemble(const double r[9], double ener[10])
{
    const double C = MCX_CHGCHARMM * MCX_KCALTOIU;
    const double elec = C * (r[0] + r[2]);
    const double vdw = 4.0 * (r[1] + r[3]);
    const double bond = r[4], ang = r[5], ub = r[6], dihe = r[7], impr = r[8];
    const double pot = elec + vdw + bond + ang + ub + dihe + impr;
    ener[0] = pot + 0.0; ener[1] = pot; ener[2] = 0.0; ener[3] = elec; ener[4] = vdw;
    ener[5] = bond; ener[6] = ang; ener[7] = ub; ener[8] = dihe; ener[9] = impr;
}

// fractional coordinate -> 32-bit fixed point, as the device computes it
int to_fixed(double x, double rL)
{
    double v = x * rL;
    v = v - std::floor(v);
    return (int)(unsigned)(long long)(v * 4294967296.0);
}

void upload_mirror(mcx_ctx *h)
{
    const int n = h->n;
    if (!h->have_coords) return;
    if (h->host_stale) sync_host_coords(h);
    std::vector<int4> m(n);
    const bool pbc = h->pbtype != MCX_PBC_NONE;
    for (int i = 0; i < n; i++) {
        float q = h->have_atoms ? (float)h->q[i] : 0.f;
        int qb;
        std::memcpy(&qb, &q, 4);
        if (pbc)
            m[i] = make_int4(to_fixed(h->hx[i], 1.0 / h->box[0]), to_fixed(h->hy[i], 1.0 / h->box[1]),
                             to_fixed(h->hz[i], 1.0 / h->box[2]), qb);
        else
            m[i] = make_int4(0, 0, 0, qb);
    }
    h->ipos.upload(m.data(), n, h->stream);
    std::vector<float2> ek(n);
    const double sc = h->pbtype == MCX_PBC_CUBIC ? 4294967296.0 / h->box[0] : 1.0;
    for (int i = 0; i < n; i++) ek[i] = make_float2((float)h->eps[i], (float)(h->sig[i] * sc));
    h->esk.upload(ek.data(), n, h->stream);
    std::vector<TileAtom> tq(n);
    for (int i = 0; i < n; i++) tq[i] = TileAtom{(double)__builtin_bit_cast(float, m[i].w), ek[i].x, ek[i].y};
    h->tq.upload(tq.data(), n, h->stream);
    HIPCHK(hipStreamSynchronize(h->stream));
}

void upload_coords(mcx_ctx *h, const double *x, const double *y, const double *z)
{
    const int n = h->n;
    h->x.upload(x, n, h->stream);
    h->y.upload(y, n, h->stream);
    h->z.upload(z, n, h->stream);
    h->hx.assign(x, x + n); h->hy.assign(y, y + n); h->hz.assign(z, z + n);
    h->have_coords = true;
    h->list_valid = false;
    upload_mirror(h);
}

// pull the device master coordinates back into the host copy (after MC runs)
void sync_host_coords(mcx_ctx *h)
{
    const int n = h->n;
    h->host_stale = false;
    h->hx.resize(n); h->hy.resize(n); h->hz.resize(n);
    HIPCHK(hipMemcpyAsync(h->hx.data(), h->x.p, n * 8, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(h->hy.data(), h->y.p, n * 8, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(h->hz.data(), h->z.p, n * 8, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
}

// List_Moves::NewMove_TRN_ROT instance tables (src/List_Moves.cpp:257-457,
// gtrsfl :630-697): one instance per selected atom (RESIDUE: the residue
// range, pivot = the atom for ROT; ATOM: the atom; ALL: one instance of the
// whole contiguous selection with a centre-of-mass pivot).
// gtrsfl's table: resLast[k] = last atom of residue k, resLast[0] = -1
std::vector<int> residue_last_table(const mcx_ctx *h)
{
    const int N = h->n;
    const int iresl = *std::max_element(h->resid.begin(), h->resid.end());
    std::vector<int> resLast(iresl + 1);
    resLast[0] = -1;
    int last = 1, idx = 0;
    for (int it = 0; it < N; it++) if (h->resid[it] > last) { last++; resLast[++idx] = it - 1; }
    resLast[iresl] = N - 1;
    return resLast;
}

// gtrsfl's binary search, bounds inclusive on both sides as in the
// reference (src/List_Moves.cpp:659-685): for the last atom of a residue
// the search may settle on the following residue, depending on its path.
void residue_range(const std::vector<int> &resLast, int atomidx, int &a1, int &a2)
{
    int iresf = 1, iresl = (int)resLast.size() - 1;
    while (iresf <= iresl) {
        const int mid = (iresl + iresf) / 2;
        if (resLast[mid - 1] <= atomidx && resLast[mid] >= atomidx) { a1 = resLast[mid - 1] + 1; a2 = resLast[mid]; return; }
        else if (resLast[mid] <= atomidx) iresf = mid + 1;
        else iresl = mid - 1;
    }
    throw mcx_error(MCX_ERR_ARG, "residue bounds not found");
}

DevMoves build_moves(mcx_ctx *h, const mcx_move_type *types, int ntypes, MoveBufs &buf, int max_atoms)
{
    if (ntypes < 1 || ntypes > kMaxMoveTypes) throw mcx_error(MCX_ERR_ARG, "1..8 move types supported");
    DevMoves mv{};
    mv.ntypes = ntypes;
    std::vector<int> F, L, P;
    const std::vector<int> resLast = residue_last_table(h);
    for (int t = 0; t < ntypes; t++) {
        const mcx_move_type &m = types[t];
        if (m.type != MCX_MV_TRN && m.type != MCX_MV_ROT && m.type != MCX_MV_FBIAS && m.type != MCX_MV_FBIAS_ACC)
            throw mcx_error(MCX_ERR_ARG, "move type must be TRN, ROT, FBIAS or FBIAS_ACC");
        if (m.mode == MCX_MODE_ATOM && m.type == MCX_MV_ROT) throw mcx_error(MCX_ERR_ARG, "rotation of a single atom");
        const bool fbias = m.type == MCX_MV_FBIAS || m.type == MCX_MV_FBIAS_ACC;
        if (fbias && !(m.dmax > 0.0)) throw mcx_error(MCX_ERR_ARG, "force-bias step A must be > 0");
        mv.type[t] = fbias ? MCX_MV_FBIAS : m.type;
        mv.fb_scale[t] = m.type == MCX_MV_FBIAS_ACC ? 1.0 : 1.0 / MCX_KCALTOIU;
        mv.dmax0[t] = m.dmax;
        mv.target[t] = m.target;
        mv.each[t] = m.each;
        mv.off[t] = (int)F.size();
        if (m.mode == MCX_MODE_ALL) {
            F.push_back(0); L.push_back(h->n - 1); P.push_back(-1);
        } else {
            for (int i = 0; i < h->n; i++) {
                if (m.sele && !m.sele[i]) continue;
                int a1 = i, a2 = i;
                if (m.mode == MCX_MODE_RESIDUE) residue_range(resLast, i, a1, a2);
                if (a2 - a1 + 1 > max_atoms)
                    throw mcx_error(MCX_ERR_UNSUPPORTED, "move instance larger than " + std::to_string(max_atoms) +
                                                             " atoms");
                F.push_back(a1); L.push_back(a2); P.push_back(m.type == MCX_MV_ROT ? i : -1);
            }
        }
        mv.ninst[t] = (int)F.size() - mv.off[t];
        if (mv.ninst[t] == 0) throw mcx_error(MCX_ERR_ARG, "empty move selection");
        if (m.mode == MCX_MODE_ALL && h->n > max_atoms)
            throw mcx_error(MCX_ERR_UNSUPPORTED, "ALL-mode moves of more than " + std::to_string(max_atoms) +
                                                     " atoms");
    }
    // every bonded term / 1-4 pair touching an instance's atom range, once,
    // as self-contained records (atoms, kind|index, type, parameters);
    // instances with the same range share one record list
    std::vector<int4> rec(F.size());
    std::vector<int2> radj(F.size()), rpairs(F.size());
    std::map<std::pair<int, int>, int2> seen, seen_p;
    std::vector<int2> pent;
    std::vector<int4> rid;
    std::vector<int> rkk, rty;
    std::vector<double4> rpar;
    for (size_t k = 0; k < F.size(); k++) {
        rec[k] = make_int4(F[k], L[k], P[k], 0);
        const auto key = std::make_pair(F[k], L[k]);
        // pair corrections of a list-free dE over all atoms: excluded
        // partners outside the range (kind 0, to subtract) and the
        // non-excluded pairs inside it (kind 1, to add); pairs of two frozen
        // atoms are in neither
        auto itp = seen_p.find(key);
        if (itp != seen_p.end()) {
            rpairs[k] = itp->second;
        } else {
            const int begin = (int)pent.size();
            for (int a = F[k]; a <= L[k]; a++) {
                const int x0 = h->xfull_off_h[a], x1 = h->xfull_off_h[a + 1];
                for (int e = x0; e < x1; e++) {
                    const int j = h->xfull_idx_h[e];
                    if (j >= F[k] && j <= L[k]) continue;
                    if (h->frozen[a] && h->frozen[j]) continue;
                    pent.push_back(make_int2(a - F[k], j));
                }
                for (int b = a + 1; b <= L[k]; b++) {
                    if (std::binary_search(h->xfull_idx_h.begin() + x0, h->xfull_idx_h.begin() + x1, b)) continue;
                    if (h->frozen[a] && h->frozen[b]) continue;
                    pent.push_back(make_int2((a - F[k]) | (1 << 16), b - F[k]));
                }
            }
            rpairs[k] = make_int2(begin, (int)pent.size());
            seen_p[key] = rpairs[k];
        }
        auto it = seen.find(key);
        if (it != seen.end()) { radj[k] = it->second; continue; }
        const int begin = (int)rid.size();
        std::vector<int> used;
        for (int a = F[k]; a <= L[k]; a++)
            for (int e = h->adj_off_h[a]; e < h->adj_off_h[a + 1]; e++) {
                const int kkv = h->adj_kk_h[e];
                if (std::find(used.begin(), used.end(), kkv) != used.end()) continue;
                used.push_back(kkv);
                const int kind = kkv >> 28, ti = kkv & 0x0fffffff;
                rid.push_back(h->terms.idx[kind][ti]);
                rkk.push_back(kkv);
                rty.push_back(h->terms.type[kind][ti]);
                rpar.push_back(h->terms.par[kind][ti]);
            }
        radj[k] = make_int2(begin, (int)rid.size());
        seen[key] = radj[k];
    }
    buf.rid.upload(rid.data(), rid.size(), h->stream);
    buf.rkk.upload(rkk.data(), rkk.size(), h->stream);
    buf.rty.upload(rty.data(), rty.size(), h->stream);
    buf.rpar.upload(rpar.data(), rpar.size(), h->stream);
    buf.inst.upload(rec.data(), rec.size(), h->stream);
    buf.inst_adj.upload(radj.data(), radj.size(), h->stream);
    buf.inst_pairs.upload(rpairs.data(), rpairs.size(), h->stream);
    buf.pair_ent.upload(pent.data(), pent.size(), h->stream);
    mv.inst_pairs = buf.inst_pairs.p;
    mv.pair_ent = buf.pair_ent.p;
    mv.inst = buf.inst.p;
    mv.inst_adj = buf.inst_adj.p;
    mv.rec_ids = buf.rid.p;
    mv.rec_kk = buf.rkk.p;
    mv.rec_type = buf.rty.p;
    mv.rec_par = buf.rpar.p;
    mv.mass = h->mass64.p;
    mv.adj_off = h->adj_off.p;
    mv.adj_ids = h->adj_ids.p;
    mv.adj_kk = h->adj_kk.p;
    mv.adj_type = h->adj_type.p;
    mv.adj_par = h->adj_par.p;
    return mv;
}

DevMoves adjacency_only(mcx_ctx *h)
{
    DevMoves mv{};
    mv.adj_off = h->adj_off.p;
    mv.adj_ids = h->adj_ids.p;
    mv.adj_kk = h->adj_kk.p;
    mv.adj_type = h->adj_type.p;
    mv.adj_par = h->adj_par.p;
    mv.mass = h->mass64.p;
    return mv;
}

void mt_seed(uint32_t *st, uint64_t seed)
{
    st[0] = (uint32_t)(seed & 0xffffffffu);
    for (int i = 1; i < 624; i++) st[i] = 1812433253u * (st[i - 1] ^ (st[i - 1] >> 30)) + (uint32_t)i;
    st[624] = 624;
}

}  // namespace mcxi

using namespace mcxi;

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" {

int mcx_create(int natom, int device, mcx_ctx **out)
{
    if (!out || natom <= 0) return MCX_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return MCX_ERR_HIP;
    auto h = std::unique_ptr<mcx_ctx>(new mcx_ctx());
    h->n = natom;
    h->device = device;
    int rc = guarded(h.get(), [&] {
        HIPCHK(hipSetDevice(device));
        HIPCHK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
        if (noalloc_init_device(device) != MCX_OK)
            throw mcx_error(MCX_ERR_HIP, "address of the device's no-allocation scratch array not found");
        h->q.assign(natom, 0); h->eps.assign(natom, 0); h->sig.assign(natom, 0);
        h->eps14.assign(natom, 0); h->sig14.assign(natom, 0); h->mass.assign(natom, 1.0);
        h->frozen.assign(natom, 0);
        h->resid.resize(natom);
        for (int i = 0; i < natom; i++) h->resid[i] = i + 1;
    });
    if (rc != MCX_OK) return rc;
    *out = h.release();
    return MCX_OK;
}

void mcx_destroy(mcx_ctx *h)
{
    if (!h) return;
    (void)hipSetDevice(h->device);
    delete h;
}

const char *mcx_last_error(const mcx_ctx *h) { return h ? h->err.c_str() : "null handle"; }

int mcx_set_box(mcx_ctx *h, int pbtype, double a, double b, double c)
{
    return guarded(h, [&] {
        if (pbtype != MCX_PBC_NONE && pbtype != MCX_PBC_CUBIC && pbtype != MCX_PBC_ORBIC)
            throw mcx_error(MCX_ERR_ARG, "unknown pbc type");
        if (pbtype == MCX_PBC_CUBIC) { b = a; c = a; }
        if (pbtype != MCX_PBC_NONE && !(a > 0 && b > 0 && c > 0)) throw mcx_error(MCX_ERR_ARG, "box lengths must be > 0");
        h->pbtype = pbtype;
        h->box[0] = a; h->box[1] = b; h->box[2] = c;
        h->list_valid = false;
        h->ntiles = 0;
        upload_mirror(h);
    });
}

int mcx_set_cut(mcx_ctx *h, int cutmode, double ctoff, double cton, double dcut, double temp)
{
    return guarded(h, [&] {
        if (cutmode != MCX_CUT_FULL && cutmode != MCX_CUT_SWITCH) throw mcx_error(MCX_ERR_ARG, "unknown cut mode");
        if (cutmode == MCX_CUT_SWITCH && !(ctoff > cton && cton >= 0 && dcut >= 0))
            throw mcx_error(MCX_ERR_ARG, "switch needs ctoff > cton >= 0 and dcut >= 0");
        h->cutmode = cutmode; h->ctoff = ctoff; h->cton = cton; h->dcut = dcut; h->temp = temp;
        h->list_valid = false;
    });
}

int mcx_set_atoms(mcx_ctx *h, const double *charge, const double *eps, const double *sig, const double *eps14,
                  const double *sig14, const double *mass, const int *frozen, const int *resid)
{
    return guarded(h, [&] {
        if (!charge || !eps || !sig) throw mcx_error(MCX_ERR_ARG, "charge, eps and sig are required");
        const int n = h->n;
        h->q.assign(charge, charge + n);
        h->eps.assign(eps, eps + n);
        h->sig.assign(sig, sig + n);
        h->eps14.assign(eps14 ? eps14 : eps, (eps14 ? eps14 : eps) + n);
        h->sig14.assign(sig14 ? sig14 : sig, (sig14 ? sig14 : sig) + n);
        if (mass) h->mass.assign(mass, mass + n);
        if (frozen) h->frozen.assign(frozen, frozen + n);
        if (resid) h->resid.assign(resid, resid + n);
        std::vector<float2> e(n), e14(n);
        for (int i = 0; i < n; i++) {
            e[i] = make_float2((float)h->eps[i], (float)h->sig[i]);
            e14[i] = make_float2((float)h->eps14[i], (float)h->sig14[i]);
        }
        h->es.upload(e.data(), n, h->stream);
        h->es14.upload(e14.data(), n, h->stream);
        std::vector<double4> qes(n);
        for (int i = 0; i < n; i++) qes[i] = make_double4(h->q[i], h->eps[i], h->sig[i], 0.0);
        h->qes.upload(qes.data(), n, h->stream);
        h->q64.upload(h->q.data(), n, h->stream);
        h->eps64.upload(h->eps.data(), n, h->stream);
        h->sig64.upload(h->sig.data(), n, h->stream);
        h->eps1464.upload(h->eps14.data(), n, h->stream);
        h->sig1464.upload(h->sig14.data(), n, h->stream);
        h->mass64.upload(h->mass.data(), n, h->stream);
        h->frozen_d.upload(h->frozen.data(), n, h->stream);
        h->have_atoms = true;
        upload_mirror(h);
        HIPCHK(hipStreamSynchronize(h->stream));
    });
}

static void check_idx(const mcx_ctx *h, int n, const int *const *arrs, int narr)
{
    for (int a = 0; a < narr; a++)
        for (int k = 0; k < n; k++)
            if (arrs[a][k] < 0 || arrs[a][k] >= h->n) throw mcx_error(MCX_ERR_ARG, "term atom index out of range");
}

int mcx_set_bonds(mcx_ctx *h, int n, const int *a, const int *b, const int *type, const double *k, const double *r0,
                  const double *beta)
{
    return guarded(h, [&] {
        const int *ar[2] = {a, b};
        if (n > 0) check_idx(h, n, ar, 2);
        auto &T = h->terms;
        T.idx[kBond].clear(); T.type[kBond].clear(); T.par[kBond].clear();
        for (int t = 0; t < n; t++) {
            T.idx[kBond].push_back(make_int4(a[t], b[t], -1, -1));
            T.type[kBond].push_back(type ? type[t] : 0);
            T.par[kBond].push_back(make_double4(k[t], r0[t], beta ? beta[t] : 0.0, 0.0));
        }
        h->have_excl = false;
    });
}

int mcx_set_ub(mcx_ctx *h, int n, const int *a, const int *b, const double *k, const double *r0)
{
    return guarded(h, [&] {
        const int *ar[2] = {a, b};
        if (n > 0) check_idx(h, n, ar, 2);
        auto &T = h->terms;
        T.idx[kUB].clear(); T.type[kUB].clear(); T.par[kUB].clear();
        for (int t = 0; t < n; t++) {
            T.idx[kUB].push_back(make_int4(a[t], b[t], -1, -1));
            T.type[kUB].push_back(0);
            T.par[kUB].push_back(make_double4(k[t], r0[t], 0.0, 0.0));
        }
    });
}

int mcx_set_angles(mcx_ctx *h, int n, const int *a, const int *b, const int *c, const double *k, const double *theta0)
{
    return guarded(h, [&] {
        const int *ar[3] = {a, b, c};
        if (n > 0) check_idx(h, n, ar, 3);
        auto &T = h->terms;
        T.idx[kAngle].clear(); T.type[kAngle].clear(); T.par[kAngle].clear();
        for (int t = 0; t < n; t++) {
            T.idx[kAngle].push_back(make_int4(a[t], b[t], c[t], -1));
            T.type[kAngle].push_back(0);
            T.par[kAngle].push_back(make_double4(k[t], theta0[t], 0.0, 0.0));
        }
        h->have_excl = false;
    });
}

static int set_torsions(mcx_ctx *h, int kind, int n, const int *a, const int *b, const int *c, const int *d,
                        const int *type, const double *k, const double *phi0, const double *mult)
{
    return guarded(h, [&] {
        const int *ar[4] = {a, b, c, d};
        if (n > 0) check_idx(h, n, ar, 4);
        auto &T = h->terms;
        T.idx[kind].clear(); T.type[kind].clear(); T.par[kind].clear();
        for (int t = 0; t < n; t++) {
            T.idx[kind].push_back(make_int4(a[t], b[t], c[t], d[t]));
            T.type[kind].push_back(type ? type[t] : 1);
            T.par[kind].push_back(make_double4(k[t], phi0[t], mult ? mult[t] : 0.0, 0.0));
        }
        h->have_excl = false;
    });
}

int mcx_set_dihedrals(mcx_ctx *h, int n, const int *a, const int *b, const int *c, const int *d, const int *type,
                      const double *k, const double *phi0, const double *mult)
{
    return set_torsions(h, kDihe, n, a, b, c, d, type, k, phi0, mult);
}

int mcx_set_impropers(mcx_ctx *h, int n, const int *a, const int *b, const int *c, const int *d, const int *type,
                      const double *k, const double *phi0, const double *mult)
{
    return set_torsions(h, kImpr, n, a, b, c, d, type, k, phi0, mult);
}

int mcx_build_exclusions(mcx_ctx *h)
{
    return guarded(h, [&] {
        build_exclusions_host(h);
        upload_terms(h);
        HIPCHK(hipStreamSynchronize(h->stream));
        h->have_excl = true;
        h->list_valid = false;
    });
}

int mcx_get_exclusions(mcx_ctx *h, int *exclPair, int *idx, long long cap, long long *nidx, int *pairs14, int cap14,
                       int *npair14)
{
    return guarded(h, [&] {
        if (!h->have_excl) throw mcx_error(MCX_ERR_STATE, "exclusions not built");
        if (exclPair) std::copy(h->excl_pair.begin(), h->excl_pair.end(), exclPair);
        if (nidx) *nidx = (long long)h->excl_idx.size();
        if (idx) for (long long k = 0; k < (long long)h->excl_idx.size() && k < cap; k++) idx[k] = h->excl_idx[k];
        if (npair14) *npair14 = (int)h->pairs14.size() / 2;
        if (pairs14) for (int k = 0; k < (int)h->pairs14.size() && k < cap14; k++) pairs14[k] = h->pairs14[k];
    });
}

int mcx_set_coords(mcx_ctx *h, const double *x, const double *y, const double *z)
{
    return guarded(h, [&] {
        if (!x || !y || !z) throw mcx_error(MCX_ERR_ARG, "null coordinate array");
        upload_coords(h, x, y, z);
    });
}

int mcx_set_coords_moved(mcx_ctx *h, const double *x, const double *y, const double *z)
{
    return guarded(h, [&] {
        if (!x || !y || !z) throw mcx_error(MCX_ERR_ARG, "null coordinate array");
        if (!h->have_coords) {
            upload_coords(h, x, y, z);
            return;
        }
        // one upload of x, y, z packed in pinned memory; the device copy is
        // the master from here on (the host copy is read back when needed)
        const size_t n = (size_t)h->n;
        if (h->stage_ev) HIPCHK(hipEventSynchronize(h->stage_ev));    // the last upload from stage_h is done
        else HIPCHK(hipEventCreateWithFlags(&h->stage_ev, hipEventDisableTiming));
        if (h->stage_cap < 3 * n) {
            if (h->stage_h) HIPCHK(hipHostFree(h->stage_h));
            h->stage_h = nullptr;
            h->stage_cap = 0;
            HIPCHK(hipHostMalloc((void **)&h->stage_h, 3 * n * sizeof(double), hipHostMallocDefault));
            h->stage_cap = 3 * n;
        }
        std::memcpy(h->stage_h, x, n * sizeof(double));
        std::memcpy(h->stage_h + n, y, n * sizeof(double));
        std::memcpy(h->stage_h + 2 * n, z, n * sizeof(double));
        h->xyz_in.alloc(3 * n);
        HIPCHK(hipMemcpyAsync(h->xyz_in.p, h->stage_h, 3 * n * sizeof(double), hipMemcpyHostToDevice, h->stream));
        HIPCHK(hipEventRecord(h->stage_ev, h->stream));
        launch_mirror_from_packed(h->sys(), h->xyz_in.p, h->stream);
        HIPCHK(hipGetLastError());
        h->host_stale = true;
        // the moves since the last list update were not recorded on the device
        h->patch_valid = false;
    });
}

int mcx_get_coords(mcx_ctx *h, double *x, double *y, double *z)
{
    return guarded(h, [&] {
        if (!h->have_coords) throw mcx_error(MCX_ERR_STATE, "no coordinates");
        HIPCHK(hipMemcpyAsync(x, h->x.p, h->n * 8, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipMemcpyAsync(y, h->y.p, h->n * 8, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipMemcpyAsync(z, h->z.p, h->n * 8, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
    });
}

int mcx_update_list(mcx_ctx *h, long long *npairs)
{
    return guarded(h, [&] {
        long long np = 0;
        if (h->cutmode != MCX_CUT_FULL) np = build_list(h);
        if (npairs) *npairs = np;
    });
}

int mcx_list_capacity(mcx_ctx *h, int set, int *current)
{
    return guarded(h, [&] {
        if (set < 0) throw mcx_error(MCX_ERR_ARG, "capacity < 0");
        if (set > 0) {
            h->cap = set;
            h->nbr.alloc((size_t)h->n * set);
            h->cnt.alloc(h->n);
            h->half.alloc(h->n);
            h->overflow.alloc(1);
            h->list_valid = false;
        }
        if (current) *current = h->cap;
    });
}

int mcx_get_list(mcx_ctx *h, int *neighPair, int *flat, long long cap)
{
    return guarded(h, [&] {
        if (!h->list_valid) throw mcx_error(MCX_ERR_STATE, "list not built");
        const int n = h->n;
        std::vector<int> c(n), hf(n), nb((size_t)n * h->cap);
        HIPCHK(hipMemcpy(c.data(), h->cnt.p, sizeof(int) * n, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(hf.data(), h->half.p, sizeof(int) * n, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(nb.data(), h->nbr.p, sizeof(int) * nb.size(), hipMemcpyDeviceToHost));
        long long o = 0;
        for (int i = 0; i < n; i++) {
            if (neighPair) neighPair[i] = c[i] - hf[i];
            for (int k = hf[i]; k < c[i]; k++) {
                if (flat && o < cap) flat[o] = nb[(size_t)i * h->cap + k];
                o++;
            }
        }
    });
}

int mcx_energy(mcx_ctx *h, double ener[10])
{
    return guarded(h, [&] {
        double r[9];
        energy_device(h, r);
        assemble(r, ener);
    });
}

int mcx_delta_energy(mcx_ctx *h, int first, int last, const double *nx, const double *ny, const double *nz, double *de)
{
    return guarded(h, [&] {
        if (first < 0 || last >= h->n || last < first) throw mcx_error(MCX_ERR_ARG, "bad atom range");
        if (last - first + 1 > kMaxMoveAtoms) throw mcx_error(MCX_ERR_UNSUPPORTED, "at most 64 moved atoms");
        if (!h->have_excl || !h->have_coords) throw mcx_error(MCX_ERR_STATE, "system not set up");
        if (h->cutmode != MCX_CUT_FULL && !h->list_valid) build_list(h);
        const int nm = last - first + 1;
        std::vector<double> buf(3 * nm);
        for (int a = 0; a < nm; a++) { buf[3 * a] = nx[a]; buf[3 * a + 1] = ny[a]; buf[3 * a + 2] = nz[a]; }
        h->newxyz.upload(buf.data(), buf.size(), h->stream);
        h->scalar.alloc(4);
        launch_delta_energy(h->sys(), h->dterms(), adjacency_only(h), first, last, h->newxyz.p, h->scalar.p, h->stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(de, h->scalar.p, sizeof(double), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
    });
}

int mcx_mc_run(mcx_ctx *h, const mcx_move_type *types, int ntypes, int nsteps, uint64_t seed, int rng_mode,
               int list_every, mcx_mc_stats *stats, unsigned char *acc_out, double *de_out, int *mvtyp_out,
               int *mvinst_out)
{
    mcx_mc_trace tr{acc_out, de_out, mvtyp_out, mvinst_out, nullptr, nullptr, 0, nullptr, nullptr};
    return mcx_mc_run_traced(h, types, ntypes, nsteps, seed, rng_mode, list_every, stats, &tr);
}

int mcx_mc_run_traced(mcx_ctx *h, const mcx_move_type *types, int ntypes, int nsteps, uint64_t seed, int rng_mode,
                      int list_every, mcx_mc_stats *stats, const mcx_mc_trace *trace)
{
    const mcx_mc_trace none{};
    const mcx_mc_trace &tr = trace ? *trace : none;
    unsigned char *acc_out = tr.acc;
    double *de_out = tr.de, *disp_out = tr.disp, *lnq_out = tr.lnq;
    int *mvtyp_out = tr.mvtyp, *mvinst_out = tr.mvinst;
    return guarded(h, [&] {
        if (nsteps < 0) throw mcx_error(MCX_ERR_ARG, "nsteps < 0");
        if (!h->have_excl || !h->have_coords || !h->have_atoms) throw mcx_error(MCX_ERR_STATE, "system not set up");
        const bool autolist = list_every == MCX_LIST_AUTO && h->cutmode != MCX_CUT_FULL;
        if (list_every < 0 && list_every != MCX_LIST_AUTO) throw mcx_error(MCX_ERR_ARG, "list_every < 0");
        if (autolist && h->pbtype == MCX_PBC_NONE)
            throw mcx_error(MCX_ERR_UNSUPPORTED, "displacement-triggered lists need a periodic box");
        // scheduled rebuilds as incremental updates (the same list as a full
        // rebuild) unless MCX_LIST_PATCH=0; they need every move since the
        // last full build recorded, so a run that does not record leaves the
        // next one a full build
        const char *lp = std::getenv("MCX_LIST_PATCH");
        const bool use_patch = list_every > 0 && h->cutmode != MCX_CUT_FULL && !(lp && lp[0] == '0') &&
                               h->list_valid && list_patch_supported(h->sys());
        if (use_patch) {
            ensure_patch(h);
            if (!h->patch_valid) build_list(h);
        } else {
            h->patch_valid = false;
        }
        if (h->cutmode != MCX_CUT_FULL && !h->list_valid) build_list(h);
        DevMoves mv = build_moves(h, types, ntypes, h->mvb, kMaxMoveAtoms);
        // device-side run state
        DBuf<double> etot, dmax;
        DBuf<uint32_t> mts;
        DBuf<long long> trials, accepted;
        DBuf<int> acc_tmp;
        DBuf<unsigned char> acc_d;
        DBuf<double> de_d;
        DBuf<int> typ_d, inst_d;
        DBuf<double> disp_d, lnq_d;
        DBuf<int> stop_d, bail_d;
        struct Cleanup {
            std::vector<std::function<void()>> f;
            ~Cleanup() { for (auto &g : f) g(); }
        } cleanup;
        cleanup.f.push_back([&] { etot.release(); dmax.release(); mts.release(); trials.release(); accepted.release();
                                  acc_tmp.release(); acc_d.release(); de_d.release(); typ_d.release(); inst_d.release();
                                  disp_d.release(); lnq_d.release(); stop_d.release(); bail_d.release(); });
        double ener[10], r9[9];
        energy_device(h, r9);
        assemble(r9, ener);
        etot.upload(&ener[0], 1, h->stream);
        std::vector<double> dm(ntypes);
        for (int t = 0; t < ntypes; t++) dm[t] = types[t].dmax;
        dmax.upload(dm.data(), ntypes, h->stream);
        std::vector<uint32_t> st(kRngWordsWide, 0);
        mt_seed(st.data(), seed);
        st[625] = 0; st[626] = 0;
        st[627] = 32;     // empty draw batch: the first draw refills it
        mts.upload(st.data(), st.size(), h->stream);
        std::vector<long long> z8(ntypes, 0);
        trials.upload(z8.data(), ntypes, h->stream);
        accepted.upload(z8.data(), ntypes, h->stream);
        std::vector<int> zi(ntypes, 0);
        acc_tmp.upload(zi.data(), ntypes, h->stream);
        if (acc_out) acc_d.alloc(std::max(nsteps, 1));
        if (de_out) de_d.alloc(std::max(nsteps, 1));
        if (mvtyp_out) typ_d.alloc(std::max(nsteps, 1));
        if (mvinst_out) inst_d.alloc(std::max(nsteps, 1));
        if (disp_out) disp_d.alloc((size_t)3 * std::max(nsteps, 1));
        if (lnq_out) lnq_d.alloc(std::max(nsteps, 1));
        MCRunArgs a{};
        a.rng_mode = rng_mode == MCX_MC_REPLAY ? 1 : 0;
        a.seed = seed;
        a.stream = h->rng_stream;
        a.beta = 1.0 / ((MCX_RBOLTZUI / MCX_KCALTOIU) * h->temp);
        // speculative rounds (several proposals evaluated together; the same
        // chain) unless MCX_SPECULATE=0; move sets with force-bias types run
        // them on the wide kernel only (one workgroup per proposal: its force
        // pass precedes its trial), else one proposal at a time
        const char *sp = std::getenv("MCX_SPECULATE");
        a.spec = !(sp && sp[0] == '0');
        if (use_patch) {
            a.patch = h->patch();
            a.record_moves = 1;
        }
        bool has_fb = false;
        for (int k = 0; k < ntypes; k++)
            if (types[k].type == MCX_MV_FBIAS || types[k].type == MCX_MV_FBIAS_ACC) has_fb = true;
        // speculative rounds over several workgroups (k_mc_wide) for the
        // same move sets; MCX_WIDE=<n> sets the workgroup count, 0 or 1
        // keeps the single-workgroup read-ahead kernel
        // MCX_WIDE=<n> proposals per round, MCX_WIDE_SPLIT=<k> workgroups per
        // proposal (its pair wavefronts shared out between them)
        int nwg = 0, split = 1;
        if (a.spec) {
            const char *wv = std::getenv("MCX_WIDE");
            const char *sv = std::getenv("MCX_WIDE_SPLIT");
            const int want = wv ? std::atoi(wv) : kDefaultWide;
            split = has_fb ? 1 : std::max(1, std::min(sv ? std::atoi(sv) : kDefaultSplit, 15));
            const int cap = std::min(kMaxWide, mc_wide_max_workgroups(h->device)) / split;
            const int W = std::min(want, cap);
            if (want > 1 && W >= 2) nwg = W * split;
        }
        if (has_fb && !nwg) a.spec = 0;
        DBuf<unsigned> sync_d, exit_d, arrive_d;
        DBuf<int> werr_d, go_d;
        DBuf<double> bred_d, bxyz_d, blnq_d;
        DBuf<int4> bmeta_d, bip_d;
        cleanup.f.push_back([&] { sync_d.release(); exit_d.release(); arrive_d.release(); go_d.release();
                                  werr_d.release(); bred_d.release(); bxyz_d.release(); blnq_d.release();
                                  bmeta_d.release(); bip_d.release(); });
        WideArgs wa{};
        if (nwg) {
            const unsigned uzero = 0u;
            sync_d.upload(&uzero, 1, h->stream);
            exit_d.upload(&uzero, 1, h->stream);
            arrive_d.upload(&uzero, 1, h->stream);
            const int zero = 0;
            werr_d.upload(&zero, 1, h->stream);
            go_d.upload(&zero, 1, h->stream);
            bred_d.alloc((size_t)kMaxWide * 16 * 3);
            bmeta_d.alloc(kMaxWide);
            bxyz_d.alloc((size_t)kMaxWide * kMaxMoveAtoms * 3);
            bip_d.alloc((size_t)kMaxWide * kMaxMoveAtoms);
            blnq_d.alloc(kMaxWide);
            int rate_khz = 0;
            HIPCHK(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, h->device));
            wa.nwg = nwg;
            wa.sync = sync_d.p;
            wa.exit_ctr = exit_d.p;
            wa.err = werr_d.p;
            wa.spin_ticks = (long long)std::max(rate_khz, 1000) * 500;    // half a second
            wa.rdv_ticks = (long long)std::max(rate_khz, 1000) * 20;      // 20 ms
            wa.arrive = arrive_d.p;
            wa.go = go_d.p;
            wa.split = split;
            wa.bred = bred_d.p; wa.bmeta = bmeta_d.p; wa.bxyz = bxyz_d.p; wa.bip = bip_d.p; wa.blnq = blnq_d.p;
        }
        // MCX_KERNEL_EVENTS=k (k >= 1): every k-th chunk and every k-th
        // in-run list update between a pair of HIP events (pairs from the
        // handle's pool), summed at the end
        const char *kev = std::getenv("MCX_KERNEL_EVENTS");
        const int ev_every = kev ? std::max(0, std::atoi(kev)) : 0;
        std::vector<std::pair<int, int>> ev_used;     // (pool index of the first event, 0 chain / 1 list)
        long long ev_seen[2] = {0, 0};
        auto ev_begin = [&](int what) -> int {
            if (!ev_every || ev_seen[what]++ % ev_every != 0) return -1;
            const size_t k = 2 * ev_used.size();
            while (h->ev_pool.size() < k + 2) {
                hipEvent_t e;
                HIPCHK(hipEventCreate(&e));
                h->ev_pool.push_back(e);
            }
            ev_used.emplace_back((int)k, what);
            HIPCHK(hipEventRecord(h->ev_pool[k], h->stream));
            return (int)k;
        };
        auto ev_end = [&](int k) {
            if (k >= 0) HIPCHK(hipEventRecord(h->ev_pool[k + 1], h->stream));
        };
        auto launch_chunk = [&]() {
            const int ev = ev_begin(0);
            if (nwg) {
                HIPCHK(launch_mc_wide(h->sys(), h->dterms(), mv, a, wa, h->stream));
            } else {
                launch_mc_chunk(h->sys(), h->dterms(), mv, a, h->stream);
            }
            ev_end(ev);
        };
        // true: a launch found its workgroups not resident together and
        // changed nothing (it set the bail-out step); it is run again
        int wide_retries = 0;
        auto check_wide = [&]() -> bool {
            if (!nwg) return false;
            int e = 0;
            HIPCHK(hipMemcpyAsync(&e, werr_d.p, sizeof(int), hipMemcpyDeviceToHost, h->stream));
            HIPCHK(hipStreamSynchronize(h->stream));
            if (e == 3) {
                if (++wide_retries > 64)
                    throw mcx_error(MCX_ERR_HIP, "the speculative kernel's workgroups were never resident together "
                                                 "(a busy device; MCX_WIDE=0 avoids it)");
                const int zero = 0;
                werr_d.upload(&zero, 1, h->stream);
                return true;
            }
            if (e == 2) throw mcx_error(MCX_ERR_HIP, "a proposal needed more draws than the speculative kernel holds");
            if (e) throw mcx_error(MCX_ERR_HIP, "a speculative round barrier gave up waiting (MCX_WIDE=0 avoids it)");
            return false;
        };
        a.etot = etot.p; a.mt_state = mts.p; a.dmax = dmax.p; a.trials = trials.p; a.accepted = accepted.p;
        a.acc_tmp = acc_tmp.p;
        a.acc_out = acc_out ? acc_d.p : nullptr;
        a.de_out = de_out ? de_d.p : nullptr;
        a.mvtyp_out = mvtyp_out ? typ_d.p : nullptr;
        a.mvinst_out = mvinst_out ? inst_d.p : nullptr;
        a.disp_out = disp_out ? disp_d.p : nullptr;
        a.lnq_out = lnq_out ? lnq_d.p : nullptr;
        if (autolist) {
            stop_d.alloc(2);
            a.ipos_ref = h->ipos_ref.p;
            a.half_skin2 = 0.25 * h->dcut * h->dcut;
            a.stop_at = stop_d.p;
        }
        const int save = (tr.save_each > 0 && tr.on_frame) ? tr.save_each : 0;
        std::vector<double> fx, fy, fz;
        auto emit_frame = [&](int step) {
            double e = 0.0;
            const size_t n = (size_t)h->n;
            fx.resize(n); fy.resize(n); fz.resize(n);
            HIPCHK(hipMemcpyAsync(&e, etot.p, sizeof(double), hipMemcpyDeviceToHost, h->stream));
            HIPCHK(hipMemcpyAsync(fx.data(), h->x.p, n * 8, hipMemcpyDeviceToHost, h->stream));
            HIPCHK(hipMemcpyAsync(fy.data(), h->y.p, n * 8, hipMemcpyDeviceToHost, h->stream));
            HIPCHK(hipMemcpyAsync(fz.data(), h->z.p, n * 8, hipMemcpyDeviceToHost, h->stream));
            HIPCHK(hipStreamSynchronize(h->stream));
            tr.on_frame(tr.user, step, e, fx.data(), fy.data(), fz.data());
        };
        HIPCHK(hipStreamSynchronize(h->stream));
        const auto t0 = std::chrono::steady_clock::now();
        long long nlist = 0;
        int done = 0;
        // displacement-triggered list: launches of up to kAutoChunk moves,
        // each ending early after an accepted move that carried an atom
        // beyond half the skin; the list is then rebuilt
        while (autolist && done < nsteps) {
            constexpr int kAutoChunk = 4096;
            a.nsteps = std::min(kAutoChunk, nsteps - done);
            if (save) a.nsteps = std::min(a.nsteps, save - done % save);
            a.step0 = done;
            launch_chunk();
            HIPCHK(hipGetLastError());
            if (check_wide()) continue;
            int st[2] = {0, 0};
            HIPCHK(hipMemcpyAsync(st, stop_d.p, sizeof(st), hipMemcpyDeviceToHost, h->stream));
            HIPCHK(hipStreamSynchronize(h->stream));
            if (st[0] < 1 && !st[1]) throw mcx_error(MCX_ERR_HIP, "Metropolis chunk made no progress");
            done += st[0];
            if (st[1]) {
                rebuild_list_fast(h);
                nlist++;
            }
            if (save && done % save == 0) emit_frame(done);
        }
        const bool rebuilds = list_every > 0 && h->cutmode != MCX_CUT_FULL;
        const int every = rebuilds ? list_every : nsteps;
        // Chunks and list rebuilds are queued without waiting on the host.
        // A rebuild that overflows the list capacity makes the following
        // chunks return untouched (the bail-out of k_mc_chunk); every
        // kCheckEvery chunks, and before a frame or the end, the host looks:
        // it then grows the list, rebuilds it and resumes at the first chunk
        // that bailed out.
        constexpr int kCheckEvery = 16;
        const int kNone = INT_MAX;
        // the bail-out record also lets a wide launch that found its
        // workgroups not resident together hand its chunk back
        if (!autolist && (rebuilds || nwg)) {
            bail_d.upload(&kNone, 1, h->stream);
            a.bail = bail_d.p;
        }
        int since_check = 0;
        // the incremental updates test every atom moved since the last full
        // build from an explicit list, which a full build empties; when a
        // check finds it past half its limit, the next scheduled update is a
        // full build (before the limit makes a patch give up and the chunks
        // after it bail out)
        bool full_next = false;
        auto check_bail = [&]() {
            // a launch whose workgroups were not resident together changed
            // nothing and only handed its chunk back: no rebuild for it (a
            // list overflow still shows in the flag and bails the relaunch)
            const bool retry = check_wide();
            if (!a.bail) return;
            int b = kNone, nst = 0;
            HIPCHK(hipMemcpyAsync(&b, bail_d.p, sizeof(int), hipMemcpyDeviceToHost, h->stream));
            if (use_patch)
                HIPCHK(hipMemcpyAsync(&nst, h->patch().stale_cnt, sizeof(int), hipMemcpyDeviceToHost, h->stream));
            HIPCHK(hipStreamSynchronize(h->stream));
            if (nst > kPatchStale / 2) full_next = true;
            if (b != kNone) {
                done = b;
                if (rebuilds && !retry) {
                    build_list(h);           // grows the capacity until the rows fit
                    nlist++;
                }
                bail_d.upload(&kNone, 1, h->stream);
            }
            since_check = 0;
        };
        while (!autolist && done < nsteps) {
            int chunk = every - (done % std::max(every, 1));
            if (save) chunk = std::min(chunk, save - done % save);
            if (chunk > nsteps - done) chunk = nsteps - done;
            a.nsteps = chunk;
            a.step0 = done;
            launch_chunk();
            HIPCHK(hipGetLastError());
            done += chunk;
            if (rebuilds && done % list_every == 0) {
                const int ev = ev_begin(1);
                if (use_patch && !full_next) {
                    h->p_newrow.alloc((size_t)std::min(h->n, kPatchRows) * h->cap);   // grows with the capacity
                    launch_list_patch(h->sys(), h->grid, h->patch(), h->stream);
                } else {
                    launch_list_build(h->sys(), h->grid, h->stream);
                    if (use_patch) launch_list_patch_reset(h->sys(), h->patch(), h->stream);
                    full_next = false;
                }
                HIPCHK(hipGetLastError());
                ev_end(ev);
                nlist++;
            }
            const bool frame = save && done % save == 0;
            const int reached = done;
            if (frame || ++since_check == kCheckEvery || done >= nsteps) check_bail();
            // a rewind (done < reached) resumes at a step whose frame, if
            // any, was written when it was first reached
            if (frame && done == reached) emit_frame(done);
        }
        if (rebuilds && done >= nsteps) {
            // a rebuild queued after the last chunk has no chunk after it to
            // notice an overflow: grow the list here, so that later energy and
            // gradient calls never see truncated rows
            int ov = 0;
            HIPCHK(hipMemcpyAsync(&ov, h->overflow.p, sizeof(int), hipMemcpyDeviceToHost, h->stream));
            HIPCHK(hipStreamSynchronize(h->stream));
            if (ov) {
                build_list(h);
                nlist++;
            }
        }
        HIPCHK(hipStreamSynchronize(h->stream));
        check_wide();
        const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (stats) {
            std::memset(stats, 0, sizeof(*stats));
            HIPCHK(hipMemcpy(&stats->etot, etot.p, sizeof(double), hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(stats->trials, trials.p, ntypes * sizeof(long long), hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(stats->accepted, accepted.p, ntypes * sizeof(long long), hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(stats->dmax, dmax.p, ntypes * sizeof(double), hipMemcpyDeviceToHost));
            stats->list_updates = nlist;
            stats->seconds = secs;
            for (const auto &u : ev_used) {
                float ms = 0.f;
                HIPCHK(hipEventElapsedTime(&ms, h->ev_pool[u.first], h->ev_pool[u.first + 1]));
                if (u.second == 0) { stats->chain_kernel_ms += ms; stats->chain_launches++; }
                else { stats->list_kernel_ms += ms; stats->list_launches++; }
            }
        }
        sync_host_coords(h);
        if (acc_out && nsteps) HIPCHK(hipMemcpy(acc_out, acc_d.p, nsteps, hipMemcpyDeviceToHost));
        if (de_out && nsteps) HIPCHK(hipMemcpy(de_out, de_d.p, nsteps * sizeof(double), hipMemcpyDeviceToHost));
        if (mvtyp_out && nsteps) HIPCHK(hipMemcpy(mvtyp_out, typ_d.p, nsteps * sizeof(int), hipMemcpyDeviceToHost));
        if (mvinst_out && nsteps) HIPCHK(hipMemcpy(mvinst_out, inst_d.p, nsteps * sizeof(int), hipMemcpyDeviceToHost));
        if (disp_out && nsteps)
            HIPCHK(hipMemcpy(disp_out, disp_d.p, (size_t)3 * nsteps * sizeof(double), hipMemcpyDeviceToHost));
        if (lnq_out && nsteps) HIPCHK(hipMemcpy(lnq_out, lnq_d.p, nsteps * sizeof(double), hipMemcpyDeviceToHost));
    });
}

int mcx_set_rng_stream(mcx_ctx *h, uint32_t stream)
{
    return guarded(h, [&] { h->rng_stream = stream; });
}

namespace {

// The adjoint of elec + vdw into the device buffer g8 (8 n values), on the
// handle's stream: SWITCH mode over the list (built if stale), FULL mode
// over all pairs; 1-4 pairs in both.
void grad_device(mcx_ctx *h, double *g8)
{
    if (!h->have_excl || !h->have_coords || !h->have_atoms) throw mcx_error(MCX_ERR_STATE, "system not set up");
    if (h->cutmode != MCX_CUT_FULL && !h->list_valid) build_list(h);
    launch_nonbonded_grad(h->sys(), g8, h->stream);
    HIPCHK(hipGetLastError());
}

void check_types(const mcx_ctx *h, int ntypes, const int *type)
{
    if (ntypes < 1 || !type) throw mcx_error(MCX_ERR_ARG, "atom types required");
    for (int i = 0; i < h->n; i++)
        if (type[i] < 0 || type[i] >= ntypes) throw mcx_error(MCX_ERR_ARG, "atom type out of range");
}

}  // namespace

int mcx_nonbonded_grad(mcx_ctx *h, double *gx, double *gy, double *gz, double *geps, double *gsig, double *gq)
{
    return guarded(h, [&] {
        const int n = h->n;
        DBuf<double> g;
        struct Release {
            DBuf<double> &a;
            ~Release() { a.release(); }
        } rel{g};
        g.alloc((size_t)8 * n);
        grad_device(h, g.p);
        std::vector<double> hg((size_t)6 * n);
        HIPCHK(hipMemcpyAsync(hg.data(), g.p, hg.size() * 8, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        double *outs[6] = {gx, gy, gz, geps, gsig, gq};
        for (int c = 0; c < 6; c++)
            if (outs[c]) std::copy(hg.begin() + (size_t)c * n, hg.begin() + (size_t)(c + 1) * n, outs[c]);
    });
}

int mcx_nonbonded_grad_types(mcx_ctx *h, int ntypes, const int *type, double *gx, double *gy, double *gz,
                             double *geps, double *gsig, double *gq)
{
    return guarded(h, [&] {
        check_types(h, ntypes, type);
        const int n = h->n;
        DBuf<double> g, out;
        DBuf<int> ty;
        struct Release {
            DBuf<double> &a, &b;
            DBuf<int> &c;
            ~Release() { a.release(); b.release(); c.release(); }
        } rel{g, out, ty};
        g.alloc((size_t)8 * n);
        out.alloc((size_t)5 * ntypes);
        ty.upload(type, n, h->stream);
        grad_device(h, g.p);
        launch_type_sums(g.p, n, ty.p, ntypes, out.p, h->stream);
        HIPCHK(hipGetLastError());
        std::vector<double> hg((size_t)3 * n), ho((size_t)5 * ntypes);
        HIPCHK(hipMemcpyAsync(hg.data(), g.p, hg.size() * 8, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipMemcpyAsync(ho.data(), out.p, ho.size() * 8, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        double *pos[3] = {gx, gy, gz};
        for (int c = 0; c < 3; c++)
            if (pos[c]) std::copy(hg.begin() + (size_t)c * n, hg.begin() + (size_t)(c + 1) * n, pos[c]);
        for (int t = 0; t < ntypes; t++) {
            if (geps) geps[t] = ho[5 * t];
            if (gsig) gsig[t] = ho[5 * t + 1];
            if (gq) gq[t] = ho[5 * t + 2];
        }
    });
}

int mcx_coords_updated(mcx_ctx *h)
{
    return guarded(h, [&] {
        if (!h->have_coords) throw mcx_error(MCX_ERR_STATE, "no coordinates");
        launch_mirror_from_coords(h->sys(), h->stream);
        HIPCHK(hipGetLastError());
        h->list_valid = false;
        h->patch_valid = false;
        h->host_stale = true;
        HIPCHK(hipStreamSynchronize(h->stream));
    });
}

int mcx_nonbonded_grad_device(mcx_ctx *h, double *gpos, int ntypes, const int *type, double *gtype)
{
    return guarded(h, [&] {
        if (!gpos) throw mcx_error(MCX_ERR_ARG, "null gradient buffer");
        if (ntypes > 0) {
            check_types(h, ntypes, type);
            if (!gtype) throw mcx_error(MCX_ERR_ARG, "null per-type gradient buffer");
        }
        const int n = h->n;
        DBuf<double> g;
        DBuf<int> ty;
        struct Release {
            DBuf<double> &a;
            DBuf<int> &b;
            ~Release() { a.release(); b.release(); }
        } rel{g, ty};
        g.alloc((size_t)8 * n);
        grad_device(h, g.p);
        HIPCHK(hipMemcpyAsync(gpos, g.p, (size_t)3 * n * 8, hipMemcpyDeviceToDevice, h->stream));
        if (ntypes > 0) {
            ty.upload(type, n, h->stream);
            launch_type_sums(g.p, n, ty.p, ntypes, gtype, h->stream);
            HIPCHK(hipGetLastError());
        }
        HIPCHK(hipStreamSynchronize(h->stream));
    });
}

int mcx_nonbonded_grad_batch(mcx_ctx *const *hs, int nsys, double *const *g6)
{
    if (!hs || nsys < 1 || !g6) return MCX_ERR_ARG;
    for (int k = 0; k < nsys; k++)
        if (!hs[k] || !g6[k]) return MCX_ERR_ARG;
    mcx_ctx *h0 = hs[0];
    return guarded(h0, [&] {
        std::vector<DevSys> d(nsys);
        std::vector<DBuf<double>> gb(nsys);
        DBuf<DevSys> dd;
        DBuf<double *> gp;
        struct Release {
            std::vector<DBuf<double>> &a;
            DBuf<DevSys> &b;
            DBuf<double *> &c;
            ~Release() { for (auto &x : a) x.release(); b.release(); c.release(); }
        } rel{gb, dd, gp};
        int nmax = 0;
        std::vector<double *> ptrs(nsys);
        for (int k = 0; k < nsys; k++) {
            mcx_ctx *h = hs[k];
            if (h->device != h0->device) throw mcx_error(MCX_ERR_ARG, "systems on different devices");
            if (h->cutmode != MCX_CUT_SWITCH) throw mcx_error(MCX_ERR_UNSUPPORTED, "batched gradients need SWITCH mode");
            if (!h->have_excl || !h->have_coords || !h->have_atoms) throw mcx_error(MCX_ERR_STATE, "system not set up");
            if (!h->list_valid) build_list(h);
            HIPCHK(hipStreamSynchronize(h->stream));
            d[k] = h->sys();
            gb[k].alloc((size_t)8 * h->n);
            ptrs[k] = gb[k].p;
            nmax = std::max(nmax, h->n);
        }
        dd.upload(d.data(), nsys, h0->stream);
        gp.upload(ptrs.data(), nsys, h0->stream);
        launch_nonbonded_grad_batch(dd.p, gp.p, nsys, nmax, h0->stream);
        HIPCHK(hipGetLastError());
        for (int k = 0; k < nsys; k++)
            HIPCHK(hipMemcpyAsync(g6[k], gb[k].p, (size_t)6 * hs[k]->n * 8, hipMemcpyDeviceToHost, h0->stream));
        HIPCHK(hipStreamSynchronize(h0->stream));
    });
}

int mcx_device_coords(mcx_ctx *h, int which, void **ptr)
{
    return guarded(h, [&] {
        if (!ptr || which < 0 || which > 2) throw mcx_error(MCX_ERR_ARG, "which must be 0..2");
        *ptr = which == 0 ? (void *)h->x.p : which == 1 ? (void *)h->y.p : (void *)h->z.p;
    });
}

int mcx_time_kernel(mcx_ctx *h, int which, int reps, double *ms)
{
    return guarded(h, [&] {
        if (reps < 1) throw mcx_error(MCX_ERR_ARG, "reps < 1");
        DevSys s = h->sys();
        DevTerms t = h->dterms();
        if (which == 1 || which == 2 || which == 4) { if (!h->list_valid) build_list(h); s = h->sys(); }
        if (which == 4 && h->cutmode != MCX_CUT_SWITCH) throw mcx_error(MCX_ERR_UNSUPPORTED, "gradients need SWITCH mode");
        if (which == 0 && h->ntiles == 0) setup_tiles(h);
        const int nbs = switch_energy_blocks(h->n);
        h->partial.alloc((size_t)std::max({h->ntiles * 2, nbs * 2, bonded_blocks(t) * 7, 16}));
        DBuf<double> g6;
        struct Release {
            DBuf<double> &a;
            ~Release() { a.release(); }
        } rel{g6};
        if (which == 4) g6.alloc((size_t)8 * h->n);
        auto launch = [&] {
            switch (which) {
            case 0: launch_full_energy(s, h->tiles.p, h->ntiles, kFullTile, h->partial.p, h->stream); break;
            case 1: launch_switch_energy(s, h->partial.p, nbs, h->stream); break;
            case 2: launch_list_build(s, h->grid, h->stream); h->patch_valid = false; break;
            case 3: launch_bonded(s, t, h->partial.p, bonded_blocks(t), h->stream); break;
            case 4: launch_nonbonded_grad(s, g6.p, h->stream); break;
            default: throw mcx_error(MCX_ERR_ARG, "unknown kernel id");
            }
        };
        launch();   // warm-up
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventRecord(e0, h->stream));
        for (int r = 0; r < reps; r++) launch();
        HIPCHK(hipEventRecord(e1, h->stream));
        HIPCHK(hipEventSynchronize(e1));
        float t_ms = 0.f;
        HIPCHK(hipEventElapsedTime(&t_ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *ms = t_ms / reps;
    });
}

}  // extern "C"
