// noalloc.hip -- mcx_energy_noalloc: the energy of mcx_energy computed with
// no device allocation, for callers that cannot allocate (code under HIP
// graph capture, serving loops, harnesses that forbid hipMalloc).
//
// mcx_energy keeps its per-block partial sums and the reduced components in
// buffers the handle allocates on first use (and a tile list in all-pair
// mode).  Here both live in one file-scope device array, `scratch`, of
// kScratchDoubles doubles (4 MiB), which every GPU the library runs on holds
// until the process ends; the all-pair kernel computes each block's tile
// instead of reading a list.  The kernels are the ones mcx_energy launches,
// in the same order with the same block counts, so the two entry points give
// the same bits.
//
// The array's address on each device is taken once, with that device
// current, by noalloc_init_device (called by mcx_create) and kept per
// device; host code never uses the array's name as an address.  Every
// partial slot a call reads was written by the same call, so nothing needs
// clearing between calls.  Calls on one GPU share its array: two that
// overlap (handles on different streams or threads) race and must be
// serialised by the caller.
#include "mcx_ctx.h"

#include <mutex>

namespace {

constexpr size_t kScratchDoubles = (size_t)1 << 19;   // 4 MiB
constexpr int kMaxDevices = 64;
constexpr int kResultSlots = 16;                       // reduced components at the end of the array

static __device__ __align__(256) double scratch[kScratchDoubles];

std::mutex g_mu;
double *g_scratch[kMaxDevices] = {};

int noalloc_fail(mcx_ctx *h, int code, const char *msg)
{
    h->err = msg;
    return code;
}

}  // namespace

// With `device` current: looks up this device's copy of `scratch` once.
int noalloc_init_device(int device)
{
    if (device < 0 || device >= kMaxDevices) return MCX_ERR_ARG;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch[device]) return MCX_OK;
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return MCX_ERR_HIP;
    g_scratch[device] = static_cast<double *>(p);
    return MCX_OK;
}

extern "C" int mcx_energy_noalloc(mcx_ctx *h, double ener[10])
{
    if (!h || !ener) return MCX_ERR_ARG;
    if (hipSetDevice(h->device) != hipSuccess) return noalloc_fail(h, MCX_ERR_HIP, "hipSetDevice failed");
    double *scr = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        scr = h->device < kMaxDevices ? g_scratch[h->device] : nullptr;
    }
    if (!scr) return noalloc_fail(h, MCX_ERR_STATE, "scratch array of this device not set up (mcx_create does it)");
    if (!h->have_coords || !h->have_atoms || !h->have_excl)
        return noalloc_fail(h, MCX_ERR_STATE, "atoms, coordinates and exclusions must be set before the energy");
    if (h->cutmode != MCX_CUT_FULL && !h->list_valid)
        return noalloc_fail(h, MCX_ERR_STATE,
                            "the pair list is stale: call mcx_update_list first (a rebuild may grow the list)");
    const DevSys s = h->sys();
    const DevTerms t = h->dterms();
    const int nbb = bonded_blocks(t);
    size_t nb_nb;
    if (h->cutmode == MCX_CUT_FULL) {
        const size_t nt = ((size_t)h->n + kFullTile - 1) / kFullTile;
        nb_nb = nt * (nt + 1) / 2;
    } else {
        nb_nb = (size_t)switch_energy_blocks(h->n);
    }
    const size_t need = std::max(nb_nb * 2, (size_t)nbb * 7) + kResultSlots;
    if (need > kScratchDoubles)
        return noalloc_fail(h, MCX_ERR_ARG,
                            h->cutmode == MCX_CUT_FULL
                                ? "mcx_energy_noalloc: all-pair mode supports at most 46,272 atoms (4 MiB scratch)"
                                : "mcx_energy_noalloc: the system needs more than its 4 MiB scratch");
    double *partial = scr;
    double *result = scr + (kScratchDoubles - kResultSlots);
    if (h->cutmode == MCX_CUT_FULL) launch_full_energy_listless(s, partial, h->stream);
    else launch_switch_energy(s, partial, (int)nb_nb, h->stream);
    launch_reduce(partial, (int)nb_nb, 2, result, h->stream);
    launch_bonded(s, t, partial, nbb, h->stream);
    launch_reduce(partial, nbb, 7, result + 2, h->stream);
    double r9[9];
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(r9, result, 9 * sizeof(double), hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess)
        return noalloc_fail(h, MCX_ERR_HIP, "mcx_energy_noalloc: HIP launch or copy failed");
    mcxi::assemble(r9, ener);
    return MCX_OK;
}
