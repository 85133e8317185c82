"""mcx_energy_noalloc against mcx_energy, through the C ABI.

The no-allocation entry point runs the same kernels with its partial sums
in a static device array, so its energies must equal mcx_energy's bit for
bit.  Two calls in a row on one GPU, the second on a smaller system (the
array keeps the first call's contents), in both cutoff modes; sizes beyond
the array are refused on the host with MCX_ERR_ARG, and a stale list with
MCX_ERR_STATE.
"""
import numpy as np
import pytest

from mc_cpp_amd import mcx, systems

pytestmark = pytest.mark.gpu


def _pair_of_systems(cut):
    big = systems.bonded_mix(seed=7)[0]                      # 1,350 atoms, orthorhombic, every bonded kind
    small = systems.water_box(seed=5, nside=6, L=18.64)[0]    # 648 atoms, cubic
    for s in (big, small):
        s.cutmode = cut
    return big, small


@pytest.mark.parametrize("cut", [systems.CUT_SWITCH, systems.CUT_FULL])
def test_noalloc_energy_equals_energy(cut):
    big, small = _pair_of_systems(cut)
    for s in (big, small):
        h = mcx.from_system(s)
        if cut == systems.CUT_SWITCH:
            h.update_list()
        e_na = h.energy_noalloc()
        e = h.energy()
        print("natom", s.natom, "cut", cut, "noalloc", e_na[0], "energy", e[0], flush=True)
        assert np.array_equal(e_na, e), (e_na, e)
        assert np.array_equal(h.energy_noalloc(), e)
        h.close()


def test_noalloc_refusals():
    s = systems.water_box(seed=5, nside=6, L=18.64)[0]
    h = mcx.from_system(s)                             # no list built yet
    with pytest.raises(mcx.MCXError) as ei:
        h.energy_noalloc()
    assert ei.value.code == 5                          # MCX_ERR_STATE
    h.close()
    # all-pair mode beyond 46,272 atoms: refused before any launch
    n = 46273
    big = systems.System(natom=n, box=[400.0] * 3, cutmode=systems.CUT_FULL,
                         x=np.linspace(-199.0, 199.0, n), y=np.zeros(n), z=np.zeros(n))
    hb = mcx.from_system(big)
    with pytest.raises(mcx.MCXError) as ei:
        hb.energy_noalloc()
    assert ei.value.code == 1                          # MCX_ERR_ARG
    hb.close()
