"""The host cell-list neighbour generator (the device builder's and every workload's reference) against
a brute-force O(N^2) search under the minimum image.

The generator's order is the callers' convention get_soap's arguments assume (src/soap_turbo.f90:139-169):
per site its own atom first (r = 0), then every other atom within the cutoff in increasing index. The
brute force restates the same FP64 operations (r_j - r_i wrapped by round-half-even images, then
x^2 + y^2 + z^2 < rcut^2), so pair lists and vectors must agree exactly.
"""
import numpy as np
import pytest

from soap_turbo_amd.workload import CONFIGS, build_neighbours, place_atoms, strided_sites


def brute_force(pos, L, rcut, site_atom):
    n_neigh, nbr, vec = [], [], []
    for i in site_atom:
        d = pos - pos[i]
        d = d - L * np.rint(d / L)
        r2 = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]
        inside = np.nonzero(r2 < rcut * rcut)[0]
        inside = inside[inside != i]
        n_neigh.append(1 + len(inside))
        nbr.extend([i, *inside])
        vec.append(np.zeros((1, 3)))
        vec.append(d[inside])
    return np.array(n_neigh, np.int32), np.array(nbr, np.int32), np.concatenate(vec)


@pytest.mark.parametrize("cfg,n_sites", [(1, None), (2, None), (3, 300), (5, 200)])
def test_cell_list_equals_brute_force(cfg, n_sites):
    spec, par = CONFIGS[cfg]
    L, pos, sp = place_atoms(spec)
    site_atom = np.arange(spec.n_atoms) if spec.central is None else np.nonzero(sp == spec.central)[0]
    if n_sites is not None:
        site_atom = site_atom[strided_sites(len(site_atom), n_sites, 3)]
    rcut = float(np.max(par.per_species()["rcut_hard"]))
    got = build_neighbours(pos, L, rcut, site_atom)
    want = brute_force(pos, L, rcut, site_atom)
    print(f"config {cfg}: {len(site_atom)} sites, {len(want[1])} pairs")
    assert np.array_equal(got[0], want[0])
    assert np.array_equal(got[1], want[1])
    assert np.array_equal(got[2], want[2])


def test_slab_box_equals_brute_force_on_both_phases():
    # config 4: sites from the dense slab and the dilute gas (0 to ~10 neighbours)
    from soap_turbo_amd.workload import slab_sites
    spec, par = CONFIGS[4]
    L, pos, sp = place_atoms(spec)
    site_atom = np.arange(spec.n_atoms)[slab_sites(4, 50, 150)]
    rcut = float(np.max(par.per_species()["rcut_hard"]))
    got = build_neighbours(pos, L, rcut, site_atom)
    want = brute_force(pos, L, rcut, site_atom)
    assert (want[0] == 1).any() and want[0].max() > 60
    for g, x in zip(got, want):
        assert np.array_equal(g, x)
