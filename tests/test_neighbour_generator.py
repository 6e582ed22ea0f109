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


def test_small_box_single_cell_branch_equals_brute_force():
    # a box under three cutoffs wide: the generator takes every atom as a candidate (one cell)
    from soap_turbo_amd.workload import BoxSpec
    spec = BoxSpec(40, 0.02, 1.2, [0.5, 0.5], [1, 8], seed=3)
    L, pos, sp = place_atoms(spec)
    assert 2 * 4.5 < L < 3 * 4.5
    site_atom = np.arange(spec.n_atoms)
    got = build_neighbours(pos, L, 4.5, site_atom)
    want = brute_force(pos, L, 4.5, site_atom)
    assert want[0].max() > 3
    for g, x in zip(got, want):
        assert np.array_equal(g, x)


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


# ---------------------------------------------------------------- general (triclinic) cells
from soap_turbo_amd.workload import SoapParams, skewed_cell, triclinic_workload  # noqa: E402

TRI_PARAMS = SoapParams(alpha_max=[4, 4], l_max=4, rcut_hard=5.0, rcut_soft=4.5)


def brute_force_cell(pos, cell, rcut, site_atom):
    # the generator's FP64 operations, elementwise: image n = rint(H^-1 d), d - H n (H: vectors as columns)
    h = cell.T  # h[c, v]: component c of vector v
    a, b, c = cell
    bc = np.array([b[1] * c[2] - b[2] * c[1], b[2] * c[0] - b[0] * c[2], b[0] * c[1] - b[1] * c[0]])
    ca = np.array([c[1] * a[2] - c[2] * a[1], c[2] * a[0] - c[0] * a[2], c[0] * a[1] - c[1] * a[0]])
    ab = np.array([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]])
    det = a[0] * bc[0] + a[1] * bc[1] + a[2] * bc[2]
    hi = np.array([bc / det, ca / det, ab / det])
    n_neigh, nbr, vec = [], [], []
    for i in site_atom:
        d = pos - pos[i]
        x, y, z = d[:, 0], d[:, 1], d[:, 2]
        n = [np.rint(hi[v, 0] * x + hi[v, 1] * y + hi[v, 2] * z) for v in range(3)]
        x = x - (h[0, 0] * n[0] + h[0, 1] * n[1] + h[0, 2] * n[2])
        y = y - (h[1, 0] * n[0] + h[1, 1] * n[1] + h[1, 2] * n[2])
        z = z - (h[2, 0] * n[0] + h[2, 1] * n[1] + h[2, 2] * n[2])
        r2 = x * x + y * y + z * z
        inside = np.nonzero(r2 < rcut * rcut)[0]
        inside = inside[inside != i]
        n_neigh.append(1 + len(inside))
        nbr.extend([i, *inside])
        vec.append(np.zeros((1, 3)))
        vec.append(np.stack([x, y, z], axis=1)[inside])
    return np.array(n_neigh, np.int32), np.array(nbr, np.int32), np.concatenate(vec)


def test_general_cell_list_equals_brute_force():
    w = triclinic_workload(1500, 0.10, [0.6, 0.4], TRI_PARAMS, seed=11)
    got = (w.n_neigh, w.nbr, w.vec)
    want = brute_force_cell(w.pos, w.L, 5.0, w.site_atom)
    print(f"triclinic cell: {w.n_sites} sites, {len(want[1])} pairs, neighbours {want[0].min() - 1}..{want[0].max() - 1}")
    for g, x in zip(got, want):
        assert np.array_equal(g, x)


def test_general_cell_list_holds_every_image_inside_the_cutoff():
    # independent of the image formula: every image r_j - r_i + H n with n in -3..3 (each component; positions lie up to a
    # translation outside the home cell) that lies inside the cutoff appears exactly once, as a pair vector that
    # differs from r_j - r_i by a lattice vector
    w = triclinic_workload(800, 0.10, [1.0], TRI_PARAMS, seed=12)
    cell, rcut = w.L, 5.0
    shifts = np.array([[i, j, k] for i in range(-3, 4) for j in range(-3, 4) for k in range(-3, 4)]) @ cell
    ptr = np.concatenate([[0], np.cumsum(w.n_neigh)])
    for s, i in enumerate(w.site_atom):
        d = (w.pos - w.pos[i])[:, None, :] + shifts[None, :, :]
        r2 = np.einsum("abc,abc->ab", d, d)
        inside = r2 < rcut * rcut
        inside[i] = False
        atoms, images = np.nonzero(inside)
        assert np.all(np.bincount(atoms, minlength=len(w.pos)) <= 1)  # widths above 2 rcut: one image each
        lst = w.nbr[ptr[s] + 1:ptr[s + 1]]
        assert np.array_equal(lst, np.sort(atoms))
        assert np.allclose(w.vec[ptr[s] + 1:ptr[s + 1]], d[atoms, images][np.argsort(atoms)], rtol=0, atol=1e-12)
        frac = (w.vec[ptr[s] + 1:ptr[s + 1]] - (w.pos[lst] - w.pos[i])) @ np.linalg.inv(cell)
        assert np.abs(frac - np.rint(frac)).max() < 1e-9


def test_cubic_cell_as_lattice_vectors_gives_the_cubic_list():
    spec, par = CONFIGS[1]
    L, pos, sp = place_atoms(spec)
    site_atom = np.arange(spec.n_atoms)
    cubic = build_neighbours(pos, L, 5.0, site_atom)
    general = build_neighbours(pos, L * np.eye(3), 5.0, site_atom)
    assert np.array_equal(cubic[0], general[0]) and np.array_equal(cubic[1], general[1])
    assert np.abs(cubic[2] - general[2]).max() <= 1e-13


def test_general_cell_refuses_thin_or_left_handed_cells():
    pos = np.zeros((4, 3))
    thin = skewed_cell(200, 0.1)  # widths near 12.6 A: under two cutoffs of 7 A
    with pytest.raises(ValueError):
        build_neighbours(pos, thin, 7.0, np.arange(4))
    with pytest.raises(ValueError):
        build_neighbours(pos, skewed_cell(2000, 0.1)[[1, 0, 2]], 5.0, np.arange(4))
