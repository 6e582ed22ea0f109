"""The device neighbour-list builder against the host cell-list generator of the workloads (cubic
boxes, and a triclinic cell given by its lattice vectors).

Integer outputs (pair counts, neighbour atoms and their order, species masks, the pairs of each
atom) and the minimum-image pair vectors and distances are bit-exact: both sides form them with
the same FP64 operations. thetas and phis come from the device's acos / atan2, held to 2e-16
relative (a few ulp). Descriptors and forces computed from the device-built list equal those
from the uploaded host list to the same bound.
"""
import numpy as np
import pytest

import soap_turbo_amd as st
from soap_turbo_amd.workload import BoxSpec, SoapParams, make_workload, normal

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return st.Context(0)


def build(ctx, w):
    p = w.params
    rcut = float(np.max(p.per_species()["rcut_hard"]))
    n_pairs = ctx.build_neighbours(w.pos, w.atom_species + 1, p.n_species, w.L, rcut, w.site_atom)
    return n_pairs, ctx.download_neighbours()


CASES = {
    "config 1 (216 atoms, every atom a site)": (1, None),
    "config 2 (4096 atoms, the 1228 Z=6 sites)": (2, None),
    "config 4 slab, 600 sites": (4, 600),
    # box side under three cutoffs: one cell holds every atom
    "small box, 40 atoms": ((BoxSpec(40, 0.02, 1.2, [0.5, 0.5], [1, 8], seed=3),
                             SoapParams(alpha_max=[3, 3], l_max=3, rcut_hard=4.5, rcut_soft=4.0)), None),
}


@pytest.mark.parametrize("case", list(CASES))
def test_device_list_matches_host_generator(ctx, case):
    cfg, n_sites = CASES[case]
    w = make_workload(cfg, n_sites=n_sites)
    n_pairs, (n_neigh, pa, rjs, th, ph, mask, vec) = build(ctx, w)
    print(f"{case}: {w.n_sites} sites, {n_pairs} pairs")
    assert n_pairs == len(w.nbr)
    assert np.array_equal(n_neigh, w.n_neigh)
    assert np.array_equal(pa, w.nbr)
    assert np.array_equal(mask, w.mask)
    assert np.array_equal(vec, w.vec)
    assert np.array_equal(rjs, w.rjs)
    assert np.abs(th - w.thetas).max() <= 2e-16 * np.pi
    assert np.abs(ph - w.phis).max() <= 2e-16 * np.pi
    # the pairs of each atom, in increasing pair index
    ap, pr = ctx.atom_pairs()
    order = np.argsort(w.nbr, kind="stable")
    assert np.array_equal(pr, order)
    assert np.array_equal(ap, np.concatenate([[0], np.cumsum(np.bincount(w.nbr, minlength=len(w.pos)))]))


def test_descriptors_and_forces_from_the_device_list(ctx):
    w = make_workload(2)
    p = w.params
    keep = st._lib.Keep()
    prm = st.make_params(keep, alpha_max=p.alpha_max, l_max=p.l_max, radial_enhancement=p.radial_enhancement,
                         basis=p.basis, scaling_mode=p.scaling_mode, **p.per_species())
    dim, nz = st.get_compress_indices("trivial", p.alpha_max, p.l_max, "get_dim")
    comp = (np.zeros(nz, np.int32), np.zeros(nz, np.int32), np.zeros(nz))
    st.get_compress_indices("trivial", p.alpha_max, p.l_max, "set_indices", *comp)
    cmp_ = st.make_compression(keep, True, nz, *comp)
    n_soap = int(comp[0].max())
    G = normal(17, (n_soap, w.n_sites))
    # host list uploaded
    ctx.upload_neighbours(w.n_neigh, w.species.T, w.species_multiplicity, w.mask.T, w.rjs, w.thetas, w.phis)
    ctx.upload_atoms(len(w.pos), w.nbr, w.vec)
    ctx.upload_dedq(np.asfortranarray(G.T))
    ctx.adjoint_f32(prm, cmp_)
    q_h = ctx.download_soap_f32(n_soap)
    f_h, v_h = ctx.download_forces(1)
    # device list
    build(ctx, w)
    ctx.upload_dedq(np.asfortranarray(G.T))
    ctx.adjoint_f32(prm, cmp_)
    q_d = ctx.download_soap_f32(n_soap)
    f_d, v_d = ctx.download_forces(1)
    scale = np.abs(f_h).max()
    print(f"device list vs host list: q {np.abs(q_d - q_h).max():.1e}, forces {np.abs(f_d - f_h).max() / scale:.1e}")
    assert np.abs(q_d - q_h).max() <= 1e-6
    assert np.abs(f_d - f_h).max() <= 1e-6 * scale
    assert np.abs(v_d - v_h).max() <= 1e-6 * np.abs(v_h).max()


def test_box_under_two_cutoffs_is_refused(ctx):
    w = make_workload(1)
    with pytest.raises(st.SoapTurboError):
        ctx.build_neighbours(w.pos, w.atom_species + 1, 1, 7.0, 3.6, w.site_atom)


# ---------------------------------------------------------------- general (triclinic) cells
from soap_turbo_amd.workload import skewed_cell, triclinic_workload  # noqa: E402
from oracle import oracle as orc  # noqa: E402
from tests.golden_cases import rel  # noqa: E402

TRI_PARAMS = SoapParams(alpha_max=[5, 5], l_max=6, rcut_hard=5.0, rcut_soft=4.5, basis="poly3gauss",
                        radial_enhancement=1)


@pytest.fixture(scope="module")
def tri():
    # 3000 atoms of two species in a skewed cell (angles near 70-80 degrees), a fifth of them outside the home cell
    return triclinic_workload(3000, 0.10, [0.6, 0.4], TRI_PARAMS, seed=21)


def test_general_cell_device_list_matches_host_generator(ctx, tri):
    w = tri
    n_pairs = ctx.build_neighbours(w.pos, w.atom_species + 1, w.params.n_species, w.L, 5.0, w.site_atom)
    n_neigh, pa, rjs, th, ph, mask, vec = ctx.download_neighbours()
    print(f"triclinic cell: {w.n_sites} sites, {n_pairs} pairs")
    assert n_pairs == len(w.nbr)
    assert np.array_equal(n_neigh, w.n_neigh)
    assert np.array_equal(pa, w.nbr)
    assert np.array_equal(mask, w.mask)
    assert np.array_equal(vec, w.vec)
    assert np.array_equal(rjs, w.rjs)
    assert np.abs(th - w.thetas).max() <= 2e-16 * np.pi
    assert np.abs(ph - w.phis).max() <= 2e-16 * np.pi
    ap, pr = ctx.atom_pairs()
    assert np.array_equal(pr, np.argsort(w.nbr, kind="stable"))


def test_general_cell_descriptors_and_forces_match_oracle(ctx, tri):
    # both device paths on the device-built triclinic list against the FP64 oracle on the host list
    w = tri
    p = w.params
    keep = st._lib.Keep()
    prm = st.make_params(keep, alpha_max=p.alpha_max, l_max=p.l_max, radial_enhancement=p.radial_enhancement,
                         basis=p.basis, scaling_mode=p.scaling_mode, legacy_seed=True, **p.per_species())
    cmp_ = st.make_compression(keep, False, 0, None, None, None)
    n_max = sum(p.alpha_max)
    n_soap = n_max * (n_max + 1) // 2 * (p.l_max + 1)
    W, S = ctx.get_basis(prm, n_max)
    G = normal(23, (n_soap, w.n_sites))
    qo, pair_o = orc.contract(w, W, S, G, 1.0, None, True)
    fo = np.zeros_like(w.pos)
    np.add.at(fo, w.nbr, -pair_o)
    vo = -np.einsum("ka,kb->ab", w.vec, pair_o)
    ctx.build_neighbours(w.pos, w.atom_species + 1, p.n_species, w.L, 5.0, w.site_atom)
    ctx.upload_dedq(np.asfortranarray(G.T))
    ctx.adjoint_f32(prm, cmp_)
    q32 = ctx.download_soap_f32(n_soap).T
    f32, v32 = ctx.download_forces(1)
    ctx.full_jacobian_forces(prm, cmp_)
    q64, _ = ctx.download(n_soap)
    f64, v64 = ctx.download_forces(0)
    e32 = [rel(q32, qo), rel(f32.T, fo), rel(v32, vo)]
    e64 = [rel(q64.T, qo), rel(f64.T, fo), rel(v64, vo)]
    print(f"triclinic FP32: q {e32[0]:.1e} forces {e32[1]:.1e} virial {e32[2]:.1e}; "
          f"FP64: q {e64[0]:.1e} forces {e64[1]:.1e} virial {e64[2]:.1e}")
    assert max(e32) <= 1e-5
    assert max(e64) <= 1e-10


def test_general_cell_under_two_cutoffs_is_refused(ctx):
    pos = np.zeros((4, 3))
    with pytest.raises(st.SoapTurboError):
        ctx.build_neighbours(pos, np.ones(4, np.int32), 1, skewed_cell(200, 0.1), 7.0, np.arange(4))
    with pytest.raises(st.SoapTurboError):  # left-handed
        ctx.build_neighbours(pos, np.ones(4, np.int32), 1, skewed_cell(2000, 0.1)[[1, 0, 2]], 5.0, np.arange(4))
