"""Edge cases of the later rows on the device: sites with no neighbour (only their own pair) and a
one-site list, for the central-atom Hessian, the GAP model and the poly3operator basis, against
the oracle on the same inputs."""
import numpy as np
import pytest

import soap_turbo_amd as st
from oracle import oracle as orc
from soap_turbo_amd.workload import BoxSpec, SoapParams, make_workload, normal
from tests.gap_cases import energy_and_grad, model
from tests.golden_cases import rel

pytestmark = pytest.mark.gpu

# a dilute box: at rcut 3.0 several sites have no neighbour at all
SPARSE = BoxSpec(60, 0.004, 1.5, [0.5, 0.5], [1, 8], seed=23)


def _setup(ctx, w, legacy):
    p = w.params
    keep = st._lib.Keep()
    prm = st.make_params(keep, alpha_max=p.alpha_max, l_max=p.l_max, radial_enhancement=p.radial_enhancement,
                         basis=p.basis, scaling_mode=p.scaling_mode, legacy_seed=legacy, **p.per_species())
    cmp_ = st.make_compression(keep, False)
    W, S = ctx.get_basis(prm, sum(p.alpha_max))
    ctx.upload_neighbours(w.n_neigh, w.species.T, w.species_multiplicity, w.mask.T, w.rjs, w.thetas, w.phis)
    return keep, prm, cmp_, W, S


@pytest.fixture(scope="module")
def ctx():
    return st.Context(0)


@pytest.mark.parametrize("basis", ["poly3gauss", "poly3operator"])
def test_isolated_sites(ctx, basis):
    p = SoapParams(alpha_max=[4, 4], l_max=3, rcut_hard=3.0, rcut_soft=2.5, basis=basis,
                   atom_sigma_r=0.4, atom_sigma_t=0.4, central_weight=[1.0, 0.7])
    w = make_workload((SPARSE, p))
    assert (w.n_neigh == 1).sum() >= 3, "the box should leave some sites alone"
    keep, prm, cmp_, W, S = _setup(ctx, w, legacy=False)
    n_soap = 8 * 9 // 2 * 4
    # descriptors and forces (FP32 adjoint) against the oracle
    G = normal(7, (n_soap, w.n_sites))
    ctx.upload_atoms(len(w.pos), w.nbr, w.vec)
    ctx.upload_dedq(np.asfortranarray(G.T))
    ctx.adjoint_f32(prm, cmp_)
    q32 = ctx.download_soap_f32(n_soap).T
    pair32 = ctx.download_pairs(1).T
    qo, pair_o = orc.contract(w, W, S, G, 1.0, None, False)
    assert rel(q32, qo) <= 1e-5
    assert rel(pair32, pair_o) <= 1e-4
    # FP64 descriptors
    ctx.compute(prm, cmp_, False, False)
    q64, _ = ctx.download(n_soap, cart=False)
    assert rel(q64.T, qo) <= 1e-10
    # the Hessian (poly3gauss only; poly3operator has no second radial derivatives)
    if basis == "poly3gauss":
        vecs = normal(9, (w.n_sites, 2, n_soap))
        h = ctx.central_hessian(prm, cmp_, vecs.transpose(2, 1, 0)).transpose(3, 2, 0, 1)
        _, ho, _, _ = orc.central_hessian(w, W, S, vecs, None)
        assert rel(h, ho) <= 1e-10
        alone = np.cumsum(np.concatenate([[0], w.n_neigh[:-1]]))[w.n_neigh == 1]
        assert np.all(h[alone] == 0.0)
    # the GAP model
    ctx.upload_neighbours(w.n_neigh, w.species.T, w.species_multiplicity, w.mask.T, w.rjs, w.thetas, w.phis)
    ctx.upload_atoms(len(w.pos), w.nbr, w.vec)
    T, alpha = model(qo, 16, 3)
    ctx.set_gap_model(T.T, alpha, 2, 0.5)
    E = ctx.gap_energy_forces(prm, cmp_)
    e_ref, _ = energy_and_grad(qo, T, alpha, 2, 0.5)
    assert abs(E - e_ref.sum()) <= 1e-5 * np.abs(e_ref).sum()


def test_one_site(ctx):
    p = SoapParams(alpha_max=[3, 3], l_max=2, rcut_hard=4.0, rcut_soft=3.5, basis="poly3gauss",
                   atom_sigma_r=0.45, atom_sigma_t=0.45)
    w = make_workload((BoxSpec(80, 0.05, 1.2, [0.5, 0.5], [1, 8], seed=5), p), n_sites=1)
    keep, prm, cmp_, W, S = _setup(ctx, w, legacy=False)
    n_soap = 6 * 7 // 2 * 3
    vecs = normal(4, (1, 1, n_soap))
    h = ctx.central_hessian(prm, cmp_, vecs.transpose(2, 1, 0)).transpose(3, 2, 0, 1)
    _, ho, _, _ = orc.central_hessian(w, W, S, vecs, None)
    assert rel(h, ho) <= 1e-10
    assert np.abs(h.sum(axis=0)).max() <= 1e-10 * np.abs(h).max()


def test_poly3operator_small_alpha(ctx):
    # alpha_max < 4: no reference output exists (the reference writes past an array there);
    # the device paths are held to the oracle
    p = SoapParams(alpha_max=[2, 3], l_max=3, rcut_hard=4.0, rcut_soft=3.4, basis="poly3operator",
                   atom_sigma_r=0.45, atom_sigma_t=0.45)
    w = make_workload((BoxSpec(90, 0.05, 1.2, [0.5, 0.5], [1, 8], seed=11), p), n_sites=20)
    keep, prm, cmp_, W, S = _setup(ctx, w, legacy=False)
    n_soap = 5 * 6 // 2 * 4
    G = normal(3, (n_soap, w.n_sites))
    qo, pair_o = orc.contract(w, W, S, G, 1.0, None, False)
    ctx.compute(prm, cmp_, False, False)
    q64, _ = ctx.download(n_soap, cart=False)
    assert rel(q64.T, qo) <= 1e-10
    ctx.upload_dedq(np.asfortranarray(G.T))
    ctx.adjoint_f32(prm, cmp_)
    assert rel(ctx.download_soap_f32(n_soap).T, qo) <= 1e-5
    assert rel(ctx.download_pairs(1).T, pair_o) <= 1e-4
