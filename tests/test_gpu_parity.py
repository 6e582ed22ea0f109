"""Parity of the HIP FP64 path with the reference (golden fixtures) and with the oracle.

Through the C-ABI (ctypes). Tolerance: 1e-10 relative (max |a-b| / max |b| over
each output array) for soap, soap_cart_der and the pair gradients, with the
reference's own W injected, since W = S^{-1/2} is only determined to ~1e-5 at
poly3 alpha_max = 8 by any FP64 route (cond(S) = 1.3e13).
"""
import numpy as np
import pytest

import soap_turbo_amd as st
from oracle import oracle as orc
from soap_turbo_amd.workload import BoxSpec, SoapParams, Workload, make_workload, normal
from tests.golden_cases import CASES, checksums, compression_indices, load, rel

pytestmark = pytest.mark.gpu

TOL = 1e-10


@pytest.fixture(scope="module")
def ctx():
    return st.Context(0)


def _params(keep, w, legacy):
    p = w.params
    ps = p.per_species()
    return st.make_params(keep, alpha_max=p.alpha_max, l_max=p.l_max, radial_enhancement=p.radial_enhancement,
                          basis=p.basis, scaling_mode=p.scaling_mode, legacy_seed=legacy, **ps)


def run_get_soap(ctx, w, comp, legacy, do_der=True, contract=False):
    """st.get_soap with the reference's argument list; returns oracle-layout (soap, cart)."""
    p = w.params
    ps = p.per_species()
    n_max = sum(p.alpha_max)
    n_soap = int(np.max(comp[0])) if comp is not None else n_max * (n_max + 1) // 2 * (p.l_max + 1)
    soap = np.zeros((n_soap, w.n_sites), order="F")
    cart = np.zeros((3, n_soap, w.n_pairs), order="F") if do_der and not contract else np.zeros((3, 1, 1), order="F")
    old = st.legacy_filter_seed
    st.legacy_filter_seed = legacy
    try:
        st.get_soap(w.n_sites, w.n_neigh, p.n_species, w.species.T, w.species_multiplicity, w.n_pairs, w.mask.T,
                    w.rjs, w.thetas, w.phis, p.alpha_max, p.l_max, ps["rcut_hard"], ps["rcut_soft"], ps["nf"],
                    ps["global_scaling"], ps["atom_sigma_r"], ps["atom_sigma_r_scaling"], ps["atom_sigma_t"],
                    ps["atom_sigma_t_scaling"], ps["amplitude_scaling"], p.radial_enhancement,
                    ps["central_weight"], p.basis, p.scaling_mode, False, do_der, comp is not None,
                    len(comp[0]) if comp is not None else 0, comp[0] if comp is not None else None,
                    comp[1] if comp is not None else None, comp[2] if comp is not None else None, soap, cart,
                    contract=contract, ctx=ctx)
    finally:
        st.legacy_filter_seed = old
    return soap.T, (cart.transpose(2, 1, 0) if do_der and not contract else None)


@pytest.mark.parametrize("name", CASES)
def test_hip_matches_reference_fixtures(ctx, name):
    w, comp, d = load(name)
    legacy = bool(d["legacy_seed"])
    ctx.set_basis(d["W"], d["S"])
    try:
        q0, _ = run_get_soap(ctx, w, comp, legacy, do_der=False)
        assert rel(q0, d["soap_desc"]) <= TOL
        q, cart = run_get_soap(ctx, w, comp, legacy, do_der=True)
        assert rel(q, d["soap"]) <= TOL
        assert rel(cart[:len(d["cart_head"])], d["cart_head"]) <= TOL
        np.testing.assert_allclose(checksums(cart)[[1, 3]], d["cart_checksums"][[1, 3]], rtol=TOL)
        np.testing.assert_allclose(np.sqrt((cart ** 2).sum(axis=(1, 2))), d["cart_pair_norm"], rtol=1e-9,
                                   atol=TOL * d["cart_pair_norm"].max())
        qc, _ = run_get_soap(ctx, w, comp, legacy, do_der=True, contract=True)
        assert rel(qc, d["soap_contract_q"]) <= TOL
        pair = np.zeros((3, w.n_pairs), order="F")
        st.soap_contract(qc.T, d["G"].T, 1.0, pair, ctx=ctx)
        assert rel(pair.T, d["pair"]) <= TOL
    finally:
        ctx.clear_basis()


@pytest.mark.parametrize("cfg", [1, 2])
def test_hip_matches_oracle_full_config(ctx, cfg):
    w = make_workload(cfg)
    p = w.params
    keep = st._lib.Keep()
    prm = _params(keep, w, True)
    n_max = sum(p.alpha_max)
    W, S = ctx.get_basis(prm, n_max)
    comp = None
    if p.compress_mode == "trivial":
        dim, nz = st.get_compress_indices("trivial", p.alpha_max, p.l_max, "get_dim")
        comp = (np.zeros(nz, np.int32), np.zeros(nz, np.int32), np.zeros(nz))
        st.get_compress_indices("trivial", p.alpha_max, p.l_max, "set_indices", *comp)
    q, cart = run_get_soap(ctx, w, comp, True)
    qo, carto, _ = orc.get_soap(w, W, S, comp, do_der=True, legacy_seed=True)
    assert rel(q, qo) <= TOL
    assert rel(cart, carto) <= TOL
    # resident path: full-Jacobian contraction, soap_contract, forces
    n_soap = q.shape[1]
    G = normal(77, (n_soap, w.n_sites))
    ctx.upload_neighbours(w.n_neigh, w.species.T, w.species_multiplicity, w.mask.T, w.rjs, w.thetas, w.phis)
    ctx.upload_atoms(len(w.pos), w.nbr, w.vec)
    ctx.upload_dedq(np.asfortranarray(G.T))
    cmp_ = st.make_compression(keep, comp is not None, len(comp[0]) if comp else 0, *(comp or (None, None, None)))
    ctx.compute(prm, cmp_, True, False)
    ctx.pairs_full_jacobian()
    pair_full = ctx.download_pairs(0).T
    _, pair_o = orc.contract(w, W, S, G, 1.0, comp, True)
    assert rel(pair_full, pair_o) <= TOL
    forces, virial = ctx.download_forces(0)
    fo = np.zeros_like(w.pos)
    np.add.at(fo, w.nbr, -pair_o)
    assert rel(forces.T, fo) <= TOL
    vo = -np.einsum("ka,kb->ab", w.vec, pair_o)
    assert rel(virial, vo) <= TOL
    ctx.compute(prm, cmp_, True, True)
    ctx.contract_resident(1.0)
    assert rel(ctx.download_pairs(0).T, pair_o) <= TOL


def _handmade_workload():
    """Edge cases: an isolated site, neighbours at and beyond rcut_hard, a neighbour of no species."""
    par = SoapParams(alpha_max=[4, 3], l_max=4, rcut_hard=[4.0, 5.0], rcut_soft=[3.5, 4.5], basis="poly3",
                     central_weight=[0.0, 1.0])
    rng = np.random.default_rng(3)
    n_neigh = np.array([1, 5, 1, 7], dtype=np.int32)
    P = int(n_neigh.sum())
    rjs = rng.uniform(0.8, 4.8, P)
    starts = np.concatenate([[0], np.cumsum(n_neigh)[:-1]])
    rjs[starts] = 0.0
    rjs[starts[1] + 1] = 4.0       # exactly rcut_hard of species 1: excluded for it
    rjs[starts[1] + 2] = 5.0       # rcut_hard of species 2 = rcut_max: excluded everywhere
    rjs[starts[3] + 3] = 5.3       # beyond every cutoff
    thetas = rng.uniform(0, np.pi, P); phis = rng.uniform(-np.pi, np.pi, P)
    thetas[starts] = 0.0; phis[starts] = 0.0
    sp = rng.integers(0, 2, P)
    mask = np.zeros((2, P), dtype=bool)
    mask[sp, np.arange(P)] = True
    mask[:, starts[3] + 4] = False  # a neighbour of no listed species
    species = np.array([[1], [2], [1], [2]], dtype=np.int32)
    mask[:, starts] = False
    for i, s0 in enumerate(starts):
        mask[species[i, 0] - 1, s0] = True
    return Workload(par, 0.0, None, None, None, n_neigh, None, None, rjs, thetas, phis, mask, species,
                    np.ones(4, dtype=np.int32))


def test_edge_cases_match_oracle(ctx):
    w = _handmade_workload()
    keep = st._lib.Keep()
    prm = _params(keep, w, True)
    W, S = ctx.get_basis(prm, 7)
    q, cart = run_get_soap(ctx, w, None, True)
    qo, carto, _ = orc.get_soap(w, W, S, None, do_der=True, legacy_seed=True)
    assert rel(q, qo) <= TOL
    assert rel(cart, carto) <= TOL
    assert np.all(q[0] == 0.0)  # empty sphere and no central weight: soap stays zero


def test_reference_stop_conditions(ctx):
    w = _handmade_workload()
    w.params.alpha_max = [11, 3]
    with pytest.raises(st.SoapTurboStop):
        run_get_soap(ctx, w, None, True)
    w = _handmade_workload()
    pair = np.zeros((3, w.n_pairs), order="F")
    run_get_soap(ctx, w, None, True, do_der=False)
    with pytest.raises(st.SoapTurboStop):
        st.soap_contract(np.zeros((150, 4), order="F"), np.zeros((4, 150), order="F"), 1.0, pair, ctx=ctx)
    w.params.l_max = 13
    with pytest.raises(st.SoapTurboError):
        run_get_soap(ctx, w, None, True)


# every block shape of the Jacobian kernel (outputs per thread x threads per block), and both of its
# compression paths (a direct output -> component map, and a list of weighted components per output)
@pytest.mark.parametrize("alpha_max,l_max,kind", [
    ([2, 2], 2, None),            # n_soap 30
    ([5, 5], 6, "scaled"),        # 256 of 385 components, one per output: 256 x 1
    ([5, 5], 6, None),            # 385: 256 x 2
    ([6, 6], 6, "multi"),         # 409 of 546 components into 202 outputs
    ([6, 6], 6, None),            # 546: 256 x 3
    ([8, 8], 8, None),            # 1224: 512 x 3
    ([8, 8, 4], 8, None),         # 1890: 512 x 4
    ([8, 8, 8], 8, "multi"),      # 2025 of 2700 components into 1013 outputs: 512 x 2
    ([8, 8, 8], 8, None),         # 2700: 512 x 6
])
def test_jacobian_kernel_shapes_match_oracle(ctx, alpha_max, l_max, kind):
    ns = len(alpha_max)
    probs = [1.0 / ns] * ns
    spec = BoxSpec(343, 0.09, 1.0, probs, [1, 6, 8][:ns], seed=11 + ns)
    par = SoapParams(alpha_max=alpha_max, l_max=l_max, rcut_hard=4.5, rcut_soft=4.0, basis="poly3gauss",
                     radial_enhancement=1)
    w = make_workload((spec, par), n_sites=6)
    n_max = sum(alpha_max)
    comp = compression_indices(kind, n_max * (n_max + 1) // 2 * (l_max + 1), 5 + n_max)
    keep = st._lib.Keep()
    prm = _params(keep, w, False)
    W, S = ctx.get_basis(prm, n_max)
    q, cart = run_get_soap(ctx, w, comp, False)
    qo, carto, _ = orc.get_soap(w, W, S, comp, do_der=True, legacy_seed=False)
    assert rel(q, qo) <= TOL
    assert rel(cart, carto) <= TOL
