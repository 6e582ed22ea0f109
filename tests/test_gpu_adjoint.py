"""The FP32 adjoint path against the FP64 oracle: descriptors, pair gradients, forces, virial.

Tolerance: 1e-5 relative (max |a-b| / max |b| over each array), the bound the
FP32 path is held to for forces; the oracle runs in FP64 with the same W.
"""
import numpy as np
import pytest

import soap_turbo_amd as st
from oracle import oracle as orc
from soap_turbo_amd.workload import BoxSpec, SoapParams, make_workload, normal
from tests.golden_cases import CASES, load, rel

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture(scope="module")
def ctx():
    return st.Context(0)


def setup(ctx, w, legacy=True, seed=5):
    p = w.params
    keep = st._lib.Keep()
    ps = p.per_species()
    prm = st.make_params(keep, alpha_max=p.alpha_max, l_max=p.l_max, radial_enhancement=p.radial_enhancement,
                         basis=p.basis, scaling_mode=p.scaling_mode, legacy_seed=legacy, **ps)
    comp = None
    if p.compress_mode == "trivial":
        dim, nz = st.get_compress_indices("trivial", p.alpha_max, p.l_max, "get_dim")
        comp = (np.zeros(nz, np.int32), np.zeros(nz, np.int32), np.zeros(nz))
        st.get_compress_indices("trivial", p.alpha_max, p.l_max, "set_indices", *comp)
    cmp_ = st.make_compression(keep, comp is not None, len(comp[0]) if comp else 0, *(comp or (None, None, None)))
    n_max = sum(p.alpha_max)
    n_soap = int(comp[0].max()) if comp else n_max * (n_max + 1) // 2 * (p.l_max + 1)
    W, S = ctx.get_basis(prm, n_max)
    G = normal(seed, (n_soap, w.n_sites))
    ctx.upload_neighbours(w.n_neigh, w.species.T, w.species_multiplicity, w.mask.T, w.rjs, w.thetas, w.phis)
    ctx.upload_atoms(len(w.pos), w.nbr, w.vec)
    ctx.upload_dedq(np.asfortranarray(G.T))
    return keep, prm, cmp_, comp, n_soap, W, S, G


@pytest.mark.parametrize("cfg,n_sites,legacy", [(1, None, True), (2, None, True), (2, 300, False), (4, 400, True)])
def test_adjoint_f32_matches_fp64_oracle(ctx, cfg, n_sites, legacy):
    w = make_workload(cfg, n_sites=n_sites)
    keep, prm, cmp_, comp, n_soap, W, S, G = setup(ctx, w, legacy)
    ctx.adjoint_f32(prm, cmp_)
    q32 = ctx.download_soap_f32(n_soap).T
    pair32 = ctx.download_pairs(1).T
    f32, v32 = ctx.download_forces(1)
    qo, pair_o = orc.contract(w, W, S, G, 1.0, comp, legacy)
    fo = np.zeros_like(w.pos)
    np.add.at(fo, w.nbr, -pair_o)
    vo = -np.einsum("ka,kb->ab", w.vec, pair_o)
    eq, ep, ef, ev = rel(q32, qo), rel(pair32, pair_o), rel(f32.T, fo), rel(v32, vo)
    print(f"cfg {cfg}: rel err q {eq:.2e} pair {ep:.2e} forces {ef:.2e} virial {ev:.2e}")
    assert eq <= TOL
    assert ef <= TOL
    assert ev <= TOL
    assert ep <= TOL


def test_adjoint_f32_forces_sum_to_zero_and_repeat(ctx):
    w = make_workload(1)
    keep, prm, cmp_, comp, n_soap, W, S, G = setup(ctx, w)
    ctx.adjoint_f32(prm, cmp_)
    f1, _ = ctx.download_forces(1)
    ctx.adjoint_f32(prm, cmp_)
    f2, _ = ctx.download_forces(1)
    scale = np.abs(f1).max()
    assert np.abs(f1.sum(axis=1)).max() <= 1e-5 * scale * np.sqrt(f1.shape[1])
    assert rel(f2, f1) <= 1e-6
    # the forces are gathered per atom in a fixed order: repeat calls agree exactly, virial included
    _, v1 = ctx.download_forces(1)
    ctx.adjoint_f32(prm, cmp_)
    f3, v3 = ctx.download_forces(1)
    assert np.array_equal(f3, f2) and np.array_equal(v3, v1)


def test_adjoint_f32_needs_dedq_for_the_current_sites(ctx):
    # a new neighbour list (other sites) invalidates dE/dq: the FP32 path refuses until it is uploaded
    w = make_workload(1)
    keep, prm, cmp_, comp, n_soap, W, S, G = setup(ctx, w)
    ctx.adjoint_f32(prm, cmp_)
    w2 = make_workload(1, n_sites=50)
    ctx.upload_neighbours(w2.n_neigh, w2.species.T, w2.species_multiplicity, w2.mask.T, w2.rjs, w2.thetas, w2.phis)
    with pytest.raises(st.SoapTurboError):
        ctx.adjoint_f32(prm, cmp_)


def test_adjoint_f32_central_differences(ctx):
    # forces = -dE/dx with E = sum_i G_i . q_i (FP64 oracle energies), step 1e-4 A, 64 seeded atoms
    spec = BoxSpec(64, 0.06, 1.2, [1.0], [6], seed=21)
    par = SoapParams(alpha_max=[6], l_max=6, rcut_hard=4.5, rcut_soft=4.0, basis="poly3")
    w = make_workload((spec, par))
    keep, prm, cmp_, comp, n_soap, W, S, G = setup(ctx, w, legacy=False)
    ctx.adjoint_f32(prm, cmp_)
    f32, _ = ctx.download_forces(1)
    from soap_turbo_amd.workload import spherical

    def energy(pos):
        site = np.repeat(w.site_atom, w.n_neigh)
        d = pos[w.nbr] - pos[site]
        d -= w.L * np.round(d / w.L)
        d[w.nbr == site] = 0.0
        wq = make_workload((spec, par))
        wq.rjs, wq.thetas, wq.phis = spherical(d)
        q, _, _ = orc.get_soap(wq, W, S, do_der=False, legacy_seed=False)
        return float(np.sum(q * G.T))

    h = 1e-4
    scale = np.abs(f32).max()
    worst = 0.0
    for a in range(0, 64, 9):
        for c in range(3):
            pp = w.pos.copy(); pp[a, c] += h
            pm = w.pos.copy(); pm[a, c] -= h
            fd = -(energy(pp) - energy(pm)) / (2 * h)
            worst = max(worst, abs(fd - f32[c, a]) / scale)
    print(f"central differences: max rel err {worst:.2e}")
    assert worst <= 1e-5


# every site-kernel shape class: n_max tiles 1-3, l_max 4-12, 1-5 species, the compression paths
# (none, one weighted component per output, several per output), and sites with more than one tile
# of 64 neighbours
@pytest.mark.parametrize("alpha_max,l_max,kind,rcut", [
    ([3, 3], 4, None, 4.5),
    ([5, 5], 6, "scaled", 4.5),
    ([8, 8, 8], 8, "multi", 4.5),
    ([11, 11, 11], 4, None, 4.5),       # n_max 33: three row tiles
    ([4], 10, None, 4.5),
    ([3, 2], 12, None, 4.5),
    ([2, 2, 2, 2, 2], 5, None, 4.5),
    ([4], 4, None, 6.5),                # about 100 neighbours per site: two tiles
    ([7, 7, 7, 7], 6, "scaled", 4.5),   # n_max 28: two row tiles with every K step (the full form)
    ([11, 11, 11, 11], 4, None, 4.5),   # n_max 44: three row tiles, the full form
])
def test_adjoint_f32_shapes_match_fp64_oracle(ctx, alpha_max, l_max, kind, rcut):
    from tests.golden_cases import compression_indices
    ns = len(alpha_max)
    spec = BoxSpec(1000, 0.09, 1.0, [1.0 / ns] * ns, [1, 6, 7, 8, 14][:ns], seed=31 + ns)
    par = SoapParams(alpha_max=alpha_max, l_max=l_max, rcut_hard=rcut, rcut_soft=rcut - 0.5, basis="poly3gauss",
                     radial_enhancement=1)
    w = make_workload((spec, par), n_sites=24)
    keep, prm, cmp_, _, n_soap, W, S, G = setup(ctx, w)
    n_max = sum(alpha_max)
    comp = compression_indices(kind, n_max * (n_max + 1) // 2 * (l_max + 1), 7 + n_max)
    if comp is not None:
        cmp_ = st.make_compression(keep, True, len(comp[0]), *comp)
        n_soap = int(comp[0].max())
        G = normal(5, (n_soap, w.n_sites))
        ctx.upload_dedq(np.asfortranarray(G.T))
    ctx.adjoint_f32(prm, cmp_)
    q32 = ctx.download_soap_f32(n_soap).T
    pair32 = ctx.download_pairs(1).T
    f32, v32 = ctx.download_forces(1)
    qo, pair_o = orc.contract(w, W, S, G, 1.0, comp, True)
    fo = np.zeros_like(w.pos)
    np.add.at(fo, w.nbr, -pair_o)
    vo = -np.einsum("ka,kb->ab", w.vec, pair_o)
    eq, ep, ef, ev = rel(q32, qo), rel(pair32, pair_o), rel(f32.T, fo), rel(v32, vo)
    print(f"{alpha_max} l {l_max} {kind} rcut {rcut} (max {w.n_neigh.max()} neighbours): "
          f"rel err q {eq:.2e} pair {ep:.2e} forces {ef:.2e} virial {ev:.2e}")
    assert eq <= TOL and ef <= TOL and ev <= TOL and ep <= TOL


def test_adjoint_f32_graph_replays_match_eager_runs(ctx):
    # graph mode: call 1 runs, call 2 records the step, later calls replay it. Every replay reads the
    # current dE/dq and neighbour data and must equal an uncaptured run on the same inputs exactly.
    w = make_workload(2, n_sites=500)
    keep, prm, cmp_, comp, n_soap, W, S, G = setup(ctx, w)
    eager = st.Context(0)
    setup(eager, w)

    def run(c, G_):
        c.upload_dedq(np.asfortranarray(G_.T))
        c.adjoint_f32(prm, cmp_)
        f, v = c.download_forces(1)
        return c.download_soap_f32(n_soap), c.download_pairs(1), f, v

    ctx.set_graphs(True)
    try:
        for call in range(5):
            G_ = normal(100 + call, G.shape)
            got, want = run(ctx, G_), run(eager, G_)
            for a, b in zip(got, want):
                assert np.array_equal(a, b), f"call {call}"
        # a new neighbour list of the same sizes (other coordinates): the recorded step is not reused
        w2 = make_workload(2, n_sites=500)
        w2.rjs = w2.rjs * 0.97
        for c in (ctx, eager):
            c.upload_neighbours(w2.n_neigh, w2.species.T, w2.species_multiplicity, w2.mask.T, w2.rjs, w2.thetas,
                                w2.phis)
            c.upload_atoms(len(w2.pos), w2.nbr, w2.vec)
        for call in range(3):
            G_ = normal(200 + call, G.shape)
            got, want = run(ctx, G_), run(eager, G_)
            for a, b in zip(got, want):
                assert np.array_equal(a, b), f"call {call} after the new list"
    finally:
        ctx.set_graphs(False)


@pytest.mark.parametrize("name", [c for c in CASES if c.startswith("darby_")])
def test_adjoint_f32_matches_reference_darby_fixtures(ctx, name):
    # the nu_R_nu_S compressions (outputs summing many weighted components) against the reference's
    # own descriptors and soap_contract pair gradients, with the reference's W and S
    w, comp, d = load(name)
    p = w.params
    keep = st._lib.Keep()
    prm = st.make_params(keep, alpha_max=p.alpha_max, l_max=p.l_max, radial_enhancement=p.radial_enhancement,
                         basis=p.basis, scaling_mode=p.scaling_mode, legacy_seed=bool(d["legacy_seed"]),
                         **p.per_species())
    cmp_ = st.make_compression(keep, True, len(comp[0]), *comp)
    n_soap = int(comp[0].max())
    ctx.set_basis(d["W"], d["S"])
    try:
        ctx.upload_neighbours(w.n_neigh, w.species.T, w.species_multiplicity, w.mask.T, w.rjs, w.thetas, w.phis)
        ctx.upload_dedq(np.asfortranarray(d["G"].T))
        ctx.adjoint_f32(prm, cmp_)
        q = ctx.download_soap_f32(n_soap).T
        pair = ctx.download_pairs(1).T
    finally:
        ctx.clear_basis()
    print(f"{name}: n_soap {n_soap}, rel err q {rel(q, d['soap_contract_q']):.2e} pair {rel(pair, d['pair']):.2e}")
    assert rel(q, d["soap_contract_q"]) <= TOL
    assert rel(pair, d["pair"]) <= TOL


@pytest.mark.parametrize("name", ["op_trivial", "op_scalings"])
def test_adjoint_f32_matches_reference_poly3operator_fixtures(ctx, name):
    # the poly3operator basis: its tables run to the largest cutoff (the reference has no cutoff
    # test for it), split where the atom's integration limits switch
    w, comp, d = load(name)
    p = w.params
    keep = st._lib.Keep()
    prm = st.make_params(keep, alpha_max=p.alpha_max, l_max=p.l_max, radial_enhancement=p.radial_enhancement,
                         basis=p.basis, scaling_mode=p.scaling_mode, **p.per_species())
    cmp_ = st.make_compression(keep, comp is not None, len(comp[0]) if comp else 0, *(comp or (None, None, None)))
    n_soap = int(comp[0].max()) if comp else d["soap"].shape[1]
    ctx.set_basis(d["W"], d["S"])
    try:
        ctx.upload_neighbours(w.n_neigh, w.species.T, w.species_multiplicity, w.mask.T, w.rjs, w.thetas, w.phis)
        ctx.upload_dedq(np.asfortranarray(d["G"].T))
        ctx.adjoint_f32(prm, cmp_)
        q = ctx.download_soap_f32(n_soap).T
        pair = ctx.download_pairs(1).T
    finally:
        ctx.clear_basis()
    print(f"{name}: rel err q {rel(q, d['soap_contract_q']):.2e} pair {rel(pair, d['pair']):.2e}")
    assert rel(q, d["soap_contract_q"]) <= TOL
    assert rel(pair, d["pair"]) <= TOL
