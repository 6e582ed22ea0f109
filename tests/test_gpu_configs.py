"""Configurations 3, 4 and 5 on the GPU: parity with the FP64 oracle on site samples cut from each
full box, and size-independent checks on the full boxes.

* Site samples keep every neighbour of a kept site, so their descriptors and pair gradients are those
  of the full box. Config 4's sample takes sites from both phases of the slab (gas sites have 0-10
  neighbours), config 5's from the whole box (two 64-neighbour tiles per site, n_soap 810).
* Full boxes: the FP32 adjoint forces and virial against the batched FP64 full-Jacobian path on the
  same box (every site, every pair), forces summing to zero, and bitwise repeats.
* Config 3's central differences: forces = -dE/dx of E = sum_i G_i . q_i on 64 seeded atoms of the
  32768-atom box (alpha_max 8, l_max 8, rcut 5.0 / 4.5, poly3, unified filter seed), step 1e-4 A,
  with the oracle's FP64 energies over the sites each moved atom touches.

Tolerances: 1e-5 relative (max |a - b| / max |b| over each array) for the FP32 path, 1e-10 for the
FP64 path against the oracle (both use the library's W).
"""
import numpy as np
import pytest

import soap_turbo_amd as st
from oracle import oracle as orc
from soap_turbo_amd.workload import (CONFIGS, Workload, make_workload, normal, slab_sites, spherical,
                                     strided_sites)
from tests.golden_cases import rel

pytestmark = pytest.mark.gpu

TOL32 = 1e-5
TOL64 = 1e-10


@pytest.fixture(scope="module")
def ctx():
    return st.Context(0)


def setup(ctx, w, legacy=True, seed=5):
    p = w.params
    keep = st._lib.Keep()
    prm = st.make_params(keep, alpha_max=p.alpha_max, l_max=p.l_max, radial_enhancement=p.radial_enhancement,
                         basis=p.basis, scaling_mode=p.scaling_mode, legacy_seed=legacy, **p.per_species())
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


def oracle_forces(w, pair):
    f = np.zeros_like(w.pos)
    np.add.at(f, w.nbr, -pair)
    return f, -np.einsum("ka,kb->ab", w.vec, pair)


SAMPLES = {
    3: lambda: strided_sites(32768, 512, offset=7),
    4: lambda: slab_sites(4, 300, 300),
    5: lambda: strided_sites(262144, 256, offset=11),
}


@pytest.mark.parametrize("cfg", [3, 4, 5])
def test_site_sample_fp32_and_fp64_match_oracle(ctx, cfg):
    w = make_workload(cfg, sites=SAMPLES[cfg]())
    nb = w.n_neigh - 1
    print(f"config {cfg}: {w.n_sites} sampled sites, {w.n_pairs} pairs, neighbours {nb.min()}..{nb.max()}")
    if cfg == 4:
        # both phases, including isolated gas sites and sites with a few neighbours
        assert (nb <= 10).sum() >= 200 and (nb >= 60).sum() >= 200 and nb.min() == 0
    if cfg == 5:
        assert nb.min() > 64  # every site has two tiles of neighbours
    keep, prm, cmp_, comp, n_soap, W, S, G = setup(ctx, w)
    qo, pair_o = orc.contract(w, W, S, G, 1.0, comp, True)
    fo, vo = oracle_forces(w, pair_o)
    # FP32 adjoint
    ctx.adjoint_f32(prm, cmp_)
    q32 = ctx.download_soap_f32(n_soap).T
    pair32 = ctx.download_pairs(1).T
    f32, v32 = ctx.download_forces(1)
    e32 = [rel(q32, qo), rel(pair32, pair_o), rel(f32.T, fo), rel(v32, vo)]
    # FP64 full Jacobian (two batches)
    nbat = ctx.full_jacobian_forces(prm, cmp_, max_pairs=(w.n_pairs + 1) // 2 + int(w.n_neigh.max()))
    q64, _ = ctx.download(n_soap)
    pair64 = ctx.download_pairs(0).T
    f64, v64 = ctx.download_forces(0)
    e64 = [rel(q64.T, qo), rel(pair64, pair_o), rel(f64.T, fo), rel(v64, vo)]
    print(f"config {cfg} FP32: q {e32[0]:.2e} pair {e32[1]:.2e} forces {e32[2]:.2e} virial {e32[3]:.2e}")
    print(f"config {cfg} FP64 ({nbat} batches): q {e64[0]:.2e} pair {e64[1]:.2e} forces {e64[2]:.2e} "
          f"virial {e64[3]:.2e}")
    assert nbat == 2
    assert max(e32) <= TOL32
    assert max(e64) <= TOL64


def test_config4_gas_sites_fp64_jacobian_matches_oracle(ctx):
    # soap_cart_der itself for sites of both phases, in one launch with uneven neighbour counts
    w = make_workload(4, sites=slab_sites(4, 40, 80))
    keep, prm, cmp_, comp, n_soap, W, S, G = setup(ctx, w)
    ctx.compute(prm, cmp_, True, False)
    soap, cart = ctx.download(n_soap, cart=True)
    qo, co, _ = orc.get_soap(w, W, S, comp, do_der=True)
    eq, ec = rel(soap.T, qo), rel(cart.transpose(2, 1, 0), co)
    print(f"config 4 (40 slab + 80 gas sites): soap {eq:.2e} soap_cart_der {ec:.2e}")
    assert eq <= TOL64 and ec <= TOL64


@pytest.mark.parametrize("cfg", [3, 4, 5])
def test_full_box_fp32_matches_batched_fp64(ctx, cfg):
    # every site of the full box through both GPU paths; the FP64 path runs in site batches
    w = make_workload(cfg)
    keep, prm, cmp_, comp, n_soap, W, S, G = setup(ctx, w)
    ctx.adjoint_f32(prm, cmp_)
    f32, v32 = ctx.download_forces(1)
    q32 = ctx.download_soap_f32(n_soap)
    ctx.adjoint_f32(prm, cmp_)
    f32b, v32b = ctx.download_forces(1)
    assert np.array_equal(f32, f32b) and np.array_equal(v32, v32b), "repeat runs differ"
    scale = np.abs(f32).max()
    assert np.abs(f32.sum(axis=1)).max() <= 1e-6 * scale * np.sqrt(f32.shape[1])
    max_pairs = 0 if cfg == 5 else w.n_pairs // 3 + int(w.n_neigh.max())
    nbat = ctx.full_jacobian_forces(prm, cmp_, max_pairs=max_pairs)
    f64, v64 = ctx.download_forces(0)
    q64, _ = ctx.download(n_soap)
    ef, ev, eq = rel(f32, f64), rel(v32, v64), rel(q32, q64)
    print(f"config {cfg} full box ({w.n_sites} sites, {w.n_pairs} pairs, FP64 in {nbat} batches): "
          f"FP32 vs FP64 q {eq:.2e} forces {ef:.2e} virial {ev:.2e}")
    assert nbat >= 2
    assert ef <= TOL32 and ev <= TOL32 and eq <= TOL32


def test_batched_fp64_equals_one_pass(ctx):
    # batches change nothing but which sites run together: bitwise equal to the one-pass path
    w = make_workload(2)
    keep, prm, cmp_, comp, n_soap, W, S, G = setup(ctx, w)
    ctx.compute(prm, cmp_, True, False)
    ctx.pairs_full_jacobian()
    pair1 = ctx.download_pairs(0)
    f1, v1 = ctx.download_forces(0)
    q1, _ = ctx.download(n_soap)
    for max_pairs in (5000, 20000, 0):
        nbat = ctx.full_jacobian_forces(prm, cmp_, max_pairs=max_pairs)
        pair2 = ctx.download_pairs(0)
        f2, v2 = ctx.download_forces(0)
        q2, _ = ctx.download(n_soap)
        print(f"max_pairs {max_pairs}: {nbat} batches")
        assert (nbat > 1) == (max_pairs > 0)
        assert np.array_equal(pair1, pair2) and np.array_equal(f1, f2) and np.array_equal(v1, v2)
        assert np.array_equal(q1, q2)


def test_batched_fp64_refuses_a_batch_below_one_site(ctx):
    w = make_workload(1)
    keep, prm, cmp_, comp, n_soap, W, S, G = setup(ctx, w)
    with pytest.raises(st.SoapTurboError):
        ctx.full_jacobian_forces(prm, cmp_, max_pairs=int(w.n_neigh.max()) - 1)


def test_config3_central_differences(ctx):
    # forces on 64 seeded atoms of the full config-3 box against central differences of the FP64
    # oracle energy; each moved atom changes only the sites whose neighbour lists hold it
    spec, par = CONFIGS[3]
    w = make_workload(3)
    keep, prm, cmp_, comp, n_soap, W, S, G = setup(ctx, w, legacy=False, seed=33)
    ctx.adjoint_f32(prm, cmp_)
    f32, _ = ctx.download_forces(1)
    rng = np.random.default_rng(3)
    atoms = rng.choice(len(w.pos), 64, replace=False)
    ptr = np.concatenate([[0], np.cumsum(w.n_neigh)])
    site_of_pair = np.repeat(np.arange(w.n_sites), w.n_neigh)
    scale = np.abs(f32).max()
    worst = 0.0
    h = 1e-4
    for a in atoms:
        sites = np.unique(site_of_pair[w.nbr == a])   # the sites that see atom a (a's own site included)
        idx = np.concatenate([np.arange(ptr[s], ptr[s + 1]) for s in sites])
        nbr, own = w.nbr[idx], w.site_atom[site_of_pair[idx]]
        sub = Workload(par, w.L, w.pos, w.atom_species, w.site_atom[sites], w.n_neigh[sites], nbr, w.vec[idx],
                       w.rjs[idx], w.thetas[idx], w.phis[idx], w.mask[:, idx], w.species[sites],
                       w.species_multiplicity[sites])
        Gs = np.ascontiguousarray(G[:, sites])

        def energy(pos):
            d = pos[nbr] - pos[own]
            d -= w.L * np.round(d / w.L)
            d[nbr == own] = 0.0
            sub.rjs, sub.thetas, sub.phis = spherical(d)
            q, _, _ = orc.get_soap(sub, W, S, do_der=False, legacy_seed=False)
            return float(np.sum(q * Gs.T))

        for c in range(3):
            pp = w.pos.copy(); pp[a, c] += h
            pm = w.pos.copy(); pm[a, c] -= h
            fd = -(energy(pp) - energy(pm)) / (2 * h)
            worst = max(worst, abs(fd - f32[c, a]) / scale)
    print(f"config 3 central differences on 64 atoms: max rel err {worst:.2e}")
    assert worst <= TOL32
