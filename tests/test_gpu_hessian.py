"""The device central-atom Hessian (st_central_hessian) against the reference and the oracle.

FP64 bar: 1e-10 relative (max |a-b| / max |b|) against fixtures the reference wrote
(tests/golden/hessian/, with the reference's W and S) and against the oracle on a whole
configuration. The oracle itself is pinned to the reference and to finite differences in
tests/test_hessian_oracle.py.
"""
import glob
import os

import numpy as np
import pytest

import soap_turbo_amd as st
from oracle import oracle as orc
from soap_turbo_amd.workload import BoxSpec, SoapParams, make_workload, normal
from tests.golden_cases import GOLDEN, load, rel

pytestmark = pytest.mark.gpu

TOL = 1e-10
HESS_CASES = sorted(os.path.splitext(os.path.basename(p))[0]
                    for p in glob.glob(os.path.join(GOLDEN, "hessian", "*.npz")))


@pytest.fixture(scope="module")
def ctx():
    return st.Context(0)


def _params(keep, w, legacy=False):
    p = w.params
    return st.make_params(keep, alpha_max=p.alpha_max, l_max=p.l_max, radial_enhancement=p.radial_enhancement,
                          basis=p.basis, scaling_mode=p.scaling_mode, legacy_seed=legacy, **p.per_species())


def _run(ctx, w, comp, vecs_c, legacy=False):
    """vecs_c (n_sites, n_vec, n_soap) C order; returns hess (n_pairs, n_vec, a, b)."""
    keep = st._lib.Keep()
    prm = _params(keep, w, legacy)
    cmp_ = st.make_compression(keep, comp is not None, len(comp[0]) if comp else 0, *(comp or (None, None, None)))
    ctx.upload_neighbours(w.n_neigh, w.species.T, w.species_multiplicity, w.mask.T, w.rjs, w.thetas, w.phis)
    h = ctx.central_hessian(prm, cmp_, np.asarray(vecs_c).transpose(2, 1, 0))
    return h.transpose(3, 2, 0, 1)


@pytest.mark.parametrize("name", HESS_CASES)
def test_hessian_matches_reference_fixtures(ctx, name):
    w, comp, d = load(os.path.join("hessian", name))
    ctx.set_basis(d["W"], d["S"])
    try:
        h = _run(ctx, w, comp, d["vecs"])
    finally:
        ctx.clear_basis()
    print(f"{name}: rel err {rel(h, d['hess']):.2e}")
    assert rel(h, d["hess"]) <= TOL


def _check_against_oracle(ctx, w, seed, label):
    p = w.params
    keep = st._lib.Keep()
    prm = _params(keep, w)
    n_max = sum(p.alpha_max)
    W, S = ctx.get_basis(prm, n_max)
    comp = None
    if p.compress_mode == "trivial":
        dim, nz = st.get_compress_indices("trivial", p.alpha_max, p.l_max, "get_dim")
        comp = (np.zeros(nz, np.int32), np.zeros(nz, np.int32), np.zeros(nz))
        st.get_compress_indices("trivial", p.alpha_max, p.l_max, "set_indices", *comp)
    n_soap = int(comp[0].max()) if comp else n_max * (n_max + 1) // 2 * (p.l_max + 1)
    vecs = normal(seed, (w.n_sites, 2, n_soap))
    h = _run(ctx, w, comp, vecs)
    _, ho, _, err = orc.central_hessian(w, W, S, vecs, comp)
    print(f"{label}: {w.n_sites} sites, {w.n_pairs} pairs, rel err {rel(h, ho):.2e}, oracle self-test {err:.1e}")
    assert rel(h, ho) <= TOL
    # every site's row of blocks sums to zero
    k3 = 0
    for n in w.n_neigh[:50]:
        assert np.abs(h[k3:k3 + n].sum(axis=0)).max() <= 1e-10 * np.abs(h[k3:k3 + n]).max()
        k3 += n


@pytest.mark.parametrize("cfg,n_sites", [(1, None), (4, 150)])
def test_hessian_matches_oracle_on_a_configuration(ctx, cfg, n_sites):
    _check_against_oracle(ctx, make_workload(cfg, n_sites=n_sites), 31 + cfg, f"config {cfg}")


# Shapes where the radial index runs past the lanes that hold an angular index: n_max 24 against
# k_max 10 (l_max 3), and against the 14 k of the second pass of 64 at l_max 11 (k_max 78).
@pytest.mark.parametrize("l_max", [3, 11])
def test_hessian_with_more_radial_than_angular_channels(ctx, l_max):
    box = BoxSpec(432, 0.09, 1.0, [0.5, 0.3, 0.2], [1, 6, 8], seed=12, central=1)
    prm = SoapParams(alpha_max=[8, 8, 8], l_max=l_max, rcut_hard=4.5, rcut_soft=4.0, basis="poly3gauss",
                     compress_mode="trivial")
    _check_against_oracle(ctx, make_workload((box, prm), n_sites=10), 40 + l_max, f"l_max {l_max}, n_max 24")


def test_hessian_refuses_the_legacy_filter_seed(ctx):
    w, comp, d = load(os.path.join("hessian", HESS_CASES[0]))
    with pytest.raises(st.SoapTurboStop):
        _run(ctx, w, comp, d["vecs"], legacy=True)


def test_hessian_refuses_poly3operator(ctx):
    # the reference stops: the poly3operator basis has no second radial derivatives
    w, comp, d = load("op_scalings")
    with pytest.raises(st.SoapTurboStop, match="poly3operator"):
        _run(ctx, w, comp, np.zeros((w.n_sites, 1, 10)))


@pytest.mark.parametrize("cfg,n_sites", [(1, 60), (2, 80)])
def test_hessian_selftest_rebuilds_first_derivatives(ctx, cfg, n_sites):
    # soap_hessian_selftest (src/soap_turbo.f90:1419-1451): the first derivatives the Hessian's
    # contraction also produces, against the independent FP64 Jacobian; with vectors other than the
    # ones the Hessian ran with, the two routes must disagree (the check compares something)
    w = make_workload(cfg, n_sites=n_sites)
    p = w.params
    keep = st._lib.Keep()
    prm = _params(keep, w)
    comp = None
    if p.compress_mode == "trivial":
        dim, nz = st.get_compress_indices("trivial", p.alpha_max, p.l_max, "get_dim")
        comp = (np.zeros(nz, np.int32), np.zeros(nz, np.int32), np.zeros(nz))
        st.get_compress_indices("trivial", p.alpha_max, p.l_max, "set_indices", *comp)
    cmp_ = st.make_compression(keep, comp is not None, len(comp[0]) if comp else 0, *(comp or (None, None, None)))
    n_max = sum(p.alpha_max)
    n_soap = int(comp[0].max()) if comp else n_max * (n_max + 1) // 2 * (p.l_max + 1)
    vecs = np.asfortranarray(normal(61, (n_soap, 2, w.n_sites)))
    ctx.upload_neighbours(w.n_neigh, w.species.T, w.species_multiplicity, w.mask.T, w.rjs, w.thetas, w.phis)
    ctx.central_hessian(prm, cmp_, vecs)
    err = ctx.hessian_selftest(prm, cmp_, vecs)
    other = ctx.hessian_selftest(prm, cmp_, np.asfortranarray(normal(62, vecs.shape)))
    print(f"config {cfg}: self-test err {err:.2e}, with other vectors {other:.2e}")
    assert err <= 1e-10
    assert other >= 1e-2
