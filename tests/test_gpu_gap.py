"""The device GAP energy model: energy, site energies, dE/dq and forces against the FP64 oracle.

E = sum_i delta^2 sum_t alpha_t (q_i . q_t)^zeta over the oracle's descriptors, formed in numpy
in FP64; dE/dq = delta^2 zeta sum_t alpha_t (q_i . q_t)^(zeta-1) q_t; the oracle's soap_contract
of that dE/dq gives the reference forces, which test_gap_oracle holds to central differences of
the oracle energy. Tolerance for the FP32 device path: 1e-5 relative (max |a-b| / max |b|), the
FP32 force bound.
"""
import numpy as np
import pytest

import soap_turbo_amd as st
from oracle import oracle as orc
from soap_turbo_amd.workload import make_workload
from tests.gap_cases import energy_and_grad, model
from tests.golden_cases import rel
from tests.test_gpu_adjoint import setup

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture(scope="module")
def ctx():
    return st.Context(0)


@pytest.mark.parametrize("cfg,n_sites,n_sparse,zeta", [(1, None, 64, 2), (2, None, 500, 4), (4, 400, 257, 3)])
def test_gap_energy_and_forces_match_oracle(ctx, cfg, n_sites, n_sparse, zeta):
    w = make_workload(cfg, n_sites=n_sites)
    keep, prm, cmp_, comp, n_soap, W, S, _ = setup(ctx, w)
    delta = 0.7
    qo, _, _ = orc.get_soap(w, W, S, comp, do_der=False)
    T, alpha = model(qo, n_sparse, 11)
    e_ref, G_ref = energy_and_grad(qo, T, alpha, zeta, delta)
    _, pair_o = orc.contract(w, W, S, G_ref.T, 1.0, comp)
    f_ref = np.zeros_like(w.pos)
    np.add.at(f_ref, w.nbr, -pair_o)
    v_ref = -np.einsum("ka,kb->ab", w.vec, pair_o)

    ctx.set_gap_model(T.T, alpha, zeta, delta)
    E = ctx.gap_energy_forces(prm, cmp_)
    e_site = ctx.gap_site_energies()
    f, v = ctx.download_forces(1)
    q32 = ctx.download_soap_f32(n_soap).T
    eE = abs(E - e_ref.sum()) / np.abs(e_ref).sum()
    print(f"cfg {cfg}, {n_sparse} sparse points, zeta {zeta}: E {E:.6f} (oracle {e_ref.sum():.6f}), "
          f"rel err E {eE:.1e} site {rel(e_site, e_ref):.1e} forces {rel(f.T, f_ref):.1e} virial {rel(v, v_ref):.1e}")
    assert rel(q32, qo) <= TOL
    assert eE <= TOL
    assert rel(e_site, e_ref) <= TOL
    assert rel(f.T, f_ref) <= TOL
    assert rel(v, v_ref) <= TOL
    # the descriptors the model read (the forward pass) are the ones the full pass wrote
    e32, _ = energy_and_grad(q32, T.astype(np.float32).astype(np.float64), alpha.astype(np.float32), zeta, delta)
    assert abs(E - e32.sum()) <= 1e-5 * np.abs(e32).sum()


def test_adjoint_after_gap_needs_a_new_dedq(ctx):
    w = make_workload(1)
    keep, prm, cmp_, comp, n_soap, W, S, G = setup(ctx, w)
    qo, _, _ = orc.get_soap(w, W, S, comp, do_der=False)
    T, alpha = model(qo, 32, 5)
    ctx.set_gap_model(T.T, alpha, 2, 1.0)
    ctx.gap_energy_forces(prm, cmp_)
    with pytest.raises(st.SoapTurboError):
        ctx.adjoint_f32(prm, cmp_)
    ctx.upload_dedq(np.asfortranarray(G.T))
    ctx.adjoint_f32(prm, cmp_)
    # a model of another descriptor size is refused
    ctx.set_gap_model(np.ones((n_soap + 1, 4)), np.ones(4), 2, 1.0)
    with pytest.raises(st.SoapTurboError):
        ctx.gap_energy_forces(prm, cmp_)


def test_gap_step_repeats_bitwise(ctx):
    # the site energies sum their partials in a fixed order: repeated steps agree exactly
    w = make_workload(2)
    keep, prm, cmp_, comp, n_soap, W, S, _ = setup(ctx, w)
    qo, _, _ = orc.get_soap(w, W, S, comp, do_der=False)
    T, alpha = model(qo, 300, 5)
    ctx.set_gap_model(T.T, alpha, 3, 0.5)
    E1 = ctx.gap_energy_forces(prm, cmp_)
    e1 = ctx.gap_site_energies()
    f1, v1 = ctx.download_forces(1)
    E2 = ctx.gap_energy_forces(prm, cmp_)
    e2 = ctx.gap_site_energies()
    f2, v2 = ctx.download_forces(1)
    assert E1 == E2 and np.array_equal(e1, e2) and np.array_equal(f1, f2) and np.array_equal(v1, v2)
