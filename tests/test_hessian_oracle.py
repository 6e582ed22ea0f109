"""The oracle's central-atom Hessian (get_soap_central_hessian, src/soap_turbo.f90:1151-1700).

Pinned three ways:
* against fixtures the reference itself wrote (tests/golden/hessian/, the compiled harness):
  Hessian blocks, descriptors and the radial second derivatives;
* against central differences of the oracle's own first derivatives over atom positions: the
  block of pair (i, j) is d/dr_jb of sum_d w_d dq_d/dr_ia, and the self pair's block is the
  same with the central atom moved (the neighbour list rebuilt at each displaced position);
* the radial second derivatives against central differences of the first derivatives in r.
A wrong sign, a dropped term of the quotient rule or a transposed (a, b) fails one of them.
"""
import glob
import os

import numpy as np
import pytest

from oracle import oracle as orc
from soap_turbo_amd.workload import BoxSpec, SoapParams, compress_trivial, make_workload, normal
from tests.gap_cases import displaced
from tests.golden_cases import GOLDEN, load, rel

HESS_CASES = sorted(os.path.splitext(os.path.basename(p))[0]
                    for p in glob.glob(os.path.join(GOLDEN, "hessian", "*.npz")))


def test_fixtures_present():
    assert len(HESS_CASES) >= 3


@pytest.mark.parametrize("name", HESS_CASES)
def test_oracle_hessian_matches_reference_fixtures(name):
    w, comp, d = load(os.path.join("hessian", name))
    q, h, r2, err = orc.central_hessian(w, d["W"], d["S"], d["vecs"], comp)
    print(f"{name}: hess {rel(h, d['hess']):.1e} soap {rel(q, d['soap']):.1e} rad_der2 {rel(r2, d['rad_der2']):.1e}")
    assert rel(h, d["hess"]) <= 1e-12
    assert rel(q, d["soap"]) <= 1e-12
    assert rel(r2, d["rad_der2"]) <= 1e-12
    assert err <= 1e-12 and float(d["selftest_err"]) <= 1e-12
    # the blocks of one site's row sum to zero (a rigid shift leaves the descriptor alone)
    k3 = 0
    for n in w.n_neigh:
        assert np.abs(h[k3:k3 + n].sum(axis=0)).max() <= 1e-12 * np.abs(h[k3:k3 + n]).max()
        k3 += n


BOX = BoxSpec(120, 0.07, 1.2, [0.5, 0.5], [1, 8], seed=17)
PARAMS = {
    "poly3gauss": SoapParams(alpha_max=[4, 3], l_max=4, rcut_hard=4.2, rcut_soft=3.6, basis="poly3gauss",
                             atom_sigma_r=[0.45, 0.5], atom_sigma_t=[0.5, 0.45], atom_sigma_r_scaling=[0.05, 0.02],
                             atom_sigma_t_scaling=[0.03, 0.0]),
    "poly3_trivial": SoapParams(alpha_max=[4, 4], l_max=3, rcut_hard=4.2, rcut_soft=3.6, basis="poly3",
                                atom_sigma_r=0.5, atom_sigma_t=0.5, radial_enhancement=1, central_weight=[0.6, 1.0],
                                compress_mode="trivial"),
}


@pytest.mark.parametrize("kind", list(PARAMS))
def test_oracle_hessian_is_the_derivative_of_the_first_derivative(kind):
    p = PARAMS[kind]
    w = make_workload((BOX, p), n_sites=6)
    W, S = orc.basis_matrices(w)
    comp = compress_trivial(p.alpha_max, p.l_max) if p.compress_mode == "trivial" else None
    q0, _, _ = orc.get_soap(w, W, S, comp, do_der=False, legacy_seed=False)
    vecs = normal(21, (w.n_sites, 1, q0.shape[1]))
    _, hess, _, err = orc.central_hessian(w, W, S, vecs, comp)
    assert err <= 1e-12
    i = 2
    k3 = int(w.n_neigh[:i].sum())

    def first(pos):
        """sum_d w_d dq_d/dr_i (3,) of site i at these positions."""
        wd = displaced(w, pos)
        _, cart, _ = orc.get_soap(wd, W, S, comp, do_der=True, legacy_seed=False)
        return cart[int(wd.n_neigh[:i].sum())].T @ vecs[i, 0]

    h = 1e-4
    scale = np.abs(hess[k3:k3 + w.n_neigh[i], 0]).max()
    for k2 in (k3, k3 + 1, k3 + 3):          # the self pair, then two neighbours
        atom = w.site_atom[i] if k2 == k3 else w.nbr[k2]
        for b in range(3):
            pp, pm = w.pos.copy(), w.pos.copy()
            pp[atom, b] += h
            pm[atom, b] -= h
            fd = (first(pp) - first(pm)) / (2 * h)
            np.testing.assert_allclose(hess[k2, 0, :, b], fd, rtol=0, atol=2e-6 * scale,
                                       err_msg=f"pair {k2}, b {b}")


def test_radial_second_derivative_is_the_derivative_of_the_first():
    w = make_workload((BOX, PARAMS["poly3gauss"]), n_sites=4)
    W, _ = orc.basis_matrices(w)
    _, d1, d2 = orc.radial_der2(w, W)
    h = 1e-5
    wp, wm = displaced(w, w.pos), displaced(w, w.pos)
    wp.rjs = w.rjs + h
    wm.rjs = w.rjs - h
    _, d1p, _ = orc.radial_der2(wp, W)
    _, d1m, _ = orc.radial_der2(wm, W)
    fd = (d1p - d1m) / (2 * h)
    inside = w.rjs < 4.2 - 2 * h
    assert rel(d2[inside], fd[inside]) <= 1e-7
