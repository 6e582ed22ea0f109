"""The GAP model's forces on the oracle: minus the gradient of its energy over atom positions.

Central differences of E = sum_i delta^2 sum_t alpha_t (q_i . q_t)^zeta, with the oracle's
descriptors at displaced positions (the neighbour list rebuilt each time), against the forces the
oracle's soap_contract gives for dE/dq = delta^2 zeta sum_t alpha_t (q_i . q_t)^(zeta-1) q_t. A
wrong sign, index or dropped factor anywhere in that chain fails.
"""
import numpy as np

from oracle import oracle as orc
from soap_turbo_amd.workload import compress_trivial, make_workload
from tests.gap_cases import displaced, energy_and_grad, model


def test_oracle_gap_forces_are_minus_the_energy_gradient():
    """Central differences of the oracle energy over atom positions pin the dE/dq -> forces chain."""
    w = make_workload(1)
    p = w.params
    W, S = orc.basis_matrices(w)
    comp = compress_trivial(p.alpha_max, p.l_max) if p.compress_mode == "trivial" else None
    qo, _, _ = orc.get_soap(w, W, S, comp, do_der=False)
    T, alpha = model(qo, 48, 3)
    zeta, delta = 3, 0.9
    _, G = energy_and_grad(qo, T, alpha, zeta, delta)
    _, pair = orc.contract(w, W, S, G.T, 1.0, comp)
    f = np.zeros_like(w.pos)
    np.add.at(f, w.nbr, -pair)

    def E(pos):
        wd = displaced(w, pos)
        q, _, _ = orc.get_soap(wd, W, S, comp, do_der=False)
        return energy_and_grad(q, T, alpha, zeta, delta)[0].sum()

    h = 1e-4
    for a in (0, 37, 150):
        for c in range(3):
            pp, pm = w.pos.copy(), w.pos.copy()
            pp[a, c] += h
            pm[a, c] -= h
            fd = -(E(pp) - E(pm)) / (2 * h)
            assert abs(fd - f[a, c]) <= 1e-6 * np.abs(f).max(), (a, c, fd, f[a, c])
