"""GAP-form energy model in FP64 numpy for the tests: E = sum_i delta^2 sum_t alpha_t (q_i . q_t)^zeta."""
import numpy as np

from soap_turbo_amd.workload import build_neighbours, spherical


def model(qo, n_sparse, seed):
    """Sparse points near the descriptors (normalised, as GAP's are) and signed weights."""
    rng = np.random.default_rng(seed)
    T = qo[rng.integers(0, len(qo), n_sparse)] + 0.05 * rng.standard_normal((n_sparse, qo.shape[1]))
    T /= np.linalg.norm(T, axis=1, keepdims=True)
    alpha = rng.standard_normal(n_sparse)
    return T, alpha


def energy_and_grad(q, T, alpha, zeta, delta):
    K = q @ T.T
    e_site = delta**2 * (K**zeta) @ alpha
    G = delta**2 * zeta * (K ** (zeta - 1) * alpha) @ T
    return e_site, G


def displaced(w, pos):
    """The workload's sites and species at new atom positions (neighbour list rebuilt)."""
    rcut = float(np.max(w.params.per_species()["rcut_hard"]))
    n_neigh, nbr, vec = build_neighbours(pos, w.L, rcut, w.site_atom)
    rjs, th, ph = spherical(vec)
    mask = np.zeros((w.params.n_species, len(nbr)), dtype=bool)
    for s in range(w.params.n_species):
        mask[s] = w.atom_species[nbr] == s
    return type(w)(w.params, w.L, pos, w.atom_species, w.site_atom, n_neigh, nbr, vec, rjs, th, ph, mask, w.species,
                   w.species_multiplicity)
