"""Loads the golden fixtures (tests/golden/*.npz) back into workloads; shared test helpers."""
from __future__ import annotations

import glob
import os

import numpy as np

from soap_turbo_amd.workload import SoapParams, Workload

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
CASES = sorted(os.path.splitext(os.path.basename(p))[0] for p in glob.glob(os.path.join(GOLDEN, "*.npz")))


def load(name):
    d = dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))
    params = SoapParams(alpha_max=[int(a) for a in d["alpha_max"]], l_max=int(d["l_max"]),
                        rcut_hard=d["rcut_hard"], rcut_soft=d["rcut_soft"], atom_sigma_r=d["atom_sigma_r"],
                        atom_sigma_t=d["atom_sigma_t"], atom_sigma_r_scaling=d["atom_sigma_r_scaling"],
                        atom_sigma_t_scaling=d["atom_sigma_t_scaling"], amplitude_scaling=d["amplitude_scaling"],
                        central_weight=d["central_weight"], nf=d["nf"], global_scaling=d["global_scaling"],
                        basis=str(d["basis"]), scaling_mode=str(d["scaling_mode"]),
                        radial_enhancement=int(d["radial_enhancement"]), compress_mode=str(d["compress_mode"]))
    n_sites = len(d["n_neigh"])
    w = Workload(params, 0.0, None, None, None, d["n_neigh"].astype(np.int32), d["nbr"], None, d["rjs"], d["thetas"],
                 d["phis"], d["mask"].astype(bool), d["species"].astype(np.int32).reshape(n_sites, 1),
                 d["species_multiplicity"].astype(np.int32))
    comp = (d["P_i"], d["P_j"], d["P_el"]) if "P_i" in d else None
    return w, comp, d


def rel(a, b):
    """max |a - b| / max |b| over the array."""
    a = np.asarray(a); b = np.asarray(b)
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / (den if den > 0 else 1.0))


def checksums(cart):
    w = np.cos(np.arange(cart.size, dtype=np.float64) * 0.37).reshape(cart.shape)
    return np.array([cart.sum(), (cart * cart).sum(), (w * cart).sum(), np.abs(cart).max()])


def compression_indices(kind, n_unc, seed):
    """Compression indices (1-based, the reference's P_i/P_j/P_el) of three kinds: None, one entry per
    output with a weight ("scaled"), and outputs summing one to three weighted components ("multi")."""
    if kind is None:
        return None
    rng = np.random.default_rng(seed)
    if kind == "scaled":
        cols = np.sort(rng.choice(n_unc, size=(2 * n_unc) // 3, replace=False))
        P_j = (cols + 1).astype(np.int32)
        P_i = np.arange(1, len(cols) + 1, dtype=np.int32)
        return P_i, P_j, rng.uniform(0.25, 2.0, len(cols))
    P_i, P_j = [], []
    out = 0
    for c in rng.permutation(n_unc)[: (3 * n_unc) // 4]:
        if not P_i or rng.random() < 0.5:
            out += 1
        P_i.append(out)
        P_j.append(c + 1)
    return np.array(P_i, np.int32), np.array(P_j, np.int32), rng.uniform(-1.5, 1.5, len(P_i))

