"""Writes the golden fixtures of tests/golden/*.npz from the reference itself.

Runs oracle/_ref/ref_harness (the reference Fortran compiled from its own
sources by ``make -C oracle ref``) on small seeded inputs and stores inputs and
outputs: W and S, the radial and angular expansions, soap without and with
derivatives, soap_cart_der and the angular expansion (every pair of the first two sites,
plus weighted checksums and per-pair norms of the whole Jacobian), and soap_contract's pair gradients for a
seeded dE/dq. Run in a container where /root/reference exists:

    python tests/golden/make_golden.py [case ...]

(no names: every case). The hess_* cases (tests/golden/hessian/) hold get_soap_central_hessian's
output for two weight vectors (src/soap_turbo.f90:1151-1700). The darby_* cases use the reference's own compression indices for the
nu_R_nu_S modes of get_compress_indices (src/soap_turbo_compress.f90:120-298).
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import refio  # noqa: E402
from soap_turbo_amd.workload import (BoxSpec, CONFIGS, SoapParams, compress_trivial, make_workload,  # noqa: E402
                                     normal)

HERE = os.path.dirname(os.path.abspath(__file__))

SMALL_BOX = BoxSpec(160, 0.08, 1.1, [0.6, 0.4], [6, 8], seed=41)

CASES = {
    # name: (workload spec, params, n_sites, legacy_seed)
    "cfg1_legacy": (CONFIGS[1][0], CONFIGS[1][1], 8, True),
    "cfg1_unified": (CONFIGS[1][0], CONFIGS[1][1], 8, False),
    "cfg2_legacy": (CONFIGS[2][0], CONFIGS[2][1], 8, True),
    "cfg2_unified": (CONFIGS[2][0], CONFIGS[2][1], 8, False),
    # scalings, amplitude_scaling /= 0, 1, central weights, radial_enhancement 2, global_scaling, nf
    "var_poly3_scalings": (SMALL_BOX, SoapParams(alpha_max=[5, 4], l_max=5, rcut_hard=[5.0, 4.6], rcut_soft=[4.2, 4.1],
                                                 atom_sigma_r=[0.45, 0.55], atom_sigma_t=[0.5, 0.4],
                                                 atom_sigma_r_scaling=[0.1, 0.05],
                                                 atom_sigma_t_scaling=[0.05, 0.08], amplitude_scaling=[2.0, 1.5],
                                                 central_weight=[0.5, 1.3], nf=[3.0, 4.5],
                                                 global_scaling=[1.0, 0.7], radial_enhancement=2, basis="poly3"),
                           6, True),
    "var_poly3gauss_amp0": (SMALL_BOX, SoapParams(alpha_max=[4, 6], l_max=3, rcut_hard=4.8, rcut_soft=4.0,
                                                  atom_sigma_r=[0.4, 0.6], atom_sigma_t=[0.45, 0.5],
                                                  atom_sigma_r_scaling=[0.08, 0.0], atom_sigma_t_scaling=0.0,
                                                  amplitude_scaling=0.0, central_weight=[0.8, 0.0],
                                                  radial_enhancement=2, basis="poly3gauss",
                                                  compress_mode="trivial"),
                            6, False),
    "var_nofilter_lmax10": (SMALL_BOX, SoapParams(alpha_max=[3, 3], l_max=10, rcut_hard=4.5, rcut_soft=4.5,
                                                  central_weight=[1.0, 0.0], basis="poly3",
                                                  scaling_mode="none"),
                            5, True),
    "var_lmax12_gauss1": (SMALL_BOX, SoapParams(alpha_max=[1, 2], l_max=12, rcut_hard=4.0, rcut_soft=3.5,
                                                basis="poly3gauss", radial_enhancement=1),
                          5, False),
}


def _darby(mode, alpha_max, l_max, basis="poly3", box=SMALL_BOX):
    return (box, SoapParams(alpha_max=alpha_max, l_max=l_max, rcut_hard=4.6, rcut_soft=4.0, basis=basis,
                            atom_sigma_r=[0.45, 0.5], atom_sigma_t=[0.5, 0.45], central_weight=[1.0, 0.6],
                            compress_mode=mode), 4, True)


CASES.update({
    "darby_0_0": _darby("0_0", [3, 3], 3),
    "darby_0_1": _darby("0_1", [3, 3], 2, "poly3gauss"),
    "darby_0_2": _darby("0_2", [3, 3], 3),
    "darby_1_0": _darby("1_0", [4, 4], 4, "poly3gauss"),
    "darby_1_1": _darby("1_1", [3, 3], 4),
    "darby_1_2": _darby("1_2", [3, 3], 3, "poly3gauss"),
    "darby_2_0": _darby("2_0", [4, 4], 3),
    "darby_2_1": _darby("2_1", [3, 3], 3),
    "darby_2_2": _darby("2_2", [4, 4], 5, "poly3gauss"),
})

# poly3operator (src/soap_turbo_radial.f90:1549-2141), with the reference's tabulated W. alpha_max >= 4:
# for smaller alpha_max the reference's soft region writes past an array sized alpha_max
# (g_aux_left_array(1:nn, 1:4, ...) into (1:nn, 1:alpha_max, ...)) and the run aborts.
CASES.update({
    "op_trivial": (SMALL_BOX, SoapParams(alpha_max=[6, 6], l_max=5, rcut_hard=4.6, rcut_soft=3.6, basis="poly3operator",
                                         atom_sigma_r=[0.45, 0.5], atom_sigma_t=[0.5, 0.45],
                                         compress_mode="trivial"), 5, True),
    "op_scalings": (SMALL_BOX, SoapParams(alpha_max=[4, 5], l_max=4, rcut_hard=[4.6, 4.2], rcut_soft=[3.6, 3.9],
                                          basis="poly3operator", atom_sigma_r=[0.45, 0.5], atom_sigma_t=[0.5, 0.45],
                                          atom_sigma_r_scaling=[0.05, 0.1], amplitude_scaling=[1.0, 2.0],
                                          central_weight=[0.5, 1.0], radial_enhancement=1, global_scaling=[1.0, 0.8]),
                    5, True),
})

CART_SITES = 2


def n_soap_of(params):
    """Descriptor length: maxval(P_i) of the compression, which is dim for every mode."""
    n_max = sum(params.alpha_max)
    L1 = params.l_max + 1
    mode = params.compress_mode
    if mode == "trivial":
        return int(compress_trivial(params.alpha_max, params.l_max)[0].max())
    if mode in (None, "", "none"):
        return n_max * (n_max + 1) // 2 * L1
    nu_R, nu_S = int(mode[0]), int(mode[2])
    a, ns = params.alpha_max[0], len(params.alpha_max)
    N1, N2 = (a if nu_R > 0 else 1), (a if nu_R == 2 else 1)
    S1, S2 = (ns if nu_S > 0 else 1), (ns if nu_S == 2 else 1)
    if nu_R % 2 == 0 and nu_S % 2 == 0:
        return N1 * S1 * (N1 * S1 + 1) // 2 * L1
    return N1 * S1 * N2 * S2 * L1


def checksums(cart):
    """Size-independent summaries of soap_cart_der (n_pairs, n_soap, 3 in C order)."""
    w = np.cos(np.arange(cart.size, dtype=np.float64) * 0.37).reshape(cart.shape)
    return np.array([cart.sum(), (cart * cart).sum(), (w * cart).sum(), np.abs(cart).max()])


def make_case(name, spec, params, n_sites, legacy):
    w = make_workload((spec, params), n_sites=n_sites)
    n_soap = n_soap_of(params)
    G = normal(1000 + n_sites, (n_soap, w.n_sites))  # C (n_soap, n_sites) == G(n_sites, n_soap)
    o = refio.run(w, G, modes=refio.MODE_DESC | refio.MODE_FULL | refio.MODE_CONTRACT | refio.MODE_EXPANSION,
                  legacy_seed=legacy)
    nc = int(np.sum(w.n_neigh[:CART_SITES]))
    ps = params.per_species()
    out = dict(
        n_neigh=w.n_neigh, species=w.species[:, 0], species_multiplicity=w.species_multiplicity,
        mask=w.mask.astype(np.int8), rjs=w.rjs, thetas=w.thetas, phis=w.phis, nbr=w.nbr,
        alpha_max=np.asarray(params.alpha_max, dtype=np.int32), l_max=params.l_max,
        radial_enhancement=params.radial_enhancement, basis=params.basis, scaling_mode=params.scaling_mode,
        compress_mode=params.compress_mode, legacy_seed=legacy, G=G,
        W=o["W"], S=o["S"], soap_desc=o["soap_desc"], soap=o["soap"], soap_contract_q=o["soap_contract_q"],
        pair=o["pair"], rad=o["rad"], rad_der=o["rad_der"], ang=o["ang"][:nc], ang_rad=o["ang_rad"][:nc],
        ang_azi=o["ang_azi"][:nc], ang_pol=o["ang_pol"][:nc], cart_head=o["soap_cart_der"][:nc], cart_checksums=checksums(o["soap_cart_der"]),
        cart_pair_norm=np.sqrt((o["soap_cart_der"] ** 2).sum(axis=(1, 2))),
        **{k: v for k, v in ps.items()})
    if "P_i" in o:
        assert int(o["P_i"].max()) == n_soap
        out.update(P_i=o["P_i"], P_j=o["P_j"], P_el=o["P_el"])
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    return w.n_sites, w.n_pairs, n_soap


# central-atom Hessian cases (tests/golden/hessian/): get_soap_central_hessian for two weight vectors
def _hess(basis, mode, alpha_max, l_max, **kw):
    return (SMALL_BOX, SoapParams(alpha_max=alpha_max, l_max=l_max, rcut_hard=4.6, rcut_soft=4.0, basis=basis,
                                  atom_sigma_r=[0.45, 0.5], atom_sigma_t=[0.5, 0.45], compress_mode=mode, **kw), 4)


HESSIAN_CASES = {
    "hess_poly3gauss": _hess("poly3gauss", "none", [4, 3], 5, atom_sigma_r_scaling=[0.05, 0.0],
                             atom_sigma_t_scaling=[0.02, 0.05], central_weight=[1.0, 0.5]),
    "hess_poly3_trivial": _hess("poly3", "trivial", [5, 4], 4, radial_enhancement=2, amplitude_scaling=[1.5, 0.0],
                                central_weight=[0.7, 1.0]),
    "hess_darby_1_1": _hess("poly3gauss", "1_1", [3, 3], 6, radial_enhancement=1),
}


def make_hessian_case(name, spec, params, n_sites, n_vec=2):
    w = make_workload((spec, params), n_sites=n_sites)
    n_soap = n_soap_of(params)
    vecs = normal(3000 + n_sites, (w.n_sites, n_vec, n_soap))
    o = refio.run(w, None, modes=refio.MODE_HESSIAN, legacy_seed=False, vecs=vecs)
    ps = params.per_species()
    out = dict(
        n_neigh=w.n_neigh, species=w.species[:, 0], species_multiplicity=w.species_multiplicity,
        mask=w.mask.astype(np.int8), rjs=w.rjs, thetas=w.thetas, phis=w.phis, nbr=w.nbr,
        alpha_max=np.asarray(params.alpha_max, dtype=np.int32), l_max=params.l_max,
        radial_enhancement=params.radial_enhancement, basis=params.basis, scaling_mode=params.scaling_mode,
        compress_mode=params.compress_mode, legacy_seed=False, vecs=vecs, W=o["W"], S=o["S"],
        soap=o["soap_hess"], hess=o["hess"], rad_der2=o["rad_der2"], selftest_err=o["hess_selftest_err"],
        **{k: v for k, v in ps.items()})
    if "P_i" in o:
        out.update(P_i=o["P_i"], P_j=o["P_j"], P_el=o["P_el"])
    os.makedirs(os.path.join(HERE, "hessian"), exist_ok=True)
    np.savez_compressed(os.path.join(HERE, "hessian", name + ".npz"), **out)
    return w.n_sites, w.n_pairs, n_soap


if __name__ == "__main__":
    if not refio.available():
        raise SystemExit("build the reference harness first: make -C oracle ref")
    for name in sys.argv[1:] or [*CASES, *HESSIAN_CASES]:
        if name in HESSIAN_CASES:
            print(name, *make_hessian_case(name, *HESSIAN_CASES[name]))
            continue
        spec, params, ns, legacy = CASES[name]
        print(name, *make_case(name, spec, params, ns, legacy))
