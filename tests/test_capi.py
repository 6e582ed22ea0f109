"""The C-ABI library: it loads, exports every symbol include/*.h declares, and its host-only logic."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

import soap_turbo_amd as st
from soap_turbo_amd import _lib
from tests.golden_cases import CASES, load

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"\b(st_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(_lib.SIGNATURES)


@pytest.mark.parametrize("name", [c for c in CASES if "cfg2" in c or "gauss_amp0" in c])
def test_compress_indices_match_reference(name):
    _, comp, d = load(name)
    dim, nz = st.get_compress_indices("trivial", d["alpha_max"], int(d["l_max"]), "get_dim")
    assert dim == nz == len(d["P_i"])
    P_i = np.zeros(nz, dtype=np.int32); P_j = np.zeros(nz, dtype=np.int32); P_el = np.zeros(nz)
    st.get_compress_indices("Trivial", d["alpha_max"], int(d["l_max"]), "set_indices", P_i, P_j, P_el)
    np.testing.assert_array_equal(P_i, d["P_i"])
    np.testing.assert_array_equal(P_j, d["P_j"])
    np.testing.assert_array_equal(P_el, d["P_el"])


def test_compress_indices_brute_force_small():
    # trivial compression keeps (n, m >= n) when n or m is the first radial index of a species
    for am, L in (([2, 3], 1), ([1, 1, 2], 2), ([4], 0)):
        dim, nz = st.get_compress_indices("trivial", am, L, "get_dim")
        P_i = np.zeros(nz, dtype=np.int32); P_j = np.zeros(nz, dtype=np.int32); P_el = np.zeros(nz)
        st.get_compress_indices("trivial", am, L, "set_indices", P_i, P_j, P_el)
        piv = set(np.cumsum([0] + am[:-1]) + 1)
        n_max = sum(am)
        expect = [k for k, (n, m, l) in enumerate(
            (n, m, l) for n in range(1, n_max + 1) for m in range(n, n_max + 1) for l in range(L + 1))
            if n in piv or m in piv]
        np.testing.assert_array_equal(P_j - 1, expect)
        np.testing.assert_array_equal(P_i, np.arange(1, nz + 1))


@pytest.mark.parametrize("name", [c for c in CASES if c.startswith("darby_")])
def test_darby_compress_indices_match_reference(name):
    # P_i, P_j, P_el of the nu_R_nu_S modes as the reference's get_compress_indices wrote them
    _, comp, d = load(name)
    mode = str(d["compress_mode"])
    dim, nz = st.get_compress_indices(mode, d["alpha_max"], int(d["l_max"]), "get_dim")
    assert nz == len(d["P_i"]) and dim == int(d["P_i"].max())
    P_i = np.zeros(nz, dtype=np.int32); P_j = np.zeros(nz, dtype=np.int32); P_el = np.zeros(nz)
    st.get_compress_indices(mode + "  ", d["alpha_max"], int(d["l_max"]), "set_indices", P_i, P_j, P_el)
    np.testing.assert_array_equal(P_i, d["P_i"])
    np.testing.assert_array_equal(P_j, d["P_j"])
    np.testing.assert_array_equal(P_el, d["P_el"])


def _indices(mode, am, L):
    dim, nz = st.get_compress_indices(mode, am, L, "get_dim")
    P_i = np.zeros(nz, dtype=np.int32); P_j = np.zeros(nz, dtype=np.int32); P_el = np.zeros(nz)
    st.get_compress_indices(mode, am, L, "set_indices", P_i, P_j, P_el)
    return dim, P_i, P_j, P_el


@pytest.mark.parametrize("am,L", [([3, 3], 2), ([2, 2, 2], 1), ([4], 3)])
def test_darby_modes_closed_forms(am, L):
    n_max, L1 = sum(am), L + 1
    n_unc = n_max * (n_max + 1) // 2 * L1
    comps = [(n, m, l) for n in range(n_max) for m in range(n, n_max) for l in range(L1)]
    # 2_2 keeps every component once, in place
    dim, P_i, P_j, P_el = _indices("2_2", am, L)
    assert dim == n_unc
    np.testing.assert_array_equal(P_i, np.arange(1, n_unc + 1))
    np.testing.assert_array_equal(P_j, np.arange(1, n_unc + 1))
    np.testing.assert_array_equal(P_el, np.ones(n_unc))
    # 0_0 sums every component of one l into channel l: weight sqrt 2 off the diagonal
    dim, P_i, P_j, P_el = _indices("0_0", am, L)
    assert dim == L1
    np.testing.assert_array_equal(P_j, np.arange(1, n_unc + 1))
    np.testing.assert_array_equal(P_i, [l + 1 for n, m, l in comps])
    np.testing.assert_array_equal(P_el, [1.0 if n == m else np.sqrt(2.0) for n, m, l in comps])
    # every mode is a sum over the full, ordered power spectrum p(a, b, l) with the radial and species
    # channels collapsed: a -> (z_a mod S1) N1 + (n_a mod N1) first, b -> (z_b mod S2) N2 + (n_b mod N2)
    # second; symmetric modes fold (I, J) onto I <= J with 1/sqrt 2 off the diagonal. u is the stored
    # (uncompressed) spectrum, sqrt 2 p off the diagonal.
    rng = np.random.default_rng(5)
    p = rng.standard_normal((n_max, n_max, L1))
    p = p + p.transpose(1, 0, 2)
    u = np.array([p[n, m, l] * (1.0 if n == m else np.sqrt(2.0)) for n, m, l in comps])
    a, ns = am[0], len(am)
    for mode in ("0_1", "0_2", "1_0", "1_1", "1_2", "2_0", "2_1"):
        nu_R, nu_S = int(mode[0]), int(mode[2])
        N1, N2 = (a if nu_R > 0 else 1), (a if nu_R == 2 else 1)
        S1, S2 = (ns if nu_S > 0 else 1), (ns if nu_S == 2 else 1)
        sym = nu_R % 2 == 0 and nu_S % 2 == 0
        dim, P_i, P_j, P_el = _indices(mode, am, L)
        c = np.zeros(dim)
        np.add.at(c, P_i - 1, P_el * u[P_j - 1])
        expect = np.zeros(dim)
        K = N1 * S1
        for x in range(n_max):
            for y in range(n_max):
                I = (x // a % S1) * N1 + x % a % N1
                J = (y // a % S2) * N2 + y % a % N2
                for l in range(L1):
                    if sym:
                        lo, hi = min(I, J), max(I, J)
                        k = (lo * (2 * K + 1 - lo) * L1) // 2 + (hi - lo) * L1 + l
                        expect[k] += p[x, y, l] * (1.0 if I == J else 1.0 / np.sqrt(2.0))
                    else:
                        expect[(I * S2 * N2 + J) * L1 + l] += p[x, y, l]
        np.testing.assert_allclose(c, expect, rtol=1e-12, atol=1e-12 * np.abs(expect).max(), err_msg=mode)


def test_darby_modes_stop_on_unequal_radial_counts():
    with pytest.raises(st.SoapTurboStop, match="requires all species to have the same number"):
        st.get_compress_indices("1_1", [3, 4], 2, "get_dim")


def test_unknown_compress_mode_and_what_to_do_fail():
    with pytest.raises(st.SoapTurboStop, match="I don't understand compress_mode"):
        st.get_compress_indices("0_0_darby", [2], 1, "get_dim")
    with pytest.raises(st.SoapTurboStop):
        st.get_compress_indices("3_0", [2], 1, "get_dim")
    with pytest.raises(st.SoapTurboStop):
        st.get_compress_indices("trivial", [2], 1, "bogus")


def test_n_soap():
    keep = _lib.Keep()
    p = st.make_params(keep, alpha_max=[8, 8, 8], l_max=8, rcut_hard=[5.0] * 3, rcut_soft=[4.5] * 3, nf=[4.0] * 3,
                       global_scaling=[1.0] * 3, atom_sigma_r=[0.5] * 3, atom_sigma_r_scaling=[0.0] * 3,
                       atom_sigma_t=[0.5] * 3, atom_sigma_t_scaling=[0.0] * 3, amplitude_scaling=[1.0] * 3,
                       radial_enhancement=1, central_weight=[1.0] * 3, basis="poly3gauss", scaling_mode="polynomial")
    assert _lib.lib().st_n_soap(ctypes.byref(p), None) == 24 * 25 // 2 * 9


@pytest.mark.parametrize("name", CASES)
def test_own_basis_is_the_exact_inverse_square_root(name):
    # W = S^{-1/2} from the library's extended-precision eigen-route (host only, no GPU) against a
    # 50-digit mpmath S^{-1/2}, and against the reference's LAPACK W, which is itself only as
    # accurate as cond(S) allows (2.8e-6 at poly3 alpha_max = 8, cond(S) = 1.3e13)
    import mpmath as mp
    w, comp, d = load(name)
    p = w.params
    keep = _lib.Keep()
    prm = st.make_params(keep, alpha_max=p.alpha_max, l_max=p.l_max, radial_enhancement=p.radial_enhancement,
                         basis=p.basis, scaling_mode=p.scaling_mode, **p.per_species())
    n = sum(p.alpha_max)
    W = np.zeros((n, n), order="F"); S = np.zeros((n, n), order="F")
    assert _lib.lib().st_get_basis(None, ctypes.byref(prm), _lib.ptr(W), _lib.ptr(S)) == 0
    np.testing.assert_array_equal(S, d["S"])
    mp.mp.dps = 50
    Wex = np.zeros((n, n))
    o = 0
    for a in p.alpha_max:
        E, Q = mp.eigsy(mp.matrix(S[o:o + a, o:o + a].tolist()))
        Wex[o:o + a, o:o + a] = np.array((Q * mp.diag([1 / mp.sqrt(E[i]) for i in range(a)]) * Q.T).tolist(),
                                         dtype=float)
        o += a
    from tests.golden_cases import rel
    assert rel(W, Wex) <= 1e-7
    assert rel(W, d["W"]) <= 1e-5


@pytest.mark.parametrize("name", [c for c in CASES if not c.startswith("op_")])
def test_lapack_route_w_is_the_reference_w(name):
    # W by the reference's own sequence (dgesvd -> U sqrt(Sigma) V^T -> dpotrf/dpotri), host-only, on the
    # host's LAPACK: bit for bit the W of the reference run on the same host and the same LAPACK (MKL
    # picks its code path per CPU, so the fixtures' W, written on another host, may differ in the
    # last bits, amplified by cond(S))
    from oracle import refio
    w, comp, d = load(name)
    p = w.params
    keep = st._lib.Keep()
    prm = st.make_params(keep, alpha_max=p.alpha_max, l_max=p.l_max, radial_enhancement=p.radial_enhancement,
                         basis=p.basis, scaling_mode=p.scaling_mode, **p.per_species())
    try:
        W, S = st.basis_matrices(prm, sum(p.alpha_max), lapack=True)
    except st.SoapTurboError as e:
        pytest.skip(f"no LAPACK on this host: {e}")
    assert np.array_equal(S, d["S"])
    if not refio.available():
        pytest.skip("the reference harness (oracle/_ref) is not built")
    Wr = refio.run(w, modes=refio.MODE_DESC, legacy_seed=bool(d["legacy_seed"]))["W"]
    print(f"{name}: W bitwise equal to the same-host reference's: {np.array_equal(W, Wr)}; "
          f"to the fixture's: {np.array_equal(W, d['W'])}")
    assert np.array_equal(W, Wr)
