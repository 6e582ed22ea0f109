"""ctypes front end of the FP64 C restatement (``libsoap_oracle.so``) -- TEST ORACLE ONLY.

Only tests, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg use
this module, as the checker. Arrays are numpy C order, transposes of the
reference's Fortran shapes (see ``soap_oracle.h``).

W is an input everywhere: the reference builds it with LAPACK
(``get_orthonormalization_matrix_poly3[gauss]``,
src/soap_turbo_radial.f90:1023-1219: dgesvd -> sqrt of the singular values ->
dgemm -> dpotrf/dpotri). ``orthonormalization`` restates those steps with
numpy's LAPACK (numpy.linalg.svd, a Cholesky factor and its inverse); its
result agrees with the reference's W only as far as the conditioning of S
allows, so parity runs inject the reference's own W.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class _Params(ctypes.Structure):
    _fields_ = [("n_species", ctypes.c_int), ("l_max", ctypes.c_int), ("radial_enhancement", ctypes.c_int),
                ("basis", ctypes.c_int), ("scaling_poly", ctypes.c_int), ("legacy_seed", ctypes.c_int)] + \
               [(k, ctypes.c_void_p) for k in ("alpha_max", "rcut_hard", "rcut_soft", "nf", "global_scaling",
                                               "atom_sigma_r", "atom_sigma_r_scaling", "atom_sigma_t",
                                               "atom_sigma_t_scaling", "amplitude_scaling", "central_weight")]


class _Inputs(ctypes.Structure):
    _fields_ = [("n_sites", ctypes.c_int), ("n_pairs", ctypes.c_int)] + \
               [(k, ctypes.c_void_p) for k in ("n_neigh", "species", "species_mult", "mask", "rjs", "thetas",
                                               "phis")]


class _Compress(ctypes.Structure):
    _fields_ = [("compress", ctypes.c_int), ("P_nonzero", ctypes.c_int), ("P_i", ctypes.c_void_p),
                ("P_j", ctypes.c_void_p), ("P_el", ctypes.c_void_p)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "libsoap_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built (make -C oracle oracle)")
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        L.so_radial.argtypes = [P, P, P, ctypes.c_int, P, P]
        L.so_angular.argtypes = [P, P, ctypes.c_int, P, P, P, P]
        L.so_get_soap.argtypes = [P, P, P, P, P, ctypes.c_int, P, P, P]
        L.so_contract.argtypes = [P, P, P, P, P, P, ctypes.c_double, P, P]
        L.so_preflm.argtypes = [ctypes.c_int, P]
        L.so_overlap_S.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, P]
        L.so_radial2.argtypes = [P, P, P, ctypes.c_int, P, P, P]
        L.so_central_hessian.argtypes = [P, P, P, P, P, ctypes.c_int, P, P, P, P, P]
        _LIB = L
    return _LIB


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


class _Keep:
    """Holds the numpy arrays the ctypes structures point into."""

    def __init__(self):
        self.arrays = []

    def __call__(self, a, dtype):
        a = np.ascontiguousarray(a, dtype=dtype)
        self.arrays.append(a)
        return ctypes.c_void_p(a.ctypes.data)


def _structs(w, legacy_seed=True, compress=None):
    p = w.params
    ps = p.per_species()
    keep = _Keep()
    prm = _Params(p.n_species, p.l_max, p.radial_enhancement, {"poly3": 0, "poly3gauss": 1, "poly3operator": 2}[p.basis],
                  1 if p.scaling_mode == "polynomial" else 0, 1 if legacy_seed else 0,
                  keep(p.alpha_max, np.int32), *[keep(ps[k], np.float64) for k in (
                      "rcut_hard", "rcut_soft", "nf", "global_scaling", "atom_sigma_r", "atom_sigma_r_scaling",
                      "atom_sigma_t", "atom_sigma_t_scaling", "amplitude_scaling", "central_weight")])
    inp = _Inputs(w.n_sites, w.n_pairs, keep(w.n_neigh, np.int32), keep(w.species[:, 0], np.int32),
                  keep(w.species_multiplicity, np.int32), keep(w.mask, np.int32), keep(w.rjs, np.float64),
                  keep(w.thetas, np.float64), keep(w.phis, np.float64))
    if compress is not None:
        P_i, P_j, P_el = compress
        cmp_ = _Compress(1, len(P_i), keep(P_i, np.int32), keep(P_j, np.int32), keep(P_el, np.float64))
    else:
        cmp_ = _Compress(0, 0, None, None, None)
    return prm, inp, cmp_, keep


def n_max(w):
    return int(sum(w.params.alpha_max))


def k_max(l_max):
    return 1 + l_max * (l_max + 1) // 2 + l_max


def overlap_S(basis, alpha_max, atom_sigma_r, rcut_hard):
    S = np.zeros((alpha_max, alpha_max))
    lib().so_overlap_S(1 if basis == "poly3gauss" else 0, alpha_max, atom_sigma_r, rcut_hard, _ptr(S))
    return S.T.copy()   # column-major -> row index first


def orthonormalization(S):
    """W = S^{-1/2} by the reference's steps (SVD, sqrt, Cholesky inverse) on numpy's LAPACK."""
    U, sv, VT = np.linalg.svd(S)
    Sh = (U * np.sqrt(sv)) @ VT
    Lc = np.linalg.cholesky(Sh)          # Sh = Lc Lc^T
    Li = np.linalg.inv(Lc)
    return Li.T @ Li


def basis_matrices(w):
    """Block-diagonal (W, S) of workload ``w`` from the restated S and numpy's LAPACK."""
    p = w.params
    ps = p.per_species()
    nm = n_max(w)
    W = np.zeros((nm, nm)); S = np.zeros((nm, nm))
    o = 0
    for s, a in enumerate(p.alpha_max):
        Ss = overlap_S(p.basis, a, ps["atom_sigma_r"][s], ps["rcut_hard"][s])
        S[o:o + a, o:o + a] = Ss
        W[o:o + a, o:o + a] = orthonormalization(Ss)
        o += a
    return W, S


def preflm(l_max):
    out = np.zeros(k_max(l_max))
    lib().so_preflm(l_max, _ptr(out))
    return out


def radial(w, W, legacy_seed=True, do_der=True):
    prm, inp, _, keep = _structs(w, legacy_seed)
    nm = n_max(w)
    Wf = np.asfortranarray(W)
    rad = np.zeros((w.n_pairs, nm)); rad_der = np.zeros((w.n_pairs, nm))
    lib().so_radial(ctypes.byref(prm), ctypes.byref(inp), ctypes.c_void_p(Wf.ctypes.data), int(do_der),
                    _ptr(rad), _ptr(rad_der))
    return rad, rad_der


def angular(w, do_der=True):
    prm, inp, _, keep = _structs(w)
    km = k_max(w.params.l_max)
    out = [np.zeros((w.n_pairs, km), dtype=np.complex128) for _ in range(4)]
    lib().so_angular(ctypes.byref(prm), ctypes.byref(inp), int(do_der), *[_ptr(a) for a in out])
    return out


def get_soap(w, W, S, compress=None, do_der=True, legacy_seed=True):
    """Returns (soap[n_sites, n_soap], soap_cart_der[n_pairs, n_soap, 3] or None, sqrt_dot_p)."""
    prm, inp, cmp_, keep = _structs(w, legacy_seed, compress)
    n_soap = len(compress[0]) if compress is not None else n_max(w) * (n_max(w) + 1) // 2 * (w.params.l_max + 1)
    if compress is not None:
        n_soap = int(np.max(compress[0]))
    soap = np.zeros((w.n_sites, n_soap))
    cart = np.zeros((w.n_pairs, n_soap, 3)) if do_der else None
    sdp = np.zeros(w.n_sites)
    Wf, Sf = np.asfortranarray(W), np.asfortranarray(S)
    lib().so_get_soap(ctypes.byref(prm), ctypes.byref(inp), ctypes.c_void_p(Wf.ctypes.data),
                      ctypes.c_void_p(Sf.ctypes.data), ctypes.byref(cmp_), int(do_der), _ptr(soap), _ptr(cart),
                      _ptr(sdp))
    return soap, cart, sdp


def contract(w, W, S, G, sgn=1.0, compress=None, legacy_seed=True):
    """get_soap(contract=.true.) + soap_contract; G is (n_soap, n_sites) C == G(n_sites, n_soap) F."""
    prm, inp, cmp_, keep = _structs(w, legacy_seed, compress)
    n_soap = G.shape[0]
    soap = np.zeros((w.n_sites, n_soap))
    pair = np.zeros((w.n_pairs, 3))
    Wf, Sf = np.asfortranarray(W), np.asfortranarray(S)
    Gc = np.ascontiguousarray(G, dtype=np.float64)
    lib().so_contract(ctypes.byref(prm), ctypes.byref(inp), ctypes.c_void_p(Wf.ctypes.data),
                      ctypes.c_void_p(Sf.ctypes.data), ctypes.byref(cmp_), _ptr(Gc), sgn, _ptr(soap), _ptr(pair))
    return soap, pair


def radial_der2(w, W):
    """The radial expansion and its first and second derivatives in r (unified form)."""
    prm, inp, _, keep = _structs(w, legacy_seed=False)
    nm = n_max(w)
    Wf = np.asfortranarray(W)
    rad, rad_der, rad_der2 = (np.zeros((w.n_pairs, nm)) for _ in range(3))
    lib().so_radial2(ctypes.byref(prm), ctypes.byref(inp), ctypes.c_void_p(Wf.ctypes.data), 1, _ptr(rad),
                     _ptr(rad_der), _ptr(rad_der2))
    return rad, rad_der, rad_der2


def central_hessian(w, W, S, vecs, compress=None):
    """get_soap_central_hessian (src/soap_turbo.f90:1151-1700), unified radial form. vecs is
    (n_sites, n_vec, n_soap); returns (soap (n_sites, n_soap), hess (n_pairs, n_vec, 3 a, 3 b),
    rad_der2 (n_pairs, n_max), the self-test's relative error)."""
    prm, inp, cmp_, keep = _structs(w, False, compress)
    vecs = np.ascontiguousarray(vecs, dtype=np.float64)
    n_sites, n_vec, n_soap = vecs.shape
    soap = np.zeros((w.n_sites, n_soap))
    hess = np.zeros((w.n_pairs, n_vec, 3, 3))
    rad2 = np.zeros((w.n_pairs, n_max(w)))
    err = ctypes.c_double()
    Wf, Sf = np.asfortranarray(W), np.asfortranarray(S)
    rc = lib().so_central_hessian(ctypes.byref(prm), ctypes.byref(inp), ctypes.c_void_p(Wf.ctypes.data),
                                  ctypes.c_void_p(Sf.ctypes.data), ctypes.byref(cmp_), n_vec, _ptr(vecs), _ptr(soap),
                                  _ptr(hess), _ptr(rad2), ctypes.byref(err))
    if rc:
        raise RuntimeError("the central-atom Hessian needs the unified radial form (legacy_seed False)")
    return soap, hess, rad2, err.value

