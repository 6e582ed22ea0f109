"""MI355X-native get_soap: the soap_turbo descriptor hot path on gfx950 HIP kernels.

Python mirror of the reference's interface (module ``soap_turbo_desc``,
src/soap_turbo.f90:139-169 and :1723-1730; ``get_compress_indices``,
src/soap_turbo_compress.f90:44-54), over the C-ABI of ``libsoapturbo_hip.so``:

* ``get_soap(...)`` takes the reference's arguments, with arrays in the
  reference's Fortran shapes (``mask`` is ``(n_atom_pairs, n_species)``,
  ``soap`` is ``(n_soap, n_sites)``, ``soap_cart_der`` is
  ``(3, n_soap, n_atom_pairs)``) and fills ``soap`` and ``soap_cart_der`` in
  place. As for the reference, both must be zero on entry.
* ``soap_contract(soap, G, sgn, pair)`` follows a ``get_soap(...,
  contract=True)``; ``G`` is ``(n_sites, n_soap)``.
* ``legacy_filter_seed`` mirrors soap_turbo_radial's flag
  (src/soap_turbo_radial.f90:67), default True.

Errors the reference stops on raise ``SoapTurboStop`` with its message; the
rest raise ``SoapTurboError``. There is no CPU fallback.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import Keep, StCompression, StNeighbours, StParams, lib, ptr

legacy_filter_seed = True

__all__ = ["get_soap", "soap_contract", "get_compress_indices", "Context", "SoapTurboError", "SoapTurboStop",
           "legacy_filter_seed"]


class SoapTurboError(RuntimeError):
    pass


class SoapTurboStop(SoapTurboError):
    """A condition on which the reference writes its message and stops."""


def _check(ctx_handle, rc):
    if rc == _lib.ST_OK:
        return
    msg = lib().st_error(ctx_handle)
    msg = msg.decode() if msg else ""
    if rc == _lib.ST_ERR_STOP or rc == _lib.ST_ERR_STATE and msg.startswith("ERROR"):
        raise SoapTurboStop(msg)
    raise SoapTurboError(f"soap_turbo_hip error {rc}: {msg}")


def make_params(keep: Keep, *, alpha_max, l_max, rcut_hard, rcut_soft, nf, global_scaling, atom_sigma_r,
                atom_sigma_r_scaling, atom_sigma_t, atom_sigma_t_scaling, amplitude_scaling, radial_enhancement,
                central_weight, basis, scaling_mode, legacy_seed=None):
    basis = basis.strip()
    codes = {"poly3": 0, "poly3gauss": 1, "poly3operator": 2}
    if basis not in codes:
        raise SoapTurboError(f'basis "{basis}" is not supported (poly3, poly3gauss, poly3operator)')
    ns = len(alpha_max)
    seed = legacy_filter_seed if legacy_seed is None else legacy_seed
    return StParams(ns, int(l_max), int(radial_enhancement), codes[basis],
                    0 if scaling_mode.strip() == "polynomial" else 1, 1 if seed else 0,
                    keep(alpha_max, np.int32), keep(rcut_hard, np.float64), keep(rcut_soft, np.float64),
                    keep(nf, np.float64), keep(global_scaling, np.float64), keep(atom_sigma_r, np.float64),
                    keep(atom_sigma_r_scaling, np.float64), keep(atom_sigma_t, np.float64),
                    keep(atom_sigma_t_scaling, np.float64), keep(amplitude_scaling, np.float64),
                    keep(central_weight, np.float64))


def make_compression(keep: Keep, compress_soap, P_nonzero=0, P_i=None, P_j=None, P_el=None):
    if not compress_soap or not P_nonzero:
        return StCompression(0, None, None, None)
    return StCompression(int(P_nonzero), keep(np.asarray(P_i)[:P_nonzero], np.int32),
                         keep(np.asarray(P_j)[:P_nonzero], np.int32), keep(np.asarray(P_el)[:P_nonzero], np.float64))


class Context:
    """One device context (stream, resident buffers, the W/S cache)."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        rc = lib().st_create(ctypes.byref(h), device)
        if rc != 0:
            raise SoapTurboError(f"st_create failed ({rc}): no usable GPU {device}")
        self.h = h

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            lib().st_destroy(h)
            self.h = None

    def check(self, rc):
        _check(self.h, rc)

    # ---- basis
    def set_basis(self, W, S):
        W = np.asfortranarray(W, dtype=np.float64); S = np.asfortranarray(S, dtype=np.float64)
        self.check(lib().st_set_basis(self.h, W.shape[0], ptr(W), ptr(S)))

    def set_basis_route(self, lapack: bool):
        """W by the reference's LAPACK sequence (True) or the extended-precision route (False, default)."""
        self.check(lib().st_set_basis_route(self.h, 1 if lapack else 0))

    def clear_basis(self):
        self.check(lib().st_set_basis(self.h, 0, None, None))

    def get_basis(self, params: StParams, n_max: int):
        W = np.zeros((n_max, n_max), order="F"); S = np.zeros((n_max, n_max), order="F")
        self.check(lib().st_get_basis(self.h, ctypes.byref(params), ptr(W), ptr(S)))
        return W, S

    # ---- resident use
    def upload_neighbours(self, n_neigh, species, species_multiplicity, mask, rjs, thetas, phis):
        keep = Keep()
        mask = np.asfortranarray(mask)
        species = np.asfortranarray(species, dtype=np.int32)
        if species.ndim == 1:
            species = species.reshape(1, -1, order="F")
        nb = StNeighbours(len(n_neigh), len(rjs), mask.shape[1], species.shape[0], keep(n_neigh, np.int32),
                          keep(species, np.int32), keep(species_multiplicity, np.int32), keep(mask, np.int32),
                          keep(rjs, np.float64), keep(thetas, np.float64), keep(phis, np.float64))
        self.check(lib().st_upload_neighbours(self.h, ctypes.byref(nb)))
        self.n_sites, self.n_pairs = len(n_neigh), len(rjs)

    def set_autotune(self, on: bool):
        """Time the candidate launch shapes of the Jacobian and site kernels on first use; keep the fastest."""
        self.check(lib().st_set_autotune(self.h, int(bool(on))))

    def jacobian_variants(self, n_soap: int) -> int:
        return lib().st_jacobian_variants(self.h, int(n_soap))

    def set_jacobian_variant(self, variant: int):
        """Fix the Jacobian kernel's launch variant (-1: default or tuned)."""
        self.check(lib().st_set_jacobian_variant(self.h, int(variant)))

    def set_graphs(self, on: bool):
        """Replay the FP32 step from a captured HIP graph while its inputs keep their shapes."""
        self.check(lib().st_set_graphs(self.h, int(bool(on))))

    def build_neighbours(self, pos, atom_species, n_species, box, rcut, site_atom):
        """Neighbour list of a periodic cell built on the device (pos (n_atoms, 3), atom_species 1-based,
        site_atom 0-based): the resident state of upload_neighbours + upload_atoms. box is a cubic side, or a
        (3, 3) array whose rows are the lattice vectors of a general cell."""
        keep = Keep()
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        n = ctypes.c_int32(0)
        if np.ndim(box) == 2:
            self.check(lib().st_build_neighbours_cell(self.h, len(pos), keep(pos, np.float64, "C"),
                                                      keep(atom_species, np.int32), int(n_species),
                                                      keep(box, np.float64, "C"), float(rcut), len(site_atom),
                                                      keep(site_atom, np.int32), ctypes.byref(n)))
        else:
            self.check(lib().st_build_neighbours(self.h, len(pos), keep(pos, np.float64, "C"),
                                                 keep(atom_species, np.int32), int(n_species), float(box), float(rcut),
                                                 len(site_atom), keep(site_atom, np.int32), ctypes.byref(n)))
        self.n_sites, self.n_pairs, self.n_atoms, self.n_species_nb = len(site_atom), n.value, len(pos), int(n_species)
        return n.value

    def download_neighbours(self):
        """(n_neigh, pair_atom, rjs, thetas, phis, mask (n_species, n_pairs) bool, vec (n_pairs, 3))."""
        P = self.n_pairs
        n_neigh = np.zeros(self.n_sites, np.int32); pa = np.zeros(P, np.int32)
        rjs, th, ph = np.zeros(P), np.zeros(P), np.zeros(P)
        mask = np.zeros((self.n_species_nb, P), np.int32); vec = np.zeros((P, 3))
        self.check(lib().st_download_neighbours(self.h, ptr(n_neigh), ptr(pa), ptr(rjs), ptr(th), ptr(ph), ptr(mask),
                                                ptr(vec)))
        return n_neigh, pa, rjs, th, ph, mask.astype(bool), vec

    def atom_pairs(self):
        """(atom_ptr, atom_pairs): the pairs whose neighbour is each atom, in increasing pair index."""
        ap = np.zeros(self.n_atoms + 1, np.int32); pr = np.zeros(self.n_pairs, np.int32)
        self.check(lib().st_atom_pairs(self.h, ptr(ap), ptr(pr)))
        return ap, pr

    def upload_atoms(self, n_atoms, pair_atom, vec=None):
        keep = Keep()
        v = keep(vec, np.float64, "C") if vec is not None else None
        self.check(lib().st_upload_atoms(self.h, int(n_atoms), keep(pair_atom, np.int32), v))
        self.n_atoms = int(n_atoms)

    def upload_dedq(self, G):
        """G(n_sites, n_soap), Fortran order (a C-ordered (n_soap, n_sites) array is the same memory)."""
        G = np.asfortranarray(G, dtype=np.float64)
        self.check(lib().st_upload_dedq(self.h, G.shape[1], ptr(G)))

    def compute(self, params, compression, do_derivatives=True, contract=False):
        self.check(lib().st_compute(self.h, ctypes.byref(params), ctypes.byref(compression), int(do_derivatives),
                                    int(contract)))

    def pairs_full_jacobian(self):
        self.check(lib().st_pairs_full_jacobian(self.h))

    def full_jacobian_forces(self, params, compression, max_pairs=0):
        """The FP64 full-Jacobian path in batches of sites (st_full_jacobian_forces); the batch count."""
        nb = ctypes.c_int32(0)
        self.check(lib().st_full_jacobian_forces(self.h, ctypes.byref(params), ctypes.byref(compression),
                                                 int(max_pairs), ctypes.byref(nb)))
        return nb.value

    def contract_resident(self, sgn=1.0):
        self.check(lib().st_contract_resident(self.h, sgn))

    def adjoint_f32(self, params, compression):
        self.check(lib().st_adjoint_f32(self.h, ctypes.byref(params), ctypes.byref(compression)))

    def download(self, n_soap, cart=False):
        soap = np.zeros((n_soap, self.n_sites), order="F")
        c = np.zeros((3, n_soap, self.n_pairs), order="F") if cart else None
        self.check(lib().st_download(self.h, ptr(soap), ptr(c)))
        return soap, c

    def download_pairs(self, which=0):
        pair = np.zeros((3, self.n_pairs), order="F")
        self.check(lib().st_download_pairs(self.h, int(which), ptr(pair)))
        return pair

    def download_forces(self, which=0):
        f = np.zeros((3, self.n_atoms), order="F")
        v = np.zeros((3, 3), order="F")
        self.check(lib().st_download_forces(self.h, int(which), ptr(f), ptr(v)))
        return f, v

    def download_soap_f32(self, n_soap):
        soap = np.zeros((n_soap, self.n_sites), order="F")
        self.check(lib().st_download_soap_f32(self.h, ptr(soap)))
        return soap

    def central_hessian(self, params: StParams, comp: StCompression, vecs):
        """get_soap_central_hessian on the resident list: vecs (n_soap, n_vec, n_sites) Fortran order;
        returns hess (3, 3, n_vec, n_pairs), hess[a, b, v, k] = sum_d vecs[d, v, i] d2 q_d / dr_ia dr_jb."""
        vecs = np.asfortranarray(vecs, dtype=np.float64)
        n_vec = vecs.shape[1]
        hess = np.zeros((3, 3, n_vec, self.n_pairs), order="F")
        self.check(lib().st_central_hessian(self.h, ctypes.byref(params), ctypes.byref(comp), n_vec, ptr(vecs),
                                            ptr(hess)))
        return hess

    def hessian_selftest(self, params: StParams, comp: StCompression, vecs) -> float:
        """soap_hessian_selftest after central_hessian with the same vecs: the Hessian route's first
        derivatives against the FP64 Jacobian's, max |difference| / max |value|."""
        vecs = np.asfortranarray(vecs, dtype=np.float64)
        err = ctypes.c_double(0.0)
        self.check(lib().st_hessian_selftest(self.h, ctypes.byref(params), ctypes.byref(comp), vecs.shape[1],
                                             ptr(vecs), ctypes.byref(err)))
        return err.value

    # ---- GAP energy model on the device
    def set_gap_model(self, sparse_q, alpha, zeta: int, delta: float):
        """sparse_q (n_soap, n_sparse), alpha (n_sparse): E = sum_i delta^2 sum_t alpha_t (q_i . q_t)^zeta."""
        T = np.asfortranarray(sparse_q, dtype=np.float64)
        a = np.ascontiguousarray(alpha, dtype=np.float64)
        if a.shape != (T.shape[1],):
            raise ValueError("alpha needs one weight per sparse point")
        self.check(lib().st_set_gap_model(self.h, T.shape[1], T.shape[0], ptr(T), ptr(a), int(zeta), float(delta)))

    def gap_energy_forces(self, params: StParams, comp: StCompression) -> float:
        """Descriptors, the model's energy and dE/dq, then forces and virial (download_forces(1))."""
        e = ctypes.c_double()
        self.check(lib().st_gap_energy_forces(self.h, ctypes.byref(params), ctypes.byref(comp), ctypes.byref(e)))
        return e.value

    def gap_site_energies(self):
        e = np.zeros(self.n_sites)
        self.check(lib().st_gap_site_energies(self.h, ptr(e)))
        return e

    def synchronize(self):
        self.check(lib().st_synchronize(self.h))

    def event_record(self, slot):
        self.check(lib().st_event_record(self.h, slot))

    def event_elapsed(self, a, b):
        ms = ctypes.c_float()
        self.check(lib().st_event_elapsed(self.h, a, b, ctypes.byref(ms)))
        return ms.value

    def kernel_timing(self, on=True):
        self.check(lib().st_kernel_timing(self.h, int(on)))

    def kernel_time(self, name):
        t = ctypes.c_double(); n = ctypes.c_int32()
        self.check(lib().st_kernel_time(self.h, name.encode(), ctypes.byref(t), ctypes.byref(n)))
        return t.value, n.value


_DEFAULT = None


def basis_matrices(params: StParams, n_max: int, lapack: bool = False):
    """W and S (n_max x n_max) on the host, no device: the reference's LAPACK route or the extended one."""
    W = np.zeros((n_max, n_max), order="F"); S = np.zeros((n_max, n_max), order="F")
    _check(None, lib().st_basis_matrices(1 if lapack else 0, ctypes.byref(params), ptr(W), ptr(S)))
    return W, S


def default_context() -> Context:
    global _DEFAULT
    if _DEFAULT is None:
        _DEFAULT = Context(0)
    return _DEFAULT


def get_soap(n_sites, n_neigh, n_species, species, species_multiplicity, n_atom_pairs, mask, rjs, thetas, phis,
             alpha_max, l_max, rcut_hard, rcut_soft, nf, global_scaling, atom_sigma_r, atom_sigma_r_scaling,
             atom_sigma_t, atom_sigma_t_scaling, amplitude_scaling, radial_enhancement, central_weight, basis,
             scaling_mode, do_timing, do_derivatives, compress_soap, compress_P_nonzero, compress_P_i, compress_P_j,
             compress_P_el, soap, soap_cart_der, contract=False, ctx: Context | None = None):
    """get_soap (src/soap_turbo.f90:139-169): fills soap(n_soap, n_sites) and, with derivatives and not
    contract, soap_cart_der(3, n_soap, n_atom_pairs), both Fortran-ordered numpy arrays, in place."""
    ctx = ctx or default_context()
    keep = Keep()
    prm = make_params(keep, alpha_max=alpha_max, l_max=l_max, rcut_hard=rcut_hard, rcut_soft=rcut_soft, nf=nf,
                      global_scaling=global_scaling, atom_sigma_r=atom_sigma_r,
                      atom_sigma_r_scaling=atom_sigma_r_scaling, atom_sigma_t=atom_sigma_t,
                      atom_sigma_t_scaling=atom_sigma_t_scaling, amplitude_scaling=amplitude_scaling,
                      radial_enhancement=radial_enhancement, central_weight=central_weight, basis=basis,
                      scaling_mode=scaling_mode)
    cmp_ = make_compression(keep, compress_soap, compress_P_nonzero, compress_P_i, compress_P_j, compress_P_el)
    mask = np.asfortranarray(mask)
    species = np.asfortranarray(species, dtype=np.int32)
    if species.ndim == 1:
        species = species.reshape(1, -1, order="F")
    if mask.shape != (n_atom_pairs, n_species):
        raise SoapTurboError("mask must be (n_atom_pairs, n_species)")
    nb = StNeighbours(int(n_sites), int(n_atom_pairs), int(n_species), species.shape[0], keep(n_neigh, np.int32),
                      keep(species, np.int32), keep(species_multiplicity, np.int32), keep(mask, np.int32),
                      keep(rjs, np.float64), keep(thetas, np.float64), keep(phis, np.float64))
    if not (isinstance(soap, np.ndarray) and soap.dtype == np.float64 and soap.flags.f_contiguous):
        raise SoapTurboError("soap must be a Fortran-ordered float64 array (n_soap, n_sites)")
    want_cart = do_derivatives and not contract
    if want_cart and not (isinstance(soap_cart_der, np.ndarray) and soap_cart_der.dtype == np.float64
                          and soap_cart_der.flags.f_contiguous and soap_cart_der.shape == (3, soap.shape[0],
                                                                                            n_atom_pairs)):
        raise SoapTurboError("soap_cart_der must be a Fortran-ordered float64 array (3, n_soap, n_atom_pairs)")
    rc = lib().st_get_soap(ctx.h, ctypes.byref(prm), ctypes.byref(nb), ctypes.byref(cmp_), int(do_derivatives),
                           int(bool(contract)), soap.shape[0], ptr(soap), ptr(soap_cart_der) if want_cart else None)
    ctx.check(rc)
    ctx.n_sites, ctx.n_pairs = int(n_sites), int(n_atom_pairs)


def soap_contract(soap, G, sgn, pair, ctx: Context | None = None):
    """soap_contract (src/soap_turbo.f90:1723-1730): pair(3, n_pairs) in place."""
    ctx = ctx or default_context()
    soap = np.asfortranarray(soap, dtype=np.float64)
    G = np.asfortranarray(G, dtype=np.float64)
    if not (isinstance(pair, np.ndarray) and pair.dtype == np.float64 and pair.flags.f_contiguous):
        raise SoapTurboError("pair must be a Fortran-ordered float64 array (3, n_pairs)")
    ctx.check(lib().st_soap_contract(ctx.h, ptr(soap), ptr(G), float(sgn), ptr(pair)))


def get_compress_indices(compress_mode, alpha_max, l_max, what_to_do, P_i=None, P_j=None, P_el=None):
    """get_compress_indices (src/soap_turbo_compress.f90:44-54). Returns (dim, P_nonzero); with
    what_to_do == "set_indices" also fills P_i, P_j, P_el in place."""
    if what_to_do not in ("get_dim", "set_indices"):
        raise SoapTurboStop("ERROR: incorrect what_to_do in get_compress_indices() subroutine!")
    am = np.asarray(alpha_max, dtype=np.int32)
    dim = ctypes.c_int32(); nz = ctypes.c_int32()
    setting = what_to_do == "set_indices"
    if setting:
        for a, dt in ((P_i, np.int32), (P_j, np.int32), (P_el, np.float64)):
            if not (isinstance(a, np.ndarray) and a.dtype == dt and a.flags.c_contiguous):
                raise SoapTurboError("P_i, P_j (int32) and P_el (float64) must be contiguous arrays")
    rc = lib().st_get_compress_indices(compress_mode.encode(), len(am), ptr(am), int(l_max), ctypes.byref(dim),
                                       ctypes.byref(nz), ptr(P_i) if setting else None,
                                       ptr(P_j) if setting else None, ptr(P_el) if setting else None, int(setting))
    if rc == _lib.ST_ERR_STOP:
        # the reference's two stops (src/soap_turbo_compress.f90:136-138, :306)
        m = compress_mode.rstrip()
        if len(m) == 3 and m[1] == "_" and m[0] in "012" and m[2] in "012":
            raise SoapTurboStop(f"ERROR: compress_mode={m} requires all species to have the same number of "
                                "radial basis functions!")
        raise SoapTurboStop(f"ERROR: I don't understand compress_mode ={compress_mode}")
    if rc != 0:
        raise SoapTurboError(f"get_compress_indices failed ({rc})")
    return dim.value, nz.value
