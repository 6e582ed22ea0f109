"""ctypes binding of ``libsoapturbo_hip.so`` (declarations in ``include/soap_turbo_hip.h``).

Loading fails loudly: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsoapturbo_hip.so")

ST_OK, ST_ERR_ARG, ST_ERR_HIP, ST_ERR_UNSUPPORTED, ST_ERR_STATE, ST_ERR_STOP = range(6)

P = ctypes.c_void_p
I32 = ctypes.c_int32


class StParams(ctypes.Structure):
    _fields_ = [("n_species", I32), ("l_max", I32), ("radial_enhancement", I32), ("basis", I32),
                ("scaling_mode", I32), ("legacy_filter_seed", I32)] + \
               [(k, P) for k in ("alpha_max", "rcut_hard", "rcut_soft", "nf", "global_scaling", "atom_sigma_r",
                                 "atom_sigma_r_scaling", "atom_sigma_t", "atom_sigma_t_scaling", "amplitude_scaling",
                                 "central_weight")]


class StNeighbours(ctypes.Structure):
    _fields_ = [("n_sites", I32), ("n_pairs", I32), ("n_species", I32), ("max_mult", I32)] + \
               [(k, P) for k in ("n_neigh", "species", "species_multiplicity", "mask", "rjs", "thetas", "phis")]


class StCompression(ctypes.Structure):
    _fields_ = [("P_nonzero", I32), ("P_i", P), ("P_j", P), ("P_el", P)]


# name -> (restype, argtypes); every symbol the header declares
SIGNATURES = {
    "st_create": (ctypes.c_int, [ctypes.POINTER(P), ctypes.c_int]),
    "st_destroy": (None, [P]),
    "st_error": (ctypes.c_char_p, [P]),
    "st_n_soap": (ctypes.c_int, [P, P]),
    "st_get_basis": (ctypes.c_int, [P, P, P, P]),
    "st_set_basis": (ctypes.c_int, [P, I32, P, P]),
    "st_set_basis_route": (ctypes.c_int, [P, I32]),
    "st_basis_matrices": (ctypes.c_int, [I32, P, P, P]),
    "st_get_soap": (ctypes.c_int, [P, P, P, P, I32, I32, I32, P, P]),
    "st_soap_contract": (ctypes.c_int, [P, P, P, ctypes.c_double, P]),
    "st_get_compress_indices": (ctypes.c_int, [ctypes.c_char_p, I32, P, I32, P, P, P, P, P, I32]),
    "st_upload_neighbours": (ctypes.c_int, [P, P]),
    "st_upload_atoms": (ctypes.c_int, [P, I32, P, P]),
    "st_build_neighbours": (ctypes.c_int, [P, I32, P, P, I32, ctypes.c_double, ctypes.c_double, I32, P,
                                           ctypes.POINTER(I32)]),
    "st_build_neighbours_cell": (ctypes.c_int, [P, I32, P, P, I32, P, ctypes.c_double, I32, P, ctypes.POINTER(I32)]),
    "st_download_neighbours": (ctypes.c_int, [P, P, P, P, P, P, P, P]),
    "st_atom_pairs": (ctypes.c_int, [P, P, P]),
    "st_upload_dedq": (ctypes.c_int, [P, I32, P]),
    "st_compute": (ctypes.c_int, [P, P, P, I32, I32]),
    "st_pairs_full_jacobian": (ctypes.c_int, [P]),
    "st_hessian_selftest": (ctypes.c_int, [P, P, P, I32, P, P]),
    "st_full_jacobian_forces": (ctypes.c_int, [P, P, P, ctypes.c_int64, ctypes.POINTER(I32)]),
    "st_pairs_full_jacobian_noalloc": (ctypes.c_int, [P, I32, I32, I32, I32, P, P, P, P, P, P, P, P, P]),
    "st_contract_resident": (ctypes.c_int, [P, ctypes.c_double]),
    "st_adjoint_f32": (ctypes.c_int, [P, P, P]),
    "st_set_graphs": (ctypes.c_int, [P, I32]),
    "st_set_autotune": (ctypes.c_int, [P, I32]),
    "st_jacobian_variants": (ctypes.c_int, [P, I32]),
    "st_set_jacobian_variant": (ctypes.c_int, [P, I32]),
    "st_download": (ctypes.c_int, [P, P, P]),
    "st_download_pairs": (ctypes.c_int, [P, I32, P]),
    "st_download_forces": (ctypes.c_int, [P, I32, P, P]),
    "st_download_soap_f32": (ctypes.c_int, [P, P]),
    "st_central_hessian": (ctypes.c_int, [P, P, P, I32, P, P]),
    "st_set_gap_model": (ctypes.c_int, [P, I32, I32, P, P, I32, ctypes.c_double]),
    "st_gap_energy_forces": (ctypes.c_int, [P, P, P, ctypes.POINTER(ctypes.c_double)]),
    "st_gap_site_energies": (ctypes.c_int, [P, P]),
    "st_synchronize": (ctypes.c_int, [P]),
    "st_event_record": (ctypes.c_int, [P, I32]),
    "st_event_elapsed": (ctypes.c_int, [P, I32, I32, ctypes.POINTER(ctypes.c_float)]),
    "st_kernel_timing": (ctypes.c_int, [P, I32]),
    "st_kernel_time": (ctypes.c_int, [P, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(I32)]),
}

_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with soap_turbo_amd.build (no CPU fallback exists)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def ptr(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


class Keep:
    """Keeps the numpy arrays that ctypes structures point into alive."""

    def __init__(self):
        self.arrays = []

    def __call__(self, a, dtype, order="F"):
        a = np.asarray(a, dtype=dtype)
        a = np.asfortranarray(a) if order == "F" else np.ascontiguousarray(a)
        self.arrays.append(a)
        return ctypes.c_void_p(a.ctypes.data)
