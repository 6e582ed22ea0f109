"""st_pairs_full_jacobian_noalloc (no allocation; static device scratch) against st_pairs_full_jacobian.

The no-alloc entry point runs the same kernels on caller-owned device arrays (hipMalloc'd here, by the
test, as a caller would), with the force gather's partial sums in the library's static scratch array.
Its pair gradients, forces and virial must equal the allocating entry point's exactly, on two calls in
a row (the second smaller, so the scratch holds the first call's partials when the second starts).
"""
import ctypes

import numpy as np
import pytest

import soap_turbo_amd as st
from soap_turbo_amd import _lib
from soap_turbo_amd.workload import make_workload, normal

pytestmark = pytest.mark.gpu


def prepare(ctx, w):
    p = w.params
    keep = st._lib.Keep()
    prm = st.make_params(keep, alpha_max=p.alpha_max, l_max=p.l_max, radial_enhancement=p.radial_enhancement,
                         basis=p.basis, scaling_mode=p.scaling_mode, **p.per_species())
    dim, nz = st.get_compress_indices("trivial", p.alpha_max, p.l_max, "get_dim")
    comp = (np.zeros(nz, np.int32), np.zeros(nz, np.int32), np.zeros(nz))
    st.get_compress_indices("trivial", p.alpha_max, p.l_max, "set_indices", *comp)
    cmp_ = st.make_compression(keep, True, nz, *comp)
    n_soap = int(comp[0].max())
    G = normal(17 + w.n_sites, (n_soap, w.n_sites))     # C (n_soap, n_sites) == G(n_sites, n_soap)
    ctx.upload_neighbours(w.n_neigh, w.species.T, w.species_multiplicity, w.mask.T, w.rjs, w.thetas, w.phis)
    ctx.upload_atoms(len(w.pos), w.nbr, w.vec)
    ctx.upload_dedq(np.asfortranarray(G.T))
    ctx.compute(prm, cmp_, True, False)
    return keep, prm, cmp_, n_soap, G


class DeviceArrays:
    """Device copies of host arrays through the HIP runtime the library itself is linked with (the
    soname is already loaded by the library, so this is the same runtime instance)."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), max(a.nbytes, 8)) == 0
        self.ptrs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0   # host to device
        return p

    def get(self, p, shape, dtype):
        out = np.empty(shape, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0   # device to host
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def noalloc(ctx, w, n_soap, G, cart):
    d = DeviceArrays()
    try:
        site_ptr = d.put(np.concatenate([[0], np.cumsum(w.n_neigh)]).astype(np.int32))
        atom_ptr, atom_pairs = ctx.atom_pairs()
        d_atom_ptr, d_atom_pairs = d.put(atom_ptr.astype(np.int32)), d.put(atom_pairs.astype(np.int32))
        d_G = d.put(G)                                  # (n_soap, n_sites) C == G(n_sites, n_soap) F
        d_cart = d.put(np.asfortranarray(cart).T)       # (n_pairs, n_soap, 3) C == soap_cart_der(3, n_soap, n_pairs) F
        d_vec = d.put(w.vec)                            # (n_pairs, 3) C == vec(3, n_pairs) F
        n_atoms = len(w.pos)
        pair = d.put(np.full((w.n_pairs, 3), np.nan))
        forces = d.put(np.full((n_atoms, 3), np.nan))
        virial = d.put(np.full((3, 3), np.nan))
        rc = _lib.lib().st_pairs_full_jacobian_noalloc(
            ctx.h, w.n_sites, w.n_pairs, n_soap, n_atoms, site_ptr, d_G, d_cart, d_atom_ptr, d_atom_pairs, d_vec,
            pair, forces, virial)
        ctx.check(rc)
        ctx.synchronize()
        return (d.get(pair, (w.n_pairs, 3), np.float64).T, d.get(forces, (n_atoms, 3), np.float64).T,
                d.get(virial, (3, 3), np.float64).T)
    finally:
        d.free()


def test_noalloc_equals_allocating_entry_point_twice():
    ctx = st.Context(0)
    for n_sites in (None, 300):   # the whole config-2 box (1228 sites), then 300 of its sites
        w = make_workload(2, n_sites=n_sites)
        keep, prm, cmp_, n_soap, G = prepare(ctx, w)
        _, cart = ctx.download(n_soap, cart=True)
        ctx.pairs_full_jacobian()
        pair_ref = ctx.download_pairs(0)
        f_ref, v_ref = ctx.download_forces(0)
        pair, f, v = noalloc(ctx, w, n_soap, G, cart)
        print(f"{w.n_sites} sites, {w.n_pairs} pairs: max |diff| pair {np.abs(pair - pair_ref).max():.1e} "
              f"forces {np.abs(f - f_ref).max():.1e} virial {np.abs(v - v_ref).max():.1e}")
        assert np.array_equal(pair, pair_ref)
        assert np.array_equal(f, f_ref)
        assert np.array_equal(v, v_ref)


def test_noalloc_refuses_more_atoms_than_its_scratch_holds():
    ctx = st.Context(0)
    big = 1 << 22
    rc = _lib.lib().st_pairs_full_jacobian_noalloc(ctx.h, 1, 1, 1, big, *([ctypes.c_void_p(8)] * 9))
    assert rc == _lib.ST_ERR_ARG
    assert b"ST_NOALLOC_MAX_ATOMS" in _lib.lib().st_error(ctx.h)
