"""Seeded synthetic inputs for ``get_soap`` in the reference's argument layout.

The arrays follow ``get_soap``'s argument list (src/soap_turbo.f90:139-169),
held in numpy C order as the transpose of the Fortran shape:

* ``n_neigh[n_sites]`` -- pairs per site, the site itself first (r = 0);
* ``rjs, thetas, phis[n_pairs]``;
* ``mask[n_species, n_pairs]`` -- Fortran ``mask(n_pairs, n_species)``;
* ``species[n_sites, 1]`` / ``species_multiplicity[n_sites]`` -- one species
  per site, 1-based as the reference expects;
* per-species hyper-parameter vectors.

The atoms come from random sequential addition in a cubic periodic box
(``csrc/workload.cpp``); the neighbour list uses the minimum image, so every
box is more than twice the cutoff across.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "_workload.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
        lib = ctypes.CDLL(path)
        P = ctypes.c_void_p
        lib.stw_place_atoms.restype = ctypes.c_int64
        lib.stw_place_atoms.argtypes = [ctypes.c_uint64, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                        P, P, P, ctypes.c_int64, P]
        lib.stw_draw_species.restype = None
        lib.stw_draw_species.argtypes = [ctypes.c_uint64, ctypes.c_int64, ctypes.c_int, P, P]
        lib.stw_normal.restype = None
        lib.stw_normal.argtypes = [ctypes.c_uint64, ctypes.c_int64, P]
        lib.stw_neighbours.restype = ctypes.c_int64
        lib.stw_neighbours.argtypes = [ctypes.c_int64, P, ctypes.c_double, ctypes.c_double, ctypes.c_int64,
                                       P, P, P, P]
        lib.stw_neighbours_cell.restype = ctypes.c_int64
        lib.stw_neighbours_cell.argtypes = [ctypes.c_int64, P, P, ctypes.c_double, ctypes.c_int64, P, P, P, P]
        lib.stw_spherical.restype = None
        lib.stw_spherical.argtypes = [ctypes.c_int64, P, P, P, P]
        _LIB = lib
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


@dataclass
class SoapParams:
    """Descriptor hyper-parameters, one entry per species where the reference has arrays."""
    alpha_max: list
    l_max: int
    rcut_hard: float
    rcut_soft: float
    atom_sigma_r: float = 0.5
    atom_sigma_t: float = 0.5
    atom_sigma_r_scaling: float = 0.0
    atom_sigma_t_scaling: float = 0.0
    amplitude_scaling: float = 1.0
    central_weight: float = 1.0
    nf: float = 4.0
    global_scaling: float = 1.0
    basis: str = "poly3"
    scaling_mode: str = "polynomial"
    radial_enhancement: int = 0
    compress_mode: str = "none"

    @property
    def n_species(self):
        return len(self.alpha_max)

    def per_species(self):
        ns = self.n_species
        f = lambda v: np.full(ns, float(v)) if np.isscalar(v) else np.asarray(v, dtype=np.float64)
        return dict(rcut_hard=f(self.rcut_hard), rcut_soft=f(self.rcut_soft), nf=f(self.nf),
                    global_scaling=f(self.global_scaling), atom_sigma_r=f(self.atom_sigma_r),
                    atom_sigma_r_scaling=f(self.atom_sigma_r_scaling), atom_sigma_t=f(self.atom_sigma_t),
                    atom_sigma_t_scaling=f(self.atom_sigma_t_scaling),
                    amplitude_scaling=f(self.amplitude_scaling), central_weight=f(self.central_weight))


@dataclass
class BoxSpec:
    n_atoms: int
    density: float                 # atoms / A^3 (uniform box)
    dmin: float
    species_prob: list
    species_z: list
    seed: int
    central: int | None = None     # 0-based species index of the sites; None = every atom
    slab: tuple | None = None      # (dense_density, gas_density): z < L/2 dense, rest gas


# The five workloads: sizes, densities and hyper-parameters as the benchmark names them.
CONFIGS = {
    1: (BoxSpec(216, 0.10, 1.2, [1.0], [6], seed=1),
        SoapParams(alpha_max=[8], l_max=8, rcut_hard=5.0, rcut_soft=4.5, basis="poly3")),
    2: (BoxSpec(4096, 0.09, 1.0, [0.5, 0.3, 0.2], [1, 6, 8], seed=2, central=1),
        SoapParams(alpha_max=[8, 8, 8], l_max=8, rcut_hard=5.0, rcut_soft=4.5, basis="poly3gauss",
                   radial_enhancement=1, compress_mode="trivial")),
    3: (BoxSpec(32768, 0.10, 1.2, [1.0], [6], seed=3),
        SoapParams(alpha_max=[8], l_max=8, rcut_hard=5.0, rcut_soft=4.5, basis="poly3")),
    4: (BoxSpec(100000, 0.0625, 1.0, [0.7, 0.3], [6, 14], seed=4, slab=(0.12, 0.005)),
        SoapParams(alpha_max=[6, 6], l_max=6, rcut_hard=5.5, rcut_soft=5.0, basis="poly3gauss")),
    5: (BoxSpec(262144, 0.10, 1.0, [0.25, 0.25, 0.25, 0.25], [1, 6, 7, 8], seed=5),
        SoapParams(alpha_max=[6, 6, 6, 6], l_max=8, rcut_hard=6.0, rcut_soft=5.5, basis="poly3gauss",
                   compress_mode="trivial")),
}


@dataclass
class Workload:
    params: SoapParams
    L: float
    pos: np.ndarray            # (n_atoms, 3)
    atom_species: np.ndarray   # (n_atoms,) 0-based
    site_atom: np.ndarray      # (n_sites,) atom index of each site
    n_neigh: np.ndarray        # (n_sites,) int32
    nbr: np.ndarray            # (n_pairs,) atom index of each pair's neighbour (the site's own atom first)
    vec: np.ndarray            # (n_pairs, 3) r_j - r_i
    rjs: np.ndarray
    thetas: np.ndarray
    phis: np.ndarray
    mask: np.ndarray           # (n_species, n_pairs) bool == Fortran mask(n_pairs, n_species)
    species: np.ndarray        # (n_sites, 1) int32, 1-based == Fortran species(1, n_sites)
    species_multiplicity: np.ndarray
    extra: dict = field(default_factory=dict)

    @property
    def n_sites(self):
        return len(self.n_neigh)

    @property
    def n_pairs(self):
        return len(self.rjs)


def place_atoms(spec: BoxSpec):
    lib = _lib()
    if spec.slab is None:
        L = (spec.n_atoms / spec.density) ** (1.0 / 3.0)
        zlo = np.array([0.0]); zhi = np.array([L]); cnt = np.array([spec.n_atoms], dtype=np.int64)
    else:
        dense, gas = spec.slab
        V = spec.n_atoms / (0.5 * (dense + gas))
        L = V ** (1.0 / 3.0)
        n_dense = int(round(0.5 * V * dense))
        zlo = np.array([0.0, 0.5 * L]); zhi = np.array([0.5 * L, L])
        cnt = np.array([n_dense, spec.n_atoms - n_dense], dtype=np.int64)
    pos = np.zeros((spec.n_atoms, 3))
    got = lib.stw_place_atoms(spec.seed, L, spec.dmin, len(cnt), _p(zlo), _p(zhi), _p(cnt), 100000, _p(pos))
    if got != spec.n_atoms:
        raise RuntimeError(f"random sequential addition placed {got} of {spec.n_atoms} atoms")
    cum = np.cumsum(np.asarray(spec.species_prob, dtype=np.float64))
    sp = np.zeros(spec.n_atoms, dtype=np.int32)
    lib.stw_draw_species(spec.seed * 7919 + 17, spec.n_atoms, len(cum), _p(cum), _p(sp))
    return L, pos, sp


def build_neighbours(pos, L, rcut, site_atom):
    """(n_neigh, nbr, vec) of the sites: L is a cubic box side, or a (3, 3) array whose rows are the
    lattice vectors of a general cell."""
    lib = _lib()
    if np.ndim(L) == 2:
        cell = np.ascontiguousarray(L, dtype=np.float64)  # rows a, b, c: the memory of cell(3, 3) by columns
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        site_atom = np.ascontiguousarray(site_atom, dtype=np.int64)
        n_neigh = np.zeros(len(site_atom), dtype=np.int32)
        n_pairs = lib.stw_neighbours_cell(len(pos), _p(pos), _p(cell), rcut, len(site_atom), _p(site_atom),
                                          _p(n_neigh), None, None)
        if n_pairs < 0:
            raise ValueError("the cell must be right-handed with every width above twice the cutoff")
        nbr = np.zeros(n_pairs, dtype=np.int32)
        vec = np.zeros((n_pairs, 3))
        lib.stw_neighbours_cell(len(pos), _p(pos), _p(cell), rcut, len(site_atom), _p(site_atom), _p(n_neigh),
                                _p(nbr), _p(vec))
        return n_neigh, nbr, vec
    if L <= 2 * rcut:
        raise ValueError("the minimum-image neighbour list needs a box wider than twice the cutoff")
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    site_atom = np.ascontiguousarray(site_atom, dtype=np.int64)
    n_neigh = np.zeros(len(site_atom), dtype=np.int32)
    n_pairs = lib.stw_neighbours(len(pos), _p(pos), L, rcut, len(site_atom), _p(site_atom), _p(n_neigh),
                                 None, None)
    nbr = np.zeros(n_pairs, dtype=np.int32)
    vec = np.zeros((n_pairs, 3))
    lib.stw_neighbours(len(pos), _p(pos), L, rcut, len(site_atom), _p(site_atom), _p(n_neigh), _p(nbr), _p(vec))
    return n_neigh, nbr, vec


def spherical(vec):
    vec = np.ascontiguousarray(vec, dtype=np.float64)
    n = len(vec)
    r = np.zeros(n); t = np.zeros(n); f = np.zeros(n)
    _lib().stw_spherical(n, _p(vec), _p(r), _p(t), _p(f))
    return r, t, f


def make_workload(cfg, n_sites: int | None = None, sites=None) -> Workload:
    """Builds configuration ``cfg`` (1..5, or a (BoxSpec, SoapParams) pair).

    ``n_sites`` keeps only the first sites (and their pairs); ``sites`` keeps the
    sites at these positions of the full site list instead (e.g. from
    ``strided_sites`` or ``slab_sites``). Either gives small parity cases cut
    from the full box, with every neighbour of a kept site still present.
    """
    spec, params = CONFIGS[cfg] if isinstance(cfg, int) else cfg
    L, pos, asp = place_atoms(spec)
    if spec.central is None:
        site_atom = np.arange(spec.n_atoms, dtype=np.int64)
    else:
        site_atom = np.nonzero(asp == spec.central)[0].astype(np.int64)
    if sites is not None:
        site_atom = site_atom[np.asarray(sites, dtype=np.int64)]
    if n_sites is not None:
        site_atom = site_atom[:n_sites]
    rcut_max = float(np.max(params.per_species()["rcut_hard"]))
    n_neigh, nbr, vec = build_neighbours(pos, L, rcut_max, site_atom)
    rjs, thetas, phis = spherical(vec)
    ns = params.n_species
    nbr_sp = asp[nbr]
    mask = np.zeros((ns, len(nbr)), dtype=bool)
    for s in range(ns):
        mask[s] = nbr_sp == s
    species = (asp[site_atom] + 1).astype(np.int32).reshape(-1, 1)
    mult = np.ones(len(site_atom), dtype=np.int32)
    return Workload(params, L, pos, asp, site_atom, n_neigh, nbr, vec, rjs, thetas, phis, mask, species, mult)


def skewed_cell(n_atoms: int, density: float):
    """Lattice vectors (rows) of a triclinic cell of volume n_atoms / density: angles near 70, 75 and 80
    degrees, every face-to-face width within 15 % of the cube's side."""
    shape = np.array([[1.0, 0.0, 0.0], [0.30, 0.95, 0.0], [-0.20, 0.25, 0.93]])
    return shape * ((n_atoms / density) / np.linalg.det(shape)) ** (1.0 / 3.0)


def triclinic_workload(n_atoms: int, density: float, species_prob, params: SoapParams, seed: int,
                       dmin: float = 1.0, n_sites: int | None = None) -> Workload:
    """Atoms in a triclinic periodic cell (skewed_cell), placed by random sequential addition with a
    minimum distance (minimum image over the lattice), a fifth of them shifted by a lattice translation
    out of the home cell; every atom a site (or the first n_sites). Workload.L holds the lattice vectors."""
    cell = skewed_cell(n_atoms, density)
    hi = np.linalg.inv(cell)
    rng = np.random.default_rng(seed)
    pos = np.zeros((n_atoms, 3))
    placed = 0
    while placed < n_atoms:
        p = rng.random(3) @ cell
        d = pos[:placed] - p
        d = d - np.rint(d @ hi) @ cell
        if placed == 0 or (np.einsum("ij,ij->i", d, d) >= dmin * dmin).all():
            pos[placed] = p
            placed += 1
    shift = rng.random(n_atoms) < 0.2
    pos[shift] += rng.integers(-1, 2, (int(shift.sum()), 3)) @ cell
    asp = rng.choice(len(species_prob), n_atoms, p=species_prob).astype(np.int32)
    site_atom = np.arange(n_atoms if n_sites is None else n_sites, dtype=np.int64)
    rcut_max = float(np.max(params.per_species()["rcut_hard"]))
    n_neigh, nbr, vec = build_neighbours(pos, cell, rcut_max, site_atom)
    rjs, thetas, phis = spherical(vec)
    mask = np.zeros((params.n_species, len(nbr)), dtype=bool)
    for q in range(params.n_species):
        mask[q] = asp[nbr] == q
    species = (asp[site_atom] + 1).astype(np.int32).reshape(-1, 1)
    mult = np.ones(len(site_atom), dtype=np.int32)
    return Workload(params, cell, pos, asp, site_atom, n_neigh, nbr, vec, rjs, thetas, phis, mask, species, mult)


def strided_sites(n_all: int, n: int, offset: int = 0):
    """n positions spread evenly over a site list of n_all entries (every (n_all // n)-th, from offset)."""
    step = max(1, n_all // n)
    return (offset + step * np.arange(n)) % n_all


def slab_sites(cfg, n_dense: int, n_gas: int):
    """Positions in the site list of a slab box (``BoxSpec.slab``; every atom a site): n_dense spread over
    the dense half and n_gas over the gas half. Random sequential addition places the dense atoms
    first, so the dense half is the first ``round(V/2 * dense_density)`` atoms."""
    spec, _ = CONFIGS[cfg] if isinstance(cfg, int) else cfg
    if spec.slab is None or spec.central is not None:
        raise ValueError("slab_sites needs a slab box whose atoms are all sites")
    dense, gas = spec.slab
    V = spec.n_atoms / (0.5 * (dense + gas))
    nd = int(round(0.5 * V * dense))
    return np.concatenate([strided_sites(nd, n_dense), nd + strided_sites(spec.n_atoms - nd, n_gas)])


def normal(seed: int, shape):
    out = np.zeros(int(np.prod(shape)))
    _lib().stw_normal(seed, out.size, _p(out))
    return out.reshape(shape)


def compress_trivial(alpha_max, l_max):
    """"trivial" compression indices (src/soap_turbo_compress.f90:86-119), 1-based like the reference."""
    pivot = np.cumsum([0] + list(alpha_max[:-1])) + 1
    n_max = int(sum(alpha_max))
    P_i, P_j = [], []
    counter = 0
    k = 1
    for n in range(1, n_max + 1):
        for m in range(n, n_max + 1):
            for _ in range(l_max + 1):
                if n in pivot or m in pivot:
                    counter += 1
                    P_i.append(counter)
                    P_j.append(k)
                k += 1
    return (np.array(P_i, dtype=np.int32), np.array(P_j, dtype=np.int32), np.ones(counter))
