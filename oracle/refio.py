"""Input/output of ``oracle/_ref/ref_harness`` (the reference compiled from its sources).

Test infrastructure only: tests, the fixture generator under ``tests/golden/``
and the CPU-baseline leg of ``bench.py`` use it as the checker. The binary
layout is the one ``oracle/ref_harness.f90`` reads and writes (Fortran stream
access, column-major arrays).
"""
from __future__ import annotations

import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
HARNESS = os.path.join(HERE, "_ref", "ref_harness")

MODE_DESC = 1       # get_soap, no derivatives
MODE_FULL = 2       # get_soap with soap_cart_der
MODE_CONTRACT = 4   # get_soap(contract=.true.) + soap_contract
MODE_EXPANSION = 8  # radial and angular expansions alone
MODE_TIMING = 16    # wall-clock of the full-Jacobian and contract paths
MODE_HESSIAN = 32   # get_soap_central_hessian for weight vectors, its self-test, radial second derivatives


def compress_code(mode):
    """The harness's header code of a compress_mode: 0 none, 1 "trivial", 2 + 3 nu_R + nu_S for "nu_R_nu_S"."""
    if mode in (None, "", "none"):
        return 0
    if mode in ("trivial", "Trivial"):
        return 1
    return 2 + 3 * int(mode[0]) + int(mode[2])


def available() -> bool:
    return os.path.exists(HARNESS)


def write_input(path, w, G=None, modes=MODE_FULL, legacy_seed=True, n_reps=1, vecs=None):
    p = w.params
    ps = p.per_species()
    ns = p.n_species
    hdr = np.zeros(16, dtype=np.int32)
    hdr[0] = 1
    hdr[1] = w.n_sites
    hdr[2] = w.n_pairs
    hdr[3] = ns
    hdr[4] = p.l_max
    hdr[5] = 1
    hdr[6] = p.radial_enhancement
    hdr[7] = {"poly3": 0, "poly3gauss": 1, "poly3operator": 2}[p.basis]
    hdr[8] = 0 if p.scaling_mode == "polynomial" else 1
    hdr[9] = compress_code(p.compress_mode)
    hdr[10] = 1 if legacy_seed else 0
    hdr[11] = modes
    hdr[12] = n_reps
    hdr[13] = 0 if vecs is None else vecs.shape[1]
    hdr[14] = 1 if G is not None else 0
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        f.write(np.asarray(p.alpha_max, dtype=np.int32).tobytes())
        for key in ("rcut_hard", "rcut_soft", "nf", "global_scaling", "atom_sigma_r", "atom_sigma_r_scaling",
                    "atom_sigma_t", "atom_sigma_t_scaling", "amplitude_scaling", "central_weight"):
            f.write(np.asarray(ps[key], dtype=np.float64).tobytes())
        f.write(np.asarray(w.n_neigh, dtype=np.int32).tobytes())
        f.write(np.asarray(w.species[:, 0], dtype=np.int32).tobytes())
        f.write(np.asarray(w.species_multiplicity, dtype=np.int32).tobytes())
        f.write(np.asarray(w.mask, dtype=np.int32).tobytes())      # (ns, n_pairs) C == mask(n_pairs, ns) F
        for a in (w.rjs, w.thetas, w.phis):
            f.write(np.asarray(a, dtype=np.float64).tobytes())
        if G is not None:
            f.write(np.ascontiguousarray(G, dtype=np.float64).tobytes())   # (n_soap, n_sites) C == G(n_sites, n_soap) F
        if vecs is not None:
            # (n_sites, n_vec, n_soap) C == vecs(n_soap, n_vec, n_sites) F
            f.write(np.ascontiguousarray(vecs, dtype=np.float64).tobytes())


def read_output(path, w, modes, compress, with_G, n_vec=0):
    out = {}
    with open(path, "rb") as f:
        buf = f.read()
    off = 0

    def take(dtype, count):
        nonlocal off
        a = np.frombuffer(buf, dtype=dtype, count=count, offset=off)
        off += a.nbytes
        return a.copy()

    n_max, k_max, n_soap, P_nonzero = take(np.int32, 4)
    out.update(n_max=int(n_max), k_max=int(k_max), n_soap=int(n_soap), P_nonzero=int(P_nonzero))
    if compress:
        out["P_i"] = take(np.int32, P_nonzero)
        out["P_j"] = take(np.int32, P_nonzero)
        out["P_el"] = take(np.float64, P_nonzero)
    out["W"] = take(np.float64, n_max * n_max).reshape(n_max, n_max).T.copy()
    out["S"] = take(np.float64, n_max * n_max).reshape(n_max, n_max).T.copy()
    ns_, np_ = w.n_sites, w.n_pairs
    if modes & MODE_DESC:
        out["soap_desc"] = take(np.float64, n_soap * ns_).reshape(ns_, n_soap)
    if modes & MODE_FULL:
        out["soap"] = take(np.float64, n_soap * ns_).reshape(ns_, n_soap)
        out["soap_cart_der"] = take(np.float64, 3 * n_soap * np_).reshape(np_, n_soap, 3)
    if modes & MODE_CONTRACT:
        out["soap_contract_q"] = take(np.float64, n_soap * ns_).reshape(ns_, n_soap)
        out["pair"] = take(np.float64, 3 * np_).reshape(np_, 3)
    if modes & MODE_EXPANSION:
        out["rad"] = take(np.float64, n_max * np_).reshape(np_, n_max)
        out["rad_der"] = take(np.float64, n_max * np_).reshape(np_, n_max)
        for key in ("ang", "ang_rad", "ang_azi", "ang_pol"):
            out[key] = take(np.complex128, k_max * np_).reshape(np_, k_max)
    if modes & MODE_HESSIAN:
        out["soap_hess"] = take(np.float64, n_soap * ns_).reshape(ns_, n_soap)
        # hess(a, b, v, k) F -> (n_pairs, n_vec, a, b)
        out["hess"] = take(np.float64, 9 * n_vec * np_).reshape(np_, n_vec, 3, 3).transpose(0, 1, 3, 2).copy()
        out["hess_selftest_err"] = float(take(np.float64, 1)[0])
        out["rad_der2"] = take(np.float64, n_max * np_).reshape(np_, n_max)
    if modes & MODE_TIMING:
        out["t_full"], out["t_contract"] = take(np.float64, 2)
        out["threads"] = int(take(np.int32, 1)[0])
    return out


def run(w, G=None, modes=MODE_FULL, legacy_seed=True, n_reps=1, threads=None, timeout=3600, vecs=None):
    """Runs the reference on workload ``w``; returns a dict of its outputs (numpy C order)."""
    if not available():
        raise RuntimeError("oracle/_ref/ref_harness is not built (make -C oracle ref)")
    env = dict(os.environ)
    if threads is not None:
        env["OMP_NUM_THREADS"] = str(threads)
    with tempfile.TemporaryDirectory() as td:
        fi, fo = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        write_input(fi, w, G, modes, legacy_seed, n_reps, vecs)
        r = subprocess.run([HARNESS, fi, fo], env=env, cwd=td, capture_output=True, text=True, timeout=timeout)
        if r.returncode != 0:
            raise RuntimeError(f"ref_harness failed ({r.returncode}): {r.stdout[-2000:]} {r.stderr[-2000:]}")
        out = read_output(fo, w, modes, compress_code(w.params.compress_mode) != 0, G is not None,
                          0 if vecs is None else vecs.shape[1])
        out["stdout"] = r.stdout
    return out
