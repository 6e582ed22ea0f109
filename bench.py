"""Throughput of get_soap descriptors + forces on one MI355X.

Workload (default, ``--config 5``): 262144 atoms of 4 species (Z = 1, 6, 7, 8,
uniform) in a periodic box at 0.10 atoms/A^3 (1.0 A minimum distance), every
atom a site; rcut_hard = 6.0, rcut_soft = 5.5 (about 90 neighbours),
alpha_max = {6, 6, 6, 6}, l_max = 8, poly3gauss, "trivial" compression
(n_soap = 810); dE/dq ~ N(0, 1) from a fixed seed.

One step = one pass of a path over the whole workload with the inputs
already resident in HBM:
* FP32 adjoint path: descriptors, then forces and virial from dE/dq through
  the adjoint kernel, without forming soap_cart_der (the headline value);
* FP64 full-Jacobian path: get_soap with soap_cart_der, then the caller's
  contraction of soap_cart_der with dE/dq to pair gradients, forces, virial;
  in batches of sites when soap_cart_der exceeds the free HBM (466 GB here).

Beside the line: the maximum relative errors of both paths against the FP64
oracle on a sample of the workload's sites; the reference (compiled from its
Fortran sources) on a sample of sites on the host cores; and the 4096-atom
three-species workload (``--config 2``) with the GAP model, the central-atom
Hessian and the neighbour-list builder (a ``config2`` block, or the line's own
extra keys when ``--config 2`` is the workload).

Prints one JSON line (rank 0, N = 1). ``--gpus N`` with N > 1 is refused: this
build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def profile_figure(cfg, key, kernel):
    """A per-launch figure of `kernel` on workload `cfg` from the committed rocprofv3 counter summaries
    (profiles/config<cfg>_hbm_traffic.md: `traffic_bytes_per_launch <kernel>: <bytes>`, from FETCH_SIZE and
    WRITE_SIZE runs each scaled by a copy kernel of known size; `mfma_util <kernel>: <fraction>`, from
    SQ_VALU_MFMA_BUSY_CYCLES over the cycles the SIMDs had); None if absent."""
    path = os.path.join(ROOT, "profiles", f"config{cfg}_hbm_traffic.md")
    try:
        for line in open(path):
            if line.startswith(f"{key} {kernel}:"):
                v = line.split(":")[1].strip()
                return int(v) if key.startswith("traffic") else float(v)
    except OSError:
        pass
    return None


def build_problem(cfg):
    import soap_turbo_amd as st
    from soap_turbo_amd.workload import make_workload, normal
    w = make_workload(cfg)
    p = w.params
    keep = st._lib.Keep()
    prm = st.make_params(keep, alpha_max=p.alpha_max, l_max=p.l_max, radial_enhancement=p.radial_enhancement,
                         basis=p.basis, scaling_mode=p.scaling_mode, legacy_seed=True, **p.per_species())
    comp, cmp_, n_soap = compression(keep, p)
    G = normal(2024 + cfg, (n_soap, w.n_sites))   # C (n_soap, n_sites) == G(n_sites, n_soap)
    return w, keep, prm, cmp_, comp, n_soap, G


def time_path(ctx, step, steps, warmup, kernels):
    """Times `steps` steps with HIP events on the library's stream: once as is (the step time), then
    again with per-kernel events around every launch (the per-kernel times)."""
    for _ in range(warmup):
        step()
    ctx.synchronize()
    ctx.event_record(0)
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    ctx.event_record(1)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    ms = ctx.event_elapsed(0, 1)
    ctx.kernel_timing(True)
    for _ in range(steps):
        step()
    ctx.synchronize()
    ktimes = {k: ctx.kernel_time(k) for k in kernels}
    ctx.kernel_timing(False)
    return ms, wall, ktimes


def time_neighbours(w, reps=20):
    """The workload's neighbour list built on the device (st_build_neighbours: positions in, the resident
    list out, its host synchronisations included), and the host cell-list generator on the same input."""
    import soap_turbo_amd as st
    from soap_turbo_amd.workload import build_neighbours, spherical
    log("timing the device neighbour-list builder")
    rcut = float(np.max(w.params.per_species()["rcut_hard"]))
    ctx = st.Context(0)
    args = (w.pos, w.atom_species + 1, w.params.n_species, w.L, rcut, w.site_atom)
    n_pairs = ctx.build_neighbours(*args)
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.build_neighbours(*args)
    dev_ms = (time.perf_counter() - t0) / reps * 1e3
    t0 = time.perf_counter()
    n_neigh, nbr, vec = build_neighbours(w.pos, w.L, rcut, w.site_atom)
    spherical(vec)
    host_ms = (time.perf_counter() - t0) * 1e3
    return {"atoms": len(w.pos), "sites": w.n_sites, "pairs": n_pairs, "device_ms": dev_ms,
            "host_cell_list_ms": host_ms, "host_threads": 1}


def time_gap(ctx, w, prm, cmp_, n_soap, steps, warmup, n_sparse=1024, zeta=4):
    """One energy-and-forces step of a GAP model on the device (st_gap_energy_forces): a forward-only
    site pass, K = Q T^T and dE/dq = W T by rocBLAS, then the adjoint pass. The sparse points are
    descriptors of this workload's sites (as a fitted model's are), the weights random."""
    log(f"timing the GAP energy model ({n_sparse} sparse points, zeta {zeta})")
    from soap_turbo_amd.workload import normal
    q = ctx.download_soap_f32(n_soap)                           # (n_soap, n_sites)
    idx = np.arange(n_sparse) * 7919 % w.n_sites
    T = q[:, idx] + 0.05 * normal(77, (n_soap, n_sparse))
    T /= np.linalg.norm(T, axis=0, keepdims=True)
    ctx.set_gap_model(T, normal(78, (n_sparse,)), zeta, 0.5)
    energy = []
    ms, _, kt = time_path(ctx, lambda: energy.append(ctx.gap_energy_forces(prm, cmp_)), steps, warmup,
                          ["adj_site_f32", "gap_model", "adj_forces"])
    return {"n_sparse": n_sparse, "zeta": zeta, "ms_per_step": ms / steps, "atoms_per_s": w.n_sites / (ms / steps * 1e-3),
            "kernels_ms_per_step": {k: v[0] / max(steps, 1) for k, v in kt.items() if v[1]},
            "energy": energy[-1]}


def time_hessian(w, comp, n_soap, steps, with_cpu, n_vec=3):
    """The central-atom Hessian (st_central_hessian) for n_vec weight vectors per site (3: the
    components of a dipole model's gradient), FP64, unified radial form. One call runs the FP64
    descriptor pass, the radial second derivatives and the three Hessian kernels, and copies
    hess (3, 3, n_vec, n_pairs) back. The oracle's OpenMP restatement on the host cores is timed
    beside it on the same input."""
    import soap_turbo_amd as st
    from soap_turbo_amd.workload import normal
    log(f"timing the central-atom Hessian ({n_vec} weight vectors)")
    p = w.params
    keep = st._lib.Keep()
    prm = st.make_params(keep, alpha_max=p.alpha_max, l_max=p.l_max, radial_enhancement=p.radial_enhancement,
                         basis=p.basis, scaling_mode=p.scaling_mode, legacy_seed=False, **p.per_species())
    cmp_ = st.make_compression(keep, comp is not None, len(comp[0]) if comp else 0, *(comp or (None, None, None)))
    vecs = normal(91, (n_soap, n_vec, w.n_sites))
    ctx = st.Context(0)
    ctx.upload_neighbours(w.n_neigh, w.species.T, w.species_multiplicity, w.mask.T, w.rjs, w.thetas, w.phis)
    ctx.central_hessian(prm, cmp_, vecs)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        h = ctx.central_hessian(prm, cmp_, vecs)
    wall = (time.perf_counter() - t0) / steps * 1e3
    ctx.kernel_timing(True)
    ctx.central_hessian(prm, cmp_, vecs)
    kt = {k: ctx.kernel_time(k)[0] for k in ("radial_f64", "angular_f64", "cnk_f64", "power_f64", "radial_d2_f64",
                                              "hessian_f64")}
    ctx.kernel_timing(False)
    out = {"n_vec": n_vec, "ms_per_call": wall, "atoms_per_s": w.n_sites / (wall * 1e-3),
           "kernels_ms": kt, "hess_bytes": int(h.nbytes)}
    if with_cpu:
        from oracle import oracle as orc
        W, S = ctx.get_basis(prm, sum(p.alpha_max))
        t0 = time.perf_counter()
        orc.central_hessian(w, W, S, np.ascontiguousarray(vecs.transpose(2, 1, 0)), comp)
        cpu = time.perf_counter() - t0
        out["cpu_port"] = {"ms_per_call": cpu * 1e3, "kind": "port",
                           "threads": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))}
    return out


def cpu_baseline(cfg, n_sample, threads):
    """The reference (oracle/_ref, compiled from its Fortran sources) on n_sample sites spread over
    workload cfg (every neighbour of a kept site present, so each site costs what it costs in the full
    box): get_soap with soap_cart_der, then its contraction with dE/dq; best of 3, OMP_NUM_THREADS =
    threads. Without the harness, the oracle's FP64 C restatement on the same sample."""
    from oracle import refio
    from soap_turbo_amd.workload import make_workload, normal, strided_sites
    n_all = site_count(cfg)
    n_sample = min(n_sample, n_all)
    w = make_workload(cfg, sites=strided_sites(n_all, n_sample, 5))
    sample = f"{w.n_sites} of {n_all} sites (every {max(1, n_all // n_sample)}-th), {w.n_pairs} pairs"
    if refio.available():
        try:
            p = w.params
            n_soap = soap_size(p)
            G = normal(4048 + cfg, (n_soap, w.n_sites))
            o = refio.run(w, G, modes=refio.MODE_TIMING, n_reps=3, threads=threads, timeout=900)
            return {"value": w.n_sites / o["t_full"], "unit": "atoms/s", "cores": o["threads"], "kind": "reference",
                    "sample": sample + ": reference get_soap with soap_cart_der + contraction with dE/dq, "
                              f"best of 3, OMP_NUM_THREADS={o['threads']}",
                    "contract_path_atoms_per_s": w.n_sites / o["t_contract"]}
        except Exception as e:  # the harness may not run on this host
            log(f"reference CPU baseline unavailable: {e}")
    from oracle import oracle as orc
    t0 = time.perf_counter()
    W, S = orc.basis_matrices(w)
    orc.get_soap(w, W, S, None, do_der=True)
    dt = time.perf_counter() - t0
    return {"value": w.n_sites / dt, "unit": "atoms/s", "cores": threads, "kind": "port",
            "sample": sample + ": FP64 C restatement, one pass"}


def site_count(cfg):
    """Sites of workload cfg (all atoms, or the atoms of the central species)."""
    from soap_turbo_amd.workload import CONFIGS, place_atoms
    spec, _ = CONFIGS[cfg]
    if spec.central is None:
        return spec.n_atoms
    _, _, sp = place_atoms(spec)
    return int((sp == spec.central).sum())


def soap_size(p):
    n_max = sum(p.alpha_max)
    if p.compress_mode == "trivial":
        import soap_turbo_amd as st
        return st.get_compress_indices("trivial", p.alpha_max, p.l_max, "get_dim")[0]
    return n_max * (n_max + 1) // 2 * (p.l_max + 1)


def accuracy(cfg, n_sample=256, n_jac=32):
    """Maximum relative errors (max |a - b| / max |b| per array) of both GPU paths against the FP64
    oracle (the same W) on n_sample sites spread over workload cfg: FP32 adjoint q, pair gradients,
    forces, virial; FP64 q, pair gradients, forces, and soap_cart_der on the first n_jac of them."""
    import soap_turbo_amd as st
    from oracle import oracle as orc
    from soap_turbo_amd.workload import make_workload, normal, strided_sites
    log(f"accuracy against the FP64 oracle on {n_sample} sites")
    n_all = site_count(cfg)
    w = make_workload(cfg, sites=strided_sites(n_all, n_sample, 11))
    p = w.params
    keep = st._lib.Keep()
    prm = st.make_params(keep, alpha_max=p.alpha_max, l_max=p.l_max, radial_enhancement=p.radial_enhancement,
                         basis=p.basis, scaling_mode=p.scaling_mode, legacy_seed=True, **p.per_species())
    comp, cmp_, n_soap = compression(keep, p)
    ctx = st.Context(0)
    W, S = ctx.get_basis(prm, sum(p.alpha_max))
    G = normal(55, (n_soap, w.n_sites))
    ctx.upload_neighbours(w.n_neigh, w.species.T, w.species_multiplicity, w.mask.T, w.rjs, w.thetas, w.phis)
    ctx.upload_atoms(len(w.pos), w.nbr, w.vec)
    ctx.upload_dedq(np.asfortranarray(G.T))
    qo, pair_o = orc.contract(w, W, S, G, 1.0, comp, True)
    fo = np.zeros_like(w.pos)
    np.add.at(fo, w.nbr, -pair_o)
    vo = -np.einsum("ka,kb->ab", w.vec, pair_o)
    rel = lambda a, b: float(np.abs(a - b).max() / np.abs(b).max())
    ctx.adjoint_f32(prm, cmp_)
    f32, v32 = ctx.download_forces(1)
    out = {"sites": w.n_sites, "pairs": w.n_pairs,
           "fp32_adjoint": {"q": rel(ctx.download_soap_f32(n_soap).T, qo), "pair_gradients": rel(ctx.download_pairs(1).T, pair_o),
                            "forces": rel(f32.T, fo), "virial": rel(v32, vo)}}
    ctx.full_jacobian_forces(prm, cmp_)
    f64, v64 = ctx.download_forces(0)
    out["fp64_full_jacobian"] = {"q": rel(ctx.download(n_soap)[0].T, qo), "pair_gradients": rel(ctx.download_pairs(0).T, pair_o),
                                 "forces": rel(f64.T, fo), "virial": rel(v64, vo)}
    wj = make_workload(cfg, sites=strided_sites(n_all, n_jac, 13))
    ctx.upload_neighbours(wj.n_neigh, wj.species.T, wj.species_multiplicity, wj.mask.T, wj.rjs, wj.thetas, wj.phis)
    ctx.compute(prm, cmp_, True, False)
    _, cart = ctx.download(n_soap, cart=True)
    _, co, _ = orc.get_soap(wj, W, S, comp, do_der=True)
    out["fp64_full_jacobian"]["soap_cart_der"] = rel(cart.transpose(2, 1, 0), co)
    out["soap_cart_der_sites"] = wj.n_sites
    return out


def compression(keep, p):
    import soap_turbo_amd as st
    comp = None
    if p.compress_mode == "trivial":
        dim, nz = st.get_compress_indices("trivial", p.alpha_max, p.l_max, "get_dim")
        comp = (np.zeros(nz, np.int32), np.zeros(nz, np.int32), np.zeros(nz))
        st.get_compress_indices("trivial", p.alpha_max, p.l_max, "set_indices", *comp)
    cmp_ = st.make_compression(keep, comp is not None, len(comp[0]) if comp else 0, *(comp or (None, None, None)))
    return comp, cmp_, soap_size(p)


def time_paths(w, prm, cmp_, n_soap, G, steps, warmup, cfg):
    """Both paths over the whole workload: the step times, per-kernel times, the FP32 path's roofline
    and its force and virial errors against the FP64 path."""
    import soap_turbo_amd as st
    ctx = st.Context(0)
    ctx.set_autotune(True)  # launch shapes chosen by timing on the first (warm-up) steps
    ctx.upload_neighbours(w.n_neigh, w.species.T, w.species_multiplicity, w.mask.T, w.rjs, w.thetas, w.phis)
    ctx.upload_atoms(len(w.pos), w.nbr, w.vec)
    ctx.upload_dedq(np.asfortranarray(G.T))
    nbat = []
    k64 = ["radial_f64", "angular_f64", "cnk_f64", "power_f64", "derivatives_f64", "normalise_f64",
           "pairs_full_jacobian_f64", "forces_f64", "fill_rad"]
    log("timing the FP64 full-Jacobian path")
    ms64, _, kt64 = time_path(ctx, lambda: nbat.append(ctx.full_jacobian_forces(prm, cmp_)), steps, warmup, k64)
    f64, v64 = ctx.download_forces(0)
    k32 = ["adj_tables", "adj_site_f32", "adj_forces"]
    log("timing the FP32 adjoint path")
    step32 = lambda: ctx.adjoint_f32(prm, cmp_)
    ms32_eager, _, kt32 = time_path(ctx, step32, steps, warmup, k32)
    # the same steps replayed from a captured HIP graph (the library's graph mode); the per-kernel
    # times above come from uncaptured launches
    ctx.set_graphs(True)
    ms32_graph, _, _ = time_path(ctx, step32, steps, warmup, [])
    ctx.set_graphs(False)
    use_graph = ms32_graph < ms32_eager   # the faster of the two modes is the path's throughput
    ms32 = ms32_graph if use_graph else ms32_eager
    f32, v32 = ctx.download_forces(1)
    p = w.params
    n_nb = (w.n_pairs - w.n_sites) / w.n_sites
    n_max, k_max = sum(p.alpha_max), 1 + p.l_max * (p.l_max + 1) // 2 + p.l_max
    tot, nl = kt32["adj_site_f32"]
    avg_ms = tot / max(nl, 1)
    # algorithmic FP32 flops per site of the adjoint formulation: per neighbour and (n, k), cnk 4
    # (real x complex FMA) and the vector-Jacobian product 16 (the four complex contractions of
    # f_rad (two), f_azi and f_pol with H, 4 each); per site and (n, n', k), spectrum 4 and H = Y c 4.
    # (The kernel factors the product as u = R H, v = R' H, 8, then a per-k contraction.)
    flops_site = n_nb * 20 * n_max * k_max + 8 * n_max * n_max * k_max
    achieved = flops_site * w.n_sites / (avg_ms * 1e-3) / 1e12
    roofline = {"bound": "mfma", "kernel": "adj_site_f32", "achieved": achieved, "peak": FP32_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS,
                "traffic": profile_figure(cfg, "traffic_bytes_per_launch", "adj_site_f32"),
                "mfma_util": profile_figure(cfg, "mfma_util", "adj_site_f32"),
                "avg_launch_ms": avg_ms, "flops_per_site": flops_site}
    # the FP64 Jacobian kernel against HBM: soap_cart_der written (24 n_soap per pair, self included), the
    # spectrum read (8 n_soap per site) and the pair geometry read (24 per pair), per step
    jac_ms = kt64["derivatives_f64"][0] / max(steps, 1)
    jac_bytes = w.n_pairs * (24 + 24 * n_soap) + 8 * n_soap * w.n_sites
    kper = lambda kt: {k: v[0] / max(steps, 1) for k, v in kt.items() if v[1]}
    fp64 = {"atoms_per_s": w.n_sites / (ms64 / steps * 1e-3), "ms_per_step": ms64 / steps, "batches": nbat[-1],
            "kernels_ms_per_step": kper(kt64),
            "jacobian_hbm": {"bytes_per_step": jac_bytes, "achieved_GBs": jac_bytes / (jac_ms * 1e-3) / 1e9,
                             "frac": jac_bytes / (jac_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                             "traffic": profile_figure(cfg, "traffic_bytes_per_launch", "derivatives_f64")}}
    fp32 = {"atoms_per_s": w.n_sites / (ms32 / steps * 1e-3), "ms_per_step": ms32 / steps,
            "ms_per_step_uncaptured": ms32_eager / steps, "ms_per_step_graph": ms32_graph / steps,
            "graph": use_graph, "kernels_ms_per_step": kper(kt32), "speedup_vs_fp64": ms64 / ms32,
            "max_rel_err_forces_vs_fp64": float(np.abs(f32 - f64).max() / np.abs(f64).max()),
            "max_rel_err_virial_vs_fp64": float(np.abs(v32 - v64).max() / np.abs(v64).max())}
    return ctx, fp32, fp64, roofline


def config2_extras(steps, warmup, with_cpu):
    """The 4096-atom three-species workload: both paths, the GAP model, the central-atom Hessian and the
    neighbour-list builder."""
    import soap_turbo_amd as st
    log("config 2: building")
    w, keep, prm, cmp_, comp, n_soap, G = build_problem(2)
    ctx, fp32, fp64, roofline = time_paths(w, prm, cmp_, n_soap, G, steps, warmup, 2)
    out = {"workload": f"config 2: {w.n_sites} sites, {w.n_pairs} pairs, n_soap {n_soap}",
           "fp32_adjoint": fp32, "fp64_full_jacobian": fp64, "roofline": roofline}
    out["gap_energy_forces"] = time_gap(ctx, w, prm, cmp_, n_soap, steps, warmup)
    out["central_hessian"] = time_hessian(w, comp, n_soap, max(3, steps // 4), with_cpu)
    out["neighbour_list"] = time_neighbours(w)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=5, help="workload 1..5 (default 5)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-extras", action="store_true", help="skip the accuracy sample and the config-2 extras")
    ap.add_argument("--cpu-sites", type=int, default=8192, help="sites of the CPU baseline's sample")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (a known-size kernel against which "
                         "HBM traffic counters are calibrated)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); multi-GPU sharding is not part of this build",
              file=sys.stderr)
        sys.exit(2)

    import soap_turbo_amd as st  # noqa: F401  (fails loudly without the HIP library)
    if args.copy_mib > 0:
        import torch
        n = args.copy_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copied {args.copy_mib} MiB device to device")
        del src, dst
    log(f"building workload {args.config}")
    w, keep, prm, cmp_, comp, n_soap, G = build_problem(args.config)
    log(f"sites {w.n_sites} pairs {w.n_pairs} n_soap {n_soap}")
    ctx, fp32, fp64, roofline = time_paths(w, prm, cmp_, n_soap, G, args.steps, args.warmup, args.config)
    extras = {}
    if args.config == 2 and not args.no_extras:
        # the rows after the path on this workload itself: GAP model, Hessian, neighbour list
        extras["gap_energy_forces"] = time_gap(ctx, w, prm, cmp_, n_soap, args.steps, args.warmup)
        extras["central_hessian"] = time_hessian(w, comp, n_soap, max(3, args.steps // 4), not args.no_cpu)
        extras["neighbour_list"] = time_neighbours(w)
    del ctx
    p = w.params
    n_max = sum(p.alpha_max)
    out = {
        "metric": "atoms/s SOAP+forces, FP32 adjoint vs FP64 full-Jacobian path, 1 GPU; max rel err",
        "value": fp32["atoms_per_s"], "unit": "atoms/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": fp32["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded random box, dE/dq ~ N(0,1))",
        "config": {"workload": f"config {args.config}: {w.n_sites} sites, {w.n_pairs} pairs, n_max {n_max}, "
                               f"l_max {p.l_max}, n_soap {n_soap}, basis {p.basis}",
                   "sites": w.n_sites, "pairs": w.n_pairs, "n_soap": n_soap, "path": "FP32 adjoint",
                   "graph": fp32["graph"], "autotune": True},
        "roofline": roofline,
        "fp32_adjoint": fp32,
        "fp64_full_jacobian": fp64,
    }
    out.update(extras)
    if not args.no_extras:
        out["max_rel_err_vs_fp64_oracle"] = accuracy(args.config)
    threads = min(os.cpu_count() or 1, 16)
    out["host"] = {"nproc": os.cpu_count(), "threads_used": threads,
                   "note": "the CPU baseline uses min(nproc, 16) threads: 16 is this GPU's share of the host"}
    if not args.no_cpu:
        log(f"CPU baseline ({args.cpu_sites} sites)")
        out["cpu_baseline"] = cpu_baseline(args.config, args.cpu_sites, threads)
    if not args.no_extras and args.config != 2:
        out["config2"] = config2_extras(args.steps, args.warmup, not args.no_cpu)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
