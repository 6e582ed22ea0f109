"""Throughput of get_soap descriptors + forces on one MI355X.

Workload (default): 3 species (Z = 1, 6, 8 at 0.5/0.3/0.2), 4096 atoms in a
periodic box at 0.09 atoms/A^3 (1.0 A minimum distance), the Z = 6 atoms as
sites; alpha_max = {8, 8, 8}, l_max = 8, poly3gauss, radial_enhancement = 1,
"trivial" compression (n_soap = 621); dE/dq ~ N(0, 1) from a fixed seed.

One step = one pass of a path over the whole workload with the inputs
already resident in HBM:
* FP32 adjoint path: descriptors, then forces and virial from dE/dq through
  the adjoint kernel, without forming soap_cart_der (the headline value);
* FP64 full-Jacobian path: get_soap with soap_cart_der, then the caller's
  contraction of soap_cart_der with dE/dq to pair gradients, forces, virial.

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


def measured_traffic(kernel, path=os.path.join(ROOT, "profiles", "config2_hbm_traffic.md")):
    """HBM bytes per launch of `kernel` on config 2, from the committed rocprofv3 counter summary
    (FETCH_SIZE and WRITE_SIZE runs, each scaled by a copy kernel of known size); None if absent."""
    try:
        for line in open(path):
            if line.startswith(f"traffic_bytes_per_launch {kernel}:"):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


def build_problem(cfg):
    import soap_turbo_amd as st
    from soap_turbo_amd.workload import make_workload, normal
    w = make_workload(cfg)
    p = w.params
    keep = st._lib.Keep()
    ps = p.per_species()
    prm = st.make_params(keep, alpha_max=p.alpha_max, l_max=p.l_max, radial_enhancement=p.radial_enhancement,
                         basis=p.basis, scaling_mode=p.scaling_mode, legacy_seed=True, **ps)
    comp = None
    if p.compress_mode == "trivial":
        dim, nz = st.get_compress_indices("trivial", p.alpha_max, p.l_max, "get_dim")
        comp = (np.zeros(nz, np.int32), np.zeros(nz, np.int32), np.zeros(nz))
        st.get_compress_indices("trivial", p.alpha_max, p.l_max, "set_indices", *comp)
    cmp_ = st.make_compression(keep, comp is not None, len(comp[0]) if comp else 0, *(comp or (None, None, None)))
    n_max = sum(p.alpha_max)
    n_soap = int(comp[0].max()) if comp else n_max * (n_max + 1) // 2 * (p.l_max + 1)
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


def cpu_baseline(w, G, cfg):
    """The reference (oracle/_ref, compiled from its sources) on the same workload, all host cores."""
    from oracle import refio
    threads = min(os.cpu_count() or 1, 16)
    if refio.available():
        try:
            o = refio.run(w, G, modes=refio.MODE_TIMING, n_reps=3, threads=threads, timeout=600)
            return {"value": w.n_sites / o["t_full"], "unit": "atoms/s", "cores": o["threads"], "kind": "reference",
                    "sample": f"whole workload ({w.n_sites} sites), reference get_soap with soap_cart_der + "
                              f"contraction with dE/dq, best of 3, OMP_NUM_THREADS={o['threads']}",
                    "contract_path_atoms_per_s": w.n_sites / o["t_contract"]}
        except Exception as e:  # the harness may not run on this host
            log(f"reference CPU baseline unavailable: {e}")
    from oracle import oracle as orc
    import soap_turbo_amd as st
    keep = st._lib.Keep()
    t0 = time.perf_counter()
    W, S = orc.basis_matrices(w)
    orc.get_soap(w, W, S, None, do_der=True)
    dt = time.perf_counter() - t0
    return {"value": w.n_sites / dt, "unit": "atoms/s", "cores": threads, "kind": "port",
            "sample": f"whole workload ({w.n_sites} sites), FP64 C restatement, one pass"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=2, help="workload 1..5 (default 2)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (a known-size kernel against which "
                         "HBM traffic counters are calibrated)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); multi-GPU sharding is not part of this build",
              file=sys.stderr)
        sys.exit(2)

    import soap_turbo_amd as st
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
    ctx = st.Context(0)
    ctx.set_autotune(True)  # launch shapes chosen by timing on the first (warm-up) steps
    ctx.upload_neighbours(w.n_neigh, w.species.T, w.species_multiplicity, w.mask.T, w.rjs, w.thetas, w.phis)
    ctx.upload_atoms(len(w.pos), w.nbr, w.vec)
    ctx.upload_dedq(np.asfortranarray(G.T))

    def step64():
        ctx.full_jacobian_forces(prm, cmp_)

    def step32():
        ctx.adjoint_f32(prm, cmp_)

    k64 = ["radial_f64", "angular_f64", "cnk_f64", "power_f64", "derivatives_f64", "normalise_f64",
           "pairs_full_jacobian_f64", "forces_f64", "fill_rad"]
    log("timing the FP64 full-Jacobian path")
    ms64, wall64, kt64 = time_path(ctx, step64, args.steps, args.warmup, k64)
    f64, v64 = ctx.download_forces(0)

    have32 = True
    try:
        step32()
    except st.SoapTurboError as e:
        log(f"FP32 adjoint path unavailable: {e}")
        have32 = False
    k32 = ["adj_tables", "adj_site_f32", "adj_forces"]
    if have32:
        log("timing the FP32 adjoint path")
        ms32_eager, _, kt32 = time_path(ctx, step32, args.steps, args.warmup, k32)
        # the same steps replayed from a captured HIP graph (the library's graph mode); the per-kernel
        # times above come from uncaptured launches
        ctx.set_graphs(True)
        ms32_graph, _, _ = time_path(ctx, step32, args.steps, args.warmup, [])
        ctx.set_graphs(False)
        # the faster of the two modes is the path's throughput; both are reported
        use_graph = ms32_graph < ms32_eager
        ms32 = ms32_graph if use_graph else ms32_eager
        f32, v32 = ctx.download_forces(1)
        ferr = float(np.abs(f32 - f64).max() / np.abs(f64).max())
        verr = float(np.abs(v32 - v64).max() / np.abs(v64).max())

    n_nb = (w.n_pairs - w.n_sites) / w.n_sites
    p = w.params
    n_max, k_max = sum(p.alpha_max), 1 + p.l_max * (p.l_max + 1) // 2 + p.l_max
    if have32:
        step_ms = ms32 / args.steps
        name = "adj_site_f32"
        tot, nl = kt32[name]
        avg_ms = tot / max(nl, 1)
        # algorithmic FP32 flops per site of the adjoint formulation: per neighbour and (n, k), cnk 4
        # (real x complex FMA) and the vector-Jacobian product 16 (the four complex contractions of
        # f_rad (two), f_azi and f_pol with H, 4 each); per site and (n, n', k), spectrum 4 and H = Y c 4.
        # (The kernel factors the product as u = R H, v = R' H, 8, then a per-k contraction.)
        flops_site = n_nb * 20 * n_max * k_max + 8 * n_max * n_max * k_max
        achieved = flops_site * w.n_sites / (avg_ms * 1e-3) / 1e12
        roofline = {"bound": "mfma", "kernel": name, "achieved": achieved, "peak": FP32_PEAK_TFLOPS,
                    "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS,
                    "traffic": measured_traffic(name) if args.config == 2 else None, "avg_launch_ms": avg_ms}
        metric_path = "FP32 adjoint"
    else:
        step_ms = ms64 / args.steps
        tot, nl = kt64["derivatives_f64"]
        avg_ms = tot / max(nl, 1)
        # algorithmic bytes of the Jacobian kernel: soap_cart_der written (24 n_soap per pair, self included),
        # the spectrum read (8 n_soap per site) and the pair geometry read (24 per pair)
        bytes_launch = w.n_pairs * (24 + 24 * n_soap) + 8 * n_soap * w.n_sites
        achieved = bytes_launch / (avg_ms * 1e-3) / 1e9
        roofline = {"bound": "hbm", "kernel": "derivatives_f64", "achieved": achieved, "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                    "traffic": measured_traffic("derivatives_f64") if args.config == 2 else None,
                    "avg_launch_ms": avg_ms}
        metric_path = "FP64 full-Jacobian (FP32 adjoint unavailable)"
    value = w.n_sites / (step_ms * 1e-3)
    out = {
        "metric": "atoms/s SOAP+forces, FP32 adjoint vs FP64 full-Jacobian path, 1 GPU; max rel err",
        "value": value, "unit": "atoms/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": step_ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if have32 else "f64", "data": "synthetic (seeded random box, dE/dq ~ N(0,1))",
        "config": {"workload": f"config {args.config}: {w.n_sites} sites, {w.n_pairs} pairs, n_max {n_max}, "
                               f"l_max {p.l_max}, n_soap {n_soap}, basis {p.basis}",
                   "sites": w.n_sites, "pairs": w.n_pairs, "n_soap": n_soap, "path": metric_path,
                   "graph": have32 and use_graph, "autotune": True},
        "roofline": roofline,
        "fp64_full_jacobian": {"atoms_per_s": w.n_sites / (ms64 / args.steps * 1e-3), "ms_per_step": ms64 / args.steps,
                               "kernels_ms_per_step": {k: v[0] / max(args.steps, 1) for k, v in kt64.items() if v[1]}},
    }
    if have32:
        out["fp32_adjoint"] = {"atoms_per_s": value, "ms_per_step": step_ms,
                               "ms_per_step_uncaptured": ms32_eager / args.steps,
                               "ms_per_step_graph": ms32_graph / args.steps,
                               "kernels_ms_per_step": {k: v[0] / max(args.steps, 1) for k, v in kt32.items() if v[1]},
                               "speedup_vs_fp64": (ms64 / ms32),
                               "max_rel_err_forces_vs_fp64": ferr, "max_rel_err_virial_vs_fp64": verr}
    if have32 and args.config <= 2:
        out["gap_energy_forces"] = time_gap(ctx, w, prm, cmp_, n_soap, args.steps, args.warmup)
    if args.config <= 2:
        out["central_hessian"] = time_hessian(w, comp, n_soap, max(3, args.steps // 4), not args.no_cpu)
    out["neighbour_list"] = time_neighbours(w)
    if not args.no_cpu:
        log("CPU baseline")
        out["cpu_baseline"] = cpu_baseline(w, G, args.config)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
