"""Benchmark: SOAP create() + position gradients (atom-envs/s) on one MI355X.

Default workload (--config water_box): periodic bulk water, 32,768 H2O = 98,304 atoms at
1.0 g/cm^3 (cubic cell L = 99.338 A), molecules on a jittered 32^3 lattice with random
orientations (seed 0), species H/O, GTO basis, r_cut=6.0, sigma=0.5, n_max=8, l_max=6,
average="off", float32 output.  One step = create() of every center plus the forces of a
seeded Gaussian linear readout w.p summed over all centers, computed by the adjoint
kernels (descriptor_and_vjp: one neighbour search, one coefficient pass, no dense dp/dR,
which would be 98,304^2 x 3 x 952 values).  Positions are resident in HBM when the timed
region starts.

The timed form is --precision/--radial (default "mixed" + "learned": the trained radial
networks, gated end to end against the exact radial functions, tabulated; float32
harmonics and accumulation).  Next to it the bench times the plain float32 form (the
reference algorithm in float32: per-edge exponentials, per-center Loewdin contraction),
the plain float64 form (the reference's internal arithmetic) and the tabulated exact
radial form, and reports every form's errors against the float64 CPU oracle on the
first --cpu-centers centers (descriptor rows, and the forces of the readout summed over
those centers).  That oracle run is also the CPU baseline (all host cores, one process
per core, linearly extrapolated to the metric's unit).

Other workloads:
  hcno_box  periodic H/C/N/O box, 20,000 atoms at 0.10 A^-3, r_cut=5, sigma=0.4, n_max=8, l_max=6
  h2o_1m    non-periodic (H2O)_333334 cluster, 1,000,002 atoms, r_cut=6, sigma=0.5, n_max=7, l_max=3
  h2o_1000  non-periodic (H2O)_1000 cluster, 3,000 atoms, r_cut=10, n_max=7, l_max=3: create plus
            the full dense (3000, 3000, 3, 420) analytical Jacobian

python bench.py [--gpus 1] [--steps K] [--warmup W] [--config NAME] [--precision P] [--radial R]
Prints one JSON line.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

METRIC = "atom-envs/s (create + dp/dR) at ≤1e-4 rel err: mixed/learned vs plain FP64 on GPU"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
HERE = os.path.dirname(os.path.abspath(__file__))
PROFILES = os.path.join(HERE, "profiles")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse_args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="water_box", choices=["water_box", "hcno_box", "h2o_1m", "h2o_1000"])
    ap.add_argument("--precision", default="mixed", choices=["float64", "mixed", "float32"])
    ap.add_argument("--radial", default="learned", choices=["exact", "table", "learned", "chebyshev"],
                    help="radial functions of the timed form: exact per edge, the tabulated exact functions, "
                         "or the learned model (trained and gated first)")
    ap.add_argument("--radial-checkpoint", default="", metavar="NPZ",
                    help="learned radial networks: load this checkpoint if it exists, else train and save it")
    ap.add_argument("--spectrum", default="exact", choices=["exact", "fp16", "fp16x3", "bf16", "bf16x3"],
                    help="power-spectrum contraction of the timed form (matrix-core modes)")
    ap.add_argument("--no-beside", action="store_true", help="time the chosen form only")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-centers", type=int, default=4096)
    ap.add_argument("--cpu-workers", type=int, default=0, help="oracle processes (default: host cores, at most 16)")
    ap.add_argument("--no-float64-reference", action="store_true",
                    help="(h2o_1000) skip timing the plain float64 form next to the chosen one")
    ap.add_argument("--copy-calibration", type=float, default=0.0, metavar="GB",
                    help="only run one plain device-to-device copy of GB gigabytes (PMC counter calibration)")
    return ap.parse_args()


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cores() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def traffic_from_profile(name):
    """HBM bytes per launch of a kernel from a committed rocprofv3 PMC summary, or None."""
    path = os.path.join(PROFILES, name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1].strip())
    return None


# ---- workloads --------------------------------------------------------------------------

ADJ_CONFIGS = {
    "water_box": ("periodic bulk water, 32768 H2O = 98,304 atoms, 1.0 g/cm^3, L=99.338 A, H/O, GTO, "
                  "r_cut=6.0, sigma=0.5, n_max=8, l_max=6, average=off; create + forces of a Gaussian "
                  "linear readout w.p by the adjoint (no dense dp/dR)",
                  dict(r_cut=6.0, n_max=8, l_max=6, sigma=0.5, species=[1, 8], periodic=True)),
    "hcno_box": ("periodic H/C/N/O box, 20,000 atoms (0.50/0.25/0.10/0.15) at 0.10 A^-3, d_min 0.9 A, GTO, "
                 "r_cut=5.0, sigma=0.4, n_max=8, l_max=6; create + adjoint forces of w.p",
                 dict(r_cut=5.0, n_max=8, l_max=6, sigma=0.4, species=[1, 6, 7, 8], periodic=True)),
    "h2o_1m": ("non-periodic (H2O)_333334 cluster, 1,000,002 atoms, H/O, GTO, r_cut=6.0, sigma=0.5, n_max=7, "
               "l_max=3, average=off; create + adjoint forces of w.p",
               dict(r_cut=6.0, n_max=7, l_max=3, sigma=0.5, species=[1, 8])),
}


def make_system(name):
    from cusoap_amd import systems
    if name == "water_box":
        return systems.water_box(32, seed=0)
    if name == "hcno_box":
        return systems.random_box(n_atoms=20000, density=0.10, seed=0)
    return systems.water_cluster_1m(seed=0)


def readout_weights(n_feat):
    return np.random.default_rng(12345).normal(size=n_feat)


# ---- CPU baseline: the float64 oracle over a slice of the centers ------------------------

_ORACLE = {}


def _oracle_chunk(task):
    """Oracle descriptor rows and readout forces of centers [c0, c1) (one worker process)."""
    from threadpoolctl import threadpool_limits
    system, params, c0, c1 = task
    if "o" not in _ORACLE:
        from oracle import soap_oracle as so
        _ORACLE["o"] = so.OracleSOAP(**params)
    o = _ORACLE["o"]
    w = readout_weights(o.get_number_of_features())
    cen = list(range(c0, c1))
    with threadpool_limits(1):
        d = o.create(system, centers=cen)
        F = o.vjp(system, w, centers=cen)
    return c0, d, F


def cpu_baseline(system, params, n_centers, workers):
    """Float64 oracle create + readout forces on the first n_centers centers, chunks of 256 over
    `workers` forked processes (started before this process touches the GPU)."""
    import multiprocessing as mp
    chunk = 256
    tasks = [(system, params, c0, min(c0 + chunk, n_centers)) for c0 in range(0, n_centers, chunk)]
    t0 = time.perf_counter()
    with mp.get_context("fork").Pool(workers) as pool:
        res = pool.map(_oracle_chunk, tasks, chunksize=1)
    dt = time.perf_counter() - t0
    res.sort(key=lambda r: r[0])
    desc = np.concatenate([r[1] for r in res])
    forces = sum(r[2] for r in res)
    return desc, forces, dt


# ---- the large workloads: create + adjoint forces ------------------------------------------

def algorithmic_flops(kernel, S, n, lmax, kbar):
    """Flops per atom environment of one kernel (the hot-path work model: FMA = 2 flops).

    coefficients: forward F = 2 k n L2 + 2 S n^2 L2 + 2 P L2 (P = S n (S n + 1) / 2);
    vjp_edges: the edge pass of the adjoint, 8 k n L2."""
    L2 = (lmax + 1) ** 2
    P = S * n * (S * n + 1) // 2
    if kernel == "coefficients":
        return 2 * kbar * n * L2 + 2 * S * n * n * L2 + 2 * P * L2
    if kernel == "vjp_edges":
        return 8 * kbar * n * L2
    return None


def bench_adjoint(args):
    desc_txt, params = ADJ_CONFIGS[args.config]
    t0 = time.perf_counter()
    system = make_system(args.config)
    t_gen = time.perf_counter() - t0
    n_env = system["positions"].shape[0]
    n_cpu = min(args.cpu_centers, n_env)
    cpu = None
    if not args.no_cpu_baseline and n_cpu > 0:
        workers = args.cpu_workers or min(16, host_cores())
        log(f"cpu baseline: float64 oracle on {n_cpu} centers, {workers} processes")
        desc_ref, F_ref, dt_cpu = cpu_baseline(system, params, n_cpu, workers)
        cpu = {"value": n_cpu / dt_cpu, "unit": "atom-envs/s", "cores": workers, "kind": "port",
               "sample": f"first {n_cpu} of {n_env} centers of the same workload: create + forces of the readout "
                         f"summed over those centers, numpy float64 oracle, {workers} processes of one thread "
                         f"({cpu_model()}), {dt_cpu:.1f} s; the rate extrapolates linearly to all centers"}
        log(f"  {dt_cpu:.1f} s")

    import torch
    from cusoap_amd import SOAP, _lib
    dev = torch.device("cuda:0")
    sys_dev = dict(system, positions=torch.tensor(system["positions"], device=dev))
    soaps = {}
    info = {}

    def soap_for(prec, radial="exact", spectrum="exact"):
        key = (prec, radial, spectrum)
        if key in soaps:
            return soaps[key]
        learned = [v for k, v in soaps.items() if k[1] == "learned"]
        if radial == "learned" and learned:
            soaps[key] = SOAP(device=dev, dtype="float32", precision=prec, radial="learned",
                              surrogate=learned[0].surrogate, spectrum=spectrum, **params)
        elif radial == "learned":
            t1 = time.perf_counter()
            ckpt = args.radial_checkpoint
            src = ckpt if ckpt and os.path.exists(ckpt) else None
            log("training the radial networks" if src is None else f"loading {src}")
            soaps[key] = SOAP(device=dev, dtype="float32", precision=prec, radial="learned", spectrum=spectrum,
                              surrogate=src, **params)
            sur = soaps[key].surrogate
            if ckpt and src is None:
                sur.save(ckpt)
            info.update({"radial_checkpoint_loaded": src is not None,
                         "radial_training_and_gate_s": round(time.perf_counter() - t1, 2),
                         "radial_pairs_learned": [int(u) for u in sur.use_pair],
                         **sur.gate_report()})
            log(f"  {info['radial_training_and_gate_s']} s, pairs learned {info['radial_pairs_learned']}")
        else:
            soaps[key] = SOAP(device=dev, dtype="float32", precision=prec, spectrum=spectrum, radial=radial,
                              **params)
        return soaps[key]

    chosen = (args.precision, args.radial, args.spectrum)
    forms = [chosen]
    if not args.no_beside:
        for f in (("float32", "exact", "exact"), ("float64", "exact", "exact"), ("mixed", "table", "exact"),
                  ("mixed", "chebyshev", "exact")):
            if f not in forms:
                forms.append(f)
    n_feat = soap_for(*chosen).get_number_of_features()
    w = torch.tensor(readout_weights(n_feat), device=dev)
    names = ("coefficients", "vjp_edges", "vjp_adjoint", "edge_reverse", "vjp_gather", "center_order", "nl_count",
             "nl_fill", "power_spectrum")

    # full-box reference: the float64 GPU form (within 1e-7 / 1e-10 of the float64 oracle on
    # its first centers), against which every form's descriptors and forces are checked on
    # all atoms as well
    log("float64 GPU reference on the full box")
    d64_full, F64_full = soap_for("float64", "exact", "exact").descriptor_and_vjp(sys_dev, w)
    d64_full = d64_full.double()
    d64_scale = d64_full.abs().amax(0).clamp_min(1e-300)
    F64_scale = F64_full.abs().max()

    def time_form(form):
        sp = soap_for(*form)
        for _ in range(args.warmup):
            d, F = sp.descriptor_and_vjp(sys_dev, w)
            del d, F
        torch.cuda.synchronize()
        _lib.profile_enable(True)
        for nm in names:
            _lib.profile_get(nm)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            d, F = sp.descriptor_and_vjp(sys_dev, w)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        _lib.profile_enable(False)
        kern = {}
        for nm in names:
            ms, cnt = _lib.profile_get(nm)
            kern[nm] = ms / max(cnt, 1)
        res = {"value": n_env * args.steps / dt, "ms_per_step": dt * 1e3 / args.steps, "kernel_ms": kern}
        if cpu is not None:
            dd = d[:n_cpu].double().cpu().numpy()
            sc = np.abs(desc_ref).max(0)
            res["desc_max_col_rel_err"] = float((np.abs(dd - desc_ref).max(0) / np.where(sc > 0, sc, 1.0)).max())
            _, Fs = sp.descriptor_and_vjp(sys_dev, w, centers=list(range(n_cpu)))
            res["force_max_rel_err"] = float(np.abs(Fs.cpu().numpy() - F_ref).max() / np.abs(F_ref).max())
        res["full_box_vs_gpu_float64"] = {
            "desc_max_col_rel_err": float(((d.double() - d64_full).abs().amax(0) / d64_scale).max()),
            "force_max_rel_err": float((F - F64_full).abs().max() / F64_scale)}
        del d, F
        return res

    results = {}
    for f in forms:
        log("timing", ":".join(f))
        results[f] = time_form(f)
    main_res = results[chosen]
    kern = main_res["kernel_ms"]
    # the same step with host (numpy) positions: each call copies them to the device
    sp = soap_for(*chosen)
    sys_host = dict(system, positions=np.ascontiguousarray(system["positions"], dtype=np.float64))
    for _ in range(max(args.warmup, 1)):
        d, F = sp.descriptor_and_vjp(sys_host, w)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        d, F = sp.descriptor_and_vjp(sys_host, w)
    torch.cuda.synchronize()
    dt_host = time.perf_counter() - t0
    del d, F
    api_host = {"value": round(n_env * args.steps / dt_host, 1), "ms_per_step": round(dt_host * 1e3 / args.steps, 3),
                "note": "positions passed as a float64 host array: includes the host-to-device copy each step"}
    # beside the chosen step (skipped with --no-beside): the atoms in random order, and the
    # fixed-order force sums
    shuffled = ordered = None
    if not args.no_beside:
        # the same step with the atoms in random order (the edge pass groups centers along a
        # Morton curve of their coordinates, so its speed should not depend on the order)
        perm = np.random.default_rng(17).permutation(n_env)
        sys_shuf = dict(system, positions=torch.tensor(np.asarray(system["positions"])[perm], device=dev),
                        atomic_numbers=np.asarray(system["atomic_numbers"])[perm])
        for _ in range(max(args.warmup, 1)):
            d, F = sp.descriptor_and_vjp(sys_shuf, w)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            d, F = sp.descriptor_and_vjp(sys_shuf, w)
        torch.cuda.synchronize()
        dt_shuf = time.perf_counter() - t0
        del d, F, sys_shuf
        shuffled = {"value": round(n_env * args.steps / dt_shuf, 1), "ms_per_step": round(dt_shuf * 1e3 / args.steps, 3),
                    "note": "the same box with its atoms in random order"}
        # the same step with fixed-order force sums (opt-in: forces the same bits on every run)
        sp.ordered_forces = True
        log("timing", ":".join(chosen), "with ordered force sums")
        r_ord = time_form(chosen)
        F1 = sp.descriptor_and_vjp(sys_dev, w)[1]
        F2 = sp.descriptor_and_vjp(sys_dev, w)[1]
        sp.ordered_forces = False
        F0 = sp.descriptor_and_vjp(sys_dev, w)[1]
        ordered = {"value": round(r_ord["value"], 1), "ms_per_step": round(r_ord["ms_per_step"], 3),
                   "kernel_ms": {k: round(v, 4) for k, v in r_ord["kernel_ms"].items()},
                   "forces_bitwise_equal_on_repeat": bool(torch.equal(F1, F2)),
                   "max_rel_diff_vs_atomic_sums": float((F1 - F0).abs().max() / F0.abs().max())}
        del F0, F1, F2
    nl = soap_for(*chosen)._neighbors(soap_for(*chosen)._context(sys_dev, None), drop_self=False)
    kbar = nl.n_edges / n_env
    dom = max(("coefficients", "vjp_edges"), key=lambda k: kern.get(k, 0.0))
    S, n, lmax = len(params["species"]), params["n_max"], params["l_max"]
    fl = algorithmic_flops(dom, S, n, lmax, kbar) * n_env
    achieved = fl / (kern[dom] * 1e-3) / 1e12
    traffic = traffic_from_profile(f"{dom}_{args.config}_traffic.txt")

    def form_name(f):
        return ":".join(x for x in f if x != "exact") or f[0]

    out = {
        "metric": METRIC,
        "value": round(main_res["value"], 1),
        "unit": "atom-envs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(main_res["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision != "float64" and args.radial != "exact" else
                 {"float64": "f64", "mixed": "f64+f32", "float32": "f32"}[args.precision],
        "data": "synthetic",
        "config": {"workload": desc_txt, "n_atoms": n_env, "n_features": n_feat, "edges_per_center": round(kbar, 2),
                   "precision": args.precision, "radial": args.radial, "spectrum": args.spectrum, **info,
                   "system_generation_s": round(t_gen, 1), "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 2), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "traffic": traffic,
                     "kernel": dom, "peak_basis": "FP32 dense (compute-bound kernel)",
                     "flops_per_launch": fl, "avg_kernel_ms": round(kern[dom], 4)},
        "kernel_ms": {k: round(v, 4) for k, v in kern.items()},
        "api_host_positions": api_host,
        **({"shuffled_atoms": shuffled, "ordered_force_sums": ordered} if shuffled is not None else {}),
        "forms": {form_name(f): {"value": round(r["value"], 1), "ms_per_step": round(r["ms_per_step"], 3),
                                 **{k: r[k] for k in ("desc_max_col_rel_err", "force_max_rel_err",
                                                      "full_box_vs_gpu_float64") if k in r},
                                 "kernel_ms": {k: round(v, 4) for k, v in r["kernel_ms"].items()}}
                  for f, r in results.items()},
    }
    if ("float32", "exact", "exact") in results:
        out["speedup_vs_plain_float32"] = round(main_res["value"] / results[("float32", "exact", "exact")]["value"], 3)
    if cpu is not None:
        out["error_reference"] = (f"float64 CPU oracle on the first {n_cpu} centers; full_box_vs_gpu_float64: "
                                  f"the float64 GPU form on all {n_env} atoms")
        out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)


# ---- (H2O)_1000: create + the dense Jacobian ----------------------------------------------

def bench_dense(args):
    import torch
    from cusoap_amd import SOAP, _lib, systems
    dev = torch.device("cuda:0")
    params = dict(r_cut=10.0, n_max=7, l_max=3, sigma=0.5, species=[1, 8], rbf="gto", average="off")
    system = systems.water_cluster_3000(seed=0)
    sys_dev = {"positions": torch.tensor(system["positions"], device=dev),
               "atomic_numbers": system["atomic_numbers"]}
    n_env = system["positions"].shape[0]
    names = ("jacobian_dense", "jacobian_edges", "coefficients", "nl_count", "nl_fill", "power_spectrum")

    def run(precision):
        soap = SOAP(device=dev, dtype="float32", precision=precision, **params)

        def step():
            d = soap.create(sys_dev)
            J = soap.derivatives(sys_dev, method="analytical", return_descriptor=False)
            return d, J

        for _ in range(args.warmup):
            d, J = step()
            del d, J
        torch.cuda.synchronize()
        _lib.profile_enable(True)
        for name in names:
            _lib.profile_get(name)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            d, J = step()
            del d, J
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        _lib.profile_enable(False)
        kern = {}
        for name in names:
            ms, n = _lib.profile_get(name)
            kern[name] = ms / max(n, 1)
        return soap, step, dt, kern

    prec = args.precision
    soap, step, dt, kern = run(prec)
    n_feat = soap.get_number_of_features()
    ref_form = None
    if prec != "float64" and not args.no_float64_reference:
        d, J = step()
        soap64, step64, dt64, kern64 = run("float64")
        d64, J64 = step64()
        err_d = float(((d.double() - d64.double()).abs().amax(0) / d64.double().abs().amax(0).clamp_min(1e-300)).max())
        num = torch.zeros(n_feat, dtype=torch.float64, device=dev)
        den = torch.zeros(n_feat, dtype=torch.float64, device=dev)
        for c0 in range(0, n_env, 100):                     # per-feature column max over (center, atom, xyz)
            a, b = J[c0:c0 + 100].double(), J64[c0:c0 + 100].double()
            num = torch.maximum(num, (a - b).abs().amax(dim=(0, 1, 2)))
            den = torch.maximum(den, b.abs().amax(dim=(0, 1, 2)))
        err_j = float((num / den.clamp_min(1e-300)).max())
        del d, J, d64, J64
        ref_form = {"precision": "float64", "value": round(n_env * args.steps / dt64, 1),
                    "ms_per_step": round(dt64 * 1e3 / args.steps, 3),
                    "kernel_ms": {k: round(v, 4) for k, v in kern64.items()},
                    "chosen_form_max_col_rel_err": {"descriptor": err_d, "jacobian": err_j},
                    "speedup_of_chosen_form": round(dt64 / dt, 3)}
    jac_bytes = float(n_env) * n_env * 3 * n_feat * 4          # algorithmic: every output element stored once
    achieved = jac_bytes / (kern["jacobian_dense"] * 1e-3) / 1e9
    out = {
        "metric": METRIC, "value": round(n_env * args.steps / dt, 1), "unit": "atom-envs/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt * 1e3 / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"float64": "f64", "mixed": "f64+f32", "float32": "f32"}[prec], "data": "synthetic",
        "config": {"workload": "(H2O)_1000 non-periodic cluster, 3000 atoms, H/O, GTO, r_cut=10.0, sigma=0.5, "
                               "n_max=7, l_max=3, average=off; create + dense analytical dp/dR "
                               "(3000x3000x3x420 float32)",
                   "n_atoms": n_env, "n_features": n_feat, "precision": prec, "output_dtype": "float32",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic_from_profile("jacobian_dense_h2o1000_traffic.txt"),
                     "kernel": "dense_fused_kernel (jacobian_dense)", "bytes_per_launch": jac_bytes,
                     "avg_kernel_ms": round(kern["jacobian_dense"], 4)},
        "kernel_ms": {k: round(v, 4) for k, v in kern.items()},
    }
    if ref_form is not None:
        out["plain_float64_form"] = ref_form
    print(json.dumps(out), flush=True)


def main():
    args = parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path (centers are independent; "
              "a spatial split with a halo exchange of forces is the design for several GPUs)", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration > 0:
        import torch
        n = int(args.copy_calibration * 1e9) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n}), flush=True)
        return
    if args.config == "h2o_1000":
        return bench_dense(args)
    return bench_adjoint(args)


if __name__ == "__main__":
    main()
