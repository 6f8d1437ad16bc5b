"""Benchmark: SOAP create() + analytical dp/dR throughput (atom-envs/s) on one MI355X.

Default workload: the non-periodic (H2O)_1000 cluster (3,000 atoms; O sites on a
jittered 10x10x10 lattice of 3.1 A, random orientations, seed 0), species H/O,
GTO basis, r_cut=10.0, sigma=0.5, n_max=7, l_max=3, average="off", float32
output.  One step = one create() of every center plus one
derivatives(method="analytical") call, i.e. the full dense (3000, 3000, 3, 420)
Jacobian, as the reference's serial driver does.  Positions are already resident
in HBM when the timed region starts.  The timed form is --precision (default
"mixed": float64 radial functions, float32 harmonics, accumulation and
single-edge Jacobian rows); the plain float64 form (the reference's internal
arithmetic) is timed next to it and the chosen form's per-feature error against
it is reported ("plain_float64_form").

Other workloads (--config), each step = create() + the adjoint forces of a
seeded Gaussian linear readout (descriptor_and_vjp; no dp/dR is formed):
  water_box  periodic bulk water, 32768 H2O = 98,304 atoms at 1.0 g/cm^3 (L=99.338 A),
             r_cut=6, sigma=0.5, n_max=8, l_max=6
  hcno_box   periodic H/C/N/O box, 20,000 atoms at 0.10 A^-3, r_cut=5, sigma=0.4, n_max=8, l_max=6
  h2o_1m     non-periodic (H2O)_333334 cluster, 1,000,002 atoms, r_cut=6, sigma=0.5, n_max=7, l_max=3

python bench.py [--gpus 1] [--steps K] [--warmup W] [--config NAME] [--precision P] [--no-cpu-baseline]
Prints one JSON line.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

METRIC = "atom-envs/s (create + dp/dR) at ≤1e-4 rel err: mixed/learned vs plain FP64 on GPU"
HBM_PEAK_GBS = 8000.0
HERE = os.path.dirname(os.path.abspath(__file__))
TRAFFIC_FILE = os.path.join(HERE, "profiles", "jacobian_dense_h2o1000_traffic.txt")


def parse_args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", default="mixed", choices=["float64", "mixed", "float32"])
    ap.add_argument("--no-float64-reference", action="store_true",
                    help="(h2o_1000) skip timing the plain float64 form next to the chosen one")
    ap.add_argument("--config", default="h2o_1000", choices=["h2o_1000", "water_box", "hcno_box", "h2o_1m"])
    ap.add_argument("--radial", default="exact", choices=["exact", "table", "learned"],
                    help="radial functions of the large configs: exact per edge, the tabulated exact functions, "
                         "or the learned model (trained first)")
    ap.add_argument("--radial-checkpoint", default="", metavar="NPZ",
                    help="learned radial networks: load this checkpoint if it exists, else train and save it")
    ap.add_argument("--spectrum", default="exact", choices=["exact", "fp16", "fp16x3", "bf16", "bf16x3"],
                    help="power-spectrum contraction of the large configs (matrix-core modes)")
    ap.add_argument("--compare", default="",
                    help="comma list of forms to compare against float64 (large configs): a precision, "
                         "optionally with ':table' or ':learned' and a spectrum mode, e.g. mixed:learned:bf16x3")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-centers", type=int, default=200)
    ap.add_argument("--check", action="store_true", help="compare the descriptor with the FP64 oracle")
    ap.add_argument("--copy-calibration", type=float, default=0.0, metavar="GB",
                    help="only run one plain device-to-device copy of GB gigabytes (PMC counter calibration)")
    return ap.parse_args()


def traffic_from_profile():
    """HBM bytes per launch of the dense Jacobian kernel from the committed rocprofv3 PMC summary."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1].strip())
    return None


def cpu_baseline(system, params, n_centers):
    """The FP64 oracle (numpy, 1 thread) on the first n_centers centers (search over all atoms)."""
    from threadpoolctl import threadpool_limits
    from oracle import soap_oracle as so
    o = so.OracleSOAP(**params)
    cen = list(range(n_centers))
    with threadpool_limits(1):
        t0 = time.perf_counter()
        o.create(system, centers=cen)
        o.derivatives(system, centers=cen, return_descriptor=False)
        dt = time.perf_counter() - t0
    return {"value": n_centers / dt, "unit": "atom-envs/s", "cores": 1, "kind": "port",
            "sample": f"first {n_centers} of 3000 centers of the same (H2O)_1000 workload: create + dense "
                      f"analytical dp/dR, numpy float64 oracle, 1 thread, {dt:.1f} s"}


def main():
    args = parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path (centers are independent; "
              "a spatial split with a halo exchange of forces is the design for several GPUs)", file=sys.stderr)
        sys.exit(2)
    import torch
    from cusoap_amd import SOAP, _lib, systems

    dev = torch.device("cuda:0")
    if args.copy_calibration > 0:
        n = int(args.copy_calibration * 1e9) // 4
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n}), flush=True)
        return
    if args.config != "h2o_1000":
        return bench_adjoint(args, dev)
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

    soap, step, dt, kern = run(args.precision)
    n_feat = soap.get_number_of_features()
    ms_step = dt * 1e3 / args.steps
    ref_form = None
    if args.precision != "float64" and not args.no_float64_reference:
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
    value = n_env * args.steps / dt
    jac_bytes = float(n_env) * n_env * 3 * n_feat * 4          # algorithmic: every output element stored once
    achieved = jac_bytes / (kern["jacobian_dense"] * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "atom-envs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "float64" else ("f64+f32" if args.precision == "mixed" else "f32"),
        "data": "synthetic",
        "config": {"workload": "(H2O)_1000 non-periodic cluster, 3000 atoms, H/O, GTO, r_cut=10.0, sigma=0.5, "
                               "n_max=7, l_max=3, average=off; create + dense analytical dp/dR "
                               "(3000x3000x3x420 float32)",
                   "n_atoms": n_env, "n_features": n_feat, "precision": args.precision,
                   "output_dtype": "float32", "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profile(),
                     "kernel": "dense_kernel (jacobian_dense)", "bytes_per_launch": jac_bytes,
                     "avg_kernel_ms": round(kern["jacobian_dense"], 4)},
        "kernel_ms": {k: round(v, 4) for k, v in kern.items()},
    }
    if ref_form is not None:
        out["plain_float64_form"] = ref_form
    if args.check:
        from oracle import soap_oracle as so
        ref = so.OracleSOAP(**params).create(system)
        d = soap.create(sys_dev).double().cpu().numpy()
        sc = np.abs(ref).max(0)
        out["max_col_rel_err_vs_oracle"] = float((np.abs(d - ref).max(0) / np.where(sc > 0, sc, 1)).max())
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(system, params, args.cpu_centers)
    print(json.dumps(out), flush=True)


ADJ_CONFIGS = {
    "water_box": ("periodic bulk water, 32768 H2O = 98,304 atoms, 1.0 g/cm^3, L=99.338 A, H/O, GTO, "
                  "r_cut=6.0, sigma=0.5, n_max=8, l_max=6, average=off; create + adjoint forces of w.p",
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


def bench_adjoint(args, dev):
    """create + adjoint forces on one of the large workloads."""
    import torch
    from cusoap_amd import SOAP, _lib
    desc_txt, params = ADJ_CONFIGS[args.config]
    t0 = time.perf_counter()
    system = make_system(args.config)
    t_gen = time.perf_counter() - t0
    n_env = system["positions"].shape[0]
    sys_dev = dict(system, positions=torch.tensor(system["positions"], device=dev))
    soaps = {}
    train = {}

    def soap_for(prec, radial="exact", spectrum="exact"):
        key = (prec, radial, spectrum)
        if key not in soaps:
            learned = [v for k, v in soaps.items() if k[1] == "learned" and k[0] == prec]
            if radial == "learned" and learned:
                soaps[key] = SOAP(device=dev, dtype="float32", precision=prec, radial="learned",
                                  surrogate=learned[0].surrogate, spectrum=spectrum, **params)
            elif radial == "learned":
                t1 = time.perf_counter()
                ckpt = args.radial_checkpoint
                src = ckpt if ckpt and os.path.exists(ckpt) else None
                soaps[key] = SOAP(device=dev, dtype="float32", precision=prec, radial="learned", spectrum=spectrum,
                                  surrogate=src, **params)
                sur = soaps[key].surrogate
                if ckpt and src is None:
                    sur.save(ckpt)
                train["radial_checkpoint_loaded"] = src is not None
                train.update({"radial_training_s": round(time.perf_counter() - t1, 2),
                              "radial_gate": [int(u) for u in sur.use_pair],
                              "radial_heldout_value_err": float(sur.errors["value"].max()),
                              "radial_heldout_slope_err": float(sur.errors["slope"].max())})
            else:
                soaps[key] = SOAP(device=dev, dtype="float32", precision=prec, spectrum=spectrum, radial=radial,
                                  **params)
        return soaps[key]

    soap = soap_for(args.precision, args.radial, args.spectrum)
    n_feat = soap.get_number_of_features()
    w = torch.tensor(np.random.default_rng(12345).normal(size=n_feat), device=dev)

    def step(sp):
        return sp.descriptor_and_vjp(sys_dev, w)

    for _ in range(args.warmup):
        d, F = step(soap)
        del d, F
    torch.cuda.synchronize()
    names = ("coefficients", "vjp_edges", "vjp_adjoint", "nl_count", "nl_fill", "power_spectrum")
    _lib.profile_enable(True)
    for nm in names:
        _lib.profile_get(nm)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        d, F = step(soap)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    _lib.profile_enable(False)
    kern = {}
    for nm in names:
        ms, n = _lib.profile_get(nm)
        kern[nm] = ms / max(n, 1)
    out = {"metric": METRIC, "value": round(n_env * args.steps / dt, 1), "unit": "atom-envs/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt * 1e3 / args.steps, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": ("f32" if args.radial != "exact" and args.precision != "float64" else
                     {"float64": "f64", "mixed": "f64+f32", "float32": "f32"}[args.precision]),
           "data": "synthetic",
           "config": {"workload": desc_txt, "n_atoms": n_env, "n_features": n_feat, "precision": args.precision,
                      "radial": args.radial, "spectrum": args.spectrum, **train,
                      "system_generation_s": round(t_gen, 1), "parallelism": "single GPU"},
           "kernel_ms": {k: round(v, 4) for k, v in kern.items()}}
    if args.compare:
        ref_d = d.double()
        ref_F = F
        if args.precision != "float64" or args.radial != "exact" or args.spectrum != "exact":
            ref_d, ref_F = step(soap_for("float64"))
            ref_d = ref_d.double()
        errs = {}
        for form in [p for p in args.compare.split(",") if p]:
            parts = form.split(":")
            rad = next((q for q in parts[1:] if q in ("learned", "table")), "exact")
            spec = next((q for q in parts[1:] if q not in ("learned", "table")), "exact")
            dd, FF = step(soap_for(parts[0], rad, spec))
            dd = dd.double()
            sc = ref_d.abs().amax(0).clamp_min(1e-300)
            errs[form] = {"desc_max_col_rel": float(((dd - ref_d).abs().amax(0) / sc).max()),
                          "force_max_rel": float((FF - ref_F).abs().max() / ref_F.abs().max())}
        out["errors_vs_gpu_float64"] = errs
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
