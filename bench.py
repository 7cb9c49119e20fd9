"""Benchmark of the H-MEC 2D per-timestep hot path on one MI355X.

Workload: the 801x201-node precision configuration (972,024 unknowns,
2,560,000 markers, 16 per cell; fault viscosity contrast up to 1e6,
log-uniform permeability), synthetic seeded input generated on the host and
uploaded once.  A step is one full timestep: marker->node scatter, then the
Picard loop of VEP update, CSR assembly, sparse LU solve, RSF fault solve and
timestep control, then the final stresses and history.

value = timesteps/s of the mixed-precision path: an FP32 multifrontal factor
        reused across Picard iterations and timesteps as the preconditioner of
        FP64 GMRES on the original system (relative inf-norm residual
        <= 1e-12 every solve; refactored in FP32 when GMRES needs more than
        its budget).  The plain FP64-factor path (refactored every Picard
        iteration, PardisoLU semantics) is timed beside it.
roofline = the dominant kernel of the FP32 path, measured with HIP events
        on the context's stream over the timed region.
sustained = the FP32 path over 200 further steps (its factorization cadence
        included), with factorizations, p50 and p99 of that window.
error_vs_oracle = per-field relative error of both GPU paths against the FP64
        oracle (MKL PARDISO) after the same two timesteps.
cpu_baseline = the CPU oracle (plain-C restatement + MKL PARDISO) on the same
        workload on this host: setting A as shipped (MKL sequential, 1 core),
        setting B with MKL threaded on every core of the process; the second
        step's cost per Picard iteration scaled to the GPU window's Picard
        iterations per step.

--serve runs the serving workload instead and prints its own JSON line: a
Poisson-arrival stream of ensemble members with grids drawn uniformly from
{101x51, 201x51, 401x101, 801x201}, a ~ U[0.008, 0.012], b ~ U[0.012, 0.020],
L ~ U[0.01, 0.05] m on the fault and one log-uniform permeability in
[1e-16, 1e-14] m^2 per member, served by the ensemble server (admission by
device footprint, continuous batching, batched same-size factorizations);
value = member-timesteps/s with p50/p99 step latency.
"""
import argparse
import ctypes as C
import importlib.util
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PEAK_HBM = 8.0e12          # B/s
PEAK_FP64 = 78.6e12        # FLOP/s
PEAK_FP32 = 157.3e12       # FLOP/s
METRIC = "timesteps/s, FP32-LU+FP64 IR vs FP64-LU at rel.res ≤1e-10; p99 step latency"


def load_hmec():
    spec = importlib.util.spec_from_file_location("hmec", os.path.join(ROOT, "h-mecs-2d_amd", "hmec.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["hmec"] = mod
    spec.loader.exec_module(mod)
    return mod


KT_NAMES = ("scatter", "assemble", "factor", "solve", "schur", "sweep_form", "sweep_fwd", "sweep_bwd")
# the timers of the throughput windows: every event pair sits in the stream
# and costs device time, so they record only the sweep (the roofline's
# kernel); a separate window with every timer gives the breakdown
ROOF_TIMERS = "sweep_fwd,sweep_bwd"
ROOF_STRIDE = 4     # the throughput windows time every fourth sweep of each direction


def kernel_times(ctx):
    kt = {}
    for name in KT_NAMES:
        n, ms, work = ctx.kernel_time(name)
        if n > 0:
            kt[name] = {"n": n, "avg_ms": ms, "work": work}
    return kt


def run_path(hm, p, soa, prec, warmup, steps):
    ctx = hm.Context(p, soa, precision=prec)
    for _ in range(warmup):
        ctx.step()
    ctx.timing(True)
    ctx.timing_select(ROOF_TIMERS)
    ctx.timing_sample(ROOF_STRIDE)
    ctx.timing_reset()
    lat = []
    infos = []
    t0 = time.perf_counter()
    for _ in range(steps):
        a = time.perf_counter()
        infos.append(ctx.step())      # returns after the step's last stream sync
        lat.append(time.perf_counter() - a)
    total = time.perf_counter() - t0
    return ctx, total, np.array(lat), infos, kernel_times(ctx)


ERR_FIELDS = ("vxs", "vys", "pt", "pf", "vxD", "vyD", "SXX", "SYY", "SXY", "VSLIPB", "ETA", "OM",
              "SigmaY", "SII_fault")


def host_facts():
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)), "model": model}


def mkl_version():
    """Version string of the MKL the oracle loads (same search as the oracle)."""
    for name in ("libmkl_rt.so", "libmkl_rt.so.1", "/opt/conda/lib/libmkl_rt.so",
                 "/opt/conda/lib/libmkl_rt.so.1"):
        try:
            L = C.CDLL(name)
        except OSError:
            continue
        buf = C.create_string_buffer(256)
        L.MKL_Get_Version_String(buf, 256)
        return buf.value.decode(errors="replace").strip()
    return None


def cpu_sample(args):
    """Child process: the CPU oracle (plain-C restatement + MKL PARDISO) on
    the first two timesteps of the workload, with MKL's threading fixed by the
    parent's environment; prints per-phase timings as JSON and, with
    --cpu-fields, saves the fields after step 2 for the error report."""
    hm = load_hmec()
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    oracle.build()
    p = hm.params(args.cfg)
    soa = hm.generate_markers(p, cfg=args.cfg, seed=args.cfg)
    o = oracle.Oracle(p, soa, hm)
    steps = []
    for _ in range(2):
        t = time.perf_counter()
        info = o.step()
        steps.append({"s": time.perf_counter() - t, "iterations": info.iterations,
                      "scatter_ms": info.ms_scatter, "assemble_ms": info.ms_assemble,
                      "solve_ms": info.ms_factor, "update_ms": info.ms_update})
    if args.cpu_fields:
        np.savez(args.cpu_fields, **{f: o.get(f) for f in ERR_FIELDS})
    print(json.dumps({"steps": steps, "pardiso": oracle.have_pardiso(), "mkl": mkl_version()}))


def cpu_baseline(args, picard_per_step, fields_path):
    """The oracle timed on this host in two MKL settings (children of this
    process): A = as shipped, MKL sequential and the single-threaded oracle
    (1 core); B = MKL threaded on every core of this process's affinity.
    The second (steady-state) step gives the cost of a scatter and of one
    Picard iteration (assembly, PARDISO analyze + factorize + solve, update);
    the value is the timesteps/s at the GPU window's mean Picard iterations
    per step."""
    out = {}
    ncores = len(os.sched_getaffinity(0))
    for name, env_over, cores in (("A", {"MKL_THREADING_LAYER": "SEQUENTIAL", "MKL_NUM_THREADS": "1"}, 1),
                                  ("B", {"MKL_THREADING_LAYER": "INTEL",
                                         "MKL_NUM_THREADS": str(ncores)}, ncores)):
        env = dict(os.environ)
        env.update(env_over)
        cmd = [sys.executable, os.path.abspath(__file__), "--cpu-sample", "--cfg", str(args.cfg)]
        if name == "A":
            cmd += ["--cpu-fields", fields_path]
        r = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=900)
        if r.returncode != 0:
            out[name] = {"error": r.stderr[-400:]}
            continue
        d = json.loads(r.stdout.strip().splitlines()[-1])
        st = d["steps"][1]
        per_it = (st["assemble_ms"] + st["solve_ms"] + st["update_ms"]) / st["iterations"]
        step_ms = st["scatter_ms"] + per_it * picard_per_step
        out[name] = {"value": 1e3 / step_ms, "cores": cores, "per_picard_iteration_s": per_it * 1e-3,
                     "scatter_s": st["scatter_ms"] * 1e-3, "sample_steps": d["steps"],
                     "pardiso": d["pardiso"], "mkl": d["mkl"]}
    a = out.get("A", {})
    res = {"value": a.get("value"), "unit": "timesteps/s", "cores": 1, "kind": "port",
           "sample": ("the first two timesteps of the 801x201 workload on the oracle (plain-C "
                      "restatement, MKL PARDISO sequential: setting A, as shipped); the second step's "
                      "scatter and per-Picard-iteration cost scaled to the GPU window's %.2f Picard "
                      "iterations per step" % picard_per_step),
           "solver": ("MKL PARDISO (%s)" % a.get("mkl")) if a.get("pardiso") else "banded LU (no MKL)",
           "host": host_facts(), "setting_A": a, "setting_B": out.get("B")}
    return res


def serve(hm, args):
    rng = np.random.default_rng(args.serve_seed)
    grids = [(101, 51), (201, 51), (401, 101), (801, 201)]
    prec = hm.PREC_FP64 if args.serve_fp64 else hm.PREC_FP32
    srv = hm.Server(precision=prec, max_active=args.serve_slots, batch_window_ms=2.0)
    t = 0.0
    mix = {}
    for k in range(args.serve_members):
        t += rng.exponential(args.serve_interarrival_ms)
        nx, ny = grids[int(rng.integers(len(grids)))]
        a = rng.uniform(0.008, 0.012)
        b = rng.uniform(0.012, 0.020)
        L = rng.uniform(0.01, 0.05)
        p = hm.params(1, Nx=nx, Ny=ny, arsfm=(0.011, a), brsfm=(0.007, b), lrsfm=(0.02, L))
        srv.submit(p, args.serve_steps, cfg=5, seed=1000 + k, arrival_ms=t)
        mix["%dx%d" % (nx, ny)] = mix.get("%dx%d" % (nx, ny), 0) + 1
    srv.run()
    st = srv.stats()
    res = [srv.result(i) for i in range(args.serve_members)]
    out = {
        "metric": "member-timesteps/s, serving mix (continuous batching)",
        "value": st.member_steps_per_s,
        "unit": "member-timesteps/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "p50_ms": st.p50_ms,
        "p99_ms": st.p99_ms,
        "wall_s": st.wall_ms * 1e-3,
        "member_steps": int(st.member_steps),
        "members_done": st.members_done,
        "members_failed": st.members_failed,
        "max_concurrent": st.max_concurrent,
        "peak_admitted_GB": st.peak_bytes / 1e9,
        "factor_batches": st.factor_batches,
        "factor_batch_members": st.factor_batch_members,
        "factorizations": int(sum(r.factorizations for r in res)),
        "fallbacks": int(sum(r.fallbacks for r in res)),
        "max_rel_residual": float(max(r.max_rel_residual for r in res)),
        "dtype": "f32-factor+f64" if prec == hm.PREC_FP32 else "f64",
        "data": "synthetic (seeded host generator per member)",
        "config": {"workload": "serving mix", "members": args.serve_members,
                   "steps_per_member": args.serve_steps, "grids": mix,
                   "mean_interarrival_ms": args.serve_interarrival_ms,
                   "slots": args.serve_slots},
    }
    srv.close()
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cfg", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--serve", action="store_true", help="serving workload instead")
    ap.add_argument("--serve-members", type=int, default=64)
    ap.add_argument("--serve-steps", type=int, default=100)
    ap.add_argument("--serve-slots", type=int, default=16)
    ap.add_argument("--serve-interarrival-ms", type=float, default=50.0)
    ap.add_argument("--serve-seed", type=int, default=5)
    ap.add_argument("--serve-fp64", action="store_true")
    ap.add_argument("--sustained", type=int, default=200, help="steps of the sustained FP32 window")
    ap.add_argument("--cpu-sample", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--cpu-fields", default="", help=argparse.SUPPRESS)
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run a plain device copy of this many bytes (counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        n = args.copy_probe // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n * 4}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each ensemble member "
              "or grid runs on one device); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.cpu_sample:
        cpu_sample(args)
        return
    hm = load_hmec()
    if args.serve:
        serve(hm, args)
        return
    p = hm.params(args.cfg)
    soa = hm.generate_markers(p, cfg=args.cfg, seed=args.cfg)
    ctx32, t32, lat32, inf32, kt32 = run_path(hm, p, soa, hm.PREC_FP32, args.warmup, args.steps)
    res = {}
    if not args.no_fp64:
        ctx64, t64, lat64, inf64, kt64 = run_path(hm, p, soa, hm.PREC_FP64, args.warmup, args.steps)
        errs = {}
        for f in ("vxs", "vys", "pt", "pf", "vxD", "vyD", "SXX", "SYY", "SXY", "VSLIPB", "ETA", "OM"):
            a, b = ctx32.get(f), ctx64.get(f)
            nb = np.linalg.norm(b)
            errs[f] = float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a - b))
        res["fp64"] = {"timesteps_per_s": args.steps / t64,
                       "p50_ms": float(np.percentile(lat64, 50) * 1e3),
                       "p99_ms": float(np.percentile(lat64, 99) * 1e3),
                       "factorizations": int(sum(i.factorizations for i in inf64)),
                       "max_rel_residual": float(max(i.max_rel_residual for i in inf64))}
        res["error_vs_fp64"] = errs
        res["speedup_vs_fp64"] = (args.steps / t32) / (args.steps / t64)
        iters_same = all(a.iterations == b.iterations for a, b in zip(inf32, inf64))
        res["same_picard_iterations"] = bool(iters_same)
    # roofline of the dominant kernel class of the FP32 path
    cands = []
    fac = kt32.get("schur")
    if fac:
        ach = fac["work"] / (fac["avg_ms"] * 1e-3)
        cands.append((fac["n"] * fac["avg_ms"], {"kernel": "k_schur<float>", "bound": "mfma",
                                                   "achieved": ach / 1e12, "peak": PEAK_FP32 / 1e12,
                                                   "unit": "TFLOP/s", "frac": ach / PEAK_FP32,
                                                   "traffic": None}))
    # measured HBM bytes per launch of the sweep kernels (rocprofv3 FETCH_SIZE /
    # WRITE_SIZE passes, copy-calibrated), committed beside the kernel summary
    traffic = {}
    tpath = os.path.join(ROOT, "profiles", "sweep_801x201_traffic.json")
    if args.cfg == 3 and os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f)
        traffic = {"sweep_fwd": tj["dir0"]["traffic_bytes_per_launch"],
                   "sweep_bwd": tj["dir1"]["traffic_bytes_per_launch"]}
    for nm, kn in (("sweep_fwd", "k_sweep<float,0> (forward, per level)"),
                   ("sweep_bwd", "k_sweep<float,1> (backward, per level)")):
        k = kt32.get(nm)
        if k:
            ach = k["work"] / (k["avg_ms"] * 1e-3)
            cands.append((k["n"] * k["avg_ms"], {"kernel": kn, "bound": "hbm", "achieved": ach / 1e9,
                                                  "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                                                  "frac": ach / PEAK_HBM,
                                                  "algorithmic_bytes": k["work"],
                                                  "traffic": traffic.get(nm)}))
    sc = kt32.get("scatter")
    if sc:
        ach = sc["work"] / (sc["avg_ms"] * 1e-3)
        cands.append((sc["n"] * sc["avg_ms"], {"kernel": "k_scatter_grid (4 grids)", "bound": "hbm",
                                                "achieved": ach / 1e9, "peak": PEAK_HBM / 1e9,
                                                "unit": "GB/s", "frac": ach / PEAK_HBM,
                                                "traffic": None}))
    roof = max(cands, key=lambda c: c[0])[1] if cands else None
    value = args.steps / t32
    # the same FP32 context over a longer window: its refactorization cadence
    # (a far dt gets its own factor; a stalled GMRES refactors) is inside it
    sustained = None
    if args.sustained > 0:
        lat_s, inf_s = [], []
        t0 = time.perf_counter()
        for _ in range(args.sustained):
            a = time.perf_counter()
            inf_s.append(ctx32.step())
            lat_s.append(time.perf_counter() - a)
        ts = time.perf_counter() - t0
        sustained = {"steps": args.sustained, "timesteps_per_s": args.sustained / ts,
                     "p50_ms": float(np.percentile(lat_s, 50) * 1e3),
                     "p99_ms": float(np.percentile(lat_s, 99) * 1e3),
                     "factorizations": int(sum(i.factorizations for i in inf_s)),
                     "gmres_iterations": int(sum(i.ir_iterations for i in inf_s)),
                     "fallbacks": int(sum(i.fallbacks for i in inf_s)),
                     "picard_iterations": int(sum(i.iterations for i in inf_s)),
                     "max_rel_residual": float(max(i.max_rel_residual for i in inf_s))}
    # per-phase breakdown: 10 further steps with every timer recording
    ctx32.timing_select(None)
    ctx32.timing_sample(1)
    ctx32.timing_reset()
    for _ in range(10):
        ctx32.step()
    breakdown = kernel_times(ctx32)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "timesteps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t32 / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-factor+f64",
        "data": "synthetic (seeded host generator, uploaded once)",
        "config": {"workload": "801x201 nodes, 972024 unknowns, 2560000 markers (precision config)",
                   "grid": [p.Nx, p.Ny], "unknowns": (p.Nx + 1) * (p.Ny + 1) * 6,
                   "markers": int(soa.shape[1]), "parallelism": "1 GPU"},
        "p50_ms": float(np.percentile(lat32, 50) * 1e3),
        "p99_ms": float(np.percentile(lat32, 99) * 1e3),
        "fp32_ir": {"factorizations": int(sum(i.factorizations for i in inf32)),
                    "ir_iterations": int(sum(i.ir_iterations for i in inf32)),
                    "fallbacks": int(sum(i.fallbacks for i in inf32)),
                    "max_rel_residual": float(max(i.max_rel_residual for i in inf32)),
                    "picard_iterations": int(sum(i.iterations for i in inf32))},
        "kernels_fp32": {"window": "10 steps after the sustained window, every timer on",
                         **{k: {"n": v["n"], "avg_ms": v["avg_ms"]} for k, v in breakdown.items()}},
        "roofline": roof,
        "sustained": sustained,
    }
    out.update(res)
    if not args.no_cpu:
        import shutil
        import tempfile
        tmp = tempfile.mkdtemp()
        fields_path = os.path.join(tmp, "oracle_fields.npz")
        try:
            picard = sum(i.iterations for i in inf32) / len(inf32)
            out["cpu_baseline"] = cpu_baseline(args, picard, fields_path)
            if os.path.exists(fields_path):
                ref = np.load(fields_path)
                ev = {}
                for name, prec in (("fp64", hm.PREC_FP64), ("fp32", hm.PREC_FP32)):
                    c = hm.Context(p, soa, precision=prec)
                    for _ in range(2):
                        c.step()
                    ev[name] = {}
                    for f in ERR_FIELDS:
                        a, b = c.get(f), ref[f]
                        nb = np.linalg.norm(b)
                        ev[name][f] = float(np.linalg.norm(a - b) / nb) if nb > 0 else float(
                            np.linalg.norm(a - b))
                    c.close()
                out["error_vs_oracle"] = ev
        except Exception as e:     # reported, never fatal for the GPU number
            out["cpu_baseline"] = {"error": str(e)}
        finally:
            shutil.rmtree(tmp, ignore_errors=True)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
