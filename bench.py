#!/usr/bin/env python3
"""Throughput of the TRiSK/HTC shallow-water step on one MI355X.

Workload (the glevel-8 configuration): icosahedral eqs/nopt grid, glevel 8
(655,362 cells, 1,966,080 edges, 1,310,720 triangles), seeded spectral state
(h' 5 % of H0 = 2.94e4/g, 20 m/s rms wind), RK4 with dt = 112.5 s, FP32
stencil with FP64 geometry, FP64 compensated mass-flux divergence and FP64
reductions. A "step" is one RK4 step; --steps 768 is one simulated day.

Prints ONE JSON line: cell-steps/s, the HBM roofline of the step (CANON-1
algorithmic bytes per step / measured step time vs 8 TB/s), per-kernel
timings, accuracy against the FP64 path (bit-identical to the reference) and
the reference CPU path timed on this host's cores.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def canon_step_bytes(nv, ne, nt, fp64=False):
    """CANON-1 algorithmic bytes of one RK4 step (each array once per stage)."""
    P = 30 * nv - 120
    if fp64:
        return 4 * (40 * nv + 66 * ne + 55 * nt + 12 * P) + 48 * (nv + ne)
    return 4 * (24 * nv + 50 * ne + 35 * nt + 8 * P) + 24 * (nv + ne)


def impl_kernel_bytes(nv, ne, nt):
    """Bytes each FP32 fused kernel must move once per launch with its layout
    (every table and state array read or written exactly once)."""
    # triangle: 2 vertex ids, edge bits, 3 eta coefficients, 3 kite weights,
    #           f, q out; gathers u (edges) and h (cells) once
    # cell:     6 neighbour ids, 2 bit words, 6 kite areas, 1/(A a) (FP64),
    #           K normaliser, h/h0/hout/acc/phi, 6 dual lengths and 6 K
    #           weights (per slot); per edge u and the Coriolis partial out;
    #           q per triangle
    # edge:     other-cell id, 1/l, 2 partials, u0, acc (r+w), u out; phi
    return {
        "fast_tri": nt * (8 + 4 + 12 + 12 + 4 + 4) + ne * 4 + nv * 4,
        "fast_cell": nv * (24 + 4 + 4 + 24 + 8 + 4 + 4 + 4 + 8 + 4 + 4 + 24 + 24) + ne * (4 + 8) + nt * 4,
        "fast_edge": ne * (4 + 4 + 4 + 4 + 4 + 8 + 4) + nv * 4,
    }


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cpu():
    """This host's CPU model and the cores this process may run on."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        ncpu = os.cpu_count() or 1
    return model, ncpu


def cpu_baseline(mesh, h0, u0, dt, nsteps, nsteps1=2, reps=3):
    """The reference Fortran loop (built from the reference sources by
    oracle/Makefile.ref) on this host's cores, from the same mesh and state:
    the timing build (-O3 -fopenmp, portable -march) best of `reps` runs of
    nsteps steps on every core, and one run of nsteps1 steps on one core.
    Falls back to the parity build, then to the C oracle."""
    from oracle import oracle as orc
    from imodel_amd import imxbin

    model, ncpu = host_cpu()
    fast = os.path.join(ROOT, "oracle", "_ref", "fast", "ref_harness")
    harness = fast if os.access(fast, os.X_OK) else (orc.REF_HARNESS if orc.ref_harness_available() else None)
    if harness is None:
        o = orc.Oracle(mesh.arrays)
        t0 = time.perf_counter()
        o.step(h0, u0, dt, nsteps)
        wall = time.perf_counter() - t0
        return {"value": mesh.nv * nsteps / wall, "unit": "cell-steps/s", "cores": ncpu, "kind": "port",
                "sample": f"glevel {mesh.glevel}, {nsteps} RK4 steps, C oracle (oracle/swm_oracle.c, OpenMP); "
                          f"{wall:.2f} s"}
    build = ("amdflang -O3 -fopenmp -march=x86-64-v3" if harness == fast else
             "amdflang -O2 -fopenmp -ffp-contract=off")
    with tempfile.TemporaryDirectory() as td:
        mesh.write(os.path.join(td, "mesh.bin"))
        imxbin.write(os.path.join(td, "state.bin"), {"h": h0, "u": u0})

        def run(threads, steps):
            env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
            subprocess.run([harness, "run", "mesh.bin", "state.bin", str(steps), repr(dt), "0", "out.bin"],
                           cwd=td, env=env, check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                           timeout=900)
            r = imxbin.read(os.path.join(td, "out.bin"))
            return float(r["wall"][0]), int(r["wall"][1])

        walls = []
        for k in range(reps):
            w, nthr = run(ncpu, nsteps)
            walls.append(w)
            log(f"cpu baseline run {k + 1}/{reps}: {nsteps} steps on {nthr} threads, {w:.2f} s")
        # one core takes about nthr times longer per step: one step only when
        # that would pass a minute (glevel 10)
        if min(walls) / nsteps * nthr > 60.0:
            nsteps1 = 1
        log(f"cpu baseline one core: {nsteps1} steps ...")
        w1, _ = run(1, nsteps1)
        log(f"cpu baseline one core: {nsteps1} steps, {w1:.2f} s")
    best = min(walls)
    return {"value": mesh.nv * nsteps / best, "unit": "cell-steps/s", "cores": nthr, "kind": "reference",
            "sample": f"glevel {mesh.glevel}, {nsteps} RK4 steps of dt={dt} s from the same seeded state, FP64, "
                      f"reference ode_rk4 loop (src/swm.f90:238-370) built by oracle/Makefile.ref ({build}); "
                      f"best of {reps}: {best:.2f} s (runs {', '.join(f'{w:.2f}' for w in walls)})",
            "one_core": {"value": mesh.nv * nsteps1 / w1, "steps": nsteps1, "wall_s": w1},
            "host": {"cpu": model, "cores_available": ncpu}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=768)
    ap.add_argument("--warmup", type=int, default=16)
    ap.add_argument("--glevel", type=int, default=8)
    ap.add_argument("--dt", type=float, default=112.5)
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64", "fp64_ref"])
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-steps", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default launch shapes instead of autotuning them before the timed run")
    ap.add_argument("--profile-steps", type=int, default=8)
    ap.add_argument("--copy-probe-mib", type=int, default=1024,
                    help="size of the streaming-copy bandwidth probe (0: skip)")
    ap.add_argument("--workload", default="single", choices=["single", "ensemble", "learned", "mlp", "scheme"],
                    help="single: one instance at --glevel; ensemble: --instances instances, "
                         "levels uniform in 4..7, CFL-0.8 time steps; learned: glevel-6 model with "
                         "the learned correction vs plain glevel 6 and 7 against restricted glevel 8; "
                         "mlp: MLP inference and corrected-step timing at --glevel (untrained network)")
    ap.add_argument("--instances", type=int, default=256)
    ap.add_argument("--compare-coarser", action="store_true",
                    help="also run the plain model one glevel coarser (2 dt, steps/2) from its own "
                         "sampling of the same spectral state and compare on the coarser grid")
    ap.add_argument("--train-seeds", type=int, default=24)
    ap.add_argument("--train-samples", type=int, default=48)
    ap.add_argument("--fit-steps", type=int, default=6000)
    ap.add_argument("--eval-steps", type=int, default=0, help="learned: coarse steps evaluated (0: one day)")
    ap.add_argument("--limiter", default="fallback", choices=["fallback", "scale"],
                    help="learned: a correction above the bound keeps the plain step, or is scaled to it")
    ap.add_argument("--coarse-glevel", type=int, default=6,
                    help="learned: the corrected glevel gc (6: config 4; 8: config 5's leg), trained on "
                         "glevel gc+2 truth, compared with plain gc+1 at half the step; dt = 450 s at glevel 6, "
                         "halved per level")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "sharding one instance across GPUs is designed but not implemented", file=sys.stderr)
        sys.exit(2)

    from imodel_amd.swm import Mesh, Swm

    if args.workload == "ensemble":
        return bench_ensemble(args)
    if args.workload == "learned":
        return bench_learned(args)
    if args.workload == "mlp":
        return bench_mlp(args)
    if args.workload == "scheme":
        return bench_scheme(args)

    t = time.perf_counter()
    mesh = Mesh.icos(args.glevel)
    log(f"mesh g{args.glevel}: nv={mesh.nv} ne={mesh.ne} nt={mesh.nt} ({time.perf_counter() - t:.1f} s)")
    t = time.perf_counter()
    h0, u0 = mesh.spectral(args.seed)
    log(f"spectral state ({time.perf_counter() - t:.1f} s)")
    nv, ne, nt = mesh.nv, mesh.ne, mesh.nt

    t = time.perf_counter()
    s = Swm(mesh, precision=args.precision, dt=args.dt)
    log(f"handle created ({time.perf_counter() - t:.1f} s)")
    s.set_state(h0, u0)
    tuned = None
    if not args.no_tune and args.precision != "fp64_ref":
        # block size and XCD remap per fused kernel, outside the timed region
        # (bit-identical results for every choice)
        tuned = s.tune(12)
        log("autotuned", tuned)
        s.set_state(h0, u0)
    s.step(args.warmup)
    s.set_state(h0, u0)
    t0 = time.perf_counter()
    ev_ms = s.step_timed(args.steps)
    wall = time.perf_counter() - t0
    hK, uK = s.get_state()
    diag = s.diagnostics()
    kern = s.profile(args.profile_steps) if args.precision != "fp64_ref" else {}
    s.close()
    ms_step = 1e3 * wall / args.steps
    ev_step = ev_ms / args.steps
    value = nv * args.steps / wall
    log(f"{args.steps} steps: wall {wall * 1e3:.2f} ms, events {ev_ms:.2f} ms, {value:.3e} cell-steps/s")

    copy_gbs = None
    if args.copy_probe_mib > 0:
        import ctypes
        from imodel_amd import capi
        g = ctypes.c_double()
        capi.check(capi.lib().imx_copy_probe(args.copy_probe_mib << 20, 10, ctypes.byref(g)), "imx_copy_probe")
        copy_gbs = g.value
        log(f"copy probe: {copy_gbs:.0f} GB/s")

    fp64 = args.precision != "fp32"
    B = canon_step_bytes(nv, ne, nt, fp64)
    achieved = B / (ev_step * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_g{args.glevel}_{args.precision}.json")
    if os.path.exists(tpath):
        traffic = json.load(open(tpath)).get("hbm_bytes_per_step")
    hbm_gbs = traffic / (ev_step * 1e-3) / 1e9 if traffic else None
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                "hbm_GBs": hbm_gbs, "hbm_frac": (hbm_gbs / HBM_PEAK_GBS) if hbm_gbs else None,
                "kernel": "RK4 step (one hipGraph launch: 4 x fast_tri, fast_cell, fast_edge)",
                "algorithmic_bytes_per_launch": B, "launch_ms": ev_step,
                "copy_probe_GBs": copy_gbs,
                # the step's measured bytes (CANON-1 bytes without counter
                # data) over the rate a streaming copy reaches on this GPU
                "frac_of_copy": ((hbm_gbs or achieved) / copy_gbs) if copy_gbs else None,
                "frac_of_copy_basis": "measured bytes" if hbm_gbs else "CANON-1 bytes"}
    kernels = {}
    if kern:
        kb = impl_kernel_bytes(nv, ne, nt)
        for k, ms in kern.items():
            kernels[k] = {"ms": ms, "impl_bytes": kb.get(k), "GBs": kb.get(k, 0) / (ms * 1e-3) / 1e9}

    accuracy = None
    if not args.no_accuracy and args.precision == "fp32":
        with Swm(mesh, precision="fp64_ref", dt=args.dt) as r:
            r.set_state(h0, u0)
            t = time.perf_counter()
            r.step(args.steps)
            log(f"FP64 reference-order run ({time.perf_counter() - t:.1f} s)")
            hR, uR = r.get_state()
            dR = r.diagnostics()
        A = mesh.arrays["hx_areag"]
        L = mesh.arrays["ed_leng"] * mesh.arrays["edhx_leng"] * 0.5
        m0 = np.sum(A * h0)
        accuracy = {
            "vs": "FP64 GPU path bit-identical to the reference (same state, same steps)",
            "h_l2_rel": float(np.sqrt(np.sum(A * (hK - hR) ** 2) / np.sum(A * hR ** 2))),
            "h_linf_rel": float(np.max(np.abs(hK - hR)) / np.max(np.abs(hR))),
            "hprime_l2_rel": float(np.sqrt(np.sum(A * (hK - hR) ** 2) / np.sum(A * (hR - hR.mean()) ** 2))),
            "u_l2_rel": float(np.sqrt(np.sum(L * (uK - uR) ** 2) / np.sum(L * uR ** 2))),
            "u_linf_rel": float(np.max(np.abs(uK - uR)) / np.max(np.abs(uR))),
            "mass_drift": float((np.sum(A * hK) - m0) / m0),
            "mass_drift_fp64": float((np.sum(A * hR) - m0) / m0),
            "simulated_s": args.steps * args.dt,
        }
        log("accuracy", accuracy)

    coarser = None
    if args.compare_coarser:
        coarser = compare_coarser(args, mesh, hK, uK)
        coarser["fine_gpu_ms_same_interval"] = ev_ms
        log("coarser grid", coarser)

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(mesh, h0, u0, args.dt, args.cpu_steps)
            log("cpu baseline", cpu)
        except Exception as e:  # reported, never silently replaced
            cpu = {"value": None, "error": str(e)[:300]}

    out = {
        "metric": "cell-steps/s at glevel 8 FP32 vs FP64; time-to-accuracy learned vs plain",
        "value": value, "unit": "cell-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32" if args.precision == "fp32" else "f64",
        "data": "synthetic (seeded spectral state, spherical harmonics l=1..16)",
        "config": {"workload": f"TRiSK/HTC SWM RK4 step, icos eqs nopt glevel {args.glevel}",
                   "glevel": args.glevel, "cells": nv, "edges": ne, "triangles": nt, "dt": args.dt,
                   "precision": args.precision, "seed": args.seed, "autotuned": tuned},
        "roofline": roofline, "cpu_baseline": cpu, "kernels": kernels, "accuracy": accuracy,
        "coarser_grid": coarser,
        "diagnostics_end": {k: float(v) for k, v in diag.items()},
    }
    print(json.dumps(out), flush=True)


def compare_coarser(args, mesh, hK, uK):
    """Plain glevel-1 run over the same simulated time, compared with this
    run's final state restricted to the coarser grid."""
    from imodel_amd.learn import restriction
    from imodel_amd.swm import Mesh, Swm

    t = time.perf_counter()
    mc = Mesh.icos(args.glevel - 1)
    hc0, uc0 = mc.spectral(args.seed)
    Rh, Ru = restriction(mesh, mc)
    log(f"coarser mesh and restriction ({time.perf_counter() - t:.1f} s)")
    nsteps = args.steps // 2
    with Swm(mc, precision=args.precision, dt=2 * args.dt) as c:
        c.set_state(hc0, uc0)
        c.step(1)
        c.set_state(hc0, uc0)
        ms = c.step_timed(nsteps)
        hc, uc = c.get_state()
    ht, ut = Rh @ hK, Ru @ uK
    A = mc.arrays["hx_areag"]
    ha = ht - np.sum(A * ht) / np.sum(A)
    return {"glevel": args.glevel - 1, "dt": 2 * args.dt, "steps": nsteps, "gpu_ms": ms,
            "fine_gpu_ms_same_interval": None,
            "hprime_l2_rel_vs_fine": float(np.sqrt(np.sum(A * (hc - ht) ** 2) / np.sum(A * ha ** 2))),
            "u_l2_rel_vs_fine": float(np.linalg.norm(uc - ut) / np.linalg.norm(ut))}


# option sets of the scheme workload (par/swm.par lines), FP64_REF path
SCHEME_SETS = {
    "trsk10": {},
    "apvm": {"pv": "apvm", "pvspar": 0.5},
    "k2const": {"k2": 1e5, "diffus": "const"},
    "k4diam": {"k4": 1e14, "hyperdiffus": "diam"},
    "pxt16": {"wrap": "pxt16"},
    "gass18": {"wrap": "gass18"},
    "pxmix": {"wrap": "pxt16", "pv": "apvm", "pvspar": 0.5, "k4": 1e14, "hyperdiffus": "const"},
}


def bench_scheme(args):
    """The FP64 reference-order step (bit-identical to the reference) with
    each option set of SCHEME_SETS at --glevel: ms per step and cell-steps/s,
    against the plain trsk10 step of the same path. value = the pxmix set
    (pxt16 + APVM + hyperdiffusion). The roofline uses the plain step's
    CANON-1 bytes (FP64), so the options' extra passes show as a lower
    fraction."""
    from imodel_amd.swm import Mesh, Swm

    mesh = Mesh.icos(args.glevel)
    h0, u0 = mesh.spectral(args.seed)
    nv, ne, nt = mesh.nv, mesh.ne, mesh.nt
    steps = min(args.steps, 96)
    res = {}
    for name, sc in SCHEME_SETS.items():
        with Swm(mesh, precision="fp64_ref", dt=args.dt) as s:
            s.set_scheme(**sc)
            s.set_state(h0, u0)
            s.step(max(args.warmup, 1))
            s.set_state(h0, u0)
            ms = s.step_timed(steps)
            d = s.diagnostics()
        res[name] = {"ms_per_step": ms / steps, "cell_steps_per_s": nv * steps / (ms * 1e-3),
                     "options": sc, "mass_end": float(d["mass"])}
        log(f"{name}: {ms / steps:.3f} ms/step")
    for name in res:
        res[name]["vs_trsk10"] = res["trsk10"]["ms_per_step"] / res[name]["ms_per_step"]
    # the options on the fused kernels, FP32 and FP64, against their plain
    # step
    fused = {}
    fused_sets = (("trsk10", {}), ("k2const", SCHEME_SETS["k2const"]),
                  ("k2diam", {"k2": 2e5, "diffus": "diam"}), ("apvm", SCHEME_SETS["apvm"]),
                  ("apvm_k2", dict(SCHEME_SETS["apvm"], **SCHEME_SETS["k2const"])),
                  ("hc", {"stag": "HC"}), ("k4diam", SCHEME_SETS["k4diam"]),
                  ("pxt16", SCHEME_SETS["pxt16"]), ("pxmix", SCHEME_SETS["pxmix"]),
                  ("gass18", SCHEME_SETS["gass18"]), ("hcgass18", {"stag": "HC", "wrap": "gass18"}),
                  ("hcpxt16", {"stag": "HC", "wrap": "pxt16"}),
                  # individual methods of wrapper none
                  ("nopv", {"nopv": True}), ("melv", {"wrap": "none", "rec": "melv"}),
                  ("gassk", {"wrap": "none", "rec": "gass", "gasscoef": 0.5}),
                  ("dtred", {"wrap": "none", "cor": "dtred"}),
                  ("gasscor", {"wrap": "none", "cor": "gass"}),
                  ("gassperhx", {"wrap": "none", "sint": "gass", "cor": "gass", "rec": "perhx"}))
    for prec in ("fp32", "fp64"):
        for name, sc in fused_sets:
            with Swm(mesh, precision=prec, dt=args.dt) as s:
                s.set_scheme(**sc)
                s.set_state(h0, u0)
                s.step(max(args.warmup, 1))
                s.set_state(h0, u0)
                ms = s.step_timed(steps)
            key = f"{prec}_{name}"
            fused[key] = {"ms_per_step": ms / steps, "cell_steps_per_s": nv * steps / (ms * 1e-3),
                          "options": sc}
            log(f"fused {key}: {ms / steps:.4f} ms/step")
        for name, _ in fused_sets[1:]:
            fused[f"{prec}_{name}"]["vs_plain"] = (fused[f"{prec}_trsk10"]["ms_per_step"]
                                                   / fused[f"{prec}_{name}"]["ms_per_step"])
    top = res["pxmix"]
    B = canon_step_bytes(nv, ne, nt, True)
    achieved = B / (top["ms_per_step"] * 1e-3) / 1e9
    out = {
        "metric": "FP64 reference-order step with scheme options (cell-steps/s)",
        "value": top["cell_steps_per_s"], "unit": "cell-steps/s", "n_gpus": 1, "steps": steps,
        "warmup": max(args.warmup, 1), "ms_per_step": top["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded spectral state)",
        "config": {"workload": f"pxt16 + APVM + hyperdiffusion, FP64 reference order, icos glevel {args.glevel}",
                   "glevel": args.glevel, "cells": nv, "dt": args.dt},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "RK4 step, reference-order kernels (plain step's CANON-1 bytes)",
                     "algorithmic_bytes_per_launch": B, "launch_ms": top["ms_per_step"]},
        "cpu_baseline": None,
        "schemes": res,
        "fused_options": fused,
    }
    print(json.dumps(out), flush=True)


def bench_ensemble(args):
    """Batched configuration: independent instances with glevels drawn
    uniformly from 4..7, each with its own seeded spectral state and a time
    step at Courant number 0.8; one launch per fused kernel and RK stage
    advances all of them; block shapes chosen by the autotuner."""
    from concurrent.futures import ThreadPoolExecutor

    from imodel_amd.swm import Ensemble, Mesh, cfl_dt

    rng = np.random.default_rng(args.seed)
    levels = rng.integers(4, 8, size=args.instances)
    t = time.perf_counter()
    meshes = {g: Mesh.icos(int(g)) for g in sorted(set(levels.tolist()))}
    ms = [meshes[g] for g in levels]
    with ThreadPoolExecutor(max_workers=16) as ex:
        states = list(ex.map(lambda i: ms[i].spectral(i + 1), range(args.instances)))
    dts = [cfl_dt(m, st[1]) for m, st in zip(ms, states)]
    cells = int(sum(m.nv for m in ms))
    log(f"ensemble: {args.instances} instances, {cells} cells, set-up {time.perf_counter() - t:.1f} s")
    t = time.perf_counter()
    E = Ensemble(ms, dts, precision=args.precision)
    for i, (h, u) in enumerate(states):
        E.set_state(i, h, u)
    log(f"ensemble create {time.perf_counter() - t:.1f} s, device bytes {E.info()[3] / 1e9:.2f} GB")
    cfg = E.tune(2)
    log("autotuned", cfg)
    E.step(args.warmup)
    for i, (h, u) in enumerate(states):
        E.set_state(i, h, u)
    t0 = time.perf_counter()
    ev_ms = E.step_timed(args.steps)
    wall = time.perf_counter() - t0
    d = [E.diagnostics(i) for i in range(args.instances)]
    bad = sum(x["badflag"] for x in d)
    E.close()
    value = cells * args.steps / wall
    nv, ne, nt = cells, 3 * cells - 6 * args.instances, 2 * cells - 4 * args.instances
    B = canon_step_bytes(nv, ne, nt, args.precision != "fp32") - 4 * 8 * (30 * nv - 120) + 4 * 8 * (30 * nv - 120 * args.instances)
    achieved = B / (ev_ms / args.steps * 1e-3) / 1e9
    out = {
        "metric": "aggregate cell-steps/s, batched ensemble", "value": value, "unit": "cell-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else "f64",
        "data": "synthetic (seeded spectral states)",
        "config": {"workload": f"{args.instances} independent SWM instances, glevel 4-7, CFL 0.8",
                   "cells": cells, "levels": np.bincount(levels, minlength=8)[4:].tolist(),
                   "precision": args.precision, "autotuned": cfg},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "ensemble RK4 step (one hipGraph launch)", "algorithmic_bytes_per_launch": B,
                     "launch_ms": ev_ms / args.steps},
        "instances_with_nan": int(bad),
    }
    print(json.dumps(out), flush=True)


MLP_FLOP_PER_EDGE = 2.0 * (48 * 128 + 2 * 128 * 128 + 128 * 2)


def bench_mlp(args):
    """Per-launch cost of the learned correction at --glevel: MLP inference
    over the handle's edge slots (HIP events over --steps launches), and the
    corrected step against the plain step (wall time per step). The network
    is the seeded untrained one: timing does not depend on the weights."""
    from imodel_amd.learn import Learn
    from imodel_amd.mlp import Mlp
    from imodel_amd.swm import H0_GEO, Mesh, Swm

    mesh = Mesh.icos(args.glevel)
    dt = args.dt
    s = Swm(mesh, precision="fp32", dt=dt, h_ref=H0_GEO)
    m = Mlp(seed=1)
    L = Learn(s, m, mesh)
    h0, u0 = mesh.spectral(args.seed)
    s.set_state(h0, u0)
    raw = L.sample(raw=True)
    L.set_scaling(raw.mean(0), raw.std(0) + 1e-3, 1e4, 1e3)  # small corrections: the run stays smooth
    ne_slots = 3 * mesh.nv
    fwd_ms = m.forward_timed(ne_slots, reps=args.steps)
    s.set_state(h0, u0)
    L.step(args.warmup, 1e30)
    t = time.perf_counter()
    nfb = L.step(args.steps, 1e30)
    learned_us = 1e6 * (time.perf_counter() - t) / args.steps
    s.set_state(h0, u0)
    s.step(args.warmup)
    t = time.perf_counter()
    s.step(args.steps)
    plain_us = 1e6 * (time.perf_counter() - t) / args.steps
    tflops = MLP_FLOP_PER_EDGE * ne_slots / (fwd_ms * 1e-3) / 1e12
    out = {
        "metric": "MLP inference TFLOP/s and corrected-step cost", "value": tflops, "unit": "TFLOP/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": learned_us / 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16 MFMA, f32 state",
        "data": "synthetic (seeded spectral state, seeded untrained network)",
        "config": {"workload": f"learned-correction step, icos glevel {args.glevel}", "glevel": args.glevel,
                   "edge_slots": ne_slots, "dt": dt},
        "roofline": {"bound": "mfma", "kernel": "mlp_fwd_lds (48-128-128-128-2, bf16 MFMA)",
                     "achieved": tflops, "peak": 2500.0, "unit": "TFLOP/s", "frac": tflops / 2500.0,
                     "traffic": None, "launch_ms": fwd_ms, "flop_per_edge": MLP_FLOP_PER_EDGE},
        "corrected_step_us": learned_us, "plain_step_us": plain_us, "limited": nfb,
    }
    print(json.dumps(out), flush=True)


def bench_learned(args):
    """Learned correction at glevel gc = --coarse-glevel (dt 450 s at glevel
    6, halved per finer level) against the plain glevel-gc and glevel-gc+1
    (half the step) models, all started from the FP64 glevel-gc+2 run
    restricted to their grids and compared on glevel gc after --eval-steps
    coarse steps (default one day). Training: --train-seeds runs x
    --train-samples one-step samples produced by this solver, stored on the
    device, --fit-steps shuffled AdamW steps. Reports errors, wall times per
    simulated interval, limited steps and the MLP kernel's MFMA throughput."""
    from imodel_amd.learn import Trainer, evaluate, restriction
    from imodel_amd.swm import H0_GEO, Mesh, Swm

    gc = args.coarse_glevel
    dt_c = 450.0 / 2 ** (gc - 6)
    nsteps = args.eval_steps or int(round(86400.0 / dt_c))
    t0 = time.perf_counter()
    tr = Trainer(gc=gc, gf=gc + 2, dt_c=dt_c, log=log)
    tr.learn.set_limiter(args.limiter)
    tr.calibrate([1000, 1001], nsamples=8, spinup=4)
    tc = time.perf_counter()
    tr.collect(list(range(1, args.train_seeds + 1)), args.train_samples, spinup=4)
    t_collect = time.perf_counter() - tc
    tc = time.perf_counter()
    fit = tr.fit(args.fit_steps, 1e-3, 5e-5, 1e-4)
    t_fit = time.perf_counter() - tc
    mm = Mesh.icos(gc + 1)
    mid = (Swm(mm, precision="fp32", dt=dt_c / 2, h_ref=H0_GEO),) + restriction(tr.fine_mesh, mm) + \
        restriction(mm, tr.coarse_mesh)
    res = [evaluate(tr, sd, nsteps, spinup=4, mid=mid) for sd in (9999, 8888)]
    for r in res:
        log(json.dumps(r))
    ne6 = 3 * tr.coarse_mesh.nv
    fwd_ms = tr.mlp.forward_timed(ne6, reps=50)
    tflops = MLP_FLOP_PER_EDGE * ne6 / (fwd_ms * 1e-3) / 1e12

    def mean(k, f):
        return float(np.mean([r[k][f] for r in res]))

    speed = mean("mid", "wall_s") / mean("learned", "wall_s")
    out = {
        "metric": f"learned glevel-{gc} vs plain glevel-{gc + 1}: error ratio and wall-time speedup "
                  f"({nsteps * dt_c / 86400:g} day)",
        "value": speed, "unit": f"x (plain glevel-{gc + 1} wall time / learned glevel-{gc} wall time)",
        "n_gpus": 1, "steps": nsteps, "warmup": 1, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 state, bf16 MFMA network",
        "data": f"synthetic (seeded spectral states; truth = FP64 glevel-{gc + 2} run of this solver)",
        "config": {"workload": f"learned correction, glevel {gc} dt {dt_c:g} s vs glevel {gc + 1} dt "
                               f"{dt_c / 2:g} s vs glevel {gc + 2} truth",
                   "train_samples": args.train_seeds * args.train_samples, "fit_steps": args.fit_steps,
                   "limiter": args.limiter,
                   "collect_s": t_collect, "fit_s": t_fit, "setup_s": time.perf_counter() - t0},
        "errors": {k: {"h": mean(k, "h"), "u": mean(k, "u"), "wall_s": mean(k, "wall_s")}
                   for k in ("plain", "learned", "mid")},
        "learned_vs_mid_error": {"h": mean("learned", "h") / mean("mid", "h"),
                                 "u": mean("learned", "u") / mean("mid", "u")},
        "limited": [r["learned"]["limited"] for r in res],
        "mass_drift_learned": [r["learned"]["mass_drift"] for r in res],
        "fit_loss_last": fit[-1].tolist(),
        "roofline": {"bound": "mfma", "kernel": "mlp_fwd (48-128-128-128-2, bf16 MFMA)",
                     "achieved": tflops, "peak": 2500.0, "unit": "TFLOP/s", "frac": tflops / 2500.0,
                     "traffic": None, "launch_ms": fwd_ms, "edges": ne6},
        "runs": res,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
