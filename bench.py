#!/usr/bin/env python3
"""Benchmark: zone-updates/s of the explicit FV MHD update on one MI355X.

Default workload -- the combined configuration the metric is quoted on:
3D MHD + Helmholtz EoS stratified slab, 512^3 zones (log rho 6 -> 2, log T
9 -> 7, abar 4 -> 16, seeded 1e-4 noise, seeded random-phase
divergence-free field at plasma beta 1e5), USE_FASTEOS + USE_WB +
EVOLVE_ETOT + ADVECT_YE_IABAR, LHLLD + CT + LIM2ND, RK2, CFL 0.8, the
generated 541 x 271 Helmholtz table.  It runs

  plain FP64     the reference algorithm (T(rho, e) Newton from the carried T);
  mixed + NN     FP32 reconstruction / Riemann on FP64 deviations, FP64 update,
                 CT and EoS residual, T(rho, e) Newton started from the
                 surrogate's corrected warm start (network trained on the GPU
                 on EoS-generated pairs over the slab's states);
  mixed + pred   the same with the energy-predictor start;
  mixed + warm   the same from the carried T;

each from the same initial data: ``--warmup`` untimed steps, ``--steps``
timed steps (device-synchronised loop, state resident in HBM), then on to
``--l1-steps`` in total, where every conserved field and each face-B
component of the mixed runs is compared with the plain FP64 state (the FP64
GPU path equals the CPU oracle bit for bit on every parity case).  One
"step" is one full time step (both RK stages, CT, CFL update) over the whole
grid; the metric follows the reference's "updated cells/s" = steps * zones /
loop wall time (source/source.F90:6834-6835).  ``value`` is the mixed + NN
path; the plain path's figures sit beside it.

Other workloads (``--workload``): ot3d (3D Orszag-Tang MHD 256^3), helm_slab
(Helmholtz slab 256^3, optionally MHD), lowmach (heated box 384^3),
hotbubble (2D 256^2), each with ``--precision`` fp64 / mixed.

Prints ONE JSON line on stdout (progress goes to stderr).  Includes
  roofline      -- the plain FP64 path's dominant kernel: its share of the
                   algorithmic bytes per zone-update B_zu = 8 (5 n_q + 4 n_e)
                   (B_zu / RK stages per zone and launch) over its average
                   launch time (HIP events on the library's stream), against
                   the 8 TB/s HBM peak; per-kernel byte models under
                   ``kernel_models``;
  step_roofline -- whole step: B_zu * zone-updates/s against the same peak;
  cpu_baseline  -- the CPU FP64 restatement (oracle/) on the host cores over
                   a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # B/s, MI355X HBM3E


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def kernel_bytes(cfg, name: str, swept: bool = False) -> float:
    """Algorithmic HBM bytes of one launch: the compulsory reads of the
    kernel's inputs (each value once) and writes of its outputs, FP64;
    swept: the update that finishes a fused-sweep stage."""
    nz, ny, nx = cfg.shape
    N = nx * ny * nz
    nv, sd = cfg.nvars, cfg.sdims
    mhd = cfg.mhd
    nf = [(nx + 1) * ny * nz, nx * (ny + 1) * nz, nx * ny * (nz + 1)]
    cell_in = 8 * (nv + (sd if mhd else 0))           # prim (+ b_cc) per cell
    helm = cfg.eos == 1
    if helm:  # + gamma_e, gamma_c per cell (FASTEOS) or T (face Newton guess)
        cell_in += 16 if cfg.fasteos else 8
    if name.startswith("flux_x"):
        d = int(name[-1]) - 1
        per_face_out = 8 * nv + (8 * 2 + 1 if mhd else 0)  # fluxes (+2 EMFs + sign)
        return N * cell_in + nf[d] * ((8 if mhd else 0) + per_face_out)
    if name == "corner":
        # 2 EMFs + sign per face and the 3 cell EMFs in (FP32 in the mixed
        # path), 3 edge EMFs out
        w = 4 if cfg.precision == 1 else 8
        return sum(nf[:sd]) * (2 * w + 1) + N * w * 3 + 3 * N * w
    if name == "faceb":    # b, b0, 4 edge EMFs in; b (+b0 at stage 1) out
        w = 4 if cfg.precision == 1 else 8
        return sum(nf[:sd]) * 8 * 3 + 3 * N * w
    if name == "sweep":
        # the fused sweep, averaged over the two RK2 stages: the cell inputs
        # and the normal face B once; stage 1 the prim, eint, b_cc of the
        # state0 rebuild, stage 2 state0 and the stage state; the new state
        # (and state0 at stage 1), 2 EMFs + sign per face (MHD) out
        st = nv + 1 + (sd if mhd else 0)
        rd = N * cell_in + (sum(nf[:sd]) * 8 if mhd else 0) + 0.5 * N * 8 * (st + 2 * nv)
        wr = (sum(nf[:sd]) * 17 if mhd else 0) + 1.5 * N * 8 * nv
        return rd + wr
    if name == "update" and swept:
        # cons -> prim after the fused sweep: U and the face B (Helmholtz: T)
        # in; prim, eint, b_cc, the cell EMFs, gammas (T) out
        gam = 2 if (helm and cfg.fasteos) else 0
        ec = sd if mhd else 0
        rd = nv + (3 if mhd else 0) + (1 if helm else 0)
        wr = nv + 1 + (sd if mhd else 0) + ec + gam + (1 if helm else 0)
        return 8.0 * N * (rd + wr)
    if name == "update":
        # compulsory traffic per launch, averaged over the two RK2 stages:
        # face fluxes (FP32 in the mixed path) and face B in; stage 1 reads
        # prim + eint + b_cc to rebuild state0 and writes it to U0 (the mixed
        # path reads U and keeps it as state0), stage 2 reads state0 and the
        # stage state; both write the new state and the cell EMFs; the FP64
        # prim / eint / b_cc (+ gammas) are written at every stage (mixed:
        # the gammas and the pressure at stage 2, the FP32 sweep inputs at
        # both); Helmholtz
        # reads and writes T at both
        mix = cfg.precision == 1
        st = nv + 1 + (sd if mhd else 0)                  # prim, eint, b_cc
        gam = 2 if (helm and cfg.fasteos) else 0
        ec = sd if mhd else 0
        rd = sum(nf[:sd]) * (4 if mix else 8) * nv + (sum(nf[:sd]) * 8 if mhd else 0)
        s1_rd = nv if mix else st
        s2_rd = 2 * nv
        s1_wr = (nv if mix else 2 * nv) + ec + (0 if mix else st + gam)
        # the mixed path leaves the FP64 prim / eint / b_cc to an on-demand
        # rebuild from U (Helmholtz: it still stores the EoS pressure)
        s2_wr = nv + ec + gam + ((1 if helm else 0) if mix else st)
        words = 0.5 * (s1_rd + s2_rd + s1_wr + s2_wr) + (2 if helm else 0)
        pf = 4 * (nv + 5) if mix else 0
        return rd + N * (8 * words + pf)
    if name == "misc" and cfg.shock_flattening:
        # shock-flattening flags: pressure in, detector bytes out and in
        # (cells -2..n+1), flag bytes out (cells -1..n)
        e = [nx + 4, ny + 4, nz + 4 if sd == 3 else 1]
        f = [nx + 2, ny + 2, nz + 2 if sd == 3 else 1]
        return N * 8 + 2 * e[0] * e[1] * e[2] + f[0] * f[1] * f[2]
    return 0.0


def diffusion_bytes(cfg, stages: int) -> float:
    """Algorithmic bytes of one thermal-diffusion step (RKL2, s stages):
    stage 1 reads T, rho, kappa, eint and writes e_0, Me_0, e_jm2, eint;
    every further stage converts eint to T (eint, rho in, T out; Helmholtz
    also Ye, 1/abar and the carried T in) and reads T, rho, kappa, eint,
    e_jm2, e_0, Me_0, writing eint and e_jm2; the end converts once more and
    writes T and p (Helmholtz also eint and the FASTEOS gammas)."""
    nz, ny, nx = cfg.shape
    N = nx * ny * nz
    helm = cfg.eos == 1
    to_t = 24 + (24 if helm else 0)
    stage = 24 + 32 + 16
    first = 24 + 8 + 32 + (0 if helm else 24)   # + T from p for the ideal gas
    final = to_t + 8 + (24 if helm and cfg.fasteos else (8 if helm else 0))
    return N * (first + (stages - 1) * (to_t + stage) + final)


def b_zu(cfg) -> int:
    """Algorithmic bytes per zone-update of the whole step, FP64, RK2:
    8 * (5 n_q + 4 n_e) with n_q evolved fields (conserved + face B) and n_e
    carried EoS fields (T, gamma_e, gamma_c for Helmholtz FASTEOS)."""
    n_q = cfg.nvars + (cfg.sdims if cfg.mhd else 0)
    n_e = (3 if cfg.fasteos else 1) if cfg.eos == 1 else 0
    stages = cfg.rk_stages
    if stages == 2:
        return 8 * (5 * n_q + 4 * n_e)
    return 8 * (8 * n_q + 6 * n_e)


def host_description(threads: int) -> dict:
    """The host the CPU baseline ran on: nproc, the lscpu model and socket
    count, OMP_NUM_THREADS and the thread binding."""
    import subprocess
    info = {"nproc": os.cpu_count(), "threads_used": threads,
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS"),
            "OMP_PROC_BIND": os.environ.get("OMP_PROC_BIND", "unset (no pinning)")}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "Socket(s)", "Thread(s) per core",
                             "Core(s) per socket"):
                info[k.strip()] = v.strip()
    except Exception as e:  # lscpu missing: report what is known
        info["lscpu"] = f"unavailable ({e})"
    try:
        info["affinity_cpus"] = len(os.sched_getaffinity(0))
    except Exception:
        pass
    return info


def cpu_baseline(pb, steps: int, threads: int):
    """The oracle (FP64 restatement, -O3 -march=x86-64-v3 -ffp-contract=off,
    OpenMP) timed over `steps` full steps of the same problem, after its own
    step-0 initialisation."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    r = oracle.OracleRun(pb, threads)
    r.init()
    t0 = time.perf_counter()
    r.run(steps)
    dt = time.perf_counter() - t0
    r.close()
    zones = pb.cfg.nx1 * pb.cfg.nx2 * (pb.cfg.nx3 if pb.cfg.sdims == 3 else 1)
    return zones * steps / dt, dt


def build_problem(args):
    from phlegethon_amd import problems
    n = args.n
    if args.workload == "ot3d":
        log(f"building 3D Orszag-Tang {n}^3 initial data")
        pb = problems.orszag_tang_3d(n=n)
        desc = f"3D Orszag-Tang ideal MHD + CT, {n}^3, LHLLD + LIM2ND, RK2, CFL 0.8, periodic"
        data = "synthetic (seeded 3D Orszag-Tang initial data, counter-based PRNG)"
    elif args.workload == "c5":
        from phlegethon_amd import helmholtz
        from phlegethon_amd.runtime import HelmEoS
        table = helmholtz.default_table()
        log(f"building the MHD + Helmholtz slab {n}^3 (table {table.nrho}x{table.nT}, beta 1e5)")
        eos = HelmEoS(table)
        pb = problems.helm_slab(eos, table, n=n, mhd_beta=1e5)
        eos.close()
        desc = (f"3D MHD + Helmholtz stratified slab, {n}^3, USE_FASTEOS + USE_WB + EVOLVE_ETOT"
                f" + ADVECT_YE_IABAR, LHLLD + CT + LIM2ND, beta 1e5, RK2, CFL 0.8")
        data = ("synthetic (plane-parallel rho/T/abar profiles, seeded 1e-4 density noise, "
                f"seeded random-phase divergence-free field; generated {table.nrho}x{table.nT} "
                "Helmholtz table)")
    elif args.workload == "helm_slab":
        from phlegethon_amd import helmholtz
        from phlegethon_amd.runtime import HelmEoS
        table = helmholtz.default_table()
        mhd = args.mhd_beta is not None
        log(f"building the Helmholtz slab {n}^3 (table {table.nrho}x{table.nT}, "
            f"MHD beta {args.mhd_beta})")
        eos = HelmEoS(table)
        pb = problems.helm_slab(eos, table, n=n, mhd_beta=args.mhd_beta,
                                fasteos=not args.no_fasteos)
        eos.close()
        eosdesc = ("per-face T(rho,P) Newton (no USE_FASTEOS)" if args.no_fasteos
                   else "USE_FASTEOS")
        desc = (f"3D stratified Helmholtz slab, {n}^3, {eosdesc} + USE_WB + EVOLVE_ETOT + "
                f"ADVECT_YE_IABAR, {'LHLLD + CT, beta ' + format(args.mhd_beta, 'g') if mhd else 'LHLLC'}"
                f" + LIM2ND, RK2, CFL 0.8")
        data = ("synthetic (plane-parallel rho/T/abar profiles, seeded 1e-4 density noise"
                f"{', seeded random-phase divergence-free field' if mhd else ''}; generated "
                f"{table.nrho}x{table.nT} Helmholtz table)")
    elif args.workload == "hotbubble":
        log(f"building the 2D hot bubble {n}^2")
        pb = problems.hot_bubble(n=n)
        desc = (f"2D hot bubble, {n}^2, ideal gas gamma 5/3 + gravity, USE_WB, LHLLC + LIM2ND, "
                f"RK2, CFL 0.8")
        data = "synthetic (polytrope + Gaussian entropy bump, seeded 1e-6 velocity noise)"
    else:
        log(f"building the low-Mach heated box {n}^3 (Mach {args.mach})")
        pb = problems.lowmach_box(n=n, heating=args.heating, mach=args.mach)
        heat = args.mach ** 3 if args.mach is not None else args.heating
        desc = (f"3D low-Mach heated box, {n}^3, ideal gas + constant gravity, USE_WB + "
                f"EVOLVE_ETOT + USE_EDOT, LHLLC + LIM2ND, RK2, CFL 0.8, heating {heat:g}"
                + (f", initial rolls at RMS Mach {args.mach:g}" if args.mach is not None else ""))
        data = ("synthetic (polytrope + bottom heating layer, seeded 1e-6 density noise"
                + (", divergence-free rolls" if args.mach is not None else "") + ")")
    return pb, desc, data


def conserved(st, cfg):
    import numpy as np
    p = st["prim"]
    rho = p[0]
    out = {"rho": rho}
    ke = np.zeros_like(rho)
    for d in range(cfg.sdims):
        out[f"m{d + 1}"] = rho * p[1 + d]
        ke += 0.5 * rho * p[1 + d] ** 2
    E = st["eint"] + ke
    if cfg.mhd:
        E += 0.5 * (st["b_cc"] ** 2).sum(axis=0)
    out["E"] = E
    for v in range(cfg.sdims + 2, cfg.nvars):
        out[f"rho_s{v - cfg.sdims - 2}"] = rho * p[v]
    return out


def deviation_errors(st, ref, pb):
    """rel. L1 of rho' = rho - rho_eq, p' = p - p_eq and of each velocity
    component, each normalised by the FP64 field's own L1 norm."""
    import numpy as np
    sd = pb.cfg.sdims
    eqr = pb.profiles["eq_rho_cc"][None, :, None]
    eqp = pb.profiles["eq_p_cc"][None, :, None]
    ip = sd + 1
    out = {}
    a, b = st["prim"][0] - eqr, ref["prim"][0] - eqr
    out["rho'"] = float(np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300))
    a, b = st["prim"][ip] - eqp, ref["prim"][ip] - eqp
    out["p'"] = float(np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300))
    for d in range(sd):
        a, b = st["prim"][1 + d], ref["prim"][1 + d]
        out[f"v{d + 1}"] = float(np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300))
    return out


def add_diffusion(pb, ratio: float):
    """Turn on THERMAL_DIFFUSION_STS with the uniform kappa whose dt_par is
    the first hyperbolic dt / ratio (dt_par = 3 kappa rho_min / (2 c sum
    1/dx^2))."""
    import numpy as np
    from phlegethon_amd.runtime import Simulation
    sim = Simulation(pb.cfg)
    sim.setup(pb)
    sim.init()
    dt = sim.scalars()["dt"]
    sim.close()
    cfg = pb.cfg
    inv_dl2 = sum(1.0 / cfg.dx[d] ** 2 for d in range(cfg.sdims))
    kap = 2.0 * 2.99792458e10 * inv_dl2 * (dt / ratio) / (3.0 * float(pb.prim[0].min()))
    nz = cfg.nx3 + 2 if cfg.sdims == 3 else 1
    pb.kappa = np.full((nz, cfg.nx2 + 2, cfg.nx1 + 2), kap)
    pb.cfg = cfg.with_(thermal_diffusion=True, eval_parabolic_dt=True)


def run_variant(pb, cfg, args, guess="warm", keep_state=False, profile=True, total_steps=None):
    """Upload, init, (train the surrogate,) warm up, time args.steps steps,
    count Newton iterations over two more, optionally profile the kernels
    for args.profile_steps, then run on to total_steps steps in all and keep
    the state."""
    from phlegethon_amd.runtime import Simulation
    sim = Simulation(cfg)
    sim.setup(pb)
    res = {}
    helm = cfg.eos == 1
    if guess == "mlp":
        t0 = time.perf_counter()
        from phlegethon_amd.problems import state_box
        box = state_box(pb) if args.mlp_box == "state" else None
        loss, rmst, rmse = sim.mlp_train(nsamples=args.mlp_samples, epochs=args.mlp_epochs,
                                         batch=4096, lr=5e-3, box=box)
        res["surrogate"] = {"train_s": time.perf_counter() - t0, "samples": args.mlp_samples,
                            "box": ([round(v, 4) for v in box] if box is not None else "table"),
                            "epochs": args.mlp_epochs, "final_batch_loss": loss,
                            "heldout_rms_log10T": rmst, "heldout_rms_log10e": rmse}
        log(f"surrogate trained in {res['surrogate']['train_s']:.2f} s, held-out RMS log10 T "
            f"{rmst:.4f}, implied log10 e {rmse:.4f}")
    if helm:
        sim.set_eos_guess(guess, resid_tol=0.25, count=False)
    sim.set_fused_sweep({"fused": True, "split": False, "auto": None}[args.sweep])
    # --graph on: phlx_run replays one captured graph per step (captured in
    # the warmup run; runs with thermal diffusion and the profiled steps
    # launch eagerly)
    sim.set_step_graph(args.graph == "on")
    sim.init()
    log(f"warmup {args.warmup} steps")
    sim.run(args.warmup)
    sim.synchronize()
    log(f"timing {args.steps} steps")
    t0 = time.perf_counter()
    sim.run(args.steps)
    sim.synchronize()
    elapsed = time.perf_counter() - t0
    zones = cfg.nx1 * cfg.nx2 * (cfg.nx3 if cfg.sdims == 3 else 1)
    res.update(value=zones * args.steps / elapsed, ms_per_step=1e3 * elapsed / args.steps)
    log(f"{res['value']:.4e} zone-updates/s, {res['ms_per_step']:.3f} ms/step")
    done = args.warmup + args.steps
    if helm:
        # Newton iterations per solve over two counted steps
        sim.set_eos_guess(guess, resid_tol=0.25, count=True)
        sim.eos_stats(reset=True)
        sim.run(2)
        done += 2
        st = sim.eos_stats()
        res["eos"] = {"newton_iters_per_solve": st["newton_iters"] / max(st["solves"], 1),
                      "fallback_rate": st["fallbacks"] / max(st["solves"], 1),
                      "guess": guess}
        log(f"EoS: {res['eos']['newton_iters_per_solve']:.3f} Newton iterations per solve, "
            f"fallback rate {res['eos']['fallback_rate']:.4f}")
        sim.set_eos_guess(guess, resid_tol=0.25, count=False)
    if profile:
        sim.profile(True)
        sim.run(args.profile_steps)
        sim.synchronize()
        done += args.profile_steps
        kern = {}
        swept = sim.profile_read("sweep")[1] > 0
        for name in ("sweep", "flux_x1", "flux_x2", "flux_x3", "corner", "faceb", "update",
                     "misc", "diffusion"):
            ms, cnt = sim.profile_read(name)
            if cnt:
                kern[name] = {"total_ms": ms, "launches": cnt, "avg_ms": ms / cnt,
                              "alg_bytes": kernel_bytes(cfg, name, swept)}
        if "diffusion" in kern:
            # per step: two diffusion steps and one dt_par pass (rho, kappa)
            s = sim.diffusion_stages()
            nz, ny, nx = cfg.shape
            per_step = 2 * diffusion_bytes(cfg, s) + 16 * nx * ny * nz
            kern["diffusion"]["alg_bytes"] = per_step * args.profile_steps / \
                kern["diffusion"]["launches"]
            kern["diffusion"]["stages"] = s
        sim.profile(False)
        res["kern"] = kern
        res["copy_gbps"] = sim.stream_copy(1 << 30, 10)
        log(f"stream copy (1 GiB x10): {res['copy_gbps']:.0f} GB/s")
    if total_steps is not None and total_steps > done:
        log(f"running on to step {total_steps}")
        sim.run(total_steps - done)
    res["scalars"] = sim.scalars()
    res["divb"] = sim.divb_max()
    if keep_state:
        res["state"] = sim.download()
    sim.close()
    return res


def rel_l1_fields(st, ref, cfg):
    """rel. L1 of every conserved field and each face-B component."""
    import numpy as np
    a, b = conserved(st, cfg), conserved(ref, cfg)
    if cfg.mhd:
        for d in range(cfg.sdims):
            k = f"b_x{d + 1}"
            a[k], b[k] = st[k], ref[k]
    return {k: float(np.abs(a[k] - b[k]).sum() / max(np.abs(b[k]).sum(), 1e-300)) for k in b}


def roofline_of(cfg, kern, value):
    """The dominant kernel's share of B_zu per launch over its average
    launch time; the whole step's B_zu rate; the per-kernel byte models."""
    dom = max(kern, key=lambda k: kern[k]["total_ms"])
    d = kern[dom]
    zones = cfg.nx1 * cfg.nx2 * (cfg.nx3 if cfg.sdims == 3 else 1)
    bzu = b_zu(cfg)
    per_launch = bzu / cfg.rk_stages * zones
    achieved = per_launch / (d["avg_ms"] * 1e-3)
    roof = {"bound": "hbm", "kernel": dom, "achieved": achieved / 1e9,
            "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": achieved / HBM_PEAK,
            "traffic": None, "alg_bytes_per_launch": per_launch,
            "alg_bytes_rule": f"B_zu / {cfg.rk_stages} stages x {zones} zones, B_zu = {bzu} B",
            "avg_launch_ms": d["avg_ms"]}
    step = {"bound": "hbm", "B_zu": bzu, "achieved": bzu * value / 1e9,
            "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": bzu * value / HBM_PEAK}
    models = {k: {"avg_ms": v["avg_ms"], "launches": v["launches"],
                  "model_bytes_per_launch": v["alg_bytes"],
                  "model_GBps": v["alg_bytes"] / (v["avg_ms"] * 1e-3) / 1e9}
              for k, v in kern.items()}
    return roof, step, models


def counter_traffic(name: str, kernel: str):
    """L2-miss bytes per dispatch of `kernel` from a committed rocprofv3
    --pmc FETCH_SIZE / WRITE_SIZE summary (profiles/<name>.json), with the
    head of the build it was measured on; None without one."""
    path = os.path.join(ROOT, "profiles", name + ".json")
    if not os.path.exists(path):
        return None, None
    with open(path) as fh:
        rec = json.load(fh)
    k = rec.get("kernels", {}).get(kernel)
    if not k:
        return None, None
    return k.get("hbm_bytes"), {"file": f"profiles/{name}.json", "head": rec.get("head"),
                                "counter": "FETCH_SIZE + WRITE_SIZE (L2-miss bytes, "
                                           "Infinity-Cache hits included), copy-calibrated"}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=None, help="zones per direction")
    ap.add_argument("--profile-steps", type=int, default=2)
    ap.add_argument("--l1-steps", type=int, default=100,
                    help="c5: steps after which the mixed runs are compared with plain FP64")
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-n", type=int, default=128,
                    help="c5: horizontal zones of the CPU baseline's full-height column")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", choices=("c5", "ot3d", "helm_slab", "lowmach", "hotbubble"),
                    default="c5")
    ap.add_argument("--variants", default="mlp,pred,warm",
                    help="c5: the mixed-precision T(rho, e) starts to run (comma list of "
                         "mlp, pred, warm; the first is the headline value)")
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="fp64")
    ap.add_argument("--sweep", choices=("auto", "fused", "split"), default="auto",
                    help="3D runs: the fused one-pass sweep, the three directional sweeps "
                         "(same results bit for bit), or the library's choice (fused without "
                         "MHD)")
    ap.add_argument("--graph", choices=("on", "off"), default="off",
                    help="run the steps as one captured HIP graph per step (bit for bit the "
                         "eager launches; measured no faster, even on the 256^2 2D grid)")
    ap.add_argument("--eos-guess", choices=("warm", "mlp", "pred"), default="warm",
                    help="Helmholtz: Newton start (carried T, the trained surrogate, or the "
                         "energy predictor)")
    ap.add_argument("--mhd-beta", type=float, default=None,
                    help="helm_slab: add a random divergence-free field at this plasma beta")
    ap.add_argument("--heating", type=float, default=1e-3,
                    help="lowmach: bottom heating amplitude")
    ap.add_argument("--compare", action="store_true",
                    help="also run the plain FP64 (warm-start) path on the same initial data "
                         "and report the speed-up and the rel. L1 error per conserved field")
    ap.add_argument("--no-fasteos", action="store_true",
                    help="helm_slab: Helmholtz without USE_FASTEOS (two T(rho, P) Newton "
                         "solves per face in the Riemann solvers)")
    ap.add_argument("--mach", type=float, default=None,
                    help="lowmach: start in convective rolls at this RMS Mach number, "
                         "heating = mach^3")
    ap.add_argument("--mlp-box", choices=("state", "table"), default="state",
                    help="surrogate training box: around the problem's states, or the "
                         "whole table range")
    ap.add_argument("--shock-flattening", type=float, default=None, metavar="EPS_SF",
                    help="USE_SHOCK_FLATTENING at this eps_sf (FP64)")
    ap.add_argument("--diffusion", type=float, default=None, metavar="RATIO",
                    help="THERMAL_DIFFUSION_STS with a uniform kappa chosen so that the "
                         "first step's dt / dt_par is RATIO")
    ap.add_argument("--mlp-samples", type=int, default=1 << 22)
    ap.add_argument("--mlp-epochs", type=int, default=8)
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(json.dumps({"error": "this build runs on one GPU only; the multi-GPU "
                          "brick decomposition is a design (not built): use --gpus 1"}))
        sys.exit(2)
    if args.n is None:
        args.n = {"c5": 512, "lowmach": 384}.get(args.workload, 256)

    pb, desc, data = build_problem(args)
    if args.shock_flattening is not None:
        pb.cfg = pb.cfg.with_(shock_flattening=True, eps_sf=args.shock_flattening)
        desc += f" | shock flattening eps_sf={args.shock_flattening:g}"
    if args.diffusion is not None:
        add_diffusion(pb, args.diffusion)
        desc += (f" | THERMAL_DIFFUSION_STS (RKL2), uniform kappa, dt/dt_par "
                 f"{args.diffusion:g} at the start")
    if args.workload == "c5":
        out = bench_c5(pb, desc, data, args)
    else:
        out = bench_single(pb, desc, data, args)
    out["cpu_baseline"] = None if args.no_cpu_baseline else run_cpu_baseline(pb, args)
    print(json.dumps(out), flush=True)


def cpu_threads(args) -> int:
    """Host threads for the CPU baseline: --cpu-threads, else the process's
    CPU share (OMP_NUM_THREADS where the host sets it, else the affinity
    mask)."""
    if args.cpu_threads:
        return args.cpu_threads
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except Exception:
        return os.cpu_count() or 1


def run_cpu_baseline(pb, args):
    threads = cpu_threads(args)
    what = "the same"
    if args.workload == "c5" and args.cpu_n and args.cpu_n < args.n:
        # a full-height column of the slab: the same vertical profile, zone
        # size, field strength and EoS work per zone, fewer horizontal zones
        from phlegethon_amd import helmholtz, problems
        from phlegethon_amd.runtime import HelmEoS
        table = helmholtz.default_table()
        eos = HelmEoS(table)
        pb = problems.helm_slab(eos, table, n=args.cpu_n, ny=args.n, mhd_beta=1e5)
        eos.close()
        what = f"a full-height {args.cpu_n}x{args.n}x{args.cpu_n} column of the {args.n}^3"
    shape = "x".join(str(v) for v in reversed(pb.cfg.shape) if v > 1)
    log(f"cpu baseline: oracle on {threads} threads, {args.cpu_steps} steps of {shape}")
    rate, wall = cpu_baseline(pb, args.cpu_steps, threads)
    zones = pb.cfg.nx1 * pb.cfg.nx2 * (pb.cfg.nx3 if pb.cfg.sdims == 3 else 1)
    # wct/cell/cycle/core in microseconds (source/source.F90:6832-6833)
    us = wall / args.cpu_steps / zones * 1e6 * threads
    log(f"cpu baseline {rate:.4e} zone-updates/s ({us:.3f} us/cell/cycle/core, "
        f"{wall:.1f} s)")
    return {"value": rate, "unit": "zone-updates/s", "cores": threads, "kind": "port",
            "wct_per_cell_cycle_core_us": us, "host": host_description(threads),
            "sample": f"{args.cpu_steps} full step(s) of {what} workload ({shape}, "
                      f"{wall:.1f} s wall) after the oracle's step-0 initialisation; C FP64 "
                      "restatement with OpenMP (-O3 -march=x86-64-v3 -ffp-contract=off)"}


def bench_single(pb, desc, data, args):
    """One configuration at one precision (the non-default workloads)."""
    import numpy as np
    from phlegethon_amd.config import PREC_MIXED
    cfg = pb.cfg
    if args.precision == "mixed":
        cfg = cfg.with_(precision=PREC_MIXED)
        desc += " | mixed precision (FP32 recon/Riemann, FP64 deviations/update/CT/EoS)"
    else:
        desc += " | FP64"
    if args.eos_guess != "warm":
        desc += f" | T(rho, e) start: {args.eos_guess}"
    res = run_variant(pb, cfg, args, guess=args.eos_guess, keep_state=args.compare)
    value = res["value"]
    roof, step, models = roofline_of(cfg, res["kern"], value)
    step["stream_copy_GBps"] = res["copy_gbps"]
    if args.precision == "fp64" and args.workload == "ot3d" and args.n == 256 and \
            args.shock_flattening is None and args.diffusion is None:
        roof["traffic"], roof["traffic_source"] = counter_traffic("ot3d256_fp64_pmc",
                                                                  roof["kernel"])
    out = {
        "metric": "zone-updates/s (" + ("plain FP64 path" if args.precision == "fp64" else
                                       "mixed precision") + "; HBM-roofline fraction)",
        "value": value, "unit": "zone-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": res["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64", "data": data,
        "config": {"workload": desc, "zones": int(np.prod(cfg.shape)),
                   "sim_time": res["scalars"]["time"], "max_divB": res["divb"],
                   "step_graph": args.graph == "on"},
        "roofline": roof, "step_roofline": step, "kernel_models": models,
    }
    if "eos" in res:
        out["eos"] = res["eos"]
    if "surrogate" in res:
        out["surrogate"] = res["surrogate"]
    if args.compare:
        log("plain FP64 warm-start run on the same initial data")
        base = run_variant(pb, pb.cfg, args, guess="warm", keep_state=True, profile=False,
                           total_steps=args.warmup + args.steps + (2 if cfg.eos == 1 else 0)
                           + args.profile_steps)
        rel = rel_l1_fields(res["state"], base["state"], pb.cfg)
        out["compare"] = {"plain_value": base["value"], "plain_ms_per_step": base["ms_per_step"],
                          "speedup": value / base["value"], "rel_l1_vs_plain_fp64": rel,
                          "plain_eos": base.get("eos")}
        log(f"speed-up {out['compare']['speedup']:.3f}x; rel L1 " +
            " ".join(f"{k} {v:.2e}" for k, v in rel.items()))
        if "eq_rho_cc" in pb.profiles:
            out["compare"]["rel_l1_deviations"] = deviation_errors(res["state"], base["state"], pb)
    return out


def bench_c5(pb, desc, data, args):
    """Plain FP64 against the mixed-precision runs with each T(rho, e) start
    on the combined workload; the first variant is the headline value."""
    import numpy as np
    from phlegethon_amd.config import PREC_MIXED
    variants = [v.strip() for v in args.variants.split(",") if v.strip()]
    log("plain FP64 (reference algorithm, Newton from the carried T)")
    plain = run_variant(pb, pb.cfg, args, guess="warm", keep_state=True,
                        total_steps=args.l1_steps)
    roof, step, models = roofline_of(pb.cfg, plain["kern"], plain["value"])
    step["stream_copy_GBps"] = plain["copy_gbps"]
    roof["path"] = "plain FP64"
    roof["traffic"], roof["traffic_source"] = counter_traffic("c5_512_fp64_pmc", roof["kernel"])
    ref_state = plain.pop("state")
    mixed_cfg = pb.cfg.with_(precision=PREC_MIXED)
    runs = {}
    for v in variants:
        log(f"mixed precision, T(rho, e) start: {v}")
        r = run_variant(pb, mixed_cfg, args, guess=v, keep_state=True,
                        total_steps=args.l1_steps)
        rel = rel_l1_fields(r.pop("state"), ref_state, pb.cfg)
        log(f"  speed-up {r['value'] / plain['value']:.3f}x; rel L1 after {args.l1_steps} "
            "steps " + " ".join(f"{k} {e:.2e}" for k, e in rel.items()))
        mroof, mstep, mmodels = roofline_of(mixed_cfg, r["kern"], r["value"])
        runs[v] = {"value": r["value"], "ms_per_step": r["ms_per_step"],
                   "speedup_vs_plain_fp64": r["value"] / plain["value"],
                   "rel_l1_vs_plain_fp64": rel, "l1_after_steps": args.l1_steps,
                   "max_rel_l1": max(rel.values()), "eos": r.get("eos"),
                   "max_divB": r["divb"], "kernels": mmodels, "step_roofline": mstep}
        if "surrogate" in r:
            runs[v]["surrogate"] = r["surrogate"]
    del ref_state
    if not variants:
        # plain FP64 only (profiling runs)
        runs["fp64"] = {"value": plain["value"], "ms_per_step": plain["ms_per_step"]}
        variants = ["fp64"]
    head = variants[0]
    names = {"mlp": "FP32 + NN-EoS guess", "pred": "FP32 + energy-predictor guess",
             "warm": "FP32 + carried-T guess", "fp64": "plain FP64"}
    out = {
        "metric": "zone-updates/s, plain FP64 vs FP32+NN-EoS guess at fixed L1 err; % HBM peak",
        "value": runs[head]["value"], "unit": "zone-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": runs[head]["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64", "data": data,
        "config": {"workload": desc + f" | value: {names[head]} (mixed precision)",
                   "zones": int(np.prod(pb.cfg.shape)), "l1_after_steps": args.l1_steps,
                   "sim_time_plain": plain["scalars"]["time"], "max_divB_plain": plain["divb"],
                   "step_graph": args.graph == "on"},
        "roofline": roof, "step_roofline": step,
        "plain_fp64": {"value": plain["value"], "ms_per_step": plain["ms_per_step"],
                       "eos": plain.get("eos"), "kernels": models},
        "mixed": runs,
    }
    return out


if __name__ == "__main__":
    main()
