#!/usr/bin/env python3
"""Benchmark: DOF-updates/s per RK stage of the DG right-hand side, Euler p=2
on 524,288 triangles (isentropic vortex, periodic [0,10]^2 split 512x512 quads
with interior vertices jittered +-0.2h from seed 20261202, LLF, SSP-RK3,
CFL 0.2), one MI355X.

A "step" is one SSP-RK3 time step = 3 RK stages, each one fused stage kernel
over all elements (inputs resident in HBM).  value = K * n_p * N * stages /
time.  roofline: algorithmic bytes per stage (state in + state out + 2/3 of a
base-state read + per-element metric/connectivity + per-side normal/length)
over the stage kernel's average duration measured with HIP events on the
library's stream.  cpu_baseline: the oracle (a C restatement of the
reference's element and edge routines) on a bounded sample of the same
workload on this host's cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
METRIC = "DOF-updates/s per RK stage (Euler p=2, 524k tris) + adjoint/forward time ratio"
SEED = 20261202
VORTEX = (5.0, 5.0, 5.0, 1.0, 1.0, 10.0)  # beta, x0, y0, u_inf, v_inf, period
EVERY = 16  # RK-stage checkpoints every 16 steps (the Maxwell TE adjoint workload)
# untimed stepping after the warm-up steps: the GPU's clock rises from about
# 2.3 to 2.55 GHz over the first tens of milliseconds of load (GRBM_GUI_ACTIVE
# per dispatch), so the timed steps start once the clocks have settled
SETTLE_S = 0.5
T_START = time.perf_counter()


def progress(msg):
    """a progress line on stderr (the legs run for minutes in total)"""
    print("[bench %6.1fs] %s" % (time.perf_counter() - T_START, msg), file=sys.stderr, flush=True)


def lib_sha256():
    """sha256 of the libdgamd.so this process loads (ties a profile to a build)"""
    import hashlib
    from dgcuda_amd import _lib as L
    with open(L.LIB_PATH, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


def measured_traffic(sha):
    """HBM bytes per stage launch measured with rocprofv3 (FETCH_SIZE and
    WRITE_SIZE, each scaled by a plain copy's reading) for the library build
    with this sha256, as recorded in the committed profile summary of this
    workload; None when no profile of this build is there"""
    path = os.path.join(ROOT, "profiles", "k_stage_euler_p2_524k.md")
    found = {}
    try:
        with open(path) as f:
            for line in f:
                for key in ("traffic_bytes_per_stage:", "libdgamd_sha256:"):
                    if line.startswith(key):
                        found[key] = line.split(":", 1)[1].strip()
    except OSError:
        return None
    if found.get("libdgamd_sha256:") != sha or "traffic_bytes_per_stage:" not in found:
        return None
    return float(found["traffic_bytes_per_stage:"])


def stage_bytes(K, S, n_p, N, r):
    """algorithmic HBM bytes of one fused RK stage"""
    return 8 * K * n_p * N * (2 + r) + K * (5 * 8 + 3 * 4) + S * (3 * 8 + 4 * 4)


BASELINE_CFLAGS = "-O3 -march=native -fPIC -fopenmp -ffp-contract=off"


def cpu_baseline(mesh, steps_all=5, steps_one=1, warmup=1):
    """the oracle on this host, same mesh and state, compiled here with
    BASELINE_CFLAGS: steps_all SSP-RK3 steps on every core of the share and
    steps_one on one thread, each after `warmup` untimed steps"""
    import tempfile
    import oracle as O
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    dof = mesh.K * 6 * 4
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        lib = O.load(O.build_baseline(tmp, BASELINE_CFLAGS))
        for th, ns in ((cores, steps_all), (1, steps_one)):
            progress("CPU baseline: %d steps on %d threads" % (ns, th))
            o = O.Oracle(O.EULER, 2, integrator=4, cfl=0.2, problem=2, problem_params=VORTEX,
                         threads=th, lib=lib)
            o.set_mesh(mesh).project_ic()
            if warmup:
                rc, _, _ = o.run(warmup)
                assert rc == 0
            t0 = time.perf_counter()
            rc, _, _ = o.run(ns)
            el = time.perf_counter() - t0
            assert rc == 0
            out[th] = dof * 3 * ns / el
            del o
    return cores, out


def host_cpu():
    """the host's CPU (/proc/cpuinfo: model name, family, model, stepping) and
    its logical CPU count"""
    info = {"model name": "unknown"}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if ":" not in line:
                    if info["model name"] != "unknown":
                        break
                    continue
                k, v = (x.strip() for x in line.split(":", 1))
                if k == "model name":
                    info["model name"] = v
                elif k in ("vendor_id", "cpu family", "model", "stepping"):
                    info[{"vendor_id": "vendor", "cpu family": "family"}.get(k, k)] = v
    except OSError:
        pass
    # the model name some hosts report is only the product line; the family
    # names the core generation (AMD family 25: Zen 3 / Zen 4, 26: Zen 5)
    gen = {"25": "Zen 3 / Zen 4", "26": "Zen 5"}.get(info.get("family", ""))
    if info.get("vendor") == "AuthenticAMD" and gen:
        info["core"] = "family %s (%s), model %s" % (info["family"], gen, info.get("model", "?"))
    return info, os.cpu_count()


def adjoint_leg(n, nsteps, fd_dirs=32):
    """the adjoint/forward time ratio: Maxwell TE (Ex, Ey, Hz), p=3, walled
    unit square split n x n quads, per-element eps ~ U[1,4] (seeded), Gaussian
    Hz pulse (sigma 0.05), exact-Riemann upwind flux, RK4 at a fixed dt (the
    CFL step of the initial state), nsteps steps; objective: the squared Hz
    over a receiver patch.  The gradient (dJ/du0 and dJ/deps) costs one
    checkpointed forward solve plus the reverse sweep (recompute + VJPs).
    ratio_gradient: the whole gradient (checkpointed forward + sweep) over a
    plain forward solve of the same steps; ratio_sweep: the sweep alone over it."""
    import numpy as np
    import torch
    from dgcuda_amd import _lib as L
    from dgcuda_amd import mesh as M
    from dgcuda_amd.solver import Solver
    mesh = M.structured(n, periodic=False, jitter=0.2, seed=SEED, bc=(-1, -1, -1, -1))
    eps = np.random.default_rng(SEED).uniform(1.0, 4.0, mesh.K)
    s = Solver(L.MAXWELL_TE, 3, integrator=L.RK4, problem=L.PROB_MAXWELL_PULSE,
               problem_params=(1.0, 0.5, 0.5, 0.05), device=0)
    s.set_mesh(mesh)
    s.set_eps(eps)
    s.project_ic()
    u0 = s.coeffs()
    _, dt = s.step(1)  # the CFL step of the initial state
    s.set_objective(L.OBJ_PATCH_SQ, (0.6, 0.7, 0.3, 0.4, 2))
    # warm-up (graphs, buffers), then the plain forward solve
    s.set_coeffs(u0)
    s.step(min(10, nsteps), dt=dt)
    s.set_coeffs(u0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.step(nsteps, dt=dt)
    torch.cuda.synchronize()
    fwd = time.perf_counter() - t0
    s.set_coeffs(u0)
    t0 = time.perf_counter()
    J, gu, ge = s.adjoint_solve(nsteps, dt, every=EVERY, want_eps=True)
    grad = time.perf_counter() - t0
    fwd_ck, rev = s.adjoint_timing()
    resident = s.info().adj_resident
    assert np.isfinite(J) and np.all(np.isfinite(gu)) and np.all(np.isfinite(ge))
    # central differences of full forward solves along unit random directions,
    # alternating initial-state and permittivity directions (outside the timing)
    rng = np.random.default_rng(SEED + 1)
    fd = {"u0": [], "eps": []}
    fd_vals = []
    progress("adjoint leg: gradient %.2f s, forward %.2f s; %d FD directions" % (grad, fwd, fd_dirs))
    for q in range(fd_dirs):
        if q % 4 == 0:
            progress("adjoint leg: FD direction %d" % q)
        if q % 2 == 0:
            d = rng.standard_normal(u0.shape)
            d /= np.linalg.norm(d)
            h = 1e-3 * np.linalg.norm(u0)
            Jpm = []
            for sg in (1.0, -1.0):
                s.set_coeffs(u0 + sg * h * d)
                s.step(nsteps, dt=dt)
                Jpm.append(s.objective())
            ad, key = float(np.vdot(gu, d)), "u0"
        else:
            # J is smooth but not quadratic in eps: fourth-order differences,
            # with a step large enough that J's rounding stays far below the
            # small directional derivatives
            d = rng.standard_normal(eps.shape)
            d /= np.linalg.norm(d)
            h = 1e-3
            Jpm = []
            for sg in (1.0, -1.0, 2.0, -2.0):
                s.set_eps(eps + sg * h * d)
                s.set_coeffs(u0)
                s.step(nsteps, dt=dt)
                Jpm.append(s.objective())
            s.set_eps(eps)
            ad, key = float(np.vdot(ge, d)), "eps"
        if key == "u0":  # J is quadratic in u0: central differences are exact
            fdv = (Jpm[0] - Jpm[1]) / (2 * h)
        else:
            fdv = (8 * (Jpm[0] - Jpm[1]) - (Jpm[2] - Jpm[3])) / (12 * h)
        fd[key].append(abs(fdv - ad) / max(abs(fdv), abs(ad), 1e-300))
        fd_vals.append((key, fdv, ad))
    s.close()
    return {"workload": "Maxwell TE p=3 RK4, walled unit square %d x %d quads -> %d triangles, "
                        "eps ~ U[1,4], Gaussian Hz pulse, %d steps at dt %.3e, receiver-patch "
                        "objective, checkpoints every %d steps"
                        % (n, n, mesh.K, nsteps, dt, EVERY),
            "forward_s": fwd, "gradient_s": grad, "checkpointed_forward_s": fwd_ck * 1e-3,
            "reverse_s": rev * 1e-3, "segments": -(-nsteps // EVERY),
            # segments whose stage data stayed in HBM from the forward pass
            "resident_segments": resident,
            # the whole gradient (checkpointed forward + sweep) over the forward solve
            "ratio_gradient": grad / fwd,
            # the reverse sweep (recompute + stage adjoints) alone over the forward solve
            "ratio_sweep": rev * 1e-3 / fwd, "J": J,
            "fd_check": {"directions": fd_dirs, "rel_err_u0": fd["u0"], "rel_err_eps": fd["eps"],
                         "max_rel_err": max(fd["u0"] + fd["eps"]) if fd_dirs else None,
                         "fd_and_adjoint": fd_vals,
                         "method": "differences of full forward solves along unit random "
                                   "directions: central with step 1e-3 |u0| for u0 (J is "
                                   "quadratic in u0), fourth-order with step 1e-3 for eps"}}


def ensemble_leg(ninst, nsteps, seq_steps=20):
    """Batched ensemble: ninst shallow-water instances (g = 9.81) on seeded
    jittered walled unit-square meshes with element counts log-uniform in
    [2k, 200k], p drawn from {1, 2, 3}, h = 1 + 0.1 exp(-|x-c|^2/sigma^2)
    with c uniform in the unit square and sigma uniform in [0.05, 0.2], zero
    velocity, Roe flux, SSP-RK3 with each instance's CFL dt, nsteps steps.
    One launch per RK stage covers every instance; the elements per block are
    picked by timing.  Beside it: the same instances stepped one at a time on
    their own handles (seq_steps steps each)."""
    import numpy as np
    import torch
    from dgcuda_amd import _lib as L
    from dgcuda_amd import mesh as M
    from dgcuda_amd.ensemble import Ensemble
    from dgcuda_amd.solver import Solver
    rng = np.random.default_rng(SEED + 7)
    Ks = np.exp(rng.uniform(np.log(2e3), np.log(2e5), ninst))
    ns = np.maximum(2, np.rint(np.sqrt(Ks / 2))).astype(int)
    ps = rng.integers(1, 4, ninst)
    prms = [(1.0, 0.1, rng.uniform(), rng.uniform(), rng.uniform(0.05, 0.2), 1.0)
            for _ in range(ninst)]
    kw = dict(physics=L.SWE, integrator=L.SSPRK3, flux=L.FLUX_ROE, g=9.81,
              problem=L.PROB_SWE_BUMP)
    E = Ensemble(**kw)
    t0 = time.perf_counter()
    meshes = []
    for i in range(ninst):
        m = M.structured(int(ns[i]), periodic=False, jitter=0.2, seed=SEED + i, bc=(-1, -1, -1, -1))
        meshes.append(m)
        E.add(m, int(ps[i]), prms[i])
    E.build(0)
    setup = time.perf_counter() - t0
    E.step(2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    E.step(nsteps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    inf = E.info()
    for i in (0, ninst - 1):
        assert np.all(np.isfinite(E.coeffs(i)))
    E.close()
    # the same instances, one handle each, one after another
    seq = 0.0
    for i in range(ninst):
        s = Solver(kw["physics"], int(ps[i]), integrator=kw["integrator"], flux=kw["flux"],
                   g=9.81, problem=kw["problem"], problem_params=prms[i])
        s.set_mesh(meshes[i]).project_ic()
        s.step(2)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        s.step(seq_steps)
        torch.cuda.synchronize()
        seq += time.perf_counter() - t1
        s.close()
    rate = inf.dofs * 3 * nsteps / wall
    # algorithmic bytes of one stage of every instance (each its own fused
    # stage's bytes; the padding to 256 elements not counted)
    b_stage = sum(stage_bytes(m.K, m.S, (int(q) + 1) * (int(q) + 2) // 2, 3, 2.0 / 3.0)
                  for m, q in zip(meshes, ps))
    ms_step = wall * 1e3 / nsteps
    return {"workload": "%d shallow-water instances, %d to %d triangles (%d in all), p in {1,2,3}, "
                        "Roe flux, SSP-RK3, %d steps" % (ninst, int(2 * ns.min() ** 2),
                                                         int(2 * ns.max() ** 2), inf.elements,
                                                         nsteps),
            "dofs": inf.dofs, "seconds": wall, "ms_per_step": ms_step,
            "dof_updates_per_s": rate, "block": inf.block,
            # the step's three stage launches and its closing launch against
            # three stages' algorithmic bytes
            "bytes_per_stage": b_stage,
            "roofline_frac": 3 * b_stage / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "tune_ms_per_step": {"64": inf.tune_ms[0], "128": inf.tune_ms[1], "256": inf.tune_ms[2],
                                 "64 split": inf.tune_split_ms[0], "128 split": inf.tune_split_ms[1],
                                 "256 split": inf.tune_split_ms[2]},
            "split": bool(inf.split),
            "block_per_order": {str(q): inf.block_per_order[q] for q in range(6)
                                if inf.block_per_order[q]},
            "tune_best_ms_per_step": inf.tune_best_ms,
            "vgprs": inf.vgprs, "launches_per_step": inf.launches_per_step,
            "setup_s": setup,
            "one_handle_each": {"steps": seq_steps, "ms_per_step": seq * 1e3 / seq_steps,
                                "dof_updates_per_s": inf.dofs * 3 * seq_steps / seq}}


def c5_leg(nsteps, n=1024, p=4, fd_dirs=8, fd_h=(1e-3, 1e-4, 1e-5), fd_nsteps=100, volume=0):
    """Adjoint at scale: the isentropic vortex of the main workload on an n x n
    periodic mesh (1024 -> 2,097,152 triangles) at p = 4, SSP-RK3, CFL 0.2.
    (1) the forward stage rate; (2) the gradient of the total entropy
    -integral(rho s) w.r.t. the initial coefficients over nsteps fixed-dt
    steps with sqrt(nsteps) checkpointing: the adjoint/forward ratios; (3)
    the gradient checked by central differences of full forward solves: an
    O(1) linear objective J = <g, u_T> (g Gaussian, scaled to |J| ~ 1) over
    fd_nsteps steps, fd_dirs unit random directions, each at every step in
    fd_h (the relative error falls with h until rounding or an LLF branch
    switch sets its floor)."""
    import math
    import numpy as np
    import torch
    from dgcuda_amd import _lib as L
    from dgcuda_amd import mesh as M
    from dgcuda_amd.solver import Solver
    mesh = M.structured(n, lx=10.0, periodic=True, jitter=0.2, seed=SEED)
    s = Solver(L.EULER, p, integrator=L.SSPRK3, cfl=0.2, problem=L.PROB_EULER_VORTEX,
               problem_params=VORTEX, device=0, volume=volume)
    s.set_mesh(mesh).project_ic()
    inf = s.info()
    u0 = s.coeffs()
    _, dt = s.step(5)
    t, _ = s.step(min(nsteps, 20), t=0.0)
    ms_stage, _ = s.timing()
    K, S, n_p, N = inf.K, inf.S, inf.n_p, inf.N
    B = stage_bytes(K, S, n_p, N, 2.0 / 3.0)
    every = int(math.ceil(math.sqrt(nsteps)))
    s.set_objective(L.OBJ_ENTROPY)
    s.set_coeffs(u0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.step(nsteps, dt=dt)
    torch.cuda.synchronize()
    fwd = time.perf_counter() - t0
    progress("Euler p=4 leg: forward %.2f s; gradient" % fwd)
    s.set_coeffs(u0)
    t0 = time.perf_counter()
    J, gu, _ = s.adjoint_solve(nsteps, dt, every=every)
    grad = time.perf_counter() - t0
    fwd_ck, rev = s.adjoint_timing()
    resident = s.info().adj_resident
    # (3) finite differences of an O(1) linear objective
    rng = np.random.default_rng(SEED + 5)
    g = rng.standard_normal(u0.shape)
    s.set_coeffs(u0)
    s.step(fd_nsteps, dt=dt)
    g *= 1.0 / abs(float(np.vdot(g, s.coeffs())))
    s.set_objective(L.OBJ_LINEAR, weights=g)
    s.set_coeffs(u0)
    Jl, gl, _ = s.adjoint_solve(fd_nsteps, dt, every=int(math.ceil(math.sqrt(fd_nsteps))))
    fd = []
    for q in range(fd_dirs):
        progress("Euler p=4 leg: FD direction %d" % q)
        d = rng.standard_normal(u0.shape)
        d /= np.linalg.norm(d)
        ad = float(np.vdot(gl, d))
        row = {"adjoint": ad}
        for h in fd_h:
            Jpm = []
            for sg in (1.0, -1.0):
                s.set_coeffs(u0 + sg * h * d)
                s.step(fd_nsteps, dt=dt)
                Jpm.append(s.objective())
            fdv = (Jpm[0] - Jpm[1]) / (2 * h)
            row["rel_err h=%g" % h] = abs(fdv - ad) / max(abs(fdv), abs(ad), 1e-300)
        fd.append(row)
    s.close()
    best = [min(v for k, v in r.items() if k.startswith("rel_err")) for r in fd]
    return {"workload": "Euler isentropic vortex p=%d SSP-RK3, %d x %d periodic quads -> %d "
                        "triangles, entropy objective, %d steps at dt %.3e, checkpoints every %d"
                        % (p, n, n, K, nsteps, dt, every),
            "dofs": K * n_p * N, "ms_per_stage": ms_stage,
            "dof_updates_per_s": K * n_p * N / (ms_stage * 1e-3),
            "roofline_frac": B / (ms_stage * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "forward_s": fwd, "checkpointed_forward_s": fwd_ck * 1e-3, "reverse_s": rev * 1e-3,
            "segments": -(-nsteps // every), "resident_segments": resident,
            # the whole gradient (checkpointed forward + sweep) over a forward solve
            "ratio_gradient": grad / fwd,
            # the reverse sweep (recompute + stage adjoints) alone over a forward solve
            "ratio_sweep": rev * 1e-3 / fwd,
            "J_entropy": J,
            "fd_check": {"objective": "<g, u_T>, g Gaussian scaled so |J| = 1, over %d steps"
                                      % fd_nsteps,
                         "J": Jl, "directions": fd_dirs, "per_direction": fd,
                         "best_rel_err_per_direction": best,
                         "max_best_rel_err": max(best) if best else None,
                         "method": "central differences of full forward solves along unit "
                                   "random directions at each step h"}}


# euler/benchmark.sh: dgsupersonic -b -t <steps> -n <p> <mesh> (meshes E and F of the
# script are not in the reference tree)
SWEEP = [("A", 1, 100000), ("A", 2, 50000), ("A", 3, 40000), ("A", 4, 30000), ("A", 5, 20000),
         ("B", 1, 80000), ("B", 2, 60000), ("B", 3, 30000), ("B", 4, 15000), ("B", 5, 10000),
         ("C", 1, 40000), ("C", 2, 30000), ("C", 3, 15000), ("C", 4, 8000), ("C", 5, 1000),
         ("D", 1, 20000), ("D", 2, 10000), ("D", 3, 5000), ("D", 4, 2000), ("D", 5, 1000)]


def sweep_leg(scale=1.0):
    """The reference's benchmark sweep (conservation/euler/benchmark.sh:3-37):
    the supersonic vortex (euler/supersonic, RK4, the reference's CFL 0.5) on
    its nested meshes A-D (180 to 11,520 triangles) at p = 1..5, each for the
    script's step count times `scale`, timed around the whole time loop as
    the -b flag times it (main.cu:824-849).  No published numbers exist."""
    import torch
    from dgcuda_amd import _lib as L
    from dgcuda_amd import mesh as M
    from dgcuda_amd.solver import Solver
    meshes = {}
    out = []
    for name, p, steps in SWEEP:
        if name not in meshes:
            meshes[name] = M.read_pmsh(os.path.join(ROOT, "tests", "golden",
                                                    "supersonic_%s.pmsh" % name))
        m = meshes[name]
        nst = max(1, int(round(steps * scale)))
        progress("supersonic sweep: mesh %s, p = %d, %d steps" % (name, p, nst))
        s = Solver(L.EULER, p, integrator=L.RK4, problem=L.PROB_EULER_SUPERSONIC,
                   problem_params=(), device=0)
        s.set_mesh(m).project_ic()
        s.step(min(nst, 64))  # graphs built, kernels loaded
        s.project_ic()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        t, _ = s.step(nst)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        err = s.l2_error(t)
        n_p = (p + 1) * (p + 2) // 2
        out.append({"mesh": name, "K": m.K, "p": p, "steps": nst, "seconds": wall,
                    "dof_updates_per_s": m.K * n_p * 4 * 4 * nst / wall,
                    "t_final": t, "l2_rho_error": float(err[0])})
        s.close()
    return {"workload": "supersonic vortex (euler/benchmark.sh): meshes A-D x p = 1..5, RK4, "
                        "CFL 0.5, the script's step counts x %g" % scale,
            "cells": out, "total_seconds": sum(c["seconds"] for c in out)}


def rothill_leg(nsteps, p=4):
    """The reference's only published timing (conservation/conserv_kernels.cu:1090-1092):
    rotating-hill advection, p = 4, 1000 RK4 steps, 15.28 s with the shipped
    kernels on an unstated Fermi-era GPU and mesh.  Timed here on the
    reference's mesh A (1,080 triangles), CFL dt every step, the
    time-dependent inflow refreshed at every stage."""
    import torch
    from dgcuda_amd import _lib as L
    from dgcuda_amd import mesh as M
    from dgcuda_amd.solver import Solver
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden",
                        "rotatinghill_A.pmsh")
    mesh = M.read_pmsh(path)
    s = Solver(L.ADVECTION, p, integrator=L.RK4, problem=L.PROB_ADV_ROTHILL, problem_params=(),
               device=0)
    s.set_mesh(mesh).project_ic()
    s.step(10)
    s.project_ic()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    t, _ = s.step(nsteps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # the hill rotated to time t (the velocity 2 pi (-y, x) turns it once per
    # unit time), against the solution at the elements' vertices
    import numpy as np
    u = np.stack(s.eval_u(0))
    x, y = mesh.V[0::2], mesh.V[1::2]
    c, sn = np.cos(2 * np.pi * t), np.sin(2 * np.pi * t)
    xr, yr = x * c + y * sn - 0.2, -x * sn + y * c
    exact = 5 * np.exp(-(xr * xr + yr * yr) / (2 * 0.15 ** 2))
    s.close()
    return {"workload": "rotating-hill advection p=%d RK4, reference mesh A (%d triangles), "
                        "%d steps, CFL 0.5" % (p, mesh.K, nsteps),
            "seconds": wall, "t_final": t,
            "vertex_max_error_vs_rotated_hill": float(np.abs(u - exact).max()),
            "hill_height": 5.0,
            "reference_seconds": 15.28,
            "reference_note": "source comment, hardware and mesh not stated"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--n", type=int, default=512, help="quads per side (512 -> 524,288 tris)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--adj-steps", type=int, default=2000,
                    help="steps of the adjoint workload (0 skips the adjoint leg)")
    ap.add_argument("--adj-n", type=int, default=316,
                    help="quads per side of the adjoint workload (316 -> 199,712 tris)")
    ap.add_argument("--adj-fd", type=int, default=32,
                    help="random directions of the finite-difference check of the adjoint "
                         "gradient (alternating u0 and eps)")
    ap.add_argument("--ens-inst", type=int, default=256,
                    help="instances of the ensemble leg (0 skips it)")
    ap.add_argument("--ens-steps", type=int, default=500, help="steps of the ensemble leg")
    ap.add_argument("--c5-steps", type=int, default=1000,
                    help="steps of the Euler p=4 2M-triangle adjoint leg (0 skips it)")
    ap.add_argument("--c5-fd", type=int, default=8,
                    help="finite-difference directions of the Euler p=4 leg (0 skips the check)")
    ap.add_argument("--c5-volume", choices=("default", "vector", "mfma"), default="default",
                    help="volume integral of the Euler p=4 leg's stage kernel")
    ap.add_argument("--rothill-steps", type=int, default=1000,
                    help="steps of the rotating-hill leg (0 skips it)")
    ap.add_argument("--sweep-scale", type=float, default=1.0,
                    help="step counts of the supersonic benchmark sweep relative to "
                         "euler/benchmark.sh (0 skips the leg)")
    ap.add_argument("--settle", type=float, default=SETTLE_S, metavar="SECONDS",
                    help="untimed stepping after the warm-up steps (GPU clock ramp)")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB five times (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters against known bytes)")
    args = ap.parse_args()
    if args.calib_copy:
        import torch
        n = args.calib_copy * (1 << 20) // 8
        x = torch.ones(n, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(5):
            y.copy_(x)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 5
        print(json.dumps({"calib_copy_bytes_each_way": n * 8, "copies": 6,
                          "copy_ms": ms, "copy_GBps_read_plus_write": 2 * n * 8 / ms / 1e6}))
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the library has no multi-GPU path "
              "(the sharding is designed, not built); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from dgcuda_amd import _lib as L
    from dgcuda_amd import mesh as M
    from dgcuda_amd.solver import Solver

    torch.cuda.init()
    mesh = M.structured(args.n, lx=10.0, periodic=True, jitter=0.2, seed=SEED)
    s = Solver(L.EULER, 2, integrator=L.SSPRK3, cfl=0.2, problem=L.PROB_EULER_VORTEX,
               problem_params=VORTEX, device=0)
    s.set_mesh(mesh).project_ic()
    inf = s.info()
    K, S, n_p, N = inf.K, inf.S, inf.n_p, inf.N
    stages = 3

    progress("main leg: K = %d, %d warm-up steps" % (K, args.warmup))
    t = 0.0
    t, _ = s.step(args.warmup, t=t)
    settle = 0
    ts = time.perf_counter()
    while time.perf_counter() - ts < args.settle:
        t, _ = s.step(20, t=t)
        settle += 20
    torch.cuda.synchronize()
    settle_s = time.perf_counter() - ts
    t0 = time.perf_counter()
    t, dt = s.step(args.steps, t=t)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    progress("main leg: %d timed steps in %.2f ms" % (args.steps, wall * 1e3))
    ms_stage_ev, nst = s.timing()  # HIP events around the stage launches only
    assert nst == args.steps * stages

    dof = K * n_p * N
    value = dof * stages * args.steps / wall
    B = stage_bytes(K, S, n_p, N, 2.0 / 3.0)
    achieved = B / (ms_stage_ev * 1e-3) / 1e9
    err = s.l2_error(t)
    assert np.all(np.isfinite(err)) and err[0] < 1e-2, err

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "settle": {"steps": settle, "seconds": settle_s,
                   "why": "untimed steps after the warm-up until the GPU clock has ramped "
                          "from idle (about 2.3 -> 2.55 GHz over the first tens of ms)"},
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded jittered periodic mesh, isentropic vortex initial state)",
        "config": {"workload": "Euler isentropic vortex p=2 SSP-RK3 CFL 0.2, %d x %d quads -> "
                               "%d triangles, %d sides, LLF flux, positivity check on"
                               % (args.n, args.n, K, S),
                   "K": K, "S": S, "n_p": n_p, "N": N, "dofs": dof, "stages_per_step": stages,
                   "t_final": t, "dt_last": dt, "l2_rho_error": float(err[0])},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(lib_sha256()) if args.n == 512 else None,
                     "bytes_per_stage": B, "ms_per_stage": ms_stage_ev,
                     "kernel": "dgk::k_stage<1,2> (Euler, p=2)"},
    }
    s.close()
    if args.adj_steps > 0:
        progress("adjoint leg (Maxwell TE p=3)")
        line["adjoint"] = adjoint_leg(args.adj_n, args.adj_steps, args.adj_fd)
    if args.rothill_steps > 0:
        progress("rotating-hill leg")
        line["rotating_hill"] = rothill_leg(args.rothill_steps)
    if args.sweep_scale > 0:
        progress("supersonic sweep leg")
        line["supersonic_sweep"] = sweep_leg(args.sweep_scale)
    if args.ens_inst > 0 and args.ens_steps > 0:
        progress("ensemble leg")
        line["ensemble"] = ensemble_leg(args.ens_inst, args.ens_steps)
    if args.c5_steps > 0:
        progress("Euler p=4 2M-triangle leg")
        line["adjoint_at_scale"] = c5_leg(args.c5_steps, fd_dirs=args.c5_fd,
                                          volume={"default": 0, "vector": 1, "mfma": 2}[args.c5_volume])
    if not args.no_cpu:
        progress("CPU baseline (build + oracle run)")
        cores, cb = cpu_baseline(mesh)
        line["cpu_baseline"] = {"value": cb[cores], "unit": "DOF-updates/s", "cores": cores,
                                "kind": "port",
                                "sample": "5 SSP-RK3 steps (15 stages) of the same 524,288-triangle "
                                          "workload after 1 warm-up step, oracle (C, OpenMP) on %d "
                                          "threads" % cores,
                                "value_1_thread": cb[1],
                                "sample_1_thread": "1 SSP-RK3 step on one thread after 1 warm-up "
                                                   "step"}
        info, ncpu = host_cpu()
        line["cpu_baseline"].update({"cpu_model": info.pop("model name"), "cpu_id": info,
                                     "nproc": ncpu, "build": "gcc " + BASELINE_CFLAGS})
    progress("done")
    print(json.dumps(line))


if __name__ == "__main__":
    main()
