#!/usr/bin/env python3
"""Throughput of the plain FP64 MPM-STD time step on one MI355X.

Workload: 512^3 cells (n1m=n2m=n3m=512), Lx=Lz=2*pi, Ly=2, periodic x/z,
no-slip walls in y with a tanh-stretched y grid (gamma 2.0), Ra=1e8,
Pr=0.71, water property polynomials with DeltaT=0.2, T = conduction profile
+ seeded noise of amplitude 0.01.  The time step is fixed (default 1e-3)
while Post_CFL is still evaluated every step.  Two properties of the
reference scheme fix these choices (the CPU oracle shows them):
  * from rest its CFL rule returns dtMax = 100*min(dx) (0.39 at 512^3), and
    a step of that size drives maxDivU to ~1e3;
  * on one y-rank its thermal boundary routine sets only the bottom ghost
    row of T (if / else if), so the top row is set here to Tup, as the top
    rank of a multi-rank run sets it;
  * its variable-density pressure correction (ExtraTerm with the in-place
    dPhat update of the projection and a wall ghost of dPhat that is never
    set) amplifies the divergence at the wall-adjacent cells by a factor
    that grows with the density contrast and the wall-normal resolution: at
    n2m=512 it diverges for DeltaT >= 0.5 and decays for DeltaT = 0.2.
One step is the loop body of the reference's main program: thermal,
momentum (dU, dV, dW, block corrections, update), pressure (RHS, FFT/TDMA
Poisson, projection), every ghost-cell update, Post_Div and Post_CFL.

Prints one JSON line (metric Mcell-steps/s = cells * steps / seconds / 1e6)
with the dominant kernel's HBM roofline (HIP events on the library stream)
and the CPU oracle timed on this host as cpu_baseline.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mcell-steps/s at 512³: FP64 plain vs FP32-mixed at stated error; HBM GB/s"
HBM_PEAK_GBS = 8000.0
# the mixed form's accuracy target: every state field and the wall Nusselt
# numbers within this relative error of the plain FP64 form
MIXED_TARGET_REL = 1e-6


def workload_config(n, props="water", deltat=0.2):
    from pascal_tcs_amd.config import Properties, TCSConfig
    p = {"water": Properties.water(), "constant": Properties.constant()}[props]
    return TCSConfig(n1m=n, n2m=n, n3m=n, uniform2=0, gamma2=2.0, Ra=1e8, Pr=0.71, DeltaT=deltat,
                     MaxCFL=0.5, dtStart=1e-3, props=p)


def kernel_bytes(label, cfg):
    """Algorithmic HBM bytes of one launch of a step kernel: each input array
    read once and each output written once (1-D metrics and (j,k) tables are
    free).  TDMA sweeps: read a,b,c,d and write x, 40 B per unknown."""
    n1, n2, n3 = cfg.n1m, cfg.n2m, cfg.n3m
    N = n1 * n2 * n3
    fv = (n2 - 1) / n2  # dV systems skip the wall row
    table = {
        # 4 non-cyclic y sweeps (T, dU, dV[n2-1 rows], dW) per step
        "ptdma": 40.0 * N * (3 + fv) / 4,
        # 8 cyclic z/x sweeps (T, dU, dW full; dV without the wall row)
        "ptdma_cyclic": 40.0 * N * (6 + 2 * fv) / 8,
        # split-line kernels: y sweeps; z sweeps (strided, cyclic); x sweeps (contiguous, cyclic)
        "ptdma_part": 40.0 * N * (3 + fv) / 4,
        "ptdma_part_cyclic": 40.0 * N * (3 + fv) / 4,
        "ptdma_part_rows_cyclic": 40.0 * N * (3 + fv) / 4,
        "ptdma_part_f32": 20.0 * N * (3 + fv) / 4,
        "ptdma_part_cyclic_f32": 20.0 * N * (3 + fv) / 4,
        "ptdma_part_rows_cyclic_f32": 20.0 * N * (3 + fv) / 4,
        # Poisson y solves: complex right-hand side in and solution out per (kx, kz, j)
        "poisson_part": 32.0 * (n1 // 2 + 1) * n3 * n2,
        "thermal_z": 8.0 * N * (4 + 4), "thermal_x": 8.0 * N * (2 + 3), "thermal_y": 8.0 * N * (2 + 3),
        "dU_z": 8.0 * N * (5 + 4), "dV_z": 8.0 * N * (5 + 4) * fv, "dW_z": 8.0 * N * (6 + 4),
        "dU_x": 8.0 * N * (1 + 3), "dU_y": 8.0 * N * (1 + 3), "dV_x": 8.0 * N * (1 + 3) * fv,
        "dV_y": 8.0 * N * (1 + 3) * fv, "dW_x": 8.0 * N * (1 + 3), "dW_y": 8.0 * N * (1 + 3),
        # block corrections with the pseudo-update: V, dW, dV in, dV and V + dV
        # out; U, dV, dW, dU, W in, dU, U + dU and W + dW out
        "blockLdV": 8.0 * N * (3 + 2) * fv, "blockLdU": 8.0 * N * (5 + 3),
        # projection with the dPhat history: U, V, W, P, dP, dPhat, dP0 in;
        # U, V, W, P, dPhat, dP0 out
        "pseudoupdate": 8.0 * N * 9, "projection": 8.0 * N * (7 + 6), "dphat_update": 8.0 * N * 5,
        "pressure_rhs": 8.0 * N * (4 + 1), "add_interior": 8.0 * N * 3, "store_interior": 8.0 * N * 2,
        "post_div_cfl": 8.0 * N * 3,
        # split line solves that compute their rows: the stencil's fields in,
        # the solution out (z: into the scratch; y: the scratch's right-hand
        # side in, the increment field out)
        "thermal_z_solve": 8.0 * N * (4 + 1), "dU_z_solve": 8.0 * N * (5 + 1),
        "dV_z_solve": 8.0 * N * (5 + 1) * fv, "dW_z_solve": 8.0 * N * (6 + 1),
        "thermal_y_solve": 8.0 * N * (3 + 1), "dU_y_solve": 8.0 * N * (2 + 1),
        "dV_y_solve": 8.0 * N * (2 + 1) * fv, "dW_y_solve": 8.0 * N * (2 + 1),
        # x solves: the velocity field and the z solution in, the solution out
        "dU_x_solve": 8.0 * N * (2 + 1), "dV_x_solve": 8.0 * N * (2 + 1) * fv, "dW_x_solve": 8.0 * N * (2 + 1),
        # Poisson transforms: the real field (8 B a cell) and the half
        # spectrum (16 B a mode, n1/2 + 1 modes a line); the step's forward x
        # transform makes the right-hand side from U, V, W, dPhat
        "fft_x_fwd": 8.0 * N + 16.0 * (n1 // 2 + 1) * n2 * n3,
        "rhs_fft_x_fwd": 8.0 * N * 4 + 16.0 * (n1 // 2 + 1) * n2 * n3,
        "fft_z<-1>": 32.0 * (n1 // 2 + 1) * n2 * n3, "fft_z<+1>": 32.0 * (n1 // 2 + 1) * n2 * n3,
        "fft_x_inv": 8.0 * N + 16.0 * (n1 // 2 + 1) * n2 * n3, "dp_subtract": 16.0 * N,
    }
    return table.get(label)


def run_gpu(args, precision, routines=False, profile=True):
    """routines=True: each step as the reference's main program issues it,
    one routine entry at a time (the drop-in sequence of the Fortran
    modules) instead of the fused tcs_step."""
    import torch
    from pascal_tcs_amd.solver import TCSSolver

    cfg = workload_config(args.n, deltat=args.deltat)
    s = TCSSolver(cfg, tdma_split=args.tdma_split, precision=precision)
    s.setup_seeded(amp=0.01, seed=20240512)
    T = s.get("T")
    T[:, -1, :] = cfg.Tup  # top wall row, as the top y-rank of a multi-rank run sets it
    s.set("T", T)
    del T
    state = {"dt": args.dt if args.dt > 0 else cfg.dtStart, "div": 0.0, "cfl": 0.0, "divs": []}

    def one_step():
        if routines:
            s.step_routines(state["dt"])
        else:
            s.step(state["dt"])
        state["div"] = s.mpi_Post_Div()
        state["divs"].append(state["div"])
        cfl, newdt = s.mpi_Post_CFL()
        if args.dt > 0:
            state["cfl"] = state["dt"] * cfl / newdt  # CFL number of the step just taken
        else:
            state["cfl"], state["dt"] = cfl, newdt

    for _ in range(args.warmup):
        one_step()
    torch.cuda.synchronize()
    s.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        one_step()
    s.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    sec = t1 - t0

    # per-stage / per-kernel HIP-event timers on the library's stream
    rep = {}
    if profile:
        s.profile(True)
        for _ in range(args.profile_steps):
            one_step()
        s.sync()
        rep = s.profile_report()
        s.profile(False)
    nu = s.nusselt()
    fields = {n: s.get(n) for n in ("T", "U", "V", "W", "P")} if args.mixed and not routines else None
    s.close()
    return cfg, sec, rep, state, nu, fields


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_step512_fp64.json")


def measured_traffic(label, cfg):
    """HBM bytes per launch of a kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE passes (512^3 workload only), or None."""
    if (cfg.n1m, cfg.n2m, cfg.n3m) != (512, 512, 512) or not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as fh:
        t = json.load(fh)["bytes_per_launch"].get(label)
    return None if t is None else t["read_bytes"] + t["write_bytes"]


def roofline(cfg, rep, psteps):
    kernels = {k: v for k, v in rep.items() if kernel_bytes(k, cfg) is not None}
    if not kernels:
        return None, None
    dom = max(kernels, key=lambda k: kernels[k][0])
    ms, cnt = kernels[dom]
    avg_s = ms / cnt / 1e3
    ach = kernel_bytes(dom, cfg) / avg_s / 1e9
    info = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": measured_traffic(dom, cfg),
            "kernel": dom, "avg_launch_ms": round(ms / cnt, 4), "launches_per_step": cnt // max(psteps, 1),
            "bytes_per_launch": kernel_bytes(dom, cfg)}
    return dom, info


def host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count()


def run_cpu_baseline(seconds, deltat):
    """The CPU oracle (C restatement, OpenMP build: every loop nest and line
    solve split over the threads the host gives this process, bit-identical
    to the serial build) on the full 512^3 workload: the same seeded state,
    one timed step (thermal, momentum, pressure, halos, Div, CFL).  More
    steps follow while the elapsed time stays under `seconds`."""
    from oracle import oracle as O
    from pascal_tcs_amd.ic import conduction_noise_T, embed_interior
    from tests.cases import make_oracle

    cfg = workload_config(512, deltat=deltat)
    o = make_oracle(cfg, omp=True)
    threads = int(O.lib(omp=True).omp_get_max_threads())  # libgomp's, reached through the oracle build
    o.f["T"][...] = embed_interior(conduction_noise_T(cfg, o.x2, 0.01, 20240512))
    o.bc["TBCup"][...] = cfg.Tup
    o.bc["TBCbt"][...] = cfg.Tbt
    o.thermal_boundary()
    o.momentum_boundary()
    o.f["T"][:, -1, :] = cfg.Tup
    for n in ("T", "U", "V", "W", "P"):
        o.halo(n)
    dt = 1e-3
    steps, t0 = 0, time.perf_counter()
    while True:
        o.step(dt)
        o.div()
        o.cfl(dt)
        steps += 1
        el = time.perf_counter() - t0
        print(f"cpu baseline: {steps} oracle step(s) at 512^3 in {el:.1f} s on {threads} threads",
              file=sys.stderr, flush=True)
        if el + el / steps > seconds or steps >= 5:
            break
    cells = cfg.n1m * cfg.n2m * cfg.n3m
    del o
    return {"value": round(cells * steps / el / 1e6, 4), "unit": "Mcell-steps/s", "cores": threads, "kind": "port",
            "sample": f"CPU oracle (C restatement of the reference's routines, OpenMP build, {threads} threads) on "
                      f"the full 512^3 workload, {steps} step(s) in {el:.1f} s; the host offers {host_cores()} cores "
                      f"to this process"}


def run_tdma_microbench(n, reps=5):
    """Batched tridiagonal solves on n^3 FP64 arrays along x, y and z in turn
    (n^2 systems of n rows per axis), non-cyclic and cyclic, with the split
    kernels and the single-rank kernels.  Seeded systems: a, c ~ U(-1, 0),
    b = |a| + |c| + 1 + U(0, 1), d ~ U(-1, 1).  Each timed solve starts from
    fresh copies (the single-rank solve overwrites A and C); HIP events on
    the stream the solve runs on bracket the solve alone.  40 algorithmic
    bytes per unknown (read a, b, c, d; write x)."""
    import torch
    from pascal_tcs_amd import ptdma

    g = torch.Generator(device="cuda")
    g.manual_seed(20240512)
    shape = (n, n, n)
    a0 = -torch.rand(shape, generator=g, device="cuda", dtype=torch.float64)
    c0 = -torch.rand(shape, generator=g, device="cuda", dtype=torch.float64)
    b0 = a0.abs() + c0.abs() + 1.0 + torch.rand(shape, generator=g, device="cuda", dtype=torch.float64)
    d0 = torch.rand(shape, generator=g, device="cuda", dtype=torch.float64) * 2.0 - 1.0
    A, C, D = torch.empty_like(a0), torch.empty_like(c0), torch.empty_like(d0)
    nbytes = 40.0 * n ** 3
    cases = []
    for axis, name in ((0, "x"), (1, "y"), (2, "z")):
        for cyclic in (False, True):
            for kind, solve in (("split", ptdma.axis_solve_split), ("single-rank", ptdma.axis_solve)):
                times = []
                for _ in range(reps + 1):
                    A.copy_(a0); C.copy_(c0); D.copy_(d0)
                    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                    e0.record()
                    solve(A, b0, C, D, axis, cyclic)
                    e1.record()
                    torch.cuda.synchronize()
                    times.append(e0.elapsed_time(e1))
                ms = float(np.median(times[1:]))
                gbs = nbytes / (ms / 1e3) / 1e9
                cases.append({"axis": name, "cyclic": cyclic, "kernel": kind, "ms": round(ms, 4),
                              "GB_per_s": round(gbs, 1), "frac_of_hbm_peak": round(gbs / HBM_PEAK_GBS, 4)})
                print(f"tdma {name} cyclic={cyclic} {kind}: {ms:.3f} ms {gbs:.0f} GB/s", file=sys.stderr, flush=True)
    return {"metric": f"batched tridiagonal solve at {n}^3 (40 B per unknown)", "unit": "GB/s",
            "systems_per_axis": n * n, "rows": n, "dtype": "f64", "data": "synthetic", "cases": cases}


def run_mixed_drift(n, steps, deltat, dt, report=(10, 100, 250, 500, 750, 1000)):
    """The mixed form and the plain FP64 form stepped side by side from the
    same seeded state: per-field max and L2 errors of the mixed run relative
    to the plain run, and the wall Nusselt numbers, at the listed steps."""
    from pascal_tcs_amd.solver import TCSSolver
    cfg = workload_config(n, deltat=deltat)
    plain = TCSSolver(cfg, tdma_split=1, precision=1)
    plain.setup_seeded(amp=0.01, seed=20240512)
    T = plain.get("T")
    T[:, -1, :] = cfg.Tup  # top wall row, as in the throughput run
    plain.set("T", T)
    del T
    mixed = TCSSolver(cfg, tdma_split=1, precision=2)
    for name in ("T", "U", "V", "W", "P", "dPhat", "dP0"):
        mixed.set(name, plain.get(name))
    for w in ("TBCup", "TBCbt", "UBCup", "UBCbt", "VBCup", "VBCbt", "WBCup", "WBCbt"):
        mixed.set_wall(w, plain.get_wall(w))
    rows = []
    for step in range(1, steps + 1):
        plain.step(dt)
        mixed.step(dt)
        if step in report or step == steps:
            row = {"step": step}
            for name in ("T", "U", "V", "W", "P"):
                a, b = plain.get(name), mixed.get(name)
                d = b - a
                row[name] = {"max": float(f"{np.abs(d).max() / max(np.abs(a).max(), 1e-300):.3e}"),
                             "l2": float(f"{np.linalg.norm(d) / max(np.linalg.norm(a), 1e-300):.3e}")}
                del a, b, d
            nbp, ntp = plain.nusselt()
            nbm, ntm = mixed.nusselt()
            row["Nu_plain"] = [nbp, ntp]
            row["Nu_rel_drift"] = [float(f"{abs(nbm - nbp) / abs(nbp):.3e}"), float(f"{abs(ntm - ntp) / abs(ntp):.3e}")]
            row["maxDivU_plain"] = plain.mpi_Post_Div()
            rows.append(row)
            print(f"drift step {step}: " + json.dumps(row), file=sys.stderr, flush=True)
    plain.close()
    mixed.close()
    return {"metric": f"mixed form vs plain FP64 at {n}^3 after {steps} steps", "dt": dt, "deltat": deltat,
            "errors": rows}


def run_adjoint_case(n, nsteps, every, ndir, deltat):
    """Gradient of the time-averaged bottom-wall Nusselt number over an
    nsteps rollout with respect to an n x n bottom-wall temperature
    perturbation (seeded Gaussian, sigma 0.01): 2 pi x 2 x 2 pi box, uniform
    grid, Ra 1e7, water properties, conduction profile plus noise 0.05,
    dt 1e-3, plain FP64 step (precision 0, split lines), checkpoints every
    `every` steps.  Checked against central differences on ndir random
    directions, and against the gradient of the single-rank-order step."""
    import torch
    from pascal_tcs_amd.adjoint import STATE as ST, NusseltRollout, nusselt_loss_function
    from pascal_tcs_amd.config import Properties, TCSConfig
    from pascal_tcs_amd.solver import TCSSolver
    cfg = TCSConfig(n1m=n, n2m=n, n3m=n, uniform2=1, Ra=1e7, Pr=0.71, DeltaT=deltat, props=Properties.water(),
                    dtStart=1e-3)
    out = {"metric": f"adjoint gradient of the mean bottom-wall Nusselt number, {n}^3, {nsteps} steps",
           "config": {"workload": f"{n}^3 uniform box, Ra 1e7, water properties DeltaT {deltat:g}, noise 0.05, "
                                  f"dt 1e-3, {nsteps} steps, checkpoints every {every}",
                      "parameter": f"{n}x{n} bottom-wall temperature perturbation, Gaussian sigma 0.01"}}
    grads = {}
    for split in (1, 0):
        s = TCSSolver(cfg, tdma_split=split, precision=0)
        s.setup_seeded(amp=0.05, seed=20240512)
        state0 = {k: s.get(k) for k in ST}
        rng = np.random.default_rng(4)
        theta0 = 0.01 * rng.standard_normal((n, n))
        roll = NusseltRollout(s, state0, nsteps=nsteps, dt=cfg.dtStart, every=every,
                              log=lambda m: print(m, file=sys.stderr, flush=True))
        F = nusselt_loss_function()
        th = torch.tensor(theta0, dtype=torch.float64, device="cuda", requires_grad=True)
        t0 = time.perf_counter()
        L = F.apply(th, roll)
        t1 = time.perf_counter()
        L.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        grads[split] = th.grad.detach().cpu().numpy()
        if split == 1:
            out["loss"] = float(L)
            out["forward_s"] = round(t1 - t0, 3)
            out["backward_s"] = round(t2 - t1, 3)
            out["grad_norm"] = float(np.linalg.norm(grads[1]))
            checks = []
            eps = 1e-5
            for r in range(ndir):
                d = rng.standard_normal(theta0.shape)
                fd = (roll.loss(theta0 + eps * d)[0] - roll.loss(theta0 - eps * d)[0]) / (2 * eps)
                ad = float(np.vdot(grads[1], d))
                checks.append({"fd": fd, "adjoint": ad, "rel": abs(fd - ad) / abs(fd)})
                print(f"adjoint direction {r}: fd {fd:.10e} adjoint {ad:.10e}", file=sys.stderr, flush=True)
            out["fd_checks"] = checks
            out["fd_max_rel"] = max(c["rel"] for c in checks)
        s.close()
    out["vs_single_rank_order_gradient_rel"] = float(np.linalg.norm(grads[1] - grads[0]) / np.linalg.norm(grads[0]))
    return out


def run_adjoint_step(n, deltat):
    """One adjoint step at n^3 (the config-4 workload, precision 0, split
    lines): wall time of one plain forward step and of one tcs_adjoint_step
    (which runs the step forward keeping what the backward pass needs, then
    the backward pass), each after one untimed call."""
    import torch
    from pascal_tcs_amd.adjoint import StepAdjoint
    from pascal_tcs_amd.config import Properties, TCSConfig
    from pascal_tcs_amd.solver import TCSSolver
    cfg = TCSConfig(n1m=n, n2m=n, n3m=n, uniform2=1, Ra=1e7, Pr=0.71, DeltaT=deltat, props=Properties.water(),
                    dtStart=1e-3)
    s = TCSSolver(cfg, tdma_split=1, precision=0)
    s.setup_seeded(amp=0.05, seed=20240512)
    a = StepAdjoint(s)
    a.set_bar("T", np.random.default_rng(1).standard_normal(s.shape))
    out = {}
    for label, fn in (("forward_step_ms", lambda: s.step(cfg.dtStart)), ("adjoint_step_ms", lambda: a.step(cfg.dtStart))):
        fn()
        s.sync()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        s.sync()
        torch.cuda.synchronize()
        out[label] = round((time.perf_counter() - t0) * 1e3, 3)
    s.close()
    out["ratio"] = round(out["adjoint_step_ms"] / out["forward_step_ms"], 2)
    return {"metric": f"one adjoint step at {n}^3 against one forward step", **out}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--profile-steps", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=30.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--dt", type=float, default=1e-3, help="fixed time step (<= 0: the reference's CFL rule)")
    ap.add_argument("--deltat", type=float, default=0.2, help="DeltaT of the property polynomials")
    ap.add_argument("--precision", type=int, default=1, choices=(0, 1, 2),
                    help="0: stencils in the reference's operation order; 1: FMA contraction and reciprocals "
                         "(within 1e-12 per step of the reference order); 2: mixed form, FP32 stage rows, line solves and "
                         "line solves for the increments, FP64 state and pressure")
    ap.add_argument("--tdma-split", type=int, default=1, choices=(0, 1),
                    help="1: ADI lines split into parts as PaScaL_TDMA splits them over ranks, reduced "
                         "systems by parallel cyclic reduction (within 1e-14 of the single-rank order); "
                         "0: single-rank order")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run one device-to-device copy of 1 GiB (the reference kernel for scaling "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--no-mixed", dest="mixed", action="store_false",
                    help="skip the mixed form's run (by default it follows the plain run on the same workload)")
    ap.add_argument("--no-dropin", dest="dropin", action="store_false",
                    help="skip the drop-in run (by default the plain form is also timed with each step issued "
                         "one routine entry at a time, as the reference's main program calls them)")
    ap.add_argument("--tdma", action="store_true",
                    help="instead of the step: the isolated batched tridiagonal solves at n^3 along x, y, z")
    ap.add_argument("--mixed-drift", type=int, default=0, metavar="STEPS",
                    help="instead of the step: run the mixed and plain forms side by side for STEPS steps "
                         "and report the mixed form's per-field errors and Nusselt drift")
    ap.add_argument("--adjoint-step", action="store_true",
                    help="instead of the step: one adjoint step against one forward step (n^3, default 128)")
    ap.add_argument("--adjoint", action="store_true",
                    help="instead of the step: the adjoint gradient case (n^3, default 128, 50 steps) checked "
                         "against central differences")
    args = ap.parse_args()
    if args.precision == 2:
        args.mixed = False
    if args.gpus != 1:
        print(f"bench.py: this build runs the whole domain on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    if args.tdma:
        print(json.dumps(run_tdma_microbench(args.n)))
        return
    if args.adjoint_step:
        print(json.dumps(run_adjoint_step(args.n if args.n != 512 else 128, args.deltat)))
        return
    if args.adjoint:
        n = args.n if args.n != 512 else 128
        print(json.dumps(run_adjoint_case(n, 50, 10, 16, args.deltat)))
        return
    if args.mixed_drift > 0:
        print(json.dumps(run_mixed_drift(args.n, args.mixed_drift, args.deltat, args.dt)))
        return
    if args.copy_calibration:
        import torch
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    cfg, sec, rep, state, nu, fields = run_gpu(args, args.precision)
    cells = cfg.n1m * cfg.n2m * cfg.n3m
    value = cells * args.steps / sec / 1e6
    dom, roof = roofline(cfg, rep, args.profile_steps)
    stages = {k: round(v[0] / args.profile_steps, 3) for k, v in rep.items() if " " in k or k == "step"}
    kernels = {k: [round(v[0] / args.profile_steps, 3), v[1] // args.profile_steps]
               for k, v in rep.items() if " " not in k and k != "step"}
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision < 2 else "f64 state / f32 stencils and ADI",
        "data": "synthetic",
        "config": {"workload": f"plain FP64 MPM-STD step, {cfg.n1m}x{cfg.n2m}x{cfg.n3m}, tanh y (gamma 2.0), "
                               f"Ra=1e8, Pr=0.71, water properties DeltaT={cfg.DeltaT:g}, T noise 0.01, "
                               + (f"fixed dt {args.dt:g} (CFL evaluated every step)" if args.dt > 0 else "CFL 0.5"),
                   "cells": cells,
                   "form": {0: "plain FP64, reference operation order",
                            1: "plain FP64, contracted/reciprocal stencil arithmetic",
                            2: "mixed: FP32 stage rows, block corrections, line solves and increments; FP64 state/pressure/reductions"}[args.precision] + (
                       "; ADI lines split into 32 parts per line as PaScaL_TDMA splits them over ranks, "
                       "reduced systems by parallel cyclic reduction"
                       if args.tdma_split else "; single-rank line solves")},
        "roofline": roof,
        "stage_ms_per_step": stages,
        "kernel_ms_per_step": kernels,
        "final": {"dt": state["dt"], "maxDivU": state["div"], "CFL": state["cfl"], "Nu_bottom": nu[0],
                  "Nu_top": nu[1], "maxDivU_per_step": [float(f"{d:.3e}") for d in state["divs"]]},
    }
    if args.dropin:
        # the same workload and step count with each step issued one routine
        # entry at a time with every ghost-cell update, in the order of the
        # reference's main program (the calls the Fortran drop-in modules make)
        _, sec3, _, state3, _, _ = run_gpu(args, args.precision, routines=True, profile=False)
        v3 = cells * args.steps / sec3 / 1e6
        out["dropin"] = {"value": round(v3, 2), "unit": "Mcell-steps/s", "ms_per_step": round(sec3 / args.steps * 1e3, 3),
                         "vs_fused_step": round(v3 / value, 4),
                         "sequence": "routine entries in src/main.f90:131-203 order (mpi_thermal_solver ... "
                                     "mpi_pressure_Projection, 14 ghost-cell updates) plus Post_Div and Post_CFL",
                         "final_maxDivU": state3["div"]}
    if args.mixed:
        # the mixed form on the same workload and step count, measured the same
        # way; its fields after warmup + steps + profile steps are compared with
        # the plain run's (relative to each field's magnitude)
        cfg2, sec2, rep2, state2, nu2, fields2 = run_gpu(args, 2)
        v2 = cells * args.steps / sec2 / 1e6
        err = {n: float(f"{np.abs(fields2[n] - fields[n]).max() / max(np.abs(fields[n]).max(), 1e-300):.3e}")
               for n in fields}
        nu_err = [float(f"{abs(nu2[q] - nu[q]) / abs(nu[q]):.3e}") for q in (0, 1)]
        over = sorted(n for n, e in err.items() if e > MIXED_TARGET_REL)
        over += [w for w, e in zip(("Nu_bottom", "Nu_top"), nu_err) if e > MIXED_TARGET_REL]
        out["mixed"] = {"value": round(v2, 2), "unit": "Mcell-steps/s", "ms_per_step": round(sec2 / args.steps * 1e3, 3),
                        "speedup_vs_plain": round(v2 / value, 3),
                        "form": "FP32 stage rows and block corrections (from FP32 copies of the arrays they read), FP32 line solves "
                                "and increments dU/dV/dW; FP64 state accumulation, pressure, "
                                "divergence/CFL/Nusselt reductions",
                        "rel_error_vs_plain_after_steps": err, "Nu_rel_error_vs_plain": nu_err,
                        "steps_compared": args.warmup + args.steps + args.profile_steps,
                        "target_rel": MIXED_TARGET_REL, "meets_target": not over, "over_target": over,
                        "Nu_bottom": nu2[0], "Nu_top": nu2[1],
                        "kernel_ms_per_step": {k: round(v[0] / args.profile_steps, 3) for k, v in rep2.items()
                                               if " " not in k and k != "step"}}
        del fields, fields2
    if not args.no_cpu:
        out["cpu_baseline"] = run_cpu_baseline(args.cpu_seconds, args.deltat)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
