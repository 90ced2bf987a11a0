"""Throughput of the explicit Godunov step on one MI355X.

Workload (default): the supersonic viscous cylinder, M = 3.5, Re_D = 1e5,
O-grid of 720 x 1445 nodes split into 2,079,360 triangles, cells in a seeded
random order, ns scheme (least-squares gradients + Venkatakrishnan, AUSM+,
Sutherland viscous fluxes, isothermal wall) at CFL 0.4, starting from the
uniform free stream.  A "step" is one explicit Euler step over all cells,
every array resident in HBM.  The FP64 step is the headline value; the
FP32-mixed step (same mesh, same start) is reported beside it with its density,
pressure and wall-drag deviation from the FP64 run after the same steps.  The
reference scheme (the bit-faithful restatement of the reference's step) is
timed beside them on a 2M-triangle periodic Taylor-Green square (``ref_scheme``).

Prints one JSON line.  Only one GPU is supported (--gpus N>1 exits).
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
PKG = importlib.import_module("2d-dns-godunov-solver-_amd")
L = PKG._lib
M = PKG.mesh
S = PKG.stage

METRIC = "cell-updates/s @2.08M tris: FP32-mixed vs FP64 at stated drag err; adjoint/fwd cost"
HBM_PEAK_GBS = 8000.0
SEED = 20261203


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_mesh(n_theta, n_r, dr0, order):
    t0 = time.perf_counter()
    mesh = M.make_cylinder_ogrid_mesh(n_theta, n_r, 0.5, 10.0, dr0)
    if order == "random":
        mesh = M.permute_cells(mesh, M.random_permutation(mesh.nCells, SEED))
    log(f"mesh {n_theta}x{n_r}: {mesh.nCells} cells, {mesh.nNodes} nodes ({time.perf_counter() - t0:.1f} s)")
    return mesh


def initial_U(case, n):
    r, u, v, p = case.free_stream(n)
    U = np.zeros((n, 4), order="F")
    U[:, 0] = r; U[:, 1] = r * u; U[:, 2] = r * v
    U[:, 3] = p / (case.gamma - 1.0) + 0.5 * r * (u * u + v * v)
    return U


def run_variant(mesh, case, precision, flux, order_dev, U0, warmup, steps, cfl):
    prm = case.params(precision=precision, flux=flux)
    t0 = time.perf_counter()
    sol = S.GodunovSolver(mesh, prm, order=order_dev)
    log(f"device mesh + state ({time.perf_counter() - t0:.1f} s), bandwidth {sol.dmesh.info().bandwidth}")
    sol.state.set_array(L.GD_ARR_U_OLD, U0)
    if warmup:
        sol.run(warmup, cfl=cfl, dt_policy=L.GD_DT_LOCAL, wall_code=-3)
    sol.state.sync()
    t0 = time.perf_counter()
    st = sol.run(steps, cfl=cfl, dt_policy=L.GD_DT_LOCAL, wall_code=-3)
    sol.state.sync()
    el = time.perf_counter() - t0
    U = sol.conservative()  # after warmup + steps
    f = sol.fields()
    # per-kernel HIP-event times from a second pass of the same number of
    # steps (HIP events around every launch on the state's stream; they add
    # ~25 us per step, so the pass that gives ms_per_step runs without them)
    sol.state.sync()
    t1 = time.perf_counter()
    kt = L.kernel_table(sol.run(steps, cfl=cfl, dt_policy=L.GD_DT_LOCAL, wall_code=-3, time_kernels=True))
    sol.state.sync()
    el_ev = time.perf_counter() - t1
    sol.close()
    return {"ms_per_step": 1e3 * el / steps, "seconds": el, "kernels": kt, "ms_per_step_with_kernel_events":
            1e3 * el_ev / steps, "U": U, "rho": f[0], "p": f[3],
            "drag": float(st.wall_force[0]), "lift": float(st.wall_force[1]), "flow_time": float(st.flow_time),
            "nonphysical": int(st.nonphysical)}


def adjoint_cost(mesh, case, order_dev, U0, steps, ckpt, cfl, precision=L.GD_FP32_MIXED, budget_gib=16):
    """Discrete adjoint of the summed wall drag over ``steps`` steps (FP32-mixed
    by default; fixed dt, checkpoints every ``ckpt``) against the same forward
    steps alone, from the state after 20 CFL steps.  The headline run keeps
    the cell records of as many trailing checkpoint segments as fit in
    ``budget_gib`` GiB of HBM and recomputes the others from their
    checkpoints; beside it the run that keeps one segment only (the least
    memory) and the one that keeps every segment (nothing recomputed)."""
    prm = case.params(precision=precision)
    sol = S.GodunovSolver(mesh, prm, order=order_dev)
    sol.state.set_array(L.GD_ARR_U_OLD, U0)
    st = sol.run(20, cfl=cfl, dt_policy=L.GD_DT_LOCAL, wall_code=-3)
    dt = st.last_dt
    U1 = sol.conservative()
    sol.run(steps, dt=dt, wall_code=-3)  # warm-up of the fixed-dt path
    sol.state.set_array(L.GD_ARR_U_OLD, U1)
    sol.state.sync()
    t0 = time.perf_counter()
    st = sol.run(steps, dt=dt, wall_code=-3)
    sol.state.sync()
    fwd = time.perf_counter() - t0
    J = st.wall_force_sum[0]
    sol.state.set_array(L.GD_ARR_U_OLD, U1)
    sol.adjoint(2, dt, checkpoint_every=1, wall_code=-3)  # warm-up

    def timed(**kw):
        sol.state.set_array(L.GD_ARR_U_OLD, U1)
        sol.state.sync()
        t0 = time.perf_counter()
        g, gb, st = sol.adjoint(steps, dt, wall_code=-3, **kw)
        sol.state.sync()
        return time.perf_counter() - t0, g, gb, st

    budget = int(budget_gib) << 30
    # dJ/dU0 left in HBM, as the forward leaves U (the host copy is timed apart)
    adj, _, gbc, ast = timed(checkpoint_every=ckpt, record_budget=budget, to_host=False)
    # one segment's records only (the least memory: every other segment recomputed)
    adj0, _, _, ast0 = timed(checkpoint_every=ckpt, to_host=False)
    # every segment's records kept in HBM (nothing recomputed)
    adj1, _, _, ast1 = timed(checkpoint_every=ckpt, record_budget=64 << 30, to_host=False)
    # with dJ/dU0 copied to the caller's host array (66.5 MB over PCIe)
    adj_h, g, _, _ = timed(checkpoint_every=ckpt, record_budget=budget)
    _, g0, _, _ = timed(checkpoint_every=ckpt)
    _, g1, _, _ = timed(checkpoint_every=ckpt, record_budget=64 << 30)
    sol.close()

    def variant(a_s, st, gv, gib):
        return {"record_budget_GiB": gib, "recomputed_steps": st.recomputed_steps, "adjoint_ms": round(a_s * 1e3, 3),
                "cost_vs_forward": round(a_s / fwd, 3), "storage_GB": round(st.storage_bytes / 1e9, 3),
                "reverse_step_vs_forward_step": round(st.ms_reverse / (fwd * 1e3), 3),
                "gradient_bit_identical": bool(np.array_equal(gv, g))}
    return {"steps": steps, "checkpoint_every": ckpt,
            "precision": "fp32-mixed" if precision == L.GD_FP32_MIXED else "fp64", "dt": dt,
            "record_budget_GiB": budget_gib,
            "forward_ms": round(fwd * 1e3, 3), "adjoint_ms": round(adj * 1e3, 3),
            "cost_vs_forward": round(adj / fwd, 3),
            "adjoint_ms_with_host_copy": round(adj_h * 1e3, 3),
            # per step: the reverse sweep (recomputation included, HIP events)
            # over the forward step
            "reverse_step_vs_forward_step": round(ast.ms_reverse / (fwd * 1e3), 3),
            "adjoint_forward_sweep_ms": round(ast.ms_forward, 3), "adjoint_reverse_ms": round(ast.ms_reverse, 3),
            "recomputed_steps": ast.recomputed_steps, "storage_GB": round(ast.storage_bytes / 1e9, 3),
            "J_sum_drag": ast.J, "J_rel_diff_vs_forward": abs(ast.J - J) / abs(J),
            "dJ_dU0_linf": float(np.abs(g).max()), "dJ_dinflow": [float(x) for x in gbc[1]],
            "one_segment": variant(adj0, ast0, g0, 0),
            "all_records": variant(adj1, ast1, g1, 64)}


def ref_adjoint_cost(sol, mesh, steps, ckpt, dt, budget_gib=16):
    """Discrete adjoint of J = <w, U_N> over ``steps`` fixed-dt steps of the
    reference loop (w seeded N(0, 1)), against the same steps forward with
    gd_run at that dt: with every state kept from the forward sweep
    (``budget_gib`` of HBM; the headline) and checkpointed every ``ckpt``
    steps with the rest recomputed.  Timed by the host clock around
    synchronized calls; dJ/dU0 stays on the device, w is uploaded and U_N read
    back for J inside the timed call."""
    n = mesh.nCells
    w = np.asfortranarray(M.normal(SEED + 5, 4 * n).reshape(n, 4, order="F"))
    U0 = sol.conservative()  # every timed run starts from this state
    sol.adjoint(2, dt, checkpoint_every=1, dJ_dUN=w, to_host=False)  # incidence lists, workspace
    sol.state.set_array(L.GD_ARR_U_OLD, U0)
    sol.state.sync()
    t0 = time.perf_counter()
    sol.run(steps, dt=dt)
    sol.state.sync()
    fwd = time.perf_counter() - t0

    def one(budget):
        sol.state.set_array(L.GD_ARR_U_OLD, U0)
        sol.state.sync()
        t0 = time.perf_counter()
        _, _, st = sol.adjoint(steps, dt, checkpoint_every=ckpt, dJ_dUN=w, to_host=False, record_budget=budget)
        sol.state.sync()
        adj = time.perf_counter() - t0
        return {"adjoint_ms": round(adj * 1e3, 3), "cost_vs_forward": round(adj / fwd, 3),
                "adjoint_forward_sweep_ms": round(st.ms_forward, 3), "adjoint_reverse_ms": round(st.ms_reverse, 3),
                "recomputed_steps": st.recomputed_steps, "storage_GB": round(st.storage_bytes / 1e9, 3), "J": st.J}
    kept = one(int(budget_gib * 2**30))
    ckpted = one(0)
    ckpted["J_rel_diff_vs_kept"] = abs(ckpted["J"] - kept["J"]) / abs(kept["J"])
    return {"steps": steps, "checkpoint_every": ckpt, "dt": dt, "forward_ms": round(fwd * 1e3, 3),
            "record_budget_GiB": budget_gib, **kept, "checkpointed": ckpted}


def ref_tgv(n, steps, warmup, adj_steps=0, adj_ckpt=10):
    """The reference scheme, FP64 (vertex-based MUSCL with the reference's
    Minmod_Limiter, AUSM+, the calcDiffTerm viscous term, updateVariablesSA and
    the node interpolation, as calcOneStageMUSCLcpuFast and its driver run
    them) on the periodic Taylor-Green square of n x n x 2 triangles: Re 1600,
    Ma 0.1, seeded N(0, 1e-6) density noise, local dt at CFL 0.5.  This is the
    reference's CPU case scaled up; the device layout is in Hilbert order."""
    t0 = time.perf_counter()
    mesh = M.make_tgv_periodic_mesh(n)
    x, y = mesh.cell_mid_points[:, 0], mesh.cell_mid_points[:, 1]
    g, ma = 1.4, 0.1
    rho = 1.0 + 1e-6 * M.normal(SEED, mesh.nCells)
    u, v = np.sin(x) * np.cos(y), -np.cos(x) * np.sin(y)
    p = 1.0 / (g * ma * ma) + (np.cos(2 * x) + np.cos(2 * y)) / 4.0
    prm = S.make_params(viscous=True, energy_law=L.GD_ENERGY_CONST, mu_const=1.0 / 1600, k_const=1.0 / 0.72,
                        RGAS=1.0, Gamma=g, bcs={})
    sol = S.GodunovSolver(mesh, prm, order=L.GD_ORDER_HILBERT)
    sol.set_primitives(rho, u, v, p)
    init = (mesh, rho, u, v, p)
    log(f"reference scheme, TGV {n}x{n}x2: {mesh.nCells} cells, {mesh.nNodes} nodes "
        f"({time.perf_counter() - t0:.1f} s to set up)")
    if warmup:
        sol.run(warmup, cfl=0.5, dt_policy=L.GD_DT_LOCAL)
    sol.state.sync()
    t0 = time.perf_counter()
    sol.run(steps, cfl=0.5, dt_policy=L.GD_DT_LOCAL)
    sol.state.sync()
    el = time.perf_counter() - t0
    rs = sol.run(steps, cfl=0.5, dt_policy=L.GD_DT_LOCAL, time_kernels=True)
    kt = L.kernel_table(rs)
    sol.state.sync()
    adjoint = ref_adjoint_cost(sol, mesh, adj_steps, adj_ckpt, rs.last_dt) if adj_steps > 0 else None
    sol.close()
    loop = [k for k in kt if k["launches"] == steps]  # ref_phi runs once per call, before the loop
    ms = 1e3 * el / steps
    b = sum(k["bytes_per_launch"] for k in loop)
    dom, roof = roofline(loop, steps)
    for k in loop:
        log(f"   {k['name']:16s} {k['ms_avg'] * 1e3:9.2f} us/launch  "
            f"{k['bytes_per_launch'] / (k['ms_avg'] * 1e-3) / 1e9:8.1f} GB/s")
    log(f"reference scheme: {ms:.4f} ms/step, {mesh.nCells / (ms * 1e-3):.4e} cell-updates/s, "
        f"step {b / (ms * 1e-3) / 1e9:.0f} GB/s")
    if adjoint:
        log(f"reference-scheme adjoint: {adjoint['steps']} steps, checkpoints every {adjoint['checkpoint_every']}: "
            f"{adjoint['adjoint_ms']:.1f} ms vs forward {adjoint['forward_ms']:.1f} ms -> "
            f"{adjoint['cost_vs_forward']:.2f}x with every state kept ({adjoint['storage_GB']:.1f} GB; "
            f"forward sweep {adjoint['adjoint_forward_sweep_ms']:.1f} ms, reverse "
            f"{adjoint['adjoint_reverse_ms']:.1f} ms), {adjoint['checkpointed']['cost_vs_forward']:.2f}x "
            f"checkpointed ({adjoint['checkpointed']['recomputed_steps']} steps recomputed)")
    return init, {"workload": f"periodic Taylor-Green square, {n}x{n}x2 = {mesh.nCells} triangles, reference scheme "
                        f"(MUSCL + Minmod_Limiter + AUSM+ + calcDiffTerm), FP64, CFL 0.5, Hilbert device order",
            "cells": mesh.nCells, "steps": steps, "warmup": warmup, "ms_per_step": round(ms, 5),
            "value": round(mesh.nCells / (ms * 1e-3), 1), "unit": "cell-updates/s",
            "kernels_us": {k["name"]: round(k["ms_avg"] * 1e3, 2) for k in loop},
            "kernel_bytes": {k["name"]: k["bytes_per_launch"] for k in loop},
            "step_hbm": {"bytes_per_step": b, "achieved_GBs": round(b / (ms * 1e-3) / 1e9, 1),
                         "frac": round(b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
            "roofline": {**roof, "kernel": dom["name"], "bytes_per_launch": dom["bytes_per_launch"],
                         "avg_launch_us": round(dom["ms_avg"] * 1e3, 2)},
            "adjoint": adjoint}


def roofline(kt, steps):
    dom = max(kt, key=lambda k: k["ms_total"])
    achieved = dom["bytes_per_launch"] / (dom["ms_avg"] * 1e-3) / 1e9
    return dom, {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None}


def measured_traffic(kernel_name, n_cells, device_order):
    """HBM bytes per launch of `kernel_name` from the committed rocprofv3 PMC
    summary, with the build it was measured on, or (None, None) when that
    summary is of another mesh size or device order (traffic depends on both)."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        d = json.load(open(path))
    except Exception:
        return None, None
    if d.get("cells") != n_cells or d.get("device_order") != device_order:
        return None, None
    return d.get(kernel_name, {}).get("bytes_per_launch"), d.get("commit")


def host_info():
    """The host the CPU baselines ran on: logical CPUs, the CPUs this process
    may use, the OpenMP threads it was given, and the CPU model."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)),
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS"), "cpu_model": model}


def ref_cpu_baseline(init, threads, warm=3, steps=20):
    """The reference's own CPU time-loop body, as the oracle restates it
    (C++/OpenMP, the reference's static partition of distibuteCellsInThreadsSA
    over ``threads`` threads): calcDiffTerm + calcOneStageMUSCLcpuFast +
    updateVariablesSA + the node interpolation, plus the local CFL step, on the
    same TGV as the GPU reference leg; median of ``steps`` steps after
    ``warm``."""
    import oracle as O
    mesh, rho, u, v, p = init
    g = 1.4
    n = mesh.nCells
    mv = O.MeshView(mesh)
    th = M.THERMOPHYSICS(RGAS=1.0, Gamma=g)
    prm = O.params(energy_law=O.ENERGY_CONST, mu_const=1.0 / 1600, k_const=1.0 / 0.72)
    f = M.fields2d.zeros(n, mesh.nNodes)
    f.densityCells[:], f.UxCells[:], f.UyCells[:], f.pressureCells[:] = rho, u, v, p
    U = O.phs2dcns(f, g, n)
    O.update_variables(f, g, U.copy(order="F"), U, n_threads=threads)  # aSound, VMAX
    O.cells2nodes(mv, f, U, n_threads=threads)
    times, t = [], 0.0
    for s_ in range(warm + steps):
        t0 = time.perf_counter()
        Ud, _ = O.calc_diff_term(mv, f, th, prm, U, n_threads=threads)
        dt = O.local_dt(mv, f, 0.5)
        _, UN = O.calc_one_stage_muscl_cpu_fast(mv, f, th, prm, U, Ud, 1.0, dt, t, n_threads=threads)
        O.update_variables(f, g, UN, U, n_threads=threads)
        O.cells2nodes(mv, f, U, n_threads=threads)
        t += dt
        if s_ >= warm:
            times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    return {"value": round(n / med, 1), "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "host": host_info(),
            "sample": f"{steps} steps after {warm} of the reference's CPU loop body (calcDiffTerm + "
                      f"calcOneStageMUSCLcpuFast + updateVariablesSA + node interpolation, oracle C++/OpenMP "
                      f"restatement, the reference's static partition over {threads} threads) on the same "
                      f"{n}-triangle TGV, {sum(times):.1f} s in all, median step {med:.3f} s"}


def cpu_baseline(mesh, case, U0, sample_steps, threads):
    """The oracle's FP64 ns step (C++/OpenMP restatement) on the same mesh."""
    import oracle as O
    O.set_threads(threads)
    mv = O.MeshView(mesh)
    bcs = {-1: (O.BC_FIXED, (case.rho_inf, case.u_inf, 0.0, case.p_inf)), -2: (O.BC_EXTRAPOLATE, ()),
           -3: (O.BC_ISOTHERMAL_WALL, (case.T_wall,))}
    prm = O.ns_params(gamma=case.gamma, R=case.R, Pr=case.Pr, mu_ref=case.mu_inf, T_ref=case.T_inf, S=case.S,
                      venkat_K=5.0, venkat_scale=(case.rho_inf, case.u_inf, case.u_inf, case.p_inf), bcs=bcs)
    U = U0.copy(order="F")
    times = []
    for _ in range(sample_steps):
        t0 = time.perf_counter()
        dt = O.ns_dt(mv, prm, U, 0.4)
        U, _, _, _ = O.ns_step(mv, prm, dt, U)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    return {"value": round(mesh.nCells / med, 1), "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "host": host_info(),
            "sample": f"{sample_steps} FP64 ns steps of the same 2.08M-cell workload from the free stream "
                      f"(oracle C++/OpenMP restatement, {sum(times):.1f} s in all, median step {med:.3f} s)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n-theta", type=int, default=720)
    ap.add_argument("--n-r", type=int, default=1445)
    ap.add_argument("--dr0", type=float, default=2e-4)
    ap.add_argument("--order", choices=["random", "given"], default="random")
    ap.add_argument("--device-order", choices=["given", "rcm", "hilbert"], default="hilbert",
                    help="cell order of the HBM layout; variants ending in -given keep the caller's order, "
                         "-rcm use reverse Cuthill-McKee")
    ap.add_argument("--flux", choices=["ausm", "roe"], default="ausm")
    ap.add_argument("--cfl", type=float, default=0.4)
    ap.add_argument("--variants", default="fp64,mixed,fp64-rcm,mixed-rcm,fp64-given,mixed-given")
    ap.add_argument("--cpu-sample-steps", type=int, default=30, help="about 10 s of CPU work on 16 threads")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: OMP_NUM_THREADS, else every CPU available")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--adjoint-steps", type=int, default=200, help="0: skip the adjoint measurement")
    ap.add_argument("--adjoint-checkpoint", type=int, default=20)
    ap.add_argument("--adjoint-precision", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--adjoint-record-budget", type=int, default=16,
                    help="GiB of HBM for the cell records of trailing checkpoint segments")
    ap.add_argument("--ref-tgv", type=int, default=1000,
                    help="n of the n x n x 2 periodic TGV for the reference-scheme leg; 0 skips it")
    ap.add_argument("--ref-steps", type=int, default=200)
    ap.add_argument("--ref-adjoint-steps", type=int, default=100,
                    help="steps of the reference-scheme adjoint on the TGV square; 0 skips it")
    ap.add_argument("--ref-adjoint-checkpoint", type=int, default=10)
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the multi-GPU path is not built",
              file=sys.stderr)
        sys.exit(2)
    if L.device_count() < 1:
        print("no HIP device", file=sys.stderr)
        sys.exit(1)
    case = S.CylinderCase()
    mesh = build_mesh(a.n_theta, a.n_r, a.dr0, a.order)
    U0 = initial_U(case, mesh.nCells)
    flux = L.GD_FLUX_ROE if a.flux == "roe" else L.GD_FLUX_AUSM_PLUS
    orders = {"given": L.GD_ORDER_AS_GIVEN, "rcm": L.GD_ORDER_RCM, "hilbert": L.GD_ORDER_HILBERT}
    dev_order = orders[a.device_order]
    res = {}
    for var in a.variants.split(","):
        prec = L.GD_FP64 if var.startswith("fp64") else L.GD_FP32_MIXED
        order = orders[var.split("-", 1)[1]] if "-" in var else dev_order
        res[var] = run_variant(mesh, case, prec, flux, order, U0, a.warmup, a.steps, a.cfl)
        r = res[var]
        bstep = sum(k["bytes_per_launch"] for k in r["kernels"])
        log(f"{var}: {r['ms_per_step']:.4f} ms/step, {mesh.nCells / (r['ms_per_step'] * 1e-3):.4e} cell-updates/s, "
            f"step {bstep / (r['ms_per_step'] * 1e-3) / 1e9:.0f} GB/s, drag {r['drag']:.6e}, "
            f"nonphysical {r['nonphysical']}")
        for k in r["kernels"]:
            log(f"   {k['name']:16s} {k['ms_avg'] * 1e3:9.2f} us/launch  "
                f"{k['bytes_per_launch'] / (k['ms_avg'] * 1e-3) / 1e9 if k['ms_avg'] > 0 else 0:8.1f} GB/s")
    head = "fp64" if "fp64" in res else list(res)[0]
    r64 = res[head]
    value = mesh.nCells / (r64["ms_per_step"] * 1e-3)
    dom, roof = roofline(r64["kernels"], a.steps)
    roof["traffic"], roof["traffic_commit"] = measured_traffic(dom["name"], mesh.nCells, a.device_order)

    def step_frac(r):  # algorithmic bytes of one whole step over its time
        b = sum(k["bytes_per_launch"] for k in r["kernels"])
        return {"bytes_per_step": b, "achieved_GBs": round(b / (r["ms_per_step"] * 1e-3) / 1e9, 1),
                "frac": round(b / (r["ms_per_step"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "cell-updates/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(r64["ms_per_step"], 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if head == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"supersonic viscous cylinder M=3.5 Re_D=1e5, O-grid {a.n_theta}x{a.n_r} nodes, "
                               f"{mesh.nCells} triangles, {a.order} cell order, ns scheme "
                               f"({'Roe' if flux == L.GD_FLUX_ROE else 'AUSM+'}, LSQ+Venkatakrishnan, Sutherland), "
                               f"CFL {a.cfl}, {head}",
                   "cells": mesh.nCells, "nodes": mesh.nNodes, "precision": head,
                   "device_order": a.device_order},
        "roofline": {**roof, "kernel": dom["name"], "bytes_per_launch": dom["bytes_per_launch"],
                     "avg_launch_us": round(dom["ms_avg"] * 1e3, 2)},
        "kernels_us": {k["name"]: round(k["ms_avg"] * 1e3, 2) for k in r64["kernels"]},
        "ms_per_step_with_kernel_events": round(r64["ms_per_step_with_kernel_events"], 5),
        "step_hbm": step_frac(r64),
        "kernel_bytes": {k["name"]: k["bytes_per_launch"] for r in res.values() for k in r["kernels"]},
    }
    for var, r in res.items():
        if var == head:
            continue
        v = mesh.nCells / (r["ms_per_step"] * 1e-3)
        d = {"value": round(v, 1), "ms_per_step": round(r["ms_per_step"], 5),
             "speedup_vs_fp64": round(r64["ms_per_step"] / r["ms_per_step"], 3),
             "kernels_us": {k["name"]: round(k["ms_avg"] * 1e3, 2) for k in r["kernels"]}}
        if "fp64" in res:
            d["rho_rel_linf_vs_fp64"] = float(np.max(np.abs(r["rho"] - r64["rho"])) / np.max(np.abs(r64["rho"])))
            d["p_rel_linf_vs_fp64"] = float(np.max(np.abs(r["p"] - r64["p"])) / np.max(np.abs(r64["p"])))
            d["rho_rel_l2_vs_fp64"] = float(np.linalg.norm(r["rho"] - r64["rho"]) / np.linalg.norm(r64["rho"]))
            d["p_rel_l2_vs_fp64"] = float(np.linalg.norm(r["p"] - r64["p"]) / np.linalg.norm(r64["p"]))
            d["drag_rel_err_vs_fp64"] = float(abs(r["drag"] - r64["drag"]) / max(abs(r64["drag"]), 1e-300))
            d["steps_compared"] = a.warmup + a.steps
        _, rf = roofline(r["kernels"], a.steps)
        d["roofline_frac"] = rf["frac"]
        d["step_hbm"] = step_frac(r)
        out[var.replace("-", "_")] = d
    out["drag_fp64"] = r64["drag"]
    # what plain streaming kernels reach on this device (a copy and a read of
    # a 1 GiB buffer), beside the dominant kernel's and the step's rates
    cp, rd = L.hbm_stream_rates(1 << 30, 20)
    out["hbm_stream"] = {"copy_GBs": round(cp, 1), "read_GBs": round(rd, 1), "bytes": 1 << 30,
                         "dominant_kernel_frac_of_copy": round(out["roofline"]["achieved"] / cp, 4) if cp else None,
                         "step_frac_of_copy": round(out["step_hbm"]["achieved_GBs"] / cp, 4) if cp else None}
    log(f"HBM streaming on this device: copy {cp:.0f} GB/s, read {rd:.0f} GB/s (1 GiB, 16 B vectors); "
        f"{dom['name']} at {out['roofline']['achieved'] / cp:.3f} of the copy rate, "
        f"the {head} step at {out['step_hbm']['achieved_GBs'] / cp:.3f}")
    if a.adjoint_steps > 0:
        out["adjoint"] = adjoint_cost(mesh, case, dev_order, U0, a.adjoint_steps, a.adjoint_checkpoint, a.cfl,
                                      L.GD_FP64 if a.adjoint_precision == "fp64" else L.GD_FP32_MIXED,
                                      a.adjoint_record_budget)
        ad = out["adjoint"]
        log(f"adjoint: {ad['steps']} {ad['precision']} steps, checkpoints every {ad['checkpoint_every']}: "
            f"{ad['adjoint_ms']:.1f} ms vs forward {ad['forward_ms']:.1f} ms -> {ad['cost_vs_forward']:.2f}x "
            f"(forward sweep {ad['adjoint_forward_sweep_ms']:.1f} ms, reverse {ad['adjoint_reverse_ms']:.1f} ms = "
            f"{ad['reverse_step_vs_forward_step']:.2f}x per step); "
            f"one segment kept: {ad['one_segment']['adjoint_ms']:.1f} ms -> {ad['one_segment']['cost_vs_forward']:.2f}x; "
            f"all records kept: {ad['all_records']['adjoint_ms']:.1f} ms -> {ad['all_records']['cost_vs_forward']:.2f}x")
    # CPU baselines: the OpenMP threads of the host's share for this GPU
    # (OMP_NUM_THREADS where the launcher sets it, else every CPU this process
    # may run on)
    threads = a.cpu_threads or int(os.environ.get("OMP_NUM_THREADS") or 0) or len(os.sched_getaffinity(0))
    if a.ref_tgv > 0:
        init, out["ref_scheme"] = ref_tgv(a.ref_tgv, a.ref_steps, min(a.warmup, 10), a.ref_adjoint_steps,
                                          a.ref_adjoint_checkpoint)
        if not a.no_cpu:
            try:
                rc = ref_cpu_baseline(init, threads)
                out["ref_scheme"]["cpu_baseline"] = rc
                log(f"reference scheme CPU baseline: {rc['value']:.4e} cell-updates/s on {threads} threads")
            except Exception as e:  # the baseline is reported, never required
                out["ref_scheme"]["cpu_baseline"] = {"value": None, "error": str(e)}
        del init
    if not a.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(mesh, case, U0, a.cpu_sample_steps, threads)
        except Exception as e:  # the baseline is reported, never required
            out["cpu_baseline"] = {"value": None, "error": str(e)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
