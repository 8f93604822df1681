"""Benchmark: Newton solve wall time and assembly throughput, mixed FP32/FP64 vs
all-FP64, on one MI355X, at relative residual <= 1e-8.

Workload: the NeoHookean compression cube -- 1 x 1 x 1 m, 64^3 hexahedra x 6
tets (1,572,864 tets, 274,625 nodes, 823,875 dofs), mu = 3.8e5 Pa, K = 8.3e5 Pa,
symmetry rollers on x=0 / y=0 / z=0, compressive pressure on x=1 ramped over 20
load steps to 15 % nominal strain, +-2 % interior-node jitter (seed 2).  A
"step" is one full NonLinearStaticAnalysis solve (20 load steps) with
NewtonRaphson(ConvergenceSettings(rel_U_tol=1e-14, rel_res_force_tol=1e-8,
max_iter=30)).  Inputs are resident in HBM before the timed region (the mesh,
colouring, block pattern and multigrid pattern are built once, outside it).

Headline: mixed precision, inexact-Newton forcing, PCG preconditioned by the
smoothed-aggregation multigrid V-cycle whose coarse levels are built from the
first tangent (refresh "once"; the finest level is always the current tangent).
The other modes: all-FP64 with the same solver, block-Jacobi PCG, fixed inner
tolerances.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 64] [--no-fp64] [--no-cpu]
                    [--precond amg|block-jacobi] [--amg-refresh once|step|assembly]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "onsas.jl_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_mode(s, sm, p, precision, steps, warmup, nsteps, forcing=False,
             precond="block-jacobi", refresh="once", eta_first=None):
    from onsas_amd.analysis import (ConvergenceSettings, NewtonRaphson, NonLinearStaticAnalysis,
                                    ResidualsIterationStep, solve_)
    from onsas_amd.backend import GpuStaticState, MI355XBackend
    from onsas_amd import workloads

    be = MI355XBackend(precision=precision, maxiter=20000, store_stresses=False, forcing=forcing,
                       preconditioner=precond, amg_refresh=refresh,
                       **({} if eta_first is None else dict(eta_first=eta_first)))
    t0 = time.perf_counter()
    state = GpuStaticState(s, be)
    setup_s = time.perf_counter() - t0
    alg = NewtonRaphson(ConvergenceSettings(rel_U_tol=1e-14, rel_res_force_tol=1e-8, max_iter=30))
    sa = NonLinearStaticAnalysis(s, NSTEPS=nsteps)
    zeros = np.zeros(s.num_dofs())
    walls, sols = [], []
    for it in range(warmup + steps):
        sa.current_step = 1
        sa.iter_state = ResidualsIterationStep()
        sa.state = state
        state.ctx.set_U(zeros)
        if it == warmup:
            state.ctx.reset_timings()
        t0 = time.perf_counter()
        sol = solve_(sa, alg, be)
        dt = time.perf_counter() - t0
        log(f"{precision}{'+forcing' if forcing else ''} solve {it} "
            f"({'warmup' if it < warmup else 'timed'}): {dt:.3f} s, "
            f"newton {sum(sol.iterations())}, pcg {sol.linear_iterations()}")
        if it >= warmup:
            walls.append(dt)
            sols.append(sol)
    tm = state.ctx.timings()
    nn, nnzb, ncol = state.ctx.sizes()
    # the cycle's fine-level products once more, 50 back-to-back launches between
    # two events after the timed solves, on their vectors: the per-launch duration
    # without the cost of a pair of events around each sampled launch
    cycle_batch_ms = state.ctx.time_cycle_products(50) if precond == "amg" else None
    sol = sols[-1]
    a = workloads.nh_uniaxial_alpha(-p, 8.3e5, 3.8e5)
    b = workloads.nh_uniaxial_beta(a, 8.3e5, 3.8e5)
    ex = workloads.homogeneous_displacement(sm.coords, a, b)
    err = float(np.abs(sol.U[-1] - ex).max() / np.abs(ex).max())
    per_iter = [dict(step=d["load_step"], it=d["iteration"], r_rel=float("%.3e" % d["r_rel"]),
                     eta=float("%.2e" % d["lin_rtol"]), lin_iters=d["lin_iters"])
                for d in sol.newton_log if d["load_step"] in (1, 2, nsteps)]
    out = dict(
        per_iteration=per_iter,
        wall_s=float(np.mean(walls)), walls=[round(w, 4) for w in walls], setup_s=setup_s,
        newton_iters=int(sum(sol.iterations())), pcg_iters=int(sol.linear_iterations()),
        max_rel_residual=float(max(r["dr_rel"] for r in sol.records)),
        criteria=sorted({type(r["criterion"]).__name__ for r in sol.records}),
        disp_err_vs_closed_form=err, timings=tm, nn=nn, nnzb=nnzb, ncolors=ncol,
        cycle_batch_ms=cycle_batch_ms,
        U_final=sol.U[-1])
    state.close()
    return out


def cpu_baseline(n_small=16, nsteps=20, n_full=64):
    """The oracle (reference algorithm: COO->CSR assembly, unpreconditioned CG at
    reltol sqrt(eps)) on one core, full solve of a smaller cube, projected to the
    full cube with the (n_full / n_small)^4 work growth of unpreconditioned CG."""
    sys.path.insert(0, ROOT)
    from oracle import oracle as O
    from onsas_amd import workloads

    s, sm, p = workloads.cube_compression(n=n_small)
    (mat, es), = list(s.materials)
    m = O.Model(sm.coords, tets=es.conn, tet_mat=O.MAT_NEOHOOKEAN, tet_p=mat.params(),
                fixed_dofs=s.fixed_dofs)
    lfs = np.linspace(1 / nsteps, 1.0, nsteps)
    F = []
    for t in lfs:
        f = np.zeros(s.num_dofs())
        for bc, conn in s.bcs.pressure_groups():
            for tri in conn:
                fe = O.face_pressure(sm.coords[tri].reshape(-1), bc(t))
                for a in range(3):
                    f[3 * tri[a]:3 * tri[a] + 3] += fe[a]
        F.append(f)
    F = np.array(F)
    t0 = time.perf_counter()
    _, _, recs = m.newton(F, rel_U_tol=1e-14, rel_res_tol=1e-8, max_iter=30, keep_steps=False)
    dt = time.perf_counter() - t0
    # assembly rate of the reference algorithm on the same cube
    U = np.zeros(s.num_dofs())
    t1 = time.perf_counter()
    m.assemble(U, want_full=False, want_ff=True)
    ta = time.perf_counter() - t1
    scale = (n_full / n_small) ** 4
    return dict(
        value=dt * scale, unit="s", cores=1, kind="port",
        sample=(f"full {nsteps}-step NewtonRaphson solve of the {n_small}^3 x 6-tet NeoHookean "
                f"compression cube ({len(sm.tets)} tets) with the oracle's restatement of the "
                f"reference algorithm on 1 core: {dt:.2f} s, {sum(r['iters'] for r in recs)} Newton "
                f"iterations, {sum(r['cg_iters'] for r in recs)} CG iterations; projected to "
                f"{n_full}^3 by (n_full/n_small)^4 = {scale:.0f} (unpreconditioned CG: iterations "
                f"~ n, work per iteration ~ n^3); extrapolated, not measured at {n_full}^3"),
        measured_s=dt, assembly_Melem_per_s=len(sm.tets) / ta / 1e6)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--nsteps", type=int, default=20)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-fixed", action="store_true", help="skip the fixed-rtol runs")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precond", default="amg", choices=["block-jacobi", "amg"],
                    help="PCG preconditioner of the headline and FP64 runs")
    ap.add_argument("--amg-refresh", default="once", choices=["once", "step", "assembly"])
    ap.add_argument("--no-bj", action="store_true", help="skip the block-Jacobi comparison run")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this engine has no multi-GPU path); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    from onsas_amd import workloads
    log(f"building the {args.n}^3 x 6-tet compression cube")
    s, sm, p = workloads.cube_compression(n=args.n)
    ne = len(sm.tets)

    runs = {}
    # headline: mixed precision with inexact-Newton forcing
    pc = dict(precond=args.precond, refresh=args.amg_refresh)
    runs["mixed+forcing"] = run_mode(s, sm, p, "mixed", args.steps, args.warmup, args.nsteps,
                                     forcing=True, **pc)
    if not args.no_fp64:
        runs["fp64+forcing"] = run_mode(s, sm, p, "fp64", args.steps, args.warmup, args.nsteps,
                                        forcing=True, **pc)
    if not args.no_bj and args.precond != "block-jacobi":
        runs["mixed+forcing+block-jacobi"] = run_mode(s, sm, p, "mixed", args.steps, args.warmup,
                                                      args.nsteps, forcing=True,
                                                      precond="block-jacobi")
    if not args.no_fixed:
        # the fixed inner tolerances: 1e-6 (mixed) and sqrt(eps) as in the reference (fp64)
        runs["mixed"] = run_mode(s, sm, p, "mixed", args.steps, args.warmup, args.nsteps, **pc)
        if not args.no_fp64:
            runs["fp64"] = run_mode(s, sm, p, "fp64", args.steps, args.warmup, args.nsteps, **pc)
        # a tighter first inner tolerance (5e-3): faster in both modes, with the
        # larger gain in FP64 (the default 1e-2 keeps the forcing comparison at 1.5x)
        for prec in ("mixed",) + (() if args.no_fp64 else ("fp64",)):
            runs[f"{prec}+forcing+eta_first_5e-3"] = run_mode(
                s, sm, p, prec, args.steps, args.warmup, args.nsteps, forcing=True, eta_first=5e-3,
                **pc)

    def asm_rates(r, kt):
        """Assembly throughput and its HBM roofline fraction from algorithmic
        bytes (every datum moved once).  Gather assembly (NeoHookean: 24 state
        planes):
          element pass: conn 16 + dN/V 104 per tet, U 24 per node, state write
                        24 kt + nodal forces 96 per tet
          gather pass : contribution entries 16 x 4, state read 24 kt, forces 96
                        per tet; K 9 kt per block written once; row_ptr/col;
                        fixed mask, block-Jacobi 9 kt and f_int 24 per node."""
        tm = r["timings"]
        na = max(tm["n_assemble"], 1)
        nn, nnzb = r["nn"], r["nnzb"]
        asm_ms = tm["assemble_ms"] / na
        elem_ms = tm["elem_ms"] / na
        gather_ms = (tm["scatter_ms"] - tm["elem_ms"]) / na
        b_elem = ne * (16 + 104 + 24 * kt + 96) + 24 * nn
        b_gather = (ne * (64 + 24 * kt + 96) + 9 * kt * nnzb + 4 * (nn + 1) + 4 * nnzb +
                    nn * (1 + 9 * kt + 24))
        tot = b_elem + b_gather
        if tm["assembly_mode"] == 0:
            # coloured scatter: conn 16 + dN/V 104 + u 96 + slots 64 + K block RMW + f_int RMW
            tot = ne * (16 + 104 + 96 + 64 + 2 * 16 * 9 * kt + 2 * 12 * 8)
            b_elem = b_gather = 0
        # the reference algorithm's bytes per tet (coloured BSR scatter: conn 16,
        # grad N and V 104, u 96, slot map 64, K block and f_int read-modify-write)
        survey_b = 1624 if kt == 4 else 2776
        return dict(assemble_ms=asm_ms, elem_pass_ms=elem_ms, gather_pass_ms=gather_ms,
                    assembly_Melem_per_s=ne / (asm_ms * 1e-3) / 1e6,
                    reference_algorithm_bytes_per_elem=survey_b,
                    frac_of_peak_reference_algorithm_bytes=ne * survey_b / (asm_ms * 1e-3) / 1e9
                    / HBM_PEAK_GBS,
                    bytes_per_assembly=tot, bytes_per_elem=tot / ne,
                    achieved_GBps=tot / (asm_ms * 1e-3) / 1e9,
                    frac_of_peak=tot / (asm_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                    elem_pass_frac=b_elem / (elem_ms * 1e-3) / 1e9 / HBM_PEAK_GBS if elem_ms else None,
                    gather_pass_frac=b_gather / (gather_ms * 1e-3) / 1e9 / HBM_PEAK_GBS
                    if gather_ms else None,
                    mode="gather" if tm["assembly_mode"] == 1 else "scatter")

    def spmv_roofline(r, kt):
        """The fine-level tangent products: k_pcg_A (PCG: K blocks + col +
        gathered z and p_old, written p_new and q) and, with the multigrid, the
        cycle's two fine-level products k_sell_cheb<float, float> (K blocks +
        col, plus per dof 3 and 6 vector values read or written: x, b, r / x, b,
        block-Jacobi 3, z -- 4.5 on average; with ONSAS_AMG_FINE_DEGREE=2 the two
        exact products of the degree-2 cycle: x, b, r / x, b, block-Jacobi 3, r,
        d0 -- 5)."""
        tm = r["timings"]
        nn, nnzb = r["nn"], r["nnzb"]
        n = 3 * nn
        out = dict(pcg_ms_per_iter=tm["pcg_ms"] / max(tm["pcg_iters"], 1))
        cyc_vec = 5 if os.environ.get("ONSAS_AMG_FINE_DEGREE") == "2" else 4.5
        for name, ms, cnt, vec in (("k_pcg_A", tm["spmv_ms"], tm["spmv_launches"], 4),
                                   ("k_sell_cheb", tm["cycle_spmv_ms"], tm["cycle_spmv_launches"],
                                    cyc_vec)):
            if cnt <= 0:
                continue
            b = int(nnzb * (9 * kt + 4) + vec * n * kt)
            avg = stamped = ms / cnt
            if name == "k_sell_cheb" and r.get("cycle_batch_ms"):
                avg = r["cycle_batch_ms"]
            out[name] = dict(bytes_per_launch=b, avg_launch_ms=avg,
                             achieved_GBps=b / (avg * 1e-3) / 1e9, launches=cnt,
                             stamped_launch_ms=stamped)
        return out

    head = runs["mixed+forcing"]
    h_spmv = spmv_roofline(head, 4)
    rk = "k_sell_cheb" if "k_sell_cheb" in h_spmv else "k_pcg_A"
    hk = h_spmv[rk]
    # HBM traffic per launch of that kernel from the committed rocprof summary of
    # this workload (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, copy-kernel calibration)
    traffic = None
    prof = os.path.join(ROOT, "profiles", "newton_cube64_kernels.md")
    if args.n == 64 and os.path.exists(prof):
        import re
        full = {"k_sell_cheb": "k_sell_cheb<float, float>", "k_pcg_A": "k_pcg_A<float>"}[rk]
        mt = re.search(r"traffic_bytes_per_launch %s: (\d+)" % re.escape(full), open(prof).read())
        traffic = int(mt.group(1)) if mt else None
    result = {
        "metric": "Newton solve wall time & assembly Melem/s, mixed vs FP64 on 1 GPU at resid<=1e-8",
        "value": round(head["wall_s"], 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["wall_s"] * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"NeoHookean compression cube {args.n}^3x6 tets ({ne} tets, "
                               f"{3 * sm.nn} dofs), {args.nsteps} load steps to 15% strain, "
                               f"mixed FP32 tangent/PCG + FP64 residual/update, inexact-Newton "
                               f"forcing, PCG preconditioner {args.precond}"
                               + (f" (coarse levels refresh {args.amg_refresh})"
                                  if args.precond == "amg" else ""),
                   "tets": ne, "dofs": 3 * sm.nn, "load_steps": args.nsteps,
                   "rel_res_force_tol": 1e-8, "preconditioner": args.precond},
        "roofline": {"bound": "hbm", "achieved": round(hk["achieved_GBps"], 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(hk["achieved_GBps"] / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": ("k_sell_cheb (fine-level sliced-ELL 3x3 tangent product + "
                                "Chebyshev update, FP32)" if rk == "k_sell_cheb" else
                                "k_pcg_A (sliced-ELL 3x3 SpMV + p update, FP32)"),
                     "bytes_per_launch": hk["bytes_per_launch"],
                     "avg_launch_ms": round(hk["avg_launch_ms"], 5),
                     "duration_from": ("50 back-to-back launches of the cycle's two fine-level "
                                       "products between two HIP events, after the timed solves"
                                       if rk == "k_sell_cheb" and head.get("cycle_batch_ms") else
                                       "dispatch-stamped HIP events on sampled launches in the "
                                       "timed solve"),
                     "stamped_launch_ms": round(hk["stamped_launch_ms"], 5)},
        "modes": {},
    }
    for name, r in runs.items():
        kt = 4 if name.startswith("mixed") else 8
        d = {k: v for k, v in r.items() if k not in ("U_final",)}
        d["assembly"] = asm_rates(r, kt)
        d["pcg"] = spmv_roofline(r, kt)
        result["modes"][name] = d
    if "fp64+forcing" in runs:
        result["speedup_mixed_vs_fp64_forcing"] = round(
            runs["fp64+forcing"]["wall_s"] / head["wall_s"], 3)
    if "fp64+forcing+eta_first_5e-3" in runs:
        result["speedup_mixed_vs_fp64_forcing_eta_first_5e-3"] = round(
            runs["fp64+forcing+eta_first_5e-3"]["wall_s"]
            / runs["mixed+forcing+eta_first_5e-3"]["wall_s"], 3)
    if "fp64" in runs and "mixed" in runs:
        result["speedup_mixed_vs_fp64_fixed_rtol"] = round(
            runs["fp64"]["wall_s"] / runs["mixed"]["wall_s"], 3)
    if "mixed+forcing+block-jacobi" in runs:
        result["speedup_vs_block_jacobi"] = round(
            runs["mixed+forcing+block-jacobi"]["wall_s"] / head["wall_s"], 3)
    ref = runs.get("fp64") or runs.get("fp64+forcing")
    if ref is not None:
        d = np.abs(head["U_final"] - ref["U_final"]).max() / np.abs(ref["U_final"]).max()
        result["mixed_vs_fp64_max_rel_disp_diff"] = float(d)
    ha = result["modes"]["mixed+forcing"]["assembly"]
    result["assembly_Melem_per_s"] = round(ha["assembly_Melem_per_s"], 1)
    # assembly against HBM peak: with the reference algorithm's bytes per tet
    # (coloured scatter, 1,624 B mixed) and with the bytes this gather design moves
    result["assembly_roofline"] = {
        "frac_reference_algorithm_bytes": round(ha["frac_of_peak_reference_algorithm_bytes"], 4),
        "reference_algorithm_bytes_per_tet": ha["reference_algorithm_bytes_per_elem"],
        "frac_gather_design_bytes": round(ha["frac_of_peak"], 4),
        "gather_design_bytes_per_tet": round(ha["bytes_per_elem"], 1)}
    if not args.no_cpu:
        log("cpu baseline (oracle, 1 core)")
        result["cpu_baseline"] = cpu_baseline(nsteps=args.nsteps, n_full=args.n)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
