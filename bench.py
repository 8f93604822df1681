"""Benchmark: Newton solve wall time and assembly throughput, mixed FP32/FP64 vs
all-FP64, on one MI355X, at relative residual <= 1e-8.

Workload: the NeoHookean compression cube -- 1 x 1 x 1 m, 64^3 hexahedra x 6
tets (1,572,864 tets, 274,625 nodes, 823,875 dofs), mu = 3.8e5 Pa, K = 8.3e5 Pa,
symmetry rollers on x=0 / y=0 / z=0, compressive pressure on x=1 ramped over 20
load steps to 15 % nominal strain, +-2 % interior-node jitter (seed 2).  A
"step" is one full NonLinearStaticAnalysis solve (20 load steps) with
NewtonRaphson(ConvergenceSettings(rel_U_tol=1e-14, rel_res_force_tol=1e-8,
max_iter=30)).  Inputs are resident in HBM before the timed region (the mesh,
block pattern and multigrid pattern are built once, outside it).

Headline: mixed precision, inexact-Newton forcing, PCG preconditioned by the
smoothed-aggregation multigrid V-cycle whose coarse levels are built from the
first tangent (refresh "once"; the finest level is always the current tangent).
Beside it: the all-FP64 mode (every stored value FP64, the multigrid's too)
with the same solver, both swept over the first-iteration inner tolerance so
that the mixed/FP64 ratio compares each mode at its fastest setting; fixed inner
tolerances (the reference's sqrt(eps) in FP64); block-Jacobi PCG.

Output: the LAST stdout line is one compact JSON object (the headline,
roofline of the dominant kernel, CPU baseline, mixed/FP64 speedup); the per-mode
detail goes to stderr as one JSON line prefixed "[bench] detail " and to
gpurun_out/bench_detail.json when that folder can be written.

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
ETA_SWEEP = (1e-2, 5e-3, 3e-3)  # first-iteration inner tolerances compared per mode


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def mode_label(precision, forcing, precond, eta_first=None):
    lab = precision + ("+forcing" if forcing else "+fixed-rtol")
    if precond == "block-jacobi":
        lab += "+block-jacobi"
    if forcing and eta_first is not None:
        lab += f"+eta_first={eta_first:g}"
    return lab


def run_mode(s, sm, p, precision, steps, warmup, nsteps, forcing=False,
             precond="block-jacobi", refresh="once", eta_first=None, kernels=False):
    """warmup + steps solves on one context; the first one is the cold solve
    (multigrid numerics, graph captures).  kernels: also time the multigrid
    levels' products in event-timed batches after the timed solves."""
    from onsas_amd.analysis import (ConvergenceSettings, NewtonRaphson, NonLinearStaticAnalysis,
                                    ResidualsIterationStep, solve_)
    from onsas_amd.backend import GpuStaticState, MI355XBackend
    from onsas_amd import workloads

    label = mode_label(precision, forcing, precond, eta_first)
    be = MI355XBackend(precision=precision, maxiter=20000, store_stresses=False, forcing=forcing,
                       preconditioner=precond, amg_refresh=refresh,
                       **({} if eta_first is None else dict(eta_first=eta_first)))
    t0 = time.perf_counter()
    state = GpuStaticState(s, be)
    setup_s = time.perf_counter() - t0
    alg = NewtonRaphson(ConvergenceSettings(rel_U_tol=1e-14, rel_res_force_tol=1e-8, max_iter=30))
    sa = NonLinearStaticAnalysis(s, NSTEPS=nsteps)
    zeros = np.zeros(s.num_dofs())
    walls, sols, cold_s = [], [], None
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
        if it == 0:
            cold_s = dt
        log(f"{label} solve {it} ({'warmup' if it < warmup else 'timed'}): {dt:.3f} s, "
            f"newton {sum(sol.iterations())}, pcg {sol.linear_iterations()}")
        if it >= warmup:
            walls.append(dt)
            sols.append(sol)
    tm = state.ctx.timings()
    nn, nnzb, ncol = state.ctx.sizes()
    layout = state.ctx.gather_layout()
    levels = []
    if kernels and precond == "amg":
        # the levels' cycle products once more, back to back between two events
        # after the timed solves, on their vectors: the per-launch duration
        # without the cost of a pair of events around each sampled launch
        nl = state.ctx.amg_levels()
        for l in range(nl):
            info = state.ctx.amg_level_info(l)
            ent = dict(level=l, n=info["n"], bs=info["bs"], nnzb=info["nnzb"])
            if l + 1 < nl:
                try:
                    ent["batch_ms_per_launch"] = state.ctx.time_level_products(
                        l, 100 if l else 50)
                except RuntimeError as e:  # a level without a sliced-ELL operator
                    ent["batch_error"] = str(e)
            levels.append(ent)
    sol = sols[-1]
    a = workloads.nh_uniaxial_alpha(-p, 8.3e5, 3.8e5)
    b = workloads.nh_uniaxial_beta(a, 8.3e5, 3.8e5)
    ex = workloads.homogeneous_displacement(sm.coords, a, b)
    err = float(np.abs(sol.U[-1] - ex).max() / np.abs(ex).max())
    per_iter = [dict(step=d["load_step"], it=d["iteration"], r_rel=float("%.3e" % d["r_rel"]),
                     eta=float("%.2e" % d["lin_rtol"]), lin_iters=d["lin_iters"])
                for d in sol.newton_log if d["load_step"] in (1, 2, nsteps)]
    out = dict(
        label=label, precision=precision, forcing=forcing, precond=precond,
        eta_first=be.eta_first if forcing else None,
        per_iteration=per_iter,
        wall_s=float(np.mean(walls)), walls=[round(w, 4) for w in walls], setup_s=setup_s,
        cold_solve_s=cold_s,
        newton_iters=int(sum(sol.iterations())), pcg_iters=int(sol.linear_iterations()),
        max_rel_residual=float(max(r["dr_rel"] for r in sol.records)),
        criteria=sorted({type(r["criterion"]).__name__ for r in sol.records}),
        disp_err_vs_closed_form=err, timings=tm, nn=nn, nnzb=nnzb, ncolors=ncol, layout=layout,
        levels=levels, U_final=sol.U[-1])
    state.close()
    return out


def asm_rates(r, kt, ne):
    """Assembly throughput and its HBM roofline fraction.  Gather assembly
    (NeoHookean: 24 state planes), every datum moved once:
      element pass: conn 16 per tet, U and the reference coordinates 24 + 24
                    per node (the gradients are recomputed from them), state
                    write 24 kt + nodal forces 96 per tet
      gather pass : contribution entries 16 x 4, state read 24 kt, forces 96
                    per tet; K 9 kt per block written once; row_ptr/col;
                    fixed mask, block-Jacobi 9 kt and f_int 24 per node.
                    Staged (ONSAS_GATHER=staged): the stored 16-bit entries (with
                    padding) instead of 16 x 4, the element lists 4 per staged
                    element, the step table 16 per step and 24 per workgroup,
                    the diagonal index 4 per node; the states once.
    Beside it the reference algorithm's bytes (coloured BSR scatter: conn 16,
    grad N and V 104, u 96, slot map 64, K block and f_int read-modify-write:
    1,624 B/tet mixed, 2,776 B/tet FP64)."""
    tm = r["timings"]
    na = max(tm["n_assemble"], 1)
    nn, nnzb = r["nn"], r["nnzb"]
    asm_ms = tm["assemble_ms"] / na
    elem_ms = tm["elem_ms"] / na
    gather_ms = (tm["scatter_ms"] - tm["elem_ms"]) / na
    b_elem = ne * (16 + 24 * kt + 96) + 48 * nn
    b_gather = (ne * (64 + 24 * kt + 96) + 9 * kt * nnzb + 4 * (nn + 1) + 4 * nnzb +
                nn * (1 + 9 * kt + 24))
    lay = r.get("layout") or {}
    if tm["assembly_mode"] == 2:
        b_gather = (ne * (24 * kt + 96) + 2 * lay["entries"] + 4 * lay["staged"] +
                    16 * lay["steps"] + 24 * lay["groups"] + 9 * kt * nnzb + 4 * (nn + 1) +
                    4 * nnzb + nn * (4 + 1 + 9 * kt + 24))
    tot = b_elem + b_gather
    if tm["assembly_mode"] == 0:
        tot = ne * (16 + 104 + 96 + 64 + 2 * 16 * 9 * kt + 2 * 12 * 8)
        b_elem = b_gather = 0
    survey_b = 1624 if kt == 4 else 2776
    gbs = lambda b, ms: b / (ms * 1e-3) / 1e9  # noqa: E731
    return dict(assemble_ms=asm_ms, elem_pass_ms=elem_ms, gather_pass_ms=gather_ms,
                assembly_Melem_per_s=ne / (asm_ms * 1e-3) / 1e6,
                reference_algorithm_bytes_per_elem=survey_b,
                frac_reference_algorithm_bytes=gbs(ne * survey_b, asm_ms) / HBM_PEAK_GBS,
                design_bytes_per_elem=tot / ne,
                frac_design_bytes=gbs(tot, asm_ms) / HBM_PEAK_GBS,
                elem_pass_frac=gbs(b_elem, elem_ms) / HBM_PEAK_GBS if elem_ms else None,
                gather_pass_frac=gbs(b_gather, gather_ms) / HBM_PEAK_GBS if gather_ms else None,
                gather_pass_bytes=b_gather,
                mode={2: "staged gather", 1: "gather"}.get(tm["assembly_mode"], "scatter"))


def kernel_rooflines(r, kt):
    """Algorithmic bytes and event-timed duration per launch of the cycle's
    operator products, and their share of a PCG iteration.
      k_sell_cheb (fine level, 2 launches per iteration with the degree-1 fine
        smoother): nnzb (9 kt + 4) + 4.5 n kt (x gathered, b, r / x, b,
        block-Jacobi 3, z: 3 and 6 values per dof, mean 4.5)
      k_sell6_cheb (level 1, sliced ELL, 4 launches per iteration with the
        degree-2 smoother): nnzb_1 (36 vb + 4) + 49.5 kt per row (vb = 2, BF16
        copy, in mixed mode; 8 in FP64 mode; per row x gathered 6, b 6, D^-1
        36, outputs 6 or 12: 60, 18, 60 and 60 values for the four products);
        with the degree-1 smoother 2 launches, 18 and 54 values per row."""
    out = {}
    lv = {e["level"]: e for e in r.get("levels", [])}
    if 0 in lv and "batch_ms_per_launch" in lv[0]:
        nn, nnzb = r["nn"], r["nnzb"]
        b = int(nnzb * (9 * kt + 4) + 4.5 * 3 * nn * kt)
        ms = lv[0]["batch_ms_per_launch"]
        out["k_sell_cheb"] = dict(
            kernel=f"k_sell_cheb<{'float' if kt == 4 else 'double'}, "
                   f"{'float' if kt == 4 else 'double'}>",
            what="fine-level sliced-ELL 3x3 tangent product + smoothing epilogue",
            bytes_per_launch=b, ms_per_launch=ms, launches_per_pcg_iter=2,
            GBps=b / (ms * 1e-3) / 1e9)
    if 1 in lv and "batch_ms_per_launch" in lv[1]:
        e = lv[1]
        vb = 2 if kt == 4 else 8
        deg1 = os.environ.get("ONSAS_AMG_COARSE_DEGREE", "2") == "1"
        b = int(e["nnzb"] * (36 * vb + 4) + (36.0 if deg1 else 49.5) * kt * e["n"])
        ms = e["batch_ms_per_launch"]
        rpw = os.environ.get("ONSAS_SELL6_RPW", "8")
        ur = os.environ.get("ONSAS_SELL6_ROUNDS", "2" if kt == 4 else "1")
        wv = os.environ.get("ONSAS_SELL6_WAVES", "8") if kt == 4 else "4"
        out["k_sell6_cheb"] = dict(
            kernel=f"k_sell6_cheb<{'float, unsigned short' if kt == 4 else 'double, double'}, "
                   f"{rpw}, {ur}, {wv}>",
            what="level-1 sliced-ELL 6x6 operator product + Chebyshev epilogue",
            bytes_per_launch=b, ms_per_launch=ms, launches_per_pcg_iter=2 if deg1 else 4,
            GBps=b / (ms * 1e-3) / 1e9)
    for k in out.values():
        k["ms_per_pcg_iter"] = k["ms_per_launch"] * k["launches_per_pcg_iter"]
        k["frac"] = k["GBps"] / HBM_PEAK_GBS
    return out


def read_traffic(kernel_name):
    """HBM bytes per launch of a kernel from the committed rocprof summary of
    this workload (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, copy-kernel
    calibration), or None."""
    import re
    prof = os.path.join(ROOT, "profiles", "newton_cube64_kernels.md")
    if not os.path.exists(prof):
        return None
    mt = re.search(r"traffic_bytes_per_launch %s: (\d+)" % re.escape(kernel_name),
                   open(prof).read())
    return int(mt.group(1)) if mt else None


# CG iterations of the oracle's full Newton solves of this workload (20 load
# steps, ConvergenceSettings(1e-14, 1e-8, 30), CG at reltol sqrt(eps)): the
# counts are properties of the algorithm, not of the host, and were measured
# with oracle/baseline_counts.py (16^3 and 32^3 on one core, 64^3 with the
# OpenMP build on 7 threads, whose reduction order moves a count by about 0.1 %).
# Every solve takes 80 Newton iterations; the CG count grows as n^0.88.
ORACLE_CG_ITERS = {16: 30635, 32: 55683, 64: 103470}
ORACLE_NEWTON_ITERS = 80


def cpu_baseline(n_full=64, cg_sample=20):
    """The reference algorithm restated by the oracle (per Newton iteration: the
    element loop and a COO->CSR assembly of the free-dof tangent; CG without a
    preconditioner at reltol sqrt(eps)), timed on this host on a bounded sample
    of the n_full^3 workload: one assembly at the first Newton iteration and
    the better of two runs of cg_sample CG iterations on its system, on one
    core and with the OpenMP build on every core of this process.  The full
    solve's time is those
    per-unit costs times the oracle's measured iteration counts
    (ORACLE_CG_ITERS, 80 Newton iterations): extrapolated from measured parts,
    and labelled so."""
    sys.path.insert(0, ROOT)
    from oracle import oracle as O
    from onsas_amd import workloads

    s, sm, p = workloads.cube_compression(n=n_full)
    (mat, es), = list(s.materials)
    f1 = np.zeros(s.num_dofs())
    for bc, conn in s.bcs.pressure_groups():
        for tri in conn:
            fe = O.face_pressure(sm.coords[tri].reshape(-1), bc(1.0 / 20))
            for a in range(3):
                f1[3 * tri[a]:3 * tri[a] + 3] += fe[a]
    cores = len(os.sched_getaffinity(0))
    out = {}
    for omp in (False, True):
        m = O.Model(sm.coords, tets=es.conn, tet_mat=O.MAT_NEOHOOKEAN, tet_p=mat.params(),
                    fixed_dofs=s.fixed_dofs, omp=omp)
        U = np.zeros(s.num_dofs())
        t0 = time.perf_counter()
        Kff = m.assemble(U, want_full=False, want_ff=True)["Kff"]
        t_asm = time.perf_counter() - t0
        t_cg = float("inf")
        for _ in range(2):  # the better of two samples: the first can pay page faults
            t0 = time.perf_counter()
            O.cg(Kff, f1[m.free], reltol=0.0, maxiter=cg_sample, omp=omp)
            t_cg = min(t_cg, (time.perf_counter() - t0) / cg_sample)
        out["omp" if omp else "serial"] = (t_asm, t_cg)
        del Kff, m
        log(f"cpu baseline ({'all cores' if omp else '1 core'}): assembly {t_asm:.2f} s, "
            f"CG {t_cg * 1e3:.1f} ms per iteration")
    n_cg = ORACLE_CG_ITERS.get(n_full)
    res = dict(unit="s", kind="port")
    if n_cg is None:
        return dict(res, value=None, cores=1, sample=f"no measured CG count for {n_full}^3")
    ta, tc = out["serial"]
    ta_all, tc_all = out["omp"]
    res.update(
        value=round(ORACLE_NEWTON_ITERS * ta + n_cg * tc, 1), cores=1,
        value_all_cores=round(ORACLE_NEWTON_ITERS * ta_all + n_cg * tc_all, 1),
        all_cores=cores,
        sample=(f"oracle restatement of the reference algorithm on the {n_full}^3x6-tet cube: "
                f"1 assembly + {cg_sample} CG iterations of the first Newton system, timed "
                f"(1 core: {ta:.2f} s + {tc * 1e3:.1f} ms/it; {cores} cores: {ta_all:.2f} s + "
                f"{tc_all * 1e3:.1f} ms/it); EXTRAPOLATED to the full solve by its measured "
                f"counts ({ORACLE_NEWTON_ITERS} Newton, {n_cg} CG iterations)"))
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--nsteps", type=int, default=20)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-fixed", action="store_true", help="skip the fixed-rtol runs")
    ap.add_argument("--no-sweep", action="store_true",
                    help="skip the eta_first sweep (headline setting only)")
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
    from onsas_amd.backend import MI355XBackend
    log(f"building the {args.n}^3 x 6-tet compression cube")
    s, sm, p = workloads.cube_compression(n=args.n)
    ne = len(sm.tets)
    eta0 = MI355XBackend().eta_first
    pc = dict(precond=args.precond, refresh=args.amg_refresh)
    side = dict(steps=min(args.steps, 3), warmup=1)  # comparison runs: fewer solves

    runs = {}
    head = run_mode(s, sm, p, "mixed", args.steps, args.warmup, args.nsteps, forcing=True,
                    kernels=True, **pc)
    runs[head["label"]] = head
    etas = [eta0] + ([] if args.no_sweep else [e for e in ETA_SWEEP if e != eta0])
    for prec in ("mixed",) + (() if args.no_fp64 else ("fp64",)):
        for eta in etas:
            if prec == "mixed" and eta == eta0:
                continue
            r = run_mode(s, sm, p, prec, args.steps, args.warmup, args.nsteps, forcing=True,
                         eta_first=eta, kernels=(prec == "fp64" and eta == eta0), **pc)
            runs[r["label"]] = r
    if not args.no_fixed:
        # fixed inner tolerances: 1e-6 (mixed) and sqrt(eps) as in the reference (fp64)
        for prec in ("mixed",) + (() if args.no_fp64 else ("fp64",)):
            r = run_mode(s, sm, p, prec, side["steps"], side["warmup"], args.nsteps, **pc)
            runs[r["label"]] = r
    if not args.no_bj and args.precond != "block-jacobi":
        r = run_mode(s, sm, p, "mixed", side["steps"], side["warmup"], args.nsteps, forcing=True,
                     precond="block-jacobi")
        runs[r["label"]] = r

    def best(prec):
        cand = [r for r in runs.values() if r["precision"] == prec and r["forcing"]
                and r["precond"] == args.precond]
        return min(cand, key=lambda r: r["wall_s"]) if cand else None

    bm, bf = best("mixed"), best("fp64")
    head_k = kernel_rooflines(head, 4)
    # the dominant kernel: the largest share of a PCG iteration
    dom_name = max(head_k, key=lambda k: head_k[k]["ms_per_pcg_iter"]) if head_k else None
    roofline = None
    if dom_name:
        d = head_k[dom_name]
        roofline = {"bound": "hbm", "achieved": round(d["GBps"], 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(d["frac"], 4),
                    "traffic": read_traffic(d["kernel"]) if args.n == 64 else None,
                    "kernel": d["kernel"], "bytes_per_launch": d["bytes_per_launch"],
                    "ms_per_launch": round(d["ms_per_launch"], 5),
                    "duration_from": "event-timed batch of back-to-back launches on the solve "
                                     "stream after the timed solves"}
    asm = asm_rates(head, 4, ne)
    tm = head["timings"]
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
        "dtype": "f32 tangent/PCG + f64 residual/update; multigrid P fp16, level-1 bf16",
        "data": "synthetic",
        "config": {"workload": f"NeoHookean compression cube {args.n}^3x6 tets ({ne} tets, "
                               f"{3 * sm.nn} dofs), {args.nsteps} load steps to 15% strain, "
                               f"inexact Newton, multigrid PCG (eta_first={eta0:g})",
                   "tets": ne, "dofs": 3 * sm.nn, "rel_res_force_tol": 1e-8},
        "roofline": roofline,
        "speedup_mixed_vs_fp64": (round(bf["wall_s"] / bm["wall_s"], 3) if bm and bf else None),
        "speedup_basis": (f"each mode at its fastest eta_first: mixed {bm['eta_first']:g} "
                          f"{bm['wall_s']:.4f} s, fp64 {bf['eta_first']:g} {bf['wall_s']:.4f} s"
                          if bm and bf else None),
        "cold_solve_s": round(head["cold_solve_s"], 4),
        "setup_s": round(head["setup_s"], 3),
        "pcg_iters": head["pcg_iters"],
        "pcg_ms_per_iter": round(tm["pcg_ms"] / max(tm["pcg_iters"], 1), 4),
        "disp_err_vs_closed_form": float("%.3e" % head["disp_err_vs_closed_form"]),
        "assembly_Melem_per_s": round(asm["assembly_Melem_per_s"], 1),
        "assembly_roofline": {"frac_reference_algorithm_bytes":
                              round(asm["frac_reference_algorithm_bytes"], 4),
                              "frac_design_bytes": round(asm["frac_design_bytes"], 4),
                              "design_bytes_per_tet": round(asm["design_bytes_per_elem"], 1),
                              "gather_pass_frac": round(asm["gather_pass_frac"] or 0, 4),
                              "gather_pass_ms": round(asm["gather_pass_ms"], 4),
                              "mode": asm["mode"]},
    }
    fp = runs.get(mode_label("fp64", False, args.precond)) or bf
    if fp is not None:
        result["mixed_vs_fp64_max_rel_disp_diff"] = float("%.3e" % (
            np.abs(head["U_final"] - fp["U_final"]).max() / np.abs(fp["U_final"]).max()))
    bj = runs.get(mode_label("mixed", True, "block-jacobi", None))
    if bj:
        result["speedup_vs_block_jacobi"] = round(bj["wall_s"] / head["wall_s"], 3)
    if not args.no_cpu:
        log("cpu baseline (oracle)")
        result["cpu_baseline"] = cpu_baseline(n_full=args.n)

    detail = {"modes": {}, "kernels_mixed": head_k}
    for name, r in runs.items():
        kt = 4 if r["precision"] == "mixed" else 8
        d = {k: v for k, v in r.items() if k != "U_final"}
        d["assembly"] = asm_rates(r, kt, ne)
        if r.get("levels"):
            d["kernels"] = kernel_rooflines(r, kt)
        detail["modes"][name] = d
    line = json.dumps(detail)
    print(f"[bench] detail {line}", file=sys.stderr, flush=True)
    try:
        os.makedirs(os.path.join(ROOT, "gpurun_out"), exist_ok=True)
        with open(os.path.join(ROOT, "gpurun_out", "bench_detail.json"), "w") as f:
            f.write(line + "\n")
    except OSError:
        pass
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
