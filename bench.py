#!/usr/bin/env python3
"""Headline benchmark: GDOF/s of the DG residual + VJP on one MI355X.

Workload (default): 3D compressible Euler, HLLC flux, 64^3 periodic hex mesh of
the unit cube, p = 3 (20 modes x 5 variables = 26 214 400 DOF), FP64.  One step
is one fused pass computing the residual R(U) and the vector-Jacobian product
J(U)^T W over the whole mesh, with U and W resident in HBM.

Prints ONE JSON line (see the keys below).  The roofline entry prices the fused
kernel's dense-contraction flops (the reference formulation, 2x the forward
count) against the MI355X FP64 peak; per-launch durations come from HIP events
on the stream the kernel runs on.  The CPU baseline is the C++ restatement of
the reference evaluator (oracle/) timed on a bounded sample on the host cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6
METRIC = "GDOF/s DG residual apply + VJP (FP64 vs mixed) at fixed accuracy, % roofline"


def cpu_baseline(form, p, budget_s=15.0, n_max=64):
    """Time the oracle's residual + VJP on a periodic n^3 sample sized to ~budget_s."""
    import numpy as np
    from bosss_amd.grid import CartesianGrid
    from bosss_amd.operator import default_quad_order
    from bosss_amd.refelement import num_modes
    from bosss_amd.workloads import euler_state, uniform_coefficients
    from oracle.bosss_ref import RefModel, params_vector, OP_EULER_HLLC, OP_EULER_LF, OP_ADVDIFF, OP_SIP, _load

    kind = {"euler_hllc": OP_EULER_HLLC, "euler_lf": OP_EULER_LF, "advdiff": OP_ADVDIFF, "sip": OP_SIP}[form]
    order = default_quad_order(form, p)
    N = num_modes(3, p)
    prm = params_vector(1.3, -0.01, (1.0, 0.5, 0.25), 1.4)

    def run(n):
        g = CartesianGrid((n,) * 3, periodic=[True] * 3)
        m = RefModel(3, (n,) * 3, p, order, periodic=[1, 1, 1])
        U = euler_state(g, N, seed=1) if form.startswith("euler") else uniform_coefficients(g.J * N, 1)
        W = uniform_coefficients(U.size, 2)
        t0 = time.perf_counter()
        m.evaluate(kind, prm, U)
        t1 = time.perf_counter()
        m.vjp(kind, prm, U, W)
        t2 = time.perf_counter()
        return U.size, t1 - t0, t2 - t1

    dof0, tr0, tv0 = run(8)                                  # calibration on 512 cells
    per_cell = (tr0 + tv0) / 512.0
    n = 8
    while (n + 2) ** 3 * per_cell < budget_s and n < n_max:
        n += 2
    dof, tr, tv = run(n)
    threads = _load().bref_num_threads()
    return {
        "value": dof / (tr + tv) / 1e9, "unit": "GDOF/s", "cores": int(threads), "kind": "port",
        "sample": "%s p=%d, %d^3 periodic cells (%d DOF): residual %.3f s + VJP %.3f s (residual alone %.4f GDOF/s)"
                  % (form, p, n, dof, tr, tv, dof / tr / 1e9),
        "host": host_cpu_layout(),
    }


def host_cpu_layout():
    """The host's CPU layout as lscpu reports it (model, sockets, cores per socket, threads per
    core) and the CPUs this process may run on."""
    import subprocess
    info = {"cpus_usable": len(os.sched_getaffinity(0))}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
    except (OSError, subprocess.SubprocessError):
        return info
    keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
            "Thread(s) per core": "threads_per_core", "CPU(s)": "cpus"}
    for line in out.splitlines():
        k, _, v = line.partition(":")
        if k.strip() in keys and keys[k.strip()] not in info:
            info[keys[k.strip()]] = v.strip()
    return info


def kernel_label(form, p, split):
    """Name of the dominant kernel of one step, as rocprofv3 lists it (template arguments shortened)."""
    if split:
        return "residual kernel + VJP kernel (%s, p=%d)" % (form, p)
    if form.startswith("euler"):
        return "kt_vjp_euler<FormEuler<3,%s>,D=3,P=%d,K1=%d,double,fused>" % (
            "HLLC" if form == "euler_hllc" else "LF", p, (3 * p + 2) // 2)   # Gauss points of order 3p
    return "kt_adjoint<%s,D=3,P=%d,double,fused>" % (form, p)


def hbm_traffic(kernel, n):
    """HBM bytes per launch of `kernel` on the n^3 mesh, from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (profiles/hbm_traffic.json), or None if none was measured for it."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    for r in rec.get("kernels", []):
        if r.get("kernel") == kernel and r.get("mesh") == n:
            return r.get("bytes_per_launch")
    return None


def fp64_issue(kernel, n, kernel_ms):
    """FP64 work the kernel issues per launch on the n^3 mesh (committed rocprofv3 counts,
    profiles/fp64_issue.json: MFMA flops = SQ_INSTS_VALU_MFMA_MOPS_F64 x 512, VALU flops =
    64 x (ADD + MUL + TRANS + 2 FMA)_F64) over this run's kernel time: the MFMA, VALU and combined
    rates as fractions of the FP64 peak, or None if no counts were taken for the kernel.  The
    dense-contraction roofline counts only the algorithm's contractions; this counts what the
    kernel issues, padding of the MFMA tiles and the pointwise flux and adjoint included."""
    path = os.path.join(ROOT, "profiles", "fp64_issue.json")
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    for r in rec.get("kernels", []):
        if r.get("kernel") == kernel and r.get("mesh") == n:
            t = kernel_ms * 1e-3 * 1e12
            m, v = r["mfma_flops"] / t, r["valu_flops"] / t
            return {"mfma_tflops": round(m, 2), "valu_tflops": round(v, 2), "mfma_frac": round(m / FP64_PEAK_TFLOPS, 4),
                    "valu_frac": round(v / FP64_PEAK_TFLOPS, 4), "combined_frac": round((m + v) / FP64_PEAK_TFLOPS, 4),
                    "mfma_flops_per_launch": r["mfma_flops"], "valu_flops_per_launch": r["valu_flops"]}
    return None


def copy_calibration(dev, nbytes=1 << 30):
    """One device-to-device copy of nbytes (torch copy kernel): the known-size reference that
    rocprofv3 FETCH_SIZE / WRITE_SIZE readings of the DG kernel are scaled by."""
    import torch
    a = torch.ones(nbytes // 8, dtype=torch.float64, device=dev)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    del a, b


def measured_ceilings(op, dev, nbytes=1 << 30):
    """The device's ceilings measured in this process: dense FP64 / FP32 MFMA and FP64 FMA
    issue rates (bosss_hip_measure_peaks), the FP64 rate of MFMA and FMA interleaved at equal
    flops (bosss_hip_measure_fp64_mix: about either single rate when the two share one FP64
    rate, about their sum if they ran side by side), and the HBM rate of a device-to-device copy
    of nbytes (read + write bytes over the best of three copies)."""
    import ctypes
    import torch
    from bosss_amd._lib import lib
    f64, f32, fma, mix = ctypes.c_double(), ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    rc = lib().bosss_hip_measure_peaks(op.ctx.h, ctypes.byref(f64), ctypes.byref(f32), ctypes.byref(fma))
    if rc != 0:
        raise RuntimeError("bosss_hip_measure_peaks: %s" % lib().bosss_hip_peaks_last_error().decode())
    rc = lib().bosss_hip_measure_fp64_mix(op.ctx.h, ctypes.byref(mix))
    if rc != 0:
        raise RuntimeError("bosss_hip_measure_fp64_mix: %s" % lib().bosss_hip_peaks_last_error().decode())
    a = torch.ones(nbytes // 8, dtype=torch.float64, device=dev)
    b = torch.empty_like(a)
    ms = min(_time_ms(lambda: b.copy_(a), 1, 1) for _ in range(3))
    del a, b
    return {"fp64_mfma_tflops": round(f64.value, 2), "fp32_mfma_tflops": round(f32.value, 2),
            "fp64_fma_tflops": round(fma.value, 2), "fp64_mfma_fma_interleaved_tflops": round(mix.value, 2),
            "hbm_copy_GBps": round(2 * nbytes / (ms * 1e-3) / 1e9, 1)}


def _time_ms(fn, steps, warmup):
    """Median duration of fn() over `steps` launches (a HIP event pair around each launch on
    torch's current stream).  A host-side stall between an event and its launch leaves the
    GPU idle inside that pair; the median keeps such a stall (one 38 ms pair has been seen
    among 0.2 ms launches) out of the figure."""
    import numpy as np
    import torch
    stream = torch.cuda.current_stream()
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    for a, b in ev:
        a.record(stream)
        fn()
        b.record(stream)
    torch.cuda.synchronize()
    return float(np.median([a.elapsed_time(b) for a, b in ev]))


def secondary_workload(args):
    """The other measured configurations of the DG operator path (one JSON line each):

    sip2d    SIP Poisson, 64^2 Cartesian, p=2, Dirichlet: residual + assembled block-sparse apply
    advdiff  upwind + SIP (nu 0.01), 64^3 periodic, p=3: matrix-free FP64 and FP32 vs assembled SpMV
    euler32  Euler LF and HLLC, 32^3 periodic, p=2: fused residual + VJP
    xdg32    XDG Poisson, sphere in 32^3, p=2: cut-cell rules, assembly, apply, block-Jacobi PCG
    blocklu  batched cell-block LU: the XDG 10/20-DOF blocks, the 100-DOF Euler p=3 blocks (with
             inverses), and multi-cell blocks of 256 to 2048 DOF (Schwarz parts)
    """
    import numpy as np
    import torch
    from bosss_amd.grid import CartesianGrid
    from bosss_amd.operator import DifferentialOperator
    from bosss_amd.workloads import euler_state, uniform_coefficients, contraction_flops
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    K, W0 = args.steps, args.warmup
    res = {}
    if args.workload == "xdg32":
        from bosss_amd.xdg import QuadLevelSet, XdgRules, XdgOperator
        grid = CartesianGrid((32,) * 3)
        t0 = time.perf_counter()
        rules = XdgRules(grid, QuadLevelSet.sphere([0.5] * 3, 0.3), n=3)
        t_rules = time.perf_counter() - t0
        mu, src = (-1.0, -10.0), (1.0, 1.0)
        op = XdgOperator(rules, 2, mu, source=src, penalty=2.0, g_D=None)
        B = op.new_matrix()
        a = torch.zeros(op.ndof, dtype=torch.float64, device=dev)
        ms_asm = _time_ms(lambda: op.assemble_dev(B, a), K, W0)
        y = torch.zeros_like(a)
        xr = torch.rand(op.ndof, dtype=torch.float64, device=dev)
        ms_mv = _time_ms(lambda: op.spmv_dev(B, xr, y), K, W0)
        b = -a
        res["rules"] = {"host_seconds": round(t_rules, 3), "cut_cells": int(rules.cut.sum()),
                        "cells_A_only": int((rules.has[:, 0] & ~rules.cut).sum()), "cells_B_only": int((rules.has[:, 1] & ~rules.cut).sum()),
                        "volume_nodes": int(len(rules.vol_w)), "face_nodes": int(len(rules.face_w)), "interface_nodes": int(len(rules.ifc_w)),
                        "min_cut_species_fraction": float(rules.volume_fraction[rules.cut].min())}
        res["assembly_ms"] = round(ms_asm, 4)
        res["spmv_ms"] = round(ms_mv, 4)
        res["matrix_bytes"] = op.nentries * 8
        for prec, graph in (("fp64", True), ("fp64", False), ("fp32", True)):
            x = torch.zeros_like(a)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            it, outer, rel = op.solve_dev(B, b, x, tol=1e-10, maxit=50000, precision=prec, use_graph=graph, check_every=16)
            torch.cuda.synchronize()
            t = time.perf_counter() - t0
            res["pcg_%s_%s" % (prec, "graph" if graph else "eager")] = {
                "seconds": round(t, 4), "iterations": it, "outer": outer, "relres": rel, "us_per_iteration": round(t / max(it, 1) * 1e6, 2)}
        for prec in ("fp64", "fp32"):
            chosen, ms = op.tiles(prec)
            res["tile_autotuner_%s" % prec] = {"chosen_groups_of_4_rows": chosen, "ms_per_apply_update": ms}
        dof, desc = op.ndof, ("XDG Poisson, sphere r=0.3 in [0,1]^3, 32^3 cells, p=2, mu_A=-1, mu_B=-10, penalty 2, "
                              "agglomeration threshold 0: assembly, block-sparse apply, block-Jacobi PCG")
        op.close()
    elif args.workload == "blocklu":
        from bosss_amd.blocks import BlockBatch
        cases = [("xdg_sphere_32cube_cells", [(10, 31008), (20, 1760)], True),
                 ("euler_p3_64cube_cells", [(100, 262144)], True),
                 ("schwarz_256", [(256, 256)], False), ("schwarz_540", [(540, 64)], False),
                 ("schwarz_1024", [(1024, 64)], False), ("schwarz_2048", [(2048, 16)], False)]
        gen = torch.Generator(device=dev)
        gen.manual_seed(0)
        for name, groups, inv in cases:
            sizes = [n for n, cnt in groups for _ in range(cnt)]
            bb = BlockBatch(sizes, device=dev)
            parts = []
            for n, cnt in groups:   # uniform[-1, 1] entries plus a dominant diagonal, as the tests' blocks
                blk = torch.rand((cnt, n, n), dtype=torch.float64, device=dev, generator=gen) * 2.0 - 1.0
                blk += 0.2 * n * torch.eye(n, dtype=torch.float64, device=dev)
                parts.append(blk.view(-1))
            A0 = torch.cat(parts)
            del parts
            A = torch.empty_like(A0)
            ipiv = torch.zeros(bb.ndof, dtype=torch.int32, device=dev)
            info = torch.zeros(bb.nblk, dtype=torch.int32, device=dev)
            per = []
            for k in range(W0 + K):
                A.copy_(A0)                       # every factorisation starts from the same blocks
                ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                ev0.record()
                bb.factor(A, None if inv else ipiv, info, inverse=inv)
                ev1.record()
                torch.cuda.synchronize()
                if k >= W0:
                    per.append(ev0.elapsed_time(ev1))
            ms = float(np.median(per))
            flops = sum(cnt * (2.0 if inv else 2.0 / 3.0) * n ** 3 for n, cnt in groups)
            nbytes = 2 * 8 * sum(cnt * n * n for n, cnt in groups)
            res[name] = {"blocks": {str(n): cnt for n, cnt in groups}, "inverse": inv, "ms": round(ms, 4),
                         "gflop": round(flops / 1e9, 4), "tflops": round(flops / ms / 1e9, 3),
                         "frac_fp64_peak": round(flops / ms / 1e9 / FP64_PEAK_TFLOPS, 4),
                         "hbm_gb_s": round(nbytes / ms / 1e6, 1), "nonzero_info": int((info != 0).sum())}
            del A, A0, ipiv, info
            torch.cuda.empty_cache()
        dof, desc = 0, ("batched cell-block LU (getf2 pivots; inverses by Gauss-Jordan up to 128 DOF, "
                        "blocked LU with a matrix-unit trailing update above 128 DOF)")
    elif args.workload == "euler32":
        grid = CartesianGrid((32,) * 3, periodic=[True] * 3)
        for form in ("euler_lf", "euler_hllc"):
            op = DifferentialOperator(grid, form, 2, gamma=1.4)
            U = torch.from_numpy(euler_state(grid, op.N, seed=0)).to(dev)
            Wt = torch.from_numpy(uniform_coefficients(op.ndof, 1)).to(dev)
            R, G = torch.zeros_like(U), torch.zeros_like(U)
            ms = _time_ms(lambda: op.evaluate_vjp_dev(U, R, Wt, G), K, W0)
            ms_r = _time_ms(lambda: op.evaluate_dev(U, R), K, W0)
            fl = contraction_flops(form, 3, 2, grid.J, grid.E)
            res[form] = {"residual_vjp_ms": round(ms, 4), "gdof_s": round(op.ndof / ms / 1e6, 4),
                         "residual_ms": round(ms_r, 4), "residual_gdof_s": round(op.ndof / ms_r / 1e6, 4),
                         "tflops_alg": round(2 * fl / ms / 1e9, 3)}
            op.close()
        dof, desc = grid.J * 50, "3D Euler LF/HLLC, 32^3 periodic, p=2 (10 modes x 5 vars), fused residual + VJP"
    else:
        if args.workload == "sip2d":
            grid = CartesianGrid((64, 64), periodic=[False, False])
            form, p, nu = "sip", 2, 1.0
            desc = "2D SIP Poisson, 64^2 Cartesian, p=2, Dirichlet: residual and assembled block-sparse apply"
        else:
            grid = CartesianGrid((64,) * 3, periodic=[True] * 3)
            form, p, nu = "advdiff", 3, -0.01
            desc = "3D upwind advection + SIP, 64^3 periodic, p=3: matrix-free FP64/FP32 vs assembled SpMV"
        vel = (1.0, 0.5, 0.25)
        op = DifferentialOperator(grid, form, p, penalty=1.3, nu=nu, velocity=vel)
        dof = op.ndof
        U = torch.from_numpy(uniform_coefficients(dof, 0)).to(dev)
        R = torch.zeros_like(U)
        ms64 = _time_ms(lambda: op.evaluate_dev(U, R), K, W0)
        fl = contraction_flops(form, grid.D, p, grid.J, grid.E)
        res["matrix_free_fp64"] = {"ms": round(ms64, 4), "gdof_s": round(dof / ms64 / 1e6, 4),
                                   "tflops_alg": round(fl / ms64 / 1e9, 3)}
        if args.workload == "advdiff":
            op32 = DifferentialOperator(grid, form, p, penalty=1.3, nu=nu, velocity=vel, precision="fp32")
            U32, R32 = U.float(), torch.zeros(dof, dtype=torch.float32, device=dev)
            ms32 = _time_ms(lambda: op32.evaluate_dev(U32, R32), K, W0)
            err = float((R32.double() - R).abs().max() / R.abs().max())
            res["matrix_free_fp32"] = {"ms": round(ms32, 4), "gdof_s": round(dof / ms32 / 1e6, 4),
                                       "tflops_alg": round(fl / ms32 / 1e9, 3), "max_rel_err_vs_fp64": err}
            op32.close()
        mb = op.GetMatrixBuilder()
        blocks = mb.new_matrix()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        mb.ComputeMatrix(blocks)
        torch.cuda.synchronize()
        t_asm = time.perf_counter() - t0
        b = torch.from_numpy(np.asarray(mb.ComputeAffine())).to(dev)      # R(U) = M U + b
        Y = torch.zeros_like(U)
        ms_mv = _time_ms(lambda: op.spmv_dev(blocks, U, Y), K, W0)
        err = float((Y + b - R).abs().max() / R.abs().max())
        nbytes = blocks.numel() * 8 + 2 * dof * 8
        res["assembled"] = {"assembly_s": round(t_asm, 4), "probe_colours": mb.ncolors, "matrix_bytes": blocks.numel() * 8,
                            "spmv_ms": round(ms_mv, 4), "spmv_gdof_s": round(dof / ms_mv / 1e6, 4),
                            "spmv_gb_s": round(nbytes / ms_mv / 1e6, 1), "spmv_frac_of_8TBs": round(nbytes / ms_mv / 1e6 / 8000, 4),
                            "max_rel_diff_MU_plus_b_vs_residual": err,
                            "spmv_tile_autotuner": op.spmv_tiles()}
        op.close()
    out = {"metric": "DG operator apply, secondary configuration", "workload": args.workload, "config": desc,
           "dof": dof, "steps": K, "warmup": W0, "dtype": "f64", "data": "synthetic", "results": res}
    print(json.dumps(out), flush=True)


def fp32_comparison(grid, form, p, Ut, Wt, R64, G64, steps, warmup):
    """The same fused residual + VJP step in FP32 (FP32 MFMA and FP32 flux, inputs rounded to
    FP32 on the device): its time and its error against the FP64 result, per field."""
    import torch
    from bosss_amd.operator import DifferentialOperator
    op32 = DifferentialOperator(grid, form, p, penalty=1.3, nu=-0.01, velocity=(1.0, 0.5, 0.25), gamma=1.4, precision="fp32")
    U32, W32 = Ut.float(), Wt.float()
    R32, G32 = torch.zeros_like(U32), torch.zeros_like(U32)
    ms = _time_ms(lambda: op32.evaluate_vjp_dev(U32, R32, W32, G32), steps, warmup)
    V, N = op32.nvars, op32.N

    def err(a32, a64):
        a, b = a32.double().view(-1, V, N), a64.view(-1, V, N)
        return max(float((a[:, f] - b[:, f]).abs().max() / b[:, f].abs().max()) for f in range(V))

    def nerr(a32, a64):
        return float((a32.double() - a64).norm() / a64.norm())

    # The HLLC Jacobian jumps where the contact speed changes sign: at the few face nodes within
    # rounding distance of such a switch the FP32 VJP follows the other branch, which sets its
    # max-norm error (the FP64 kernel fed the FP32-rounded state differs from FP64 the same way).
    return {"ms": round(ms, 4), "gdof_s": round(op32.ndof / (ms * 1e-3) / 1e9, 4),
            "max_rel_err_residual_vs_fp64": err(R32, R64), "normwise_rel_err_residual_vs_fp64": nerr(R32, R64),
            "max_rel_err_vjp_vs_fp64": err(G32, G64), "normwise_rel_err_vjp_vs_fp64": nerr(G32, G64)}


def newton_comparison(grid, op64, Un, form, p):
    """One implicit Euler step solved to ||F|| <= 1e-10: FP64 Newton-GMRES vs FP64 defect
    correction around FP32 JVPs (same driver, bosss_amd.newton), both with the device-resident
    GMRES (Arnoldi steps captured as HIP graphs); and the FP64 solve once more with the Givens
    rotations on the host (one synchronisation per Krylov step) for comparison."""
    import torch
    from bosss_amd.newton import ImplicitStep
    from bosss_amd.operator import DifferentialOperator
    op32 = DifferentialOperator(grid, form, p, gamma=1.4, precision="fp32")
    dt = 0.5 / grid.ncells[0] / (2 * p + 1)
    st = ImplicitStep(op64, Un, dt, op32=op32)
    sth = ImplicitStep(op64, Un, dt, krylov="host")
    res = {}
    for name, fn in (("fp64", st.solve_fp64), ("mixed", st.solve_mixed), ("fp64_host_krylov", sth.solve_fp64)):
        fn(Un, tol=1e-10)                # warm-up (allocations, first launches, graph capture)
        best = None
        for _ in range(2):               # best of two solves
            torch.cuda.synchronize()
            U = fn(Un, tol=1e-10)
            s = dict((st if fn != sth.solve_fp64 else sth).stats)
            if best is None or s["seconds"] < best["seconds"]:
                best = s
            del U
        res[name] = {"seconds": round(best["seconds"], 4), "newton_steps": best["newton"], "jvps": best["jvps"],
                     "final_residual": best["history"][-1],
                     "ms_per_krylov_step": round(best["seconds"] / max(best["jvps"], 1) * 1e3, 4)}
    res["dt"] = dt
    res["krylov"] = "device-resident GMRES, Arnoldi steps as HIP graphs, residual estimate read every 8 steps"
    res["speedup_mixed_vs_fp64"] = round(res["fp64"]["seconds"] / res["mixed"]["seconds"], 3)
    return res


def implicit_cfl_comparison(grid, op64, Un, form, p, mult=10.0):
    """The same implicit step at mult times the time step (a CFL number mult times larger), where the
    Krylov iterations grow: plain FP64 Newton-GMRES against the cell-block preconditioned variants
    (blocks of I/dt + dR_j/dU_j from own-cell JVP probes, inverted in FP64): FP64 with the FP64
    preconditioner, and the FP64 defect correction around FP32 GMRES with the FP32 preconditioner
    (FP32 probes and apply).  One solve each after a warm-up; the preconditioner set-up (probes,
    batched LU with inverse, transposition) is timed separately."""
    import torch
    from bosss_amd.blocks import CellBlockPreconditioner
    from bosss_amd.newton import ImplicitStep
    from bosss_amd.operator import DifferentialOperator
    op32 = DifferentialOperator(grid, form, p, gamma=1.4, precision="fp32")
    dt = mult * 0.5 / grid.ncells[0] / (2 * p + 1)
    st = ImplicitStep(op64, Un, dt, op32=op32)
    res = {"dt": dt, "dt_multiple_of_bench_step": mult}
    for name in ("fp64_plain", "fp64_cellblock_fp64", "mixed_cellblock_fp32"):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if name == "fp64_plain":
            pc = None
            fn = lambda: st.solve_fp64(Un, tol=1e-10, maxit=2000)
        elif name == "fp64_cellblock_fp64":
            pc = CellBlockPreconditioner(op64, Un, 1.0 / dt, precision="fp64")
            fn = lambda: st.solve_fp64(Un, tol=1e-10, precond=pc)
        else:
            pc = CellBlockPreconditioner(op32, Un.float(), 1.0 / dt, precision="fp32")
            fn = lambda: st.solve_mixed(Un, tol=1e-10, precond=pc)
        setup = time.perf_counter() - t0
        fn()                                   # warm-up (graph capture)
        U = fn()
        s = dict(st.stats)
        res[name] = {"seconds": round(s["seconds"], 4), "setup_seconds": round(setup, 4), "newton_steps": s["newton"],
                     "jvps": s["jvps"], "final_residual": s["history"][-1], "converged": s["converged"]}
        del U, pc
        torch.cuda.empty_cache()
    op32.close()
    return res


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--settle-s", type=float, default=2.0,
                    help="seconds of untimed steps before the warm-up steps (GPU clock settle)")
    ap.add_argument("--n", type=int, default=64, help="cells per direction")
    ap.add_argument("--p", type=int, default=3)
    ap.add_argument("--form", default="euler_hllc", choices=["euler_hllc", "euler_lf", "advdiff", "sip"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--split", action="store_true", help="time separate residual and VJP launches instead of the fused one")
    ap.add_argument("--no-newton", action="store_true", help="skip the implicit-step Newton comparison (FP64 vs mixed)")
    ap.add_argument("--workload", default="euler64", choices=["euler64", "sip2d", "advdiff", "euler32", "xdg32", "blocklu"],
                    help="euler64: the headline line; the others print the secondary configurations")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run one 1 GiB device copy first (known-size reference for HBM counter readings)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (the multi-GPU sharding is a design, not a code path); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload != "euler64":
        secondary_workload(args)
        return

    import numpy as np
    import torch
    from bosss_amd.grid import CartesianGrid
    from bosss_amd.operator import DifferentialOperator
    from bosss_amd.workloads import euler_state, uniform_coefficients, contraction_flops

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    n, p, form = args.n, args.p, args.form
    grid = CartesianGrid((n,) * 3, periodic=[True] * 3)
    op = DifferentialOperator(grid, form, p, penalty=1.3, nu=-0.01, velocity=(1.0, 0.5, 0.25), gamma=1.4)
    U = euler_state(grid, op.N, seed=0) if form.startswith("euler") else uniform_coefficients(op.ndof, 0)
    W = uniform_coefficients(op.ndof, 1)
    Ut = torch.from_numpy(U).to(dev)
    Wt = torch.from_numpy(W).to(dev)
    Rt = torch.zeros_like(Ut)
    Gt = torch.zeros_like(Ut)
    stream = torch.cuda.current_stream()
    if args.copy_calibration:
        copy_calibration(dev)

    def step():
        if args.split:
            op.evaluate_dev(Ut, Rt)
            op.vjp_dev(Ut, Wt, Gt)
        else:
            op.evaluate_vjp_dev(Ut, Rt, Wt, Gt)

    # Clock settle: a GPU that has just been idle runs its first milliseconds of work below its
    # sustained clock (on a fresh process 25 launches of this step measured 6.6 ms each against
    # 4.5 ms once warm).  Launch the step for settle_s seconds of GPU time before the warm-up
    # steps, so that the timed steps see the sustained clock whatever ran before.
    settle_steps, t_settle = 0, time.perf_counter()
    while time.perf_counter() - t_settle < args.settle_s:
        for _ in range(10):
            step()
        torch.cuda.synchronize()
        settle_steps += 10
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        starts[i].record(stream)
        step()
        ends[i].record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    per_launch = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    kernel_ms = sum(per_launch) / args.steps
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = op.ndof / (ms_per_step * 1e-3) / 1e9

    flops = 2 * contraction_flops(form, 3, p, grid.J, grid.E)  # residual + VJP
    achieved = flops / (kernel_ms * 1e-3) / 1e12
    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "GDOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "3D %s, %d^3 periodic hex mesh of the unit cube, p=%d (%d modes x %d vars, %d DOF), "
                               "fused residual + VJP per step" % (form, n, p, op.N, op.nvars, op.ndof),
                   "mesh": [n, n, n], "degree": p, "form": form, "dof": op.ndof},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
                     "traffic": hbm_traffic(kernel_label(form, p, args.split), n),
                     "kernel": kernel_label(form, p, args.split),
                     "kernel_ms": round(kernel_ms, 4), "alg_flops_per_launch": flops,
                     "kernel_ms_median": round(float(np.median(per_launch)), 4),
                     "kernel_ms_min_max": [round(min(per_launch), 4), round(max(per_launch), 4)],
                     "fp64_issue": fp64_issue(kernel_label(form, p, args.split), n, kernel_ms)},
        "clock_settle": {"seconds": args.settle_s, "steps": settle_steps},
    }
    print("bench: %s kernels, %s, %.3f ms/step, %.3f ms kernel" % (
        "generic-table" if os.environ.get("BOSSS_HIP_GENERIC_TABLES") == "1" else "tensor-table",
        "split" if args.split else "fused", ms_per_step, kernel_ms), file=sys.stderr, flush=True)
    if not args.no_newton:
        out["roofline"]["measured_ceilings"] = measured_ceilings(op, dev)
    if form.startswith("euler") and not args.split and not args.no_newton:
        f32 = fp32_comparison(grid, form, p, Ut, Wt, Rt, Gt, args.steps, args.warmup)
        f32["speedup_vs_fp64"] = round(kernel_ms / f32["ms"], 3)
        out["fp32_fused"] = f32
    if form.startswith("euler") and not args.no_newton:
        out["newton_implicit_step"] = newton_comparison(grid, op, Ut, form, p)
        out["newton_implicit_step_cfl10"] = implicit_cfl_comparison(grid, op, Ut, form, p)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(form, p)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
