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
    }


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


def newton_comparison(grid, op64, Un, form, p):
    """One implicit Euler step solved to ||F|| <= 1e-10: FP64 Newton-GMRES vs FP64 defect
    correction around FP32 JVPs (same driver, bosss_amd.newton)."""
    import torch
    from bosss_amd.newton import ImplicitStep
    from bosss_amd.operator import DifferentialOperator
    op32 = DifferentialOperator(grid, form, p, gamma=1.4, precision="fp32")
    dt = 0.5 / grid.ncells[0] / (2 * p + 1)
    st = ImplicitStep(op64, Un, dt, op32=op32)
    st.solve_fp64(Un, tol=1e-10)        # warm-up (allocations, first launches)
    res = {}
    for name, fn in (("fp64", st.solve_fp64), ("mixed", st.solve_mixed)):
        torch.cuda.synchronize()
        U = fn(Un, tol=1e-10)
        s = dict(st.stats)
        res[name] = {"seconds": round(s["seconds"], 4), "newton_steps": s["newton"], "jvps": s["jvps"],
                     "final_residual": s["history"][-1]}
        del U
    res["dt"] = dt
    res["speedup_mixed_vs_fp64"] = round(res["fp64"]["seconds"] / res["mixed"]["seconds"], 3)
    return res


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=64, help="cells per direction")
    ap.add_argument("--p", type=int, default=3)
    ap.add_argument("--form", default="euler_hllc", choices=["euler_hllc", "euler_lf", "advdiff", "sip"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--split", action="store_true", help="time separate residual and VJP launches instead of the fused one")
    ap.add_argument("--no-newton", action="store_true", help="skip the implicit-step Newton comparison (FP64 vs mixed)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run one 1 GiB device copy first (known-size reference for HBM counter readings)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (the multi-GPU sharding is a design, not a code path); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

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
    kernel_ms = sum(s.elapsed_time(e) for s, e in zip(starts, ends)) / args.steps
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
                     "kernel_ms": round(kernel_ms, 4), "alg_flops_per_launch": flops},
    }
    print("bench: %s kernels, %s, %.3f ms/step, %.3f ms kernel" % (
        "generic-table" if os.environ.get("BOSSS_HIP_GENERIC_TABLES") == "1" else "tensor-table",
        "split" if args.split else "fused", ms_per_step, kernel_ms), file=sys.stderr, flush=True)
    if form.startswith("euler") and not args.no_newton:
        out["newton_implicit_step"] = newton_comparison(grid, op, Ut, form, p)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(form, p)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
