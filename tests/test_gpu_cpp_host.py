"""The library driven by a plain C++ host (tests/cpp/abi_harness.cpp, g++ only: no torch, no
Python, no HIP headers).  The harness binds every entry point of include/bosss_hip.h by name, as
BoSSS's DynLibLoader.LoadAllDelegates binds its native libraries (DynLibLoader.cs:365-410), and
passes host arrays only: the host-pointer entry points, or device buffers from the buffer calls
plus the *_dev entry points.  Its outputs are checked against the oracle:

  dg        evaluate / jvp / vjp (host pointers) and the fused residual + VJP through device
            buffers, alpha/beta accumulation, the NaN/Inf check (Euler HLLC p = 3, SIP 2D p = 2)
  blocks    block LU pivots, INFO and factors (getf2), block solve, inverse + block gemv, for
            uneven batches in LDS (n <= 20) and in global memory (n = 300, 257)
  assemble  the block-ELL operator matrix (upwind + SIP) against the oracle's assembled matrix,
            M x through the host-pointer apply, and M x + b against the residual
  xdg       XDG Poisson (sphere, p = 2): matrix and affine vector against the restated assembly,
            the apply, FP64 and FP32 + refinement solves to 1e-10
"""
import os
import subprocess

import numpy as np
import pytest

import recfile
from bosss_amd.grid import CartesianGrid
from bosss_amd.operator import default_quad_order
from bosss_amd.refelement import Tables
from bosss_amd.workloads import euler_state
from oracle.bosss_ref import (RefModel, params_vector, ref_block_lu, ref_block_solve, xdg_assemble, xdg_to_csr,
                              OP_ADVDIFF, OP_EULER_HLLC, OP_SIP)

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HARNESS = os.path.join(ROOT, "tests", "cpp", "abi_harness")
LIB = os.path.join(ROOT, "bosss_amd", "libbosss_hip.so")
KINDS = {"sip": 0, "advdiff": 1, "euler_lf": 2, "euler_hllc": 3}


def run(scenario, recs, tmp_path):
    if not os.path.exists(HARNESS):
        pytest.fail("tests/cpp/abi_harness is not built (run __graft_entry__.build())")
    fin, fout = str(tmp_path / "in.rec"), str(tmp_path / "out.rec")
    recfile.write(fin, recs)
    print("case harness %s" % scenario, flush=True)
    r = subprocess.run([HARNESS, LIB, scenario, fin, fout], capture_output=True, text=True, timeout=240)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr[-2000:]
    return recfile.read(fout)


def grid_recs(g, form, p, prm):
    t = Tables(g.D, p, default_quad_order(form, p))
    i32 = lambda x: np.asarray(x, np.int32)
    return dict(D=i32(g.D), J=i32(g.J), E=i32(g.E), detJ=g.detJ, invJ=g.invJ, scaling=g.scaling, lenscale=g.lenscale,
                cells=i32(g.cells), trafo=i32(g.trafo), normals=g.normals, sqrtgram=g.sqrtgram,
                degree=i32(p), N=i32(t.N), Kv=i32(t.Kv), Ke=i32(t.Ke), ntrafo=i32(t.ntrafo),
                vw=t.vw, phi=t.phi, dphi=t.dphi, fw=t.fw, trace=t.trace, dtrace=t.dtrace,
                kind=i32(KINDS[form]), params=np.asarray(prm, np.float64))


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("form,D,n,p", [("euler_hllc", 3, 3, 3), ("sip", 2, 5, 2)])
def test_dg_operator_from_cpp_host(form, D, n, p, tmp_path):
    periodic = form.startswith("euler")
    g = CartesianGrid((n,) * D, periodic=[periodic] * D)
    prm = params_vector(1.3, 1.0, (0.0, 0.0, 0.0), 1.4)
    ref = RefModel(D, (n,) * D, p, default_quad_order(form, p), periodic=[int(periodic)] * D)
    kind = {"euler_hllc": OP_EULER_HLLC, "sip": OP_SIP}[form]
    Nt = ref.nvars(kind) * ref.N
    rng = np.random.default_rng(5)
    U = euler_state(g, ref.N, seed=3) if periodic else rng.uniform(-1, 1, g.J * Nt)
    dU, W, R0 = (rng.uniform(-1, 1, U.size) for _ in range(3))
    out = run("dg", dict(grid_recs(g, form, p, prm), U=U, dU=dU, W=W, R0=R0), tmp_path)
    Rr = ref.evaluate(kind, prm, U)
    Jr = ref.jvp(kind, prm, U, dU)
    Gr = ref.vjp(kind, prm, U, W)
    assert rel(out["R"], Rr) < 1e-12
    assert rel(out["JdU"], Jr) < 1e-12
    assert rel(out["JtW"], Gr) < 1e-12
    assert rel(out["Rab"], 0.5 * R0 + 2.0 * Rr) < 1e-12
    assert rel(out["Rdev"], Rr) < 1e-12          # fused residual + VJP kernel through device buffers
    assert rel(out["Gdev"], Gr) < 1e-12
    assert out["rc_finite"][0] == 0
    assert out["rc_nan"][0] == -4 and out["rc_nan_dev"][0] == -4     # BOSSS_ERR_NONFINITE


@pytest.mark.parametrize("sizes", [[10] * 40 + [20] * 8 + [7] * 3 + [1], [300, 257]])
def test_block_batch_from_cpp_host(sizes, tmp_path):
    rng = np.random.default_rng(len(sizes))
    blocks = [rng.uniform(-1, 1, (s, s)) for s in sizes]
    blocks[0][:, 0] = 0.0                        # a zero first column: INFO = 1 for block 0
    blocks[1][0, 0] = 0.0                        # zero diagonal, non-zero pivot below: no INFO
    n = np.array(sizes, np.int32)
    off = np.zeros(len(n), np.int64); off[1:] = np.cumsum(n.astype(np.int64) ** 2)[:-1]
    poff = np.zeros(len(n), np.int64); poff[1:] = np.cumsum(n.astype(np.int64))[:-1]
    A = np.concatenate([b.ravel() for b in blocks])
    b = rng.uniform(-1, 1, int(n.sum()))
    out = run("blocks", dict(n=n, off=off, pivoff=poff, xoff=poff, A=A, b=b, nmax=np.int32(n.max()),
                             npiv=np.int32(n.sum())), tmp_path)
    LUr, pivr, infor, _, _ = ref_block_lu(sizes, blocks)
    assert (out["info"] == infor).all() and infor[0] == 1 and infor[1] == 0
    assert (out["ipiv"] == pivr + 1).all()       # LAPACK's 1-based pivots; the oracle returns 0-based rows
    assert rel(out["LU"], LUr) < 1e-12
    ok = np.repeat(infor == 0, n)
    xr = ref_block_solve(sizes, LUr, pivr, off, poff, b)
    assert rel(out["x"][ok], xr[ok]) < 1e-11
    if "inv" in out:
        y = out["y"]
        for k in range(len(sizes)):
            if infor[k] == 0:
                s = slice(poff[k], poff[k] + n[k])
                assert rel(y[s], np.linalg.solve(blocks[k], b[s])) < 1e-10


def test_assembled_matrix_from_cpp_host(tmp_path):
    D, n, p = 3, 3, 2
    g = CartesianGrid((n,) * D, periodic=[True] * D)
    prm = params_vector(1.3, -0.01, (1.0, 0.5, 0.25), 1.4)
    ref = RefModel(D, (n,) * D, p, default_quad_order("advdiff", p), periodic=[1] * D)
    x = np.random.default_rng(2).uniform(-1, 1, g.J * ref.N)
    out = run("assemble", dict(grid_recs(g, "advdiff", p, prm), U=x), tmp_path)
    Nt, S = ref.N, int(out["S"][0])
    blocks, cols = out["blocks"].reshape(g.J, S, Nt, Nt), out["cols"].reshape(g.J, S)
    A = np.zeros((g.J * Nt, g.J * Nt))
    for j in range(g.J):
        for s in range(S):
            if cols[j, s] >= 0:
                A[j * Nt:(j + 1) * Nt, cols[j, s] * Nt:(cols[j, s] + 1) * Nt] += blocks[j, s]
    Ar = ref.dense_matrix(OP_ADVDIFF, prm)
    assert rel(A, Ar) < 1e-12
    assert rel(out["y"], Ar @ x) < 1e-12
    assert rel(out["y"] + out["b"], ref.evaluate(OP_ADVDIFF, prm, x)) < 1e-12


def test_xdg_from_cpp_host(tmp_path):
    import scipy.sparse.linalg as spla
    from bosss_amd.xdg import QuadLevelSet, XdgRules, desc_arrays
    p = 2
    g = CartesianGrid((6,) * 3)
    rules = XdgRules(g, QuadLevelSet.sphere([0.5] * 3, 0.3), n=p + 1)
    mu, src = (-1.0, -10.0), (1.0, 0.5)
    g_D = lambda X: 0.1 * X[:, 0] - 0.2 * X[:, 1] * X[:, 2]
    k = desc_arrays(rules, p, g_D)
    lay = rules.block_layout(len(k["ex"]))
    i32 = lambda a: np.asarray(a, np.int32)
    v = np.random.default_rng(4).uniform(-1, 1, lay["ndof"])
    recs = dict(k, D=i32(3), degree=i32(p), N=i32(len(k["ex"])), J=i32(g.J), penalty=np.float64(2.0),
                mu=np.array(mu), source=np.array(src), S=i32(lay["S"]), v=v)
    out = run("xdg", recs, tmp_path)
    Br, ar, _ = xdg_assemble(rules, p, mu, src, 2.0, rules.boundary_values(g_D))
    assert rel(out["blocks"], Br) < 1e-12
    assert rel(out["affine"], ar) < 1e-12
    M = xdg_to_csr(rules, lay, Br)
    assert rel(out["y"], M @ v) < 1e-12
    xs = spla.spsolve(M.tocsc(), -ar)
    for prec in ("fp64", "fp32"):
        sol = out["sol_" + prec]
        assert out["relres_" + prec][0] <= 1e-10
        assert np.linalg.norm(-ar - M @ sol) <= 1.01e-10 * np.linalg.norm(ar)
        assert rel(sol, xs) < 1e-5        # residual 1e-10 times the condition number of the cut-cell matrix
