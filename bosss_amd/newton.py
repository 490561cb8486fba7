"""Newton-Krylov for one implicit Euler step, device resident.

Mirrors Newton.GlobalizedNewton (Newton.cs:189-365) with the matrix-free GMRES
of Newton.MatrixFreeGMRES (Newton.cs:1128-1500), except that the Jacobian
action is the exact JVP kernel instead of the forward-difference `dirder`.

Problem (orthonormal DG basis => mass matrix I):
    F(U) = (U - Un)/dt + R(U) - S = 0,
with R the DG operator (dU/dt = -R(U)) and S a source in coefficient space.

Two drivers:
* ``solve_fp64``: Newton with FP64 residual and FP64 JVP inside GMRES;
* ``solve_mixed``: FP64 defect correction — the outer residual F(U) in FP64,
  the correction solved in FP32 (FP32 JVP on MFMA, FP32 vectors).  FP64 is used
  only for the outer residual and the state update.

Both globalise the step like GlobalizationOption.LineSearch (Newton.cs:807-860): Armijo
backtracking on ||F(U + lambda d)|| with alpha = 1e-4, first reduction sigma1 = 0.5, then
Kelley's safeguarded three-point parabola ``parab3p`` (Newton.cs:1622-1656), at most
``maxStep`` = 30 reductions; the accepted trial residual is the next iterate's residual.
``globalization="none"`` takes full steps (GlobalizationOption.None).  ``stats["converged"]``
says whether ||F|| <= tol was reached (GlobalizedNewton's `success`, Newton.cs:189-241).

GMRES is restarted, right-preconditioned (optional block-diagonal preconditioner
callable), classical Gram-Schmidt with one re-orthogonalisation, in the dtype of its
right-hand side.  The drivers use ``DeviceGmres`` by default: the Arnoldi step, the Hessenberg
matrix and the Givens rotations stay on the GPU, the host reads the residual estimate every
few steps, and each step can run as a captured HIP graph.  ``gmres`` is the same method with
the Givens rotations on the host (one synchronisation per step), kept as the plain
restatement the device version is tested against.
"""
import math
import time

import numpy as np


def _sync(t):
    if t.is_cuda:
        import torch
        torch.cuda.synchronize(t.device)


def gmres(apply_A, b, tol, restart=30, maxit=300, precond=None, true_residual=True):
    """Solve A x = b to ||b - A x|| <= tol * ||b||; returns (x, iterations, relative residual).

    A cycle that ends unconverged restarts from the true residual b - A x.  One that meets the
    tolerance by the Hessenberg least-squares estimate is confirmed with the true residual too
    (one more apply), unless true_residual is False: then the estimate is returned (inexact
    Newton, whose next nonlinear residual checks the step anyway)."""
    import torch
    n = b.numel()
    x = torch.zeros_like(b)
    bnorm = float(torch.linalg.vector_norm(b))
    if bnorm == 0.0:
        return x, 0, 0.0
    r = b.clone()
    it = 0
    rel = 1.0
    while it < maxit:
        beta = float(torch.linalg.vector_norm(r))
        rel = beta / bnorm
        if rel <= tol:
            break
        m = min(restart, maxit - it)
        Vb = torch.empty((m + 1, n), dtype=b.dtype, device=b.device)
        Zb = torch.empty((m, n), dtype=b.dtype, device=b.device) if precond is not None else None
        Vb[0] = r / beta
        H = np.zeros((m + 1, m))
        cs = np.zeros(m); sn = np.zeros(m)
        g = np.zeros(m + 1); g[0] = beta
        k_done = 0
        w = torch.empty_like(b)
        for k in range(m):
            if precond is not None:
                precond(Vb[k], Zb[k])
                apply_A(Zb[k], w)
            else:
                apply_A(Vb[k], w)
            # classical Gram-Schmidt, twice (one batched dot product per pass)
            V = Vb[:k + 1]
            h = V @ w
            w -= V.t() @ h
            h2 = V @ w
            w -= V.t() @ h2
            hh = (h + h2).double().cpu().numpy()
            hn = float(torch.linalg.vector_norm(w))
            H[:k + 1, k] = hh
            H[k + 1, k] = hn
            Vb[k + 1] = w / (hn if hn > 0 else 1.0)
            for i in range(k):
                t = cs[i] * H[i, k] + sn[i] * H[i + 1, k]
                H[i + 1, k] = -sn[i] * H[i, k] + cs[i] * H[i + 1, k]
                H[i, k] = t
            den = math.hypot(H[k, k], H[k + 1, k])
            cs[k], sn[k] = (1.0, 0.0) if den == 0 else (H[k, k] / den, H[k + 1, k] / den)
            H[k, k] = cs[k] * H[k, k] + sn[k] * H[k + 1, k]
            H[k + 1, k] = 0.0
            g[k + 1] = -sn[k] * g[k]
            g[k] = cs[k] * g[k]
            it += 1
            k_done = k + 1
            if abs(g[k + 1]) / bnorm <= tol or hn == 0:
                break
        y = np.linalg.solve(np.triu(H[:k_done, :k_done]), g[:k_done]) if k_done else np.zeros(0)
        yt = torch.from_numpy(y).to(device=b.device, dtype=b.dtype)
        dx = (Zb[:k_done] if precond is not None else Vb[:k_done]).t() @ yt
        x += dx
        if not true_residual and abs(g[k_done]) / bnorm <= tol:
            return x, it, abs(g[k_done]) / bnorm
        # true residual for the restart
        apply_A(x, w)
        r = b - w
    return x, it, float(torch.linalg.vector_norm(r)) / bnorm


class DeviceGmres:
    """Restarted, right-preconditioned GMRES whose Arnoldi process stays on the GPU
    (``bosss_hip_gmres_arnoldi_dev`` / ``bosss_hip_gmres_update_dev``): the Krylov vectors, the
    Hessenberg matrix, the Givens rotations and the residual estimates live in device memory, and
    the host reads the estimates only every ``check_every`` steps.  With ``use_graph`` each
    Arnoldi step k (preconditioner, operator, orthogonalisation, rotation) is captured once into a
    HIP graph and replayed; the operator closure must then keep reading the same buffers (the
    Newton driver passes its fixed linearisation-point copy).  Steps run past convergence inside
    a batch change nothing before them: the update uses the first k at which the estimate met the
    tolerance, so the result is the one a solver checking after every step returns.

    Mirrors Newton.MatrixFreeGMRES (Newton.cs:1128-1500) with the exact operator action in place
    of the forward-difference dirder."""

    def __init__(self, n, dtype, device, restart=40, preconditioned=False, ctx=None):
        import ctypes
        import torch
        from ._lib import lib
        from .operator import Context
        self.n, self.m, self.dtype = n, restart, dtype
        self.ctx = ctx or Context.get(device.index or 0)
        self.prec = 0 if dtype == torch.float64 else 1
        self.V = torch.zeros((restart + 1, n), dtype=dtype, device=device)
        self.Z = torch.zeros((restart, n), dtype=dtype, device=device) if preconditioned else None
        self.w = torch.zeros(n, dtype=dtype, device=device)
        f64 = dict(dtype=torch.float64, device=device)
        self.H = torch.zeros((restart, restart + 1), **f64)       # column-major (m+1) x m
        self.cs, self.sn = torch.zeros(restart, **f64), torch.zeros(restart, **f64)
        self.g, self.y = torch.zeros(restart + 1, **f64), torch.zeros(restart, **f64)
        self.res = torch.zeros(restart, **f64)                  # residual estimate after each step
        nw = ctypes.c_int64()
        self._check(lib().bosss_hip_gmres_workspace(restart, ctypes.byref(nw)))
        self.work = torch.zeros(nw.value, **f64)
        self.graphs = {}
        self.graph_key = None

    @staticmethod
    def _check(rc):
        from ._lib import lib, BosssError
        if rc != 0:
            raise BosssError(rc, lib().bosss_hip_krylov_last_error().decode())

    @staticmethod
    def _p(t):
        import ctypes
        return ctypes.c_void_p(t.data_ptr())

    def _step(self, k, apply_A, precond):
        import ctypes
        import torch
        from ._lib import lib
        if precond is not None:
            precond(self.V[k], self.Z[k])
            apply_A(self.Z[k], self.w)
        else:
            apply_A(self.V[k], self.w)
        p = self._p
        self._check(lib().bosss_hip_gmres_arnoldi_dev(self.ctx.h, self.n, k, self.m, self.prec, p(self.V), self.n, p(self.w),
                                                      p(self.H), p(self.cs), p(self.sn), p(self.g), p(self.res), p(self.work),
                                                      ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))

    def _run_step(self, k, apply_A, precond, use_graph, key):
        import torch
        if not use_graph:
            self._step(k, apply_A, precond)
            return
        if key != self.graph_key:
            self.graphs, self.graph_key = {}, key
        gr = self.graphs.get(k)
        if gr is None:
            # capture records the step without running it; the replay below runs it
            gr = torch.cuda.CUDAGraph()
            torch.cuda.synchronize()
            with torch.cuda.graph(gr):
                self._step(k, apply_A, precond)
            self.graphs[k] = gr
        gr.replay()

    def solve(self, apply_A, b, tol, maxit=300, precond=None, true_residual=True, check_every=8, use_graph=False, key=None):
        """Solve A x = b to ||b - A x|| <= tol ||b|| (apply_A(v, out), precond(v, z) as for gmres);
        returns (x, Krylov iterations used, relative residual or its estimate).  key identifies the
        operator closure for the graph cache (a new key recaptures)."""
        import ctypes
        import torch
        from ._lib import lib
        if precond is not None and self.Z is None:
            raise ValueError("DeviceGmres was built without preconditioned vectors")
        m = self.m
        x = torch.zeros_like(b)
        bnorm = float(torch.linalg.vector_norm(b))
        if bnorm == 0.0:
            return x, 0, 0.0
        r = b.clone()
        it, rel = 0, 1.0
        stream = lambda: ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        while it < maxit:
            beta = float(torch.linalg.vector_norm(r))
            rel = beta / bnorm
            if rel <= tol:
                break
            mm = min(m, maxit - it)
            torch.div(r, beta, out=self.V[0])
            self.g.zero_()
            self.g[0] = beta
            k1, est = 0, rel
            k = 0
            while k < mm:
                kend = min(mm, k + check_every)
                for kk in range(k, kend):
                    self._run_step(kk, apply_A, precond, use_graph, key)
                est_all = (self.res[:kend] / bnorm).cpu()               # one synchronisation per batch
                hit = (est_all[k:kend] <= tol).nonzero()
                if len(hit):
                    k1 = k + int(hit[0]) + 1
                    est = float(est_all[k1 - 1])
                    break
                k1, est = kend, float(est_all[kend - 1])
                k = kend
            Zs = self.Z if precond is not None else self.V
            self._check(lib().bosss_hip_gmres_update_dev(self.ctx.h, self.n, k1, m, self.prec, self._p(Zs), self.n, self._p(self.H),
                                                         self._p(self.g), self._p(self.y), self._p(x), stream()))
            it += k1
            if not true_residual and est <= tol:
                return x, it, est
            apply_A(x, self.w)
            torch.sub(b, self.w, out=r)
        return x, it, float(torch.linalg.vector_norm(r)) / bnorm


def parab3p(lambdac, lambdam, ff0, ffc, ffm, sigma0=0.1, sigma1=0.5):
    """Three-point safeguarded parabolic step-length model (Newton.cs:1622-1656, after
    C. T. Kelley): the minimiser of the parabola through (0, ff0), (lambdac, ffc),
    (lambdam, ffm), clipped to [sigma0, sigma1] * lambdac; sigma1 * lambdac if the parabola
    has no minimum."""
    c2 = lambdam * (ffc - ff0) - lambdac * (ffm - ff0)
    if c2 >= 0:
        return sigma1 * lambdac
    c1 = lambdac * lambdac * (ffm - ff0) - lambdam * lambdam * (ffc - ff0)
    lp = -c1 * 0.5 / c2
    return min(max(lp, sigma0 * lambdac), sigma1 * lambdac)


def line_search(F, U, d, nf0, Ut, rt, alpha=1e-4, sigma1=0.5, max_step=30):
    """Armijo backtracking of LineSearch (Newton.cs:807-860): Ut <- U + lambda d with the first
    lambda in 1, sigma1, parab3p(...) ... for which ||F(Ut)|| < (1 - alpha lambda) ||F(U)||, or
    the last one tried after max_step reductions.  F(x, out) writes the residual; rt holds
    F(Ut) on return.  Returns (lambda, ||F(Ut)||, number of residual evaluations)."""
    import torch
    lam, lamm, lamc = 1.0, 1.0, 1.0
    torch.add(U, d, out=Ut)
    F(Ut, rt)
    nft = float(torch.linalg.vector_norm(rt))
    ff0, ffc, ffm = nf0 * nf0, nft * nft, nft * nft
    iarm, nev = 0, 1
    while not (nft < (1.0 - alpha * lam) * nf0) and iarm < max_step:
        lam = sigma1 * lam if iarm == 0 else parab3p(lamc, lamm, ff0, ffc, ffm)
        torch.add(U, d, alpha=lam, out=Ut)
        lamm, lamc = lamc, lam
        F(Ut, rt)
        nft = float(torch.linalg.vector_norm(rt))
        ffm, ffc = ffc, nft * nft
        iarm += 1
        nev += 1
    return lam, nft, nev


class ImplicitStep:
    """One implicit Euler step of the DG operator, solved by Newton-Krylov on the GPU."""

    def __init__(self, op64, Un, dt, S=None, op32=None, krylov="device", use_graph=True, check_every=8):
        """krylov: "device" (DeviceGmres, Arnoldi step on the GPU, optionally graph-captured) or
        "host" (gmres, Givens rotations on the host)."""
        import torch
        if krylov not in ("device", "host"):
            raise ValueError("krylov must be 'device' or 'host'")
        self.op64, self.op32 = op64, op32
        self.dt = float(dt)
        self.Un = Un
        self.S = S if S is not None else torch.zeros_like(Un)
        self.work = torch.empty_like(Un)
        self.stats = {}
        self.krylov, self.use_graph, self.check_every = krylov, use_graph, check_every
        # fixed linearisation points: captured Krylov steps keep reading these buffers
        self._Ulin64 = torch.empty_like(Un)
        self._Ulin32 = torch.empty(Un.shape, dtype=torch.float32, device=Un.device) if op32 is not None else None
        self._solvers = {}

    def _new_solve(self):
        """Forget the captured Krylov steps: they hold the buffers of the previous solve's
        preconditioner, which need not be alive any more."""
        for sv in self._solvers.values():
            sv.graphs, sv.graph_key = {}, None

    def _krylov(self, apply_A, b, tol, restart, maxit, precond, true_residual, key):
        """One linear solve with the configured GMRES.  Within one Newton solve the operator
        closures read fixed buffers and the preconditioner is one object, so captured steps are
        replayed across its linear solves."""
        if self.krylov == "host" or not b.is_cuda:        # (CPU tensors: the host restatement)
            return gmres(apply_A, b, tol, restart, maxit, precond=precond, true_residual=true_residual)
        sk = (b.dtype, restart, precond is not None)
        if sk not in self._solvers:
            self._solvers[sk] = DeviceGmres(b.numel(), b.dtype, b.device, restart, preconditioned=precond is not None,
                                            ctx=self.op64.ctx)
        return self._solvers[sk].solve(apply_A, b, tol, maxit, precond=precond, true_residual=true_residual,
                                       check_every=self.check_every, use_graph=self.use_graph, key=key + (id(precond),))

    def F(self, U, out):
        """out <- (U - Un)/dt + R(U) - S   (FP64)."""
        self.op64.evaluate_dev(U, out)
        out.add_(U - self.Un, alpha=1.0 / self.dt).sub_(self.S)
        return out

    def _jac(self, op, U):
        inv_dt = 1.0 / self.dt
        def apply(v, out):
            op.jvp_dev(U, v, out)
            out.add_(v, alpha=inv_dt)
        return apply

    def _jacT(self, op, U):
        inv_dt = 1.0 / self.dt
        def apply(v, out):
            op.vjp_dev(U, v, out)
            out.add_(v, alpha=inv_dt)
        return apply

    def _newton(self, U0, tol, max_newton, step, globalization):
        """Common loop of both drivers: step(U, r, rn) -> (correction d, Krylov iterations)."""
        import torch
        if globalization not in ("linesearch", "none"):
            raise ValueError("globalization must be 'linesearch' or 'none'")
        self._new_solve()
        U = U0.clone()
        r = torch.empty_like(U)
        Ut, rt = torch.empty_like(U), torch.empty_like(U)
        hist, lams, n_jvp, n_res = [], [], 0, 1
        t0 = time.perf_counter()
        self.F(U, r)
        rn = float(torch.linalg.vector_norm(r))
        hist.append(rn)
        for k in range(max_newton):
            if rn <= tol:
                break
            d, its = step(U, r, rn)
            n_jvp += its
            if globalization == "linesearch":
                lam, rn, nev = line_search(self.F, U, d, rn, Ut, rt)
                n_res += nev
                U, Ut = Ut, U
                r, rt = rt, r
            else:
                lam = 1.0
                U += d
                self.F(U, r)
                rn = float(torch.linalg.vector_norm(r))
                n_res += 1
            lams.append(lam)
            hist.append(rn)
        _sync(U)
        self.stats = dict(history=hist, jvps=n_jvp, residuals=n_res, step_lengths=lams, converged=rn <= tol,
                          seconds=time.perf_counter() - t0, newton=len(hist) - 1)
        return U

    def solve_fp64(self, U0, tol=1e-10, inner_tol=1e-4, max_newton=20, restart=40, maxit=400, precond=None,
                   globalization="linesearch"):
        """Newton-GMRES in FP64.  precond(v, z): optional right preconditioner, z ~ (I/dt + J)^-1 v
        (e.g. the cell-block Jacobi inverse of the assembled diagonal blocks, frozen).  Returns the
        last iterate; stats["converged"] is False if max_newton steps did not reach tol."""
        def step(U, r, rn):
            self._Ulin64.copy_(U)
            d, its, _ = self._krylov(self._jac(self.op64, self._Ulin64), -r, inner_tol, restart, maxit, precond, False,
                                     ("jvp64", id(self.op64)))
            return d, its
        return self._newton(U0, tol, max_newton, step, globalization)

    def solve_mixed(self, U0, tol=1e-10, inner_tol=1e-4, max_outer=20, restart=40, maxit=400, precond=None,
                    globalization="linesearch"):
        """FP64 defect correction around an FP32 Newton-GMRES correction.  precond(v, z) acts on
        the FP32 Krylov vectors."""
        if self.op32 is None:
            raise ValueError("solve_mixed needs an FP32 operator")

        def step(U, r, rn):
            self._Ulin32.copy_(U)
            b32 = (-r / rn).float()                        # scaled so FP32 never underflows
            d32, its, _ = self._krylov(self._jac(self.op32, self._Ulin32), b32, inner_tol, restart, maxit, precond, False,
                                       ("jvp32", id(self.op32)))
            return d32.double() * rn, its
        return self._newton(U0, tol, max_outer, step, globalization)

    def adjoint(self, U, g, tol=1e-10, restart=40, maxit=600):
        """Solve (I/dt + J(U))^T lam = g with the VJP kernel (FP64)."""
        self._new_solve()
        self._Ulin64.copy_(U)
        lam, its, rel = self._krylov(self._jacT(self.op64, self._Ulin64), g, tol, restart, maxit, None, True,
                                     ("vjp64", id(self.op64)))
        self.stats["adjoint_iterations"] = its
        self.stats["adjoint_rel_residual"] = rel
        return lam


def pressure_functional(op, U_t):
    """J(U) = integral of the pressure over the domain, and dJ/dU (FP64, torch on the device).

    Pressure p = (gamma - 1)(E - gamma Ma^2 |m|^2 / (2 rho)) (IdealGas.GetPressure); the
    integral uses the operator's volume rule: sum_j detJ_j sum_k w_k p(U_j(xi_k)).
    """
    import torch
    t = op.tables
    g = op.grid
    D, V, N = g.D, op.nvars, op.N
    gamma, mach = op.params["gamma"], op.params["mach"]
    ms = gamma * mach * mach
    dev = U_t.device
    phi = torch.from_numpy(t.phi).to(dev)                         # [K, N]
    w = torch.from_numpy(t.vw).to(dev)                            # [K]
    sc = torch.from_numpy(g.scaling).to(dev)                      # [J]
    detJ = torch.from_numpy(g.detJ).to(dev)
    C = U_t.view(g.J, V, N) * sc[:, None, None]
    Q = torch.einsum("jvn,kn->jkv", C, phi)                       # node states [J, K, V]
    rho, m, E = Q[..., 0], Q[..., 1:1 + D], Q[..., D + 1]
    msq = (m * m).sum(-1)
    p = (gamma - 1.0) * (E - 0.5 * ms * msq / rho)
    wk = w[None, :] * detJ[:, None]
    Jval = float((wk * p).sum())
    dp = torch.empty_like(Q)
    dp[..., 0] = (gamma - 1.0) * 0.5 * ms * msq / (rho * rho)
    dp[..., 1:1 + D] = -(gamma - 1.0) * ms * m / rho[..., None]
    dp[..., D + 1] = gamma - 1.0
    grad = torch.einsum("jkv,kn->jvn", dp * wk[..., None], phi) * sc[:, None, None]
    return Jval, grad.reshape(-1).contiguous()
