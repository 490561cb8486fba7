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

GMRES here is restarted, right-preconditioned (optional block-diagonal
preconditioner callable), classical Gram-Schmidt with one re-orthogonalisation,
in the dtype of its right-hand side; Givens rotations on the host.
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

    def __init__(self, op64, Un, dt, S=None, op32=None):
        import torch
        self.op64, self.op32 = op64, op32
        self.dt = float(dt)
        self.Un = Un
        self.S = S if S is not None else torch.zeros_like(Un)
        self.work = torch.empty_like(Un)
        self.stats = {}

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
            d, its, _ = gmres(self._jac(self.op64, U), -r, inner_tol, restart, maxit, precond=precond, true_residual=False)
            return d, its
        return self._newton(U0, tol, max_newton, step, globalization)

    def solve_mixed(self, U0, tol=1e-10, inner_tol=1e-4, max_outer=20, restart=40, maxit=400, precond=None,
                    globalization="linesearch"):
        """FP64 defect correction around an FP32 Newton-GMRES correction.  precond(v, z) acts on
        the FP32 Krylov vectors."""
        if self.op32 is None:
            raise ValueError("solve_mixed needs an FP32 operator")

        def step(U, r, rn):
            U32 = U.float()
            b32 = (-r / rn).float()                        # scaled so FP32 never underflows
            d32, its, _ = gmres(self._jac(self.op32, U32), b32, inner_tol, restart, maxit, precond=precond,
                                true_residual=False)
            return d32.double() * rn, its
        return self._newton(U0, tol, max_outer, step, globalization)

    def adjoint(self, U, g, tol=1e-10, restart=40, maxit=600):
        """Solve (I/dt + J(U))^T lam = g with the VJP kernel (FP64)."""
        lam, its, rel = gmres(self._jacT(self.op64, U), g, tol, restart, maxit)
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
