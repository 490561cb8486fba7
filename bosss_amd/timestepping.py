"""Explicit Runge-Kutta time stepping on the device-resident operator (host mirror of
BoSSS.Solution.TimeStepping.ExplicitEuler / RungeKutta).

Convention of the reference: the operator gives the change rate R(u) of du/dt + R(u) = 0.
`ComputeChangeRate(k, U)` accumulates k += R(U) (`ExplicitEuler.cs:223-229`, an
Evaluate(1, 1, k) of the evaluator); a stage sets y = y0 - dt sum_r a[s][r] k[r]
(`RungeKutta.cs:297-310`), the step ends with y0 - dt sum_s b[s] k[s] (`RungeKutta.cs:242-285`).
Every vector stays on the GPU; one residual launch per stage.
"""


class RungeKuttaScheme:
    """Butcher tableau (a strictly lower triangular, b, c), as RungeKuttaScheme.cs."""

    def __init__(self, a, b, c):
        self.a, self.b, self.c = a, b, c
        self.stages = len(b)


# ExplicitEuler; TVD3 of Gottlieb and Shu (Math. Comp. 67 (1998) 73-85), the reference's
# RungeKuttaScheme.TVD3; the classic fourth-order scheme.
EXPLICIT_EULER = RungeKuttaScheme([[0.0]], [1.0], [0.0])
TVD3 = RungeKuttaScheme([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.25, 0.25, 0.0]], [1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0], [0.0, 1.0, 0.5])
RK4 = RungeKuttaScheme([[0.0] * 4, [0.5, 0.0, 0.0, 0.0], [0.0, 0.5, 0.0, 0.0], [0.0, 0.0, 1.0, 0.0]],
                       [1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0], [0.0, 0.5, 0.5, 1.0])


class RungeKutta:
    """Explicit Runge-Kutta stepper over a DifferentialOperator's device evaluate."""

    def __init__(self, op, U, scheme=TVD3):
        import torch
        self.op, self.scheme = op, scheme
        self.U = U                                   # current state, updated in place
        self.time = 0.0
        self.k = [torch.empty_like(U) for _ in range(scheme.stages)]
        self.y0 = torch.empty_like(U)

    def ComputeChangeRate(self, k, U):
        """k += R(U)  (ExplicitEuler.ComputeChangeRate)."""
        self.op.evaluate_dev(U, k, alpha=1.0, beta=1.0)

    def Perform(self, dt):
        """One time step of size dt; returns dt (RungeKutta.Perform)."""
        sch = self.scheme
        self.y0.copy_(self.U)
        for s in range(sch.stages):
            if s > 0:
                # PerformStage: U = y0 - dt sum_r a[s][r] k[r]
                self.U.copy_(self.y0)
                for r in range(s):
                    if sch.a[s][r] != 0.0:
                        self.U.add_(self.k[r], alpha=-dt * sch.a[s][r])
            self.k[s].zero_()
            self.ComputeChangeRate(self.k[s], self.U)
        self.U.copy_(self.y0)
        for s in range(sch.stages):
            if sch.b[s] != 0.0:
                self.U.add_(self.k[s], alpha=-dt * sch.b[s])
        self.time += dt
        return dt
