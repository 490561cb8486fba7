"""Synthetic workloads of the benchmark configurations (seeded, deterministic).

Euler states follow the stated distribution: per cell, density ~ U[0.8, 1.2],
velocity ~ U[-0.1, 0.1]^D, pressure ~ U[0.9, 1.1]; converted to conservative
variables (rho, rho u, E = p/(gamma-1) + gamma Ma^2 rho |u|^2 / 2) and written as
the cell-mean (mode 0) DG coefficient, plus a seeded perturbation of 1% of the
mean on the higher modes so the nodal states vary smoothly inside each cell.
"""
import numpy as np


def euler_state(grid, N, seed=0, gamma=1.4, mach=None, perturb=0.01):
    D, J = grid.D, grid.J
    mach = 1.0 / np.sqrt(gamma) if mach is None else mach
    ms = gamma * mach * mach
    rng = np.random.default_rng(seed)
    V = D + 2
    rho = rng.uniform(0.8, 1.2, J)
    vel = rng.uniform(-0.1, 0.1, (J, D))
    p = rng.uniform(0.9, 1.1, J)
    E = p / (gamma - 1) + 0.5 * ms * rho * (vel ** 2).sum(1)
    # coefficient of the constant function 1 on mode 0: c0 = sqrt(|K|) = sqrt(detJ 2^D)
    s0 = np.sqrt(grid.detJ * 2.0 ** D)
    U = np.zeros((J, V, N))
    U[:, 0, 0] = rho * s0
    U[:, 1:1 + D, 0] = rho[:, None] * vel * s0[:, None]
    U[:, D + 1, 0] = E * s0
    if N > 1 and perturb:
        U[:, :, 1:] = perturb * rng.uniform(-1, 1, (J, V, N - 1)) * U[:, :, :1]
    return U.reshape(-1)


def uniform_coefficients(ndof, seed=0):
    """Coefficient vector i.i.d. uniform[-1, 1] (configs 1 and 2)."""
    return np.random.default_rng(seed).uniform(-1.0, 1.0, ndof)


def contraction_flops(form, D, p, J, E):
    """Dense modal<->nodal contraction flops of one operator apply (reference formulation).

    per cell: 2 V N Kv (a_val + D a_grad + t_val + D t_grad)
    per edge: 4 V N Ke (e_val + D e_grad + et_val + D et_grad)
    """
    from .refelement import num_modes, gauss_points_for_order
    from .operator import default_quad_order
    N = num_modes(D, p)
    K1 = gauss_points_for_order(default_quad_order(form, p))
    Kv, Ke = K1 ** D, K1 ** (D - 1)
    if form.startswith("euler"):
        V = D + 2
        vol = 1 + D          # values in, test gradients
        edge = 1 + 1         # values in, test values
    elif form == "sip":
        V = 1
        vol = D + D          # gradients in, test gradients
        edge = (1 + D) + (1 + D)
    else:  # advdiff: upwind (values) + SIP (gradients)
        V = 1
        vol = (1 + D) + D
        edge = (1 + D) + (1 + D)
    return 2 * V * N * Kv * vol * J + 4 * V * N * Ke * edge * E
