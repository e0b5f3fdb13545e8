"""Host logic of the steady adjoint's GMRES (dndsr_amd.adjoint.SteadyAdjoint._gmres): with the
transposed operator and the preconditioner replaced by dense matrices on the CPU, the
multipliers must solve A^T x = -b to the tolerance, with and without restarts, from zero
and from a starting guess."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from dndsr_amd.adjoint import SteadyAdjoint


def _mock(seed=0, N=40, R=30):
    rng = np.random.default_rng(seed)
    n = 5 * N + R
    A = -2.0 * np.eye(n) + 0.3 * rng.normal(size=(n, n)) / np.sqrt(n)
    adj = SteadyAdjoint.__new__(SteadyAdjoint)
    At = torch.tensor(A.T)

    def apply(u, rec, lam, mu):
        y = At @ torch.cat([lam.reshape(-1), mu.reshape(-1)])
        return y[: 5 * N].view(5, N), y[5 * N:]

    # the u block of A^T is about -2 I: P^-T = I / 2 plays the LU-SGS sweeps
    adj._apply = apply
    adj._precond = lambda r, u, jd: 0.5 * r
    u, rec = torch.zeros(5, N, dtype=torch.float64), torch.zeros(R, dtype=torch.float64)
    bu, br = torch.tensor(rng.normal(size=(5, N))), torch.tensor(rng.normal(size=R))
    ref = np.linalg.solve(A.T, -np.concatenate([bu.reshape(-1).numpy(), br.numpy()]))
    return adj, u, rec, bu, br, ref, N


@pytest.mark.parametrize("restart", [4, 40])
def test_gmres_solves_the_transposed_system(restart):
    adj, u, rec, bu, br, ref, N = _mock()
    lam, mu, it = adj._gmres(u, rec, bu, br, None, 1e-12, 500, restart, None, None)
    x = np.concatenate([lam.reshape(-1).numpy(), mu.numpy()])
    assert np.abs(x - ref).max() <= 1e-10 * np.abs(ref).max()
    assert adj.history[-1] <= 1e-12 * np.linalg.norm(np.concatenate([bu.reshape(-1).numpy(), br.numpy()]))
    # a start at the solution needs no iteration
    lam2, mu2, it2 = adj._gmres(u, rec, bu, br, None, 1e-9, 500, restart, lam, mu)
    assert it2 == 0 and torch.equal(lam2, lam)
