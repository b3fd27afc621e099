"""Pins the Davidson and learned-preconditioner oracles (CPU only)."""
import numpy as np
import pytest

from oracle import davidson_oracle as DO
from oracle import mlp_oracle as MO


def small_problem(n=300, naux=12, seed=0):
    rng = np.random.default_rng(seed)
    d = np.sort(rng.uniform(0.5, 5.0, n))
    B = 0.05 * rng.standard_normal((n, naux))
    A = np.diag(d) + 2 * B @ B.T
    return A


def run(A, nroots, precond=None, **kw):
    hd = np.diag(A).copy()
    x0 = np.zeros(len(hd))
    x0[np.argmin(hd)] = 1.0
    p = DO.DavidsonParams(tol=1e-12, maxcyc=300, nroots=nroots, **kw)
    return DO.davidson_solve(lambda xs: [A @ x for x in xs], x0, hd, p, precond=precond)


@pytest.mark.parametrize("nroots", [1, 3, 6])
def test_davidson_matches_dense_eigensolver(nroots):
    A = small_problem()
    r = run(A, nroots, maxspace=20)
    assert all(r.conv)
    w = np.linalg.eigvalsh(A)[:nroots]
    assert np.abs(r.e - w).max() < 1e-10
    for x, e in zip(r.x, r.e):
        assert np.linalg.norm(A @ x - e * x) < 1e-6
        assert abs(np.linalg.norm(x) - 1) < 1e-10


def test_davidson_permutation_equivariant():
    A = small_problem(seed=3)
    perm = np.random.default_rng(4).permutation(A.shape[0])
    Ap = A[np.ix_(perm, perm)]
    r1 = run(A, 4, maxspace=16)
    r2 = run(Ap, 4, maxspace=16)
    assert r1.iterations == r2.iterations and r1.matvecs == r2.matvecs
    for (i1, e1, d1), (i2, e2, d2) in zip(r1.history, r2.history):
        assert np.allclose(e1, e2, atol=1e-12)


def test_davidson_preconditioner_shift_quirk():
    # the reference shifts every root's correction by e[0] (davidson.rs:224)
    hd = np.array([1.0, 2.0, 3.0])
    r = np.array([1.0, 1.0, 1.0])
    assert np.allclose(DO.precond_diag(r, hd, 1.0), [1e8, 1.0, 0.5])


def test_davidson_rejects_too_many_roots():
    A = small_problem(n=8)
    with pytest.raises(ValueError):
        run(A, 6)


def test_qr_drops_dependent_vectors():
    v = np.array([1.0, 0.0, 0.0])
    q = DO.qr_mgs([v, 2 * v, np.array([1.0, 1.0, 0.0])], 1e-14)
    assert len(q) == 2 and abs(q[1] @ v) < 1e-15


# ---- learned preconditioner oracle --------------------------------------------
def mlp_samples(n=40, ns=3, seed=5, aux=False, proj=False):
    """Random samples (r, x0, hd, ek, e0[, a]), unit targets and, with proj,
    an orthonormal subspace per sample that the target is orthogonal to."""
    rng = np.random.default_rng(seed)
    hd = np.sort(rng.uniform(0.2, 3.0, n))
    out, taus, Qs = [], [], []
    for s in range(ns):
        x0 = rng.standard_normal(n)
        x0 /= np.linalg.norm(x0)
        r = 0.01 * rng.standard_normal(n)
        smp = (r, x0, hd, 0.3 + 0.1 * s, 0.3)
        if aux:
            smp = smp + (rng.standard_normal(n),)
        out.append(smp)
        tau = rng.standard_normal(n)
        Q = None
        if proj:
            Q = np.linalg.qr(np.column_stack([x0, rng.standard_normal((n, 2 + s))]))[0]
            tau = tau - Q @ (Q.T @ tau)
        Qs.append(Q)
        taus.append(tau / np.linalg.norm(tau))
    return out, taus, Qs


@pytest.mark.parametrize("aux,proj", [(False, False), (True, True)])
def test_mlp_gradient_finite_difference(aux, proj):
    samples, taus, Qs = mlp_samples(aux=aux, proj=proj)
    rng = np.random.default_rng(9)
    p = MO.init_params(3).astype(np.float64) + 0.1 * rng.standard_normal(MO.NPARAM)
    wts = np.array([1.0, 0.5, 1.0])
    Qs = Qs if proj else None
    L, g = MO.loss_grad(p, samples, taus, wts, Qs)
    idx = rng.choice(MO.NPARAM, size=40, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros(MO.NPARAM)
        e[i] = h
        fd = (MO.loss_grad(p + e, samples, taus, wts, Qs)[0] -
              MO.loss_grad(p - e, samples, taus, wts, Qs)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-8, i


def test_mlp_projected_loss_ignores_subspace_components():
    """Adding a multiple of a subspace vector to the correction changes
    neither the projected loss nor its value's meaning (Davidson removes it)."""
    samples, taus, Qs = mlp_samples(aux=True, proj=True)
    p = MO.init_params(2).astype(np.float64)
    L, _ = MO.loss_grad(p, samples, taus, np.ones(3), Qs)
    # a network whose output is the Ritz vector x0 (in every subspace) plus the
    # reference correction gives the same projected loss as the correction alone
    p2 = p.copy()
    p2[MO.H * MO.F + MO.H + MO.O * MO.H + 2] = 5.0     # c2[2]: weight of x0 sqrt(n)
    L2, _ = MO.loss_grad(p2, samples, taus, np.ones(3), Qs)
    assert abs(L - L2) < 1e-12


def test_mlp_initial_network_is_reference_preconditioner():
    samples, _, _ = mlp_samples(aux=True)
    p = MO.init_params(1)
    for (r, x0, hd, ek, e0, a) in samples:
        t, _ = MO.forward(p, r, x0, hd, ek, e0, a)
        ref = DO.precond_diag(r, hd, e0)
        cos = t @ ref / (np.linalg.norm(t) * np.linalg.norm(ref))
        assert cos > 1 - 1e-12
