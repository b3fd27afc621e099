"""GPU Davidson, RI linear-response operator and learned preconditioner
against the oracles (oracle/davidson_oracle.py, oracle/mlp_oracle.py)."""
import numpy as np
import pytest

from oracle import davidson_oracle as DO
from oracle import mlp_oracle as MO
from oracle import ri_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def D():
    from rest_tensors_amd import device
    assert device.device_count() >= 1, "no GPU visible"
    return device


@pytest.fixture(scope="module")
def problem(D):
    """The Davidson workload's construction at the CPU-reference size:
    nbas=128, naux=384, nocc=16 -> n = 16*112 = 1792."""
    from rest_tensors_amd import davidson as DV
    nbas, naux, nocc = 128, 384, 16
    dA = D.DeviceArray(nbas * nbas * naux)
    D.gen_ri3ao(dA, nbas, naux, seed=11)
    Cm = O.gen_orthonormal(nbas, seed=12)
    dC = D.DeviceArray.from_host(Cm)
    dB = D.DeviceArray(nocc * nbas * naux)
    D.ao2mo_occ(dC, nbas, nbas, nocc, dA, naux, dB)
    Bf = dB.to_host((nocc, nbas, naux))
    eps = DO.casida_eps(nbas, nocc, seed=13)
    dvec, Bov = DO.casida_problem(Bf, eps, nocc)
    n = len(dvec)
    dd = D.DeviceArray.from_host(dvec)
    op = DV.CasidaOp(dB, nocc * nbas, n, naux, 0.05, dd, offset=nocc * nocc)
    A = np.diag(dvec) + 2 * Bov @ Bov.T
    hd = D.DeviceArray(n)
    op.hdiag(hd)
    return dict(op=op, A=A, n=n, hd=hd, dvec=dvec, Bov=Bov, keep=(dA, dC, dB, dd))


def test_casida_matvec_and_diagonal(D, problem):
    A, n, op = problem["A"], problem["n"], problem["op"]
    assert np.abs(problem["hd"].to_host() - np.diag(A)).max() < 1e-13
    rng = np.random.default_rng(1)
    for k in (1, 5, 8, 11):
        X = rng.standard_normal((n, k))
        dX, dY = D.DeviceArray.from_host(X), D.DeviceArray(n * k)
        op.apply(dX, dY, k)
        Y = dY.to_host((n, k))
        ref = A @ X
        assert np.linalg.norm(Y - ref) / np.linalg.norm(ref) < 1e-13


@pytest.mark.parametrize("n,ldb,offset", [(302, 302, 0), (301, 303, 0), (302, 304, 1),
                                          (4098, 4100, 0)])
def test_casida_matvec_layouts(D, n, ldb, offset):
    """A = diag(d) + 2 s^2 B B^T against the dense product for even strides
    (16-byte passes) and odd or misaligned ones (8-byte passes)."""
    from rest_tensors_amd import davidson as DV
    rng = np.random.default_rng(n + ldb + offset)
    naux, scale = 37, 0.05
    Bfull = rng.standard_normal(offset + ldb * naux)
    B = Bfull[offset:offset + ldb * naux].reshape((ldb, naux), order="F")[:n]
    d = rng.uniform(0.5, 2.0, n)
    dB, dd = D.DeviceArray.from_host(Bfull), D.DeviceArray.from_host(d)
    op = DV.CasidaOp(dB, ldb, n, naux, scale, dd, offset=offset)
    A = np.diag(d) + 2 * scale ** 2 * B @ B.T
    for k in (1, 3, 8, 10):
        X = rng.standard_normal((n, k))
        dX, dY = D.DeviceArray.from_host(X), D.DeviceArray(n * k)
        op.apply(dX, dY, k)
        ref = A @ X
        assert np.linalg.norm(dY.to_host((n, k)) - ref) / np.linalg.norm(ref) < 1e-13


def x0_guess(D, hd_host):
    x0 = np.zeros(len(hd_host))
    x0[np.argmin(hd_host)] = 1.0
    return D.DeviceArray.from_host(x0), x0


@pytest.mark.parametrize("precond", ["diag_ref", "diag_root"])
def test_davidson_matches_oracle(D, problem, precond):
    from rest_tensors_amd import davidson as DV
    A, n, op, hd = problem["A"], problem["n"], problem["op"], problem["hd"]
    hdh = hd.to_host()
    dx0, x0 = x0_guess(D, hdh)
    nroots = 8
    res = DV.solve(op.matvec_fn, op.handle, dx0, hd, n, nroots, tol=1e-12, maxcyc=300,
                   maxspace=12, precond=precond)
    pc = None if precond == "diag_ref" else \
        (lambda k, rk, e, x: DO.precond_diag(rk, hdh, e[k]))
    ref = DO.davidson_solve(lambda xs: [A @ x for x in xs], x0, hdh,
                            DO.DavidsonParams(tol=1e-12, maxcyc=300, nroots=nroots,
                                              maxspace=12), precond=pc)
    assert all(res.conv) and all(ref.conv)
    assert abs(res.stats["iterations"] - ref.iterations) <= 1
    w = np.linalg.eigvalsh(A)[:nroots]
    assert np.abs(res.e - w).max() < 1e-10
    X = res.x.to_host((n, nroots))
    resid = np.linalg.norm(A @ X - X * res.e[None, :], axis=0)
    assert resid.max() < 1e-6     # the residual judges convergence
    print("iterations gpu/oracle:", res.stats["iterations"], ref.iterations,
          "matvecs:", res.stats["matvecs"], ref.matvecs, "seconds:", res.stats["seconds"])


def test_davidson_recorder_counts(D, problem):
    from rest_tensors_amd import davidson as DV
    n, op, hd = problem["n"], problem["op"], problem["hd"]
    dx0, _ = x0_guess(D, hd.to_host())
    rec = DV.Recorder(n, 2000)
    res = DV.solve(op.matvec_fn, op.handle, dx0, hd, n, 4, tol=1e-10, maxcyc=200,
                   maxspace=12, recorder=rec)
    assert all(res.conv)
    assert rec.count >= res.stats["iterations"]
    assert set(rec.root[:rec.count].tolist()) <= set(range(4))
    # a recorded residual is A x - e x of its recorded Ritz vector
    q = rec.count - 1
    r = rec.r.to_host()[q * n:(q + 1) * n]
    x = rec.x.to_host()[q * n:(q + 1) * n]
    A = problem["A"]
    assert np.linalg.norm(A @ x - rec.e_k[q] * x - r) < 1e-12


def test_shift_invert_matches_dense_solve(D, problem):
    """(A - sigma)^-1 R through the Woodbury form against a dense solve."""
    from rest_tensors_amd import davidson as DV
    A, n, op = problem["A"], problem["n"], problem["op"]
    si = DV.ShiftInvert(op, shift=0.05)
    info = si.info()
    assert abs(info["sigma"] - (problem["dvec"].min() - 0.05)) < 1e-15
    assert info["residual"] < 1e-8
    rng = np.random.default_rng(4)
    for k in (1, 3, 8, 13):
        R = rng.standard_normal((n, k))
        dR, dT = D.DeviceArray.from_host(R), D.DeviceArray(n * k)
        si.apply(dR, dT, k)
        ref = np.linalg.solve(A - info["sigma"] * np.eye(n), R)
        T = dT.to_host((n, k))
        assert np.linalg.norm(T - ref) / np.linalg.norm(ref) < 1e-8
    print("capacitance inverse residual", info["residual"])


def test_davidson_shift_invert_preconditioner(D, problem):
    from rest_tensors_amd import davidson as DV
    A, n, op, hd = problem["A"], problem["n"], problem["op"], problem["hd"]
    dx0, _ = x0_guess(D, hd.to_host())
    si = DV.ShiftInvert(op, shift=0.05)
    res = DV.solve(op.matvec_fn, op.handle, dx0, hd, n, 8, tol=1e-12, maxcyc=300,
                   maxspace=12, precond="aux", aux=si)
    assert all(res.conv)
    w = np.linalg.eigvalsh(A)[:8]
    assert np.abs(res.e - w).max() < 1e-10
    X = res.x.to_host((n, 8))
    assert np.linalg.norm(A @ X - X * res.e[None, :], axis=0).max() < 1e-6
    print("shift-invert iterations", res.stats["iterations"])


def mlp_batch(D, problem, ns=6, seed=3):
    n, hd = problem["n"], problem["hd"].to_host()
    rng = np.random.default_rng(seed)
    R = 1e-3 * rng.standard_normal((n, ns))
    X = rng.standard_normal((n, ns))
    X /= np.linalg.norm(X, axis=0)
    T = rng.standard_normal((n, ns))
    T /= np.linalg.norm(T, axis=0)
    S = rng.standard_normal((n, ns))
    ek = 0.4 + 0.01 * np.arange(ns)
    e0 = np.full(ns, 0.4)
    return R, X, T, S, ek, e0, hd


def test_mlp_forward_matches_oracle(D, problem):
    from rest_tensors_amd import davidson as DV
    n = problem["n"]
    R, X, _, S, ek, e0, hd = mlp_batch(D, problem)
    ns = R.shape[1]
    net = DV.Mlp(seed=1)
    p0 = net.get_params()
    assert np.array_equal(p0, MO.init_params(1))
    rng = np.random.default_rng(2)
    for p in (p0, p0 + 0.2 * rng.standard_normal(p0.size).astype(np.float32)):
        net.set_params(p)
        for aux in (None, S):
            dT = D.DeviceArray(n * ns)
            net.apply(D.DeviceArray.from_host(R), D.DeviceArray.from_host(X), problem["hd"],
                      ek, e0, n, ns, dT,
                      S=None if aux is None else D.DeviceArray.from_host(aux))
            T = dT.to_host((n, ns))
            for s in range(ns):
                ref, _ = MO.forward(p, R[:, s], X[:, s], hd, ek[s], e0[s],
                                    None if aux is None else aux[:, s])
                # FP32 network on the device, FP64 oracle
                assert np.linalg.norm(T[:, s] - ref) / np.linalg.norm(ref) < 1e-5


def test_mlp_gradient_matches_oracle(D, problem):
    """Projected loss and gradient (subspaces of 2..6 vectors, one sample
    without) with the auxiliary basis vector, against the FP64 oracle."""
    from rest_tensors_amd import davidson as DV
    n = problem["n"]
    ns = 5
    R, X, Tau, S, ek, e0, hd = mlp_batch(D, problem, ns=ns, seed=8)
    rng = np.random.default_rng(6)
    Qs, blocks, off, cnt, pos = [], [], [], [], 0
    for s in range(ns):
        if s == 2:
            Qs.append(None)
            off.append(-1)
            cnt.append(0)
            continue
        m = 2 + s
        Q = np.linalg.qr(np.column_stack([X[:, s], rng.standard_normal((n, m - 1))]))[0]
        Tau[:, s] -= Q @ (Q.T @ Tau[:, s])
        Tau[:, s] /= np.linalg.norm(Tau[:, s])
        Qs.append(Q)
        blocks.append(Q.reshape(-1, order="F"))
        off.append(pos)
        cnt.append(m)
        pos += n * m
    dQ = D.DeviceArray.from_host(np.concatenate(blocks))
    wts = np.array([1.0, 1.0, 0.7, 0.5, 1.0])
    net = DV.Mlp(seed=4)
    p = net.get_params() + 0.1 * np.random.default_rng(5).standard_normal(
        net.nparam).astype(np.float32)
    net.set_params(p)
    L, g = net.loss_grad(D.DeviceArray.from_host(R), D.DeviceArray.from_host(X),
                         problem["hd"], ek, e0, D.DeviceArray.from_host(Tau), wts, n, ns,
                         S=D.DeviceArray.from_host(S), q=(dQ, off, cnt))
    samples = [(R[:, s], X[:, s], hd, ek[s], e0[s], S[:, s]) for s in range(ns)]
    Lr, gr = MO.loss_grad(p, samples, [Tau[:, s] for s in range(ns)], wts, Qs)
    assert abs(L - Lr) < 1e-5 * max(1.0, abs(Lr))
    assert np.linalg.norm(g - gr) / np.linalg.norm(gr) < 1e-3


def test_mlp_training_and_preconditioned_solve(D, problem):
    """Record plain-Davidson samples (with their subspaces), train the network
    on the GPU with the shift-invert correction in its basis, then solve with
    it: same eigenvalues, the residual still the judge, and fewer iterations
    than the plain diagonal preconditioner."""
    from rest_tensors_amd import davidson as DV
    A, n, op, hd = problem["A"], problem["n"], problem["op"], problem["hd"]
    dx0, _ = x0_guess(D, hd.to_host())
    si = DV.ShiftInvert(op, shift=0.05)
    rec = DV.Recorder(n, 600, subspace_doubles=n * 40 * 120)
    plain = DV.solve(op.matvec_fn, op.handle, dx0, hd, n, 8, tol=1e-12, maxcyc=300,
                     maxspace=12, recorder=rec)
    ns = rec.count
    assert (rec.xs_off[:ns] >= 0).sum() > 10
    tau = D.DeviceArray(n * ns)
    wts = DV.mlp_targets(rec.x, rec.root[:ns], ns, plain.x, n, tau, recorder=rec)
    assert wts.sum() > 10
    Sa = D.DeviceArray(n * ns)
    si.apply(rec.r, Sa, ns)
    net = DV.Mlp(seed=1)
    args = (rec.r, rec.x, hd, rec.e_k[:ns], rec.e_0[:ns], tau, wts, n, ns)
    L0, _ = net.loss_grad(*args, S=Sa, recorder=rec)
    hist = net.train(*args, steps=300, batch=32, lr=3e-3, S=Sa, recorder=rec)
    L1, _ = net.loss_grad(*args, S=Sa, recorder=rec)
    print("training loss", L0, "->", L1, "(first/last minibatch", hist[0], hist[-1], ")")
    assert L1 < L0
    res = DV.solve(op.matvec_fn, op.handle, dx0, hd, n, 8, tol=1e-12, maxcyc=300,
                   maxspace=12, precond="mlp", mlp=net, aux=si)
    assert all(res.conv)
    w = np.linalg.eigvalsh(A)[:8]
    assert np.abs(res.e - w).max() < 1e-10
    X = res.x.to_host((n, 8))
    assert np.linalg.norm(A @ X - X * res.e[None, :], axis=0).max() < 1e-6
    sio = DV.solve(op.matvec_fn, op.handle, dx0, hd, n, 8, tol=1e-12, maxcyc=300,
                   maxspace=12, precond="aux", aux=si)
    print("iterations: plain", plain.stats["iterations"], "shift-invert only",
          sio.stats["iterations"], "mlp", res.stats["iterations"])
    assert res.stats["iterations"] < plain.stats["iterations"]
