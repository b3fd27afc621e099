"""CPU oracle for the Davidson solver — TEST INFRASTRUCTURE ONLY.

A numpy restatement of davidson_solve (src/davidson.rs:48-266) and its helpers
_qr (:287-325), fill_heff (:327-390), _gen_x0 (:392-404), _normalize_xt_
(:268-285), the diagonal preconditioner (:63-70) and the subspace
eigenproblem (lapack_dspevx, matrix_blas_lapack.rs:1075-1095, here
numpy.linalg.eigh on the same symmetric heff).  The reference's quirks are
kept (davidson.rs:224 preconditions every root with e[0]; the "large |r|"
restore at :210 can never fire because :259 shadows max_dx_last; _sort_elast
at :406-432 returns its inputs unchanged).  One deviation: when _qr drops
vectors from a non-fresh expansion, the reference keeps using the pre-QR count
(davidson.rs:112-117, 134-136) and would index past xs; here the post-QR
count is used.

The product never imports this module.

Pinning: davidson_solve is pinned against numpy.linalg.eigh on small dense
matrices (the lowest nroots eigenvalues to 1e-9, residual norms below
sqrt(tol)) and by the property that the iteration history is unchanged by a
random orthogonal similarity of the problem; exact iteration-by-iteration
parity with the Rust code is unpinned (the reference cannot be built here).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class DavidsonParams:
    """DavidsonParams (davidson.rs:28-46) with its defaults."""
    tol: float = 1e-5
    maxcyc: int = 50
    maxspace: int = 12
    lindep: float = 1e-14
    nroots: int = 3


@dataclass
class DavidsonResult:
    conv: list
    e: np.ndarray
    x: list
    iterations: int = 0
    matvecs: int = 0
    history: list = field(default_factory=list)   # (icyc, e, dx_norm) per cycle


def precond_diag(dx: np.ndarray, hdiag: np.ndarray, e: float) -> np.ndarray:
    """davidson.rs:63-70: dx / (hdiag - e), |hdiag - e| < 1e-8 -> 1e-8."""
    h = hdiag - e
    h = np.where(np.abs(h) < 1e-8, 1e-8, h)
    return dx / h


def qr_mgs(x: list, lindep: float) -> list:
    """_qr (davidson.rs:287-325): modified Gram-Schmidt, dropping vectors with
    squared norm <= lindep after projection."""
    qs = []
    for xi in x:
        xi = xi.copy()
        for q in qs:
            prod = float(q @ xi)
            xi -= q * prod
        norm = float(np.sqrt(xi @ xi))
        if norm * norm > lindep:
            qs.append(xi / norm)
    return qs


def normalize_xt(xt: list, lindep: float):
    """_normalize_xt_ (davidson.rs:268-285)."""
    out, norm_min = [], 1.0
    for v in xt:
        norm = float(np.sqrt(v @ v))
        if norm * norm > lindep:
            out.append(v / norm)
            norm_min = min(norm_min, norm)
    return out, norm_min


def fill_heff(heff_old, xs, ax, nnew, fresh):
    """fill_heff (davidson.rs:327-390): extend heff by the rows/columns of the
    nnew newest vectors, heff[i,j] = xs[i] . ax[j] (symmetrized)."""
    space = len(ax)
    row0 = space - nnew
    heff = np.zeros((space, space))
    if not fresh and heff_old is not None:
        so = heff_old.shape[0]
        heff[:so, :so] = heff_old
    for i in range(row0, space):
        for j in range(row0, space):
            v = float(xs[i] @ ax[j])
            heff[i, j] = v
            heff[j, i] = v
    for j in range(row0, space):
        for i in range(row0):
            v = float(ax[i] @ xs[j])
            heff[i, j] = v
            heff[j, i] = v
    return heff


def davidson_solve(a_x, x0_1d, hdiag, params: DavidsonParams, precond=None,
                   max_dx_last_fixed=True, record=None) -> DavidsonResult:
    """davidson_solve (davidson.rs:48-266).  a_x maps a list of vectors to the
    list of A x.  precond(k, r_k, e, x0_k) -> t_k overrides the reference's
    diagonal preconditioner (which uses e[0] for every root).  record(icyc, e,
    x0, residuals, xs) is called once per cycle after the residuals."""
    tol, maxcyc, lindep, nroots = params.tol, params.maxcyc, params.lindep, params.nroots
    maxspace = params.maxspace + (nroots - 1) * 3
    n = len(x0_1d)
    x0 = [np.asarray(x0_1d, dtype=np.float64).copy()]
    fresh_start = True
    xt, xs, ax = [], [], []
    space = 0
    e = np.zeros(0)
    conv = [False] * nroots
    heff = None
    max_dx_last = 1e9
    res = DavidsonResult(conv=conv, e=e, x=x0)
    for icyc in range(maxcyc):
        if fresh_start:
            xs, ax, space = [], [], 0
            xt = qr_mgs(x0, lindep)
            if n - nroots < 3:
                raise ValueError("Too much nroots required")
            if len(xt) == 0:
                raise ValueError("No more linear independent basis found")
            max_dx_last = 1e9
        else:
            if len(xt) > 1:
                xt = qr_mgs(xt, lindep)
                xt = xt[:40]
        xtlen = len(xt)
        axt = a_x(xt)
        res.matvecs += xtlen
        xs.extend(xt)
        ax.extend(axt)
        space += xtlen
        elast = e.copy()
        heff = fill_heff(heff, xs, ax, xtlen, fresh_start)
        w, V = np.linalg.eigh(heff)
        if len(w) < nroots:
            e, v = w.copy(), V
        else:
            e, v = w[:nroots].copy(), V[:, :nroots]
        nroots_current = min(len(w), nroots)
        X = np.stack(xs, axis=1)
        AX = np.stack(ax, axis=1)
        x0 = list((X @ v).T)
        ax0 = list((AX @ v).T)
        xt = []
        dx_norm = np.zeros(nroots_current)
        de = np.zeros(nroots_current)
        for k in range(nroots_current):
            de[k] = e[k] - elast[k] if k < len(elast) else e[k]
            xt_k = ax0[k] - e[k] * x0[k]
            xt.append(xt_k)
            dx_norm[k] = float(np.sqrt(xt_k @ xt_k))
            conv[k] = bool(abs(de[k]) < tol and dx_norm[k] < np.sqrt(tol))
        res.history.append((icyc, e.copy(), dx_norm.copy()))
        if record is not None:
            record(icyc, e.copy(), x0[:nroots_current], xt, xs)
        res.iterations = icyc + 1
        if all(conv):
            break
        max_dx_norm = float(dx_norm.max())
        if max_dx_norm > 1.0 and max_dx_norm / max_dx_last > 3.0 and space > nroots + 2:
            # davidson.rs:210-217 (unreachable in practice: max_dx_last stays 1e9)
            x0 = list((X @ v).T)
            fresh_start = True
            continue
        xt_new = []
        for k in range(nroots_current):
            if dx_norm[k] ** 2 > lindep:
                if precond is None:
                    t = precond_diag(xt[k], hdiag, e[0])
                else:
                    t = precond(k, xt[k], e, x0[k])
                t = t / float(np.sqrt(t @ t))
                xt_new.append(t)
        xt = xt_new
        for i in range(space):
            for k in range(len(xt)):
                xt[k] = xt[k] - xs[i] * float(xs[i] @ xt[k])
        xt, _ = normalize_xt(xt, lindep)
        if len(xt) == 0:
            for k in range(len(dx_norm)):
                conv[k] = conv[k] or dx_norm[k] < np.sqrt(tol)
            break
        if not max_dx_last_fixed:
            max_dx_last = max_dx_norm
        fresh_start = space + nroots > maxspace
    res.conv, res.e, res.x = conv, e, x0
    return res


# ---- the RI linear-response operator of the Davidson workload -------------------
def casida_problem(B_full: np.ndarray, eps: np.ndarray, nocc: int, scale: float = 0.05):
    """A = diag(eps_a - eps_i) + 2 * Bov Bov^T with Bov[i + nocc*a, P] =
    scale * B_full[i, nocc + a, P] (B_full = ri3mo(occ, all, aux), [nocc, nmo,
    naux]).  Returns (dvec, Bov)."""
    nocc_, nmo, naux = B_full.shape
    assert nocc_ == nocc
    nvir = nmo - nocc
    Bov = scale * B_full[:, nocc:, :].reshape((nocc * nvir, naux), order="F")
    ei = eps[:nocc]
    ea = eps[nocc:]
    dvec = (ea[None, :] - ei[:, None]).reshape(-1, order="F")
    return dvec, Bov


def casida_eps(nbas: int, nocc: int, seed: int, gap: float = 0.3) -> np.ndarray:
    """Sorted N(0,1) orbital energies with the virtual block shifted up by gap."""
    from oracle.ri_oracle import gen_normal
    eps = np.sort(gen_normal(nbas, seed))
    eps[nocc:] += gap
    return eps
