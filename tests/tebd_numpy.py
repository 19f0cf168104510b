"""Independent dense-linear-algebra TEBD (numpy/LAPACK SVD) of the upstream
algorithm (cuda/src/mps.cpp:41-145, 295-356): a third implementation used to
measure how well-conditioned a truncated run is."""
import math

import numpy as np


def _entropy0(sites):
    """Qubit-0 entropy (log2) of the MPS contracted to a dense vector (small n)."""
    psi = sites[0].reshape(2, -1)
    for t in sites[1:]:
        psi = np.einsum("xa,asb->xsb", psi, t).reshape(psi.shape[0] * 2, -1)
    psi = psi.reshape(2, -1)  # site 0 is the leading index
    rho = psi @ psi.conj().T
    ev = np.clip(np.linalg.eigvalsh(rho / np.trace(rho).real), 0, 1)
    return -sum(l * math.log2(l) for l in ev if l > 1e-12)


def tebd_numpy(model, sched, dt, chi, want_entropy=False, svd=None):
    """(max discarded weight, max bond after each step[, qubit-0 entropy trace]);
    `svd` replaces numpy's gesdd (e.g. scipy's gesvd driver) to measure how far
    two LAPACK factorisations move a truncated run."""
    svd = svd or (lambda a: np.linalg.svd(a, full_matrices=False))
    n = model.n
    sites = [np.zeros((1, 2, 1), complex) for _ in range(n)]
    for t in sites:
        t[0, 0, 0] = t[0, 1, 0] = 1 / math.sqrt(2)
    trunc = 0.0
    bonds = []
    ents = []
    swap = np.zeros((4, 4))
    swap[0, 0] = swap[1, 2] = swap[2, 1] = swap[3, 3] = 1

    def gate2(p, g):
        nonlocal trunc
        a, c = sites[p], sites[p + 1]
        dl, dr = a.shape[0], c.shape[2]
        th = np.einsum("asc,ctb->astb", a, c).reshape(dl, 4, dr)
        th = np.einsum("rs,asb->arb", g, th).reshape(dl * 2, 2 * dr)
        u, s, vh = svd(th)
        keep = min(chi, len(s))
        while keep > 1 and s[keep - 1] < 1e-14:
            keep -= 1
        tot = (s ** 2).sum()
        if tot > 0:
            trunc = max(trunc, (s[keep:] ** 2).sum() / tot)
        sites[p] = u[:, :keep].reshape(dl, 2, keep)
        sites[p + 1] = (s[:keep, None] * vh[:keep]).reshape(keep, 2, dr)

    def layer(A, B):
        for i in range(n):
            g = np.array([[-B * model.h[i], -A], [-A, B * model.h[i]]])
            w, v = np.linalg.eigh(g)
            u1 = v @ np.diag(np.exp(-1j * dt / 2 * w)) @ v.conj().T
            sites[i] = np.einsum("st,atb->asb", u1, sites[i])

    for A, B in zip(sched.a, sched.b):
        layer(A, B)
        for i, j, w in zip(model.ei, model.ej, model.ew):
            i, j = sorted((int(i), int(j)))
            if i == j:
                continue
            zz = np.diag(np.exp(-1j * (-B * w * dt) * np.array([1, -1, -1, 1])))
            for k in range(j, i + 1, -1):
                gate2(k - 1, swap)
            gate2(i, zz)
            for k in range(i + 1, j):
                gate2(k, swap)
        layer(A, B)
        bonds.append(max(t.shape[2] for t in sites))
        if want_entropy:
            ents.append(_entropy0(sites))
    if want_entropy:
        return trunc, bonds, np.array(ents)
    return trunc, bonds
