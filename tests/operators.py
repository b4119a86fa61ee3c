"""Projection operators assembled from their mathematical definitions, for
pinning the oracle and the GPU solvers (scipy.sparse).

MAC: the cell-centred finite-volume operator div((1/rho) grad q): each
interior face f between cells a and b contributes k_f (q_b - q_a) / h^2 to
cell a; wall faces contribute nothing (pure Neumann). The face coefficients
come in the reference's irho layout (mgs_mac_varying_rho::irho,
fluid_2d.cc:1140-1160): entry 2(j m + i) is the face left of cell (i,j),
entry 2(j m + i)+1 the face below it.

FEM: the Galerkin stiffness matrix of bilinear elements, K_ij = sum over
cells c of k_c * integral over c of grad(phi_i).grad(phi_j), integrated
here by 2x2 Gauss quadrature (exact for these polynomials). Cell (ci,cj)
carries the coefficient the reference stores at its upper-right node,
irho[(cj+1) m + ci + 1] (fluid_2d.cc:1163-1170, mgs_fem.hh:99-106).
"""
import numpy as np
import scipy.sparse as sp


def mac_matrix(m, n, dx, dy, irho=None):
    rows, cols, vals = [], [], []
    kx, ky = 1.0 / dx**2, 1.0 / dy**2

    def add(a, b, k):
        rows.extend([a, a, b, b])
        cols.extend([b, a, a, b])
        vals.extend([k, -k, k, -k])

    for j in range(n):
        for i in range(m):
            c = j * m + i
            if i > 0:
                k = irho[2 * c] if irho is not None else 1.0
                add(c, c - 1, kx * k)
            if j > 0:
                k = irho[2 * c + 1] if irho is not None else 1.0
                add(c, c - m, ky * k)
    return sp.csr_matrix((vals, (rows, cols)), shape=(m * n, m * n))


def bilinear_element(dx, dy):
    """4x4 stiffness of one dx x dy bilinear element, nodes (0,0),(1,0),(0,1),(1,1)."""
    g = np.array([-1.0, 1.0]) / np.sqrt(3.0)
    K = np.zeros((4, 4))
    corners = [(0, 0), (1, 0), (0, 1), (1, 1)]
    for gx in g:
        for gy in g:
            s, t = 0.5 * (gx + 1), 0.5 * (gy + 1)
            grads = []
            for (a, b) in corners:
                fx = (1 if a else -1) * (t if b else 1 - t) / dx
                fy = (1 if b else -1) * (s if a else 1 - s) / dy
                grads.append((fx, fy))
            for p in range(4):
                for q in range(4):
                    K[p, q] += 0.25 * dx * dy * (grads[p][0] * grads[q][0] + grads[p][1] * grads[q][1])
    return K


def fem_matrix(m, n, dx, dy, irho=None):
    Ke = bilinear_element(dx, dy)
    rows, cols, vals = [], [], []
    for cj in range(n - 1):
        for ci in range(m - 1):
            k = irho[(cj + 1) * m + ci + 1] if irho is not None else 1.0
            nodes = [cj * m + ci, cj * m + ci + 1, (cj + 1) * m + ci, (cj + 1) * m + ci + 1]
            for p in range(4):
                for q in range(4):
                    rows.append(nodes[p])
                    cols.append(nodes[q])
                    vals.append(k * Ke[p, q])
    return sp.csr_matrix((vals, (rows, cols)), shape=(m * n, m * n))


def fem_matrix_sparse(m, n, dx, dy):
    """fem_matrix for constant density, assembled with array operations
    (for grids of a million nodes)."""
    Ke = bilinear_element(dx, dy)
    ci, cj = np.meshgrid(np.arange(m - 1), np.arange(n - 1))
    ci, cj = ci.ravel(), cj.ravel()
    nodes = [cj * m + ci, cj * m + ci + 1, (cj + 1) * m + ci, (cj + 1) * m + ci + 1]
    rows = np.concatenate([nodes[p] for p in range(4) for q in range(4)])
    cols = np.concatenate([nodes[q] for p in range(4) for q in range(4)])
    vals = np.concatenate([np.full(ci.size, Ke[p, q]) for p in range(4) for q in range(4)])
    return sp.csr_matrix((vals, (rows, cols)), shape=(m * n, m * n))


def neumann_solve(A, b):
    """Mean-free solution of a singular pure-Neumann system (one node pinned,
    then the mean removed)."""
    import scipy.sparse.linalg as spl
    N = A.shape[0]
    b = b - b.mean()
    A = A.tolil()
    A[0, :] = 0
    A[0, 0] = 1
    bb = b.copy()
    bb[0] = 0
    z = spl.spsolve(A.tocsr(), bb)
    return z - z.mean()


def mac_irho_from_rho(R, m, n):
    """Edge inverse densities from a cell density array R with one ghost
    layer, R[j+1, i+1] = rho(i,j), as set_edge_irho forms them
    (fluid_2d.cc:1140-1160)."""
    irho = np.zeros(2 * m * n + n + m)
    j, i = np.meshgrid(np.arange(n), np.arange(m), indexing="ij")
    ij = (j * m + i).ravel()
    irho[2 * ij] = (0.5 / R[j + 1, i + 1] + 0.5 / R[j + 1, i]).ravel()
    irho[2 * ij + 1] = (0.5 / R[j + 1, i + 1] + 0.5 / R[j, i + 1]).ravel()
    irho[2 * m * n: 2 * m * n + n] = 0.5 / R[1:n + 1, m] + 0.5 / R[1:n + 1, m + 1]
    irho[2 * m * n + n:] = 0.5 / R[n, 1:m + 1] + 0.5 / R[n + 1, 1:m + 1]
    return irho


def fem_irho_from_rho(R, m, n):
    """Corner inverse densities irho[j*(m+1)+i] = 1/rho(i-1,j-1) for the
    (m+1)x(n+1) node grid (fluid_2d.cc:1163-1170), plus one padding row."""
    mf, nf = m + 1, n + 1
    out = np.zeros(mf * nf + mf)
    out[: mf * nf] = (1.0 / R[0:nf, 0:mf]).ravel()
    return out


def blob_density(m, n, ratio=3.0, seed=None):
    """A cell density with one ghost layer: a disk of density `ratio` in a
    unit fluid, optionally with random smooth variations."""
    x = (np.arange(m + 2) - 0.5) / m
    y = (np.arange(n + 2) - 0.5) / n
    X, Y = np.meshgrid(x, y)
    R = 1 + (ratio - 1) * (np.sqrt((X - 0.45) ** 2 + (Y - 0.55) ** 2) < 0.3)
    if seed is not None:
        rng = np.random.default_rng(seed)
        R = R * (1 + 0.2 * np.sin(2 * np.pi * (X * rng.uniform(1, 3) + Y * rng.uniform(1, 3))))
    return R
