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


def mac_matrix(m, n, dx, dy, irho=None, px=False, py=False, zpr=False):
    """Row c gets k_f (q_nb - q_c) / h^2 for each of its faces. Along a
    periodic direction the wall faces become the seam, whose neighbour is
    the opposite edge's cell; the right side of the seam uses the
    right-edge coefficient irho[2mn + j] and the top side irho[2mn + n + i]
    (mgs_mac.hh:81-96), so the matrix is symmetric only when those agree
    with the left and bottom entries. zpr: q = 0 on the right edge face,
    half a cell from the last cell's centre, so that face contributes the
    flux k_f (0 - q_c) / (h^2 / 2) with k_f = irho[2mn + j]."""
    rows, cols, vals = [], [], []
    kx, ky = 1.0 / dx**2, 1.0 / dy**2
    mn = m * n

    def coef(e):
        return irho[e] if irho is not None else 1.0

    def face(c, nb, k):
        rows.extend([c, c])
        cols.extend([nb, c])
        vals.extend([k, -k])

    for j in range(n):
        for i in range(m):
            c = j * m + i
            if i > 0 or px:
                face(c, j * m + (i - 1) % m, kx * coef(2 * c))
            if i < m - 1 or px:
                face(c, j * m + (i + 1) % m, kx * coef(2 * (c + 1) if i < m - 1 else 2 * mn + j))
            elif zpr:
                rows.append(c)
                cols.append(c)
                vals.append(-2 * kx * coef(2 * mn + j))
            if j > 0 or py:
                face(c, ((j - 1) % n) * m + i, ky * coef(2 * c + 1))
            if j < n - 1 or py:
                face(c, ((j + 1) % n) * m + i, ky * coef(2 * (c + m) + 1 if j < n - 1 else 2 * mn + n + i))
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


def fem_matrix(m, n, dx, dy, irho=None, px=False, py=False, zpr=False):
    """m x n nodes; along a periodic direction there are m cells, the last
    one joining node m-1 to node 0 (mgs_fem.hh:99-106 wraps the corner
    lookup the same way). zpr: node column m-1 is a Dirichlet column held
    at zero: its rows keep only their diagonal, the other rows are those of
    the Galerkin matrix."""
    Ke = bilinear_element(dx, dy)
    rows, cols, vals = [], [], []
    for cj in range(n if py else n - 1):
        for ci in range(m if px else m - 1):
            i1, j1 = (ci + 1) % m, (cj + 1) % n
            k = irho[j1 * m + i1] if irho is not None else 1.0
            nodes = [cj * m + ci, cj * m + i1, j1 * m + ci, j1 * m + i1]
            for p in range(4):
                for q in range(4):
                    rows.append(nodes[p])
                    cols.append(nodes[q])
                    vals.append(k * Ke[p, q])
    A = sp.csr_matrix((vals, (rows, cols)), shape=(m * n, m * n))
    if zpr:
        A = A.tolil()
        for j in range(n):
            r = j * m + m - 1
            d = A[r, r]
            A[r, :] = 0
            A[r, r] = d
        A = A.tocsr()
    return A


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


def periodic_density(m, n, px, py, ratio=3.0, seed=1):
    """A smooth cell density with one ghost layer that is periodic along the
    periodic directions (ghost cells equal the opposite edge), so the MAC
    seam faces agree on both sides."""
    R = blob_density(m, n, ratio=ratio, seed=seed)
    if px:
        R[:, 0], R[:, m + 1] = R[:, m], R[:, 1]
    if py:
        R[0, :], R[n + 1, :] = R[n, :], R[1, :]
    return R


def fem_irho_periodic(R, m, n, px, py):
    """Corner inverse densities of an m x n node grid with m (n) cells along
    a periodic direction: node (i,j) holds 1/rho of cell (i-1,j-1), the
    ghost cell -1 standing for the seam cell (fluid_2d.cc:1163-1170)."""
    return (1.0 / R[0:n, 0:m]).ravel().copy()
