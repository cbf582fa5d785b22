"""Test matrices shared by the CPU and GPU tests (plain scipy, seeded)."""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def laplacian_2d(n: int) -> sp.csr_matrix:
    """The 5-point Laplacian of tests/test_amg_pressure_solver.py:61-86 (reference)."""
    rows, cols, vals = [], [], []

    def idx(i, j):
        return i * n + j

    for i in range(n):
        for j in range(n):
            p = idx(i, j)
            d = 0.0
            for di, dj in ((-1, 0), (1, 0), (0, -1), (0, 1)):
                ii, jj = i + di, j + dj
                if 0 <= ii < n and 0 <= jj < n:
                    rows.append(p); cols.append(idx(ii, jj)); vals.append(-1.0); d += 1.0
            rows.append(p); cols.append(p); vals.append(d)
    return sp.csr_matrix((vals, (rows, cols)), shape=(n * n, n * n))


def apply_reference_fix(A: sp.csr_matrix, anchor: int) -> sp.csr_matrix:
    """Dirichlet anchor of classical_amg.py:660-686: drop row/column `anchor`
    except the diagonal, which becomes 1."""
    A = sp.coo_matrix(A)
    keep = ~((A.row == anchor) | (A.col == anchor)) | ((A.row == anchor) & (A.col == anchor))
    r, c, v = A.row[keep], A.col[keep], A.data[keep].copy()
    v[(r == anchor) & (c == anchor)] = 1.0
    if not np.any((r == anchor) & (c == anchor)):
        r = np.append(r, anchor); c = np.append(c, anchor); v = np.append(v, 1.0)
    out = sp.csr_matrix((v, (r, c)), shape=A.shape)
    out.sort_indices()
    return out


def impes_like(nx: int, ny: int, nz: int, seed: int = 0, wells: int = 2, sigma: float = 1.0):
    """Symmetric 7-point M-matrix shaped like an IMPES pressure matrix:
    lognormal face coefficients ~1e-10, accumulation ~1e-12 on the diagonal and
    `wells` BHP-like rows with a diagonal ~1e5. Returns (A csr, d, ox, oy, oz)
    with the stencil arrays in face form (z fastest)."""
    rng = np.random.default_rng(seed)
    n = nx * ny * nz
    shape = (nx, ny, nz)
    tx = 1e-10 * np.exp(sigma * rng.standard_normal(shape))
    ty = 1e-10 * np.exp(sigma * rng.standard_normal(shape))
    tz = 1e-11 * np.exp(sigma * rng.standard_normal(shape))
    tx[-1, :, :] = 0.0
    ty[:, -1, :] = 0.0
    tz[:, :, -1] = 0.0
    ox, oy, oz = -tx.ravel(), -ty.ravel(), -tz.ravel()
    d = np.full(n, 1.2e-12)
    sx, sy = ny * nz, nz
    idx = np.arange(n)
    d += tx.ravel() + ty.ravel() + tz.ravel()
    d[idx >= sx] += tx.ravel()[idx[idx >= sx] - sx]
    jj = (idx // nz) % ny
    d[jj > 0] += ty.ravel()[idx[jj > 0] - sy]
    kk = idx % nz
    d[kk > 0] += tz.ravel()[idx[kk > 0] - 1]
    for w in range(wells):
        c = int(rng.integers(0, n))
        d[c] += 1.8e5
    A = stencil_to_csr(nx, ny, nz, d, ox, oy, oz)
    return A, d, ox, oy, oz


def stencil_to_csr(nx, ny, nz, d, ox, oy, oz) -> sp.csr_matrix:
    n = nx * ny * nz
    sx, sy = ny * nz, nz
    idx = np.arange(n)
    i = idx // sx
    j = (idx // nz) % ny
    k = idx % nz
    r = [idx]
    c = [idx]
    v = [d]
    for off, arr, ok in ((sx, ox, i < nx - 1), (sy, oy, j < ny - 1), (1, oz, k < nz - 1)):
        a = idx[ok]
        r += [a, a + off]
        c += [a + off, a]
        v += [arr[a], arr[a]]
    A = sp.csr_matrix((np.concatenate(v), (np.concatenate(r), np.concatenate(c))), shape=(n, n))
    A.sort_indices()
    return A
