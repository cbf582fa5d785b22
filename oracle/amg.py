"""ORACLE (test infrastructure only): FP64 numpy/scipy restatement of the
reference's Ruge-Stueben AMG and of the FGMRES method that replaces its
stationary V-cycle iteration.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
may import this module, and only as the checker.

Reference: src/solver/classical_amg.py (file:line cited per function).
The reference's unseeded ``torch.rand`` C/F tie-break (classical_amg.py:125) is
restated as the deterministic (degree, hash(index), index) order the GPU uses,
so C/F splittings and interpolation patterns can be compared exactly.
The reference's energy-minimising blend of interpolation weights
(classical_amg.py:389-418) indexes the fine-level basis with coarse column
numbers; it is not restated (documented deviation).

Six preconditioner choices are not in the reference and default to values
that differ from it (``REFERENCE`` holds the reference behaviour); they change
how fast the outer solver converges, not the converged answer:
  equil_mode  1 = scale the fine operator by 1/median|a_ii| (frozen at setup)
              instead of the symmetric D^-1/2 A D^-1/2 (mode 0). The symmetric
              form moves the near-nullspace to D^1/2 * 1 while the row-normalised
              interpolation reproduces only constants.
  coarse_pin  2 = pin DOF 0 of the coarsest level only when that operator is
              singular (every row |row sum| <= 1e-8 |a_ii|); 1 = reference.
  trunc/pmax  truncation of each F row of P (weights below trunc * row max,
              then all but the pmax largest), rescaled to the row's sum.
  coarsening  1 = iterated independent-set C/F split (rounds of the same
              local-max selection among undecided points); 0 = the reference's
              single pass with lift, which coarsens poorly on dense deep levels.
  dominant    rows with |a_ii| > 10 sum|a_ij| (BHP wells) are C points.
  pre_smooth/post_smooth/omega
              V(2,2) L1-Jacobi with omega 0.8 instead of V(3,2) with 0.7: as
              many outer iterations on the IMPES matrices, one sweep fewer per level.
These are pinned by invariants in tests/test_oracle_amg.py (scaling
invariance, row sums kept, pin decision on Neumann vs. nonsingular operators,
a maximal independent set checked round by round).

Parity pinning of each function (checked in tests/test_oracle_amg.py):
  strength            brute-force per-row loop on small matrices
  cf_split            brute-force definition (strict local maxima + lift)
  interpolation       P*1 = 1 on F rows, C rows identity, P >= 0, brute-force
                      per-row restatement on small matrices
  galerkin            dense R*A*P
  vcycle / setup      reference KAT: 12x12 anchored Laplacian, 3 cycles of
                      apply() reach ||Az-v||/||v|| < 5e-2 (test_amg_pressure_solver.py:89-106)
  fgmres              scipy.sparse.linalg.spsolve on the same system (both
                      orthogonalisation variants)
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

C_PT, F_PT = 0, 1


def hash32(x: np.ndarray) -> np.ndarray:
    """lowbias32 integer hash (same constants as the HIP kernel)."""
    x = np.asarray(x, dtype=np.uint64) & np.uint64(0xFFFFFFFF)
    m = np.uint64(0xFFFFFFFF)
    x ^= x >> np.uint64(16)
    x = (x * np.uint64(0x7FEB352D)) & m
    x ^= x >> np.uint64(15)
    x = (x * np.uint64(0x846CA68B)) & m
    x ^= x >> np.uint64(16)
    return x.astype(np.uint32)


def _rows(A: sp.csr_matrix) -> np.ndarray:
    return np.repeat(np.arange(A.shape[0]), np.diff(A.indptr))


def _diag_first(A: sp.csr_matrix) -> np.ndarray:
    """a_ii (first diagonal entry of each row, 0 if absent)."""
    rows = _rows(A)
    m = A.indices == rows
    d = np.zeros(A.shape[0])
    idx = np.nonzero(m)[0][::-1]  # reversed so the first occurrence is written last
    d[rows[idx]] = A.data[idx]
    return d


def strength(A: sp.csr_matrix, theta: float) -> np.ndarray:
    """classical_amg.py:34-82 — strong(i,j) iff a_ij<0, j!=i and
    -a_ij/|a_ii| >= theta * max_k(-a_ik/|a_ii|)."""
    rows = _rows(A)
    d = np.abs(_diag_first(A))
    sv = A.data / (d[rows] + 1e-30)
    cand = (A.indices != rows) & (sv < 0)
    neg_abs = np.where(cand, -sv, 0.0)
    mx = np.zeros(A.shape[0])
    np.maximum.at(mx, rows, neg_abs)
    return cand & (neg_abs >= theta * mx[rows])


def transpose_pos(A: sp.csr_matrix) -> np.ndarray:
    n = A.shape[0]
    rows = _rows(A).astype(np.int64)
    keys = rows * n + A.indices
    tkeys = A.indices.astype(np.int64) * n + rows
    pos = np.searchsorted(keys, tkeys)
    pos = np.minimum(pos, len(keys) - 1)
    ok = keys[pos] == tkeys
    return np.where(ok, pos, -1)


def cf_split(A: sp.csr_matrix, strong: np.ndarray):
    """classical_amg.py:85-146 with the deterministic tie-break. Returns
    (cf markers before orphan promotion, number of C points)."""
    n = A.shape[0]
    rows = _rows(A)
    tp = transpose_pos(A)
    if np.any(tp < 0):
        raise ValueError("matrix pattern is not structurally symmetric")
    ing = (strong | strong[tp]) & (A.indices != rows)
    deg = np.bincount(rows[ing], minlength=n)
    h = hash32(np.arange(n))
    order = np.lexsort((np.arange(n), h, deg))  # ascending by (deg, hash, idx)
    rank = np.empty(n, dtype=np.int64)
    rank[order] = np.arange(n)
    nm = np.full(n, -1, dtype=np.int64)
    np.maximum.at(nm, rows[ing], rank[A.indices[ing]])
    is_max = rank > nm
    has_c = np.zeros(n, dtype=bool)
    np.logical_or.at(has_c, rows[ing], is_max[A.indices[ing]])
    cf = np.where(is_max, C_PT, np.where(has_c, F_PT, C_PT)).astype(np.int8)
    return cf, int(np.sum(cf == C_PT))


def cf_split_iterated(A: sp.csr_matrix, strong: np.ndarray):
    """Iterated independent-set split (not in the reference): rounds of the
    cf_split selection restricted to undecided points; the undecided graph
    neighbours of each new C point become F. Returns (cf, number of C points)."""
    n = A.shape[0]
    rows = _rows(A)
    tp = transpose_pos(A)
    if np.any(tp < 0):
        raise ValueError("matrix pattern is not structurally symmetric")
    ing = (strong | strong[tp]) & (A.indices != rows)
    deg = np.bincount(rows[ing], minlength=n)
    h = hash32(np.arange(n))
    order = np.lexsort((np.arange(n), h, deg))
    rank = np.empty(n, dtype=np.int64)
    rank[order] = np.arange(n)
    r_e, c_e = rows[ing], A.indices[ing]
    und = np.ones(n, dtype=bool)
    cf = np.full(n, F_PT, dtype=np.int8)
    while und.any():
        m = und[r_e] & und[c_e]
        nm = np.full(n, -1, dtype=np.int64)
        np.maximum.at(nm, r_e[m], rank[c_e[m]])
        newc = und & (rank > nm)
        cf[newc] = C_PT
        und &= ~newc
        cov = np.zeros(n, dtype=bool)
        mm = und[r_e] & newc[c_e]
        cov[r_e[mm]] = True
        und &= ~cov
    return cf, int(np.sum(cf == C_PT))


def force_dominant(A: sp.csr_matrix, cf: np.ndarray, ratio: float) -> np.ndarray:
    """Rows with |a_ii| > ratio * sum_j!=i |a_ij| become C points (not in the
    reference): a BHP well row is Dirichlet-like, and as an F point its
    row-normalised interpolation would spread its diagonal over several coarse
    unknowns as large positive couplings."""
    if ratio <= 0:
        return cf
    rows = _rows(A)
    off = A.indices != rows
    offsum = np.bincount(rows[off], weights=np.abs(A.data[off]), minlength=A.shape[0])
    d = np.zeros(A.shape[0])
    np.add.at(d, rows[~off], A.data[~off])
    out = cf.copy()
    out[np.abs(d) > ratio * offsum] = C_PT
    return out


def promote_orphans(A: sp.csr_matrix, strong: np.ndarray, cf: np.ndarray) -> np.ndarray:
    """classical_amg.py:177-189 — F points without a strong C neighbour become C."""
    rows = _rows(A)
    m = strong & (A.indices != rows) & (cf[A.indices] == C_PT)
    has = np.bincount(rows[m], minlength=A.shape[0]) > 0
    out = cf.copy()
    out[(cf == F_PT) & ~has] = C_PT
    return out


def interpolation(A: sp.csr_matrix, cf: np.ndarray, strong: np.ndarray) -> sp.csr_matrix:
    """classical_amg.py:191-346 — RS direct weights with the weak-sum correction,
    distance-two F-F contributions, clamping, F-row normalisation and the
    zero-row fallback; explicit zeros are dropped."""
    n = A.shape[0]
    rows = _rows(A)
    cols = A.indices
    v = A.data
    isC = cf == C_PT
    isF = ~isC
    f2c = np.cumsum(isC) - 1
    nc = int(isC.sum())
    off = cols != rows
    m_sC = strong & off & isC[cols] & isF[rows]
    m_wC = ~strong & off & isC[cols] & isF[rows]
    m_wF = ~strong & off & isF[cols] & isF[rows]
    m_sF = strong & off & isF[cols] & isF[rows]
    sum_strong = np.bincount(rows[m_sC], weights=-v[m_sC], minlength=n)
    cnt_c = np.bincount(rows[m_sC], minlength=n)
    valid = (cnt_c > 0) & (sum_strong > 1e-14)
    diag = np.zeros(n)
    dm = cols == rows
    np.add.at(diag, rows[dm], v[dm])
    diag = np.where(np.abs(diag) > 1e-14, diag, 1.0)
    inval = m_sF & ~valid[cols]
    weak_c = np.bincount(rows[m_wC], weights=-v[m_wC], minlength=n)
    f_weak = np.bincount(rows[m_wF], weights=-v[m_wF], minlength=n) + np.bincount(
        rows[inval], weights=-v[inval], minlength=n)
    corr = np.zeros(n)
    mc = (sum_strong > 1e-14) & isF
    corr[mc] = (weak_c[mc] + f_weak[mc]) / sum_strong[mc]
    with np.errstate(divide="ignore", invalid="ignore"):
        d_rows = rows[m_sC]
        d_vals = np.nan_to_num(-v[m_sC] / diag[d_rows], nan=0.0, posinf=0.0, neginf=0.0)
        d_vals = np.maximum(d_vals * (1.0 + corr[d_rows]), 0.0)
    d_cols = f2c[cols[m_sC]]
    # distance-two: strong C entries of each F row j, in row order
    sc_idx = np.nonzero(m_sC)[0]
    sc_ptr = np.concatenate([[0], np.cumsum(cnt_c)])
    ff = np.nonzero(m_sF & valid[cols])[0]
    if ff.size:
        with np.errstate(divide="ignore", invalid="ignore"):
            factor = np.nan_to_num(-v[ff] / diag[rows[ff]], nan=0.0, posinf=0.0, neginf=0.0)
        j = cols[ff]
        cnt = cnt_c[j]
        rep_rows = np.repeat(rows[ff], cnt)
        rep_fac = np.repeat(factor, cnt)
        rep_j = np.repeat(j, cnt)
        base = np.repeat(sc_ptr[j], cnt)
        intra = np.arange(cnt.sum()) - np.repeat(np.cumsum(cnt) - cnt, cnt)
        e2 = sc_idx[base + intra]
        jw = -v[e2] / sum_strong[rep_j]
        f_vals = np.maximum(rep_fac * jw, 0.0)
        f_cols = f2c[cols[e2]]
        f_rows = rep_rows
    else:
        f_vals = np.zeros(0)
        f_cols = np.zeros(0, dtype=np.int64)
        f_rows = np.zeros(0, dtype=np.int64)
    c_rows = np.nonzero(isC)[0]
    R = np.concatenate([c_rows, d_rows, f_rows])
    Cc = np.concatenate([f2c[c_rows], d_cols, f_cols])
    V = np.concatenate([np.ones(c_rows.size), d_vals, f_vals])
    P = sp.coo_matrix((V, (R, Cc)), shape=(n, max(nc, 1))).tocsr()
    P.sum_duplicates()
    P.sort_indices()
    rs = np.asarray(P.sum(axis=1)).ravel()
    scale = np.ones(n)
    good = isF & (rs > 1e-14)
    scale[good] = 1.0 / rs[good]
    zero_rows = np.nonzero(isF & ~(rs > 1e-14))[0]
    if zero_rows.size:
        P = P.tolil()
        for i in zero_rows:
            nb = A.indices[A.indptr[i]:A.indptr[i + 1]]
            cn = nb[isC[nb]]
            if cn.size == 0:
                continue
            t = f2c[cn[0]]
            P[i, t] = P[i, t] + 1.0
            s = P[i].sum()
            scale[i] = 1.0 / s if s > 1e-14 else 1.0
        P = P.tocsr()
        P.sort_indices()
    P = sp.diags(scale) @ P
    P.data = np.maximum(P.data, 0.0)
    P.eliminate_zeros()
    P.sort_indices()
    return P.tocsr()


def truncate(P: sp.csr_matrix, cf: np.ndarray, trunc: float, pmax: int) -> sp.csr_matrix:
    """Truncation of the F rows of P (not in the reference). Per F row with more
    than one weight: tot = sequential sum, drop w < trunc * max(w), keep the pmax
    largest (ties: lower coarse column), scale the kept weights by tot / their
    sequential sum. C rows are untouched. Weights within 1e-10 * max(w) of each
    other count as tied, so ulp-level rounding differences do not change the
    pattern."""
    if trunc <= 0.0 and pmax <= 0:
        return P
    P = sp.csr_matrix(P)
    ptr, idx, val = P.indptr, P.indices, P.data
    rows_out, cols_out, vals_out = [], [], []
    for i in range(P.shape[0]):
        s, e = ptr[i], ptr[i + 1]
        c, v = idx[s:e], val[s:e]
        if cf[i] == C_PT or e - s <= 1:
            rows_out.append(np.full(e - s, i)); cols_out.append(c); vals_out.append(v)
            continue
        tot = np.cumsum(v)[-1]
        tie = 1e-10 * v.max()
        keep = v + tie >= trunc * v.max()
        c, v = c[keep], v[keep]
        if pmax > 0 and v.size > pmax:
            k = v.size
            above = (v[None, :] > v[:, None] + tie) | (
                (np.abs(v[None, :] - v[:, None]) <= tie) & (np.arange(k)[None, :] < np.arange(k)[:, None]))
            sel = above.sum(axis=1) < pmax
            c, v = c[sel], v[sel]
        v = v * (tot / np.cumsum(v)[-1])
        rows_out.append(np.full(v.size, i)); cols_out.append(c); vals_out.append(v)
    out = sp.csr_matrix((np.concatenate(vals_out), (np.concatenate(rows_out), np.concatenate(cols_out))),
                        shape=P.shape)
    out.sort_indices()
    return out


def inv_relax(A: sp.csr_matrix) -> np.ndarray:
    """classical_amg.py:885-889: 1/max(|a_ii|, 0.3*sum_j|a_ij|), capped at 1e2."""
    d = np.abs(_diag_first(A))
    ra = np.add.reduceat(np.abs(A.data), A.indptr[:-1]) if A.nnz else np.zeros(A.shape[0])
    ra[np.diff(A.indptr) == 0] = 0.0
    den = np.maximum(np.maximum(d, 0.3 * ra), 1e-30)
    return np.minimum(1.0 / den, 1e2)


class Hierarchy:
    """classical_amg.py:699-938 (setup) and :1037-1180 (cycles)."""

    REFERENCE = {"equil_mode": 0, "coarse_pin": 1, "trunc": 0.0, "pmax": 0, "coarsening": 0, "dominant": 0.0,
                 "pre_smooth": 3, "post_smooth": 2, "omega": 0.7}

    def __init__(self, A0: sp.csr_matrix, theta=0.25, max_levels=10, coarsest_size=200, pre_smooth=2,
                 post_smooth=2, omega=0.8, equil_threshold=1e-6, dense_max=500, coarse_sweeps=50, equil_mode=1, coarse_pin=2,
                 trunc=0.2, pmax=4, coarsening=1, dominant=10.0):
        self.theta, self.max_levels, self.coarsest_size = theta, max_levels, coarsest_size
        self.pre, self.post, self.omega = pre_smooth, post_smooth, omega
        self.equil_threshold, self.dense_max, self.coarse_sweeps = equil_threshold, dense_max, coarse_sweeps
        self.equil_mode, self.coarse_pin, self.trunc, self.pmax = equil_mode, coarse_pin, trunc, pmax
        self.coarsening, self.dominant = coarsening, dominant
        A0 = sp.csr_matrix(A0, dtype=np.float64)
        A0.sort_indices()
        self._setup(A0)

    def _equilibrate(self, A0, decide=True):
        d = np.maximum(np.abs(_diag_first(A0)), 1e-30)
        self.diag0 = _diag_first(A0)
        if self.equil_mode == 1:
            if decide:
                med = np.sort(d)[(d.size - 1) // 2]
                self.equilibrated = bool(med < self.equil_threshold)
                self.scalar = 1.0 / np.sqrt(med) if self.equilibrated else 1.0
            self.dh = np.full(A0.shape[0], self.scalar)
            A = A0.copy()
            A.data = A0.data * self.scalar * self.scalar
            return A
        med = np.sort(d)[(d.size - 1) // 2]  # torch.median: lower median
        self.equilibrated = bool(med < self.equil_threshold)
        if self.equilibrated:
            self.dh = 1.0 / np.sqrt(d)
            Dh = sp.diags(self.dh)
            A = (Dh @ A0 @ Dh).tocsr()
            A.sort_indices()
        else:
            self.dh = np.ones(A0.shape[0])
            A = A0.copy()
        return A

    def _setup(self, A0):
        A = self._equilibrate(A0)
        self.levels = []
        for _lvl in range(self.max_levels):
            n = A.shape[0]
            if n <= self.coarsest_size:
                break
            s = strength(A, self.theta)
            cf0, ncoarse = (cf_split_iterated if self.coarsening == 1 else cf_split)(A, s)
            cf0 = force_dominant(A, cf0, self.dominant)
            ncoarse = int(np.sum(cf0 == C_PT))
            if ncoarse >= 0.9 * n:
                break
            cf = promote_orphans(A, s, cf0)
            P = truncate(interpolation(A, cf, s), cf, self.trunc, self.pmax)
            R = P.T.tocsr()
            R.sort_indices()
            Ac = (R @ (A @ P)).tocsr()
            Ac.sort_indices()
            self.levels.append({"A": A, "P": P, "R": R, "cf": cf, "strong": s, "ir": inv_relax(A),
                                "coarsest": False})
            A = Ac
        self.levels.append({"A": A, "P": None, "R": None, "cf": None, "ir": inv_relax(A), "coarsest": True})
        self._coarse_factor()

    def _coarse_factor(self):
        """classical_amg.py:1045-1069: dense solve, DOF 0 pinned per coarse_pin.
        A single-level hierarchy is never pinned: alone the pin would freeze
        DOF 0 of a nonsingular system."""
        L = self.levels[-1]
        n = L["A"].shape[0]
        L["dense"] = n <= self.dense_max
        mode = self.coarse_pin if len(self.levels) > 1 else 0
        if mode == 2:
            A = L["A"]
            rs = np.asarray(A.sum(axis=1)).ravel()
            L["pin"] = not bool(np.any(np.abs(rs) > 1e-8 * np.abs(A.diagonal())))
        else:
            L["pin"] = mode == 1
        if L["dense"]:
            Ad = L["A"].toarray()
            if L["pin"]:
                Ad[0, :] = 0.0
                Ad[:, 0] = 0.0
                Ad[0, 0] = 1.0
            L["Ainv"] = np.linalg.inv(Ad)

    def update(self, A0: sp.csr_matrix):
        """classical_amg.py:1308-1392: new values, frozen P."""
        A0 = sp.csr_matrix(A0, dtype=np.float64)
        A0.sort_indices()
        A = self._equilibrate(A0, decide=False)
        for L in self.levels:
            L["A"] = A
            L["ir"] = inv_relax(A)
            if L["coarsest"]:
                break
            A = (L["R"] @ (A @ L["P"])).tocsr()
            A.sort_indices()
        self._coarse_factor()

    # ------------------------------------------------------------- cycles
    def _smooth(self, L, x, b, nu):
        for _ in range(nu):
            x = x + self.omega * L["ir"] * (b - L["A"] @ x)
        return x

    def _coarse(self, L, b):
        bl = b.copy()
        if L["pin"]:
            bl[0] = 0.0
        if L["dense"]:
            return L["Ainv"] @ bl
        return self._smooth(L, np.zeros_like(b), bl, self.coarse_sweeps)

    def vcycle(self, b, x=None, lvl=0):
        L = self.levels[lvl]
        if x is None:
            x = np.zeros_like(b)
        if L["coarsest"]:
            return self._coarse(L, b)
        x = self._smooth(L, x, b, self.pre)
        r = b - L["A"] @ x
        ec = self.vcycle(L["R"] @ r, None, lvl + 1)
        x = x + L["P"] @ ec
        return self._smooth(L, x, b, self.post)

    def precondition(self, v):
        """Preconditioner of D^-1 A: z = S V(S D v), S = D^-1/2 or I."""
        return self.dh * self.vcycle(self.dh * self.diag0 * v)

    def apply(self, b, cycles=1, tol=5e-2, max_cycles=8):
        """classical_amg.py:1157-1180 (equilibrated internally, physical in/out)."""
        bs = self.dh * b
        x = np.zeros_like(bs)
        A = self.levels[0]["A"]
        c = 0
        while c < max_cycles:
            x = self.vcycle(bs, x)
            c += 1
            rel = np.linalg.norm(bs - A @ x) / (np.linalg.norm(bs) + 1e-30)
            if c >= cycles and rel < tol:
                break
        return self.dh * x


ERR_ACCEPT = 0.5  # the error check accepts an estimate up to this fraction of etol
ERR_TARGET = 0.25  # residual targets aim the estimate at this fraction of etol


def fgmres(A: sp.csr_matrix, b: np.ndarray, x0=None, precond=None, rtol=1e-8, restart=30, maxit=300, etol=0.0,
           memo=None):
    """Right-preconditioned restarted FGMRES on D^-1 A x = D^-1 b (same algorithm
    as the HIP solver): classical Gram-Schmidt with a second, always-run pass
    (CGS2). Returns (x, stats).

    etol > 0 adds a pressure-error check once the relative residual is at rtol:
    z = M r (one preconditioner application to the scaled residual) estimates the
    error A^-1 r, since M approximates A^-1 D; the estimate is applied (x += z, a
    refinement step whose error is (I - M A) e) and accepted when
    ||z||_inf <= ERR_ACCEPT * etol * ||x||_inf. Otherwise the residual target is
    lowered by the ratio of estimate to residual and the iteration continues
    (restarting from the corrected x). ``memo`` (a dict kept across solves)
    holds the ratios of the last two checks; the next solve starts from the
    target their larger one predicts."""
    A = sp.csr_matrix(A)
    dinv = 1.0 / A.diagonal()
    bs = dinv * b
    bnorm = np.linalg.norm(bs)
    x = np.zeros_like(b) if x0 is None else x0.astype(np.float64).copy()
    if bnorm == 0:
        return np.zeros_like(b), {"iterations": 0, "converged": True, "rel_residual": 0.0, "checks": 0,
                                  "err_estimate": 0.0}
    check = etol > 0 and precond is not None
    memo = {} if memo is None else memo
    tol = rtol * bnorm
    target = tol
    ratio = max(memo.get("ratios", [0.0]))
    if check and ratio > 0:
        target = min(tol, ERR_TARGET * etol / ratio * bnorm)
    r = dinv * (b - A @ x)
    beta = np.linalg.norm(r)
    its = 0
    checks, est, ok = 0, 0.0, True
    m = restart
    while True:
        while beta > target and its < maxit:
            x, k = _fgmres_cycle(A, dinv, x, r, beta, target, precond, m, maxit - its)
            its += k
            r = dinv * (b - A @ x)
            beta = np.linalg.norm(r)
        if not check or not beta <= tol:
            ok = bool(beta <= tol)
            break
        z = precond(r)
        x = x + z
        est = float(np.max(np.abs(z)) / max(np.max(np.abs(x)), 1e-300))
        checks += 1
        memo["ratios"] = [est / max(beta / bnorm, 1e-300)] + memo.get("ratios", [])[:1]
        r = dinv * (b - A @ x)
        beta = np.linalg.norm(r)
        ok = est <= ERR_ACCEPT * etol
        if ok or its >= maxit:
            break
        target = min(target, beta * ERR_TARGET * etol / est)
    return x, {"iterations": its, "converged": bool(ok and beta <= tol), "rel_residual": beta / bnorm,
               "checks": checks, "err_estimate": est}


def _fgmres_cycle(A, dinv, x, r, beta, tol, precond, m, maxit):
    """One restart cycle of at most min(m, maxit) Arnoldi steps from residual r
    (||r|| = beta). Returns the updated x and the number of steps."""
    n = x.size
    V = np.zeros((m + 1, n))
    Z = np.zeros((m, n))
    H = np.zeros((m + 1, m))
    cs = np.zeros(m)
    sn = np.zeros(m)
    g = np.zeros(m + 1)
    g[0] = beta
    V[0] = r / beta
    k = 0
    for j in range(m):
        Z[j] = precond(V[j]) if precond is not None else V[j]
        w = dinv * (A @ Z[j])
        h = V[:j + 1] @ w
        w = w - V[:j + 1].T @ h
        h2 = V[:j + 1] @ w
        w = w - V[:j + 1].T @ h2
        H[:j + 1, j] = h + h2
        hn = np.linalg.norm(w)
        H[j + 1, j] = hn
        for i in range(j):
            a, bb = H[i, j], H[i + 1, j]
            H[i, j] = cs[i] * a + sn[i] * bb
            H[i + 1, j] = -sn[i] * a + cs[i] * bb
        a, bb = H[j, j], H[j + 1, j]
        if bb == 0:
            c, s = 1.0, 0.0
        elif abs(bb) > abs(a):
            t = a / bb
            s = 1 / np.sqrt(1 + t * t)
            c = s * t
        else:
            t = bb / a
            c = 1 / np.sqrt(1 + t * t)
            s = c * t
        cs[j], sn[j] = c, s
        H[j, j] = c * a + s * bb
        H[j + 1, j] = 0.0
        g[j + 1] = -s * g[j]
        g[j] = c * g[j]
        V[j + 1] = w / (hn if hn > 0 else 1.0)
        k = j + 1
        if not abs(g[j + 1]) > tol or k >= maxit:
            break
    y = np.zeros(k)
    for i in range(k - 1, -1, -1):
        y[i] = (g[i] - H[i, i + 1:k] @ y[i + 1:k]) / H[i, i] if H[i, i] != 0 else 0.0
    return x + Z[:k].T @ y, k
