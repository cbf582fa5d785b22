"""Classical Ruge-Stueben AMG (reference: src/solver/classical_amg.py).

``ClassicalAMG`` keeps the reference constructor, ``solve``, ``apply``,
``update_matrix`` and the ``levels`` / ``Dhalf_inv`` / ``device`` attributes;
the hierarchy is built and cycled on the ROCm device by the HIP library
(strength, C/F split, interpolation, Galerkin products and V-cycles are all
device kernels). Differences: the C/F tie-break is a deterministic hash
instead of the unseeded torch.rand (classical_amg.py:125); the
energy-minimising blend of interpolation weights (:389-418) is not applied,
so ``nullspace_dim`` / ``nullspace_blend`` / ``nullspace_reg`` are accepted
and ignored (``basis`` is None on every level); a changed sparsity pattern in
``update_matrix`` rebuilds the hierarchy instead of re-multiplying through the
frozen P.

``hierarchy="default"`` builds this package's hierarchy (scalar scaling,
iterated independent-set split, truncated interpolation, BHP rows as C points,
V(2,2) with omega 0.8); ``hierarchy="reference"`` restores the reference's
choices (symmetric equilibration, single-pass split, untruncated
interpolation, unconditional coarsest pin, V(3,2) with omega 0.7). Both change
iteration counts, not converged answers.
"""
from __future__ import annotations

import torch


def _device(dev=None):
    if dev is None or dev == "auto":
        if not torch.cuda.is_available():
            raise RuntimeError("the AMG solver runs on a ROCm GPU and none is visible")
        return torch.device("cuda")
    d = torch.device(dev)
    if d.type != "cuda":
        raise RuntimeError("the AMG solver runs on a ROCm GPU (device='cuda')")
    return d


def to_device_csr(A, device):
    """torch sparse (COO/CSR/dense) or scipy CSR -> (rowptr int32, col int32, val float64) on device,
    columns sorted within rows."""
    if hasattr(A, "indptr"):  # scipy sparse
        import scipy.sparse as sp

        A = sp.csr_matrix(A)
        A.sort_indices()
        return (torch.as_tensor(A.indptr, dtype=torch.int32, device=device),
                torch.as_tensor(A.indices, dtype=torch.int32, device=device),
                torch.as_tensor(A.data, dtype=torch.float64, device=device), A.shape)
    if A.layout == torch.sparse_coo:
        A = A.coalesce().to_sparse_csr()
    elif A.layout != torch.sparse_csr:
        A = A.to_sparse().coalesce().to_sparse_csr()
    return (A.crow_indices().to(device=device, dtype=torch.int32).contiguous(),
            A.col_indices().to(device=device, dtype=torch.int32).contiguous(),
            A.values().to(device=device, dtype=torch.float64).contiguous(), tuple(A.shape))


def _apply_reference_fix(A_csr, anchor_idx: int, anchor_value: float = 0.0):
    """classical_amg.py:660-686: Dirichlet anchor; row and column `anchor_idx`
    are removed except the diagonal, which becomes 1."""
    dev, dt = A_csr.device, A_csr.dtype
    A = A_csr.to_sparse_coo().coalesce()
    row, col = A.indices()
    val = A.values().clone()
    keep = ~((row == anchor_idx) | (col == anchor_idx)) | ((row == anchor_idx) & (col == anchor_idx))
    row, col, val = row[keep], col[keep], val[keep]
    dmask = (row == anchor_idx) & (col == anchor_idx)
    if bool(dmask.any()):
        val[dmask] = 1.0
    else:
        row = torch.cat([row, torch.tensor([anchor_idx], device=dev)])
        col = torch.cat([col, torch.tensor([anchor_idx], device=dev)])
        val = torch.cat([val, torch.tensor([1.0], dtype=dt, device=dev)])
    return torch.sparse_coo_tensor(torch.stack([row, col]), val, A_csr.size(), device=dev,
                                   dtype=dt).coalesce().to_sparse_csr()


def _diag_abs(rowptr, col, val):
    """|a_ii| per row of a CSR matrix (0 where the row has no diagonal entry)."""
    n = rowptr.numel() - 1
    rows = torch.repeat_interleave(torch.arange(n, device=val.device), (rowptr[1:] - rowptr[:-1]).long())
    out = torch.zeros(n, dtype=val.dtype, device=val.device)
    on = rows == col.long()
    out[rows[on]] = val[on].abs()
    return out


def prolongation_metrics(P, coarse_nodes, F_nodes, coords, worst: int = 5) -> dict:
    """Diagnostics of a prolongation (the reference's _analyze_prolongation,
    classical_amg.py:1182-1239): how well P reproduces constants (row sums
    against 1) and a linear field (P applied to the C points' coordinates
    against the fine coordinates), the sign counts and range of the weights,
    and the F rows with the largest errors of each kind."""
    Pc = P.to_sparse_coo().coalesce()
    rows, w = Pc.indices()[0], Pc.values().to(torch.float64)
    row_sum = torch.zeros(P.size(0), dtype=torch.float64, device=w.device).index_add_(0, rows, w)
    const_err = row_sum - 1.0
    lin = coords[coarse_nodes.long()]
    lin_err = torch.sparse.mm(P.to(torch.float64), lin.unsqueeze(1)).squeeze(1) - coords
    n = float(P.size(0))

    def top(err):
        if F_nodes.numel() == 0:
            return []
        v, i = torch.topk(err.abs()[F_nodes], k=min(worst, F_nodes.numel()))
        return [(int(F_nodes[j]), float(x)) for j, x in zip(i.tolist(), v.tolist())]

    return {"row_err_max": float(const_err.abs().max()) if w.numel() else 0.0,
            "const_rel": float(const_err.norm() / (n ** 0.5 + 1e-30)),
            "grad_rel": float(lin_err.norm() / (coords.norm() + 1e-30)),
            "row_sum_min": float(row_sum.min()), "row_sum_max": float(row_sum.max()),
            "neg_count": int((w < -1e-12).sum()), "pos_count": int((w > 1e-12).sum()),
            "min_w": float(w.min()) if w.numel() else 0.0, "max_w": float(w.max()) if w.numel() else 0.0,
            "worst_const_rows": top(const_err), "worst_grad_rows": top(lin_err)}


class ClassicalAMG:
    def __init__(self, A_csr_np, theta: float = 0.25, max_levels: int = 10, coarsest_size: int = 100,
                 anchor_idx: int | None = None, nullspace_dim: int = 3, nullspace_blend: float = 0.6,
                 nullspace_reg: float = 1e-6, device=None, precision: int = 64, hierarchy: str = "default"):
        from impes_hip import AmgHierarchy

        if hierarchy not in ("default", "reference"):
            raise ValueError(f"hierarchy must be 'default' or 'reference', got {hierarchy!r}")
        self.theta = theta
        self.max_levels = max_levels
        self.coarsest_size = coarsest_size
        self.anchor_idx = anchor_idx
        self.nullspace_dim = max(0, nullspace_dim)
        self.nullspace_blend = float(max(0.0, min(1.0, nullspace_blend)))
        self.nullspace_reg = float(max(1e-10, nullspace_reg))
        self.apply_tol = 5e-2
        self.apply_max_cycles = 8
        self.equilibration_threshold = 1e-6
        self.precision = int(precision)
        self.device = _device(device)
        self.hierarchy_mode = hierarchy
        extra = dict(AmgHierarchy.REFERENCE) if hierarchy == "reference" else {}
        self._h = AmgHierarchy(theta=theta, max_levels=max_levels, coarsest_size=coarsest_size,
                               equil_threshold=self.equilibration_threshold, anchor=anchor_idx, **extra)
        self._setup(A_csr_np)

    def _setup(self, A):
        rp, ci, v, shape = to_device_csr(A, self.device)
        if self.anchor_idx is not None:
            fixed = _apply_reference_fix(torch.sparse_csr_tensor(rp.long(), ci.long(), v, size=shape), self.anchor_idx)
            rp, ci, v, shape = to_device_csr(fixed, self.device)
        self._pattern = (rp, ci)
        self.n = shape[0]
        self._h.setup_csr(rp, ci, v)
        self._levels = None

    # ------------------------------------------------------------ attributes
    @property
    def use_equilibration(self) -> bool:
        return self._h.equilibrated

    @property
    def Dhalf_inv(self):
        return self._h.dhalf_inv() if self._h.equilibrated else None

    @property
    def levels(self) -> list:
        """Level list in the reference's shape (classical_amg.py:891-904, :925-933),
        exported from the device: A, n, P, diag (|a_ii|), inv_relax, is_coarsest
        and basis (None) on every level; on the others also cf_marker (0 = C),
        coarse_nodes and F_nodes (fine indices of the C and F points),
        fine2coarse (coarse index of each C point, -1 on F points), node_coords
        (the fine index carried down: 0..n-1 on level 0) and interp_metrics (the
        prolongation diagnostics of classical_amg.py:1182-1239)."""
        if self._levels is None:
            out = []
            coords = None
            for lv in range(self._h.num_levels):
                e = self._h.export_level(lv)
                n = e["n"]
                A = torch.sparse_csr_tensor(e["rowptr"].long(), e["col"].long(), e["val"], size=(n, n))
                if coords is None:
                    coords = torch.arange(n, dtype=torch.float64, device=self.device)
                d = {"A": A, "n": n, "P": None, "diag": _diag_abs(e["rowptr"], e["col"], e["val"]),
                     "inv_relax": e["inv_relax"], "basis": None, "is_coarsest": e["is_coarsest"]}
                if not e["is_coarsest"]:
                    nc = self._h.level_info(lv + 1)["n"]
                    P = torch.sparse_csr_tensor(e["p_rowptr"].long(), e["p_col"].long(), e["p_val"], size=(n, nc))
                    cf = e["cf"]
                    coarse = torch.nonzero(cf == 0).view(-1)
                    fine = torch.nonzero(cf != 0).view(-1)
                    f2c = torch.full((n,), -1, dtype=torch.int64, device=self.device)
                    f2c[coarse] = torch.arange(coarse.numel(), dtype=torch.int64, device=self.device)
                    d.update({"P": P, "cf_marker": cf, "coarse_nodes": coarse, "F_nodes": fine, "fine2coarse": f2c,
                              "node_coords": coords, "interp_metrics": prolongation_metrics(P, coarse, fine, coords)})
                    coords = coords[coarse]
                out.append(d)
            self._levels = out
        return self._levels

    @property
    def hierarchy(self):
        return self._h

    # ------------------------------------------------------------ cycles
    def solve(self, b, x0=None, tol: float = 1e-6, max_iter: int = 10):
        """classical_amg.py:1108-1155: V-cycles until ||r||/||b|| < tol (equilibrated space)."""
        b64 = b.to(self.device, dtype=torch.float64).contiguous()
        x0 = None if x0 is None else x0.to(self.device, dtype=torch.float64).contiguous()
        x, rel, _ = self._h.cycles(b64, x0, min_cycles=1, tol=tol, max_cycles=max_iter, precision=self.precision)
        self.last_rel_res = rel
        return x

    def apply(self, b, cycles: int = 1, pre_smooth: int = 3, post_smooth: int = 2):
        """classical_amg.py:1157-1180: at least `cycles` V-cycles, at most apply_max_cycles,
        until the relative residual is below apply_tol."""
        b64 = b.to(self.device, dtype=torch.float64).contiguous()
        x, rel, _ = self._h.cycles(b64, None, min_cycles=cycles, tol=self.apply_tol,
                                   max_cycles=self.apply_max_cycles, precision=self.precision)
        self.last_rel_res = rel
        return x

    def update_matrix(self, A_csr_np) -> None:
        """classical_amg.py:1308-1392: new values through the frozen hierarchy."""
        rp, ci, v, shape = to_device_csr(A_csr_np, self.device)
        if shape[0] != self.n:
            raise ValueError(f"AMG hierarchy built for n={self.n}, got a matrix with n={shape[0]}")
        if self.anchor_idx is not None:
            fixed = _apply_reference_fix(torch.sparse_csr_tensor(rp.long(), ci.long(), v, size=shape), self.anchor_idx)
            rp, ci, v, shape = to_device_csr(fixed, self.device)
        same = (rp.numel() == self._pattern[0].numel() and ci.numel() == self._pattern[1].numel()
                and bool(torch.equal(rp, self._pattern[0])) and bool(torch.equal(ci, self._pattern[1])))
        if same:
            self._h.update_csr(v)
            self._levels = None
        else:
            self._setup(torch.sparse_csr_tensor(rp.long(), ci.long(), v, size=shape))
