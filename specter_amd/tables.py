"""FC-Gram tables: Dirichlet (A{C}-{d}.dat, Q{d}.dat) and Neumann (Q1n{d}.dat,
Q2n{d}.dat).

The files under ``specter_amd/tables`` are data taken unchanged from the
reference's ``tables/`` directory (raw float64 streams in Fortran order, read
as in fftp/fcgram_mod.f90:221-244).  The raw tables are handed to the plan,
which forms dir = A*Q^T itself.
"""
from __future__ import annotations

import os

import numpy as np

TABLE_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tables")


def load_dirichlet_tables(C: int, d: int, tdir: str | None = None):
    """Return (A, Q) as Fortran-ordered float64 arrays of shape (C, d) and (d, d)."""
    tdir = tdir or TABLE_DIR
    apath = os.path.join(tdir, f"A{C}-{d}.dat")
    qpath = os.path.join(tdir, f"Q{d}.dat")
    for path in (apath, qpath):
        if not os.path.exists(path):
            raise FileNotFoundError(f"Could not find table {path}")
    A = np.fromfile(apath, dtype="<f8")
    Q = np.fromfile(qpath, dtype="<f8")
    if A.size != C * d or Q.size != d * d:
        raise ValueError(f"table size mismatch for C={C}, d={d}")
    return (np.asfortranarray(A.reshape((C, d), order="F")),
            np.asfortranarray(Q.reshape((d, d), order="F")))


def load_neumann_tables(d: int, ordn: int, tdir: str | None = None):
    """Return (Qn, dxp): Q{ordn}n{d}.dat holds a float64 dxp, then Qn (d x d,
    Fortran order) (fcgram_mod.f90:319-345); ordn 1 (first normal
    derivative) or 2 (second)."""
    if ordn not in (1, 2):
        raise ValueError("ordn must be 1 or 2")
    tdir = tdir or TABLE_DIR
    path = os.path.join(tdir, f"Q{ordn}n{d}.dat")
    if not os.path.exists(path):
        raise FileNotFoundError(f"Could not find table {path}")
    raw = np.fromfile(path, dtype="<f8")
    if raw.size != 1 + d * d:
        raise ValueError(f"table size mismatch for Q{ordn}n{d}")
    return np.asfortranarray(raw[1:].reshape((d, d), order="F")), float(raw[0])
