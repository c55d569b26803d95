"""Shared helpers for the parity tests (oracle <-> device comparisons)."""
import numpy as np

import oracle as O


def parity_mask(dims, p):
    return O.parity(dims) == p


def rel_err(a, b, mask=None):
    """||a - b|| / ||b|| over the sites selected by mask."""
    a = np.asarray(a)
    b = np.asarray(b)
    if mask is not None:
        a = a[mask]
        b = b[mask]
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def cplx(v):
    """(..., 2n) real pairs -> (..., n) complex"""
    return v[..., 0::2] + 1j * v[..., 1::2]


def mats(u):
    """(..., 18) column-major reals -> (..., 3, 3) complex [row, col]"""
    return cplx(u).reshape(u.shape[:-1] + (3, 3)).swapaxes(-1, -2)


def unmats(m):
    """(..., 3, 3) complex [row, col] -> (..., 18) column-major reals"""
    c = m.swapaxes(-1, -2).reshape(m.shape[:-2] + (9,))
    out = np.empty(c.shape[:-1] + (18,))
    out[..., 0::2] = c.real
    out[..., 1::2] = c.imag
    return out
