"""Oracle: the ``top_n`` coefficient selection (test infrastructure only).

Reference: ``forward_solver_utils.py:65-89``::

    abs_c = jnp.abs(coeff)
    idx_unsorted = jnp.argpartition(abs_c, abs_c.size - N)[-N:]
    idx = idx_unsorted[jnp.argsort(abs_c[idx_unsorted])]

Two formulations are kept so that each checks the other:

* ``top_n_argpartition`` emulates jax's ``argpartition`` (bottom set from
  ``lax.top_k(-a, kth+1)`` whose ties list the lower index first, the top set
  from ``lax.top_k`` on a 0/1 proxy) followed by a stable ``argsort``.
* ``top_n`` states the resulting rule directly: the N largest entries under
  the lexicographic order (|c|, index), returned in ascending (|c|, index).

The key |c| is formed as sqrt(re*re + im*im) with separately rounded IEEE
operations (the HIP kernel does the same, without fused multiply-add).  Which
of two entries whose magnitudes agree to the last bit wins is the tie rule
above; that this matches jax 0.10 bit for bit is parity unpinned.
"""

from __future__ import annotations

import numpy as np


def magnitude(c):
    """|c| as sqrt(re*re + im*im), each operation rounded separately."""
    c = np.asarray(c)
    re = np.ascontiguousarray(c.real, dtype=np.float64)
    im = np.ascontiguousarray(c.imag, dtype=np.float64)
    return np.sqrt(re * re + im * im)


def top_n(coeff, val, key=None):
    """N largest by (|c|, index), ascending; returns (idx int64, coeff[idx])."""
    coeff = np.asarray(coeff)
    a = magnitude(coeff) if key is None else np.asarray(key)
    n = a.size
    N = min(int(val), n)
    if N <= 0:
        raise ValueError("top_n threshold must be a positive integer.")
    order = np.lexsort((np.arange(n), a))      # ascending (a, idx)
    idx = order[n - N:].astype(np.int64)
    return idx, coeff[idx]


def top_n_argpartition(coeff, val, key=None):
    """Literal emulation of jnp.argpartition + stable argsort (see module doc)."""
    coeff = np.asarray(coeff)
    a = magnitude(coeff) if key is None else np.asarray(key)
    n = a.size
    N = min(int(val), n)
    kth = n - N
    # lax.top_k(-a, kth+1): largest -a first; among equal values lower index first
    neg = -a
    order = sorted(range(n), key=lambda i: (-neg[i], i))
    bottom = np.asarray(order[:kth + 1], dtype=np.int64)
    proxy = np.ones(n, dtype=bool)
    proxy[bottom] = False
    # lax.top_k(proxy, n-kth-1): all ones; lower index first
    top = np.flatnonzero(proxy)[: n - kth - 1]
    part = np.concatenate([bottom, top])[-N:]
    idx = part[np.argsort(a[part], kind="stable")]
    return idx.astype(np.int64), coeff[idx]
