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


# ---------------------------------------------------------------------------
# the other strategies (forward_solver_utils.py:14-178); indices ascending,
# padding of the reference (index -1, value 0) dropped
# ---------------------------------------------------------------------------
def threshold_hard(coeff, val):
    """_threshold_hard (14-33): |c| > val."""
    coeff = np.asarray(coeff)
    idx = np.flatnonzero(magnitude(coeff) > val)
    return idx, coeff[idx]


def percentile_linear(a, q):
    """jnp.percentile(a, q) with the default 'linear' method:
    pos = q/100 (n-1); a_lo (1 - w) + a_hi w on the sorted values."""
    s = np.sort(np.asarray(a, dtype=np.float64))
    n = s.size
    pos = (q / 100.0) * (n - 1)
    lo = int(np.floor(pos))
    hi = min(int(np.ceil(pos)), n - 1)
    w = pos - lo
    return s[lo] * (1.0 - w) + s[hi] * w


def threshold_percentile(coeff, val):
    """_threshold_percentile (36-62): |c| > percentile(|c|, val)."""
    coeff = np.asarray(coeff)
    a = magnitude(coeff)
    idx = np.flatnonzero(a > percentile_linear(a, val))
    return idx, coeff[idx]


def threshold_hard_reassign(coeff, val):
    """_threshold_hard_reassign (92-120): keep |c| >= val max|c|, scale by
    sqrt(total energy / retained energy)."""
    coeff = np.asarray(coeff)
    a = magnitude(coeff)
    mx = a.max()
    keep = (mx > 0) & (a >= val * mx)
    thr = np.where(keep, coeff, 0)
    kept = np.sum(np.abs(thr) ** 2)
    total = np.sum(a ** 2)
    ratio = np.sqrt(total / kept) if kept > 0 else 0.0
    re = thr * ratio
    idx = np.flatnonzero(np.abs(re) > 0)
    return idx, re[idx]


def threshold_bao_energy(coeff, val, wpt):
    """_threshold_bao_energy (123-153): |c |centre|^d| > val."""
    from .beams import decode_indices
    coeff = np.asarray(coeff)
    t = wpt.tiling
    lvl, multi = decode_indices(wpt, np.arange(coeff.size))
    box = multi[0] + t.boxes_cumsum[lvl]
    w = np.linalg.norm(t.centres[box].astype(float), axis=1) ** t.ndim
    idx = np.flatnonzero(magnitude(coeff * w) > val)
    return idx, coeff[idx]


def threshold_perc_max_abs(coeff, val):
    """_threshold_perc_max_abs (156-178): |c| > val max|c|."""
    coeff = np.asarray(coeff)
    a = magnitude(coeff)
    idx = np.flatnonzero(a > a.max() * val)
    return idx, coeff[idx]
