"""ORACLE -- test infrastructure only: compare a result set with the oracle's.

Error measures per target k (a = acc or jdot, 3-vectors):
  rel   = |a_k - a_k^ref| / |a_k^ref|                 (plain relative error)
  norm  = |a_k - a_k^ref| / sum_j |term_jk|            (error per unit of the
          summed magnitudes; insensitive to cancellation in the sum)
Neighbour lists must be ascending and equal to the oracle's except for pairs
whose r^2 lies within `band` (relative) of r_nb^2; those pairs are counted.
"""
from __future__ import annotations

import numpy as np

from . import band_pairs


def _stats(err):
    if err.size == 0:
        return {"max": 0.0, "p99": 0.0, "median": 0.0}
    return {"max": float(np.max(err)), "p99": float(np.quantile(err, 0.99)),
            "median": float(np.median(err))}


def force_errors(acc, adot, ref_acc, ref_adot, scale_a=None, scale_j=None):
    da = np.linalg.norm(acc - ref_acc, axis=1)
    dj = np.linalg.norm(adot - ref_adot, axis=1)
    na = np.linalg.norm(ref_acc, axis=1)
    nj = np.linalg.norm(ref_adot, axis=1)
    out = {
        "a_rel": _stats(da / np.where(na > 0, na, 1.0)),
        "jdot_rel": _stats(dj / np.where(nj > 0, nj, 1.0)),
    }
    if scale_a is not None:
        out["a_norm"] = _stats(da / np.where(scale_a > 0, scale_a, 1.0))
        out["jdot_norm"] = _stats(dj / np.where(scale_j > 0, scale_j, 1.0))
    return out


def compare_lists(nnb, lists, ref_nnb, ref_lists, x, rnb2, targets, band=1e-6):
    """Returns dict(rows_checked, rows_differing, band_pairs_flipped,
    hard_mismatches, not_ascending, overflow_mismatch)."""
    nt = len(targets)
    nnb = np.asarray(nnb)
    ref_nnb = np.asarray(ref_nnb)
    cnt = np.abs(nnb)
    rcnt = np.abs(ref_nnb)
    cap = lists.shape[1]
    w = int(min(cap, max(cnt.max(initial=0), rcnt.max(initial=0))))
    col = np.arange(w)[None, :]
    g = lists[:, :w]
    r = ref_lists[:, :w]
    valid_g = col < np.minimum(cnt, cap)[:, None]
    valid_r = col < np.minimum(rcnt, cap)[:, None]
    same = (nnb == ref_nnb) & np.all((g == r) | ~valid_g, axis=1)
    # ascending check on the GPU lists
    if w > 1:
        asc = np.all((np.diff(g, axis=1) > 0) | ~valid_g[:, 1:], axis=1)
    else:
        asc = np.ones(nt, dtype=bool)
    bad_rows = np.nonzero(~same)[0]
    flipped = 0
    hard = 0
    ovf_mismatch = int(np.sum((nnb < 0) != (ref_nnb < 0)))
    if bad_rows.size:
        bt = np.asarray(targets)[bad_rows]
        bcnt, blist = band_pairs(x, rnb2, bt, rel=band, bcap=256)
        for q, k in enumerate(bad_rows):
            sg = set(g[k][valid_g[k]].tolist())
            sr = set(r[k][valid_r[k]].tolist())
            diff = sg ^ sr
            allowed = set(blist[q, :min(bcnt[q], 256)].tolist())
            flipped += len(diff & allowed)
            hard += len(diff - allowed)
            # counts can only differ by flipped band pairs
            if len(diff - allowed) == 0 and abs(int(cnt[k]) - int(rcnt[k])) > len(diff):
                hard += 1
    return {"rows_checked": int(nt), "rows_differing": int(bad_rows.size),
            "band_pairs_flipped": int(flipped), "hard_mismatches": int(hard),
            "not_ascending": int(np.sum(~asc)), "overflow_mismatch": ovf_mismatch}
