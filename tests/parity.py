"""Hit-record comparison used by the parity tests.

Closest-hit parity (the stated contract): primitive ids must match except on
near-ties, where both results hit and |t_a - t_b| <= 1e-5 * t; where ids
match, t and the barycentrics must agree within 1e-5 relative (the HIP
kernels use the reference arithmetic, so the tests also check that they are
in fact bit-identical).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

MISS = 0xFFFFFFFF
TIE_REL = 1e-5


@dataclass
class HitDiff:
    n: int
    id_mismatch: int
    near_ties: int
    non_tie: int
    value_mismatch_bits: int
    max_rel_t: float
    max_rel_uv: float

    def ok(self) -> bool:
        return self.non_tie == 0 and self.max_rel_t <= TIE_REL and self.max_rel_uv <= TIE_REL


def compare(gpu: np.ndarray, ref: np.ndarray) -> HitDiff:
    g = np.ascontiguousarray(gpu).view(np.uint32).reshape(-1, 4)
    r = np.ascontiguousarray(ref).view(np.uint32).reshape(-1, 4)
    assert g.shape == r.shape
    gf = g.view(np.float32)
    rf = r.view(np.float32)
    same = g[:, 0] == r[:, 0]
    diff = ~same
    both_hit = (g[:, 0] != MISS) & (r[:, 0] != MISS)
    dt = np.abs(gf[:, 1].astype(np.float64) - rf[:, 1].astype(np.float64))
    tie = diff & both_hit & (dt <= TIE_REL * np.abs(rf[:, 1].astype(np.float64)))
    hit_same = same & (r[:, 0] != MISS)
    vals_bits = int(np.sum(np.any(g[hit_same, 1:] != r[hit_same, 1:], axis=1)))

    def rel(a, b):
        a = a.astype(np.float64)
        b = b.astype(np.float64)
        den = np.maximum(np.abs(b), 1e-30)
        return float(np.max(np.abs(a - b) / den)) if a.size else 0.0

    rt = rel(gf[hit_same, 1], rf[hit_same, 1])
    # barycentrics: relative to max(|v|, 1e-6) so tiny coordinates do not blow up
    uv_g = gf[hit_same, 2:4].astype(np.float64)
    uv_r = rf[hit_same, 2:4].astype(np.float64)
    ruv = float(np.max(np.abs(uv_g - uv_r) / np.maximum(np.abs(uv_r), 1e-6))) if uv_r.size else 0.0
    return HitDiff(n=g.shape[0], id_mismatch=int(diff.sum()), near_ties=int(tie.sum()),
                   non_tie=int((diff & ~tie).sum()), value_mismatch_bits=vals_bits, max_rel_t=rt, max_rel_uv=ruv)
