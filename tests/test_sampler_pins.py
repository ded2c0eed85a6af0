"""Pins of the AO pixel sampler (scene/rendering/gpu-ray-tracing/src/sampler.rs).

The sampler steps a 32-bit LCG (multiplier 747796405, increment 2891336453,
sampler.rs:11-16), hashes the new state with PCG's RXS-M-XS output function
and maps the top 23 hash bits to a float in [0, 1) (sampler.rs:64-71).  The
oracle's sequence is checked here by routes other than its own formula:

* the LCG state after n steps against the jump-ahead closed form
  a^n s + c (a^n - 1) / (a - 1)  (mod 2^32), with exact integer arithmetic;
* each hash inverted (xorshift-22 is an involution on 32 bits, the multiplier
  is odd, the variable xorshift keeps the top 4 bits that pick its shift)
  back to the state it was made from;
* each float times 2^23 equal to the hash's top 23 bits.

The seed hash xxhash32 (sampler.rs:24-41) has no published vectors for its
three-word form; its device result equals the oracle's through the AO images.
"""

import numpy as np
import pytest

from oracle import oracle as orc

A, C, M = 747796405, 2891336453, 277803737
MASK = (1 << 32) - 1


def _jump(seed: int, n: int) -> int:
    an = A ** n
    return (an * seed + C * ((an - 1) // (A - 1))) & MASK


def _unhash(out: int) -> int:
    w = (out ^ (out >> 22)) & MASK
    x1 = (w * pow(M, -1, 1 << 32)) & MASK
    s = (x1 >> 28) + 4
    x = x1
    for _ in range(32 // s + 1):
        x = x1 ^ (x >> s)
    return x & MASK


@pytest.mark.parametrize("pix", [(0, 0, 0), (1, 0, 0), (681, 383, 7), (12345, 678, 4_000_000_000)])
def test_lcg_jump_ahead(pix):
    n = 40
    seed, states, _, _ = orc.ao_sampler(*pix, n)
    for k in range(n):
        assert int(states[k]) == _jump(seed, k + 1)


@pytest.mark.parametrize("pix", [(0, 0, 0), (3, 5, 1), (681, 383, 7), (99, 1, 123456)])
def test_hash_inverts_to_state(pix):
    _, states, bits, _ = orc.ao_sampler(*pix, 64)
    for st, b in zip(states.tolist(), bits.tolist()):
        assert _unhash(b) == st


def test_float_mapping_is_top_23_bits():
    _, _, bits, vals = orc.ao_sampler(7, 9, 2, 256)
    assert np.all(vals >= 0.0) and np.all(vals < 1.0)
    assert np.array_equal((vals.astype(np.float64) * (1 << 23)).astype(np.uint64), (bits >> 9).astype(np.uint64))


def test_seeds_differ_across_pixels_and_samples():
    seeds = {orc.ao_sampler(x, y, s, 1)[0] for x in range(16) for y in range(16) for s in range(4)}
    assert len(seeds) == 16 * 16 * 4
