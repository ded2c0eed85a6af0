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
three-word form.  It is a chain of bijections on 32 bits: add, rotate-left
17, multiply by an odd prime, xorshift-right, and the three-word form of
the public xxHash32 round and avalanche (PRIME32_2 .. PRIME32_5 of the
xxHash specification).  Its output is inverted here, step by step with
those published constants, back to the third word p.z given (p.x, p.y): a
transposed prime, a wrong rotation or shift, or a swapped word fails.  Its
device result equals the oracle's through the AO images and the path
tracer's scrambles.
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


# xxHash32 primes (xxHash specification, XXH_PRIME32_2 .. XXH_PRIME32_5)
P32_2, P32_3, P32_4, P32_5 = 0x85EBCA77, 0xC2B2AE3D, 0x27D4EB2F, 0x165667B1


def _inv(m: int) -> int:
    return pow(m, -1, 1 << 32)


def _unxorshift(h: int, s: int) -> int:
    x = h
    for _ in range(32 // s + 1):
        x = h ^ (x >> s)
    return x & MASK


def _rotr(h: int, r: int) -> int:
    return ((h >> r) | (h << (32 - r))) & MASK


def _xxhash_third_word(out: int, x: int, y: int) -> int:
    """p.z from xxhash32(p) and (p.x, p.y), undoing the steps in reverse."""
    h = _unxorshift(out, 16)
    h = (h * _inv(P32_3)) & MASK
    h = _unxorshift(h, 13)
    h = (h * _inv(P32_2)) & MASK
    h = _unxorshift(h, 15)
    h = _rotr((h * _inv(P32_4)) & MASK, 17)  # second round: p4 * rotl(h + y * p3, 17)
    h = (h - y * P32_3) & MASK
    h = _rotr((h * _inv(P32_4)) & MASK, 17)  # first round: p4 * rotl(z + p5 + x * p3, 17)
    return (h - P32_5 - x * P32_3) & MASK


def test_xxhash32_inverts_to_third_word():
    rng = np.random.default_rng(11)
    words = rng.integers(0, 1 << 32, size=(2000, 3), dtype=np.uint64).tolist()
    words += [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [MASK, MASK, MASK], [1919, 1079, 65536 * 7 + 3]]
    for x, y, z in words:
        assert _xxhash_third_word(orc.xxhash32(x, y, z), x, y) == z, (x, y, z)


def test_xxhash32_is_the_ao_seed_and_the_scramble():
    # the AO sampler seeds with xxhash32(x, y, sample) (sampler.rs:48-55)
    for x, y, s in [(0, 0, 0), (3, 5, 1), (681, 383, 7)]:
        assert orc.ao_sampler(x, y, s, 1)[0] == orc.xxhash32(x, y, s)
    # the third word is a bijection for fixed (x, y): distinct outputs
    outs = {orc.xxhash32(17, 23, z) for z in range(4096)}
    assert len(outs) == 4096
