"""CPU: the wavefront-cooperative wide-integer algorithms of
parallelac-xmp_amd/csrc/pac_wide.h, emulated lane by lane with numpy
(one uint32 limb per lane, TPI lanes per integer, ballots as bit masks),
against Python's exact integers modulo 2^(32 TPI).

Pins the carry-lookahead identity (carries = ((G << 1) + P) ^ P over the
group's generate / propagate bits), subtraction as a + ~b + 1, and the
column-sum product with its three-word split and two carry-resolved
additions, on random operands of every length, signs, and the carry
chains that stress the resolution (all-ones limbs, borrows through zero).
The device code runs the same steps; its results are checked through the
pipeline against the oracle on inputs whose decisions reach the exact
stage (lattice ties, SoS).
"""
import random

import numpy as np
import pytest

M32 = (1 << 32) - 1


def to_limbs(v, tpi):
    v %= 1 << (32 * tpi)
    return np.array([(v >> (32 * i)) & M32 for i in range(tpi)], dtype=np.uint64)


def from_limbs(a):
    return sum(int(x) << (32 * i) for i, x in enumerate(a))


def signed(v, tpi):
    v %= 1 << (32 * tpi)
    return v - (1 << (32 * tpi)) if v >> (32 * tpi - 1) else v


def ballot(pred):
    return sum(1 << i for i, p in enumerate(pred) if p)


def resolve(s, tpi):
    """cw_resolve: s = per-lane partial sums (< 2^33)"""
    lo = s & M32
    g = ballot(s >> 32)
    p = ballot(lo == M32)
    mask = (1 << tpi) - 1
    c = (((g << 1) + p) ^ p) & mask
    carry = np.array([(c >> i) & 1 for i in range(tpi)], dtype=np.uint64)
    return (lo + carry) & M32


def add(a, b, tpi):
    return resolve(a + b, tpi)


def sub(a, b, tpi):
    cin = np.zeros(tpi, dtype=np.uint64)
    cin[0] = 1
    return resolve(a + (~b & M32) + cin, tpi)


def neg(a, tpi):
    return sub(np.zeros(tpi, dtype=np.uint64), a, tpi)


def is_neg(a, tpi):
    return (int(a[tpi - 1]) >> 31) != 0


def length(a):
    nz = ballot(a != 0)
    return nz.bit_length()


def mul(a, b, tpi):
    na, nb = is_neg(a, tpi), is_neg(b, tpi)
    ma = neg(a, tpi) if na else a
    mb = neg(b, tpi) if nb else b
    la, lb = length(ma), length(mb)
    lo = np.zeros(tpi, dtype=np.uint64)
    hi = np.zeros(tpi, dtype=np.uint64)
    for i in range(la):
        ai = int(ma[i])  # broadcast from lane i
        for k in range(tpi):  # lane k
            j = k - i
            if 0 <= j < lb:
                p = ai * int(mb[j])
                lo[k] += p & M32
                hi[k] += p >> 32
    t = (lo >> np.uint64(32)) + (hi & np.uint64(M32))
    w0 = lo & np.uint64(M32)
    w1 = t & np.uint64(M32)
    w2 = ((t >> np.uint64(32)) + (hi >> np.uint64(32))) & np.uint64(M32)
    up1 = np.concatenate([[0], w1[:-1]]).astype(np.uint64)
    up2 = np.concatenate([[0, 0], w2[:-2]]).astype(np.uint64)
    r = add(add(w0, up1, tpi), up2, tpi)
    return neg(r, tpi) if na != nb else r


def _operands(rng, tpi, count):
    bits = 32 * tpi
    out = []
    for _ in range(count):
        kind = rng.randrange(6)
        if kind == 0:  # random length
            v = rng.getrandbits(rng.randrange(1, bits))
        elif kind == 1:  # all-ones runs: long carry chains
            n = rng.randrange(1, bits)
            v = (1 << n) - 1
        elif kind == 2:  # powers of two: borrows through zero limbs
            v = 1 << rng.randrange(bits - 1)
        elif kind == 3:
            v = 0
        elif kind == 4:  # limbs of 0 and all ones
            v = sum((M32 if rng.random() < 0.5 else 0) << (32 * i)
                    for i in range(rng.randrange(1, tpi)))
        else:
            v = rng.getrandbits(bits // 2 - 1)
        if rng.random() < 0.5:
            v = -v
        out.append(v)
    return out


@pytest.mark.parametrize("tpi", [8, 16, 32])
def test_add_sub_exact(tpi):
    rng = random.Random(tpi)
    ops = _operands(rng, tpi, 400)
    mod = 1 << (32 * tpi)
    for x, y in zip(ops, ops[1:] + ops[:1]):
        a, b = to_limbs(x, tpi), to_limbs(y, tpi)
        assert from_limbs(add(a, b, tpi)) == (x + y) % mod
        assert from_limbs(sub(a, b, tpi)) == (x - y) % mod
        assert from_limbs(neg(a, tpi)) == (-x) % mod


@pytest.mark.parametrize("tpi", [8, 16, 32])
def test_mul_exact_mod(tpi):
    rng = random.Random(100 + tpi)
    ops = _operands(rng, tpi, 300)
    mod = 1 << (32 * tpi)
    for x, y in zip(ops, ops[7:] + ops[:7]):
        a, b = to_limbs(x, tpi), to_limbs(y, tpi)
        assert from_limbs(mul(a, b, tpi)) == (x * y) % mod, (x, y)


def test_polynomial_sign_exact():
    """a degree-8 product whose value fits, evaluated through intermediates
    that wrap around (a term far above 2^(32 TPI) added and taken away
    again), has the exact value and sign"""
    tpi = 8
    rng = random.Random(7)
    for _ in range(200):
        xs = [rng.randrange(-(1 << 28), 1 << 28) for _ in range(8)]
        big = [rng.randrange(1 << 150, 1 << 200) for _ in range(2)]
        want = 1
        acc = to_limbs(1, tpi)
        for x in xs:
            want *= x
            acc = mul(acc, to_limbs(x, tpi), tpi)
        huge = mul(to_limbs(big[0], tpi), to_limbs(big[1], tpi), tpi)
        acc = sub(add(acc, huge, tpi), huge, tpi)
        assert signed(from_limbs(acc), tpi) == want
