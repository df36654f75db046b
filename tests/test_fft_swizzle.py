"""Index algebra of the bank-spread stage orders of the compiled-length FFT
(mr_fft_ct.hpp, mr_ct_stage SZ bits), on the CPU.

The kernels' first stage (radix 8, NS = 1) stores output P at
P ^ 2((P >> 4) & 3) and the second stage reads there; for transforms 8, 8, R
with N / R = 64 the second stage stores at P ^ 8((P >> 6) & 1) and the third
reads there. The per-lane addresses the kernels compute for those writes and
reads are restated here, and a three-stage Stockham transform driven by them
must equal numpy.fft at every x length with a compiled fused kernel; the
writes must spread over the LDS banks as intended. The GPU parity tests check
the kernels themselves."""

import numpy as np
import pytest


def factor(n):
    r, m = [], n
    for s in (8, 10, 6, 4, 2, 3, 5, 7):
        while m % s == 0:
            r.append(s)
            m //= s
    assert m == 1
    return r


def can1(n):
    R = factor(n)
    return len(R) >= 2 and R[0] == 8 and (n // R[1]) % 16 == 0


def can1b(n):
    R = factor(n)
    return not can1(n) and len(R) >= 2 and R[0] == 8 and (n // R[1]) % 8 == 0


def can2(n):
    R = factor(n)
    return len(R) == 3 and R[0] == 8 and R[1] == 8 and n // R[2] == 64


def phi1(P):
    return P ^ (2 * ((P >> 4) & 3))


def phi2(P):
    return P ^ (8 * ((P >> 6) & 1))


def phi1b(P):
    return P ^ (2 * ((P >> 3) & 3))


def stage1_write(j, q, b=False):
    """Address of first-stage output q of butterfly j (16-byte pairs)."""
    f = 2 * (j & 3) if b else 2 * ((j >> 1) & 3)
    return 8 * j + ((2 * (q // 2)) ^ f) + (q & 1)


def stage2_read(j, r, m, b=False):
    if b:
        return r * m + (j ^ (2 * (((j >> 3) + r * (m // 8)) & 3)))
    return r * m + (j ^ (2 * (((j >> 4) + r * (m // 16)) & 3)))


def stage2_write(j, q):
    jn, kk = j // 8, j % 8
    s = 8 * (jn & 1)
    return jn * 64 + kk + (s if q % 2 == 0 else -s) + 8 * q


def stage3_read(j, r):
    return 64 * r + (j ^ (8 * (r & 1)))


def stockham(x, inverse=False):
    n = len(x)
    R = factor(n)
    a = x.astype(np.complex128).copy()
    ns = 1
    sgn = 1 if inverse else -1
    for s, rad in enumerate(R):
        m = n // rad
        b = np.zeros(n, np.complex128)
        for j in range(m):
            if s == 1 and (can1(n) or can1b(n)):
                v = np.array([a[stage2_read(j, r, m, can1b(n))] for r in range(rad)])
            elif s == 2 and can2(n):
                v = np.array([a[stage3_read(j, r)] for r in range(rad)])
            else:
                v = np.array([a[j + r * m] for r in range(rad)])
            jn, kk = j // ns, j % ns
            if ns > 1:
                v = v * np.exp(sgn * 2j * np.pi * kk / (ns * rad)) ** np.arange(rad)
            V = np.exp(sgn * 2j * np.pi * np.outer(np.arange(rad), np.arange(rad)) / rad) @ v
            for q in range(rad):
                if s == 0 and (can1(n) or can1b(n)):
                    b[stage1_write(j, q, can1b(n))] = V[q]
                elif s == 1 and can2(n):
                    b[stage2_write(j, q)] = V[q]
                else:
                    b[jn * ns * rad + kk + q * ns] = V[q]
        a = b
        ns *= rad
    return a


@pytest.mark.parametrize("n", [256, 320, 384, 512, 640])
def test_swizzled_stockham_is_the_dft(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert np.abs(stockham(x) - np.fft.fft(x)).max() < 1e-9 * n
    assert np.abs(stockham(x, True) - n * np.fft.ifft(x)).max() < 1e-9 * n


@pytest.mark.parametrize("n", [256, 384, 512, 640])
def test_addresses_are_the_layout_and_spread_banks(n):
    R = factor(n)
    m0, m1 = n // R[0], n // R[1]
    # the per-lane addresses are the permutation phi1 of the plain positions
    for j in range(m0):
        for q in range(8):
            assert stage1_write(j, q) == phi1(8 * j + q)
    for j in range(m1):
        for r in range(R[1]):
            assert stage2_read(j, r, m1) == phi1(j + r * m1)
    # 16-byte first-stage writes of 8 neighbouring lanes: 8 distinct 16-byte
    # bank groups (128 bytes), against 2 in plain order
    for j0 in range(0, m0 - 7, 8):
        for c in range(4):
            groups = {(stage1_write(j, 2 * c) // 2) % 8 for j in range(j0, j0 + 8)}
            assert len(groups) == 8
    if can2(n):
        for j in range(m1):
            for q in range(8):
                assert stage2_write(j, q) == phi2((j // 8) * 64 + j % 8 + 8 * q)
        for j in range(64):
            for r in range(R[2]):
                assert stage3_read(j, r) == phi2(j + 64 * r)
        # 8-byte second-stage writes of 16 neighbouring lanes: 16 distinct
        # 8-byte slots of a 128-byte bank row
        for j0 in range(0, m1 - 15, 16):
            for q in range(8):
                slots = {stage2_write(j, q) % 16 for j in range(j0, j0 + 16)}
                assert len(slots) == 16


def test_spread_order_at_320():
    """x length 320 (8, 8, 5): the second stage's N / R = 40 admits the
    order P ^ 2((P >> 3) & 3), which halves the first stage's conflicts."""
    n = 320
    R = factor(n)
    assert can1b(n) and not can1(n)
    m0, m1 = n // R[0], n // R[1]
    for j in range(m0):
        for q in range(8):
            assert stage1_write(j, q, True) == phi1b(8 * j + q)
    for j in range(m1):
        for r in range(R[1]):
            assert stage2_read(j, r, m1, True) == phi1b(j + r * m1)
    for j0 in range(0, m0 - 7, 8):
        for c in range(4):
            groups = [(stage1_write(j, 2 * c, True) // 2) % 8 for j in range(j0, j0 + 8)]
            assert len(set(groups)) == 4 and all(groups.count(g) == 2 for g in set(groups))
            plain = [(8 * j + 2 * c) // 2 % 8 for j in range(j0, j0 + 8)]
            assert len(set(plain)) == 2
