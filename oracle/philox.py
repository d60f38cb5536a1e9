"""Philox4x32-10 and the integrator's Box-Muller normals, in plain Python/NumPy.

TEST INFRASTRUCTURE ONLY: the checker of the device noise stream.

Philox4x32-10 is the counter-based generator of Salmon, Moraes, Dror and
Shaw, "Parallel random numbers: as easy as 1, 2, 3" (SC'11); its published
known-answer vectors pin this implementation (tests/test_oracle.py). The
stream layout follows the engine: counter (local atom index, step, 0, domain)
with domain 0 for the O-step noise and 1 for the initial momenta, key = the
system's 64-bit seed split into two 32-bit words; u = (x + 0.5) 2^-32 and
(n0, n1) = sqrt(-2 ln u0) (cos, sin)(2 pi u1), n2 = sqrt(-2 ln u2) cos(2 pi u3).
"""

from __future__ import annotations

import math

import numpy as np

M0, M1 = 0xD2511F53, 0xCD9E8D57
W0, W1 = 0x9E3779B9, 0xBB67AE85
MASK = 0xFFFFFFFF


def philox4x32_10(counter, key):
    c0, c1, c2, c3 = (int(x) & MASK for x in counter)
    k0, k1 = (int(x) & MASK for x in key)
    for _ in range(10):
        p0 = M0 * c0
        p1 = M1 * c2
        c0, c1, c2, c3 = ((p1 >> 32) ^ c1 ^ k0) & MASK, p1 & MASK, ((p0 >> 32) ^ c3 ^ k1) & MASK, p0 & MASK
        k0 = (k0 + W0) & MASK
        k1 = (k1 + W1) & MASK
    return c0, c1, c2, c3


def normal3(seed: int, step: int, local: int, domain: int = 0):
    r = philox4x32_10((local, step, 0, domain), (seed & MASK, (seed >> 32) & MASK))
    u = [(x + 0.5) * 2.0**-32 for x in r]
    ra = math.sqrt(-2.0 * math.log(u[0]))
    rb = math.sqrt(-2.0 * math.log(u[2]))
    return (ra * math.cos(2 * math.pi * u[1]), ra * math.sin(2 * math.pi * u[1]), rb * math.cos(2 * math.pi * u[3]))


def initial_momenta(seed: int, masses, kT: float):
    """jax_md initialize_momenta with this stream: p = sqrt(m kT) xi, minus the mean."""
    m = np.asarray(masses, np.float64)
    xi = np.array([normal3(seed, 0, i, 1) for i in range(len(m))])
    p = np.sqrt(m[:, None] * kT) * xi
    if len(m) > 1:
        p -= p.mean(axis=0, keepdims=True)
    return p
