"""Golden vectors (tests/golden/*.npz, written by tests/golden/make_golden.py):
the oracle and its C++ restatement reproduce them, and central finite
differences of the oracle energy reproduce the stored forces for the cases
that carry every bonded kind."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import gb_oracle as O
from tests.golden import make_golden as G

CASES = ("peptide138", "peptide138_cut9", "chain8", "chain8_harmonic", "ions3", "peptide138_gbcut12",
         "chain8_gbcut4")


@pytest.mark.parametrize("name", CASES)
def test_oracle_reproduces_golden(name):
    sysd, cutoff, gbc, terms, forces, born = G.load(name)
    t, f, b = G.evaluate(sysd, cutoff, gbc)
    assert np.allclose(t, terms, rtol=1e-12, atol=1e-10)
    assert np.allclose(f, forces, rtol=1e-11, atol=1e-10)
    assert np.allclose(b, born, rtol=1e-13)


@pytest.mark.parametrize("name", CASES)
def test_cpu_restatement_reproduces_golden(name):
    from oracle import cpu as C

    sysd, cutoff, gbc, terms, forces, born = G.load(name)
    cs = C.CpuSystem(sysd, cutoff=cutoff, gb_cutoff=gbc)
    e, f = cs.energy_forces(sysd["positions"])
    scale = np.sum(np.abs(terms[:-1]))
    assert np.max(np.abs(e - terms)) <= 1e-12 * scale
    assert np.max(np.abs(f - forces)) <= 1e-10 * max(1.0, np.max(np.abs(forces)))


@pytest.mark.parametrize("name", ("chain8", "chain8_harmonic", "ions3", "chain8_gbcut4"))
def test_golden_forces_are_energy_gradients(name):
    sysd, cutoff, gbc, terms, forces, _ = G.load(name)
    x0 = sysd["positions"]
    h = 1e-6
    fd = np.zeros_like(x0)
    for i in range(x0.shape[0]):
        for d in range(3):
            xp = x0.copy(); xp[i, d] += h
            xm = x0.copy(); xm[i, d] -= h
            fd[i, d] = -(G.evaluate({**sysd, "positions": xp}, cutoff, gbc)[0][-1]
                         - G.evaluate({**sysd, "positions": xm}, cutoff, gbc)[0][-1]) / (2 * h)
    assert np.max(np.abs(fd - forces)) < 2e-6 * max(1.0, np.max(np.abs(forces)))


def test_asymmetric_exclusion_row_takes_the_mean_scale():
    """Row 0 lists atom 4 at 1-4 scales and row 4 does not: the pair's LJ and
    Coulomb are 0.5 * (s_04 + s_40) of the unscaled pair (optimization.py:55-57
    sums 0.5 * S[i, j] over ordered pairs)."""
    sym, *_ = G.load("chain8")
    asym, *_ = G.load("chain8_harmonic")
    x = sym["positions"]
    r = np.linalg.norm(x[0] - x[4])
    q = sym["charges"]
    k = O.COULOMB_CONSTANT
    t_sym = O.energy_terms(sym, x)
    t_asym = O.energy_terms({**asym, "improper_params": sym["improper_params"]}, x)
    coul_pair = k * q[0] * q[4] / r
    assert t_asym["coulomb"] - t_sym["coulomb"] == pytest.approx(0.5 * (0.8333333333 - 1.0) * coul_pair, rel=1e-6)
