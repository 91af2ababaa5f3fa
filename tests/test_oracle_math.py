"""The oracle against mathematics (CPU): closed forms and dense brute force.

These checks do not use the reference: they fix what the algorithm must
compute, so a dropped term, a wrong sign or index, or a transposed operand
in the restatement fails one of them.
  - Chebyshev moments of a diagonal H:  mu_m = sum_n T_m(V_n / E_max)
  - moments of a small sparse H against dense eigendecomposition
  - U(t) = exp(-i H t) and [X, U(t)] against dense expm
  - unitarity and U(t) U(-t) = 1
  - the summation integrates to 2 mu_0 / volume (Gauss-Chebyshev)
  - Jackson g_0 = 1, 0 < g_k < 1, decreasing
  - Bessel term counts of the evolution (|J_m(t)| < 1e-15 cut)
"""
import os

import numpy as np
import pytest
import scipy.special

import oracle
from gpuqt_amd import inputs as gi


def _model(tmp_path, name, writer):
    d = tmp_path / name
    d.mkdir()
    writer(str(d))
    return oracle.OracleModel(str(d))


def _dense(m):
    """Dense H (not scaled) and X-commutator weights from the ELL arrays."""
    a = m.arrays()
    N, K = m.n, m.max_neighbor
    H = np.diag(a["potential"]).astype(complex)
    XX = np.zeros((N, N))
    for i in range(N):
        for k in range(a["neighbor_number"][i]):
            q = i * K + k
            j = a["neighbor_list"][q]
            H[i, j] += a["hopping_real"][q] + 1j * a["hopping_imag"][q]
            XX[i, j] = a["xx"][q]
    return H, XX


def _random_complex(n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


def _general_ring(d, n=30, seed=11, box=1000.0, complex_hops=True):
    """open chain with complex hoppings, positions far from the box edge so
    that the minimum-image convention never folds a distance"""
    rng = np.random.default_rng(seed)
    nb, hop = [], []
    t = [complex(-1.0 + 0.3 * rng.standard_normal(), 0.4 * rng.standard_normal())
         for _ in range(n)]
    for i in range(n):
        row, hrow = [], []
        if i > 0:
            row.append(i - 1)
            hrow.append(np.conj(t[i - 1]))
        if i < n - 1:
            row.append(i + 1)
            hrow.append(t[i])
        if i + 3 < n and i % 4 == 0:
            row.append(i + 3)
            hrow.append(complex(0.2, 0.1))
        if i - 3 >= 0 and (i - 3) % 4 == 0:
            row.append(i - 3)
            hrow.append(complex(0.2, -0.1))
        nb.append(row)
        hop.append(hrow)
    pos = [10.0 + 1.3 * i for i in range(n)]
    pot = rng.uniform(-1, 1, n)
    gi.write_general(d, nb, 3, pos, box=box, volume=float(n), potential=pot, hoppings=hop,
                     hopping_kind="complex" if complex_hops else "real")
    gi.write_para(d, model=0, moments=40, energy_max=4.0, msd=True)
    gi.write_energy(d, gi.energy_grid(-3, 3, 7))
    gi.write_time_steps(d, [1.7])


def test_diagonal_hamiltonian_closed_form(tmp_path):
    def writer(d):
        n = 50
        pot = np.linspace(-3.9, 3.7, n)
        gi.write_general(d, [[] for _ in range(n)], 1, [0.0] * n, box=1.0, volume=1.0,
                         potential=pot)
        gi.write_para(d, model=0, moments=64, energy_max=4.0)
        gi.write_energy(d, [0.0])
    m = _model(tmp_path, "diag", writer)
    r = m.random_state()
    mr, mi = m.moments(r, r, 64)
    x = np.linspace(-3.9, 3.7, 50) / 4.0
    want = np.array([np.cos(k * np.arccos(x)).sum() for k in range(64)])
    np.testing.assert_allclose(mr, want, rtol=0, atol=1e-12 * 50)
    np.testing.assert_allclose(mi, 0.0, atol=1e-12 * 50)
    # mu_0 = <r|r> = N for unit-modulus phases
    assert abs(mr[0] - 50) < 1e-12


def test_moments_against_dense(tmp_path):
    m = _model(tmp_path, "ring", _general_ring)
    H, _ = _dense(m)
    Ht = H / m.energy_max
    lam, V = np.linalg.eigh(Ht)
    assert np.abs(lam).max() < 1
    L = _random_complex(m.n, 1)
    R = _random_complex(m.n, 2)
    M = 40
    mr, mi = m.moments((L.real, L.imag), (R.real, R.imag), M)
    cL = V.conj().T @ L
    cR = V.conj().T @ R
    want = np.array([np.vdot(cL, np.cos(k * np.arccos(lam)) * cR) for k in range(M)])
    scale = np.abs(L) @ np.abs(R)
    np.testing.assert_allclose(mr + 1j * mi, want, atol=1e-12 * scale, rtol=0)


def test_operators_against_dense(tmp_path):
    m = _model(tmp_path, "ring", _general_ring)
    H, XX = _dense(m)
    psi = _random_complex(m.n, 3)
    o = m.apply(psi.real, psi.imag)
    np.testing.assert_allclose(o[0] + 1j * o[1], H @ psi / m.energy_max, atol=1e-13)
    x = np.array([10.0 + 1.3 * i for i in range(m.n)])
    X = np.diag(x)
    comm = (X @ H - H @ X) / m.energy_max
    o = m.apply_commutator(psi.real, psi.imag)
    np.testing.assert_allclose(o[0] + 1j * o[1], comm @ psi, atol=1e-12)
    # apply_current gives -i sum_m H_nm (x_m - x_n) psi_m = i [X, H] psi, not
    # scaled by E_max (src/hamiltonian.cu:354-383); the manual writes
    # V = i [H, X] (doc/gpuqt.tex), the opposite sign, which VAC, being
    # quadratic in V, does not see
    o = m.apply_current(psi.real, psi.imag)
    np.testing.assert_allclose(o[0] + 1j * o[1], 1j * (X @ H - H @ X) @ psi, atol=1e-11)


@pytest.mark.parametrize("t", [0.3, 1.7, 12.0])
def test_evolution_against_dense(tmp_path, t):
    m = _model(tmp_path, "ring", _general_ring)
    H, _ = _dense(m)
    lam, V = np.linalg.eigh(H / m.energy_max)
    U = V @ np.diag(np.exp(-1j * lam * t)) @ V.conj().T
    psi = _random_complex(m.n, 4)
    for direction, Ud in ((1, U), (-1, U.conj().T)):
        o = m.evolve(direction, t, psi.real, psi.imag)
        np.testing.assert_allclose(o[0] + 1j * o[1], Ud @ psi, atol=1e-12)
    x = np.array([10.0 + 1.3 * i for i in range(m.n)])
    X = np.diag(x)
    o = m.evolvex(1, t, psi.real, psi.imag)
    np.testing.assert_allclose(o[0] + 1j * o[1], (X @ U - U @ X) @ psi, atol=1e-10)
    Um = U.conj().T
    o = m.evolvex(-1, t, psi.real, psi.imag)
    np.testing.assert_allclose(o[0] + 1j * o[1], (Um @ X - X @ Um) @ psi, atol=1e-10)


def test_unitarity_and_inverse(tmp_path):
    m = _model(tmp_path, "ring", _general_ring)
    psi = _random_complex(m.n, 5)
    a = m.evolve(1, 7.5, psi.real, psi.imag)
    assert abs(np.linalg.norm(a[0] + 1j * a[1]) - np.linalg.norm(psi)) < 1e-12
    b = m.evolve(-1, 7.5, *a)
    np.testing.assert_allclose(b[0] + 1j * b[1], psi, atol=1e-12)


def test_summation_integral(tmp_path):
    """Gauss-Chebyshev: sum over nodes x_k of rho(x_k) sqrt(1-x_k^2) pi/K
    = int rho dE / E_max... = 2 mu_0 / (volume E_max) for undamped moments."""
    M, K = 30, 200
    mu = np.random.default_rng(0).standard_normal(M)
    xk = np.cos(np.pi * (np.arange(K) + 0.5) / K)
    emax, vol = 3.0, 7.0
    rho = oracle.summation(mu, xk * emax, emax, vol)
    integral = (rho * np.sqrt(1 - xk ** 2)).sum() * np.pi / K * emax
    assert abs(integral - 2 * mu[0] / vol) < 1e-12
    # first moment: int x rho(x) dx recovers mu_1
    first = (rho * xk * np.sqrt(1 - xk ** 2)).sum() * np.pi / K * emax
    assert abs(first - 2 * mu[1] / vol) < 1e-12


def test_jackson_kernel_properties():
    M = 100
    g, _ = oracle.damping(np.ones(M))
    assert abs(g[0] - 1) < 1e-15
    assert np.all(np.diff(g) < 0) and g[-1] > 0
    # the compile-time Lorentz variant computes k / M in integer arithmetic
    # (src/sigma.cu:70), so every factor is sinh(lambda) / sinh(lambda) = 1
    # (to the last bit of the two evaluations of sinh(4))
    gl, _ = oracle.damping(np.ones(M), lorentz=True)
    assert np.all(gl == gl[0]) and abs(gl[0] - 1.0) <= 2.3e-16


@pytest.mark.parametrize("t,last", [(10.0, 33), (20.0, 49), (1.22, 14), (410.0, 486)])
def test_bessel_stop(t, last):
    """the evolution keeps terms m >= 2 while |J_m(t)| >= 1e-15: the first
    dropped index is last + 1 (term counts measured for the reference)"""
    stop = oracle.bessel_stop(t)
    assert stop == last + 1
    assert abs(scipy.special.jv(stop, t)) < 1e-15 * 1.01
    assert abs(scipy.special.jv(last, t)) >= 1e-15 * 0.99


def test_spin_state_and_ldos_state(tmp_path):
    m = _model(tmp_path, "ring", _general_ring)
    re, im = m.random_state(orbital=5)
    assert re[5] == np.sqrt(m.n) and np.count_nonzero(re) == 1 and not im.any()
