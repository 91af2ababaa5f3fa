"""The benchmarked workloads as input directories and seeded vectors.

Shared by bench.py, the full-size parity tests and the script that writes
their fixtures, so that every one of them builds the same systems:

* the FP64 / precision job: an 8.0M-site honeycomb (2000 x 1000 four-orbital
  cells, a_cc = 1, t = -2.7 eV, periodic), Anderson disorder U[-1, 1] eV,
  3000 moments, E_max = 10 eV, 201 energies in [-8, 8] eV, MSD over time
  steps of 2 hbar/eV;
* the block job: the 160^3 simple-cubic Anderson model with a Peierls phase
  (inputs.cubic_peierls_arrays), E_max = 15, vectors of seeded random phases;
* the uneven batch: honeycomb realisations with (Nc1, Nc2) cells drawn
  uniformly in [200, 1500]^2, a_cc = 1.42 A, charged impurities at 0.1 %
  (amplitudes U[-2.8, 2.8] eV, xi = 4.26 A), R random vectors each, VAC and
  MSD over time steps of 2 hbar/eV, E_max = 13.5 eV.  E_max must exceed the
  spectral radius of every realisation: the Gershgorin bound max|V| + 3|t|
  is at most 5.21 + 8.1 = 13.31 eV over the 64 realisations (realisation 62
  has an eigenvalue at 12.01 eV, so E_max = 12 eV would leave its Chebyshev
  expansion outside [-1, 1]).
"""
from __future__ import annotations

import os

import numpy as np

from . import inputs as gi

CONFIG2_CELLS = (2000, 1000)
CUBE_L = 160
CUBE_EMAX = 15.0
CUBE_ENERGIES = np.linspace(-14, 14, 201)
BATCH_EMAX = 13.5
BATCH_DT = 2.0


def write_config2(d, nx=2000, ny=1000, moments=3000, time_steps=20, dt=2.0, energies=201,
                  msd=True):
    """The honeycomb DOS (+ MSD) job's input directory."""
    gi.write_honeycomb(d, nx, ny)
    gi.write_para(d, anderson=2.0, msd=msd, moments=moments, energy_max=10.0)
    gi.write_energy(d, gi.energy_grid(-8, 8, energies))
    gi.write_time_steps(d, [dt] * time_steps)


def cube_phases(R, n, seed=7):
    """R vectors of unit-modulus entries with seeded random phases
    (re, im), each of shape (R, n)."""
    ph = np.random.default_rng(seed).uniform(0, 2 * np.pi, (R, n))
    return np.cos(ph), np.sin(ph)


def batch_cells(count, lo=200, hi=1500, seed=5):
    """Cell counts of the uneven batch: (Nc1, Nc2) uniform in [lo, hi]^2 from
    a fixed seed; a four-orbital cell holds two two-site cells, so Nx = Nc1,
    Ny = Nc2 / 2 (Nc2 rounded up to even)."""
    c = np.random.default_rng(seed).integers(lo, hi + 1, size=(count, 2))
    return [(int(a), int((b + 1) // 2)) for a, b in c]


def write_batch_realisation(d, nx, ny, R=8, moments=1500, n_steps=16, dt=BATCH_DT,
                            emax=BATCH_EMAX):
    """One realisation of the uneven batch as an input directory."""
    os.makedirs(d, exist_ok=True)
    n = 4 * nx * ny
    gi.write_honeycomb(d, nx, ny, ax=2.4595, ay=4.26)
    gi.write_para(d, charged=(int(round(0.001 * n)), 5.6, 4.26), moments=moments,
                  energy_max=emax, vac=True, msd=True, random_vectors=R)
    gi.write_energy(d, gi.energy_grid(-8, 8, 201))
    gi.write_time_steps(d, [dt] * n_steps)


def gershgorin_bound(arrays, max_neighbor):
    """max_n (|V_n| + sum_m |t_nm|): an upper bound of the spectral radius of
    the Hamiltonian of row-major ELL arrays."""
    nn = np.asarray(arrays["neighbor_number"])
    K = max_neighbor
    valid = np.arange(K)[None, :] < nn[:, None]
    t = np.hypot(np.asarray(arrays["hopping_real"]).reshape(-1, K),
                 np.asarray(arrays["hopping_imag"]).reshape(-1, K))
    return float((np.abs(arrays["potential"]) + np.where(valid, t, 0.0).sum(axis=1)).max())


def lattice_gershgorin_bound(lattice, potential):
    """The same bound for a lattice model: max |V| plus the largest sum of
    |t| over the hoppings of one orbital (a site keeps at most those)."""
    hop = np.hypot(np.asarray(lattice["hop_re"]), np.asarray(lattice["hop_im"]))
    counts = np.asarray(lattice["hop_count"])
    per_orbital = [hop[m, :counts[m]].sum() for m in range(lattice["orbitals"])]
    return float(np.abs(potential).max() + max(per_orbital))
