"""Randomised small inputs (CPU): lattices of random shape, orbitals,
Hermitian complex hoppings, boundaries and disorder (Anderson, charged
impurities, vacancies), and general-model systems (ragged neighbour graphs,
complex or real hopping files, optional potentials, spin), with random
moment counts, energies, time steps and quantities.  For every case the oracle's .out files equal, text for text,
those of the reference program compiled from source (oracle/_ref), and the
drop-in's host Model builds the oracle's arrays and random vectors bit for
bit.  Skipped where the reference program has not been built."""
import os
import shutil

import numpy as np
import pytest

import oracle
from gpuqt_amd import hostlib
from gpuqt_amd import inputs as gi
from parity import assert_bits

pytestmark = pytest.mark.skipif(oracle.reference_binary() is None,
                                reason="the reference program is not built (make -C oracle ref)")

SEEDS = list(range(16))


def _random_case(d, seed):
    """Writes a random, valid input directory (inputs on which the reference
    itself misbehaves are avoided: charged-impurity cell lists of fewer than
    3 cells per periodic direction)."""
    rng = np.random.default_rng(1000 + seed)
    No = int(rng.integers(1, 4))
    cells = [int(rng.integers(3, 7)), int(rng.integers(1, 6)), int(rng.integers(1, 4))]
    pbc = [int(rng.integers(0, 2)) for _ in range(3)]
    consts = [float(rng.uniform(1.0, 3.0)) for _ in range(3)]
    orbitals = [tuple(float(v) for v in rng.uniform(0, 1, 3) * consts) for _ in range(No)]
    hops = [[] for _ in range(No)]
    K = int(rng.integers(2, 5))
    for _ in range(int(rng.integers(No, 3 * No + 1))):
        o1, o2 = int(rng.integers(0, No)), int(rng.integers(0, No))
        dvec = [int(rng.integers(-1, 2)) if cells[k] >= 3 else 0 for k in range(3)]
        if o1 == o2 and dvec == [0, 0, 0]:
            continue
        if len(hops[o1]) >= K or len(hops[o2]) >= K:
            continue
        if any((a, b, c, o) == (*dvec, o2) for (a, b, c, o, _, _) in hops[o1]):
            continue
        t = complex(rng.uniform(-2, 2), rng.uniform(-1, 1))
        hops[o1].append((*dvec, o2, t.real, t.imag))
        if (o1, dvec) != (o2, [-v for v in dvec]):
            hops[o2].append((-dvec[0], -dvec[1], -dvec[2], o1, t.real, -t.imag))
    if sum(len(h) for h in hops) == 0:
        hops[0].append((1, 0, 0, 0, -1.0, 0.0))
        hops[0].append((-1, 0, 0, 0, -1.0, 0.0))
    transport = int(rng.integers(0, 3))
    pbc[transport] = 1  # the transport direction must be periodic
    gi.write_lattice(d, cells, pbc, transport, consts, orbitals, hops,
                     max_neighbor=max(len(h) for h in hops))
    n = No * cells[0] * cells[1] * cells[2]
    anderson = float(rng.uniform(0.2, 3.0)) if rng.random() < 0.6 else None
    charged = None
    box = [cells[k] * consts[k] for k in range(3)]
    periodic = [box[k] for k in range(3) if pbc[k]]
    xi = min(periodic) / 16 if periodic else 0.3
    if rng.random() < 0.4 and xi >= 0.05:
        charged = (int(rng.integers(1, 4)), float(rng.uniform(0.5, 3.0)), float(xi))
    vacancies = int(rng.integers(1, max(2, n // 8))) if rng.random() < 0.3 and n >= 16 else None
    ldos = rng.random() < 0.3
    tmax = max(sum(abs(complex(h[4], h[5])) for h in row) for row in hops)
    emax = 1.2 * (tmax + (anderson or 0) / 2 + (charged[1] if charged else 0)) + 0.5
    flags = dict(vac0=rng.random() < 0.5, vac=rng.random() < 0.5, msd=rng.random() < 0.5)
    # (the lattice model has no spin option: src/model.cu rejects it)
    gi.write_para(d, anderson=anderson, charged=charged, vacancies=vacancies, ldos=ldos,
                  random_vectors=int(rng.integers(1, 3)), moments=int(rng.integers(2, 41)),
                  energy_max=emax, **flags)
    gi.write_energy(d, np.sort(rng.uniform(-0.9, 0.9, int(rng.integers(1, 6))) * emax))
    gi.write_time_steps(d, rng.uniform(0.1, 2.0, int(rng.integers(1, 4))))
    if ldos:
        gi.write_local_orbitals(d, sorted(set(int(v) for v in rng.integers(0, 4, 2))))


def _random_general_case(d, seed):
    """A random general-model directory: symmetric neighbour graph with
    ragged degrees, Hermitian hoppings (complex or real files), positions in
    a periodic box, an optional potential file; spin (even site counts) and
    LDOS sometimes."""
    rng = np.random.default_rng(5000 + seed)
    n = 2 * int(rng.integers(2, 60))
    K = int(rng.integers(1, 6))
    nb = [[] for _ in range(n)]
    hop = {}
    real = rng.random() < 0.4
    pairs = [(i, j) for i in range(n) for j in range(i + 1, n)]
    rng.shuffle(pairs)
    for i, j in pairs[: 2 * n]:
        if len(nb[i]) < K and len(nb[j]) < K:
            t = complex(rng.uniform(-2, 2), 0.0 if real else rng.uniform(-1, 1))
            nb[i].append(j)
            nb[j].append(i)
            hop[(i, j)], hop[(j, i)] = t, t.conjugate()
    box = float(rng.uniform(n / 2, 2 * n))
    potential = rng.uniform(-1, 1, n) if rng.random() < 0.6 else None
    gi.write_general(d, nb, K, rng.uniform(0, box, n), box, float(n), potential=potential,
                     hoppings=[[hop[(i, j)] for j in nb[i]] for i in range(n)],
                     hopping_kind="real" if real else "complex")
    tmax = max([sum(abs(hop[(i, j)]) for j in nb[i]) for i in range(n)] + [0.0])
    emax = 1.2 * (tmax + 1.0) + 0.5
    ldos = rng.random() < 0.3
    flags = dict(vac0=rng.random() < 0.5, vac=rng.random() < 0.5, msd=rng.random() < 0.5,
                 spin=rng.random() < 0.3)
    gi.write_para(d, model=0, ldos=ldos, random_vectors=int(rng.integers(1, 3)),
                  moments=int(rng.integers(2, 41)), energy_max=emax, **flags)
    gi.write_energy(d, np.sort(rng.uniform(-0.9, 0.9, int(rng.integers(1, 6))) * emax))
    gi.write_time_steps(d, rng.uniform(0.1, 2.0, int(rng.integers(1, 4))))
    if ldos:
        gi.write_local_orbitals(d, sorted(set(int(v) for v in rng.integers(0, n, 2))))


def make_case(d, case):
    """case: seed for a random lattice, ("general", seed) for a random
    general-model system"""
    if isinstance(case, tuple):
        _random_general_case(d, case[1])
    else:
        _random_case(d, case)


CASES = SEEDS + [("general", s) for s in range(12)]


@pytest.mark.parametrize("seed", CASES)
def test_random_case(tmp_path, seed):
    base = tmp_path / "case"
    base.mkdir()
    make_case(str(base), seed)
    ref_dir, orc_dir = tmp_path / "ref", tmp_path / "orc"
    shutil.copytree(base, ref_dir)
    shutil.copytree(base, orc_dir)
    driver = tmp_path / "input.txt"
    driver.write_text(f"{ref_dir}\n")
    r = oracle.run_reference(str(driver), timeout=120)
    assert r.returncode == 0, r.stdout[-2000:]
    oracle.lsqt(str(orc_dir))
    outs = sorted(f for f in os.listdir(ref_dir) if f.endswith(".out"))
    assert outs and outs == sorted(f for f in os.listdir(orc_dir) if f.endswith(".out"))
    for name in outs:
        assert (orc_dir / name).read_text() == (ref_dir / name).read_text(), name
    # the drop-in's host model: the oracle's arrays and random vectors
    m = hostlib.HostModel(str(base))
    o = oracle.OracleModel(str(base))
    a, b = m.arrays(), o.arrays()
    np.testing.assert_array_equal(a["neighbor_number"], b["neighbor_number"])
    vm, vo = m.valid_pairs(), o.valid_pairs()
    np.testing.assert_array_equal(vm["neighbor_list"], vo["neighbor_list"])
    for k in ("hopping_real", "hopping_imag", "xx"):
        assert_bits(vm[k], vo[k], k)
    assert_bits(a["potential"], b["potential"], "potential")
    for _ in range(m.number_of_random_vectors):
        x, y = m.random_state(), o.random_state()
        assert_bits(x[0], y[0], "random re")
        assert_bits(x[1], y[1], "random im")
