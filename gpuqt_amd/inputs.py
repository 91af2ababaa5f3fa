"""Writers for GPUQT's input files.

The formats are the reference's (doc/gpuqt.tex:327-496): ``para.in``
keyword lines (src/model.cu:243-310), ``energy.in`` / ``time_step.in`` as a
count followed by values (src/model.cu:312-334, 363-384), ``lattice.in``
(src/model_lattice.cu:165-341), and the general-model files ``neighbor.in``,
``position.in``, ``potential.in``, ``hopping.in`` (src/model_general.cu:32-185)
and ``local_orbitals.in`` (src/model.cu:336-361).

Also provides the concrete lattices used by the workloads: the 4-orbital
rectangular honeycomb cell of examples/review/graphene/vacancy/lattice.in
(nearest-neighbour hopping -2.7 eV), the single-orbital square lattice of
examples/fan2018cpc/lattice/diffusive/lattice.in, and a simple-cubic general
model with a Peierls phase on y-bonds.
"""
from __future__ import annotations

import math
import os
from typing import Iterable, Sequence

import numpy as np


def _write(path: str, text: str) -> None:
    with open(path, "w") as f:
        f.write(text)


def write_para(dirname: str, *, model: int = 1, anderson: float | None = None,
               charged: tuple[int, float, float] | None = None,
               vacancies: int | None = None, vac0: bool = False, vac: bool = False,
               msd: bool = False, spin: bool = False, ldos: bool = False,
               random_vectors: int | None = None, moments: int | None = None,
               energy_max: float | None = None, extra_lines: Sequence[str] = ()) -> None:
    lines = [f"model {model}"]
    if anderson is not None:
        lines.append(f"anderson_disorder {anderson!r}")
    if charged is not None:
        lines.append(f"charged_impurity {charged[0]} {charged[1]!r} {charged[2]!r}")
    if vacancies is not None:
        lines.append(f"vacancy_disorder {vacancies}")
    if random_vectors is not None:
        lines.append(f"number_of_random_vectors {random_vectors}")
    if moments is not None:
        lines.append(f"number_of_moments {moments}")
    if energy_max is not None:
        lines.append(f"energy_max {energy_max!r}")
    for flag, name in ((vac0, "calculate_vac0"), (vac, "calculate_vac"),
                       (msd, "calculate_msd"), (spin, "calculate_spin"),
                       (ldos, "calculate_ldos")):
        if flag:
            lines.append(name)
    lines.extend(extra_lines)
    _write(os.path.join(dirname, "para.in"), "\n".join(lines) + "\n")


def write_counted(path: str, values: Iterable[float], fmt: str = "{:.6f}") -> None:
    vals = list(values)
    _write(path, f"{len(vals)}\n" + "".join(fmt.format(v) + "\n" for v in vals))


def write_energy(dirname: str, energies: Iterable[float]) -> None:
    write_counted(os.path.join(dirname, "energy.in"), energies)


def write_time_steps(dirname: str, steps: Iterable[float]) -> None:
    write_counted(os.path.join(dirname, "time_step.in"), steps)


def write_local_orbitals(dirname: str, orbitals: Iterable[int]) -> None:
    orbs = list(orbitals)
    _write(os.path.join(dirname, "local_orbitals.in"),
           f"{len(orbs)}\n" + "".join(f"{o}\n" for o in orbs))


def energy_grid(lo: float, hi: float, n: int) -> list[float]:
    return [lo + (hi - lo) * i / (n - 1) for i in range(n)]


# --- lattice model ----------------------------------------------------------

# 4-orbital rectangular honeycomb cell with a_cc = 1 (ax = sqrt(3), ay = 3),
# as examples/review/graphene/vacancy/lattice.in.
HONEYCOMB_ORBITALS = [(0.866, 0.0, 0.0), (0.0, 0.5, 0.0), (0.0, 1.5, 0.0), (0.866, 2.0, 0.0)]
HONEYCOMB_HOPS = [
    [(1, 0, 0, 1), (0, 0, 0, 1), (0, -1, 0, 3)],
    [(0, 0, 0, 0), (-1, 0, 0, 0), (0, 0, 0, 2)],
    [(0, 0, 0, 3), (-1, 0, 0, 3), (0, 0, 0, 1)],
    [(1, 0, 0, 2), (0, 0, 0, 2), (0, 1, 0, 0)],
]


def write_lattice(dirname: str, cells: Sequence[int], pbc: Sequence[int], transport: int,
                  constants: Sequence[float], orbitals: Sequence[Sequence[float]],
                  hops: Sequence[Sequence[tuple]], max_neighbor: int | None = None) -> None:
    """hops[o] is a list of (nx, ny, nz, orbital, h_re, h_im)."""
    if max_neighbor is None:
        max_neighbor = max(len(h) for h in hops)
    lines = [" ".join(str(c) for c in cells),
             " ".join(str(p) for p in pbc) + f" {transport}",
             " ".join(repr(float(c)) for c in constants),
             f"{len(orbitals)} {max_neighbor}"]
    for o in orbitals:
        lines.append(" ".join(repr(float(c)) for c in o))
    for h in hops:
        lines.append(str(len(h)))
        for (nx, ny, nz, orb, hr, hi) in h:
            lines.append(f"{nx} {ny} {nz} {orb} {hr!r} {hi!r}")
    _write(os.path.join(dirname, "lattice.in"), "\n".join(lines) + "\n")


def read_lattice(path: str) -> dict:
    """The contents of a lattice.in (src/model_lattice.cu:165-341) as the
    tables of lsqt_ham_create_lattice."""
    tok = open(path).read().split()
    pos = 0

    def take(kind, count=1):
        nonlocal pos
        vals = [kind(t) for t in tok[pos:pos + count]]
        pos += count
        return vals if count > 1 else vals[0]

    cells = take(int, 3)
    pbc = take(int, 3)
    transport = take(int)
    a = take(float, 3)
    No, K = take(int), take(int)
    xyz = np.array([take(float, 3) for _ in range(No)])
    count = np.zeros(No, dtype=np.int32)
    cell = np.zeros((No, K, 3), dtype=np.int32)
    orb = np.zeros((No, K), dtype=np.int32)
    re = np.zeros((No, K))
    im = np.zeros((No, K))
    for m in range(No):
        count[m] = take(int)
        for i in range(count[m]):
            cell[m, i] = take(int, 3)
            orb[m, i] = take(int)
            re[m, i], im[m, i] = take(float), take(float)
    return {"cells": cells, "pbc": pbc, "transport": transport, "lattice_constant": a,
            "orbitals": No, "max_neighbor": K, "orbital_x": xyz[:, 0], "hop_count": count,
            "hop_cell": cell, "hop_orbital": orb, "hop_re": re, "hop_im": im}


def write_honeycomb(dirname: str, nx: int, ny: int, *, pbc=(1, 1, 0), transport: int = 0,
                    t: float = -2.7, ax: float = 1.7321, ay: float = 3.0) -> None:
    hops = [[(a, b, c, o, t, 0.0) for (a, b, c, o) in row] for row in HONEYCOMB_HOPS]
    orbitals = [(x * ax / 1.7321, y * ay / 3.0, z) for (x, y, z) in HONEYCOMB_ORBITALS]
    write_lattice(dirname, (nx, ny, 1), pbc, transport, (ax, ay, 1.0), orbitals, hops)


def write_square(dirname: str, nx: int, ny: int, *, pbc=(1, 1, 0), transport: int = 0,
                 t: float = -1.0) -> None:
    hops = [[(-1, 0, 0, 0, t, 0.0), (1, 0, 0, 0, t, 0.0), (0, 1, 0, 0, t, 0.0),
             (0, -1, 0, 0, t, 0.0)]]
    write_lattice(dirname, (nx, ny, 1), pbc, transport, (1.0, 1.0, 1.0), [(0.0, 0.0, 0.0)],
                  hops)


# --- general model ----------------------------------------------------------

def write_general(dirname: str, neighbors: Sequence[Sequence[int]], max_neighbor: int,
                  positions: Sequence[float], box: float, volume: float,
                  potential: Sequence[float] | None = None,
                  hoppings: Sequence[Sequence[complex]] | None = None,
                  hopping_kind: str = "complex") -> None:
    n = len(neighbors)
    with open(os.path.join(dirname, "neighbor.in"), "w") as f:
        f.write(f"{n} {max_neighbor}\n")
        for row in neighbors:
            f.write(" ".join([str(len(row))] + [str(j) for j in row]) + "\n")
    with open(os.path.join(dirname, "position.in"), "w") as f:
        f.write(f"{box!r} {volume!r}\n")
        for x in positions:
            f.write(f"{float(x)!r}\n")
    if potential is not None:
        with open(os.path.join(dirname, "potential.in"), "w") as f:
            for v in potential:
                f.write(f"{float(v):.17g}\n")
    if hoppings is not None:
        with open(os.path.join(dirname, "hopping.in"), "w") as f:
            f.write(hopping_kind + "\n")
            for row in hoppings:
                if hopping_kind == "complex":
                    f.write(" ".join(f"{h.real:.17g} {h.imag:.17g}" for h in row) + "\n")
                else:
                    f.write(" ".join(f"{h.real:.17g}" for h in row) + "\n")


def cubic_peierls(L: int, *, t: float = -1.0, flux: float = 0.01):
    """Simple-cubic L^3 lattice, periodic, neighbour order -x,+x,-y,+y,-z,+z.

    y-bonds from a site in cell column x carry t*exp(+i 2 pi flux x) going +y
    and the Hermitian conjugate going -y; site index is (x*L + y)*L + z.
    Returns (neighbors, hoppings, positions)."""
    idx = lambda x, y, z: ((x % L) * L + (y % L)) * L + (z % L)
    neighbors, hoppings, positions = [], [], []
    for x in range(L):
        for y in range(L):
            for z in range(L):
                ph = complex(math.cos(2 * math.pi * flux * x), math.sin(2 * math.pi * flux * x))
                neighbors.append([idx(x - 1, y, z), idx(x + 1, y, z), idx(x, y - 1, z),
                                  idx(x, y + 1, z), idx(x, y, z - 1), idx(x, y, z + 1)])
                hoppings.append([complex(t, 0), complex(t, 0), t * ph.conjugate(), t * ph,
                                 complex(t, 0), complex(t, 0)])
                positions.append(float(x))
    return neighbors, hoppings, positions


def cubic_peierls_arrays(L: int = 160, W: float = 16.5, flux: float = 0.01, seed: int = 2024):
    """Simple-cubic L^3 Anderson model, hopping -1, on-site U[-W/2, W/2],
    Peierls phase 2 pi flux x on y-bonds (+y: -exp(+i 2 pi flux x), -y: its
    conjugate), periodic, neighbour order -x, +x, -y, +y, -z, +z, site index
    (x L + y) L + z; xx = transport (x) offset with the minimum image.
    Returns the row-major ELL arrays the general model would build."""
    n = L ** 3
    idx = np.arange(n, dtype=np.int64)
    x, y, z = idx // (L * L), (idx // L) % L, idx % L
    site = lambda a, b, c: ((a % L) * L + (b % L)) * L + (c % L)
    nl = np.stack([site(x - 1, y, z), site(x + 1, y, z), site(x, y - 1, z), site(x, y + 1, z),
                   site(x, y, z - 1), site(x, y, z + 1)], axis=1).astype(np.int32)
    ph = np.exp(2j * np.pi * flux * x)
    hop = np.empty((n, 6), dtype=complex)
    hop[:, [0, 1, 4, 5]] = -1.0
    hop[:, 2] = -np.conj(ph)
    hop[:, 3] = -ph
    xx = np.zeros((n, 6))
    xx[:, 0], xx[:, 1] = -1.0, 1.0
    pot = np.random.default_rng(seed).uniform(-W / 2, W / 2, n)
    return dict(neighbor_number=np.full(n, 6, np.int32), neighbor_list=nl.reshape(-1),
                potential=pot, hopping_real=hop.real.reshape(-1).copy(),
                hopping_imag=hop.imag.reshape(-1).copy(), xx=xx.reshape(-1)), n


def anderson_potential(n: int, W: float, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return rng.uniform(-W / 2, W / 2, size=n)
