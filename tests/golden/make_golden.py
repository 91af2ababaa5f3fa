"""Generates the golden fixtures under tests/golden/ from the reference.

Run in a container where /root/reference exists, after `make -C oracle ref`
(which compiles the reference's CPU program and oracle/ref_harness.cpp from
the reference sources into oracle/_ref/):

    python tests/golden/make_golden.py

For every case below it writes the case's input files into
tests/golden/cases/<case>/, runs the reference program `oracle/_ref/lsqt_cpu`
on a scratch copy (its .out files, 6 significant digits, are kept as
expected/<name>.out), and runs `oracle/_ref/ref_harness`, whose hex-float
dump of every intermediate result (model arrays, random vectors, operator
applications, evolutions, raw moments, damped-and-summed rows) is stored as
dump.npz.  Failure cases keep the reference's exit status and the last line
it printed.

The fixtures are data: inputs and the reference's outputs.
"""
from __future__ import annotations

import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from gpuqt_amd import inputs as gi  # noqa: E402

REF_BIN = os.path.join(ROOT, "oracle", "_ref", "lsqt_cpu")
REF_HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
CASES_DIR = os.path.join(HERE, "cases")
FAIL_DIR = os.path.join(HERE, "failures")


def case_honeycomb_small(d):
    gi.write_honeycomb(d, 6, 3)
    gi.write_para(d, anderson=1.0, vac0=True, vac=True, msd=True, random_vectors=2,
                  moments=64, energy_max=10.0)
    gi.write_energy(d, gi.energy_grid(-8, 8, 11))
    gi.write_time_steps(d, [1.0, 1.0, 1.0])


def case_square_vacancy_ldos(d):
    gi.write_square(d, 8, 8)
    gi.write_para(d, anderson=2.0, vacancies=6, vac=True, msd=True, ldos=True, moments=48,
                  energy_max=6.0)
    gi.write_energy(d, gi.energy_grid(-5, 5, 13))
    gi.write_time_steps(d, [0.5, 1.5])
    gi.write_local_orbitals(d, [0, 17, 57])


def case_strip_open(d):
    gi.write_square(d, 10, 4, pbc=(1, 0, 0))
    gi.write_para(d, anderson=1.0, msd=True, vac0=True, moments=40, energy_max=5.5)
    gi.write_energy(d, gi.energy_grid(-4, 4, 9))
    gi.write_time_steps(d, [2.0, 0.3])


def case_honeycomb_transport_y(d):
    gi.write_honeycomb(d, 4, 4, pbc=(1, 1, 0), transport=1)
    gi.write_para(d, anderson=0.5, vac=True, msd=True, moments=32, energy_max=9.0)
    gi.write_energy(d, gi.energy_grid(-7, 7, 8))
    gi.write_time_steps(d, [1.3, 0.7, 2.1])


def case_charged_vacancy_anderson(d):
    # 12x6 cells: box 20.79 x 18 > 3 * 5 * xi in both periodic directions
    gi.write_honeycomb(d, 12, 6)
    gi.write_para(d, anderson=1.0, charged=(5, 2.0, 1.0), vacancies=4, vac0=True,
                  moments=32, energy_max=10.0)
    gi.write_energy(d, gi.energy_grid(-8, 8, 9))


def case_charged_only(d):
    gi.write_honeycomb(d, 12, 6)
    gi.write_para(d, charged=(7, 2.8, 1.2), msd=True, moments=32, energy_max=10.0)
    gi.write_energy(d, gi.energy_grid(-8, 8, 9))
    gi.write_time_steps(d, [1.0, 2.0])


def case_cube_peierls(d):
    L = 4
    nb, hop, pos = gi.cubic_peierls(L, flux=0.07)
    pot = gi.anderson_potential(L ** 3, 8.25 * 2 / 2, seed=7)
    gi.write_general(d, nb, 6, pos, box=float(L), volume=float(L ** 3), potential=pot,
                     hoppings=hop, hopping_kind="complex")
    gi.write_para(d, model=0, vac0=True, vac=True, msd=True, moments=40, energy_max=15.0)
    gi.write_energy(d, gi.energy_grid(-12, 12, 7))
    gi.write_time_steps(d, [0.4, 0.8])


def case_general_defaults(d):
    # ring of 20 sites, no hopping.in / potential.in: uniform -1, zero potential
    n = 20
    nb = [[(i - 1) % n, (i + 1) % n] for i in range(n)]
    gi.write_general(d, nb, 3, [float(i) for i in range(n)], box=float(n), volume=float(n))
    gi.write_para(d, model=0, msd=True, vac=True, moments=24, energy_max=2.5)
    gi.write_energy(d, gi.energy_grid(-2, 2, 5))
    gi.write_time_steps(d, [3.0])


def case_general_real_spin(d):
    # 2-orbital (spin up/down interleaved) ring with a real spin-mixing hop
    cells = 12
    n = 2 * cells
    nb, hop = [], []
    for c in range(cells):
        for s in range(2):
            i = 2 * c + s
            left, right = 2 * ((c - 1) % cells) + s, 2 * ((c + 1) % cells) + s
            flip = 2 * c + (1 - s)
            nb.append([left, right, flip])
            hop.append([complex(-1.0), complex(-1.0), complex(0.3)])
    pot = gi.anderson_potential(n, 1.0, seed=3)
    gi.write_general(d, nb, 3, [float(i // 2) for i in range(n)], box=float(cells),
                     volume=float(cells), potential=pot, hoppings=hop, hopping_kind="real")
    gi.write_para(d, model=0, spin=True, moments=32, energy_max=3.5, random_vectors=2)
    gi.write_energy(d, gi.energy_grid(-3, 3, 7))
    gi.write_time_steps(d, [0.5, 1.0, 1.5])


def case_config1(d):
    """Honeycomb 20,000 sites (100 x 50 four-orbital cells), t = -2.7 eV,
    periodic, Anderson W = 1 eV (U[-0.5, 0.5]), 1 random vector, 1000
    moments, E_max = 10 eV, 201 energies in [-8, 8] eV, DOS + VAC + MSD over
    10 steps of 1 hbar/eV."""
    gi.write_honeycomb(d, 100, 50)
    gi.write_para(d, anderson=1.0, vac=True, msd=True, moments=1000, energy_max=10.0)
    gi.write_energy(d, gi.energy_grid(-8, 8, 201))
    gi.write_time_steps(d, [1.0] * 10)


CASES = {
    "honeycomb_small": (case_honeycomb_small, True),
    "square_vacancy_ldos": (case_square_vacancy_ldos, True),
    "strip_open": (case_strip_open, True),
    "honeycomb_transport_y": (case_honeycomb_transport_y, True),
    "charged_vacancy_anderson": (case_charged_vacancy_anderson, True),
    "charged_only": (case_charged_only, True),
    "cube_peierls": (case_cube_peierls, True),
    "general_defaults": (case_general_defaults, True),
    "general_real_spin": (case_general_real_spin, True),
    # moments and rows only: state dumps of 20,000 sites would be too large
    "config1_honeycomb_20k": (case_config1, False),
}


# --- failure paths: (inputs writer) -> expected exit status + last line ------

def fail_unknown_keyword(d):
    case_honeycomb_small(d)
    with open(os.path.join(d, "para.in"), "a") as f:
        f.write("calculate_everything\n")


def fail_spin_lattice(d):
    case_honeycomb_small(d)
    gi.write_para(d, spin=True, moments=8)


def fail_spin_with_msd(d):
    case_general_real_spin(d)
    gi.write_para(d, model=0, spin=True, msd=True, moments=8)


def fail_general_anderson(d):
    case_general_defaults(d)
    gi.write_para(d, model=0, anderson=1.0, moments=8)


def fail_missing_energy(d):
    case_honeycomb_small(d)
    os.remove(os.path.join(d, "energy.in"))


def fail_bad_boundary(d):
    case_honeycomb_small(d)
    gi.write_lattice(d, (2, 2, 1), (1, 2, 0), 0, (1.0, 1.0, 1.0), [(0.0, 0.0, 0.0)],
                     [[(1, 0, 0, 0, -1.0, 0.0)]])


def fail_open_transport(d):
    case_honeycomb_small(d)
    gi.write_lattice(d, (2, 2, 1), (0, 1, 0), 0, (1.0, 1.0, 1.0), [(0.0, 0.0, 0.0)],
                     [[(1, 0, 0, 0, -1.0, 0.0)]])


def fail_bad_local_orbital(d):
    case_square_vacancy_ldos(d)
    gi.write_local_orbitals(d, [3, 9999])


def fail_zero_moments(d):
    case_honeycomb_small(d)
    gi.write_para(d, moments=0)


def fail_negative_energy_max(d):
    case_honeycomb_small(d)
    gi.write_para(d, energy_max=-1.0)


def fail_bad_hopping_kind(d):
    case_general_defaults(d)
    with open(os.path.join(d, "hopping.in"), "w") as f:
        f.write("imaginary\n")


def fail_zero_vacancies(d):
    case_honeycomb_small(d)
    gi.write_para(d, vacancies=0, moments=8)


def fail_missing_lattice(d):
    case_honeycomb_small(d)
    os.remove(os.path.join(d, "lattice.in"))


FAILURES = {
    "unknown_keyword": fail_unknown_keyword,
    "spin_lattice": fail_spin_lattice,
    "spin_with_msd": fail_spin_with_msd,
    "general_anderson": fail_general_anderson,
    "missing_energy": fail_missing_energy,
    "bad_boundary": fail_bad_boundary,
    "open_transport": fail_open_transport,
    "bad_local_orbital": fail_bad_local_orbital,
    "zero_moments": fail_zero_moments,
    "negative_energy_max": fail_negative_energy_max,
    "bad_hopping_kind": fail_bad_hopping_kind,
    "zero_vacancies": fail_zero_vacancies,
    "missing_lattice": fail_missing_lattice,
}

INPUT_NAMES = ("para.in", "energy.in", "time_step.in", "lattice.in", "neighbor.in",
               "position.in", "potential.in", "hopping.in", "local_orbitals.in")


def parse_dump(path):
    arrays = {}
    with open(path) as f:
        lines = f.read().split("\n")
    i = 0
    while i < len(lines) and lines[i]:
        name, count = lines[i].split()
        count = int(count)
        vals = lines[i + 1:i + 1 + count]
        if any(("p" in v or "x" in v) for v in vals[:4]) or name in ("fmeta",):
            arr = np.array([float.fromhex(v) for v in vals], dtype=np.float64)
        else:
            arr = np.array([int(v) for v in vals], dtype=np.int64)
        arrays[name] = arr
        i += 1 + count
    return arrays


def run_reference(case_dir, with_states):
    with tempfile.TemporaryDirectory() as tmp:
        work = os.path.join(tmp, "case")
        shutil.copytree(case_dir, work, ignore=shutil.ignore_patterns("expected", "dump.npz"))
        driver = os.path.join(tmp, "input.txt")
        with open(driver, "w") as f:
            f.write(work + "\n")
        r = subprocess.run([REF_BIN, driver], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"reference failed on {case_dir}: {r.stdout[-500:]}")
        outs = {}
        for name in ("dos.out", "ldos.out", "vac0.out", "vac.out", "msd.out", "S.out"):
            p = os.path.join(work, name)
            if os.path.exists(p):
                with open(p) as f:
                    outs[name] = f.read()
        dump = os.path.join(tmp, "dump.txt")
        env = dict(os.environ)
        r = subprocess.run([REF_HARNESS, work, dump], capture_output=True, text=True, env=env)
        if r.returncode != 0:
            raise RuntimeError(f"harness failed on {case_dir}")
        arrays = parse_dump(dump)
    if not with_states:
        arrays = {k: v for k, v in arrays.items()
                  if not any(k.endswith(s) for s in ("_re", "_im"))
                  or "_mu_" in k}
    return outs, arrays


def main():
    if not (os.path.exists(REF_BIN) and os.path.exists(REF_HARNESS)):
        sys.exit("build the reference first: make -C oracle ref")
    os.makedirs(CASES_DIR, exist_ok=True)
    for name, (writer, with_states) in CASES.items():
        d = os.path.join(CASES_DIR, name)
        if os.path.exists(d):
            shutil.rmtree(d)
        os.makedirs(d)
        writer(d)
        outs, arrays = run_reference(d, with_states)
        os.makedirs(os.path.join(d, "expected"))
        for fname, text in outs.items():
            with open(os.path.join(d, "expected", fname), "w") as f:
                f.write(text)
        np.savez_compressed(os.path.join(d, "dump.npz"), **arrays)
        print(f"{name}: {sorted(outs)} {len(arrays)} arrays")

    os.makedirs(FAIL_DIR, exist_ok=True)
    expected = {}
    for name, writer in FAILURES.items():
        d = os.path.join(FAIL_DIR, name)
        if os.path.exists(d):
            shutil.rmtree(d)
        os.makedirs(d)
        writer(d)
        with tempfile.TemporaryDirectory() as tmp:
            work = os.path.join(tmp, "case")
            shutil.copytree(d, work)
            driver = os.path.join(tmp, "input.txt")
            with open(driver, "w") as f:
                f.write(work + "\n")
            r = subprocess.run([REF_BIN, driver], capture_output=True, text=True)
            lines = [ln for ln in r.stdout.replace(work, "<dir>").split("\n") if ln.strip()]
            # the first line that reports the failure
            err = next((ln for ln in lines if ln.startswith(("Error", "Could not", "- Hoppings"))),
                       lines[-1] if lines else "")
            expected[name] = {"returncode": r.returncode, "message": err}
        print(f"failure {name}: {expected[name]}")
    with open(os.path.join(FAIL_DIR, "expected.json"), "w") as f:
        json.dump(expected, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
