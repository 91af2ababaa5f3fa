"""The oracle against the reference's own results (CPU).

Pins every oracle function to the reference compiled from source: the
golden dumps hold the reference's model arrays, random vectors, operator
applications, evolutions, raw moments and summed rows as exact bits, and
its .out files.  Also pins the RNG restatement to the C++ standard library.
"""
import json
import os
import shutil

import numpy as np
import pytest

import oracle
from conftest import CASES_DIR, GOLDEN, golden_cases
from parity import assert_bits, read_out


def _load(case):
    d = os.path.join(CASES_DIR, case)
    return d, np.load(os.path.join(d, "dump.npz")), oracle.OracleModel(d)


@pytest.mark.parametrize("case", golden_cases())
def test_model_arrays_bitwise(case):
    d, g, m = _load(case)
    assert m.n == g["meta"][0] and m.max_neighbor == g["meta"][1]
    assert m.number_of_moments == g["meta"][2]
    assert (m.energy_max, m.volume) == tuple(g["fmeta"])
    vp = m.valid_pairs()
    a = m.arrays()
    np.testing.assert_array_equal(a["neighbor_number"], g["neighbor_number"])
    np.testing.assert_array_equal(vp["neighbor_list"], g["neighbor_list"])
    for k in ("hopping_real", "hopping_imag", "xx"):
        assert_bits(vp[k], g[k], k)
    assert_bits(a["potential"], g["potential"], "potential")


@pytest.mark.parametrize("case", golden_cases(with_states=True))
def test_operators_and_evolution_bitwise(case):
    d, g, m = _load(case)
    r = m.random_state()
    assert_bits(r[0], g["v0_random_re"], "random re")
    assert_bits(r[1], g["v0_random_im"], "random im")
    for name, fn in (("apply", m.apply), ("commutator", m.apply_commutator),
                     ("current", m.apply_current)):
        o = fn(*r)
        assert_bits(o[0], g[f"v0_{name}_re"], name)
        assert_bits(o[1], g[f"v0_{name}_im"], name)
    s1 = m.apply(*r)
    o = m.kernel_polynomial(r, s1)
    assert_bits(o[0], g["v0_kpoly_re"], "kpoly")
    if m.number_of_steps_correlation:
        t = m.time_step[0] * m.energy_max
        for name, fn, direction in (("evolve_p", m.evolve, 1), ("evolve_m", m.evolve, -1),
                                    ("evolvex_p", m.evolvex, 1), ("evolvex_m", m.evolvex, -1)):
            o = fn(direction, t, *r)
            assert_bits(o[0], g[f"v0_{name}_re"], name)
            assert_bits(o[1], g[f"v0_{name}_im"], name)


@pytest.mark.parametrize("case", golden_cases())
def test_drivers_bitwise(case):
    """find_dos / vac0 / vac / msd / spin and LDOS: raw moments and rows."""
    d, g, m = _load(case)
    Nt = m.number_of_steps_correlation
    for i in range(int(g["meta"][3])):
        v = f"v{i}"
        r = m.random_state()
        rows, mu = m.find_dos(r)
        assert_bits(mu[0].real, g[f"{v}_dos_mu_re"], "dos mu")
        assert_bits(mu[0].imag, g[f"{v}_dos_mu_im"], "dos mu im")
        assert_bits(rows[0], g[f"{v}_dos_row"], "dos row")
        if m.calculate_vac0:
            rows, mu = m.find_vac0(r)
            assert_bits(mu[0].real, g[f"{v}_vac0_mu_re"], "vac0 mu")
            assert_bits(rows[0], g[f"{v}_vac0_row"], "vac0 row")
        for name, fn, flag in (("vac", m.find_vac, m.calculate_vac),
                               ("msd", m.find_msd, m.calculate_msd),
                               ("spin", m.find_spin, m.calculate_spin)):
            if not flag:
                continue
            rows, mu = fn(r)
            for s in range(Nt):
                assert_bits(mu[s].real, g[f"{v}_{name}{s}_mu_re"], f"{name}{s} mu re")
                assert_bits(mu[s].imag, g[f"{v}_{name}{s}_mu_im"], f"{name}{s} mu im")
                assert_bits(rows[s], g[f"{v}_{name}{s}_row"], f"{name}{s} row")
    if m.calculate_ldos:
        for i, orb in enumerate(m.local_orbitals):
            rows, _ = m.find_dos(m.random_state(int(orb)))
            assert_bits(rows[0], g[f"ldos{i}_row"], f"ldos{i}")


@pytest.mark.parametrize("case", golden_cases())
def test_out_files_text(case, tmp_path):
    """The oracle's whole-directory driver writes the reference's .out text."""
    d = os.path.join(CASES_DIR, case)
    work = tmp_path / case
    shutil.copytree(d, work, ignore=shutil.ignore_patterns("expected", "dump.npz"))
    oracle.lsqt(str(work))
    expected = sorted(os.listdir(os.path.join(d, "expected")))
    written = sorted(f for f in os.listdir(work) if f.endswith(".out"))
    assert written == expected
    for name in expected:
        with open(os.path.join(d, "expected", name)) as f:
            want = f.read()
        assert (work / name).read_text() == want, name


def test_out_rows_append(tmp_path):
    """Running twice appends identical rows (fixed seed, std::ios::app)."""
    d = os.path.join(CASES_DIR, "strip_open")
    work = tmp_path / "w"
    shutil.copytree(d, work, ignore=shutil.ignore_patterns("expected", "dump.npz"))
    oracle.lsqt(str(work))
    oracle.lsqt(str(work))
    rows = read_out(work / "dos.out")
    assert rows.shape[0] == 2 and np.array_equal(rows[0], rows[1])


@pytest.fixture(scope="module")
def rng_fixture():
    with open(os.path.join(GOLDEN, "rng_draws.json")) as f:
        return json.load(f)


def test_mt19937_standard_value(rng_fixture):
    # [rand.predef]: the 10000th draw of a default-constructed mt19937
    assert oracle.mt_first_draws(5489, 10000)[-1] == 4123659995
    assert rng_fixture["default_10000th"] == 4123659995


def test_mt19937_draws(rng_fixture):
    np.testing.assert_array_equal(oracle.mt_first_draws(12345678, 2000),
                                  np.array(rng_fixture["raw_12345678"], dtype=np.uint64))


def test_uniform_real_draws(rng_fixture):
    want = np.array([float.fromhex(x) for x in rng_fixture["phase_12345678"]])
    got = oracle.uniform_real_draws(12345678, 0.0, 2 * 3.141592653589793, len(want))
    assert_bits(got, want, "phase draws")
    want = np.array([float.fromhex(x) for x in rng_fixture["anderson_777"]])
    got = oracle.uniform_real_draws(777, -0.5 * 8.25, 0.5 * 8.25, len(want))
    assert_bits(got, want, "anderson draws")


def test_uniform_int_draws(rng_fixture):
    want = np.array(rng_fixture["int_12345678"])
    np.testing.assert_array_equal(oracle.uniform_int_draws(12345678, len(want)), want)


def test_failure_messages(tmp_path):
    """Input errors: the oracle reports the reference's message."""
    fdir = os.path.join(GOLDEN, "failures")
    with open(os.path.join(fdir, "expected.json")) as f:
        expected = json.load(f)
    for name, exp in expected.items():
        work = tmp_path / name
        shutil.copytree(os.path.join(fdir, name), work)
        with pytest.raises(oracle.OracleError) as e:
            oracle.OracleModel(str(work))
        msg = str(e.value).replace(str(work), "<dir>")
        assert msg.split("\n")[0] == exp["message"], name


@pytest.mark.slow
def test_reference_example_localized_dos(tmp_path):
    """The oracle reproduces the reference's checked-in fixed-seed DOS of
    examples/fan2018cpc/lattice/localized (4.0M-site chain, W = 4, M = 500,
    two random vectors: pins the RNG order across vectors) at 6 digits."""
    src = os.path.join(GOLDEN, "fan2018cpc", "localized")
    work = tmp_path / "localized"
    shutil.copytree(src, work, ignore=shutil.ignore_patterns("expected"))
    # DOS only: drop the MSD request
    para = (work / "para.in").read_text().replace("calculate_msd", "")
    (work / "para.in").write_text(para)
    oracle.set_threads(min(8, os.cpu_count() or 1))
    oracle.lsqt(str(work))
    got = (work / "dos.out").read_text()
    want = open(os.path.join(src, "expected", "dos.out")).read()
    assert got == want
