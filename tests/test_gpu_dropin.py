"""The drop-in program (gpuqt_amd/lsqt_gpu: C++ Model / Hamiltonian / Vector
/ sigma drivers over the HIP library) end to end on the GPU.

  - every golden case: the .out files against the reference's text, and,
    with 17 significant digits, every row against the reference's
    full-precision rows (tolerances in parity.py);
  - the reference's own fixed-seed examples (examples/fan2018cpc/lattice,
    4.0M sites): DOS identical at 6 digits, VAC/MSD identical at 6 digits
    wherever |value| > 1e-6 of the row maximum (the checked-in VAC/MSD were
    written by a build with another reduction order; below that floor they
    differ from the reference CPU program too).
"""
import os
import shutil

import numpy as np
import pytest

from conftest import CASES_DIR, GOLDEN, golden_cases
from gpuqt_amd import hostlib
from parity import assert_rows, read_out

pytestmark = pytest.mark.gpu


def _run(tmp_path, src, digits=None, name="case"):
    work = tmp_path / name
    shutil.copytree(src, work, ignore=shutil.ignore_patterns("expected", "dump.npz"))
    driver = tmp_path / f"{name}_input.txt"
    driver.write_text(f"\n{work}\n\n")
    print(f"[gpu] lsqt_gpu {src} digits={digits}", flush=True)
    r = hostlib.run_cli(str(driver), digits=digits, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return work


def _compare_text(got, want, what, floor=1e-6):
    """equal at 6 digits wherever |value| > floor * row max; normwise 1e-5"""
    assert got.shape == want.shape, what
    for r in range(want.shape[0]):
        rmax = np.abs(want[r]).max()
        mask = np.abs(want[r]) > floor * rmax
        bad = np.nonzero(mask & (got[r] != want[r]))[0]
        assert bad.size == 0, f"{what} row {r}: {bad.size} values differ, e.g. " \
                              f"{got[r][bad[0]]!r} vs {want[r][bad[0]]!r}"
        assert np.abs(got[r] - want[r]).max() <= 1e-5 * rmax, what


@pytest.mark.parametrize("case", golden_cases())
def test_out_files_match_reference(tmp_path, case):
    src = os.path.join(CASES_DIR, case)
    work = _run(tmp_path, src)
    for name in sorted(os.listdir(os.path.join(src, "expected"))):
        _compare_text(read_out(work / name), read_out(os.path.join(src, "expected", name)),
                      f"{case}/{name}")


@pytest.mark.parametrize("case", golden_cases())
def test_full_precision_rows(tmp_path, case):
    src = os.path.join(CASES_DIR, case)
    g = np.load(os.path.join(src, "dump.npz"))
    work = _run(tmp_path, src, digits=17)
    nvec = int(g["meta"][3])
    files = {"dos.out": "dos", "vac0.out": "vac0", "vac.out": "vac", "msd.out": "msd",
             "S.out": "spin"}
    for fname, tag in files.items():
        p = work / fname
        if not p.exists():
            continue
        got = read_out(p)
        want = []
        for i in range(nvec):
            if tag in ("dos", "vac0"):
                want.append(g[f"v{i}_{tag}_row"])
            else:
                s = 0
                while f"v{i}_{tag}{s}_row" in g:
                    want.append(g[f"v{i}_{tag}{s}_row"])
                    s += 1
        assert_rows(got, np.array(want), f"{case}/{fname}")
    if (work / "ldos.out").exists():
        got = read_out(work / "ldos.out")
        want = [g[f"ldos{i}_row"] for i in range(got.shape[0])]
        assert_rows(got, np.array(want), f"{case}/ldos.out")


@pytest.mark.parametrize("example", ["diffusive", "localized"])
def test_reference_examples(tmp_path, example):
    src = os.path.join(GOLDEN, "fan2018cpc", example)
    work = _run(tmp_path, src, name=example)
    exp = os.path.join(src, "expected")
    for name in sorted(os.listdir(exp)):
        got, want = read_out(work / name), read_out(os.path.join(exp, name))
        if name == "dos.out":
            assert np.array_equal(got, want), f"{example}/dos.out differs"
        else:
            _compare_text(got, want, f"{example}/{name}")


@pytest.mark.parametrize("path,tol", [("f32", 1e-5), ("block", 2e-5)])
@pytest.mark.parametrize("example", ["diffusive", "localized"])
def test_reference_examples_other_paths(tmp_path, example, path, tol):
    """The reference's own fixed-seed 4.0M-site examples through LSQT_PATH=f32
    and LSQT_PATH=block: every row of dos.out / vac.out / msd.out within `tol`
    (normwise) of the reference's checked-in 6-digit rows."""
    src = os.path.join(GOLDEN, "fan2018cpc", example)
    work = tmp_path / example
    shutil.copytree(src, work, ignore=shutil.ignore_patterns("expected", "dump.npz"))
    driver = tmp_path / "input.txt"
    driver.write_text(f"{work}\n")
    print(f"[gpu] lsqt_gpu LSQT_PATH={path} fan2018cpc/{example}", flush=True)
    r = hostlib.run_cli(str(driver), timeout=600, env=dict(os.environ, LSQT_PATH=path))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    exp = os.path.join(src, "expected")
    for name in sorted(os.listdir(exp)):
        got, want = read_out(work / name), read_out(os.path.join(exp, name))
        assert got.shape == want.shape, f"{example}/{name}"
        for k in range(want.shape[0]):
            err = np.abs(got[k] - want[k]).max() / np.abs(want[k]).max()
            assert err <= tol, f"{example}/{name} row {k}: {err:.2e}"


def test_append_and_ldos(tmp_path):
    src = os.path.join(CASES_DIR, "square_vacancy_ldos")
    work = tmp_path / "twice"
    shutil.copytree(src, work, ignore=shutil.ignore_patterns("expected", "dump.npz"))
    driver = tmp_path / "input.txt"
    driver.write_text(f"{work}\n{work}\n")
    r = hostlib.run_cli(str(driver), timeout=600)
    assert r.returncode == 0
    rows = read_out(work / "ldos.out")
    assert rows.shape[0] == 6 and np.array_equal(rows[:3], rows[3:])


@pytest.mark.parametrize("path,tol", [("f32", 1e-5), ("block", 2e-5)])
@pytest.mark.parametrize("case", golden_cases())
def test_precision_and_block_paths(tmp_path, case, path, tol):
    """LSQT_PATH=f32 (FP32-stored states) and LSQT_PATH=block (all random
    vectors at once) through the unchanged program: the same .out files with
    the same rows, each within `tol` of the reference's row (normwise)."""
    src = os.path.join(CASES_DIR, case)
    work = tmp_path / case
    shutil.copytree(src, work, ignore=shutil.ignore_patterns("expected", "dump.npz"))
    driver = tmp_path / "input.txt"
    driver.write_text(f"{work}\n")
    print(f"[gpu] lsqt_gpu LSQT_PATH={path} {case}", flush=True)
    env = dict(os.environ, LSQT_PATH=path)
    r = hostlib.run_cli(str(driver), digits=17, timeout=600, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    expected = sorted(os.listdir(os.path.join(src, "expected")))
    assert sorted(f for f in os.listdir(work) if f.endswith(".out")) == expected
    for name in expected:
        got = read_out(work / name)
        want = read_out(os.path.join(src, "expected", name))
        assert got.shape == want.shape, name
        for k in range(want.shape[0]):
            err = np.abs(got[k] - want[k]).max() / np.abs(want[k]).max()
            assert err <= tol, f"{case}/{name} row {k}: {err:.2e}"


def test_unknown_path_is_an_input_error(tmp_path):
    src = os.path.join(CASES_DIR, "honeycomb_small")
    work = tmp_path / "w"
    shutil.copytree(src, work, ignore=shutil.ignore_patterns("expected", "dump.npz"))
    driver = tmp_path / "input.txt"
    driver.write_text(f"{work}\n")
    r = hostlib.run_cli(str(driver), timeout=120, env=dict(os.environ, LSQT_PATH="f16"))
    assert r.returncode == 1 and "LSQT_PATH" in r.stdout


def test_batch_path_all_golden_cases(tmp_path):
    """LSQT_PATH=batch with every golden case in one driver file: the
    compatible ones run together, the others (spin, LDOS, other settings)
    one by one; every directory's .out files hold the reference's rows
    within 2e-5 (normwise)."""
    cases = golden_cases()
    dirs = []
    for case in cases:
        work = tmp_path / case
        shutil.copytree(os.path.join(CASES_DIR, case), work,
                        ignore=shutil.ignore_patterns("expected", "dump.npz"))
        dirs.append(work)
    driver = tmp_path / "input.txt"
    driver.write_text("".join(f"{d}\n" for d in dirs))
    print("[gpu] lsqt_gpu LSQT_PATH=batch, all golden cases", flush=True)
    r = hostlib.run_cli(str(driver), digits=17, timeout=900,
                        env=dict(os.environ, LSQT_PATH="batch"))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    for case, work in zip(cases, dirs):
        src = os.path.join(CASES_DIR, case)
        expected = sorted(os.listdir(os.path.join(src, "expected")))
        assert sorted(f for f in os.listdir(work) if f.endswith(".out")) == expected, case
        for name in expected:
            got, want = read_out(work / name), read_out(os.path.join(src, "expected", name))
            assert got.shape == want.shape, f"{case}/{name}"
            for k in range(want.shape[0]):
                err = np.abs(got[k] - want[k]).max() / np.abs(want[k]).max()
                assert err <= 2e-5, f"{case}/{name} row {k}: {err:.2e}"


def test_batch_path_uneven_realisations(tmp_path):
    """Honeycomb realisations of different sizes with the same settings (one
    directory listed twice) through LSQT_PATH=batch against the default FP64
    program on copies of the same directories."""
    from gpuqt_amd import inputs as gi
    dirs = []
    for i, (nx, ny) in enumerate(((9, 5), (14, 7), (6, 12))):
        d = tmp_path / f"r{i}"
        d.mkdir()
        gi.write_honeycomb(str(d), nx, ny)
        gi.write_para(str(d), anderson=1.5, moments=96, energy_max=10.0, vac=True, msd=True,
                      vac0=True, random_vectors=3)
        gi.write_energy(str(d), gi.energy_grid(-8, 8, 41))
        gi.write_time_steps(str(d), [0.5, 1.0, 0.75])
        dirs.append(d)
    listed = dirs + [dirs[1]]
    copies = []
    for d in dirs:
        c = tmp_path / f"{d.name}_f64"
        shutil.copytree(d, c)
        copies.append(c)
    for path, where in (("batch", listed), ("f64", copies + [copies[1]])):
        driver = tmp_path / f"input_{path}.txt"
        driver.write_text("".join(f"{d}\n" for d in where))
        print(f"[gpu] lsqt_gpu LSQT_PATH={path}, 3 realisations", flush=True)
        r = hostlib.run_cli(str(driver), digits=17, timeout=600,
                            env=dict(os.environ, LSQT_PATH=path))
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    for d, c in zip(dirs, copies):
        for name in ("dos.out", "vac0.out", "vac.out", "msd.out"):
            got, want = read_out(d / name), read_out(c / name)
            assert got.shape == want.shape, f"{d.name}/{name}"
            for k in range(want.shape[0]):
                err = np.abs(got[k] - want[k]).max() / np.abs(want[k]).max()
                assert err <= 2e-5, f"{d.name}/{name} row {k}: {err:.2e}"


def test_block_path_more_than_256_vectors(tmp_path):
    """LSQT_PATH=block with 260 random vectors (two blocks: 256 and 4, drawn in
    the reference's order): every row of every .out file within 2e-5 of the
    default FP64 program's."""
    src = os.path.join(CASES_DIR, "honeycomb_small")
    runs = {}
    for path in ("f64", "block"):
        work = tmp_path / path
        shutil.copytree(src, work, ignore=shutil.ignore_patterns("expected", "dump.npz"))
        text = (work / "para.in").read_text()
        (work / "para.in").write_text(text.replace("number_of_random_vectors 2",
                                                   "number_of_random_vectors 260")
                                          .replace("number_of_moments 64", "number_of_moments 24"))
        driver = tmp_path / f"input_{path}.txt"
        driver.write_text(f"{work}\n")
        print(f"[gpu] lsqt_gpu LSQT_PATH={path}, 260 random vectors", flush=True)
        r = hostlib.run_cli(str(driver), digits=17, timeout=600,
                            env=dict(os.environ, LSQT_PATH=path))
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        runs[path] = work
    names = sorted(f for f in os.listdir(runs["f64"]) if f.endswith(".out"))
    assert names == sorted(f for f in os.listdir(runs["block"]) if f.endswith(".out"))
    for name in names:
        want, got = read_out(runs["f64"] / name), read_out(runs["block"] / name)
        assert got.shape == want.shape and want.shape[0] >= 260, name
        for k in range(want.shape[0]):
            err = np.abs(got[k] - want[k]).max() / np.abs(want[k]).max()
            assert err <= 2e-5, f"{name} row {k}: {err:.2e}"


@pytest.mark.parametrize("seed", list(range(16)) + [("general", s) for s in range(12)])
def test_random_inputs_against_oracle(tmp_path, seed):
    """The drop-in program (FP64 path) on the random lattices of
    test_random_inputs.py against the oracle's .out files (which equal the
    reference program's text): equal at 6 digits wherever |value| > 1e-6 of
    the row maximum, normwise within 1e-5."""
    import oracle
    from test_random_inputs import make_case
    base = tmp_path / "case"
    base.mkdir()
    make_case(str(base), seed)
    orc = tmp_path / "orc"
    shutil.copytree(base, orc)
    oracle.lsqt(str(orc))
    print(f"[gpu] lsqt_gpu random lattice {seed}", flush=True)
    work = _run(tmp_path, str(base), name="gpu")
    outs = sorted(f for f in os.listdir(orc) if f.endswith(".out"))
    assert outs == sorted(f for f in os.listdir(work) if f.endswith(".out"))
    for name in outs:
        _compare_text(read_out(work / name), read_out(orc / name), f"seed {seed} {name}")


@pytest.mark.parametrize("path,tol", [("f32", 1e-5), ("block", 2e-5)])
@pytest.mark.parametrize("seed", list(range(16)) + [("general", s) for s in range(12)])
def test_random_inputs_other_paths(tmp_path, seed, path, tol):
    """LSQT_PATH=f32 and LSQT_PATH=block on the random lattices and general
    systems: every row within `tol` of the oracle's, normwise over the row
    (rows whose maximum is below 1e-6 of the file's maximum are compared
    against that floor instead)."""
    import oracle
    from test_random_inputs import make_case
    base = tmp_path / "case"
    base.mkdir()
    make_case(str(base), seed)
    orc = tmp_path / "orc"
    shutil.copytree(base, orc)
    oracle.lsqt(str(orc))
    driver = tmp_path / "input.txt"
    driver.write_text(f"{base}\n")
    print(f"[gpu] lsqt_gpu LSQT_PATH={path} random case {seed}", flush=True)
    r = hostlib.run_cli(str(driver), digits=17, timeout=300, env=dict(os.environ, LSQT_PATH=path))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    outs = sorted(f for f in os.listdir(orc) if f.endswith(".out"))
    assert outs == sorted(f for f in os.listdir(base) if f.endswith(".out"))
    for name in outs:
        got, want = read_out(base / name), read_out(orc / name)
        assert got.shape == want.shape, name
        floor = 1e-6 * np.abs(want).max()
        if floor == 0.0:  # an identically zero quantity (no bond along the transport axis)
            assert np.abs(got).max() == 0.0, f"{seed} {name}"
            continue
        for k in range(want.shape[0]):
            scale = max(np.abs(want[k]).max(), floor)
            err = np.abs(got[k] - want[k]).max() / scale
            assert err <= tol, f"{seed} {name} row {k}: {err:.2e}"


def test_batch_path_random_realisations(tmp_path):
    """LSQT_PATH=batch on random lattices of different shapes, orbitals,
    boundaries and disorder that share their settings (moments, energies,
    time steps, energy_max, quantities, vector count): one segmented
    Hamiltonian of ragged systems; every directory's rows within 2e-5 of the
    oracle's (normwise, with the floor of test_random_inputs_other_paths)."""
    import oracle
    from test_random_inputs import make_case
    from gpuqt_amd import inputs as gi
    dirs, refs = [], []
    for seed in range(6):
        d = tmp_path / f"r{seed}"
        d.mkdir()
        make_case(str(d), seed)
        # shared settings (the disorder of each lattice stays its own)
        text = [ln for ln in (d / "para.in").read_text().splitlines()
                if ln.split()[0] in ("model", "anderson_disorder", "charged_impurity",
                                     "vacancy_disorder")]
        text += ["number_of_random_vectors 3", "number_of_moments 30", "energy_max 14.0",
                 "calculate_vac0", "calculate_vac", "calculate_msd"]
        (d / "para.in").write_text("\n".join(text) + "\n")
        gi.write_energy(str(d), gi.energy_grid(-6, 6, 9))
        gi.write_time_steps(str(d), [0.5, 1.0])
        if (d / "local_orbitals.in").exists():
            (d / "local_orbitals.in").unlink()
        o = tmp_path / f"o{seed}"
        shutil.copytree(d, o)
        oracle.lsqt(str(o))
        dirs.append(d)
        refs.append(o)
    driver = tmp_path / "input.txt"
    driver.write_text("".join(f"{d}\n" for d in dirs))
    print("[gpu] lsqt_gpu LSQT_PATH=batch, 6 random lattices", flush=True)
    r = hostlib.run_cli(str(driver), digits=17, timeout=300, env=dict(os.environ, LSQT_PATH="batch"))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "simulation of 6 systems together" in r.stdout, r.stdout[-2000:]
    for d, o in zip(dirs, refs):
        outs = sorted(f for f in os.listdir(o) if f.endswith(".out"))
        assert outs == sorted(f for f in os.listdir(d) if f.endswith(".out")), d.name
        for name in outs:
            got, want = read_out(d / name), read_out(o / name)
            assert got.shape == want.shape, f"{d.name}/{name}"
            floor = 1e-6 * np.abs(want).max()
            if floor == 0.0:
                assert np.abs(got).max() == 0.0, f"{d.name}/{name}"
                continue
            for k in range(want.shape[0]):
                err = np.abs(got[k] - want[k]).max() / max(np.abs(want[k]).max(), floor)
                assert err <= 2e-5, f"{d.name}/{name} row {k}: {err:.2e}"


def test_host_and_device_lattice_builds_agree(tmp_path):
    """The drop-in expands lattice models on the GPU by default
    (LSQT_DEVICE_LATTICE=0: on the host, as the reference): the .out files of
    lattice cases with vacancies, impurities and open boundaries are the same
    at 17 digits either way."""
    cases = ["square_vacancy_ldos", "charged_vacancy_anderson", "strip_open"]
    outs = {}
    for flag in ("1", "0"):
        work = []
        for c in cases:
            w = tmp_path / f"{c}_{flag}"
            shutil.copytree(os.path.join(CASES_DIR, c), w,
                            ignore=shutil.ignore_patterns("expected", "dump.npz"))
            work.append(w)
        driver = tmp_path / f"input_{flag}.txt"
        driver.write_text("".join(f"{w}\n" for w in work))
        print(f"[gpu] lsqt_gpu LSQT_DEVICE_LATTICE={flag}", flush=True)
        r = hostlib.run_cli(str(driver), digits=17, timeout=600,
                            env=dict(os.environ, LSQT_DEVICE_LATTICE=flag))
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        outs[flag] = {(c, f): (w / f).read_text() for c, w in zip(cases, work)
                      for f in sorted(os.listdir(w)) if f.endswith(".out")}
    assert outs["1"].keys() == outs["0"].keys() and outs["1"]
    for k in outs["1"]:
        assert outs["1"][k] == outs["0"][k], k
