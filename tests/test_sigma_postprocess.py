"""Conductivity post-processing (host C++, lsqt_host_conductivity and the
--sigma mode of the program) against closed forms, the numpy restatement of
the reference's example scripts and the reference's own fixed-seed outputs
of examples/fan2018cpc/lattice/diffusive (CPU)."""
import os
import shutil

import numpy as np
import pytest

from conftest import GOLDEN
from oracle import sigma as osig
from gpuqt_amd import hostlib

DIFFUSIVE = os.path.join(GOLDEN, "fan2018cpc", "diffusive")


def _diffusive():
    ts = np.loadtxt(os.path.join(DIFFUSIVE, "time_step.in"))[1:]
    dos = np.loadtxt(os.path.join(DIFFUSIVE, "expected", "dos.out"), ndmin=2)
    vac = np.loadtxt(os.path.join(DIFFUSIVE, "expected", "vac.out"))
    msd = np.loadtxt(os.path.join(DIFFUSIVE, "expected", "msd.out"))
    return ts, dos, vac, msd


def test_closed_forms_uneven_steps():
    """Trapezoids are exact for a linear VAC; differences of a linear MSD
    give a constant and of a quadratic MSD give pi a (t_k + t_{k-1})."""
    ts = np.array([0.5, 1.0, 0.25, 2.0, 1.5])
    t = np.concatenate([[0.0], np.cumsum(ts[:-1])])   # VAC sample times
    T = np.cumsum(ts)                                 # MSD sample times
    a, b = 0.7, -0.3
    vac = np.stack([a + b * t, np.full_like(t, 2.0)], axis=1)
    msd = np.stack([a * T, a * T ** 2], axis=1)
    r = hostlib.conductivity(ts, vac=vac, msd=msd)
    np.testing.assert_allclose(r["sigma_vac"][:, 0], 2 * np.pi * (a * t + b * t ** 2 / 2),
                               rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(r["sigma_vac"][:, 1], 2 * np.pi * 2.0 * t, rtol=1e-14)
    np.testing.assert_allclose(r["sigma_msd"][:, 0], np.pi * a, rtol=1e-14)
    Tprev = np.concatenate([[0.0], T[:-1]])
    np.testing.assert_allclose(r["sigma_msd"][:, 1], np.pi * a * (T + Tprev), rtol=1e-14)


def test_vector_average():
    """Blocks of Nt rows are averaged; the DOS rows are averaged; the true
    VAC / MSD divide by the averaged DOS."""
    rng = np.random.default_rng(5)
    Nt, Ne, Ns = 4, 3, 3
    ts = np.full(Nt, 0.5)
    dos = rng.uniform(0.5, 1.0, (Ns, Ne))
    vac = rng.normal(size=(Ns * Nt, Ne))
    msd = rng.uniform(size=(Ns * Nt, Ne))
    r = hostlib.conductivity(ts, dos=dos, vac=vac, msd=msd)
    one = hostlib.conductivity(ts, dos=dos.mean(0), vac=vac.reshape(Ns, Nt, Ne).mean(0),
                               msd=msd.reshape(Ns, Nt, Ne).mean(0))
    for k in ("sigma_vac", "sigma_msd", "vac", "msd"):
        np.testing.assert_allclose(r[k], one[k], rtol=1e-13, atol=1e-15, err_msg=k)
    np.testing.assert_allclose(r["vac"], vac.reshape(Ns, Nt, Ne).mean(0) / dos.mean(0),
                               rtol=1e-13)


def test_against_restatement_on_reference_outputs():
    ts, dos, vac, msd = _diffusive()
    got = hostlib.conductivity(ts, dos=dos, vac=vac, msd=msd)
    want = osig.conductivity(ts, dos=dos, vac=vac, msd=msd)
    for k in ("sigma_vac", "sigma_msd", "vac", "msd"):
        np.testing.assert_allclose(got[k], want[k], rtol=1e-12,
                                   atol=1e-12 * np.abs(want[k]).max(), err_msg=k)


def test_vac_and_msd_estimators_agree_diffusive():
    """Physics: in the diffusive example both estimators reach the same
    semiclassical conductivity (maximum over time) inside the band."""
    ts, dos, vac, msd = _diffusive()
    r = hostlib.conductivity(ts, dos=dos, vac=vac, msd=msd)
    e = np.loadtxt(os.path.join(DIFFUSIVE, "energy.in"))[1:]
    band = np.abs(e) < 3.0
    sv = r["sigma_vac"].max(axis=0)[band]
    sm = r["sigma_msd"].max(axis=0)[band]
    assert sv.min() > 1.0
    assert np.abs(sv - sm).max() / sv.max() < 0.06


def test_postprocess_files_and_cli(tmp_path):
    ts, dos, vac, msd = _diffusive()
    work = tmp_path / "diffusive"
    shutil.copytree(DIFFUSIVE, work, ignore=shutil.ignore_patterns("expected"))
    for name in ("dos.out", "vac.out", "msd.out"):
        shutil.copy(os.path.join(DIFFUSIVE, "expected", name), work / name)
    want = osig.conductivity(ts, dos=dos, vac=vac, msd=msd)
    hostlib.postprocess(str(work))
    for name, key in (("sigma_vac.out", "sigma_vac"), ("sigma_msd.out", "sigma_msd")):
        got = np.loadtxt(work / name)
        np.testing.assert_allclose(got, want[key], rtol=5e-6, atol=1e-6, err_msg=name)
        (work / name).unlink()
    driver = tmp_path / "input.txt"
    driver.write_text(str(work) + "\n\n")
    p = hostlib.run_cli(str(driver), sigma=True, timeout=60)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "sigma_vac.out" in p.stdout and "sigma_msd.out" in p.stdout
    np.testing.assert_allclose(np.loadtxt(work / "sigma_msd.out"), want["sigma_msd"],
                               rtol=5e-6, atol=1e-6)


def test_errors(tmp_path):
    ts = np.full(4, 1.0)
    with pytest.raises(hostlib.InputError, match="whole blocks"):
        hostlib.conductivity(ts, vac=np.ones((6, 2)))
    with pytest.raises(hostlib.InputError):
        hostlib.postprocess(str(tmp_path))  # no energy.in
    (tmp_path / "energy.in").write_text("2\n0.0\n1.0\n")
    (tmp_path / "time_step.in").write_text("2\n1.0\n1.0\n")
    with pytest.raises(hostlib.InputError, match="vac.out"):
        hostlib.postprocess(str(tmp_path))
    (tmp_path / "msd.out").write_text("1 2 \n1 2 3 \n")
    with pytest.raises(hostlib.InputError, match="has 3 values"):
        hostlib.postprocess(str(tmp_path))
