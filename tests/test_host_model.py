"""The drop-in's host Model (liblsqt_host.so, C++) against the reference's
bits (CPU): arrays, random vectors, inputs and failure messages."""
import json
import os
import shutil

import numpy as np
import pytest

from conftest import CASES_DIR, GOLDEN, golden_cases
from gpuqt_amd import hostlib
from parity import assert_bits


@pytest.mark.parametrize("case", golden_cases())
def test_model_bitwise(case):
    d = os.path.join(CASES_DIR, case)
    g = np.load(os.path.join(d, "dump.npz"))
    m = hostlib.HostModel(d)
    assert (m.n, m.max_neighbor, m.number_of_moments, m.number_of_random_vectors) == \
        tuple(int(x) for x in g["meta"])
    assert (m.energy_max, m.volume) == tuple(g["fmeta"])
    a = m.arrays()
    vp = m.valid_pairs()
    np.testing.assert_array_equal(a["neighbor_number"], g["neighbor_number"])
    np.testing.assert_array_equal(vp["neighbor_list"], g["neighbor_list"])
    for k in ("hopping_real", "hopping_imag", "xx"):
        assert_bits(vp[k], g[k], k)
    assert_bits(a["potential"], g["potential"], "potential")
    for i in range(m.number_of_random_vectors):
        r = m.random_state()
        if f"v{i}_random_re" in g:
            assert_bits(r[0], g[f"v{i}_random_re"], f"v{i} re")
            assert_bits(r[1], g[f"v{i}_random_im"], f"v{i} im")


def test_model_matches_oracle_inputs():
    import oracle
    d = os.path.join(CASES_DIR, "square_vacancy_ldos")
    m = hostlib.HostModel(d)
    o = oracle.OracleModel(d)
    e, t, orbs = m.inputs()
    assert_bits(e, o.energy, "energy")
    assert_bits(t, o.time_step, "time_step")
    np.testing.assert_array_equal(orbs, o.local_orbitals)
    assert m.calculate_ldos and m.calculate_vac and m.calculate_msd and not m.calculate_spin


def test_failure_messages(tmp_path):
    fdir = os.path.join(GOLDEN, "failures")
    with open(os.path.join(fdir, "expected.json")) as f:
        expected = json.load(f)
    for name, exp in expected.items():
        work = tmp_path / name
        shutil.copytree(os.path.join(fdir, name), work)
        for on_device in (False, True):  # the lattice kept for the GPU as well
            with pytest.raises(hostlib.InputError) as e:
                hostlib.HostModel(str(work), lattice_on_device=on_device)
            msg = str(e.value).replace(str(work), "<dir>")
            assert msg.split("\n")[0] == exp["message"], name


@pytest.mark.parametrize("offset,open_y,fails", [(4, False, False), (5, False, True),
                                                 (5, True, False), (-6, False, True)])
def test_hopping_beyond_one_period(tmp_path, offset, open_y, fails):
    """A hopping reaching further than one period of a periodic direction
    leaves the lattice (an error) -- unless an open direction drops it at
    every site -- whether the lattice is expanded on the host or kept."""
    from gpuqt_amd import inputs as gi
    d = str(tmp_path)
    dy = 9 if open_y else 0  # beyond the 3 cells of the open y direction
    hops = [[(1, 0, 0, 0, -1.0, 0.0), (offset, dy, 0, 0, -0.5, 0.0)]]
    gi.write_lattice(d, (4, 3, 1), (1, 0 if open_y else 1, 0), 0, (1.0, 1.0, 1.0),
                     [(0.0, 0.0, 0.0)], hops)
    gi.write_para(d, moments=8, energy_max=4.0)
    gi.write_energy(d, [0.0])
    for on_device in (False, True):
        if fails:
            with pytest.raises(hostlib.InputError, match="more than one cell"):
                hostlib.HostModel(d, lattice_on_device=on_device)
        else:
            hostlib.HostModel(d, lattice_on_device=on_device)


def test_correlation_row_matches_reference():
    d = os.path.join(CASES_DIR, "honeycomb_small")
    g = np.load(os.path.join(d, "dump.npz"))
    m = hostlib.HostModel(d)
    e, _, _ = m.inputs()
    for tag in ("v0_dos", "v0_vac0", "v1_msd2"):
        mu = g[f"{tag}_mu_re"] + 1j * g[f"{tag}_mu_im"]
        assert_bits(hostlib.correlation_row(mu, e, m.energy_max, m.volume), g[f"{tag}_row"], tag)
