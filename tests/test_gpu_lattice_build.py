"""Device-side construction of lattice Hamiltonians (lsqt_ham_create_lattice)
against Hamiltonians uploaded from the Model's host arrays: every operator,
moment set and evolution equal bit for bit (GPU)."""
import os

import numpy as np
import pytest

from conftest import CASES_DIR

pytestmark = pytest.mark.gpu

PRISTINE_LATTICE_CASES = ["honeycomb_small", "strip_open", "honeycomb_transport_y",
                          "charged_only", "config1_honeycomb_20k"]


def _pair(ctx, case):
    import oracle
    from gpuqt_amd import inputs as gi
    from gpuqt_amd import native
    d = os.path.join(CASES_DIR, case)
    m = oracle.OracleModel(d)
    a = m.arrays()
    H_host = native.Hamiltonian.from_arrays(ctx, a, m.n, m.max_neighbor, m.energy_max)
    H_dev = native.Hamiltonian.from_lattice(ctx, gi.read_lattice(os.path.join(d, "lattice.in")),
                                            a["potential"], m.energy_max)
    return m, H_host, H_dev


def _bits(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), what


@pytest.mark.parametrize("case", PRISTINE_LATTICE_CASES)
def test_device_build_matches_host_arrays(gpu_ctx, case):
    print(f"[gpu] lattice build {case}", flush=True)
    m, Hh, Hd = _pair(gpu_ctx, case)
    assert Hd.n == m.n and Hd.info().max_neighbor == m.max_neighbor
    re, im = m.random_state()
    for name in ("apply", "apply_commutator", "apply_current"):
        outs = []
        for H in (Hh, Hd):
            v, o = H.vector(re, im), H.vector()
            getattr(H, name)(v, o)
            outs.append(o.download())
        _bits(outs[0][0], outs[1][0], name)
        _bits(outs[0][1], outs[1][1], name)
    M = min(m.number_of_moments, 200)
    mus = [H.moments(H.vector(re, im), H.vector(re, im), M) for H in (Hh, Hd)]
    _bits(mus[0].real, mus[1].real, "moments")
    _bits(mus[0].imag, mus[1].imag, "moments")
    for fn in ("evolve", "evolvex"):
        outs = []
        for H in (Hh, Hd):
            v = H.vector(re, im)
            getattr(H, fn)(1, 3.0, v)
            outs.append(v.download())
        _bits(outs[0][0], outs[1][0], fn)
        _bits(outs[0][1], outs[1][1], fn)


def test_device_build_large_honeycomb(gpu_ctx, tmp_path):
    """The 8.0M-site honeycomb of the FP64 benchmark: same moments bit for
    bit; construction times printed."""
    import time
    from gpuqt_amd import hostlib
    from gpuqt_amd import inputs as gi
    from gpuqt_amd import native
    d = str(tmp_path)
    gi.write_honeycomb(d, 2000, 1000)
    gi.write_para(d, anderson=2.0, moments=16, energy_max=10.0)
    gi.write_energy(d, [0.0])
    devnull = os.open(os.devnull, os.O_WRONLY)
    saved = os.dup(1)
    os.dup2(devnull, 1)
    try:
        m = hostlib.HostModel(d)
    finally:
        os.dup2(saved, 1)
        os.close(devnull)
        os.close(saved)
    a = m.arrays()
    t = time.perf_counter()
    Hh = native.Hamiltonian.from_arrays(gpu_ctx, a, m.n, m.max_neighbor, m.energy_max)
    th = time.perf_counter() - t
    lat = gi.read_lattice(os.path.join(d, "lattice.in"))
    t = time.perf_counter()
    Hd = native.Hamiltonian.from_lattice(gpu_ctx, lat, a["potential"], m.energy_max)
    td = time.perf_counter() - t
    print(f"\n  upload of host arrays {th:.2f} s, device build {td:.3f} s", flush=True)
    re, im = m.random_state()
    mus = [H.moments(H.vector(re, im), H.vector(re, im), 16) for H in (Hh, Hd)]
    _bits(mus[0].real, mus[1].real, "moments")


def test_device_build_errors(gpu_ctx):
    from gpuqt_amd import inputs as gi
    from gpuqt_amd import native
    lat = gi.read_lattice(os.path.join(CASES_DIR, "honeycomb_small", "lattice.in"))
    n = int(np.prod(lat["cells"])) * lat["orbitals"]
    bad = dict(lat)
    bad["hop_orbital"] = lat["hop_orbital"].copy()
    bad["hop_orbital"][0, 0] = lat["orbitals"]
    with pytest.raises(native.LsqtError, match="out of range"):
        native.Hamiltonian.from_lattice(gpu_ctx, bad, np.zeros(n), 10.0)
    bad = dict(lat)
    bad["hop_cell"] = lat["hop_cell"].copy()
    d = int(np.argmax(lat["pbc"]))
    assert lat["pbc"][d] == 1
    bad["hop_cell"][0, 0, d] = lat["cells"][d] + 1
    with pytest.raises(native.LsqtError, match="beyond one period"):
        native.Hamiltonian.from_lattice(gpu_ctx, bad, np.zeros(n), 10.0)
