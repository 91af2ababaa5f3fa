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
    # both get the same two-step sweep (slabs of one x cell), so the same
    # partial sums
    assert Hh.info().sweep_slab == Hd.info().sweep_slab == 4000
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


def _quiet(fn):
    devnull = os.open(os.devnull, os.O_WRONLY)
    saved = os.dup(1)
    os.dup2(devnull, 1)
    try:
        return fn()
    finally:
        os.dup2(saved, 1)
        os.close(devnull)
        os.close(saved)


LATTICE_CASES = ["honeycomb_small", "strip_open", "honeycomb_transport_y", "charged_only",
                 "square_vacancy_ldos", "charged_vacancy_anderson", "config1_honeycomb_20k"]


def _same_operators(Hh, Hd, re, im, what):
    for name in ("apply", "apply_commutator", "apply_current"):
        outs = []
        for H in (Hh, Hd):
            v, o = H.vector(re, im), H.vector()
            getattr(H, name)(v, o)
            outs.append(o.download())
            v.close()
            o.close()
        _bits(outs[0][0], outs[1][0], f"{what} {name}")
        _bits(outs[0][1], outs[1][1], f"{what} {name}")


@pytest.mark.parametrize("case", LATTICE_CASES)
def test_lattices_with_vacancies_match_host_arrays(gpu_ctx, case):
    """lsqt_ham_create_lattices on the description the drop-in's Model keeps
    (vacancy re-indexing, disorder from the host) against the Hamiltonian of
    the host-expanded Model's arrays: operators, moments and evolutions bit
    for bit."""
    from gpuqt_amd import hostlib, native
    print(f"[gpu] device lattice with site index: {case}", flush=True)
    d = os.path.join(CASES_DIR, case)
    a = _quiet(lambda: hostlib.HostModel(d))
    b = _quiet(lambda: hostlib.HostModel(d, lattice_on_device=True))
    lat, site_index = b.lattice()
    arr = a.arrays()
    Hh = native.Hamiltonian.from_arrays(gpu_ctx, arr, a.n, a.max_neighbor, a.energy_max)
    Hd = native.Hamiltonian.from_lattices(gpu_ctx, [lat], [site_index], b.potential(),
                                          b.energy_max)
    assert Hd.n == a.n
    re, im = a.random_state()
    _same_operators(Hh, Hd, re, im, case)
    M = min(a.number_of_moments, 128)
    mus = []
    for H in (Hh, Hd):
        v = H.vector(re, im)
        mus.append(H.moments(v, v, M))
        v.close()
    _bits(mus[0].real, mus[1].real, "moments")
    _bits(mus[0].imag, mus[1].imag, "moments")
    for fn in ("evolve", "evolvex"):
        outs = []
        for H in (Hh, Hd):
            v = H.vector(re, im)
            getattr(H, fn)(1, 2.5, v)
            outs.append(v.download())
            v.close()
        _bits(outs[0][0], outs[1][0], fn)
        _bits(outs[0][1], outs[1][1], fn)
    for o in (Hh, Hd):
        o.close()


def test_segmented_lattices_match_host_arrays(gpu_ctx):
    """Four lattice systems (two with vacancies, different slot counts) side
    by side built on the device against lsqt_ham_create_segmented on the
    concatenated host arrays: the operators and the block path's per-segment
    moments bit for bit."""
    from gpuqt_amd import hostlib, native
    print("[gpu] segmented device lattices", flush=True)
    cases = ["honeycomb_small", "square_vacancy_ldos", "charged_vacancy_anderson", "strip_open"]
    host, dev = [], []
    for c in cases:
        d = os.path.join(CASES_DIR, c)
        host.append(_quiet(lambda: hostlib.HostModel(d)))
        dev.append(_quiet(lambda: hostlib.HostModel(d, lattice_on_device=True)))
    sizes = [m.n for m in host]
    starts = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    n = int(starts[-1])
    K = max(m.max_neighbor for m in host)
    nn, nl, pot, hr, hi, xx = [], [], [], [], [], []
    for k, m in enumerate(host):
        a = m.arrays()
        Km = m.max_neighbor
        valid = np.arange(Km)[None, :] < a["neighbor_number"][:, None]
        pad = lambda x, fill=0: np.pad(x.reshape(m.n, Km), ((0, 0), (0, K - Km)),
                                       constant_values=fill)
        nn.append(a["neighbor_number"])
        nl.append(pad(np.where(valid, a["neighbor_list"].reshape(m.n, Km) + starts[k], 0)))
        pot.append(a["potential"])
        hr.append(pad(a["hopping_real"]))
        hi.append(pad(a["hopping_imag"]))
        xx.append(pad(a["xx"]))
    arrays = dict(neighbor_number=np.concatenate(nn).astype(np.int32),
                  neighbor_list=np.concatenate(nl).reshape(-1).astype(np.int32),
                  potential=np.concatenate(pot), hopping_real=np.concatenate(hr).reshape(-1),
                  hopping_imag=np.concatenate(hi).reshape(-1), xx=np.concatenate(xx).reshape(-1))
    emax = host[0].energy_max
    Hh = native.Hamiltonian.from_arrays(gpu_ctx, arrays, n, K, emax, segment_starts=starts)
    lats = [m.lattice() for m in dev]
    Hd = native.Hamiltonian.from_lattices(gpu_ctx, [l for l, _ in lats], [i for _, i in lats],
                                          np.concatenate([m.potential() for m in dev]), emax)
    assert Hd.n == n and Hd.segments == len(cases)
    rng = np.random.default_rng(3)
    ph = rng.uniform(0, 2 * np.pi, (8, n))
    _same_operators(Hh, Hd, np.cos(ph[0]), np.sin(ph[0]), "segmented")
    res = []
    for H in (Hh, Hd):
        b = native.Block(gpu_ctx, n, 8)
        b.upload(np.cos(ph), np.sin(ph))
        H.block_set_config(8, 1, 256)
        res.append(H.moments_block_self(b, 48))
        b.close()
    np.testing.assert_array_equal(res[0], res[1])
    for o in (Hh, Hd):
        o.close()
