"""Slot patterns (GPU): on a periodic lattice the moment kernels compute the
packed words of regular sites from a per-orbital pattern and load only the
words of irregular sites (boundaries, vacancies' neighbours).  The words are
the same either way, so every result must be bit-identical to the same
Hamiltonian without the pattern (LSQT_PATTERN=0), and the operators equal to
the oracle's bits."""
import numpy as np
import pytest

from parity import assert_bits, assert_moments

pytestmark = pytest.mark.gpu


def _honeycomb(d, nx, ny, vacancies):
    from gpuqt_amd import inputs as gi
    gi.write_honeycomb(d, nx, ny)
    gi.write_para(d, anderson=1.0, vacancies=vacancies, moments=64, energy_max=10.0)
    gi.write_energy(d, [0.0])
    gi.write_time_steps(d, [1.0])


def _compare(gpu_ctx, monkeypatch, arrays, n, K, emax, vectors, expect_pattern):
    """Builds the Hamiltonian with and without the slot pattern; every state
    and moment set bit-identical.  Returns the Hamiltonian with the pattern
    (and its info) for further checks."""
    from gpuqt_amd import native
    H = native.Hamiltonian.from_arrays(gpu_ctx, arrays, n, K, emax)
    info = H.info()
    if expect_pattern:
        assert info.pattern_period > 0, "no slot pattern detected"
        assert 0.0 < info.pattern_irregular <= 0.1
        assert info.bytes_per_site_f64 < 8 + 64 + 4 * K
    monkeypatch.setenv("LSQT_PATTERN", "0")
    H0 = native.Hamiltonian.from_arrays(gpu_ctx, arrays, n, K, emax)
    monkeypatch.delenv("LSQT_PATTERN")
    assert H0.info().pattern_period == 0
    v, l = H.vector(*vectors[0]), H.vector(*vectors[1])
    out, out0 = H.vector(), H0.vector()
    H.apply(v, out)
    H0.apply(v, out0)
    a, b = out.download(), out0.download()
    assert_bits(a[0], b[0], "apply re, pattern on / off")
    assert_bits(a[1], b[1], "apply im, pattern on / off")
    M = 64
    np.testing.assert_array_equal(H.moments(v, v, M), H0.moments(v, v, M))
    np.testing.assert_array_equal(H.moments(l, v, M), H0.moments(l, v, M))
    np.testing.assert_array_equal(H.moments_f32(v, v, M), H0.moments_f32(v, v, M))
    np.testing.assert_array_equal(H.moments_f32(l, v, M), H0.moments_f32(l, v, M))
    for o in (v, l, out, out0, H0):
        o.close()
    return H, info


def test_pattern_honeycomb(gpu_ctx, tmp_path, monkeypatch):
    """Periodic honeycomb (irregular sites at the x and y wraps only): the
    pattern is found, results equal the oracle's and the pattern-off ones."""
    import oracle
    print("[gpu] slot pattern, honeycomb 100x60", flush=True)
    d = str(tmp_path)
    _honeycomb(d, 100, 60, None)
    m = oracle.OracleModel(d)
    r = m.random_state()
    H, info = _compare(gpu_ctx, monkeypatch, m.arrays(), m.n, m.max_neighbor, m.energy_max,
                       [r, m.random_state()], True)
    assert info.pattern_period == 4  # orbitals of the cell, the fastest index
    v, out = H.vector(*r), H.vector()
    H.apply(v, out)
    got, want = out.download(), m.apply(*r)
    assert_bits(got[0], want[0], "apply re")
    assert_bits(got[1], want[1], "apply im")
    ref = m.moments(r, r, 64)
    assert_moments(H.moments(v, v, 64), ref[0] + 1j * ref[1], "moments")
    for o in (v, out, H):
        o.close()


def test_pattern_interior_irregular_sites(gpu_ctx, tmp_path, monkeypatch):
    """The same honeycomb with 50 sites whose first hopping takes another
    value (another class) and 20 sites with their last slot dropped: irregular
    sites inside the system, loaded instead of computed."""
    import oracle
    print("[gpu] slot pattern, honeycomb 100x60 with irregular interior sites", flush=True)
    d = str(tmp_path)
    _honeycomb(d, 100, 60, None)
    m = oracle.OracleModel(d)
    arrays = {k: np.array(v, copy=True) for k, v in m.arrays().items()}
    K = m.max_neighbor
    rng = np.random.default_rng(5)
    sites = rng.choice(m.n, 70, replace=False)
    for s in sites[:50]:
        arrays["hopping_real"][s * K] = -2.5
    for s in sites[50:]:
        k = int(arrays["neighbor_number"][s]) - 1
        arrays["neighbor_number"][s] = k
        arrays["neighbor_list"][s * K + k] = 0
        arrays["hopping_real"][s * K + k] = 0.0
        arrays["hopping_imag"][s * K + k] = 0.0
        arrays["xx"][s * K + k] = 0.0
    H, info = _compare(gpu_ctx, monkeypatch, arrays, m.n, K, m.energy_max,
                       [m.random_state(), m.random_state()], True)
    H.close()


def test_pattern_odd_site_count(gpu_ctx, tmp_path, monkeypatch):
    """A 67 x 63 square lattice (4221 sites, one orbital: period 1): the
    two-site FP32 step's last pair ends in a padding site, which the mask
    marks irregular so that its (empty) words are loaded."""
    import oracle
    from gpuqt_amd import inputs as gi
    print("[gpu] slot pattern, square 67x63 (odd site count)", flush=True)
    d = str(tmp_path)
    gi.write_square(d, 67, 63)
    gi.write_para(d, anderson=1.0, moments=64, energy_max=6.0)
    gi.write_energy(d, [0.0])
    gi.write_time_steps(d, [1.0])
    m = oracle.OracleModel(d)
    assert m.n == 67 * 63 and m.n % 2 == 1
    r = m.random_state()
    H, info = _compare(gpu_ctx, monkeypatch, m.arrays(), m.n, m.max_neighbor, m.energy_max,
                       [r, m.random_state()], True)
    assert info.pattern_period == 1
    v = H.vector(*r)
    ref = m.moments(r, r, 64)
    assert_moments(H.moments(v, v, 64), ref[0] + 1j * ref[1], "moments")
    for o in (v, H):
        o.close()


def test_pattern_vacancies(gpu_ctx, tmp_path, monkeypatch):
    """Vacancy disorder removes sites and renumbers the rest, which shifts
    the orbital residues; whatever the detection decides, results equal the
    pattern-off ones."""
    import oracle
    print("[gpu] slot pattern, honeycomb 100x60 with 40 vacancies", flush=True)
    d = str(tmp_path)
    _honeycomb(d, 100, 60, 40)
    m = oracle.OracleModel(d)
    H, info = _compare(gpu_ctx, monkeypatch, m.arrays(), m.n, m.max_neighbor, m.energy_max,
                       [m.random_state(), m.random_state()], False)
    H.close()


def test_no_pattern_without_periodicity(gpu_ctx):
    """A random graph has no pattern; the kernels then load every word."""
    from gpuqt_amd import native
    n, K = 8192, 4
    rng = np.random.default_rng(3)
    nl = rng.integers(0, n, (n, K)).astype(np.int32)
    arrays = dict(neighbor_number=np.full(n, K, np.int32), neighbor_list=nl.reshape(-1),
                  potential=rng.uniform(-1, 1, n), hopping_real=np.full(n * K, -1.0),
                  hopping_imag=np.zeros(n * K), xx=np.zeros(n * K))
    H = native.Hamiltonian.from_arrays(gpu_ctx, arrays, n, K, 10.0)
    assert H.info().pattern_period == 0
    H.close()


def _oracle_checks(H, arrays, n, K, emax, vectors):
    """apply bit for bit and moments against the oracle on the same arrays."""
    import oracle
    oh = oracle.OracleHam(arrays, n, K, emax)
    r = vectors[0]
    v, out = H.vector(*r), H.vector()
    H.apply(v, out)
    got, want = out.download(), oh.apply(*r)
    assert_bits(got[0], want[0], "apply re")
    assert_bits(got[1], want[1], "apply im")
    ref = oh.moments(r, r, 64)
    assert_moments(H.moments(v, v, 64), ref[0] + 1j * ref[1], "moments")
    for o in (v, out):
        o.close()


def test_pattern_with_global_class_table(gpu_ctx, tmp_path, monkeypatch):
    """A periodic honeycomb whose hoppings take 600 extra distinct values on
    200 sites (under 1 % of them): more than 512 classes, so the class table
    stays in global memory, and the slot pattern is still kept.  At the
    wrapped boundary sites the pattern's computed words carry the neighbour
    index into the class bits before the mask is read; the class index must
    stay inside the table.  Results equal the pattern-off ones and the
    oracle's."""
    import oracle
    print("[gpu] slot pattern, honeycomb 100x60 with a global class table", flush=True)
    d = str(tmp_path)
    _honeycomb(d, 100, 60, None)
    m = oracle.OracleModel(d)
    arrays = {k: np.array(v, copy=True) for k, v in m.arrays().items()}
    K = m.max_neighbor
    rng = np.random.default_rng(11)
    sites = rng.choice(m.n, 200, replace=False)
    for s in sites:
        for q in range(int(arrays["neighbor_number"][s])):
            arrays["hopping_real"][s * K + q] = -2.7 + rng.uniform(-0.5, 0.5)
    H, info = _compare(gpu_ctx, monkeypatch, arrays, m.n, K, m.energy_max,
                       [m.random_state(), m.random_state()], True)
    assert info.layout == 2 and info.classes > 512, (info.layout, info.classes)
    # the index bits leave room for class indices far beyond the table
    assert (1 << (32 - info.index_bits)) > 64 * info.classes
    _oracle_checks(H, arrays, m.n, K, m.energy_max, [m.random_state()])
    H.close()


def test_pattern_eight_slots(gpu_ctx, tmp_path, monkeypatch):
    """A square lattice with nearest and next-nearest neighbours (8 slots,
    outside the slot counts with their own kernels): the generic slot loop
    with the pattern, bit-identical to the pattern-off build and to the
    oracle."""
    import oracle
    from gpuqt_amd import inputs as gi
    print("[gpu] slot pattern, square 70x64 with 8 neighbours", flush=True)
    d = str(tmp_path)
    hops = [[(-1, 0, 0, 0, -1.0, 0.0), (1, 0, 0, 0, -1.0, 0.0), (0, 1, 0, 0, -1.0, 0.0),
             (0, -1, 0, 0, -1.0, 0.0), (1, 1, 0, 0, 0.25, 0.1), (-1, -1, 0, 0, 0.25, -0.1),
             (1, -1, 0, 0, 0.3, 0.0), (-1, 1, 0, 0, 0.3, 0.0)]]
    gi.write_lattice(d, (70, 64, 1), (1, 1, 0), 0, (1.0, 1.0, 1.0), [(0.0, 0.0, 0.0)], hops)
    gi.write_para(d, anderson=1.0, moments=64, energy_max=8.0)
    gi.write_energy(d, [0.0])
    gi.write_time_steps(d, [1.0])
    m = oracle.OracleModel(d)
    assert m.max_neighbor == 8
    H, info = _compare(gpu_ctx, monkeypatch, m.arrays(), m.n, m.max_neighbor, m.energy_max,
                       [m.random_state(), m.random_state()], True)
    assert info.pattern_period == 1
    _oracle_checks(H, m.arrays(), m.n, m.max_neighbor, m.energy_max, [m.random_state()])
    H.close()
