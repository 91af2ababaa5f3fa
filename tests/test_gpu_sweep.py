"""Several Chebyshev steps per launch on slab lattices (GPU).

Where every neighbour of a site lies within one slab and a few positions of
it (periodically), the moment sets advance two (k_sweep2) or four (k_sweepn)
recursion steps per launch from LDS (six for FP32 states).  Checked here:

  - which Hamiltonians get the sweep: periodic honeycomb (whole-slab and
    chunked slabs), square and triangular lattices, a ladder whose orbitals
    leave a slot unused in a periodic pattern (the exact sum over the used
    slots; the other lattices use every slot); none for a Peierls-phase
    square lattice (its hoppings change from slab to slab), the 3D cube (its
    in-slab neighbours are a row of 32 positions away), vacancies
    (renumbered sites), open boundaries (empty slots at the edges) or
    LSQT_SWEEP=0;
  - lsqt_kernel_polynomial_pair against two lsqt_kernel_polynomial calls,
    bit for bit (the per-site arithmetic is the same, in the same order);
  - FP64 and FP32-state moment sets (six steps per launch from moment 4,
    two steps before and for the remainder) against the same Hamiltonian without the
    sweep and with two steps per launch only (the partial sums group the
    sites differently: the FP64 moment tolerance) and against the oracle,
    for moment counts 2 .. 5, 10 .. 15 (every remainder of the six-step
    launches) and odd / even large ones; FP32-state moment sets against
    their one-step counterparts within 1e-6 of mu_0;
  - U(t) with six Bessel terms per launch against one term per launch, bit
    for bit in FP64.
"""
import numpy as np
import pytest

from parity import assert_bits, assert_moments

pytestmark = pytest.mark.gpu


def _lattice(nx, ny, offsets, rng, phase=None, disorder=1.0):
    """A periodic 2D lattice, site x * ny + y, neighbours at the cell
    offsets (dx, dy); hopping -1, or exp(i phase(x)) on dy != 0 bonds."""
    n = nx * ny
    idx = np.arange(n)
    x, y = idx // ny, idx % ny
    K = len(offsets)
    nl = np.empty((n, K), np.int32)
    hr = np.full((n, K), -1.0)
    hi = np.zeros((n, K))
    xx = np.zeros((n, K))
    for q, (dx, dy) in enumerate(offsets):
        nl[:, q] = ((x + dx) % nx) * ny + (y + dy) % ny
        xx[:, q] = dx
        if phase is not None and dy != 0:
            p = phase(x) * dy
            hr[:, q] = -np.cos(p)
            hi[:, q] = -np.sin(p)
    arrays = dict(neighbor_number=np.full(n, K, np.int32), neighbor_list=nl.reshape(-1),
                  potential=rng.uniform(-disorder, disorder, n), hopping_real=hr.reshape(-1),
                  hopping_imag=hi.reshape(-1), xx=xx.reshape(-1))
    return arrays, n, K


def _ladder(nx, ny, rng):
    """Rows y = 0, 1 and 2, 3 (mod 4) bonded in pairs, row 3 to the next
    row 0, every row to its x neighbours: rows 0 and 3 have four bonds, rows 1
    and 2 three, so the fourth slot is unused on half the orbitals (a
    periodic pattern of period 4 with empty slots)."""
    assert ny % 4 == 0
    n = nx * ny
    idx = np.arange(n)
    x, y = idx // ny, idx % ny
    site = lambda xx, yy: (xx % nx) * ny + yy % ny
    up = {0: 1, 1: -1, 2: 1, 3: -1}      # the pair bond, slot 2
    nl = np.empty((n, 4), np.int32)
    nl[:, 0] = site(x - 1, y)
    nl[:, 1] = site(x + 1, y)
    nl[:, 2] = site(x, y + np.vectorize(up.get)(y % 4))
    link = np.where(y % 4 == 0, -1, np.where(y % 4 == 3, 1, 0))  # rows 0 and 3, slot 3
    nl[:, 3] = np.where(link != 0, site(x, y + link), idx)
    nn = np.where(link != 0, 4, 3).astype(np.int32)
    hr = np.full((n, 4), -1.0)
    xx = np.zeros((n, 4))
    xx[:, 0], xx[:, 1] = -1, 1
    arrays = dict(neighbor_number=nn, neighbor_list=nl.reshape(-1),
                  potential=rng.uniform(-1, 1, n), hopping_real=hr.reshape(-1),
                  hopping_imag=np.zeros(n * 4), xx=xx.reshape(-1))
    return arrays, n, 4


CHAINS = [(-1, 0), (1, 0)]
SQUARE = [(-1, 0), (1, 0), (0, -1), (0, 1)]
TRIANGULAR = [(-1, 0), (1, 0), (0, -1), (0, 1), (1, -1), (-1, 1)]


def _honeycomb_model(tmp_path, name, nx, ny, vacancies=None):
    import oracle
    from gpuqt_amd import inputs as gi
    d = tmp_path / name
    d.mkdir()
    gi.write_honeycomb(str(d), nx, ny)
    gi.write_para(str(d), anderson=1.0, vacancies=vacancies, moments=64, energy_max=10.0)
    gi.write_energy(str(d), [0.0])
    gi.write_time_steps(str(d), [1.0])
    return oracle.OracleModel(str(d))


def _systems(tmp_path):
    rng = np.random.default_rng(11)
    out = []
    m = _honeycomb_model(tmp_path, "hc_whole", 100, 60)      # slab 240 sites: one chunk
    out.append(("honeycomb 100x60", m.arrays(), m.n, 3, m.energy_max, 240, 0, m))
    m = _honeycomb_model(tmp_path, "hc_chunks", 24, 250)     # slab 1000: 4 chunks, halo 1
    out.append(("honeycomb 24x250", m.arrays(), m.n, 3, m.energy_max, 1000, 1, m))
    a, n, K = _lattice(130, 300, SQUARE, rng)                 # slab 300: 2 chunks
    out.append(("square 130x300", a, n, K, 5.0, 300, 1, None))
    a, n, K = _lattice(70, 777, TRIANGULAR, rng)              # 777: 4 chunks, uneven last
    out.append(("triangular 70x777", a, n, K, 7.0, 777, 1, None))
    a, n, K = _lattice(160, 256, CHAINS, rng)                 # 2 slots, no in-slab bonds
    out.append(("chains 160x256", a, n, K, 4.0, 256, 0, None))
    a, n, K = _ladder(120, 400, rng)                          # empty slots in the pattern
    out.append(("ladder 120x400, three and four bonds", a, n, K, 5.0, 400, 1, None))
    a, n, K = _lattice(90, 400, SQUARE, rng,                  # complex, the same in every slab
                       phase=lambda x: 0.7 + 0 * x)
    out.append(("square 90x400 with a uniform complex phase", a, n, K, 5.0, 400, 1, None))
    a, n, K = _lattice(640, 300, SQUARE, rng,                 # phases change per slab
                       phase=lambda x: 2 * np.pi * 0.001 * x)
    out.append(("peierls square 640x300", a, n, K, 5.0, 0, 0, None))
    return out


def test_sweep_geometry(gpu_ctx, tmp_path, monkeypatch):
    from gpuqt_amd import native
    for name, arrays, n, K, emax, slab, halo, _ in _systems(tmp_path):
        print(f"[gpu] sweep geometry: {name}", flush=True)
        H = native.Hamiltonian.from_arrays(gpu_ctx, arrays, n, K, emax)
        info = H.info()
        if not slab:
            assert info.sweep_slab == 0, name
        else:
            assert (info.sweep_slab, info.sweep_halo) == (slab, halo), name
            assert info.sweep_tiles > 0 and info.sweep_tiles_f32 > 0
            assert info.sweep_steps == 6 and info.sweep_tiles_deep > 0
            assert 0 < info.bytes_per_site_sweep_f64 < 2 * info.bytes_per_site_f64
        H.close()
    # the cube: y-neighbours are a row (32 positions) away within a slab
    from gpuqt_amd import inputs as gi
    arrays, n = gi.cubic_peierls_arrays(32)
    H = native.Hamiltonian.from_arrays(gpu_ctx, arrays, n, 6, 15.0)
    assert H.info().sweep_slab == 0
    vs = [H.vector() for _ in range(4)]
    with pytest.raises(native.LsqtError):  # no two-step launch without a sweep
        H.kernel_polynomial_pair(*vs)
    for v in vs:
        v.close()
    H.close()
    # vacancies renumber the sites
    m = _honeycomb_model(tmp_path, "hc_vac", 100, 60, vacancies=40)
    H = native.Hamiltonian.from_arrays(gpu_ctx, m.arrays(), m.n, 3, m.energy_max)
    assert H.info().sweep_slab == 0
    H.close()
    # open boundaries in y: the edge sites have empty slots
    rng = np.random.default_rng(4)
    a, n, K = _lattice(100, 300, SQUARE, rng)
    nl = a["neighbor_list"].reshape(n, K)
    nn = np.full(n, K, np.int32)
    y = np.arange(n) % 300
    for edge, slot in ((0, 2), (299, 3)):
        rows = np.nonzero(y == edge)[0]
        nl[rows, slot] = nl[rows, 3 if slot == 2 else 2]  # keep the other y bond in slot 2
        nn[rows] = 3
    a["neighbor_list"] = nl.reshape(-1)
    a["neighbor_number"] = nn
    H = native.Hamiltonian.from_arrays(gpu_ctx, a, n, K, 5.0)
    assert H.info().sweep_slab == 0
    H.close()
    monkeypatch.setenv("LSQT_SWEEP", "0")
    m = _honeycomb_model(tmp_path, "hc_off", 100, 60)
    H = native.Hamiltonian.from_arrays(gpu_ctx, m.arrays(), m.n, 3, m.energy_max)
    assert H.info().sweep_slab == 0
    H.close()


def _random_vector(H, rng):
    ph = rng.uniform(0, 2 * np.pi, H.n)
    return H.vector(np.cos(ph) * rng.uniform(0.5, 1.5, H.n), np.sin(ph))


def test_pair_bitwise(gpu_ctx, tmp_path):
    from gpuqt_amd import native
    rng = np.random.default_rng(3)
    systems = [s for s in _systems(tmp_path) if s[5]]
    for name, arrays, n, K, emax, _, _, _ in systems:
        print(f"[gpu] kernel_polynomial_pair: {name}", flush=True)
        H = native.Hamiltonian.from_arrays(gpu_ctx, arrays, n, K, emax)
        info = H.info()
        assert info.sweep_slab > 0, name
        s0, s1 = _random_vector(H, rng), _random_vector(H, rng)
        s2, s3 = H.vector(), H.vector()
        H.kernel_polynomial_pair(s0, s1, s2, s3)
        r2, r3 = H.vector(), H.vector()
        H.kernel_polynomial(s0, s1, r2)
        H.kernel_polynomial(s1, r2, r3)
        for got, want, what in ((s2, r2, "s2"), (s3, r3, "s3")):
            g, w = got.download(), want.download()
            assert_bits(g[0], w[0], f"{name}: {what} re")
            assert_bits(g[1], w[1], f"{name}: {what} im")
        with pytest.raises(native.LsqtError):
            H.kernel_polynomial_pair(s0, s1, s2, s1)
        for o in (s0, s1, s2, s3, r2, r3, H):
            o.close()


def test_sweep_moments(gpu_ctx, tmp_path, monkeypatch):
    from gpuqt_amd import native
    rng = np.random.default_rng(9)
    for name, arrays, n, K, emax, slab, _, om in _systems(tmp_path):
        if not slab:
            continue
        print(f"[gpu] sweep moments: {name}", flush=True)
        H = native.Hamiltonian.from_arrays(gpu_ctx, arrays, n, K, emax)
        monkeypatch.setenv("LSQT_SWEEP", "0")
        H0 = native.Hamiltonian.from_arrays(gpu_ctx, arrays, n, K, emax)
        monkeypatch.delenv("LSQT_SWEEP")
        assert H.info().sweep_slab > 0 and H0.info().sweep_slab == 0
        monkeypatch.setenv("LSQT_SWEEP_STEPS", "2")
        H2 = native.Hamiltonian.from_arrays(gpu_ctx, arrays, n, K, emax)
        monkeypatch.delenv("LSQT_SWEEP_STEPS")
        assert (H.info().sweep_steps, H.info().sweep_steps_f32) == (6, 6), name
        assert (H2.info().sweep_steps, H2.info().sweep_steps_f32) == (2, 2), name
        assert H0.info().sweep_steps == 0
        v, l = _random_vector(H, rng), _random_vector(H, rng)
        # 2 .. 5: one and two steps per launch only; 10 .. 15: one six-step
        # launch, then every remainder; 16, 17: two; 64, 129: many
        for M in (2, 3, 4, 5, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 64, 129):
            ref_self, ref_pair = H0.moments(v, v, M), H0.moments(l, v, M)
            assert_moments(H.moments(v, v, M), ref_self, f"{name} M={M} self")
            assert_moments(H.moments(l, v, M), ref_pair, f"{name} M={M} pair")
            assert_moments(H2.moments(l, v, M), ref_pair, f"{name} M={M} pair, two steps")
            a, b = H.moments_f32(l, v, M), H0.moments_f32(l, v, M)
            assert np.abs(a - b).max() <= 1e-6 * n, f"{name} M={M} f32"
            c = H2.moments_f32(l, v, M)
            assert np.abs(c - b).max() <= 1e-6 * n, f"{name} M={M} f32, two steps"
        if om is not None:
            r = om.random_state()
            ov = H.vector(*r)
            ref_re, ref_im = om.moments(r, r, 64)
            assert_moments(H.moments(ov, ov, 64), ref_re + 1j * ref_im, f"{name} oracle")
            ov.close()
        for o in (v, l, H, H0, H2):
            o.close()


def test_sweep_evolution(gpu_ctx, tmp_path, monkeypatch):
    """U(t) with six Bessel terms per launch (FP64 and FP32 states) against
    one term per launch (LSQT_SWEEP_STEPS=2 leaves the evolution one term
    per launch): FP64 states bit for bit, FP32-state evolution within FP32
    rounding; series lengths with every remainder."""
    from gpuqt_amd import native
    rng = np.random.default_rng(21)
    for name, arrays, n, K, emax, slab, _, _ in _systems(tmp_path):
        if not slab:
            continue
        print(f"[gpu] sweep evolution: {name}", flush=True)
        H = native.Hamiltonian.from_arrays(gpu_ctx, arrays, n, K, emax)
        monkeypatch.setenv("LSQT_SWEEP_STEPS", "2")
        H2 = native.Hamiltonian.from_arrays(gpu_ctx, arrays, n, K, emax)
        monkeypatch.delenv("LSQT_SWEEP_STEPS")
        ph = rng.uniform(0, 2 * np.pi, n)
        re, im = np.cos(ph), np.sin(ph)
        for t in (0.4, 2.5, 7.0, 20.0):
            for direction in (1, -1):
                a, b = H.vector(re, im), H2.vector(re, im)
                H.evolve(direction, t, a)
                H2.evolve(direction, t, b)
                ga, gb = a.download(), b.download()
                assert_bits(ga[0], gb[0], f"{name} t={t} dir={direction} re")
                assert_bits(ga[1], gb[1], f"{name} t={t} dir={direction} im")
                a.close()
                b.close()
            a, b = H.vector(re, im), H2.vector(re, im)
            H.evolve_f32(1, t, a)
            H2.evolve_f32(1, t, b)
            ga, gb = a.download(), b.download()
            err = max(np.abs(ga[0] - gb[0]).max(), np.abs(ga[1] - gb[1]).max())
            assert err <= 2e-6, f"{name} t={t} f32: {err}"
            a.close()
            b.close()
        for o in (H, H2):
            o.close()
