"""The block path (R vectors per sparse sweep, FP32 states, FP64 moments)
against per-vector FP64 moments (GPU).

Stated tolerance: each vector's DOS row within 1e-6 (normwise) of the FP64
single-vector row; kernel shapes agree with each other to 1e-12."""
import os

import numpy as np
import pytest

from conftest import CASES_DIR

pytestmark = pytest.mark.gpu


def _setup(ctx, case):
    import oracle
    from gpuqt_amd import native
    d = os.path.join(CASES_DIR, case)
    m = oracle.OracleModel(d)
    H = native.Hamiltonian.from_arrays(ctx, m.arrays(), m.n, m.max_neighbor, m.energy_max)
    return m, H


def _block_of(ctx, m, R):
    from gpuqt_amd import native
    vecs = [m.random_state() for _ in range(R)]
    re = np.array([v[0] for v in vecs])
    im = np.array([v[1] for v in vecs])
    b = native.Block(ctx, m.n, R)
    b.upload(re, im)
    return b, re, im


def scale_of(mu):
    return np.abs(mu).max()


def _row(m, mu):
    from gpuqt_amd import hostlib
    return hostlib.correlation_row(mu, m.energy, m.energy_max, m.volume)


@pytest.mark.parametrize("case,R", [("honeycomb_small", 8), ("cube_peierls", 16),
                                    ("config1_honeycomb_20k", 32), ("square_vacancy_ldos", 4)])
def test_block_dos_against_f64(gpu_ctx, case, R):
    print(f"[gpu] block {case} R={R}", flush=True)
    m, H = _setup(gpu_ctx, case)
    M = min(m.number_of_moments, 400)
    b, re, im = _block_of(gpu_ctx, m, R)
    mu = H.moments_block(b, b, M)
    for r in range(R):
        v = H.vector(re[r], im[r])
        ref = H.moments(v, v, M)
        a, w = _row(m, mu[r]), _row(m, ref)
        err = np.abs(a - w).max() / np.abs(w).max()
        assert err <= 1e-6, f"vector {r}: normwise {err:.2e}"


@pytest.mark.parametrize("R", [8, 32])
def test_block_shapes_agree(gpu_ctx, R):
    m, H = _setup(gpu_ctx, "config1_honeycomb_20k")
    b, re, im = _block_of(gpu_ctx, m, R)
    results = {}
    for W in (1, 2, 4, 8):
        for rows in (64, 128, 256, 1024, 4096):
            for depth in (1, 2):
                H.block_set_config(R, W, rows, depth=depth)
                results[(W, rows, depth)] = H.moments_block(b, b, 64)
    H.block_set_config(R, 0)
    ref = results[(1, 256, 1)]
    for k, mu in results.items():
        np.testing.assert_allclose(mu, ref, rtol=0, atol=1e-12 * np.abs(ref).max(),
                                   err_msg=str(k))


@pytest.mark.parametrize("L,R", [(16, 16), (32, 8)])
def test_block_blocked_schedule(gpu_ctx, L, R):
    """Lattice detection on the simple-cubic Peierls model (slab L^2, row L)
    and the blocked tile schedule against the natural one and FP64."""
    from gpuqt_amd import inputs as gi
    from gpuqt_amd import native
    print(f"[gpu] blocked schedule cube L={L}", flush=True)
    arrays, n = gi.cubic_peierls_arrays(L, seed=11)
    H = native.Hamiltonian.from_arrays(gpu_ctx, arrays, n, 6, 15.0)
    info = H.info()
    assert (info.slab, info.row) == (L * L, L)
    M = 96
    rng = np.random.default_rng(3)
    ph = rng.uniform(0, 2 * np.pi, (R, n))
    b = native.Block(gpu_ctx, n, R)
    b.upload(np.cos(ph), np.sin(ph))
    H.block_set_config(R, 1, 256)
    ref = H.moments_block(b, b, M)
    tol = 1e-12 * np.abs(ref).max()
    for depth in (1, 2):
        H.block_set_config(R, 2, 256, depth=depth)
        np.testing.assert_allclose(H.moments_block(b, b, M), ref, rtol=0, atol=tol,
                                   err_msg=f"depth {depth}")
        np.testing.assert_allclose(H.moments_block_self(b, M), ref, rtol=0, atol=1e-5 * scale_of(ref),
                                   err_msg=f"self, depth {depth}")
    walked = 0
    for W in (1, 2, 4, 8):
        for rows in (64, 128, 256, 512, 1024):
            if (L * L) % rows:
                continue
            for sched in (1, 2):
                if sched == 2 and (rows % L or L % (256 * W // R)):
                    with pytest.raises(native.LsqtError):
                        H.block_set_config(R, W, rows, blocked=2)
                    continue
                walked += sched == 2
                H.block_set_config(R, W, rows, blocked=sched)
                np.testing.assert_allclose(H.moments_block(b, b, M), ref, rtol=0, atol=tol,
                                           err_msg=f"W={W} rows={rows} schedule {sched}")
    assert walked > 0
    H.block_set_config(R, 0)
    v = H.vector(np.cos(ph[R - 1]), np.sin(ph[R - 1]))
    mu64 = H.moments(v, v, M)
    assert np.abs(ref[R - 1] - mu64).max() <= 2e-6 * np.abs(mu64).max()
    for o in (v, b, H):
        o.close()


def _lattice3d(Lx, Ly, Lz, open_y=False, seed=5):
    """x-slowest 3D lattice (site (x Ly + y) Lz + z), neighbours -x, +x, -y,
    +y, -z, +z (periodic; open in y if asked: absent slots at the end),
    hopping -1 with a Peierls phase 2 pi 0.013 x on y-bonds, on-site
    U[-4, 4]; row-major ELL arrays."""
    n = Lx * Ly * Lz
    idx = np.arange(n)
    x, y, z = idx // (Ly * Lz), (idx // Lz) % Ly, idx % Lz
    site = lambda a, b, c: ((a % Lx) * Ly + (b % Ly)) * Lz + (c % Lz)
    ph = np.exp(2j * np.pi * 0.013 * x)
    cand = [(site(x - 1, y, z), np.full(n, -1.0 + 0j), np.ones(n, bool)),
            (site(x + 1, y, z), np.full(n, -1.0 + 0j), np.ones(n, bool)),
            (site(x, y - 1, z), -np.conj(ph), (y > 0) if open_y else np.ones(n, bool)),
            (site(x, y + 1, z), -ph, (y < Ly - 1) if open_y else np.ones(n, bool)),
            (site(x, y, z - 1), np.full(n, -1.0 + 0j), np.ones(n, bool)),
            (site(x, y, z + 1), np.full(n, -1.0 + 0j), np.ones(n, bool))]
    nn = np.zeros(n, np.int32)
    nl = np.zeros((n, 6), np.int32)
    hop = np.zeros((n, 6), complex)
    for j, t, ok in cand:
        nl[ok, nn[ok]] = j[ok]
        hop[ok, nn[ok]] = t[ok]
        nn[ok] += 1
    pot = np.random.default_rng(seed).uniform(-4, 4, n)
    return dict(neighbor_number=nn, neighbor_list=nl.reshape(-1), potential=pot,
                hopping_real=hop.real.reshape(-1).copy(), hopping_imag=hop.imag.reshape(-1).copy(),
                xx=np.zeros(n * 6)), n


@pytest.mark.parametrize("dims,open_y", [((6, 10, 32), False), ((5, 8, 64), True),
                                         ((4, 16, 16), False)])
def test_block_sliding_window(gpu_ctx, dims, open_y):
    """The sliding-window schedule (rows y - 1 .. y + 1 of a chunk in LDS) on
    3D lattices of uneven extents, periodic and open, every tile height that
    divides the rows per slab and every W: the same moments as the natural
    schedule, direct (<L|psi> with L != R) and by the product identities."""
    from gpuqt_amd import native
    Lx, Ly, Lz = dims
    print(f"[gpu] sliding window {dims} open_y={open_y}", flush=True)
    arrays, n = _lattice3d(Lx, Ly, Lz, open_y)
    H = native.Hamiltonian.from_arrays(gpu_ctx, arrays, n, 6, 12.0)
    info = H.info()
    assert (info.slab, info.row) == (Ly * Lz, Lz)
    R, M = 32, 80
    rng = np.random.default_rng(9)
    pl, pr = rng.uniform(0, 2 * np.pi, (2, R, n))
    bl, br = native.Block(gpu_ctx, n, R), native.Block(gpu_ctx, n, R)
    bl.upload(np.cos(pl), np.sin(pl))
    br.upload(np.cos(pr), np.sin(pr))
    H.block_set_config(R, 2, 256)
    ref = H.moments_block(bl, br, M)
    ref_self = H.moments_block_self(br, M)
    tol = 1e-12 * np.abs(ref).max()
    tol_self = 1e-12 * np.abs(ref_self).max()
    tried = 0
    for W in (1, 2, 4):
        chunk = 256 * W // R
        for by in range(1, Ly + 1):
            if Ly % by or Lz % chunk:
                continue
            H.block_set_config(R, W, by, blocked=3)
            tried += 1
            np.testing.assert_allclose(H.moments_block(bl, br, M), ref, rtol=0, atol=tol,
                                       err_msg=f"W={W} by={by}")
            np.testing.assert_allclose(H.moments_block_self(br, M), ref_self, rtol=0,
                                       atol=tol_self, err_msg=f"self W={W} by={by}")
            if W <= 2:  # pairs: 2W vectors per lane on the same window
                np.testing.assert_allclose(H.moments_block_pair(bl, br, M), ref, rtol=0,
                                           atol=2e-6 * np.abs(ref).max(),
                                           err_msg=f"pairs W={W} by={by}")
    assert tried > 0
    with pytest.raises(native.LsqtError):
        H.block_set_config(R, 8, 1, blocked=3)  # W > 4
    with pytest.raises(native.LsqtError):
        H.block_set_config(R, 2, 3, blocked=3)  # 3 does not divide the rows per slab
    H.block_set_config(R, 0)
    for o in (bl, br, H):
        o.close()


@pytest.mark.parametrize("case,R", [("cube_peierls", 8), ("config1_honeycomb_20k", 32),
                                    ("square_vacancy_ldos", 4)])
def test_block_self_products(gpu_ctx, case, R):
    """DOS moments from the product identities (M/2 steps) against the direct
    block moments and the FP64 single-vector path, for M across the 64-slot
    partial ring's boundaries."""
    print(f"[gpu] block self products {case} R={R}", flush=True)
    m, H = _setup(gpu_ctx, case)
    b, re, im = _block_of(gpu_ctx, m, R)
    Mmax = 400
    direct = H.moments_block(b, b, Mmax)
    for M in (2, 3, 4, 64, 65, 66, 129, 130, 131, Mmax):
        mu = H.moments_block_self(b, M)
        assert mu.shape == (R, M)
        scale = np.abs(direct[:, :1]).max()
        np.testing.assert_allclose(mu, direct[:, :M], rtol=0, atol=2e-5 * scale,
                                   err_msg=f"M={M}")
    for r in (0, R - 1):
        v = H.vector(re[r], im[r])
        ref = H.moments(v, v, Mmax)
        a, w = _row(m, mu[r].real.astype(complex)), _row(m, ref)
        err = np.abs(a - w).max() / np.abs(w).max()
        assert err <= 1e-6, f"vector {r}: DOS normwise {err:.2e}"


def test_block_left_right_differ(gpu_ctx):
    """<L_r|T_m|R_r> with L != R (the VAC shape), directly and from the left
    and right recursions side by side, against the FP64 path"""
    m, H = _setup(gpu_ctx, "cube_peierls")
    R = 4
    bl, lre, lim = _block_of(gpu_ctx, m, R)
    br, rre, rim = _block_of(gpu_ctx, m, R)
    mu = H.moments_block(bl, br, 41)
    for M in (2, 3, 40, 41):
        pair = H.moments_block_pair(bl, br, M)
        assert pair.shape == (R, M)
        np.testing.assert_allclose(pair, mu[:, :M], rtol=0, atol=2e-6 * np.abs(mu).max(),
                                   err_msg=f"M={M}")
    for r in range(R):
        ref = H.moments(H.vector(lre[r], lim[r]), H.vector(rre[r], rim[r]), 41)
        assert np.abs(mu[r] - ref).max() <= 2e-6 * np.abs(ref).max()
        assert np.abs(pair[r] - ref).max() <= 2e-6 * np.abs(ref).max()


@pytest.mark.parametrize("case,R", [("config1_honeycomb_20k", 8), ("cube_peierls", 32),
                                    ("honeycomb_small", 128)])
def test_block_pairs_across_shapes(gpu_ctx, case, R):
    """Pair moment sets (2R columns) for every block shape the pair route can
    take, across the 64-slot partial ring, against the direct moments."""
    print(f"[gpu] block pairs {case} R={R}", flush=True)
    m, H = _setup(gpu_ctx, case)
    bl, _, _ = _block_of(gpu_ctx, m, R)
    br, _, _ = _block_of(gpu_ctx, m, R)
    M = 150
    H.block_set_config(R, max(1, R // 64), 256)
    ref = H.moments_block(bl, br, M)
    tol = 2e-6 * np.abs(ref).max()
    for W in (1, 2, 4):
        for rows in (256, 1024):
            for depth in (1, 2):
                if R % W or R // W > 64:
                    continue
                H.block_set_config(R, W, rows, depth=depth)
                got = H.moments_block_pair(bl, br, M)
                np.testing.assert_allclose(got, ref, rtol=0, atol=tol,
                                           err_msg=f"W={W} rows={rows} depth={depth}")
    H.block_set_config(R, 0)
    got = H.moments_block_pair(bl, br, M)
    np.testing.assert_allclose(got, ref, rtol=0, atol=tol, err_msg="tuned shape")
    for o in (bl, br, H):
        o.close()


def test_block_round_trip_and_errors(gpu_ctx):
    from gpuqt_amd import native
    m, H = _setup(gpu_ctx, "honeycomb_small")
    b, re, im = _block_of(gpu_ctx, m, 8)
    got_re, got_im = b.download()
    np.testing.assert_array_equal(got_re, re.astype(np.float32).astype(np.float64))
    np.testing.assert_array_equal(got_im, im.astype(np.float32).astype(np.float64))
    with pytest.raises(native.LsqtError):
        native.Block(gpu_ctx, m.n, 12)  # not a power of two
    Hf = native.Hamiltonian.from_arrays(gpu_ctx, m.arrays(), m.n, m.max_neighbor, m.energy_max,
                                        layout=native.LAYOUT_FULL)
    with pytest.raises(native.LsqtError):
        Hf.moments_block(b, b, 10)
    with pytest.raises(native.LsqtError):
        H.block_set_config(8, 3, 256)
    if H.info().slab % 256 != 0:
        with pytest.raises(native.LsqtError):
            H.block_set_config(8, 2, 256, blocked=True)  # no slab that 256 divides


def test_autotuner(gpu_ctx, tmp_path, monkeypatch):
    monkeypatch.setenv("LSQT_TUNE_CACHE", str(tmp_path / "tune.txt"))
    m, H = _setup(gpu_ctx, "config1_honeycomb_20k")
    b, _, _ = _block_of(gpu_ctx, m, 16)
    W, rows, blocked, us = H.block_tune(b, b, force=True)
    assert W in (1, 2, 4, 8) and rows in (64, 128, 256, 1024, 4096) and not blocked and us > 0
    assert (tmp_path / "tune.txt").read_text().startswith("gfx950 16 3")
    assert H.block_tune(b, b) == (W, rows, blocked, us)


def test_block_large_streaming(gpu_ctx, tmp_path):
    import oracle
    from gpuqt_amd import inputs as gi
    from gpuqt_amd import native
    d = str(tmp_path)
    gi.write_honeycomb(d, 520, 520)
    gi.write_para(d, anderson=1.5, moments=64, energy_max=10.0)
    gi.write_energy(d, gi.energy_grid(-8, 8, 21))
    m = oracle.OracleModel(d)
    H = native.Hamiltonian.from_arrays(gpu_ctx, m.arrays(), m.n, m.max_neighbor, m.energy_max)
    b, re, im = _block_of(gpu_ctx, m, 32)
    mu = H.moments_block(b, b, 64)
    for r in (0, 31):
        ref = H.moments(H.vector(re[r], im[r]), H.vector(re[r], im[r]), 64)
        err = np.abs(_row(m, mu[r]) - _row(m, ref)).max() / np.abs(_row(m, ref)).max()
        assert err <= 1e-6
