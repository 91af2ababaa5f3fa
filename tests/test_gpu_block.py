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
        for rows in (256, 1024, 4096):
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
    """<L_r|T_m|R_r> with L != R (the VAC shape), against the FP64 path"""
    m, H = _setup(gpu_ctx, "cube_peierls")
    R = 4
    bl, lre, lim = _block_of(gpu_ctx, m, R)
    br, rre, rim = _block_of(gpu_ctx, m, R)
    mu = H.moments_block(bl, br, 40)
    for r in range(R):
        ref = H.moments(H.vector(lre[r], lim[r]), H.vector(rre[r], rim[r]), 40)
        assert np.abs(mu[r] - ref).max() <= 2e-6 * np.abs(ref).max()


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
    assert W in (1, 2, 4, 8) and rows in (256, 1024, 4096) and not blocked and us > 0
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
