"""Time-correlation drivers of the block path (VAC and MSD for R vectors at
once, FP32 recursion states, FP64 evolved states and moments) against the
oracle's find_vac / find_msd per vector, and segmented Hamiltonians (several
independent systems side by side in one launch per step) against the same
systems run one by one (GPU).

Stated tolerances: every VAC / MSD row (Jackson-damped Chebyshev sum) within
2e-5 of the FP64 oracle's row, normwise over the row; segmented and
one-by-one runs of the same systems agree to 1e-9 of the moment scale (the
same per-site arithmetic; only the grouping of the sums over tiles differs)."""
import os

import numpy as np
import pytest

from conftest import CASES_DIR

pytestmark = pytest.mark.gpu

ROW_TOL = 2e-5


def _model(case):
    import oracle
    return oracle.OracleModel(os.path.join(CASES_DIR, case))


def _ham(ctx, m, **kw):
    from gpuqt_amd import native
    return native.Hamiltonian.from_arrays(ctx, m.arrays(), m.n, m.max_neighbor, m.energy_max,
                                          **kw)


def _rows(m, mu):
    from gpuqt_amd import hostlib
    return np.array([hostlib.correlation_row(x, m.energy, m.energy_max, m.volume) for x in mu])


def _vectors(m, R):
    vecs = [m.random_state() for _ in range(R)]
    return np.array([v[0] for v in vecs]), np.array([v[1] for v in vecs]), vecs


def _normwise(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("case,R", [("honeycomb_small", 4), ("cube_peierls", 8),
                                    ("honeycomb_transport_y", 2), ("square_vacancy_ldos", 4),
                                    ("config1_honeycomb_20k", 4)])
def test_block_vac_against_oracle(gpu_ctx, case, R):
    print(f"[gpu] block VAC {case} R={R}", flush=True)
    m = _model(case)
    H = _ham(gpu_ctx, m)
    re, im, vecs = _vectors(m, R)
    ts = m.time_step * m.energy_max
    mu = H.block_vac(re, im, m.number_of_moments, ts)
    assert mu.shape == (len(ts), 1, R, m.number_of_moments)
    worst = 0.0
    for r in range(R):
        want_rows, _ = m.find_vac(vecs[r])
        got_rows = _rows(m, mu[:, 0, r])
        for t in range(len(ts)):
            worst = max(worst, _normwise(got_rows[t], want_rows[t]))
    print(f"    worst VAC row error {worst:.2e}", flush=True)
    assert worst <= ROW_TOL


@pytest.mark.parametrize("case,R", [("honeycomb_small", 4), ("cube_peierls", 8),
                                    ("strip_open", 2), ("charged_only", 4),
                                    ("config1_honeycomb_20k", 4)])
def test_block_msd_against_oracle(gpu_ctx, case, R):
    print(f"[gpu] block MSD {case} R={R}", flush=True)
    m = _model(case)
    H = _ham(gpu_ctx, m)
    re, im, vecs = _vectors(m, R)
    ts = m.time_step * m.energy_max
    mu = H.block_msd(re, im, m.number_of_moments, ts)
    worst = 0.0
    for r in range(R):
        want_rows, _ = m.find_msd(vecs[r])
        got_rows = _rows(m, mu[:, 0, r])
        for t in range(len(ts)):
            worst = max(worst, _normwise(got_rows[t], want_rows[t]))
    print(f"    worst MSD row error {worst:.2e}", flush=True)
    assert worst <= ROW_TOL


@pytest.mark.parametrize("R", [1, 4])
def test_block_spin_against_oracle(gpu_ctx, R):
    """find_spin_polarization for R vectors at once: S_z flips the sign of
    every odd site of the recursion block, states evolve forward in time."""
    print(f"[gpu] block spin general_real_spin R={R}", flush=True)
    m = _model("general_real_spin")
    H = _ham(gpu_ctx, m)
    re, im, vecs = _vectors(m, R)
    ts = m.time_step * m.energy_max
    mu = H.block_spin(re, im, m.number_of_moments, ts)
    assert mu.shape == (len(ts), 1, R, m.number_of_moments)
    worst = 0.0
    for r in range(R):
        want_rows, _ = m.find_spin(vecs[r])
        got_rows = _rows(m, mu[:, 0, r])
        for t in range(len(ts)):
            worst = max(worst, _normwise(got_rows[t], want_rows[t]))
    print(f"    worst spin row error {worst:.2e}", flush=True)
    assert worst <= ROW_TOL


def _concat(models):
    """row-major ELL arrays of independent systems side by side"""
    K = max(m.max_neighbor for m in models)
    starts = [0]
    parts = {k: [] for k in ("neighbor_number", "neighbor_list", "potential", "hopping_real",
                             "hopping_imag", "xx")}
    for m in models:
        a = m.arrays()
        n, k = m.n, m.max_neighbor
        nn = a["neighbor_number"]
        pad = lambda x, fill: np.hstack([x.reshape(n, k), np.full((n, K - k), fill, x.dtype)])
        nl = pad(a["neighbor_list"], 0)
        valid = np.arange(K)[None, :] < nn[:, None]
        nl = np.where(valid, nl + starts[-1], 0)
        parts["neighbor_number"].append(nn)
        parts["neighbor_list"].append(nl.reshape(-1))
        parts["potential"].append(a["potential"])
        for key in ("hopping_real", "hopping_imag", "xx"):
            parts[key].append(pad(a[key], 0.0).reshape(-1))
        starts.append(starts[-1] + n)
    arrays = {k: np.concatenate(v) for k, v in parts.items()}
    return arrays, starts, K


def test_segmented_matches_one_by_one(gpu_ctx):
    """Three honeycomb systems of different sizes (one energy_max): block
    moments, product-identity moments, VAC and MSD of the segmented
    Hamiltonian equal those of each system alone."""
    import oracle
    from gpuqt_amd import inputs as gi
    from gpuqt_amd import native
    print("[gpu] segmented block drivers", flush=True)
    import tempfile
    models = []
    with tempfile.TemporaryDirectory() as tmp:
        for i, (nx, ny) in enumerate(((6, 4), (11, 3), (4, 9))):
            d = os.path.join(tmp, f"s{i}")
            os.makedirs(d)
            gi.write_honeycomb(d, nx, ny)
            gi.write_para(d, anderson=1.0 + i, moments=48, energy_max=10.0, msd=True, vac=True)
            gi.write_energy(d, gi.energy_grid(-8, 8, 21))
            gi.write_time_steps(d, [1.0, 0.5])
            models.append(oracle.OracleModel(d))
    arrays, starts, K = _concat(models)
    n = starts[-1]
    Hs = native.Hamiltonian.from_arrays(gpu_ctx, arrays, n, K, 10.0, segment_starts=starts)
    assert Hs.segments == 3
    R, M = 4, 48
    rng = np.random.default_rng(9)
    ph = rng.uniform(0, 2 * np.pi, (R, n))
    re, im = np.cos(ph), np.sin(ph)
    ts = np.array([1.0, 0.5]) * 10.0
    W, rows = 1, 256
    Hs.block_set_config(R, W, rows)
    b = native.Block(gpu_ctx, n, R)
    b.upload(re, im)
    seg = {"moments": Hs.moments_block(b, b, M), "self": Hs.moments_block_self(b, M),
           "vac": Hs.block_vac(re, im, M, ts), "msd": Hs.block_msd(re, im, M, ts)}
    for k, m in enumerate(models):
        lo, hi = starts[k], starts[k + 1]
        H = _ham(gpu_ctx, m)
        H.block_set_config(R, W, rows)
        bk = native.Block(gpu_ctx, m.n, R)
        bk.upload(re[:, lo:hi], im[:, lo:hi])
        one = {"moments": H.moments_block(bk, bk, M), "self": H.moments_block_self(bk, M),
               "vac": H.block_vac(re[:, lo:hi], im[:, lo:hi], M, ts)[:, 0],
               "msd": H.block_msd(re[:, lo:hi], im[:, lo:hi], M, ts)[:, 0]}
        for key in one:
            got = seg[key][k] if key in ("moments", "self") else seg[key][:, k]
            want = one[key]
            scale = np.abs(want).max()
            np.testing.assert_allclose(got, want, rtol=0, atol=1e-9 * scale,
                                       err_msg=f"system {k} {key}")
        for o in (bk, H):
            o.close()
    for o in (b, Hs):
        o.close()


def test_segmented_errors(gpu_ctx):
    from gpuqt_amd import native
    m = _model("honeycomb_small")
    a = m.arrays()
    with pytest.raises(native.LsqtError, match="outside"):
        native.Hamiltonian.from_arrays(gpu_ctx, a, m.n, m.max_neighbor, m.energy_max,
                                       segment_starts=[0, m.n // 2, m.n])
    with pytest.raises(native.LsqtError):
        native.Hamiltonian.from_arrays(gpu_ctx, a, m.n, m.max_neighbor, m.energy_max,
                                       segment_starts=[0, m.n + 1])


def test_graph_replays_match_eager(gpu_ctx):
    """Moment sets replayed from captured hipGraphs equal the same kernels
    launched one by one, bit for bit, on every replay, with eager work in
    between."""
    from gpuqt_amd import native
    print("[gpu] hipGraph replays", flush=True)
    m = _model("cube_peierls")
    H = _ham(gpu_ctx, m)
    R, M = 8, 130
    re, im, _ = _vectors(m, R)
    b = native.Block(gpu_ctx, m.n, R)
    b.upload(re, im)
    ts = m.time_step * m.energy_max
    H.block_set_graphs(False)
    eager = (H.moments_block(b, b, M), H.moments_block_self(b, M),
             H.block_vac(re, im, 40, ts), H.block_msd(re, im, 40, ts))
    for rep in range(3):
        H.block_set_graphs(True)
        got = (H.moments_block(b, b, M), H.moments_block_self(b, M),
               H.block_vac(re, im, 40, ts), H.block_msd(re, im, 40, ts))
        for g, e in zip(got, eager):
            np.testing.assert_array_equal(g, e, err_msg=f"replay {rep}")
        H.block_set_graphs(False)
        np.testing.assert_array_equal(H.moments_block(b, b, M), eager[0])
    for o in (b, H):
        o.close()


def test_block_time_sliding_window(gpu_ctx):
    """The block operators, VAC / MSD / spin drivers with the sliding-window
    evolution kernels on a 3D lattice against the same with natural tiles:
    operators bit for bit (the same per-site arithmetic), driver moments to
    1e-9 of their scale (the norms are summed over differently cut tiles)."""
    from gpuqt_amd import native
    from test_gpu_block import _lattice3d
    print("[gpu] block time evolution, sliding window", flush=True)
    arrays, n = _lattice3d(6, 10, 32)
    arrays["xx"] = np.tile(np.array([-1.0, 1.0, 0.0, 0.0, 0.0, 0.0]), n)
    H = native.Hamiltonian.from_arrays(gpu_ctx, arrays, n, 6, 12.0)
    assert (H.info().slab, H.info().row) == (320, 32)
    R, M = 32, 48
    rng = np.random.default_rng(21)
    ph = rng.uniform(0, 2 * np.pi, (R, n))
    re, im = np.cos(ph), np.sin(ph)
    re[:, 1::2] = im[:, 1::2] = 0.0  # spin vectors: even sites only
    ts = np.array([1.5, 2.0, 1.0]) * 12.0
    b_in = native.Block(gpu_ctx, n, R)
    b_in.upload(np.cos(ph), np.sin(ph))
    b_out = native.Block(gpu_ctx, n, R)
    results = {}
    for name, shape in (("natural", (2, 256, 0)), ("window", (2, 5, 3)), ("window W4", (4, 2, 3))):
        W, rows, sched = shape
        H.block_set_config(R, W, rows, blocked=sched)
        ops = []
        for op in (0, 1, 2):
            H.block_apply(op, b_in, b_out)
            ops.append(b_out.download())
        results[name] = (ops, H.block_vac(re, im, M, ts), H.block_msd(re, im, M, ts),
                         H.block_spin(re, im, M, ts))
    H.block_set_config(R, 0)
    ref = results["natural"]
    for name in ("window", "window W4"):
        got = results[name]
        for op in range(3):
            np.testing.assert_array_equal(got[0][op][0], ref[0][op][0], err_msg=f"{name} op {op}")
            np.testing.assert_array_equal(got[0][op][1], ref[0][op][1], err_msg=f"{name} op {op}")
        for k, what in ((1, "VAC"), (2, "MSD"), (3, "spin")):
            scale = np.abs(ref[k]).max()
            assert np.abs(got[k] - ref[k]).max() <= 1e-9 * scale, f"{name} {what}"
    for o in (b_in, b_out, H):
        o.close()
