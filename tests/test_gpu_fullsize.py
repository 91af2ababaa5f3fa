"""The benchmarked workloads at their real sizes against the oracle (GPU).

The fixtures under tests/golden/fullsize/ come from the oracle run on the
same inputs at the same sizes (tests/golden/make_fullsize.py; the DOS of the
8.0M-site honeycomb was also checked there against the reference program's
own dos.out).  Here the same inputs are built by the drop-in's host Model and
run through the C ABI and the lsqt_gpu program:

  - 8.0M-site honeycomb, M = 3000: DOS moments within the FP64 moment
    tolerance (parity.py), the state after evolve / evolvex at t~ = 20 bit
    for bit, find_msd over two steps (moments and rows), the same with the
    slot pattern off (bit-identical), and the drop-in program's dos.out /
    msd.out rows at 17 digits within 1e-10;
  - 160^3 Peierls cube, M = 2048: FP64 moments of vector 0 within the moment
    tolerance; the block path (R = 32, direct and product identities) and
    the precision path: DOS rows of vectors 0 and 1 within 1e-5 (normwise);
  - realisation 8 of the uneven batch (106,128 sites, charged impurities),
    M = 1500, 16 steps: the FP64 program's VAC / MSD rows of vectors 0 and 1
    within 1e-10, and the batch path (segmented with realisation 7) within
    1e-5.
"""
import hashlib
import os

import numpy as np
import pytest

from conftest import GOLDEN
from parity import assert_moments, assert_rows, read_out

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

FIX = os.path.join(GOLDEN, "fullsize")
BLOCK_RTOL = 1e-5
# element-wise FP64 row checks apply where |value| > ROW_FLOOR x the row's
# largest: the moment sums differ from the oracle's by their reduction order
# (~1e-14 of the row's scale at these sizes), which is large against values
# near the zero crossings of a VAC row
ROW_FLOOR = 1e-3


def _fixture(name):
    return np.load(os.path.join(FIX, f"{name}.npz"), allow_pickle=False)


def _host_model(d):
    """The drop-in's Model of directory d (its banner goes to a pipe)."""
    from gpuqt_amd import hostlib
    devnull = os.open(os.devnull, os.O_WRONLY)
    saved = os.dup(1)
    os.dup2(devnull, 1)
    try:
        return hostlib.HostModel(d)
    finally:
        os.dup2(saved, 1)
        os.close(devnull)
        os.close(saved)


def _check_state(vec, fx, key):
    re, im = vec.download()
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(re, dtype="<f8").tobytes())
    h.update(np.ascontiguousarray(im, dtype="<f8").tobytes())
    idx = fx[f"{key}_idx"]
    # sampled entries first: they name the first difference if there is one
    np.testing.assert_array_equal(re[idx], fx[f"{key}_re"], err_msg=f"{key} re (sampled)")
    np.testing.assert_array_equal(im[idx], fx[f"{key}_im"], err_msg=f"{key} im (sampled)")
    assert h.hexdigest() == str(fx[f"{key}_sha"]), f"{key}: state differs from the oracle's"


def _normwise(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(b).max())


def _msd_moments(H, r_vec, t, M):
    """find_msd's two moment sets (src/sigma.cu:333-393) over the C ABI."""
    from gpuqt_amd import native
    n = H.n
    state, sx, copy = r_vec.clone(), r_vec.clone(), native.Vector(H.ctx, n)
    H.evolve(1, t, state)
    H.evolvex(1, t, sx)
    out = [H.moments(sx, sx, M)]
    copy.copy_from(state)
    H.evolvex(1, t, copy)
    H.evolve(1, t, sx)
    sx.add(copy)
    H.evolve(1, t, state)
    out.append(H.moments(sx, sx, M))
    for v in (state, sx, copy):
        v.close()
    return out


def test_config2_honeycomb_8m(gpu_ctx, tmp_path, monkeypatch):
    from gpuqt_amd import hostlib, native
    from gpuqt_amd import workloads as wl
    fx = _fixture("config2")
    d = str(tmp_path / "config2")
    os.makedirs(d)
    wl.write_config2(d, time_steps=2)
    print("[gpu] full size: 8.0M-site honeycomb, M = 3000", flush=True)
    m = _host_model(d)
    assert m.n == int(fx["n"])
    arrays = m.arrays()
    r = m.random_state()
    energies, _, _ = m.inputs()
    M = int(fx["moments"])
    t = float(fx["t_scaled"])
    H = native.Hamiltonian.from_arrays(gpu_ctx, arrays, m.n, m.max_neighbor, m.energy_max)
    assert H.info().pattern_period > 0
    v = H.vector(*r)
    mu = H.moments(v, v, M)
    assert_moments(mu, fx["dos_mu"], "DOS moments")
    assert_rows(hostlib.correlation_row(mu, energies, m.energy_max, m.volume), fx["dos_row"],
                "DOS row")
    s = H.vector(*r)
    H.evolve(1, t, s)
    _check_state(s, fx, "evolve")
    sx = H.vector(*r)
    H.evolvex(1, t, sx)
    _check_state(sx, fx, "evolvex")
    print("[gpu] full size: MSD over two steps", flush=True)
    msd = _msd_moments(H, v, t, M)
    for k in range(2):
        assert_moments(msd[k], fx["msd_mu"][k], f"MSD moments {k}")
        assert_rows(hostlib.correlation_row(msd[k], energies, m.energy_max, m.volume),
                    fx["msd_rows"][k], f"MSD row {k}", floor=ROW_FLOOR)
    # the same Hamiltonian without the slot pattern: bit-identical
    print("[gpu] full size: slot pattern off", flush=True)
    monkeypatch.setenv("LSQT_PATTERN", "0")
    H0 = native.Hamiltonian.from_arrays(gpu_ctx, arrays, m.n, m.max_neighbor, m.energy_max)
    monkeypatch.delenv("LSQT_PATTERN")
    assert H0.info().pattern_period == 0
    del arrays
    np.testing.assert_array_equal(H0.moments(v, v, M), mu)
    s0 = H0.vector(*r)
    H0.evolve(1, t, s0)
    _check_state(s0, fx, "evolve")
    for o in (v, s, sx, s0, H, H0):
        o.close()
    m.close()


def test_config2_dropin_program_8m(tmp_path):
    """The lsqt_gpu program on the same directory: dos.out and msd.out at 17
    digits against the oracle's rows."""
    from gpuqt_amd import hostlib
    from gpuqt_amd import workloads as wl
    fx = _fixture("config2")
    d = tmp_path / "config2"
    d.mkdir()
    wl.write_config2(str(d), time_steps=2)
    driver = tmp_path / "input.txt"
    driver.write_text(f"{d}\n")
    print("[gpu] full size: lsqt_gpu on the 8.0M-site honeycomb", flush=True)
    r = hostlib.run_cli(str(driver), digits=17, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert_rows(read_out(d / "dos.out"), fx["dos_row"][None, :], "dos.out")
    assert_rows(read_out(d / "msd.out"), fx["msd_rows"], "msd.out", floor=ROW_FLOOR)


def test_cube_160_block_and_precision(gpu_ctx):
    from gpuqt_amd import hostlib, native
    from gpuqt_amd import inputs as gi
    from gpuqt_amd import workloads as wl
    fx = _fixture("cube")
    M = int(fx["moments"])
    print("[gpu] full size: 160^3 Peierls cube, R = 32, M = 2048", flush=True)
    arrays, n = gi.cubic_peierls_arrays(wl.CUBE_L)
    assert n == int(fx["n"])
    H = native.Hamiltonian.from_arrays(gpu_ctx, arrays, n, 6, wl.CUBE_EMAX)
    del arrays
    re, im = wl.cube_phases(32, n)
    blk = native.Block(gpu_ctx, n, 32)
    blk.upload(re, im)
    vol = float(wl.CUBE_L ** 3)
    row = lambda mu: hostlib.correlation_row(mu, fx["energies"], wl.CUBE_EMAX, vol)
    v0 = H.vector(re[0], im[0])
    assert_moments(H.moments(v0, v0, M), fx["mu"][0], "FP64 moments, vector 0")
    mub = H.moments_block(blk, blk, M)
    mus = H.moments_block_self(blk, M)
    for r in range(2):
        want = fx["rows"][r]
        vr = H.vector(re[r], im[r])
        errs = {"block": _normwise(row(mub[r]), want), "block_self": _normwise(row(mus[r]), want),
                "f32": _normwise(row(H.moments_f32(vr, vr, M)), want)}
        print(f"[gpu] cube vector {r}: DOS normwise errors {errs}", flush=True)
        for k, e in errs.items():
            assert e <= BLOCK_RTOL, f"{k} DOS row of vector {r}: {e:.2e}"
        vr.close()
    for o in (v0, blk, H):
        o.close()


def test_batch_realisation_8(tmp_path):
    """Realisation 8 of the uneven batch at its real size through the FP64
    program and through LSQT_PATH=batch (segmented with realisation 7)."""
    from gpuqt_amd import hostlib
    from gpuqt_amd import workloads as wl
    fx = _fixture("batch8")
    cells = wl.batch_cells(64)
    k = int(fx["realisation"])
    assert tuple(cells[k]) == tuple(fx["cells"])
    dirs = {}
    for path in ("f64", "batch"):
        for i in ((k,) if path == "f64" else (k, 7)):
            d = tmp_path / f"{path}_r{i}"
            wl.write_batch_realisation(str(d), *cells[i])
            dirs[path, i] = d
        driver = tmp_path / f"input_{path}.txt"
        driver.write_text("".join(f"{d}\n" for (p, _), d in dirs.items() if p == path))
        print(f"[gpu] full size: batch realisation {k}, LSQT_PATH={path}", flush=True)
        r = hostlib.run_cli(str(driver), digits=17, timeout=600,
                            env=dict(os.environ, LSQT_PATH=path))
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    nt = len(fx["time_step"])
    for path, tol in (("f64", None), ("batch", BLOCK_RTOL)):
        d = dirs[path, k]
        for q in ("vac", "msd"):
            got = read_out(d / f"{q}.out")
            assert got.shape[0] == 8 * nt, f"{path} {q}.out rows"
            for v in range(2):
                want = fx[f"{q}_rows"][v]
                rows = got[v * nt:(v + 1) * nt]
                if tol is None:
                    assert_rows(rows, want, f"{path} {q}.out vector {v}", floor=ROW_FLOOR)
                else:
                    err = max(_normwise(a, b) for a, b in zip(rows, want))
                    print(f"[gpu] batch {q} vector {v}: normwise error {err:.2e}", flush=True)
                    assert err <= tol, f"{path} {q}.out vector {v}: {err:.2e}"
