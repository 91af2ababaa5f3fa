"""The general-model files: the C++ writer and the fast reader of the drop-in
Model (src/model_general.cu:32-185) -- round trips from built arrays through
the files back to the same arrays, and the fast reader against the stream
reader on ordinary and unusual inputs (CPU)."""
import os
import shutil

import numpy as np
import pytest

from conftest import CASES_DIR
from gpuqt_amd import hostlib
from gpuqt_amd import inputs as gi


def _load(d, stream=False):
    old = os.environ.get("LSQT_STREAM_READER")
    if stream:
        os.environ["LSQT_STREAM_READER"] = "1"
    else:
        os.environ.pop("LSQT_STREAM_READER", None)
    try:
        devnull = os.open(os.devnull, os.O_WRONLY)
        saved = os.dup(1)
        os.dup2(devnull, 1)
        try:
            return hostlib.HostModel(str(d))
        finally:
            os.dup2(saved, 1)
            os.close(devnull)
            os.close(saved)
    finally:
        if old is None:
            os.environ.pop("LSQT_STREAM_READER", None)
        else:
            os.environ["LSQT_STREAM_READER"] = old


def _general_dir(d, moments=16):
    (d / "para.in").write_text(f"model 0\nnumber_of_moments {moments}\nenergy_max 15.0\n")
    gi.write_energy(str(d), [0.0])


def _same_bits(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, what
    assert np.array_equal(a.view(np.uint64) if a.dtype == np.float64 else a,
                          b.view(np.uint64) if b.dtype == np.float64 else b), what


def test_cube_round_trip(tmp_path):
    """Peierls cube arrays -> files -> Model arrays, bit for bit (positions
    are whole numbers, so xx = reduced x differences is exact)."""
    L = 10
    arrays, n = gi.cubic_peierls_arrays(L, seed=4)
    x = (np.arange(n) // (L * L)).astype(float)
    _general_dir(tmp_path)
    hostlib.write_general(str(tmp_path), arrays, n, 6, x, float(L), float(L ** 3))
    m = _load(tmp_path)
    got = m.arrays()
    assert m.n == n and m.max_neighbor == 6 and m.volume == float(L ** 3)
    for k in ("neighbor_number", "neighbor_list", "potential", "hopping_real", "hopping_imag",
              "xx"):
        _same_bits(got[k], arrays[k], k)


def test_lattice_round_trip(tmp_path):
    """A lattice-built honeycomb with Anderson disorder -> general files ->
    the same neighbour lists, hoppings and potential bit for bit; xx from the
    positions agrees to rounding (the lattice builds it from cell offsets)."""
    src = tmp_path / "lattice"
    src.mkdir()
    nx, ny = 7, 5
    gi.write_honeycomb(str(src), nx, ny)
    gi.write_para(str(src), anderson=2.0, moments=16, energy_max=10.0)
    gi.write_energy(str(src), [0.0])
    m = _load(src)
    a = m.arrays()
    n = m.n
    # site n = ((x Ny + y) Nz + z) N_orbital + orbital with Nz = 1, 4 orbitals
    cell_x = np.array([o[0] for o in gi.HONEYCOMB_ORBITALS])
    idx = np.arange(n)
    x = cell_x[idx % 4] + 1.7321 * (idx // (4 * ny))
    dst = tmp_path / "general"
    dst.mkdir()
    _general_dir(dst)
    hostlib.write_general(str(dst), a, n, m.max_neighbor, x, 1.7321 * nx, m.volume)
    g = _load(dst).arrays()
    for k in ("neighbor_number", "neighbor_list", "potential", "hopping_real", "hopping_imag"):
        _same_bits(g[k], a[k], k)
    np.testing.assert_allclose(g["xx"], a["xx"], rtol=0, atol=1e-12)


@pytest.mark.parametrize("case", ["cube_peierls", "general_defaults", "general_real_spin"])
def test_fast_reader_matches_stream_on_golden_cases(case, tmp_path):
    d = tmp_path / case
    shutil.copytree(os.path.join(CASES_DIR, case), d,
                    ignore=shutil.ignore_patterns("expected", "dump.npz"))
    fast, slow = _load(d).arrays(), _load(d, stream=True).arrays()
    for k in fast:
        _same_bits(fast[k], slow[k], k)


UNUSUAL = ["+1.5", "1e", "1.5x", "0x1p3", "inf", "nan", ".5", "5.", "-0", "1E+2", "1e-400",
           "1e400", "--1", "+-2", "7,5", "３"]


@pytest.mark.parametrize("token", UNUSUAL)
def test_fast_reader_matches_stream_on_unusual_tokens(token, tmp_path):
    """A token outside the plain grammar in the hoppings, the potential or the
    positions gives the same Model (or the same error) as the stream reader."""
    for where in ("hopping", "potential", "position", "neighbor"):
        d = tmp_path / f"{where}"
        shutil.copytree(os.path.join(CASES_DIR, "cube_peierls"), d,
                        ignore=shutil.ignore_patterns("expected", "dump.npz"))
        path = d / f"{where}.in"
        words = path.read_text().split(" ")
        k = min(7, len(words) - 1)
        words[k] = token + words[k][len(words[k].rstrip("\n")):]
        path.write_text(" ".join(words))
        results = []
        for stream in (False, True):
            try:
                results.append(("ok", _load(d, stream=stream).arrays()))
            except hostlib.InputError as e:
                results.append(("error", str(e)))
        assert results[0][0] == results[1][0], (where, token, results)
        if results[0][0] == "ok":
            for k in results[0][1]:
                _same_bits(results[0][1][k], results[1][1][k], f"{where} {token} {k}")
        else:
            assert results[0][1] == results[1][1]


def test_truncated_and_whitespace(tmp_path):
    d = tmp_path / "t"
    shutil.copytree(os.path.join(CASES_DIR, "cube_peierls"), d,
                    ignore=shutil.ignore_patterns("expected", "dump.npz"))
    text = (d / "hopping.in").read_text()
    (d / "hopping.in").write_text(text.replace(" ", "\t  \r\n")[: len(text) // 2])
    (d / "potential.in").write_text("")
    fast, slow = _load(d).arrays(), _load(d, stream=True).arrays()
    for k in fast:
        _same_bits(fast[k], slow[k], k)


@pytest.mark.slow
def test_large_cube_round_trip_speed(tmp_path):
    """The 160^3 Peierls cube (4.1M sites, ~1 GB of text): written and read
    back bit for bit; prints the times."""
    import time
    L = 160
    arrays, n = gi.cubic_peierls_arrays(L, seed=1)
    x = (np.arange(n) // (L * L)).astype(float)
    _general_dir(tmp_path)
    t = time.perf_counter()
    hostlib.write_general(str(tmp_path), arrays, n, 6, x, float(L), float(L ** 3))
    tw = time.perf_counter() - t
    size = sum(os.path.getsize(tmp_path / f) for f in
               ("neighbor.in", "hopping.in", "position.in", "potential.in"))
    t = time.perf_counter()
    got = _load(tmp_path).arrays()
    tr = time.perf_counter() - t
    print(f"\n  {size / 1e6:.0f} MB written in {tw:.1f} s, read in {tr:.1f} s")
    for k in ("neighbor_list", "hopping_real", "hopping_imag", "potential", "xx"):
        _same_bits(got[k], arrays[k], k)
