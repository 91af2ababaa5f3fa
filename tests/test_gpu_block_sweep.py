"""Several Chebyshev steps per launch on the block path (GPU).

Periodic lattices built on the device -- one, or several of uneven size side
by side as a segmented Hamiltonian -- take the product-identity moment sets
(self products of R vectors, pairs of a left and a right recursion) four
(k_block_sweepn, up to 16 columns) or two (k_block_sweep2) steps per launch,
one launch for all systems.  The per-site
arithmetic is that of the one-step block kernels, so the states agree bit
for bit and the moments only by the grouping of their FP64 partial sums:
every moment set is checked against the same Hamiltonian built without the
sweep (LSQT_SWEEP=0) within 1e-12 of the set's scale -- a state that differed
in one FP32 rounding would miss that by orders of magnitude.  Also checked:
which Hamiltonians get the sweep, column counts 2 .. 32 (two and four
columns per lane), moment counts from 6 (one two-step launch) and 10 (one
four-step launch) up, odd and even, two steps per launch only
(LSQT_SWEEP_STEPS=2) against four, the VAC / MSD drivers, and hipGraph
replays against eager launches.
"""
import os

import numpy as np
import pytest

from parity import assert_moments

pytestmark = pytest.mark.gpu


def _lattice(tmp_path, name, kind, nx, ny, pbc=(1, 1, 0)):
    from gpuqt_amd import inputs as gi
    d = tmp_path / name
    d.mkdir()
    if kind == "honeycomb":
        gi.write_honeycomb(str(d), nx, ny, pbc=pbc)
    else:
        gi.write_square(str(d), nx, ny, pbc=pbc)
    return gi.read_lattice(os.path.join(str(d), "lattice.in"))


def _sites(lat):
    return int(np.prod(lat["cells"]) * lat["orbitals"])


# uneven systems side by side: whole-slab (68 and 120 sites per slab at 16
# columns: 32 positions per row) and chunked slabs, two runs for the longest
UNEVEN = [("honeycomb", 40, 30), ("honeycomb", 64, 90), ("honeycomb", 25, 200),
          ("honeycomb", 100, 17), ("honeycomb", 7, 333)]


def _pair_of_hams(ctx, tmp_path, monkeypatch, systems, emax=9.0, seed=0, tag="u", env=None):
    """(with the sweep, without it) on the same lattices and potential; env:
    settings for the first"""
    from gpuqt_amd import native
    lats = [_lattice(tmp_path, f"{tag}{k}", *s) for k, s in enumerate(systems)]
    n = sum(_sites(l) for l in lats)
    pot = np.random.default_rng(seed).uniform(-1.0, 1.0, n)
    for k, v in (env or {}).items():
        monkeypatch.setenv(k, v)
    Hs = native.Hamiltonian.from_lattices(ctx, lats, None, pot, emax)
    for k in (env or {}):
        monkeypatch.delenv(k)
    monkeypatch.setenv("LSQT_SWEEP", "0")
    H1 = native.Hamiltonian.from_lattices(ctx, lats, None, pot, emax)
    monkeypatch.delenv("LSQT_SWEEP")
    return Hs, H1, n


def _phases(n, R, seed):
    ph = np.random.default_rng(seed).uniform(0, 2 * np.pi, (R, n))
    return np.cos(ph), np.sin(ph)


def _check_sets(got, ref, what):
    got, ref = np.asarray(got), np.asarray(ref)
    assert got.shape == ref.shape
    flat_g = got.reshape(-1, got.shape[-1])
    flat_r = ref.reshape(-1, ref.shape[-1])
    for i in range(flat_g.shape[0]):
        assert_moments(flat_g[i], flat_r[i], f"{what} row {i}", rtol=0.0, scale_tol=1e-12)


def test_block_sweep_geometry(gpu_ctx, tmp_path, monkeypatch):
    from gpuqt_amd import native
    print("[gpu] block sweep geometry", flush=True)
    Hs, H1, n = _pair_of_hams(gpu_ctx, tmp_path, monkeypatch, UNEVEN)
    assert Hs.info().block_sweep_systems == len(UNEVEN)
    assert Hs.info().block_sweep_steps == 4
    assert H1.info().block_sweep_systems == 0 and H1.info().block_sweep_steps == 0
    for o in (Hs, H1):
        o.close()
    # honeycombs and square lattices (3 and 4 slots) side by side
    lats = [_lattice(tmp_path, "m0", "honeycomb", 12, 40), _lattice(tmp_path, "m1", "square", 9, 70)]
    pot = np.zeros(sum(_sites(l) for l in lats))
    H = native.Hamiltonian.from_lattices(gpu_ctx, lats, None, pot, 9.0)
    assert H.info().block_sweep_systems == 2
    H.close()
    # open y boundaries leave empty slots at the edges: no periodic table
    lats = [_lattice(tmp_path, "o0", "honeycomb", 12, 40),
            _lattice(tmp_path, "o1", "honeycomb", 10, 30, pbc=(1, 0, 0))]
    pot = np.zeros(sum(_sites(l) for l in lats))
    H = native.Hamiltonian.from_lattices(gpu_ctx, lats, None, pot, 9.0)
    assert H.info().block_sweep_systems == 0
    H.close()
    # vacancies renumber the sites
    lat = _lattice(tmp_path, "v0", "honeycomb", 12, 40)
    idx = np.arange(_sites(lat), dtype=np.int32)
    idx[7] = -1
    idx[8:] -= 1
    lats = [lat, _lattice(tmp_path, "v1", "honeycomb", 10, 30)]
    pot = np.zeros(_sites(lats[0]) - 1 + _sites(lats[1]))
    H = native.Hamiltonian.from_lattices(gpu_ctx, lats, [idx, None], pot, 9.0)
    assert H.info().block_sweep_systems == 0
    H.close()
    # one lattice (the single-vector sweep's geometry serves the block path)
    lat = _lattice(tmp_path, "s0", "honeycomb", 60, 100)
    H = native.Hamiltonian.from_lattices(gpu_ctx, [lat], None, np.zeros(_sites(lat)), 9.0)
    assert H.info().block_sweep_systems == 1 and H.info().sweep_slab == 400
    H.close()


@pytest.mark.parametrize("R,steps", [(2, "4"), (4, "4"), (8, "4"), (8, "2"), (16, "4"),
                                     (32, "4")])
def test_block_sweep_self_products(gpu_ctx, tmp_path, monkeypatch, R, steps):
    from gpuqt_amd import native
    print(f"[gpu] block sweep self products R={R}, {steps} steps", flush=True)
    Hs, H1, n = _pair_of_hams(gpu_ctx, tmp_path, monkeypatch, UNEVEN, seed=R,
                              env={"LSQT_SWEEP_STEPS": steps})
    assert Hs.info().block_sweep_steps == int(steps)
    re, im = _phases(n, R, R)
    b = native.Block(gpu_ctx, n, R)
    b.upload(re, im)
    for M in ([6, 7, 8, 9, 10, 11, 12, 14, 15, 16, 18, 26, 64, 131] if R == 8 else [10, 129]):
        got = Hs.moments_block_self(b, M)
        ref = H1.moments_block_self(b, M)
        _check_sets(got, ref, f"self R={R} M={M}")
    # the sweep's partial sums group the sites by its own tiles: the two
    # paths differ in the last bits somewhere (so the sweep did run)
    assert not np.array_equal(got, ref)
    for o in (b, Hs, H1):
        o.close()


@pytest.mark.parametrize("R,steps", [(1, "4"), (2, "4"), (8, "4"), (8, "2"), (16, "4")])
def test_block_sweep_pair_products(gpu_ctx, tmp_path, monkeypatch, R, steps):
    from gpuqt_amd import native
    print(f"[gpu] block sweep pair products R={R}, {steps} steps", flush=True)
    Hs, H1, n = _pair_of_hams(gpu_ctx, tmp_path, monkeypatch, UNEVEN, seed=10 + R,
                              env={"LSQT_SWEEP_STEPS": steps})
    re, im = _phases(n, 2 * R, 20 + R)
    L = native.Block(gpu_ctx, n, R)
    Rb = native.Block(gpu_ctx, n, R)
    L.upload(re[:R], im[:R])
    Rb.upload(re[R:], im[R:])
    for M in ([6, 7, 10, 11, 13, 14, 15, 100, 131] if R == 8 else [12, 65]):
        got = Hs.moments_block_pair(L, Rb, M)
        ref = H1.moments_block_pair(L, Rb, M)
        _check_sets(got, ref, f"pair R={R} M={M}")
    assert not np.array_equal(got, ref)
    for o in (L, Rb, Hs, H1):
        o.close()


def test_block_sweep_mixed_slots_and_single_lattice(gpu_ctx, tmp_path, monkeypatch):
    """honeycomb (3 slots) and square (4 slots) systems in one launch, and a
    single lattice with two runs of slabs per chunk"""
    from gpuqt_amd import native
    print("[gpu] block sweep mixed slot counts, single lattice", flush=True)
    for tag, systems in (("mix", [("honeycomb", 12, 40), ("square", 9, 70), ("square", 31, 64)]),
                         ("one", [("honeycomb", 300, 150)])):
        Hs, H1, n = _pair_of_hams(gpu_ctx, tmp_path, monkeypatch, systems, tag=tag)
        assert Hs.info().block_sweep_systems == len(systems)
        re, im = _phases(n, 8, 5)
        b = native.Block(gpu_ctx, n, 8)
        b.upload(re, im)
        _check_sets(Hs.moments_block_self(b, 96), H1.moments_block_self(b, 96), f"{tag} self")
        half = native.Block(gpu_ctx, n, 4)
        half.upload(re[:4], im[:4])
        other = native.Block(gpu_ctx, n, 4)
        other.upload(re[4:], im[4:])
        _check_sets(Hs.moments_block_pair(half, other, 97), H1.moments_block_pair(half, other, 97),
                    f"{tag} pairs")
        for o in (b, half, other, Hs, H1):
            o.close()


def test_block_sweep_drivers_and_graphs(gpu_ctx, tmp_path, monkeypatch):
    """VAC and MSD drivers (pair and self moment sets between the time
    steps) with and without the sweep; hipGraph replays equal eager launches
    bit for bit on every replay"""
    from gpuqt_amd import native
    print("[gpu] block sweep drivers and graph replays", flush=True)
    Hs, H1, n = _pair_of_hams(gpu_ctx, tmp_path, monkeypatch, UNEVEN[:3], seed=4, tag="d")
    re, im = _phases(n, 8, 6)
    ts = np.array([0.5, 1.0]) * 9.0
    M = 40
    for name in ("block_vac", "block_msd"):
        got = getattr(Hs, name)(re, im, M, ts)
        ref = getattr(H1, name)(re, im, M, ts)
        _check_sets(got, ref, name)
    b = native.Block(gpu_ctx, n, 8)
    b.upload(re, im)
    Hs.block_set_graphs(False)
    eager = (Hs.moments_block_self(b, 70), Hs.block_msd(re, im, M, ts))
    for rep in range(3):
        Hs.block_set_graphs(True)
        got = (Hs.moments_block_self(b, 70), Hs.block_msd(re, im, M, ts))
        for g, e in zip(got, eager):
            np.testing.assert_array_equal(g, e, err_msg=f"replay {rep}")
        Hs.block_set_graphs(False)
        np.testing.assert_array_equal(Hs.moments_block_self(b, 70), eager[0])
    for o in (b, Hs, H1):
        o.close()
