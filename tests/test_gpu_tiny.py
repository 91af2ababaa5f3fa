"""Tiny and ragged systems through every device path (GPU): general-model
inputs with 1 .. 130 sites (less than one wave, exactly one, one more),
random Hermitian complex hoppings and ragged neighbour counts from 0 to
max_neighbor (absent slots), against the oracle.

Tolerances: FP64 states bit-identical, FP64 moments as in parity.py; the
precision path's moments within 1e-6 and the block path's (FP32 recursion)
within 2e-5 of the moment scale."""
import numpy as np
import pytest

from parity import assert_bits, assert_moments

pytestmark = pytest.mark.gpu

SIZES = [1, 2, 3, 5, 63, 64, 65, 130]


def _tiny_dir(d, n, max_neighbor=4, seed=0):
    """A general-model directory: random symmetric graph with degrees <=
    max_neighbor (some sites isolated), Hermitian hoppings, positions in a
    periodic box, random on-site energies."""
    from gpuqt_amd import inputs as gi
    rng = np.random.default_rng(1000 + n + seed)
    nb = [[] for _ in range(n)]
    hop = {}
    pairs = [(i, j) for i in range(n) for j in range(i + 1, n)]
    rng.shuffle(pairs)
    for i, j in pairs[: 3 * n]:
        if len(nb[i]) < max_neighbor and len(nb[j]) < max_neighbor:
            t = complex(rng.uniform(-1.5, 1.5), rng.uniform(-1.0, 1.0))
            nb[i].append(j)
            nb[j].append(i)
            hop[(i, j)] = t
            hop[(j, i)] = t.conjugate()
    hoppings = [[hop[(i, j)] for j in nb[i]] for i in range(n)]
    box = float(n + 3)
    gi.write_general(d, nb, max_neighbor, rng.uniform(0, box, n), box, float(n),
                     potential=rng.uniform(-1, 1, n), hoppings=hoppings)
    gi.write_para(d, model=0, moments=40, energy_max=9.0, msd=True)
    gi.write_energy(d, gi.energy_grid(-5, 5, 11))
    gi.write_time_steps(d, [1.5])
    return sum(len(r) for r in nb)


@pytest.mark.parametrize("n", SIZES)
def test_tiny_systems(gpu_ctx, tmp_path, n):
    import oracle
    from gpuqt_amd import native
    print(f"[gpu] tiny general model n={n}", flush=True)
    d = str(tmp_path)
    _tiny_dir(d, n)
    m = oracle.OracleModel(d)
    assert m.n == n
    for layout in (native.LAYOUT_FULL, native.LAYOUT_AUTO):
        H = native.Hamiltonian.from_arrays(gpu_ctx, m.arrays(), n, m.max_neighbor,
                                           m.energy_max, layout=layout)
        r = m.random_state()
        v = H.vector(*r)
        out = H.vector()
        for name, fn, ofn in (("apply", H.apply, m.apply),
                              ("commutator", H.apply_commutator, m.apply_commutator),
                              ("current", H.apply_current, m.apply_current)):
            fn(v, out)
            got, want = out.download(), ofn(*r)
            assert_bits(got[0], want[0], f"n={n} {name} re")
            assert_bits(got[1], want[1], f"n={n} {name} im")
        t = 1.5 * m.energy_max
        for fn, ofn in ((H.evolve, m.evolve), (H.evolvex, m.evolvex)):
            s = H.vector(*r)
            fn(-1, t, s)
            got, want = s.download(), ofn(-1, t, *r)
            assert_bits(got[0], want[0], f"n={n} {fn.__name__} re")
            assert_bits(got[1], want[1], f"n={n} {fn.__name__} im")
        M = 40
        ref = m.moments(r, r, M)
        ref = ref[0] + 1j * ref[1]
        assert_moments(H.moments(v, v, M), ref, f"n={n} moments")
        scale = np.abs(ref).max()
        assert np.abs(H.moments_f32(v, v, M) - ref).max() <= 1e-6 * scale, f"n={n} f32"
        if layout == native.LAYOUT_AUTO:
            R = 4
            vecs = [r] + [m.random_state() for _ in range(R - 1)]
            b = native.Block(gpu_ctx, n, R)
            b.upload(np.array([x[0] for x in vecs]), np.array([x[1] for x in vecs]))
            direct = H.moments_block(b, b, M)
            selfp = H.moments_block_self(b, M)
            for k, x in enumerate(vecs):
                want = m.moments(x, x, M)
                want = want[0] + 1j * want[1]
                sc = np.abs(want).max()
                assert np.abs(direct[k] - want).max() <= 2e-5 * sc, f"n={n} block {k}"
                assert np.abs(selfp[k] - want).max() <= 2e-5 * sc, f"n={n} block self {k}"
            b.close()
        for o in (v, out, H):
            o.close()
