"""lsqt_moments_noalloc (GPU): the moments of lsqt_moments from the fixed
per-GPU scratch, with no device allocation.

Two calls in a row -- a 24,000-site honeycomb (3 slots, two sites per
thread) at 200 moments (three full rings of partials and a partial one),
then a smaller 4,221-site square lattice (4 slots) at 64 moments -- each
bit-identical to lsqt_moments on the same inputs; a 1.0M-site lattice near
the size limit leaves the device's free memory unchanged; inputs beyond the
limits are rejected.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _model(tmp_path, name, write):
    import oracle
    from gpuqt_amd import inputs as gi
    d = tmp_path / name
    d.mkdir()
    write(gi, str(d))
    gi.write_para(str(d), anderson=1.0, moments=64, energy_max=10.0)
    gi.write_energy(str(d), [0.0])
    gi.write_time_steps(str(d), [1.0])
    return oracle.OracleModel(str(d))


def _free_device_memory():
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def test_noalloc_equals_moments(gpu_ctx, tmp_path):
    from gpuqt_amd import native
    # multi-step launches whose moments cross the 64-moment partial ring
    # (200, 131) and a set that fits it (64)
    cases = [("honeycomb", lambda gi, d: gi.write_honeycomb(d, 100, 60), 200),
             ("honeycomb131", lambda gi, d: gi.write_honeycomb(d, 100, 60), 131),
             ("square", lambda gi, d: gi.write_square(d, 67, 63), 64)]
    for name, write, M in cases:
        m = _model(tmp_path, name, write)
        print(f"[gpu] lsqt_moments_noalloc: {name}, {m.n} sites, {M} moments", flush=True)
        H = native.Hamiltonian.from_arrays(gpu_ctx, m.arrays(), m.n, m.max_neighbor,
                                           m.energy_max)
        left, right = H.vector(*m.random_state()), H.vector(*m.random_state())
        got = H.moments_noalloc(left, right, M)
        want = H.moments(left, right, M)
        np.testing.assert_array_equal(got, want)
        np.testing.assert_array_equal(H.moments_noalloc(right, right, M),
                                      H.moments(right, right, M))
        for o in (left, right, H):
            o.close()


def test_noalloc_allocates_nothing(gpu_ctx):
    """A fresh 1,000,000-site Hamiltonian: lsqt_moments would allocate its
    two recursion vectors (32 MB) and the partials on first use; the
    no-allocation entry point leaves the free memory as it was."""
    from gpuqt_amd import native
    L = 1000
    n = L * L
    idx = np.arange(n)
    x, y = idx // L, idx % L
    nl = np.stack([((x - 1) % L) * L + y, ((x + 1) % L) * L + y, x * L + (y - 1) % L,
                   x * L + (y + 1) % L], axis=1).astype(np.int32).reshape(-1)
    arrays = dict(neighbor_number=np.full(n, 4, np.int32), neighbor_list=nl,
                  potential=np.random.default_rng(1).uniform(-1, 1, n),
                  hopping_real=np.full(4 * n, -1.0), hopping_imag=np.zeros(4 * n),
                  xx=np.tile([-1.0, 1.0, 0.0, 0.0], n))
    print(f"[gpu] lsqt_moments_noalloc: {n} sites, free memory", flush=True)
    H = native.Hamiltonian.from_arrays(gpu_ctx, arrays, n, 4, 6.0)
    ph = np.random.default_rng(2).uniform(0, 2 * np.pi, n)
    v = H.vector(np.cos(ph), np.sin(ph))
    gpu_ctx.sync()
    before = _free_device_memory()
    got = H.moments_noalloc(v, v, 96)
    after = _free_device_memory()
    assert before - after < (1 << 20), f"free memory fell by {before - after} bytes"
    np.testing.assert_array_equal(got, H.moments(v, v, 96))
    for o in (v, H):
        o.close()


def test_noalloc_limits(gpu_ctx, tmp_path):
    from gpuqt_amd import native
    m = _model(tmp_path, "small", lambda gi, d: gi.write_square(d, 16, 16))
    H = native.Hamiltonian.from_arrays(gpu_ctx, m.arrays(), m.n, m.max_neighbor, m.energy_max)
    v = H.vector(*m.random_state())
    with pytest.raises(native.LsqtError, match="moments"):
        H.moments_noalloc(v, v, (1 << 16) + 1)
    with pytest.raises(native.LsqtError):
        H.moments_noalloc(v, v, 1)
    v.close()
    H.close()
    n = (1 << 20) + 64
    ring = np.stack([(np.arange(n) - 1) % n, (np.arange(n) + 1) % n], axis=1)
    arrays = dict(neighbor_number=np.full(n, 2, np.int32),
                  neighbor_list=ring.astype(np.int32).reshape(-1), potential=np.zeros(n),
                  hopping_real=np.full(2 * n, -1.0), hopping_imag=np.zeros(2 * n),
                  xx=np.zeros(2 * n))
    H = native.Hamiltonian.from_arrays(gpu_ctx, arrays, n, 2, 4.0)
    v = H.vector(np.ones(n), np.zeros(n))
    with pytest.raises(native.LsqtError, match="sites"):
        H.moments_noalloc(v, v, 16)
    v.close()
    H.close()
