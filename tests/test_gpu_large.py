"""Maximum sizes (GPU): a moment set whose per-block partials exceed 2^31
bytes.  The reference keeps the partial sums of find_moments_chebyshev in a
Vector of ceil(N/512) * M complex values whose byte count is an int
(src/vector.h:46, src/vector.cu:47, src/sigma.cu:39-40), which overflows once
ceil(N/512) * M * 8 B >= 2^31, i.e. about 45M sites at 3000 moments.  Here
every size is a size_t; the test runs 48M sites at 3000 moments.

Oracle: the closed form of a periodic ring with hopping -1 and no on-site
energy.  Its eigenvectors are plane waves, eps_k = -2 cos(2 pi k / N), so
mu_m = sum_k |c^_k|^2 / N * T_m(eps_k / E_max) with c^ the discrete Fourier
transform of the random vector (T_m(x) = cos(m arccos x)).  A wrong index,
stride or overflow in the partial buffers changes these values by O(mu_0).

Tolerances (scale mu_0 = N): FP64 1e-11 mu_0 (3000 recursion steps of
rounding, ~1e-13 mu_0, plus the closed form's own sum), the precision path
1e-6 mu_0, the block path 2e-5 mu_0 on its moments."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 48_000_000
M = 3000
EMAX = 2.5
CHECK = (0, 1, 2, 3, 1499, 2998, 2999)


def _ring_arrays(n):
    i = np.arange(n, dtype=np.int64)
    nl = np.stack([(i - 1) % n, (i + 1) % n], axis=1).astype(np.int32).reshape(-1)
    xx = np.tile(np.array([-1.0, 1.0]), n)
    return dict(neighbor_number=np.full(n, 2, np.int32), neighbor_list=nl,
                potential=np.zeros(n), hopping_real=np.full(2 * n, -1.0),
                hopping_imag=np.zeros(2 * n), xx=xx)


def _closed_form(c, ms):
    n = len(c)
    w = np.abs(np.fft.fft(c)) ** 2 / n
    theta = np.arccos(-2.0 * np.cos(2.0 * np.pi * np.arange(n) / n) / EMAX)
    return {m: float(np.dot(w, np.cos(m * theta))) for m in ms}


def test_moment_partials_beyond_2gb(gpu_ctx):
    from gpuqt_amd import native
    assert -(-N // 512) * M * 8 >= 2 ** 31  # past the reference's int byte count
    print(f"[gpu] ring of {N} sites, {M} moments", flush=True)
    H = native.Hamiltonian.from_arrays(gpu_ctx, _ring_arrays(N), N, 2, EMAX)
    rng = np.random.default_rng(11)
    phases = rng.uniform(0.0, 2.0 * np.pi, N)
    re, im = np.cos(phases), np.sin(phases)
    del phases
    want = _closed_form(re + 1j * im, CHECK)
    v = H.vector(re, im)
    mu64 = H.moments(v, v, M)
    mu32 = H.moments_f32(v, v, M)
    for m in CHECK:
        assert abs(mu64[m] - want[m]) <= 1e-11 * N, f"FP64 mu_{m}: {mu64[m]} vs {want[m]}"
        assert abs(mu32[m] - want[m]) <= 1e-6 * N, f"FP32 mu_{m}: {mu32[m]} vs {want[m]}"
    print("[gpu] ring: FP64 and FP32 moment sets checked", flush=True)
    # the block path (R = 4, product identities): vector 0 against the closed
    # form, every vector's mu_0 = N (unit-modulus entries)
    R = 4
    blk_re = np.empty((R, N))
    blk_im = np.empty((R, N))
    blk_re[0], blk_im[0] = re, im
    for r in range(1, R):
        ph = rng.uniform(0.0, 2.0 * np.pi, N)
        blk_re[r], blk_im[r] = np.cos(ph), np.sin(ph)
    b = native.Block(gpu_ctx, N, R)
    b.upload(blk_re, blk_im)
    del blk_re, blk_im
    mub = H.moments_block_self(b, M)
    for m in CHECK:
        assert abs(mub[0][m] - want[m]) <= 2e-5 * N, f"block mu_{m}: {mub[0][m]} vs {want[m]}"
    np.testing.assert_allclose(mub[:, 0].real, N, rtol=1e-9)
    for o in (b, v, H):
        o.close()
