"""The reference's method-level interface driven term by term (GPU).

The fused entry points run a whole evolution or moment set per call; the
reference's own functions instead call Hamiltonian::chebyshev_01 / _2 /
_1x / _2x once per Bessel term, rotating the states with Vector::swap
(src/sigma.cu:116-200), and compute moments with kernel_polynomial plus
Vector::inner_product_1 / inner_product_2 (src/sigma.cu:33-61).  Here those
sequences run (1) through the C++ host classes (lsqt_host_model_evolve /
_moments with by_terms) and (2) through the per-term C entry points from
Python, and must reproduce the reference's dumped states bit for bit and
its moments within the moment tolerance.
"""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import CASES_DIR, golden_cases
from parity import assert_bits, assert_moments

pytestmark = pytest.mark.gpu

_libm = C.CDLL("libm.so.6")
for _f in ("j0", "j1"):
    getattr(_libm, _f).restype = C.c_double
    getattr(_libm, _f).argtypes = [C.c_double]
_libm.jn.restype = C.c_double
_libm.jn.argtypes = [C.c_int, C.c_double]

EVOLUTIONS = (("evolve_p", 0, 1), ("evolve_m", 0, -1), ("evolvex_p", 1, 1),
              ("evolvex_m", 1, -1))


def _cases_with_time():
    out = []
    for case in golden_cases(with_states=True):
        g = np.load(os.path.join(CASES_DIR, case, "dump.npz"))
        if "v0_evolve_p_re" in g.files:
            out.append(case)
    return out


def _host_model(d):
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


@pytest.mark.parametrize("case", _cases_with_time())
def test_host_classes_term_by_term(case):
    print(f"[gpu] host classes term by term: {case}", flush=True)
    d = os.path.join(CASES_DIR, case)
    g = np.load(os.path.join(d, "dump.npz"))
    m = _host_model(d)
    _, ts, _ = m.inputs()
    t = float(ts[0] * m.energy_max)
    r = (g["v0_random_re"], g["v0_random_im"])
    for name, kind, direction in EVOLUTIONS:
        for by_terms in (True, False):
            re, im = m.evolve(kind, direction, t, *r, by_terms=by_terms)
            assert_bits(re, g[f"v0_{name}_re"], f"{case} {name} re (by_terms={by_terms})")
            assert_bits(im, g[f"v0_{name}_im"], f"{case} {name} im (by_terms={by_terms})")
    want = g["v0_dos_mu_re"] + 1j * g["v0_dos_mu_im"]
    assert m.number_of_moments == len(want)
    by_terms = m.moments(r, r, by_terms=True)
    fused = m.moments(r, r, by_terms=False)
    assert_moments(by_terms, want, f"{case} moments by terms")
    assert_moments(fused, want, f"{case} moments fused")
    m.close()


def _bessel_terms(kind, direction, t):
    """(2 J_m(t), phase label) for m >= 2 while |J_m| >= 1e-15, with the
    labels of src/sigma.cu:136-146 (evolve) and :180-190 (evolvex)."""
    out = []
    m = 2
    while True:
        b = _libm.jn(m, t)
        if -1e-15 < b < 1e-15:
            return out
        k = m % 4
        if kind == 0:
            label = {0: 1, 2: 2}.get(k, 3 if (k == 1) == (direction == 1) else 4)
        else:
            label = {1: 3, 3: 4}.get(k, 1 if (k == 0) == (direction == 1) else 2)
        out.append((2.0 * b, label))
        m += 1


@pytest.mark.parametrize("case", _cases_with_time())
def test_c_abi_term_by_term(gpu_ctx, case):
    """lsqt_chebyshev_01 / _2 / _1x / _2x called once per term from Python,
    states rotated by lsqt_vec_swap."""
    import oracle
    from gpuqt_amd import native
    print(f"[gpu] C ABI term by term: {case}", flush=True)
    d = os.path.join(CASES_DIR, case)
    g = np.load(os.path.join(d, "dump.npz"))
    om = oracle.OracleModel(d)
    H = native.Hamiltonian.from_arrays(gpu_ctx, om.arrays(), om.n, om.max_neighbor,
                                       om.energy_max)
    n = om.n
    t = float(om.time_step[0] * om.energy_max)
    r = (g["v0_random_re"], g["v0_random_im"])
    for name, kind, direction in EVOLUTIONS:
        state = H.vector(*r)
        p0, p1, p2 = state.clone(), native.Vector(gpu_ctx, n), native.Vector(gpu_ctx, n)
        H.apply(state, p1)
        if kind == 0:
            H.chebyshev_01(p0, p1, state, _libm.j0(t), 2.0 * _libm.j1(t), direction)
            for bm, label in _bessel_terms(kind, direction, t):
                H.chebyshev_2(p0, p1, p2, state, bm, label)
                p0.swap(p1)
                p1.swap(p2)
        else:
            x0, x1, x2 = (native.Vector(gpu_ctx, n) for _ in range(3))
            H.apply_commutator(state, x1)
            H.chebyshev_1x(x1, state, 2.0 * _libm.j1(t))
            for bm, label in _bessel_terms(kind, direction, t):
                H.chebyshev_2x(p0, x0, p1, x1, p2, x2, state, bm, label)
                p0.swap(p1)
                p1.swap(p2)
                x0.swap(x1)
                x1.swap(x2)
            for v in (x0, x1, x2):
                v.close()
        re, im = state.download()
        assert_bits(re, g[f"v0_{name}_re"], f"{case} {name} re")
        assert_bits(im, g[f"v0_{name}_im"], f"{case} {name} im")
        for v in (state, p0, p1, p2):
            v.close()
    H.close()


@pytest.mark.parametrize("case", ["honeycomb_small", "config1_honeycomb_20k"])
def test_inner_products_and_moment_vector(gpu_ctx, case):
    """Vector::inner_product_1 / inner_product_2 through the C ABI against
    per-block sums in numpy, and find_moments_chebyshev into a device vector
    against the host-array entry point (the same numbers)."""
    import oracle
    from gpuqt_amd import native
    print(f"[gpu] inner products: {case}", flush=True)
    d = os.path.join(CASES_DIR, case)
    om = oracle.OracleModel(d)
    n = om.n
    H = native.Hamiltonian.from_arrays(gpu_ctx, om.arrays(), n, om.max_neighbor, om.energy_max)
    rng = np.random.default_rng(4)
    a = rng.normal(size=n) + 1j * rng.normal(size=n)
    b = rng.normal(size=n) + 1j * rng.normal(size=n)
    va, vb = H.vector(a.real, a.imag), H.vector(b.real, b.imag)
    G = (n - 1) // 512 + 1
    M = 5
    parts = native.Vector(gpu_ctx, G * M + 3)
    for m in range(M):
        # moment m: the block sums of a_n conj(b_n) scaled by (m + 1)
        s = H.vector((m + 1) * a.real, (m + 1) * a.imag)
        s.inner_product_1(n, vb, parts, m * G)
        s.close()
    pr, pi = parts.download()
    blocks = np.add.reduceat(a * np.conj(b), np.arange(0, n, 512))
    for m in range(M):
        got = pr[m * G:(m + 1) * G] + 1j * pi[m * G:(m + 1) * G]
        np.testing.assert_allclose(got, (m + 1) * blocks, rtol=1e-12, atol=1e-12 * abs(blocks).max())
    out = native.Vector(gpu_ctx, M)
    parts.inner_product_2(n, M, out)
    orr, oi = out.download()
    total = np.sum(a * np.conj(b))
    np.testing.assert_allclose(orr + 1j * oi, total * np.arange(1, M + 1), rtol=1e-12)
    # sizes are checked
    with pytest.raises(native.LsqtError):
        va.inner_product_1(n, vb, parts, G * M + 3 - G + 1)
    with pytest.raises(native.LsqtError):
        parts.inner_product_2(n, M + 1, out)
    # device storage: one address per vector, exchanged by swap
    pa, pb = va.device_data(), vb.device_data()
    assert pa and pb and pa != pb
    va.swap(vb)
    assert (va.device_data(), vb.device_data()) == (pb, pa)
    # find_moments_chebyshev into a device vector
    r = om.random_state()
    v = H.vector(*r)
    Mm = 32
    mv = native.Vector(gpu_ctx, Mm + 1)
    H.moments_to_vec(v, v, Mm, mv)
    re, im = mv.download()
    np.testing.assert_array_equal(re[:Mm] + 1j * im[:Mm], H.moments(v, v, Mm))
    for o in (va, vb, parts, out, v, mv, H):
        o.close()
