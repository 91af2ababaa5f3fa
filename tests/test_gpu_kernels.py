"""GPU parity of the FP64 kernels against the reference's own bits.

The golden dumps (tests/golden/cases/*/dump.npz) were written by the
reference compiled from source (tests/golden/make_golden.py).  Inputs here
are the dumped random vectors and the oracle's model arrays (themselves
pinned bit-for-bit to the dumped arrays by test_oracle_golden.py).  States
must match bit for bit, moments and rows to the tolerances in parity.py.
"""
import os

import numpy as np
import pytest

from conftest import CASES_DIR, golden_cases
from parity import assert_bits, assert_moments, assert_rows

pytestmark = pytest.mark.gpu

LAYOUTS = {"packed": 2, "full": 1}


def _load(case):
    import oracle
    d = os.path.join(CASES_DIR, case)
    g = np.load(os.path.join(d, "dump.npz"))
    m = oracle.OracleModel(d)
    return d, g, m


def _ham(ctx, m, layout):
    from gpuqt_amd import native
    return native.Hamiltonian.from_arrays(ctx, m.arrays(), m.n, m.max_neighbor, m.energy_max,
                                          layout=layout)


@pytest.mark.parametrize("layout", list(LAYOUTS))
@pytest.mark.parametrize("case", golden_cases(with_states=True))
def test_operators_bitwise(gpu_ctx, case, layout):
    print(f"[gpu] operators {case} {layout}", flush=True)
    d, g, m = _load(case)
    H = _ham(gpu_ctx, m, LAYOUTS[layout])
    r = (g["v0_random_re"], g["v0_random_im"])
    v = H.vector(*r)
    out = H.vector()
    for name, fn in (("apply", H.apply), ("commutator", H.apply_commutator),
                     ("current", H.apply_current)):
        fn(v, out)
        re, im = out.download()
        assert_bits(re, g[f"v0_{name}_re"], f"{case} {name} re")
        assert_bits(im, g[f"v0_{name}_im"], f"{case} {name} im")
    s1 = H.vector()
    H.apply(v, s1)
    s2 = H.vector()
    H.kernel_polynomial(v, s1, s2)
    re, im = s2.download()
    assert_bits(re, g["v0_kpoly_re"], f"{case} kpoly re")
    assert_bits(im, g["v0_kpoly_im"], f"{case} kpoly im")
    # in place over s0
    s0 = v.clone()
    H.kernel_polynomial(s0, s1, s0)
    re, im = s0.download()
    assert_bits(re, g["v0_kpoly_re"], f"{case} kpoly in-place re")


@pytest.mark.parametrize("layout", list(LAYOUTS))
@pytest.mark.parametrize("case", golden_cases(with_states=True))
def test_evolution_bitwise(gpu_ctx, case, layout):
    d, g, m = _load(case)
    if m.number_of_steps_correlation == 0:
        pytest.skip("case has no time steps")
    print(f"[gpu] evolution {case} {layout}", flush=True)
    H = _ham(gpu_ctx, m, LAYOUTS[layout])
    t = m.time_step[0] * m.energy_max
    r = (g["v0_random_re"], g["v0_random_im"])
    for name, fn, direction in (("evolve_p", H.evolve, 1), ("evolve_m", H.evolve, -1),
                                ("evolvex_p", H.evolvex, 1), ("evolvex_m", H.evolvex, -1)):
        s = H.vector(*r)
        fn(direction, t, s)
        re, im = s.download()
        assert_bits(re, g[f"v0_{name}_re"], f"{case} {name} re")
        assert_bits(im, g[f"v0_{name}_im"], f"{case} {name} im")


def _moment_row(oracle_mod, m, mu):
    M = len(mu)
    re, im = oracle_mod.damping(mu.real.copy(), mu.imag.copy())
    return oracle_mod.summation(re, m.energy, m.energy_max, m.volume)


@pytest.mark.parametrize("layout", list(LAYOUTS))
@pytest.mark.parametrize("case", golden_cases())
def test_driver_sequences(gpu_ctx, case, layout):
    """The moment sets of find_dos / find_vac0 / find_vac / find_msd /
    find_spin_polarization and LDOS (src/sigma.cu:204-442), composed from
    the fused device entry points in the drivers' order."""
    import oracle
    print(f"[gpu] drivers {case} {layout}", flush=True)
    d, g, m = _load(case)
    H = _ham(gpu_ctx, m, LAYOUTS[layout])
    M = m.number_of_moments
    Nt = m.number_of_steps_correlation
    nvec = int(g["meta"][3])

    def check(tag, mu):
        assert_moments(mu, g[f"{tag}_mu_re"] + 1j * g[f"{tag}_mu_im"], tag)
        assert_rows(_moment_row(oracle, m, mu), g[f"{tag}_row"], tag)

    for i in range(nvec):
        v = f"v{i}"
        # random vectors come from the oracle's generator (pinned to the dump
        # where it holds the vectors)
        r = m.random_state()
        if f"{v}_random_re" in g:
            assert_bits(r[0], g[f"{v}_random_re"], "random state")
        rs = H.vector(*r)
        check(f"{v}_dos", H.moments(rs, rs, M))
        if m.calculate_vac0:
            st = H.vector()
            H.apply_current(rs, st)
            check(f"{v}_vac0", H.moments(st, st, M))
        if m.calculate_vac:
            left, left_copy, right = rs.clone(), H.vector(), H.vector()
            H.apply_current(left, right)
            for s in range(Nt):
                H.apply_current(left, left_copy)
                check(f"{v}_vac{s}", H.moments(right, left_copy, M))
                if s < Nt - 1:
                    t = m.time_step[s] * m.energy_max
                    H.evolve(-1, t, left)
                    H.evolve(-1, t, right)
        if m.calculate_msd:
            state, state_x, copy = rs.clone(), rs.clone(), H.vector()
            t = m.time_step[0] * m.energy_max
            H.evolve(1, t, state)
            H.evolvex(1, t, state_x)
            for s in range(Nt):
                check(f"{v}_msd{s}", H.moments(state_x, state_x, M))
                if s < Nt - 1:
                    t = m.time_step[s + 1] * m.energy_max
                    copy.copy_from(state)
                    H.evolvex(1, t, copy)
                    H.evolve(1, t, state_x)
                    state_x.add(copy)
                    H.evolve(1, t, state)
        if m.calculate_spin:
            state, sz = rs.clone(), H.vector()
            for s in range(Nt):
                sz.apply_sz(state)
                check(f"{v}_spin{s}", H.moments(state, sz, M))
                if s < Nt - 1:
                    H.evolve(1, m.time_step[s] * m.energy_max, state)
    if m.calculate_ldos:
        for i, orb in enumerate(m.local_orbitals):
            r = m.random_state(int(orb))
            rs = H.vector(*r)
            check(f"ldos{i}", H.moments(rs, rs, M))


def test_layout_choice(gpu_ctx):
    d, g, m = _load("honeycomb_small")
    H = _ham(gpu_ctx, m, 0)
    info = H.info()
    assert info.layout == 2 and info.classes <= 12
    Hf = _ham(gpu_ctx, m, 1)
    assert Hf.info().layout == 1


def test_vector_ops(gpu_ctx):
    from gpuqt_amd import native
    rng = np.random.default_rng(5)
    n = 1000
    a = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    va = native.Vector.from_host(gpu_ctx, a.real, a.imag)
    vb = native.Vector.from_host(gpu_ctx, b.real, b.imag)
    z = native.Vector(gpu_ctx, n)
    re, im = z.download()
    assert not re.any() and not im.any()
    va.add(vb)
    re, im = va.download()
    assert_bits(re, a.real + b.real, "add re")
    assert_bits(im, a.imag + b.imag, "add im")
    z.apply_sz(vb)
    re, im = z.download()
    sign = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    assert_bits(re, sign * b.real, "sz")
    va.swap(vb)
    re, _ = va.download()
    assert_bits(re, b.real, "swap")
    dot = vb.dot(va)  # conj(a + b) . b
    ref = np.vdot(a + b, b)
    assert abs(dot - ref) <= 1e-12 * np.abs(a + b).dot(np.abs(b))


def test_error_paths(gpu_ctx):
    from gpuqt_amd import native
    d, g, m = _load("honeycomb_small")
    H = _ham(gpu_ctx, m, 0)
    v = H.vector()
    with pytest.raises(native.LsqtError):
        H.apply(v, v)
    with pytest.raises(native.LsqtError):
        H.moments(v, v, 1)
    w = native.Vector(gpu_ctx, m.n + 1)
    with pytest.raises(native.LsqtError):
        H.apply(v, w)
    with pytest.raises(native.LsqtError):
        H.evolve(2, 1.0, v)
    a = m.arrays()
    bad = dict(a)
    bad["neighbor_list"] = a["neighbor_list"].copy()
    bad["neighbor_list"][0] = m.n + 5
    with pytest.raises(native.LsqtError):
        native.Hamiltonian.from_arrays(gpu_ctx, bad, m.n, m.max_neighbor, m.energy_max)


@pytest.mark.parametrize("layout", list(LAYOUTS))
def test_streaming_path_large(gpu_ctx, tmp_path, layout):
    """Systems above the streaming threshold (2^20 sites) take the
    non-temporal load/store kernels: still bit-identical states."""
    import oracle
    from gpuqt_amd import inputs as gi
    print(f"[gpu] streaming path 1.08M sites {layout}", flush=True)
    d = str(tmp_path)
    gi.write_honeycomb(d, 520, 520)
    gi.write_para(d, anderson=1.5, moments=24, energy_max=10.0, msd=True)
    gi.write_energy(d, [0.0])
    gi.write_time_steps(d, [2.0])
    m = oracle.OracleModel(d)
    assert m.n >= 1 << 20
    H = _ham(gpu_ctx, m, LAYOUTS[layout])
    r = m.random_state()
    v = H.vector(*r)
    mu = H.moments(v, v, 24)
    ref = m.moments(r, r, 24)
    assert_moments(mu, ref[0] + 1j * ref[1], "moments 1.08M")
    t = 2.0 * m.energy_max
    for fn, ofn in ((H.evolve, m.evolve), (H.evolvex, m.evolvex)):
        s = H.vector(*r)
        fn(1, t, s)
        got = s.download()
        want = ofn(1, t, *r)
        assert_bits(got[0], want[0], f"{fn.__name__} re")
        assert_bits(got[1], want[1], f"{fn.__name__} im")
    out = H.vector()
    H.apply_current(v, out)
    want = m.apply_current(*r)
    assert_bits(out.download()[0], want[0], "current")


def test_context_outlives_its_objects():
    """A context refuses destruction while a Hamiltonian, vector or block
    uses it (no handle may outlive its device stream); it goes once they are
    destroyed."""
    import oracle
    from gpuqt_amd import native
    lib = native.load()
    m = oracle.OracleModel(os.path.join(CASES_DIR, "honeycomb_small"))
    ctx = native.Context(0)
    H = native.Hamiltonian.from_arrays(ctx, m.arrays(), m.n, m.max_neighbor, m.energy_max)
    v = H.vector()
    b = native.Block(ctx, m.n, 4)
    for obj in (b, v, H):
        assert lib.lsqt_ctx_destroy(ctx.h) == 1  # LSQT_ERR_INVALID
        assert "still use this context" in lib.lsqt_last_error().decode()
        obj.close()
    assert lib.lsqt_ctx_destroy(ctx.h) == 0
    ctx.h = native._vp()
