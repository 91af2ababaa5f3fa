"""The precision path (FP32-stored Chebyshev states, FP64 arithmetic and
accumulation) against the FP64 results of the reference (GPU).

Stated tolerances, relative to the FP64 rows of the reference (golden dumps):
  - DOS / VAC0 / LDOS rows (no time evolution): normwise <= 1e-6;
  - VAC / MSD / spin rows (time-evolved states): normwise <= 1e-5;
  - a single evolve / evolvex: |psi32 - psi64| <= 1e-6 |psi64|.
"""
import os

import numpy as np
import pytest

from conftest import CASES_DIR, golden_cases

pytestmark = pytest.mark.gpu


def _normwise(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def _setup(ctx, case):
    import oracle
    from gpuqt_amd import native
    d = os.path.join(CASES_DIR, case)
    g = np.load(os.path.join(d, "dump.npz"))
    m = oracle.OracleModel(d)
    H = native.Hamiltonian.from_arrays(ctx, m.arrays(), m.n, m.max_neighbor, m.energy_max)
    return g, m, H


def _row(m, mu):
    from gpuqt_amd import hostlib
    return hostlib.correlation_row(mu, m.energy, m.energy_max, m.volume)


@pytest.mark.parametrize("case", golden_cases())
def test_f32_drivers(gpu_ctx, case):
    print(f"[gpu] precision drivers {case}", flush=True)
    g, m, H = _setup(gpu_ctx, case)
    M, Nt = m.number_of_moments, m.number_of_steps_correlation
    worst = {}

    def check(tag, mu, tol):
        err = _normwise(_row(m, mu), g[f"{tag}_row"])
        worst[tag] = err
        assert err <= tol, f"{tag}: normwise {err:.2e} > {tol}"

    for i in range(int(g["meta"][3])):
        v = f"v{i}"
        r = m.random_state()
        rs = H.vector(*r)
        check(f"{v}_dos", H.moments_f32(rs, rs, M), 1e-6)
        if m.calculate_vac0:
            st = H.vector()
            H.apply_current(rs, st)
            check(f"{v}_vac0", H.moments_f32(st, st, M), 1e-6)
        if m.calculate_vac:
            left, left_copy, right = rs.clone(), H.vector(), H.vector()
            H.apply_current(left, right)
            for s in range(Nt):
                H.apply_current(left, left_copy)
                check(f"{v}_vac{s}", H.moments_f32(right, left_copy, M), 1e-5)
                if s < Nt - 1:
                    t = m.time_step[s] * m.energy_max
                    H.evolve_f32(-1, t, left)
                    H.evolve_f32(-1, t, right)
        if m.calculate_msd:
            state, state_x, copy = rs.clone(), rs.clone(), H.vector()
            t = m.time_step[0] * m.energy_max
            H.evolve_f32(1, t, state)
            H.evolvex_f32(1, t, state_x)
            for s in range(Nt):
                check(f"{v}_msd{s}", H.moments_f32(state_x, state_x, M), 1e-5)
                if s < Nt - 1:
                    t = m.time_step[s + 1] * m.energy_max
                    copy.copy_from(state)
                    H.evolvex_f32(1, t, copy)
                    H.evolve_f32(1, t, state_x)
                    state_x.add(copy)
                    H.evolve_f32(1, t, state)
        if m.calculate_spin:
            state, sz = rs.clone(), H.vector()
            for s in range(Nt):
                sz.apply_sz(state)
                check(f"{v}_spin{s}", H.moments_f32(state, sz, M), 1e-5)
                if s < Nt - 1:
                    H.evolve_f32(1, m.time_step[s] * m.energy_max, state)
    print(f"[gpu]   worst normwise error {max(worst.values()):.2e}", flush=True)


@pytest.mark.parametrize("case", ["honeycomb_small", "cube_peierls", "config1_honeycomb_20k"])
def test_f32_evolution(gpu_ctx, case):
    g, m, H = _setup(gpu_ctx, case)
    r = m.random_state()
    t = 2.0 * m.energy_max
    for name, f64, f32 in (("evolve", H.evolve, H.evolve_f32),
                           ("evolvex", H.evolvex, H.evolvex_f32)):
        a, b = H.vector(*r), H.vector(*r)
        f64(1, t, a)
        f32(1, t, b)
        x = np.array(a.download())
        y = np.array(b.download())
        err = np.abs(y - x).max() / np.abs(x).max()
        assert err <= 1e-6, f"{case} {name}: {err:.2e}"


def test_f32_renormalisation(gpu_ctx):
    """U is unitary: with renormalisation the FP64 state keeps its norm to
    FP64 rounding across many steps."""
    g, m, H = _setup(gpu_ctx, "config1_honeycomb_20k")
    r = m.random_state()
    s = H.vector(*r)
    n0 = np.linalg.norm(np.array(r))
    for _ in range(10):
        H.evolve_f32(1, 20.0, s)
    n1 = np.linalg.norm(np.array(s.download()))
    assert abs(n1 - n0) <= 1e-12 * n0


def test_f32_large_streaming(gpu_ctx, tmp_path):
    import oracle
    from gpuqt_amd import inputs as gi
    from gpuqt_amd import native
    d = str(tmp_path)
    gi.write_honeycomb(d, 520, 520)
    gi.write_para(d, anderson=1.5, moments=200, energy_max=10.0)
    gi.write_energy(d, gi.energy_grid(-8, 8, 41))
    m = oracle.OracleModel(d)
    H = native.Hamiltonian.from_arrays(gpu_ctx, m.arrays(), m.n, m.max_neighbor, m.energy_max)
    r = m.random_state()
    v = H.vector(*r)
    mu64 = H.moments(v, v, 200)
    mu32 = H.moments_f32(v, v, 200)
    err = _normwise(_row(m, mu32), _row(m, mu64))
    assert err <= 1e-6, f"normwise {err:.2e}"
