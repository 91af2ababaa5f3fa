"""Benchmark: Chebyshev site-updates/s of the LSQT hot path on one MI355X.

Workload (one "step"): the plain FP64 DOS + MSD job on an 8.0M-site
honeycomb lattice -- 2000 x 1000 four-orbital rectangular cells (a_cc = 1,
t = -2.7 eV, periodic), Anderson disorder U[-1, 1] eV, one random-phase
vector, 3000 Chebyshev moments, E_max = 10 eV, 201 energies in [-8, 8] eV,
MSD over 20 time steps of 2 hbar/eV (Bessel-expanded evolve / evolvex).
Per step: 21 moment sets + 59 evolutions, i.e. ~5.3e11 site-updates, all on
the GPU through the C ABI (lsqt_moments / lsqt_evolve / lsqt_evolvex), with
the host-side damping and Chebyshev summation of every row.  The
Hamiltonian and the random vector are resident in HBM before timing starts.

A site-update is one recursion step T_m = 2 H~ T_{m-1} - T_{m-2} at one site
(moment product included): N (M - 1) per moment set, N (1 + terms) per
evolve and 2 N (1 + terms) per evolvex.

Prints one JSON line (see the fields below).  --gpus N > 1 is refused: this
build runs on one GPU.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from gpuqt_amd import workloads as wl  # noqa: E402

HBM_PEAK_GBS = 8000.0
TRAFFIC_PROFILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                               "hbm_traffic_probe.md")


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` (template arguments without spaces)
    from the committed FETCH_SIZE/WRITE_SIZE profile of the default
    workloads, scaled by the plain-copy calibration; None if absent."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "traffic" and parts[1] == kernel:
                    return float(parts[2])
    except OSError:
        pass
    return None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


class HipEvents:
    """hipEvent timing on the context's stream (libamdhip64 via ctypes)."""

    def __init__(self, stream):
        self.hip = C.CDLL("libamdhip64.so")
        self.stream = C.c_void_p(stream)
        for f in ("hipEventCreate", "hipEventRecord", "hipEventSynchronize",
                  "hipEventElapsedTime", "hipEventDestroy"):
            getattr(self.hip, f).restype = C.c_int

    def new(self):
        ev = C.c_void_p()
        assert self.hip.hipEventCreate(C.byref(ev)) == 0
        return ev

    def record(self, ev):
        assert self.hip.hipEventRecord(ev, self.stream) == 0

    def elapsed_ms(self, a, b):
        assert self.hip.hipEventSynchronize(b) == 0
        ms = C.c_float()
        assert self.hip.hipEventElapsedTime(C.byref(ms), a, b) == 0
        return ms.value


def copy_bandwidth(ctx, n=1 << 27, reps=10):
    """GB/s of the library's plain copy kernel (k_copy: 16 B read and 16 B
    written per element) between two FP64 vectors of n elements (2.1 GB
    each), timed with hipEvents on the context's stream: what a pure stream
    kernel reaches on this GPU (one element per thread, no cache hints),
    beside which the fused step's rate reads."""
    from gpuqt_amd import native
    a, b = native.Vector(ctx, n), native.Vector(ctx, n)
    b.copy_from(a)
    ctx.sync()
    ev = HipEvents(ctx.stream)
    e0, e1 = ev.new(), ev.new()
    ev.record(e0)
    for _ in range(reps):
        b.copy_from(a)
    ev.record(e1)
    ms = ev.elapsed_ms(e0, e1) / reps
    for v in (a, b):
        v.close()
    return 32.0 * n / (ms * 1e-3) / 1e9


def run_job(H, model, r_vec, energies, time_steps, emax, volume, M, msd, ev=None, acc=None,
            precision="f64"):
    """One DOS (+ MSD) job on the device.  Returns (site_updates, rows).
    If ev is given, accumulates the device time of the moment sets in acc.
    precision "f32" runs the precision path (FP32-stored states)."""
    from gpuqt_amd import hostlib, native
    n = H.n
    f32 = precision == "f32"
    mom_fn = H.moments_f32 if f32 else H.moments
    evo_fn = H.evolve_f32 if f32 else H.evolve
    evx_fn = H.evolvex_f32 if f32 else H.evolvex
    rows = []
    updates = 0

    def moments(L, R):
        nonlocal updates
        if ev is not None:
            a, b = ev.new(), ev.new()
            ev.record(a)
        mu = mom_fn(L, R, M)
        if ev is not None:
            ev.record(b)
            acc["moment_ms"] += ev.elapsed_ms(a, b)
            acc["moment_sets"] += 1
        updates += n * (M - 1)
        return mu

    def row(mu):
        return hostlib.correlation_row(mu, energies, emax, volume)

    def evolve(direction, t, v):
        nonlocal updates
        evo_fn(direction, t, v)
        updates += n * (1 + terms(t))

    def evolvex(direction, t, v):
        nonlocal updates
        evx_fn(direction, t, v)
        updates += 2 * n * (1 + terms(t))

    rows.append(row(moments(r_vec, r_vec)))
    if msd:
        state, state_x, copy = r_vec.clone(), r_vec.clone(), native.Vector(H.ctx, n)
        t = time_steps[0] * emax
        evolve(1, t, state)
        evolvex(1, t, state_x)
        for s in range(len(time_steps)):
            rows.append(row(moments(state_x, state_x)))
            if s < len(time_steps) - 1:
                t = time_steps[s + 1] * emax
                copy.copy_from(state)
                evolvex(1, t, copy)
                evolve(1, t, state_x)
                state_x.add(copy)
                evolve(1, t, state)
        H.ctx.sync()
        for v in (state, state_x, copy):
            v.close()
    return updates, np.array(rows)


_terms_cache = {}


def terms(t):
    from gpuqt_amd import native
    if t not in _terms_cache:
        # chebyshev_2 launches = last term index - 1
        _terms_cache[t] = native.bessel_last_term(t) - 1
    return _terms_cache[t]


# block-step tile schedules (BlockShape.blocked in capi_block.hip)
SCHEDULES = ("natural", "blocked", "column-walk", "sliding-window")


def bench_block(ctx, L=160, R=32, M=2048, reps=2, check_vectors=4):
    """Config 4: block-FP32 KPM (R vectors per sweep) against the plain FP64
    single-vector kernel on the same system, and the DOS error of the block
    path against FP64."""
    from gpuqt_amd import hostlib, native
    from gpuqt_amd import inputs as gi
    t0 = time.perf_counter()
    arrays, n = gi.cubic_peierls_arrays(L)
    H = native.Hamiltonian.from_arrays(ctx, arrays, n, 6, 15.0)
    arrays_full = arrays
    del arrays
    info = H.info()
    log(f"cubic {L}^3: {n} sites, layout {'packed' if info.layout == 2 else 'full'}, "
        f"{info.classes} classes, built in {time.perf_counter() - t0:.1f} s")
    rng = np.random.default_rng(7)
    ph = rng.uniform(0, 2 * np.pi, (R, n))
    blk = native.Block(ctx, n, R)
    blk.upload(np.cos(ph), np.sin(ph))
    W, rows, blocked, us = H.block_tune(blk, blk)
    log(f"lattice slab {info.slab} row {info.row}; autotuner: W={W} (VPW={R // W}) "
        f"rows={rows} {SCHEDULES[blocked]} schedule: {us:.1f} us/step")
    v0 = H.vector(np.cos(ph[0]), np.sin(ph[0]))

    def timed(fn):
        fn()
        ctx.sync()
        best = 1e30
        out = None
        for _ in range(reps):
            t = time.perf_counter()
            out = fn()
            ctx.sync()
            best = min(best, time.perf_counter() - t)
        return best, out

    t64, mu64 = timed(lambda: H.moments(v0, v0, M))
    t32, mu32 = timed(lambda: H.moments_f32(v0, v0, M))
    # the FP64 kernel on the reference's data model (int32 index and FP64
    # re / im / xx per pair, no packing): 76 + 20 z bytes per site-update
    Hf = native.Hamiltonian.from_arrays(ctx, arrays_full, n, 6, 15.0, layout=native.LAYOUT_FULL)
    del arrays_full
    vf = Hf.vector(np.cos(ph[0]), np.sin(ph[0]))
    t64f, mu64f = timed(lambda: Hf.moments(vf, vf, M))
    assert np.abs(mu64f - mu64).max() <= 1e-10 * abs(mu64[0])
    bpu64f = Hf.info().bytes_per_site_f64
    vf.close()
    Hf.close()
    tb, mub = timed(lambda: H.moments_block(blk, blk, M))
    ts, mus = timed(lambda: H.moments_block_self(blk, M))
    # <L|T_m|R> with L != R (the VAC shape) by pair moment sets
    ph2 = rng.uniform(0, 2 * np.pi, (R, n))
    blk2 = native.Block(ctx, n, R)
    blk2.upload(np.cos(ph2), np.sin(ph2))
    v0b = H.vector(np.cos(ph2[0]), np.sin(ph2[0]))
    del ph2
    tp, mup = timed(lambda: H.moments_block_pair(blk, blk2, M))
    ref_pair = H.moments(v0, v0b, M)
    err_pair = float(np.abs(mup[0] - ref_pair).max() / np.abs(ref_pair).max())
    # the same error on the scale |L||R| = n of unit-modulus vectors (random
    # pairs have |<L|T_m|R>| ~ sqrt(n), which inflates the normwise figure)
    err_pair_n = float(np.abs(mup[0] - ref_pair).max() / n)
    e = np.linspace(-14, 14, 201)
    vol = float(L ** 3)
    errs, errs_self = [], []
    for r in range(check_vectors):
        ref = mu64 if r == 0 else H.moments(*(2 * [H.vector(np.cos(ph[r]), np.sin(ph[r]))]), M)
        b = hostlib.correlation_row(ref, e, 15.0, vol)
        for got, out in ((mub, errs), (mus, errs_self)):
            a = hostlib.correlation_row(got[r], e, 15.0, vol)
            out.append(float(np.abs(a - b).max() / np.abs(b).max()))
    u64 = n * (M - 1) / t64
    u64f = n * (M - 1) / t64f
    u32 = n * (M - 1) / t32
    ub = n * R * (M - 1) / tb
    # measured HBM bytes per launch, where the profile covers this workload
    profiled = L == 160 and R == 32
    db = "true" if getattr(H, "block_depth", 1) == 2 else "false"
    kname = ((lambda mode: f"k_block_win<1,6,{W},true,{mode}>") if blocked == 3 else
             (lambda mode: f"k_block_step<1,6,{W},true,{mode},false,{db}>"))
    traffic_left = profiled_traffic(kname(1)) if profiled else None
    traffic_self = profiled_traffic(kname(2)) if profiled else None
    steps_self = M // 2
    us_ = n * R * steps_self / ts          # recursion site-updates/s
    um_ = n * R * (M - 1) / ts             # moments x sites x vectors per second
    up_ = n * R * (M - 1) / tp             # the same for R pairs L != R
    bpu_blk = (4.0 + 4.0 * 6) / R + 32.0
    bpu64 = info.bytes_per_site_f64
    res = {
        "workload": f"simple cubic {L}^3 ({n} sites), hopping -1, Peierls phase 2pi*0.01*x on "
                    f"y-bonds, on-site U[-8.25,8.25], E_max=15, {M} moments, DOS",
        "f64_single": {"value": u64, "unit": "site-updates/s", "bytes_per_site_update": bpu64,
                       "achieved_GBs": u64 * bpu64 / 1e9,
                       "what": "the drop-in's FP64 single-vector kernel (packed layout): the "
                               "comparator of every speedup_vs_f64 below"},
        "f64_single_reference_layout": {
            "value": u64f, "unit": "site-updates/s", "bytes_per_site_update": bpu64f,
            "achieved_GBs": u64f * bpu64f / 1e9,
            "what": "the same FP64 kernel on the reference's data model (int32 index + FP64 "
                    "re/im/xx per pair, LSQT_LAYOUT_FULL)",
            "block_raw_speedup": None, "block_self_time_to_moments_speedup": None},
        "f32_single": {"value": u32, "unit": "site-updates/s", "speedup_vs_f64": u32 / u64},
        "block_f32": {"R": R, "W": W, "vectors_per_wavefront": R // W, "rows_per_tile": rows,
                      "tile_schedule": SCHEDULES[blocked],
                      "tile_rows": ("lattice rows" if blocked == 3 else "sites"),
                      "value": ub, "unit": "site-updates/s", "speedup_vs_f64": ub / u64,
                      "bytes_per_site_update": bpu_blk,
                      "achieved_GBs": ub * bpu_blk / 1e9,
                      "frac_of_hbm_peak": ub * bpu_blk / 1e9 / HBM_PEAK_GBS,
                      "bytes_per_launch_algorithmic": bpu_blk * n * R,
                      "traffic_per_launch": traffic_left,
                      "dos_normwise_error_vs_f64": errs},
        "block_f32_product_identities": {
            "R": R, "steps": steps_self, "moments": M,
            "value": us_, "unit": "site-updates/s",
            "bytes_per_site_update": (4.0 + 4.0 * 6) / R + 24.0,
            "achieved_GBs": us_ * ((4.0 + 4.0 * 6) / R + 24.0) / 1e9,
            "frac_of_hbm_peak": us_ * ((4.0 + 4.0 * 6) / R + 24.0) / 1e9 / HBM_PEAK_GBS,
            "traffic_per_launch": traffic_self,
            "moment_site_updates_per_s": um_,
            "time_to_moments_speedup_vs_f64": um_ / u64,
            "dos_normwise_error_vs_f64": errs_self},
        "block_f32_pair_products": {
            "R": R, "steps": steps_self, "moments": M,
            "what": "<L_r|T_m|R_r> for R independent pairs L_r != R_r (the VAC / spin shape)",
            "moment_site_updates_per_s": up_,
            "time_to_moments_speedup_vs_f64": up_ / u64,
            "moment_normwise_error_vs_f64": err_pair,
            "moment_error_vs_f64_over_norm_product": err_pair_n},
    }
    ref_layout = res["f64_single_reference_layout"]
    ref_layout["block_raw_speedup"] = us_ / u64f
    ref_layout["block_self_time_to_moments_speedup"] = um_ / u64f
    ref_layout["block_direct_speedup"] = ub / u64f
    log(f"FP64 single on the reference data model: {u64f:.3e} site-updates/s; block product "
        f"identities {us_ / u64f:.2f}x raw, {um_ / u64f:.2f}x time to moments")
    log(f"block R={R}: {ub:.3e} site-updates/s = {ub / u64:.2f}x FP64 single "
        f"({u64:.3e}); DOS error {max(errs):.2e}")
    log(f"block R={R} product identities: {us_:.3e} site-updates/s in {steps_self} steps, "
        f"{um_:.3e} moment site-updates/s = {um_ / u64:.2f}x FP64 single; "
        f"DOS error {max(errs_self):.2e}")
    log(f"block R={R} pairs L!=R: {up_:.3e} moment site-updates/s = {up_ / u64:.2f}x FP64 "
        f"single; moment error {err_pair:.2e} ({err_pair_n:.2e} of |L||R|)")
    for o in (H, blk, blk2, v0, v0b):
        o.close()
    return res


def fp64_vac_msd(H, re, im, M, ts):
    """find_vac and find_msd of one vector on the FP64 single-vector path
    (src/sigma.cu:278-393 over the C ABI): the moment sets of every step."""
    from gpuqt_amd import native
    n = H.n
    vac, msd = [], []
    left, right, copy = H.vector(re, im), native.Vector(H.ctx, n), native.Vector(H.ctx, n)
    H.apply_current(left, right)
    for m in range(len(ts)):
        H.apply_current(left, copy)
        vac.append(H.moments(right, copy, M))
        if m < len(ts) - 1:
            H.evolve(-1, ts[m], left)
            H.evolve(-1, ts[m], right)
    state, sx = H.vector(re, im), H.vector(re, im)
    H.evolve(1, ts[0], state)
    H.evolvex(1, ts[0], sx)
    for m in range(len(ts)):
        msd.append(H.moments(sx, sx, M))
        if m < len(ts) - 1:
            copy.copy_from(state)
            H.evolvex(1, ts[m + 1], copy)
            H.evolve(1, ts[m + 1], sx)
            sx.add(copy)
            H.evolve(1, ts[m + 1], state)
    for v in (left, right, copy, state, sx):
        v.close()
    return np.array(vac), np.array(msd)


class BatchMember:
    """One realisation of the uneven batch: its host Model (lattice kept for
    the device), lattice description, vacancy map, potential and R random
    vectors."""

    def __init__(self, d, R):
        from gpuqt_amd import hostlib
        self.model = hostlib.HostModel(d, lattice_on_device=True)
        self.n = self.model.n
        self.lattice, self.site_index = self.model.lattice()
        self.potential = self.model.potential()
        self.vecs = [self.model.random_state() for _ in range(R)]
        energies, _, _ = self.model.inputs()
        self.energies = energies
        self.volume = self.model.volume


def build_batch(realisations, R, M, n_steps, dt, emax, threads=16):
    """The first `realisations` honeycomb realisations of the uneven batch:
    host Models built in parallel with their lattices kept for the device
    (the neighbour lists are expanded on the GPU), side by side.  Returns
    (members, potential, starts, n, re, im, build seconds): the segmented
    Hamiltonian's potential and the vector-major random vectors of all."""
    import concurrent.futures as cf
    cells = wl.batch_cells(64)[:realisations]
    t0 = time.perf_counter()
    with tempfile.TemporaryDirectory() as tmp:
        dirs = []
        for i, (nx, ny) in enumerate(cells):
            d = os.path.join(tmp, f"r{i}")
            wl.write_batch_realisation(d, nx, ny, R=R, moments=M, n_steps=n_steps, dt=dt,
                                       emax=emax)
            dirs.append(d)
        devnull = os.open(os.devnull, os.O_WRONLY)
        saved = os.dup(1)
        os.dup2(devnull, 1)
        try:
            with cf.ThreadPoolExecutor(max(1, min(threads, len(dirs)))) as pool:
                members = list(pool.map(lambda d: BatchMember(d, R), dirs))
        finally:
            os.dup2(saved, 1)
            os.close(devnull)
            os.close(saved)
    sizes = [b.n for b in members]
    bound = max(wl.lattice_gershgorin_bound(b.lattice, b.potential) for b in members)
    if not bound < emax:
        raise RuntimeError(f"energy_max {emax} does not exceed the spectral bound {bound:.3f} "
                           f"of the batch")
    starts = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    n = int(starts[-1])
    potential = np.concatenate([b.potential for b in members])
    re = np.concatenate([np.array([v[0] for v in b.vecs]) for b in members], axis=1)
    im = np.concatenate([np.array([v[1] for v in b.vecs]) for b in members], axis=1)
    build_s = time.perf_counter() - t0
    log(f"batch: {realisations} realisations, {min(sizes)}..{max(sizes)} sites, {n} in all, "
        f"host models in {build_s:.1f} s; spectral bound {bound:.2f} < E_max {emax}")
    return members, potential, starts, n, re, im, build_s


def batch_hamiltonian(ctx, members, potential, emax):
    """The segmented Hamiltonian of the batch, expanded on the device."""
    from gpuqt_amd import native
    return native.Hamiltonian.from_lattices(ctx, [b.lattice for b in members],
                                            [b.site_index for b in members], potential, emax)


def bench_batch(ctx, realisations=64, R=8, M=1500, n_steps=16, dt=2.0, emax=wl.BATCH_EMAX,
                threads=16, check=True):
    """Uneven batch: honeycomb realisations of uneven size (a_cc = 1.42 A,
    charged impurities at 0.1 %, amplitudes U[-2.8, 2.8] eV, xi = 4.26 A),
    R random vectors each, VAC + MSD over n_steps steps of dt, all
    realisations side by side in one segmented Hamiltonian: one launch per
    recursion step for the whole batch, moment sets replayed as hipGraphs."""
    from gpuqt_amd import hostlib, native
    members, potential, starts, n, re, im, build_s = build_batch(realisations, R, M, n_steps,
                                                                 dt, emax, threads)
    sizes = [b.n for b in members]
    t0 = time.perf_counter()
    H = batch_hamiltonian(ctx, members, potential, emax)
    ctx.sync()
    ham_s = time.perf_counter() - t0
    log(f"batch: segmented Hamiltonian expanded on the device in {ham_s:.2f} s")
    del potential
    ts = np.full(n_steps, dt * emax)
    # untimed: tuning and graph capture on a short run of the same shape
    H.block_vac(re, im, M, ts[:1])
    H.block_msd(re, im, M, ts[:1])
    ctx.sync()
    log(f"batch: warm-up (one time step of VAC and MSD) done; timing {n_steps} steps")
    t = time.perf_counter()
    vac = H.block_vac(re, im, M, ts)
    t_vac = time.perf_counter() - t
    log(f"batch: VAC over {n_steps} steps in {t_vac:.1f} s")
    t = time.perf_counter()
    msd = H.block_msd(re, im, M, ts)
    t_msd = time.perf_counter() - t
    log(f"batch: MSD over {n_steps} steps in {t_msd:.1f} s")
    elapsed = t_vac + t_msd
    K_t = terms(dt * emax)
    # recursion site-updates per vector and site: VAC (M - 1) per step, MSD
    # M / 2 per step (product identities); evolve 1 + K, evolvex 2 (1 + K)
    per_site = (n_steps * (M - 1) + 2 * (n_steps - 1) * (1 + K_t)
                + n_steps * (M // 2) + 3 * (1 + K_t) + (n_steps - 1) * 4 * (1 + K_t))
    updates = float(n) * R * per_site
    moment_updates = float(n) * R * 2 * n_steps * (M - 1)
    log(f"batch VAC {t_vac:.1f} s + MSD {t_msd:.1f} s: {updates / elapsed:.3e} site-updates/s, "
        f"{moment_updates / elapsed:.3e} moment site-updates/s")
    res = {
        "workload": f"{realisations} honeycomb realisations of the 64-realisation uneven batch "
                    f"(cells U[200,1500]^2 from seed 5; {min(sizes)}..{max(sizes)} sites, {n} "
                    f"in all), charged impurities 0.1 % U[-2.8,2.8] eV xi=4.26 A, {R} vectors "
                    f"each, {M} moments, VAC + MSD over {n_steps} steps of dt={dt}, E_max={emax}",
        "realisations": realisations, "sites": n, "R": R,
        "value": updates / elapsed, "unit": "site-updates/s",
        "moment_site_updates_per_s": moment_updates / elapsed,
        "seconds": {"vac": t_vac, "msd": t_msd, "host_models": build_s,
                    "hamiltonian_on_device": ham_s},
        "launches": "one per recursion step for all realisations (segmented Hamiltonian), "
                    "four steps per launch in the moment sets (k_block_sweepn: every "
                    "realisation is a periodic honeycomb); moment sets replayed from captured "
                    "hipGraphs",
        "block_sweep_steps": H.info().block_sweep_steps,
        "block_sweep_systems": H.info().block_sweep_systems,
    }
    # counted HBM traffic of the batch's kernels over their compulsory bytes
    # (profiled on the first 8 realisations, bench.py --probe)
    n8 = sum(b.n for b in members[:8]) if len(members) >= 8 else None
    if n8 == 8973112 and R == 8:
        ratios = {}
        for kname, per_site in (("k_block_sweepn<3,2,3,4,1,false>", 516.0),
                                ("k_block_sweepn<3,2,2,4,1,false>", 260.0),
                                ("k_block_sweepn<3,2,2,4,1,true>", 516.0),
                                ("k_block_sweep2<3,2,3,4>", 516.0),
                                ("k_block_sweep2<3,2,2,4>", 260.0),
                                ("k_block_step<1,3,2,true,3,false,false>", 400.0),
                                ("k_block_step<1,3,2,true,2,false,false>", 208.0),
                                ("k_block_evo<1,3,2,1,true>", 336.0),
                                ("k_block_evo<1,3,2,4,true>", 528.0)):
            t = profiled_traffic(kname)
            if t:
                ratios[kname] = t / (per_site * n8)
        res["traffic_ratio_to_algorithmic"] = ratios
    if check:
        # vector 0 of every realisation against the FP64 single-vector path
        t_chk = time.perf_counter()
        per = []
        for k, b in enumerate(members):
            H1 = native.Hamiltonian.from_lattices(ctx, [b.lattice], [b.site_index], b.potential,
                                                  emax)
            v64, m64 = fp64_vac_msd(H1, b.vecs[0][0], b.vecs[0][1], M, ts)
            H1.close()
            row = lambda mu: hostlib.correlation_row(mu, b.energies, emax, b.volume)
            e = {"realisation": k, "sites": b.n}
            for name, got, want in (("vac", vac[:, k, 0], v64), ("msd", msd[:, k, 0], m64)):
                g = np.array([row(x) for x in got])
                w = np.array([row(x) for x in want])
                e[f"{name}_rows"] = float(
                    max(np.abs(a_ - b_).max() / np.abs(b_).max() for a_, b_ in zip(g, w)))
                sg = hostlib.conductivity(ts / emax, **{name: g})[f"sigma_{name}"]
                sw = hostlib.conductivity(ts / emax, **{name: w})[f"sigma_{name}"]
                e[f"sigma_{name}"] = float(
                    max(np.abs(a_ - b_).max() / np.abs(b_).max()
                        for a_, b_ in zip(sg, sw) if np.abs(b_).max() > 0))
            per.append(e)
            if k % 8 == 7 or k == len(members) - 1:
                log(f"batch FP64 check: {k + 1} of {len(members)} realisations")
        keys = ("vac_rows", "msd_rows", "sigma_vac", "sigma_msd")
        res["error_vs_f64"] = {
            "what": "normwise error of every VAC / MSD row and sigma_VAC / sigma_MSD row of "
                    "random vector 0 of every realisation against the FP64 single-vector "
                    "drivers on the same realisation",
            **{f"{key}_normwise_max": max(e[key] for e in per) for key in keys},
            "per_realisation": per,
        }
        log(f"batch errors vs FP64 (vector 0 of all {len(per)} realisations, "
            f"{time.perf_counter() - t_chk:.1f} s): "
            + ", ".join(f"{key} {res['error_vs_f64'][key + '_normwise_max']:.2e}"
                        for key in keys))
    H.close()
    return res


probe_arrays = None


def probe(ctx, H, rv, args, moments=10):
    """Short runs of the hot kernels for counter profiles: a plain copy of
    one vector (k_copy: n x 16 B read and written, the counters' scale), the
    FP64 and FP32 fused moment steps on the honeycomb workload, and the block
    steps on the simple-cubic one."""
    from gpuqt_amd import inputs as gi
    from gpuqt_amd import native
    w = H.vector()
    w.copy_from(rv)
    ctx.sync()
    log(f"probe: copy of {H.n} x 16 B")
    H.moments(rv, rv, moments)
    H.moments_f32(rv, rv, moments)
    ctx.sync()
    log(f"probe: {moments} FP64 and FP32 moments on {H.n} sites")
    if H.info().sweep_slab:
        # the one-step kernels on the same system, for comparison
        os.environ["LSQT_SWEEP"] = "0"
        H1 = native.Hamiltonian.from_arrays(ctx, probe_arrays, H.n, H.info().max_neighbor,
                                            H.info().energy_max)
        del os.environ["LSQT_SWEEP"]
        v1 = H1.vector()
        v1.copy_from(rv)
        H1.moments(v1, v1, moments)
        H1.moments_f32(v1, v1, moments)
        ctx.sync()
        log(f"probe: the same with one step per launch")
        v1.close()
        H1.close()
    if args.block:
        arrays, n = gi.cubic_peierls_arrays(args.cubic_L)
        Hc = native.Hamiltonian.from_arrays(ctx, arrays, n, 6, 15.0)
        del arrays
        R = args.block_R
        ph = np.random.default_rng(7).uniform(0, 2 * np.pi, (R, n))
        blk = native.Block(ctx, n, R)
        blk.upload(np.cos(ph), np.sin(ph))
        W, rows, blocked, _ = Hc.block_tune(blk, blk)
        # the tuned shape, and the sliding window and the column walk at both
        # loop depths with W = 2, whichever the tuner picks
        shapes = [(W, rows, blocked, getattr(Hc, "block_depth", 1))]
        info = Hc.info()
        if info.slab and info.row:
            ny = info.slab // info.row
            by = max(d for d in range(1, 33) if ny % d == 0)
            shapes.append((2, by, 3, 1))
            span = next((s for s in range(info.row, info.slab + 1, info.row)
                         if info.slab % s == 0 and s % 64 == 0 and s >= 512), 0)
            if span:
                shapes += [(2, span, 2, 1), (2, span, 2, 2)]
        for (w, r, b, depth) in shapes:
            try:
                Hc.block_set_config(R, w, r, blocked=b, depth=depth)
            except native.LsqtError:
                continue
            Hc.moments_block(blk, blk, moments)
            Hc.moments_block_self(blk, 2 * moments)
        Hc.block_set_config(R, 0)
        ctx.sync()
        log(f"probe: block R={R} moments on {n} sites")
    if args.batch:
        # the uneven batch's kernels (pair and self block steps, evolution
        # terms) on its first 8 realisations, one time step of VAC and MSD
        members, potential, starts, n, re, im, _ = build_batch(8, 8, moments, 1, wl.BATCH_DT,
                                                               wl.BATCH_EMAX)
        Hb = batch_hamiltonian(ctx, members, potential, wl.BATCH_EMAX)
        ts = np.full(1, wl.BATCH_DT * wl.BATCH_EMAX)
        Hb.block_vac(re, im, moments, ts)
        Hb.block_msd(re, im, moments, ts)
        ctx.sync()
        log(f"probe: uneven batch VAC + MSD ({moments} moments, one step) on {n} sites")
        Hb.close()


def host_cpu():
    """The host CPU model and the threads this process may use."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None
    return {"model": model, "logical_cpus_usable": usable}


def reference_provenance(binary):
    """Where the reference program came from: the build record written next
    to it by the oracle's recipe, and the binary's sha256."""
    import hashlib
    info = {"binary": os.path.relpath(binary, ROOT)}
    with open(binary, "rb") as f:
        info["sha256"] = hashlib.sha256(f.read()).hexdigest()
    rec = os.path.join(os.path.dirname(binary), "buildinfo.txt")
    try:
        with open(rec) as f:
            for line in f:
                key, _, val = line.strip().partition(" ")
                if key:
                    info[key] = val
    except OSError:
        info["build_record"] = "absent"
    return info


def cpu_baseline(nx, ny, moments_sample, threads):
    """Times the reference CPU program (compiled from the reference's sources
    into oracle/_ref by `make -C oracle ref`) on a bounded sample of the same
    workload: the DOS moment set of the same lattice with fewer moments, on
    one core and as `threads` concurrent processes.  Where that program is
    absent (a tree built without the reference sources) the oracle's C
    restatement is timed instead, with kind "port" and a note saying so.
    Returns the cpu_baseline dict."""
    import oracle
    with tempfile.TemporaryDirectory() as tmp:
        d = os.path.join(tmp, "cpu")
        os.makedirs(d)
        wl.write_config2(d, nx=nx, ny=ny, moments=moments_sample, msd=False)
        n_sites = 4 * nx * ny
        updates = n_sites * (moments_sample - 1)
        ref = oracle.reference_binary()
        parallel = None
        if ref is not None:
            driver = os.path.join(tmp, "input.txt")
            with open(driver, "w") as f:
                f.write(d + "\n")
            r = subprocess.run([ref, driver], capture_output=True, text=True, timeout=600)
            m = re.search(r"Time used for finding DOS = ([0-9.eE+-]+) s", r.stdout)
            if r.returncode != 0 or not m:
                raise RuntimeError("reference CPU run failed")
            secs = float(m.group(1))
            kind, cores = "reference", 1
            parallel = parallel_reference(ref, tmp, nx, ny, moments_sample, threads, updates)
        else:
            oracle.set_threads(threads)
            om = oracle.OracleModel(d)
            rr = om.random_state()
            t0 = time.perf_counter()
            om.moments(rr, rr, moments_sample)
            secs = time.perf_counter() - t0
            kind, cores = "port", oracle.num_threads()
    out = {"value": updates / secs, "unit": "site-updates/s", "cores": cores, "kind": kind,
           "sample": f"DOS moment set, {moments_sample} moments, {n_sites}-site honeycomb "
                     f"(same lattice as the workload), {secs:.2f} s",
           "host_cpu": host_cpu()}
    if kind == "reference":
        out["program"] = reference_provenance(ref)
        out["timer"] = "the program's own DOS phase timer (model build and I/O excluded)"
        if parallel:
            out["parallel_processes"] = parallel
    else:
        out["note"] = ("the reference program is not built in this tree (oracle/_ref/lsqt_cpu; "
                       "make -C oracle ref needs the reference sources): the oracle's C "
                       f"restatement of the same recursion was timed on {cores} OpenMP threads")
    return out


def parallel_reference(ref, tmp, nx, ny, moments_sample, procs, updates):
    """P copies of the single-threaded reference program at once, each on
    its own copy of the job (the reference has no threads; independent
    vectors or realisations are how it uses P cores).  The DOS phase time is
    the wall time of the whole concurrent run minus that of the same runs
    with 2 moments (model build, I/O)."""
    procs = max(1, min(procs, os.cpu_count() or 1))

    def wall(moments):
        drivers = []
        for p in range(procs):
            d = os.path.join(tmp, f"par{moments}_{p}")
            os.makedirs(d)
            wl.write_config2(d, nx=nx, ny=ny, moments=moments, msd=False)
            drv = os.path.join(tmp, f"par{moments}_{p}.txt")
            with open(drv, "w") as f:
                f.write(d + "\n")
            drivers.append(drv)
        t0 = time.perf_counter()
        running = [subprocess.Popen([ref, drv], stdout=subprocess.DEVNULL,
                                    stderr=subprocess.DEVNULL) for drv in drivers]
        codes = [r.wait(timeout=900) for r in running]
        if any(codes):
            raise RuntimeError("parallel reference run failed")
        return time.perf_counter() - t0

    setup = wall(2)
    total = wall(moments_sample)
    secs = max(total - setup, 1e-9)
    return {"value": procs * updates / secs, "unit": "site-updates/s", "processes": procs,
            "sample": f"{procs} concurrent reference runs of the same DOS moment set; "
                      f"{total:.1f} s wall, {setup:.1f} s of it model build and I/O"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nx", type=int, default=2000)
    ap.add_argument("--ny", type=int, default=1000)
    ap.add_argument("--moments", type=int, default=3000)
    ap.add_argument("--time-steps", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-f32", dest="precision_f32", action="store_false",
                    help="skip the precision-path (FP32 states) job")
    ap.add_argument("--f32-steps", type=int, default=5,
                    help="timed steps of the precision-path job (its own count)")
    ap.add_argument("--f32-warmup", type=int, default=1)
    ap.add_argument("--no-block", dest="block", action="store_false",
                    help="skip the block-path (R vectors per sweep) measurement")
    ap.add_argument("--cubic-L", type=int, default=160)
    ap.add_argument("--block-R", type=int, default=32)
    ap.add_argument("--block-moments", type=int, default=2048)
    ap.add_argument("--cpu-moments", type=int, default=250)
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--no-batch", dest="batch", action="store_false",
                    help="skip the uneven-batch (VAC + MSD over many realisations) section")
    ap.add_argument("--batch-realisations", type=int, default=64)
    ap.add_argument("--batch-moments", type=int, default=1500)
    ap.add_argument("--batch-steps", type=int, default=16)
    ap.add_argument("--probe", action="store_true",
                    help="run each hot kernel a few times on the workloads' inputs (for "
                         "hardware-counter profiles) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): the LSQT path "
              f"here has no multi-GPU mode", file=sys.stderr)
        sys.exit(2)

    from gpuqt_amd import hostlib, native

    with tempfile.TemporaryDirectory() as tmp:
        d = os.path.join(tmp, "config2")
        os.makedirs(d)
        wl.write_config2(d, nx=args.nx, ny=args.ny, moments=args.moments,
                      time_steps=args.time_steps)
        t0 = time.perf_counter()
        # the drop-in's Model (host C++): lattice, Anderson draws, random phases
        devnull = os.open(os.devnull, os.O_WRONLY)
        saved = os.dup(1)
        os.dup2(devnull, 1)
        try:
            model = hostlib.HostModel(d)
            r = model.random_state()
            arrays = model.arrays()
        finally:
            os.dup2(saved, 1)
            os.close(devnull)
            os.close(saved)
        energies, time_steps, _ = model.inputs()
        build_s = time.perf_counter() - t0
        log(f"model: {model.n} sites built in {build_s:.1f} s")

    ctx = native.Context(0)
    H = native.Hamiltonian.from_arrays(ctx, arrays, model.n, model.max_neighbor,
                                       model.energy_max)
    if args.probe:
        global probe_arrays
        probe_arrays = arrays
    del arrays
    info = H.info()
    rv = H.vector(*r)
    ctx.sync()
    if args.probe:
        probe(ctx, H, rv, args)
        return
    log(f"Hamiltonian on device: layout {'packed' if info.layout == 2 else 'full'}, "
        f"{info.classes} hopping classes, {info.bytes_per_site_f64:.0f} B per site-update")

    M = args.moments
    job = dict(energies=energies, time_steps=time_steps, emax=model.energy_max,
               volume=model.volume, M=M, msd=args.time_steps > 0)

    sweep = info.sweep_slab > 0
    # steps per launch of the moment sets' dominant kernel: four (k_sweepn)
    # or two (k_sweep2) on a periodic lattice, else one (k_hpsi)
    steps64 = max(info.sweep_steps, 1)
    steps32 = max(info.sweep_steps_f32, 1)

    def launches_per_set(steps):
        # the moment set's launches of its dominant kernel: M - 1 one-step
        # launches, (M - 2) // 2 two-step launches, or (M - 4) // steps
        # launches of more steps after the two-step one at moment 2 (the set's
        # other launches -- mu_0, mu_1, the remainder, the reduction -- are
        # counted into their time, which overstates it by under 0.5 %)
        if steps >= 3 and M >= 4 + steps:
            return (M - 4) // steps
        if steps >= 2 and M >= 4:
            return (M - 2) // 2
        return M - 1

    def timed(precision, steps, warmup):
        for w in range(warmup):
            run_job(H, model, rv, precision=precision, **job)
            log(f"{precision} warmup {w} done")
        ev = HipEvents(ctx.stream)
        acc = {"moment_ms": 0.0, "moment_sets": 0}
        ctx.sync()
        t0 = time.perf_counter()
        total = 0
        rows = None
        for s in range(steps):
            u, rows = run_job(H, model, rv, ev=ev, acc=acc, precision=precision, **job)
            total += u
            log(f"{precision} step {s} done")
        ctx.sync()
        elapsed = time.perf_counter() - t0
        per_set = launches_per_set(steps32 if precision == "f32" else steps64)
        avg_launch_ms = acc["moment_ms"] / max(acc["moment_sets"] * per_set, 1)
        return total, elapsed, avg_launch_ms, rows

    total_updates, elapsed, avg_launch_ms, rows = timed("f64", args.steps, args.warmup)
    rows = np.array(rows)
    value = total_updates / elapsed
    # dominant kernel: the multi-step sweep (steps64 site-updates per site
    # per launch) where the lattice has one, else the fused one-step kernel
    updates_per_site = steps64
    bytes_per_site = info.bytes_per_site_sweep_f64 if sweep else info.bytes_per_site_f64
    bytes_per_launch = bytes_per_site * model.n
    achieved = bytes_per_launch / (avg_launch_ms * 1e-3) / 1e9
    default_workload = (args.nx, args.ny) == (2000, 1000) and info.layout == 2 \
        and info.max_neighbor == 3
    kname = {6: "k_sweepn<double2,3,6,1,false,false,true>", 2: "k_sweep2<double2,3,4,false>"}.get(
        steps64, "k_hpsi<1,3,2,true,true,true>")
    kname32 = {6: "k_sweepn<float2,3,6,1,true,false,false>", 2: "k_sweep2<float2,3,4,true>"}.get(
        steps32, "k_hpsi32")
    traffic = profiled_traffic(kname) if default_workload else None
    # the same launch time priced on other byte counts: the one-step kernel's
    # (what it moves for the same site-updates on this layout) and the
    # reference's data model (int32 index + FP64 re / im per pair, separate
    # inner-product pass: 76 + 20 z + 16 bytes per site-update)
    one_step_GBs = info.bytes_per_site_f64 * updates_per_site * model.n / (avg_launch_ms * 1e-3) / 1e9
    ref_model_bpu = 76.0 + 20.0 * info.max_neighbor + 16.0
    ref_model_GBs = ref_model_bpu * updates_per_site * model.n / (avg_launch_ms * 1e-3) / 1e9
    stream = copy_bandwidth(ctx)
    log(f"plain copy: {stream:.0f} GB/s; fused step {achieved:.0f} GB/s")
    dos = rows[0]
    log(f"DOS(E=0) = {dos[len(dos) // 2]:.6g}, {len(rows)} rows")

    f32 = None
    if args.precision_f32:
        u32, e32, l32, rows32 = timed("f32", args.f32_steps, args.f32_warmup)
        # FP32 path: the sweep moves potential 4 + five vector streams x 8 per
        # site and launch; the one-step kernel potential 4 + 4 x 8 + the same
        # pair bytes as FP64 (which has potential 8 + 4 x 16)
        bpu32 = (info.bytes_per_site_sweep_f32 if sweep
                 else 36.0 + (info.bytes_per_site_f64 - 72.0))
        ach32 = bpu32 * model.n / (l32 * 1e-3) / 1e9
        errs = [float(np.abs(a - b).max() / np.abs(b).max()) for a, b in zip(rows32, rows)]
        elem = []
        for a, b in zip(rows32, rows):
            mask = np.abs(b) > 1e-3 * np.abs(b).max()
            elem.append(float((np.abs(a - b)[mask] / np.abs(b)[mask]).max()))
        sigma_err = None
        if len(rows) > 1:
            s64 = hostlib.conductivity(time_steps, msd=rows[1:])["sigma_msd"]
            s32 = hostlib.conductivity(time_steps, msd=np.array(rows32)[1:])["sigma_msd"]
            sigma_err = float(max(np.abs(a - b).max() / np.abs(b).max()
                                  for a, b in zip(s32, s64) if np.abs(b).max() > 0))
        f32 = {
            "value": u32 / e32,
            "unit": "site-updates/s",
            "ms_per_step": e32 / args.f32_steps * 1e3,
            "steps": args.f32_steps,
            "warmup": args.f32_warmup,
            "speedup_vs_f64": (u32 / e32) / value,
            "dtype": "f32 states, f64 arithmetic and accumulation",
            "roofline": {"bound": "hbm",
                         "kernel": kname32 + (f" ({steps32} steps per launch)" if sweep else ""),
                         "achieved": ach32, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": ach32 / HBM_PEAK_GBS,
                         "traffic": (profiled_traffic(kname32)
                                     if sweep and default_workload else None),
                         "bytes_per_site_per_launch": bpu32,
                         "site_updates_per_site_per_launch": steps32,
                         "bytes_per_launch_algorithmic": bpu32 * model.n,
                         "avg_launch_us": l32 * 1e3},
            "error_vs_f64": {
                "dos_normwise": errs[0],
                "msd_normwise_max": max(errs[1:]) if len(errs) > 1 else None,
                "all_rows_normwise": errs,
                "elementwise_max_above_1e-3_of_row_max": max(elem),
                "sigma_msd_normwise_max": sigma_err,
            },
        }
        log(f"f32: {u32 / e32:.3e} site-updates/s, DOS normwise error {errs[0]:.2e}, "
            f"MSD max {max(errs):.2e}")

    result = {
        "metric": "Chebyshev site-updates/s + HBM GB/s; block-FP32 vs plain FP64 at stated "
                  "error",
        "metric_detail": "value: FP64 KPM + Bessel evolution site-updates/s of the whole job; "
                         "roofline: HBM GB/s of the fused recursion kernel; precision_f32 and "
                         "batching_block_f32: the FP32 and block-FP32 paths against plain FP64 "
                         "with their errors",
        "value": value,
        "unit": "site-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded lattice, Anderson disorder and random phases of the "
                "reference's RNG)",
        "config": {
            "workload": f"honeycomb {args.nx}x{args.ny} four-orbital cells "
                        f"({model.n} sites), t=-2.7 eV, Anderson U[-1,1] eV, 1 random vector, "
                        f"{M} moments, E_max=10 eV, 201 energies, DOS + MSD over "
                        f"{args.time_steps} steps of dt=2 hbar/eV",
            "sites": model.n,
            "moments": M,
            "site_updates_per_step": total_updates // args.steps,
            "layout": "packed" if info.layout == 2 else "full",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": (kname + f" ({steps64} Chebyshev steps per launch: psi_m .. "
                       f"psi_{{m+{steps64 - 1}}} from psi_{{m-2}}, psi_{{m-1}}, each level one "
                       "slab behind the one before in LDS rings, the last two written, fused "
                       "with <L|psi> of every level)"
                       if steps64 > 2 else
                       kname + " (two Chebyshev steps per launch: psi_m and psi_{m+1} from "
                       "psi_{m-2}, psi_{m-1}, four slabs at a time from LDS, fused with "
                       "<L|psi_m> and <L|psi_{m+1}>)" if sweep else
                       "k_hpsi (2 H~ psi_{m-1} - psi_{m-2} fused with <L|psi_m>; slot words "
                       "computed from the lattice pattern, 2 sites per thread)"),
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": "rocprofv3 FETCH_SIZE + WRITE_SIZE of this kernel on this workload, "
                              "calibrated on a plain copy (profiles/hbm_traffic_probe.md)"
                              if traffic else None,
            "bytes_per_launch_algorithmic": bytes_per_launch,
            "bytes_per_site_per_launch": bytes_per_site,
            "site_updates_per_site_per_launch": updates_per_site,
            "avg_launch_us": avg_launch_ms * 1e3,
            "plain_copy_GBs": stream,
            "ratio_to_plain_copy": achieved / stream,
            "one_step_equivalent_GBs": one_step_GBs,
            "one_step_equivalent_note": "the same launch time priced at the one-step kernel's "
                                        f"{info.bytes_per_site_f64:.2f} B per site-update: "
                                        "what a one-step kernel would have to stream",
            "reference_data_model_GBs": ref_model_GBs,
            "reference_data_model_note": f"priced at {ref_model_bpu:.0f} B per site-update "
                                         "(int32 index + FP64 re / im per pair and a separate "
                                         "inner-product pass); not a roofline: above the peak "
                                         "where the layout moves fewer bytes",
        },
    }
    if f32 is not None:
        result["precision_f32"] = f32
    if args.block:
        result["batching_block_f32"] = bench_block(ctx, L=args.cubic_L, R=args.block_R,
                                                   M=args.block_moments)
    if args.batch:
        result["batching_uneven"] = bench_batch(ctx, realisations=args.batch_realisations,
                                                M=args.batch_moments,
                                                n_steps=args.batch_steps)
    if not args.no_cpu_baseline:
        log("CPU baseline ...")
        result["cpu_baseline"] = cpu_baseline(args.nx, args.ny, args.cpu_moments,
                                              args.cpu_threads)
    rv.close()
    H.close()
    ctx.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
