"""bench.py's contract on the CPU: more than one GPU is refused with a
message, and the committed HBM-traffic profile holds the kernels whose
`traffic` the JSON line reports (the FP64 and FP32 four- and two-step
sweeps and one-step kernels of the default workload, the tuned block kernels
of the 160^3 cube and the uneven batch's block steps and sweeps)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""


def test_traffic_profile_entries():
    import bench
    fp64 = bench.profiled_traffic("k_hpsi<1,3,2,true,true,true>")
    # HBM bytes per launch within a few per cent of the algorithmic 72.1 B x
    # 8.0M (slot pattern: words of the 0.1 % irregular sites, 1 mask bit)
    assert fp64 is not None and abs(fp64 / (72.137 * 8_000_000) - 1.0) < 0.05
    assert bench.profiled_traffic("k_hpsi32<1,3,true,true,true>") is not None
    # the two- and six-step sweeps: 88 B (FP64) and 44 B (FP32) x 8.0M sites
    # per launch, within 7 %; the FP32 six-step sweep within 15 %: its four
    # workgroups per CU make runs of about 33 slabs, each of which re-reads
    # 2 x 5 slabs at its ends (mostly, not always, from L2)
    for name, per_site, tol in (("k_sweep2<double2,3,4,false>", 88.0, 0.07),
                                ("k_sweep2<float2,3,4,true>", 44.0, 0.07),
                                ("k_sweepn<double2,3,6,1,false,false,true>", 88.0, 0.07),
                                ("k_sweepn<float2,3,6,1,true,false,false>", 44.0, 0.15)):
        t = bench.profiled_traffic(name)
        assert t is not None and abs(t / (per_site * 8_000_000) - 1.0) < tol, name
    for mode in (1, 2):  # <L|psi> and product-identity block steps, sliding window
        t = bench.profiled_traffic(f"k_block_win<1,6,2,true,{mode}>")
        assert t is not None and t > 0
    # the uneven batch's pair and self-product steps (8 realisations, 8.97M
    # sites, 400 and 208 B per site) at their compulsory traffic within 7 %,
    # their two- and four-step sweeps (516 and 260 B per site and launch)
    # within 10 %: a chunk holds 24 of a row's 32 positions, and the 8 halo
    # positions both neighbouring tiles read come from L2 only where the two
    # tiles run close together in time
    for name, per_site, tol in (("k_block_step<1,3,2,true,3,false,false>", 400.0, 0.07),
                                ("k_block_step<1,3,2,true,2,false,false>", 208.0, 0.07),
                                ("k_block_sweep2<3,2,3,4>", 516.0, 0.10),
                                ("k_block_sweep2<3,2,2,4>", 260.0, 0.10),
                                ("k_block_sweepn<3,2,3,4,1,false>", 516.0, 0.10),
                                ("k_block_sweepn<3,2,2,4,1,false>", 260.0, 0.10),
                                ("k_block_sweepn<3,2,2,4,1,true>", 516.0, 0.10)):
        t = bench.profiled_traffic(name)
        assert t is not None and abs(t / (per_site * 8_973_112) - 1.0) < tol, name
    assert bench.profiled_traffic("no_such_kernel<1>") is None
