"""Builds tests/golden/rng_draws.cpp with the host g++ and stores its output
as tests/golden/rng_draws.json (draws of std::mt19937 and libstdc++'s
uniform_real / uniform_int distributions)."""
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))

with tempfile.TemporaryDirectory() as tmp:
    exe = os.path.join(tmp, "rng_draws")
    subprocess.run(["g++", "-O2", "-std=c++11", "-o", exe, os.path.join(HERE, "rng_draws.cpp")],
                   check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
with open(os.path.join(HERE, "rng_draws.json"), "w") as f:
    f.write(out)
print("wrote rng_draws.json")
