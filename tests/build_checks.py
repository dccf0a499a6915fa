"""Builds the check programs under tests/ (tests/_bin/, git-ignored): device
programs with hipcc, host programs over liblps.so and the oracle with g++.
Called by the package build so that the GPU tests find them prebuilt."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
BIN = os.path.join(HERE, "_bin")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

CHECKS = {"div_rn_check": "div_rn_check.hip", "bf16_check": "bf16_check.hip"}
# host programs over the libraries: the C++ layer (include/lps_solver.hpp) against the oracle
HOST_CHECKS = {"cpp_solver_check": "cpp_solver_check.cpp"}


def path(name: str) -> str:
    return os.path.join(BIN, name)


def build(verbose: bool = False) -> None:
    os.makedirs(BIN, exist_ok=True)
    for name, src in CHECKS.items():
        cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-I", os.path.join(ROOT, "liquidsimulation_amd", "csrc"),
               "-o", path(name) + ".tmp", os.path.join(HERE, src)]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError(f"hipcc failed building {name}")
        os.replace(path(name) + ".tmp", path(name))
        if verbose:
            print(path(name))
    for name, src in HOST_CHECKS.items():
        lib_dir = os.path.join(ROOT, "liquidsimulation_amd")
        orc_dir = os.path.join(ROOT, "oracle")
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", "-ffp-contract=off",
               "-I", os.path.join(ROOT, "include"), "-I", orc_dir,
               "-o", path(name) + ".tmp", os.path.join(HERE, src),
               "-L", lib_dir, "-llps", "-L", orc_dir, "-lorc",
               "-Wl,-rpath,$ORIGIN/../../liquidsimulation_amd:$ORIGIN/../../oracle"]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError(f"g++ failed building {name}")
        os.replace(path(name) + ".tmp", path(name))
        if verbose:
            print(path(name))


if __name__ == "__main__":
    build(verbose=True)
