"""Build recipe for the oracle (test infrastructure only).

Compiles
  oracle/liboracle.so      -- the CPU restatement (vpmc_oracle.c), gcc;
  oracle/_ref/libvpmc_ref.so -- the reference's own CPU path, compiled from its
                              sources where they lie under /root/reference
                              (src/cpu/{hash,evaluate_vorticity,
                              monte_carlo_sampling,monte_carlo_bs}.cpp,
                              src/omp/hash_omp.cpp) together with this
                              directory's ref_harness.cpp, which only exposes
                              C entry points.  Skipped when the reference tree
                              is absent; the built .so travels with the repo.

Run:  python oracle/build.py
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("oracle build failed: " + " ".join(cmd))


def build_oracle():
    out = os.path.join(HERE, "liboracle.so")
    _run(["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
          "-o", out, os.path.join(HERE, "vpmc_oracle.c"), "-lm"])
    return out


def build_ref():
    if not os.path.isdir(os.path.join(REF, "src", "cpu")):
        return None
    outdir = os.path.join(HERE, "_ref")
    os.makedirs(outdir, exist_ok=True)
    out = os.path.join(outdir, "libvpmc_ref.so")
    srcs = [os.path.join(REF, "src", "cpu", f) for f in
            ("hash.cpp", "evaluate_vorticity.cpp", "monte_carlo_sampling.cpp",
             "monte_carlo_bs.cpp")]
    srcs.append(os.path.join(REF, "src", "omp", "hash_omp.cpp"))
    srcs.append(os.path.join(HERE, "ref_harness.cpp"))
    _run(["g++", "-O2", "-std=c++20", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
          "-w", "-I", os.path.join(REF, "include"), "-o", out] + srcs)
    return out


def build():
    a = build_oracle()
    b = build_ref()
    return a, b


if __name__ == "__main__":
    print(build())
