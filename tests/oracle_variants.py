"""Builds copies of the oracle source with one planted mistake (CPU only),
so a test can show that its checks fail on a plausible slip."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(ROOT, "oracle", "bqp_oracle.c")


def build_variant(tag: str, old: str, new: str, tmp: str) -> str:
    """Compiles the oracle with the single occurrence of `old` replaced by
    `new`; returns the shared library's path."""
    with open(SRC) as fh:
        src = fh.read()
    assert src.count(old) == 1, f"planting {tag}: pattern not found exactly once in the oracle source"
    path = os.path.join(tmp, f"oracle_{tag}.c")
    with open(path, "w") as fh:
        fh.write(src.replace(old, new))
    lib = os.path.join(tmp, f"liboracle_{tag}.so")
    subprocess.run(["gcc", "-O2", "-march=x86-64-v3", "-ffp-contract=off", "-fno-math-errno", "-fopenmp", "-fPIC",
                    "-std=gnu11", "-I", os.path.join(ROOT, "oracle"), "-shared", "-o", lib, path, "-lm"],
                   check=True)
    return lib
