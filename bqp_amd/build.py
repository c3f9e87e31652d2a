"""Compile the gfx950 HIP library ``bqp_amd/libbqp.so`` in-tree with hipcc.

Flags: ``-ffp-contract=off`` so only the explicit fma() calls fuse (the CPU
restatement uses the same rule), and the packed-FP32 target feature disabled.
The second-order tableau constants (csrc/second_order_tables.h) are
regenerated by the host program csrc/gen_second_order.cpp first.
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
SRCS = [os.path.join(HERE, "csrc", f) for f in
        ("prop_sur.hip", "prop_sur_fp8.hip", "prop_encke.hip", "prop_fp64.hip", "prop_split.hip", "bqp.hip",
         "mlp.hip")]
OUT = os.path.join(HERE, "libbqp.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Wno-unused-value",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def gen_second_order() -> str:
    """Regenerate csrc/second_order_tables.h: the second-order tableau
    rearrangement (csrc/second_order.h), formed in long double on the host by
    csrc/gen_second_order.cpp and printed as hexadecimal literals."""
    src = os.path.join(HERE, "csrc", "gen_second_order.cpp")
    out = os.path.join(HERE, "csrc", "second_order_tables.h")
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "gen_second_order")
        subprocess.run([os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-o", exe, src], check=True)
        subprocess.run([exe, out], check=True)
    return out


def build(verbose: bool = False) -> str:
    """Compile libbqp.so (always recompiles); returns its path.  The
    tableau constants are regenerated first; the translation units compile
    in parallel, then one link."""
    gen_second_order()
    objs = [os.path.join(HERE, "csrc", os.path.basename(src) + ".o") for src in SRCS]
    compile_flags = [f for f in FLAGS if f != "-shared"]
    procs = [subprocess.Popen([HIPCC, *compile_flags, "-c", "-o", obj, src], stdout=subprocess.PIPE,
                              stderr=subprocess.PIPE, text=True) for src, obj in zip(SRCS, objs)]
    failed = []
    for src, p in zip(SRCS, procs):
        out, err = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out + err)
            failed.append(src)
        elif verbose:
            sys.stderr.write(err)
    if failed:
        raise RuntimeError(f"hipcc failed building {' '.join(failed)}")
    tmp = OUT + ".tmp"
    res = subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs],
                         capture_output=True, text=True)
    for obj in objs:
        try:
            os.remove(obj)
        except OSError:
            pass
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc link failed ({res.returncode})")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
