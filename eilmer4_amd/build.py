"""In-tree build of libe4hip.so for gfx950 with hipcc.

Translation units, compiled in parallel:
  stage_exact.hip   -ffp-contract=off: the FP64 ideal-gas kernels with the
                    reference's roundings (bit-identical to the CPU oracle)
  tp_exact.hip      -ffp-contract=off: the same for the thermally perfect mixture
  stage_fast64.hip  -ffp-contract=fast: the plain FP64 ideal-gas kernels
  stage_mixed.hip   -ffp-contract=fast: the mixed FP32/FP64 ideal-gas kernels
  tp_fast.hip       -ffp-contract=fast: the plain and mixed mixture kernels and
                    the learned-guess inference
  gen_exact.hip     -ffp-contract=off: blocks built from their grid (general
                    structured geometry), reference roundings; their CFL kernel
  gen_fast.hip      -ffp-contract=fast: the same in the plain and mixed forms
  e4hip_api.hip     -ffp-contract=off: C ABI, host bookkeeping, CFL/decode kernels
  mlp_train.hip     -ffp-contract=fast: training kernels of the temperature-guess network
Packed FP32 instructions are disabled in all of them (-packed-fp32-ops).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
OUT = os.path.join(HERE, "libe4hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", "-Wno-unused-result", "-Wno-unused-value"]
UNITS = [("stage_exact.hip", ["-ffp-contract=off"]),
         ("tp_exact.hip", ["-ffp-contract=off"]),
         ("stage_fast64.hip", ["-ffp-contract=fast"]),
         ("stage_mixed.hip", ["-ffp-contract=fast"]),
         ("tp_fast.hip", ["-ffp-contract=fast"]),
         ("gen_exact.hip", ["-ffp-contract=off"]),
         ("gen_fast.hip", ["-ffp-contract=fast"]),
         ("e4hip_api.hip", ["-ffp-contract=off"]),
         ("mlp_train.hip", ["-ffp-contract=fast"])]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("hipcc failed: %s" % " ".join(cmd))
    return r.stdout


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    objs, cmds = [], []
    for src, flags in UNITS:
        obj = os.path.join(OBJ, src.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([HIPCC] + COMMON + flags + ["-c", "-o", obj, os.path.join(CSRC, src)])
    with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", OUT + ".tmp"] + objs, verbose)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
