"""Builds libmuspin_hip.so in-tree with hipcc for gfx950.

Usage: python -m muspinsim_amd.build   (or muspinsim_amd.build.build()).
"""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmuspin_hip.so")
SOURCES = ["msp_kernels.hip", "msp_tridiag_lp.hip", "msp_tridiag_rs.hip", "msp_backtransform_mfma.hip", "msp_sbr.hip", "msp_lindblad.hip",
           "msp_mixed.hip", "msp_timesum_mfma.hip",
           "msp_adjoint.hip", "msp_gemm_mfma.hip", "msp_celio.hip", "msp_capi.hip"]
# per-source code generation options (none at present)
EXTRA = {}
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed FP32 instructions are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def build(verbose=False):
    """Compile every HIP source of the library (one hipcc per source, in
    parallel) and link muspinsim_amd/libmuspin_hip.so."""
    import tempfile
    tmpdir = tempfile.mkdtemp(prefix="msp_build_")
    objs, procs = [], []
    for s in SOURCES:
        obj = os.path.join(tmpdir, s.replace(".hip", ".o"))
        cmd = [HIPCC] + [f for f in FLAGS if f != "-shared"] + EXTRA.get(s, []) + \
            ["-c", os.path.join(CSRC, s), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((s, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                          text=True)))
        objs.append(obj)
    failed = []
    for s, pr in procs:
        out, _ = pr.communicate()
        if pr.returncode != 0:
            sys.stderr.write(out)
            failed.append(s)
    if failed:
        raise RuntimeError("hipcc failed on " + ", ".join(failed))
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed linking libmuspin_hip.so")
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    os.rmdir(tmpdir)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
