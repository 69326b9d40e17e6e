"""Build recipe of the HIP library (libbroadcast_amd.so, gfx950).

    python -m broadcast_amd.build
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("BCA_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-shared", "-std=c++17",
    # packed FP32 forms are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    return subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)


def build(verbose=False, jobs=None):
    """Compile the C-ABI object and the seven face-kernel variants in parallel,
    then link libbroadcast_amd.so."""
    csrc = os.path.join(HERE, "csrc")
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    units = [("capi", os.path.join(csrc, "capi.hip"), [])]
    units += [(f"blocks{v}", os.path.join(csrc, "blocks.hip"), [f"-DBCA_VARIANT={v}"]) for v in range(7)]
    procs = []
    for name, src, defs in units:
        obj = os.path.join(objdir, name + ".o")
        cmd = [HIPCC, *[f for f in FLAGS if f != "-shared"], *defs, "-c", "-o", obj, src]
        procs.append((name, obj, _run(cmd, verbose)))
    failed = []
    for name, obj, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append(name)
            sys.stderr.write(out)
    if failed:
        raise RuntimeError(f"hipcc failed for {failed}")
    out = os.path.join(HERE, "libbroadcast_amd.so")
    tmp = out + ".tmp"
    link = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *[o for _, o, _ in procs]]
    r = subprocess.run(link, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed for libbroadcast_amd.so")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
