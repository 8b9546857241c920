"""Builds libncs_hip.so (gfx950) in-tree with hipcc.

    python -m ncs_amd.build          # build (always recompiles)

The library is written next to this file so that it travels with the source
tree; ncs_amd._lib loads it from there and refuses to run without it.
"""
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libncs_hip.so")
ARCH = "gfx950"

# packed FP32 VALU ops are disabled: the compiler must not form
# v_pk_{add,mul,fma}_f32 from scalar float code.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-ffp-contract=fast"]

SOURCES = ["ncs_solver.hip", "ncs_convnet.hip", "ncs_camera.hip", "ncs_subregion.hip",
           "ncs_capi.cpp"]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found; libncs_hip.so cannot be built")


def build(verbose=False, jobs=4):
    """Compile every source for gfx950 and link libncs_hip.so."""
    cc = hipcc()
    tmp = tempfile.mkdtemp(prefix="ncs_build_")
    try:
        procs = []
        objs = []
        for src in SOURCES:
            obj = os.path.join(tmp, os.path.splitext(src)[0] + ".o")
            cmd = [cc, "-x", "hip", "--offload-arch=" + ARCH] + HIP_FLAGS + \
                  ["-I", os.path.join(HERE, "..", "include"), "-c",
                   os.path.join(CSRC, src), "-o", obj]
            if verbose:
                print(" ".join(cmd), flush=True)
            procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                                stderr=subprocess.STDOUT, text=True)))
            objs.append(obj)
        failed = []
        for src, p in procs:
            out, _ = p.communicate()
            if p.returncode != 0:
                failed.append(src)
                sys.stderr.write(out)
            elif verbose and out.strip():
                sys.stderr.write(out)
        if failed:
            raise RuntimeError("hipcc failed on " + ", ".join(failed))
        tmp_lib = LIB + ".tmp"
        cmd = [cc, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp_lib] + objs
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
        os.replace(tmp_lib, LIB)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
