"""Compiles the HIP sources of nngpr_amd into libnngp_hip.so (gfx950).

Run as ``python -m nngpr_amd.build``.  The library is built in-tree next to this
file; the package loads it from there and refuses to run without it.
"""
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libnngp_hip.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

ARCH = os.environ.get("NNGP_OFFLOAD_ARCH", "gfx950")


def hipcc():
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    return os.path.join(rocm, "bin", "hipcc")


def sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")))


def build(quiet=False, jobs=None):
    """Compile every .hip source to an object, then link the shared library."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    # -amdgpu-mfma-vgpr-form: matrix instructions write their results to VGPRs, so
    # the screen reads them without an AGPR-to-VGPR copy per result
    flags = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH, "-fno-slp-vectorize",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-mllvm", "-amdgpu-mfma-vgpr-form",
             "-Wno-unused-command-line-argument", "-I", INCLUDE]
    procs = []
    objs = []
    for src in sources():
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        objs.append(obj)
        cmd = [hipcc()] + flags + ["-c", src, "-o", obj]
        if not quiet:
            print(" ".join(cmd), file=sys.stderr)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        text = "\n".join(l for l in text.splitlines() if "packed-fp32-ops" not in l)
        if p.returncode != 0:
            failed.append((src, text))
        elif text.strip() and not quiet:
            print(text, file=sys.stderr)
    if failed:
        for src, text in failed:
            print("=== %s ===\n%s" % (src, text), file=sys.stderr)
        raise RuntimeError("HIP compilation failed: " + ", ".join(os.path.basename(s) for s, _ in failed))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    if not quiet:
        print(" ".join(cmd), file=sys.stderr)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
