"""Builds libspdepy_amd.so (HIP for gfx950 + host C++) in place.

    python -m spdepy_amd.build          # or: python spdepy_amd/build.py
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libspdepy_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_SOURCES = ["dense.hip", "mf.hip", "numeric.hip", "assemble.hip", "grad.hip", "fields.hip", "kron.hip", "loglik.hip", "pcg.hip", "learned.hip"]
CXX_SOURCES = ["symbolic.cpp"]
COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function", "-Wno-unused-value", "-Wno-unused-result"]
# packed FP32 instructions are disabled for every device compile
HIP_FLAGS = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=900)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build failed: " + " ".join(cmd))
    return out


def build(verbose=False):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(objdir, src + ".o")
        _run([HIPCC] + COMMON + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in CXX_SOURCES:
        obj = os.path.join(objdir, src + ".o")
        _run(["g++"] + COMMON + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = OUT + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, OUT)
    build_shims()
    if verbose:
        print("built", OUT)
    return OUT


# The reference loads one stencil library per kind and boundary condition
# (ccode/lib_{AcH,AH,Acw,Aw}_2D_b{1,2,3}.so); these carry its C interface over
# libspdepy_amd.so (csrc/ccode_shim.cpp).
SHIMS = {"AcH": (1, 0), "AH": (1, 1), "Acw": (0, 0), "Aw": (0, 1)}
SHIM_DIR = os.path.join(HERE, "ccode")


def build_shims():
    os.makedirs(SHIM_DIR, exist_ok=True)
    src = os.path.join(CSRC, "ccode_shim.cpp")
    for kind, (diffusion, variable) in SHIMS.items():
        for bc in (1, 2, 3):
            out = os.path.join(SHIM_DIR, "lib_%s_2D_b%d.so" % (kind, bc))
            _run([HIPCC] + COMMON + ["-D__HIP_PLATFORM_AMD__", "-DSHIM_DIFFUSION=%d" % diffusion,
                                     "-DSHIM_VARIABLE=%d" % variable, "-DSHIM_BC=%d" % bc, "-shared", src,
                                     "-o", out + ".tmp", "-L" + HERE, "-lspdepy_amd", "-Wl,-rpath,$ORIGIN/.."])
            os.replace(out + ".tmp", out)


if __name__ == "__main__":
    build(verbose=True)
