"""Build acousticsim_amd/libbwcloth.so (HIP for gfx950 + host C++) in-tree.

    python -m acousticsim_amd.build

Objects go to acousticsim_amd/_build/.  The assembly kernels are compiled
with floating-point contraction off (their FP64 W must round like the
reference's CPU arithmetic); so is the tolerance-mode solver (cg.hip), whose
fused multiply-adds are written out, so that every kernel variant rounds
alike; the reference-semantics and GPU-form solvers keep contraction on.
Packed FP32 instructions are disabled for all device code.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libbwcloth.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-std=c++17", "-fPIC", "-I" + CSRC, "-I" + INCLUDE, "-Wall",
          "-Wno-unused-result"]
DEVICE = ["-x", "hip", "--offload-arch=" + ARCH,
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

SOURCES = [
    ("json_lite.cpp", ["-ffp-contract=off"]),
    ("model.cpp", ["-ffp-contract=off"]),
    ("mesh_gen.cpp", ["-ffp-contract=off"]),
    ("batch.cpp", ["-ffp-contract=off"]),
    ("capi.cpp", ["-ffp-contract=off"]),
    ("server.cpp", ["-ffp-contract=off", "-pthread"]),
    ("assemble.hip", DEVICE + ["-ffp-contract=off"]),
    ("pcg.hip", DEVICE + ["-ffp-contract=fast"]),
    ("cg.hip", DEVICE + ["-ffp-contract=off"]),
    ("normals.hip", DEVICE + ["-ffp-contract=off"]),
    ("contact.hip", DEVICE + ["-ffp-contract=off"]),
    ("collide.hip", DEVICE + ["-ffp-contract=off"]),
    ("body.cpp", ["-ffp-contract=off"]),
    ("gpuform.hip", DEVICE + ["-ffp-contract=fast"]),
    ("noalloc.hip", DEVICE + ["-ffp-contract=off"]),
]


def _compile(src, flags):
    obj = os.path.join(OUT, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + flags + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(verbose=False):
    os.makedirs(OUT, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        objs = list(ex.map(lambda a: _compile(*a), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "-pthread", "--offload-arch=" + ARCH, "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
