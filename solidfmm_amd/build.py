"""Build the HIP translation library in-tree: solidfmm_amd/lib/libsolidfmm_gpu.so.

Every source is compiled for gfx950 only, with packed FP32 instructions
disabled, then linked into one shared library exporting the C ABI of
include/solidfmm_gpu.h.  Run ``python -m solidfmm_amd.build``.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "lib")
OBJ_DIR = os.path.join(HERE, "lib", "obj")
LIB = os.path.join(OUT_DIR, "libsolidfmm_gpu.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    "-O3", "-std=c++20", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fconstexpr-steps=100000000", "-Wall", "-Wno-unused-function",
]

# Translation units; the kernel instantiations are split so they compile in
# parallel, the longest first.
SOURCES = ["mfma_prec_hi.hip", "mfma_prec_mid.hip", "lane_prec_dw.hip", "mfma_m2l_rt.hip", "mfma_m2l.hip", "mfma_m2m_rt.hip", "mfma_l2l_rt.hip", "mfma_m2m.hip",
           "mfma_l2l.hip", "lane_prec_f32.hip", "lane_rt.hip", "lane_fixed.hip", "prec_generic.hip", "mfma_prec.hip", "reduce.hip", "particles.hip", "translate.hip", "generic.hip", "arith_probe.hip", "swap_host.cpp", "capi.cpp", "dropin.cpp",
           "sphere.cpp", "server.cpp"]

# Host-only C++ (the drop-in API over the C ABI).
HOST_ONLY = {"dropin.cpp", "sphere.cpp"}
INCLUDE = os.path.join(os.path.dirname(HERE), "include")


def _compile(src):
    obj = os.path.join(OBJ_DIR, os.path.splitext(src)[0] + ".o")
    if src in HOST_ONLY:
        cmd = [HIPCC, "-O3", "-std=c++20", "-fPIC", "-Wall", "-x", "c++", "-I", INCLUDE,
               "-c", os.path.join(CSRC, src), "-o", obj]
    else:
        lang = [] if src.endswith(".hip") else ["-x", "hip"]
        cmd = [HIPCC] + COMMON + lang + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def _stale(src):
    """True when src's object is missing or older than src or any header."""
    obj = os.path.join(OBJ_DIR, os.path.splitext(src)[0] + ".o")
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    deps = [os.path.join(CSRC, src)] + [os.path.join(CSRC, h) for h in os.listdir(CSRC) if h.endswith(".hpp")]
    deps += [os.path.join(INCLUDE, "solidfmm_gpu.h")]
    deps += [os.path.join(INCLUDE, "solidfmm", h) for h in os.listdir(os.path.join(INCLUDE, "solidfmm"))]
    return any(os.path.getmtime(d) > t for d in deps)


def build(verbose=True, incremental=False):
    """Compile every source (incremental: only stale ones) and link."""
    os.makedirs(OBJ_DIR, exist_ok=True)
    todo = [s for s in SOURCES if _stale(s)] if incremental else list(SOURCES)
    if todo:
        with cf.ThreadPoolExecutor(max_workers=min(os.cpu_count() or 8, len(todo))) as ex:
            list(ex.map(_compile, todo))
    objs = [os.path.join(OBJ_DIR, os.path.splitext(s)[0] + ".o") for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(incremental="--incremental" in sys.argv)
    sys.exit(0)
