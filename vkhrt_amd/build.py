"""Builds vkhrt_amd/libvkhrt.so (host C++ + gfx950 HIP kernels) with hipcc, in-tree."""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libvkhrt.so")
SOURCES = ["api.cpp", "asset.cpp", "noalloc.cpp", "bvh_build.hip", "geometry.hip", "voxel.hip", "trace.hip", "trace_phantom_const.hip",
           "trace_phantom_point.hip", "trace_lss.hip", "trace_dots.hip", "trace_tris.hip", "trace_mixed.hip", "mlp.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

# -ffp-contract=off: intersection arithmetic rounds like the reference GLSL text.
# -packed-fp32-ops: no packed FP32 instructions are generated.
FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-but-set-variable", "-Wno-unused-value", "-Wno-unused-result"]


def _compile(src):
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    lang = ["-x", "hip"] if src.endswith(".cpp") else []
    cmd = [HIPCC] + lang + FLAGS + ["-I", INCLUDE, "-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return out, r.stderr


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    sources = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(os.cpu_count() or 8, len(sources))) as ex:
        results = list(ex.map(_compile, sources))
    if verbose:
        for (_, log), s in zip(results, sources):
            warn = "\n".join(line for line in log.splitlines() if "packed-fp32-ops" not in line
                             and "not a recognized feature" not in line)
            if warn.strip():
                print(f"[{s}]\n{warn}")
    objs = [o for o, _ in results]
    tmp = LIB + ".tmp"
    r = subprocess.run([HIPCC, "-shared", "--offload-arch=gfx950", "-o", tmp] + objs,
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
