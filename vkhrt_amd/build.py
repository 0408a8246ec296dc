"""Builds vkhrt_amd/libvkhrt.so (host C++ + gfx950 HIP kernels) with hipcc, in-tree, and the
counting build vkhrt_amd/libvkhrt_count.so: the same sources with VKHRT_WORK_COUNTERS, whose
traversal kernels also sum work counters (vkhrt_work_counters) for the measurement scripts."""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libvkhrt.so")
OBJ_COUNT = os.path.join(HERE, "_build_count")
LIB_COUNT = os.path.join(HERE, "libvkhrt_count.so")
SOURCES = ["api.cpp", "asset.cpp", "noalloc.cpp", "bvh_build.hip", "geometry.hip", "voxel.hip", "trace.hip", "trace_phantom_const.hip",
           "trace_phantom_point.hip", "trace_lss.hip", "trace_dots.hip", "trace_tris.hip", "trace_mixed.hip", "mlp.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

# -ffp-contract=off: intersection arithmetic rounds like the reference GLSL text.
# -packed-fp32-ops: no packed FP32 instructions are generated.
FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-but-set-variable", "-Wno-unused-value", "-Wno-unused-result"]

# Per-source flags. mlp.hip: MFMA results in ordinary VGPRs, so the bias, ReLU and bf16 packing
# read them without v_accvgpr_read copies (the gate's VALU count per 32 samples falls by 13 %).
SOURCE_FLAGS = {"mlp.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form=1"]}


def _compile(job):
    src, counting = job
    out = os.path.join(OBJ_COUNT if counting else OBJ, os.path.splitext(src)[0] + ".o")
    lang = ["-x", "hip"] if src.endswith(".cpp") else []
    extra = ["-DVKHRT_WORK_COUNTERS"] if counting else []
    cmd = [HIPCC] + lang + FLAGS + SOURCE_FLAGS.get(src, []) + extra + ["-I", INCLUDE, "-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return out, r.stderr


def _link(objs, lib):
    tmp = lib + ".tmp"
    r = subprocess.run([HIPCC, "-shared", "--offload-arch=gfx950", "-o", tmp] + objs, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, lib)


def build(verbose=False, counting=True):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(OBJ_COUNT, exist_ok=True)
    sources = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    jobs = [(s, False) for s in sources] + ([(s, True) for s in sources] if counting else [])
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(os.cpu_count() or 8, len(jobs))) as ex:
        results = list(ex.map(_compile, jobs))
    if verbose:
        for (_, log), (s, c) in zip(results, jobs):
            warn = "\n".join(line for line in log.splitlines() if "packed-fp32-ops" not in line
                             and "not a recognized feature" not in line)
            if warn.strip() and not c:
                print(f"[{s}]\n{warn}")
    _link([o for (o, _), (_, c) in zip(results, jobs) if not c], LIB)
    if counting:
        _link([o for (o, _), (_, c) in zip(results, jobs) if c], LIB_COUNT)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
