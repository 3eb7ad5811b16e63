"""Build recipe for the gfx950 worker: libbsf.so and the BetaSeedFinderWorker executable.

Everything is compiled in-tree with hipcc into betaseedfinder_amd/lib/ (git-ignored, but it
travels with the tree to the GPU host).  FP flags follow the reference's exactness flags
(native/CMakeLists.txt:18-22, scripts/build-native.ps1:205-209): no FMA contraction, no
fast-math, no reassociation.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")
ROOT = os.path.dirname(PKG)

ARCH = os.environ.get("BSF_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = [
    "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math", "-fno-associative-math",
    "-Wall", "-Wno-unused-parameter", "-Wno-unused-function",
    "-I" + os.path.join(ROOT, "include"),
]
DEVICE = [
    "--offload-arch=" + ARCH,
    # packed FP32 ops stay off on this fleet; the host pass warns it does not know the feature
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

LIB_SOURCES = ["bsf_kernels.hip", "bsf_stage0.hip", "bsf_island.hip", "bsf_flood.hip", "bsf_region.hip", "bsf_api.cpp",
               "bsf_engine.cpp", "bsf_island_api.cpp", "bsf_server.cpp", "bsf_stage0_api.cpp", "bsf_region_api.cpp", "bsf_run.cpp"]
WORKER_SOURCES = ["worker_main.cpp"]


def _run(cmd):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), out))
    return out


def _obj(src):
    return os.path.join(LIB, "obj", os.path.splitext(src)[0] + ".o")


def build(verbose=False):
    os.makedirs(os.path.join(LIB, "obj"), exist_ok=True)
    sources = [s for s in LIB_SOURCES if os.path.exists(os.path.join(CSRC, s))]

    def compile_one(src):
        cmd = [HIPCC] + COMMON + DEVICE + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", _obj(src)]
        return _run(cmd)

    with ThreadPoolExecutor(max_workers=min(8, len(sources))) as ex:
        for out in ex.map(compile_one, sources):
            if verbose and out.strip():
                print(out)
    so = os.path.join(LIB, "libbsf.so")
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-o", so + ".tmp"] + [_obj(s) for s in sources])
    os.replace(so + ".tmp", so)
    # worker executable, linked against libbsf.so with an rpath to its own directory
    worker_src = [os.path.join(CSRC, s) for s in WORKER_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    if worker_src:
        exe = os.path.join(LIB, "BetaSeedFinderWorker")
        _run(["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include")] + worker_src +
             ["-L" + LIB, "-lbsf", "-Wl,-rpath,$ORIGIN", "-o", exe + ".tmp"])
        os.replace(exe + ".tmp", exe)
        # the Java launcher resolves backend/<vendor>/BetaSeedFinderWorker.exe (GpuBackendLocator.java:12-20)
        shutil.copyfile(exe, exe + ".exe")
        os.chmod(exe + ".exe", 0o755)
    return so


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print("built", os.path.join(LIB, "libbsf.so"))
