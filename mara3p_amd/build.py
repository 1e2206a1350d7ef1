"""Build recipe for the HIP library (gfx950) and the test oracle.

    python mara3p_amd/build.py          # HIP library + oracle (+ reference harness when present)

The HIP library is compiled in-tree so the .so travels with the repository
snapshot.  Flags:
  -ffp-contract=off        no implicit FMA: the FP64 kernels round exactly as
                           the reference and shared face fluxes are bit-identical
  -packed-fp32-ops off     no packed FP32 instructions
"""
import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "mara3p_amd")
CSRC = os.path.join(PKG, "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
REFERENCE = os.environ.get("MARA3P_REFERENCE", "/root/reference")

# capi.hip holds the entry points; the k_*.hip units compile the stage kernels
# (one scheme family each), in parallel.
HIP_SOURCES = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
HIP_FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC",
    "-ffp-contract=off", "-Wno-pass-failed",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]
OBJ_DIR = os.path.join(PKG, "_build")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in res.stdout.splitlines() if "packed-fp32-ops" not in l and "argument unused" not in l)
    if res.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), out))
    if verbose and out.strip():
        print(out)


def build_hip(verbose=False, jobs=None):
    """Compile every unit under csrc (in parallel) and link the library."""
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(OBJ_DIR, exist_ok=True)
    target = os.path.join(PKG, "libmara3p_amd.so")
    tmp = target + ".tmp"
    objs = [os.path.join(OBJ_DIR, s[:-4] + ".o") for s in HIP_SOURCES]
    jobs = jobs or min(len(HIP_SOURCES), max(1, min(os.cpu_count() or 1, 16)))

    def compile_one(k):
        _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, HIP_SOURCES[k]), "-o", objs[k]], verbose)
    with ThreadPoolExecutor(jobs) as ex:
        list(ex.map(compile_one, range(len(HIP_SOURCES))))
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-o", tmp], verbose)
    os.replace(tmp, target)
    return target


def build_oracle(verbose=False):
    oracle = os.path.join(ROOT, "oracle")
    _run(["make", "-s", "-C", oracle, os.path.join(oracle, "liboracle.so")], verbose)
    if os.path.isdir(REFERENCE):
        _run(["make", "-s", "-C", oracle, "REFERENCE=" + REFERENCE, "ref"], verbose)


def build_host_tests(verbose=False):
    """C++ test of the host API against the oracle (tests/cpp)."""
    src = os.path.join(ROOT, "tests", "cpp", "test_host_api.cpp")
    out = os.path.join(ROOT, "tests", "cpp", "test_host_api")
    _run([HIPCC, "-std=c++17", "-O2", "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(PKG, "host"),
          src, "-o", out + ".tmp",
          "-L" + PKG, "-lmara3p_amd", "-L" + os.path.join(ROOT, "oracle"), "-loracle",
          "-Wl,-rpath,$ORIGIN/../../mara3p_amd:$ORIGIN/../../oracle"], verbose)
    os.replace(out + ".tmp", out)


def build(verbose=False):
    build_hip(verbose)
    if shutil.which("make") and shutil.which("gcc"):
        build_oracle(verbose)
        build_host_tests(verbose)


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
