"""Builds polatory_amd/libpolatory_amd.so from csrc/*.hip for gfx950 with hipcc.

Every translation unit is compiled to polatory_amd/_build/*.o and linked into
one shared library exporting the C ABI declared in include/polatory_amd.h.
"""
import concurrent.futures
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpolatory_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -packed-fp32-ops: no v_pk_{add,mul,fma}_f32 in any kernel.
CFLAGS = ["-O3", "-std=c++20", f"--offload-arch={ARCH}", "-fPIC", "-Wall", "-fopenmp",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          f"-I{os.path.join(ROOT, 'include')}"]


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC] + CFLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                     if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{msgs}")
    return obj, msgs


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    jobs = min(len(srcs), max(1, (os.cpu_count() or 4)), 16)
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(_compile, srcs))
    for obj, msgs in results:
        if verbose and msgs:
            print(msgs)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-fopenmp", "-o", tmp] + [o for o, _ in results]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}{r.stderr}")
    os.replace(tmp, LIB)
    build_cpp_tests()
    return LIB


def build_cpp_tests():
    """Compiles tests/cpp/*.cpp (C++ API mirror users) against the library."""
    tdir = os.path.join(ROOT, "tests", "cpp")
    for src in sorted(glob.glob(os.path.join(tdir, "*.cpp"))):
        exe = src[:-4]
        cmd = ["g++", "-O2", "-std=c++17", f"-I{os.path.join(ROOT, 'include')}",
               f"-I{os.path.join(HERE, 'include')}", src, "-o", exe, f"-L{HERE}", "-lpolatory_amd",
               "-Wl,-rpath,$ORIGIN/../../polatory_amd"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"g++ failed for {src}:\n{r.stdout}{r.stderr}")


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
