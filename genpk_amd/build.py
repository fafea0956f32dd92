"""Builds libgenpk_amd.so (HIP for gfx950) in-tree with hipcc.

Every HIP source under csrc/ is compiled separately (in parallel) and
linked into genpk_amd/libgenpk_amd.so.  Also builds the gen-pk command line
tool (genpk_amd/gen-pk) on top of the library.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgenpk_amd.so")
CLI = os.path.join(HERE, "gen-pk")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src):
    obj = os.path.join(OBJ, src.replace(".hip", ".o"))
    cmd = [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{out}")
    return obj, out


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = _sources()
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        results = list(ex.map(_compile, srcs))
    for obj, out in results:
        if out and verbose:
            print(out, file=sys.stderr)
    objs = [o for o, _ in results]
    tmp = LIB + ".tmp"
    subprocess.run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp], check=True)
    os.replace(tmp, LIB)
    cli_src = os.path.join(HERE, "cli", "gen_pk.cpp")
    if os.path.exists(cli_src):
        rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
        subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(rocm, "include"),
                        "-I", os.path.join(HERE, "..", "include"), cli_src, "-o", CLI + ".tmp", "-L", HERE,
                        "-lgenpk_amd", "-L", os.path.join(rocm, "lib"), "-lamdhip64", "-Wl,-rpath,$ORIGIN",
                        "-Wl,-rpath," + os.path.join(rocm, "lib")], check=True)
        os.replace(CLI + ".tmp", CLI)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    print(LIB)
