"""Builds the test helper programs in place (test infrastructure, not part
of the library):

* hdf5_snap: writes and dumps Gadget-HDF5 sets through the HDF5 C library
  (when the library and its high-level API are present);
* compat_driver: the reference driver's forward through the drop-in shims
  of include/genpk_compat.h, linked against libgenpk_amd.so.

gadget_dump and bigfile_dump are built by the CPU tests that use them.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def build():
    from genpk_amd import build as lib_build
    pkg = os.path.join(ROOT, "genpk_amd")
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    hdf5 = lib_build.hdf5_flags()
    if hdf5["cflags"]:
        out = os.path.join(HERE, "hdf5_snap")
        subprocess.run(["g++", "-O1", "-std=c++17", "-Wall", "-I", os.path.join(pkg, "cli"), *hdf5["cflags"],
                        out + ".cpp", "-o", out + ".tmp", *hdf5["libs"],
                        "-Wl,-rpath,$ORIGIN/../../genpk_amd/_hdf5lib"], check=True)
        os.replace(out + ".tmp", out)
    out = os.path.join(HERE, "compat_driver")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), out + ".cpp", "-o",
                    out + ".tmp", "-L", pkg, "-lgenpk_amd", "-L", os.path.join(rocm, "lib"), "-lamdhip64",
                    "-Wl,-rpath,$ORIGIN/../../genpk_amd", "-Wl,-rpath," + os.path.join(rocm, "lib")], check=True)
    os.replace(out + ".tmp", out)


if __name__ == "__main__":
    build()
