"""In-tree build of the engine's shared libraries (run by __graft_entry__.build()).

  libgst.so      HIP (gfx950) kernels + the C ABI of include/gst.h
  libgstdata.so  CPU generators and reference file formats (include/gst_data.h)
  gst_cli        reference-compatible command-line driver (exe type path name T [D] s e)
"""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = "gfx950"
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             # packed FP32 instructions stay disabled for all device code
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd):
    r = subprocess.run(cmd, cwd=ROOT, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build_data():
    out = os.path.join(HERE, "libgstdata.so")
    _run(["g++", "-O3", "-std=c++17", "-fopenmp", "-fPIC", "-shared", "-o", out,
          os.path.join(HERE, "csrc", "gst_data.cpp")])
    return out


def build_engine():
    out = os.path.join(HERE, "libgst.so")
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", out, os.path.join(HERE, "csrc", "gst_engine.hip")])
    return out


def build_cli():
    out = os.path.join(HERE, "gst_cli")
    _run(["g++", "-O2", "-std=c++17", "-o", out, os.path.join(HERE, "host", "gst_cli.cpp"),
          f"-L{HERE}", "-lgst", "-lgstdata", f"-Wl,-rpath,$ORIGIN", f"-L{ROCM}/lib", "-lamdhip64",
          f"-Wl,-rpath,{ROCM}/lib"])
    return out


def build_all():
    build_data()
    build_engine()
    if os.path.exists(os.path.join(HERE, "host", "gst_cli.cpp")):
        build_cli()


if __name__ == "__main__":
    build_all()
    print("built")
