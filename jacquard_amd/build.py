"""Build the native libraries in-tree.

  jacquard_amd/lib/libjacquard_hip.so   gfx950 kernels + C ABI (hipcc)
  jacquard_amd/lib/libjacquard_host.so  script compiler + generators (g++)

Every HIP source is compiled for gfx950 only, with packed-FP32 instruction
selection disabled.  Run ``python -m jacquard_amd.build``.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "lib")
INC = os.path.join(ROOT, "include")

DEVICE_SOURCES = ["jq_scan.hip", "jq_batched.hip", "jq_mlp.hip", "jq_vcd.hip",
                  "jq_probe.hip"]
HOST_SOURCES = ["aig.cpp", "netgen.cpp", "compile.cpp", "verilog.cpp", "capi.cpp", "vcd.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + INC,
]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build_device(jobs=8):
    os.makedirs(LIB, exist_ok=True)
    objs = []
    procs = []
    for src in DEVICE_SOURCES:
        path = os.path.join(HERE, "csrc", "device", src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(LIB, src.replace(".hip", ".o"))
        cmd = [HIPCC] + HIP_FLAGS + ["-c", path, "-o", obj]
        print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        objs.append(obj)
    for p in procs:
        if p.wait() != 0:
            raise RuntimeError("hipcc failed")
    out = os.path.join(LIB, "libjacquard_hip.so")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", out] + objs)
    for o in objs:
        os.remove(o)
    return out


def build_host():
    os.makedirs(LIB, exist_ok=True)
    srcs = [os.path.join(HERE, "csrc", "host", s) for s in HOST_SOURCES
            if os.path.exists(os.path.join(HERE, "csrc", "host", s))]
    out = os.path.join(LIB, "libjacquard_host.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-Wall",
          "-Wno-unused-function", "-o", out] + srcs)
    return out


def main():
    build_host()
    build_device()


if __name__ == "__main__":
    sys.exit(main())
