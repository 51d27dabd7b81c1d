"""Build recipe for the k-CDS libraries (hipcc for gfx950, g++ for host-only code).

  libkcds.so      product library: HIP kernels + C ABI (include/kcds.h)
  libkcdsgen.so   seeded graph generators (host only)
  kcds            drop-in CLI: kcds <graph> <k> [pSize cpSize glBufferSize partitionSize]

Objects go to test_amd/_build/; outputs sit next to this file.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]
LIB_SOURCES = ["kcds_api.cpp", "kcds_solve.cpp", "kcds_serve.cpp", "k_util.hip", "k_orient.hip", "k_list.hip", "k_peel.hip", "k_epeel.hip", "k_dense.hip",
               "k_comp.hip", "k_flow.hip", "k_gflow.hip", "k_edges.hip"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def _obj(src):
    return os.path.join(OBJ, os.path.basename(src) + ".o")


def build_lib(jobs=8):
    os.makedirs(OBJ, exist_ok=True)
    srcs = [os.path.join(SRC, s) for s in LIB_SOURCES]

    def comp(s):
        lang = ["-x", "hip"] if s.endswith(".hip") else []
        return _run([HIPCC] + HIP_FLAGS + lang + ["-c", s, "-o", _obj(s)])

    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(comp, srcs))
    out = os.path.join(HERE, "libkcds.so")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", out + ".tmp"] + [_obj(s) for s in srcs])
    os.replace(out + ".tmp", out)
    return out


def build_cli():
    """The drop-in CLI: Graph loader (graph.cpp), its device wrappers (graph_device.cpp) and main.cpp, linked to
    libkcds.so."""
    os.makedirs(OBJ, exist_ok=True)
    out = os.path.join(HERE, "kcds")
    srcs = [os.path.join(SRC, "graph.cpp"), os.path.join(SRC, "graph_device.cpp"), os.path.join(SRC, "main.cpp")]
    _run([HIPCC, "-O3", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), "-o", out + ".tmp"] + srcs +
         ["-L", HERE, "-lkcds", "-Wl,-rpath,$ORIGIN"])
    os.replace(out + ".tmp", out)
    return out


def build_gen():
    out = os.path.join(HERE, "libkcdsgen.so")
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-Wall", "-o", out + ".tmp",
          os.path.join(SRC, "graphgen.cpp")])
    os.replace(out + ".tmp", out)
    return out


def build_all(jobs=8):
    return [build_gen(), build_lib(jobs), build_cli()]


if __name__ == "__main__":
    for p in build_all(int(sys.argv[1]) if len(sys.argv) > 1 else 8):
        print(p)
