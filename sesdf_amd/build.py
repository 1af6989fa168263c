"""Build steps for the native libraries (run on the CPU; hipcc cross-compiles gfx950).

* ``libsesdf_synth.so`` -- host-only synthetic atom-cloud generator (g++).
* ``libsesdf_hip.so``   -- the HIP kernels and the C-ABI (hipcc, --offload-arch=gfx950).
* ``tests/cpp/host_api`` -- a C++ program over the header-only host classes (include/bcs), g++.

Everything is built in-tree so the shared objects travel with the source tree.
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")

HIP_SOURCES = [
    "grid.hip",
    "circles.hip",
    "intersections.hip",
    "sdf.hip",
    "components.hip",
    "sort.hip",
    "contour_buildup.hip",
    "mmtf.cpp",
    "capi.cpp",
]
HIP_HEADERS = ["common.h", "kernels.h", "handle.h", "components.h", "sort.h"]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, cwd=cwd, check=True)


def build_synth():
    out = os.path.join(PKG, "libsesdf_synth.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", out, os.path.join(CSRC, "synth.cpp")])
    return out


def hip_flags():
    return [
        "--offload-arch=gfx950",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", os.path.join(ROOT, "include"),
        "-I", CSRC,
        "-Wno-unused-command-line-argument",
    ]


def build_hip(jobs=4):
    """Compile every HIP/C++ source of the product library and link libsesdf_hip.so."""
    objdir = os.path.join(PKG, "_objs")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    procs = []
    for src in HIP_SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmd = ["hipcc"] + hip_flags() + ["-c", path, "-o", obj]
        if src.endswith(".cpp"):
            cmd = ["hipcc", "-x", "hip"] + hip_flags() + ["-c", path, "-o", obj]
        print("+", " ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        if len(procs) >= jobs:
            for p in procs:
                if p.wait() != 0:
                    raise RuntimeError("hipcc failed")
            procs = []
    for p in procs:
        if p.wait() != 0:
            raise RuntimeError("hipcc failed")
    out = os.path.join(PKG, "libsesdf_hip.so")
    _run(["hipcc", "--offload-arch=gfx950", "-shared", "-fPIC", "-o", out] + objs)
    return out


CPP_PROGRAMS = ["host_api", "circles_census", "detailed_stages", "load_molecule", "contour_buildup"]


def build_host_api():
    """The C++ programs over the host classes (bcs::Sesdf / bcs::Ssesdf over the C ABI) that the tests
    run: tests/cpp/<name>.cpp -> tests/cpp/<name>."""
    outs = []
    for name in CPP_PROGRAMS:
        src = os.path.join(ROOT, "tests", "cpp", name + ".cpp")
        out = os.path.join(ROOT, "tests", "cpp", name)
        _run(["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
              "-I", os.path.join(ROOT, "include"), src, "-o", out, "-L", PKG, "-lsesdf_hip",
              "-Wl,-rpath,$ORIGIN/../../sesdf_amd"])
        outs.append(out)
    return outs


def build_oracle():
    _run(["make", "-C", os.path.join(ROOT, "oracle"), "-s"])


def build_all():
    build_synth()
    build_hip()
    build_host_api()


if __name__ == "__main__":
    what = sys.argv[1:] or ["synth", "hip"]
    for w in what:
        {"synth": build_synth, "hip": build_hip, "oracle": build_oracle, "host_api": build_host_api}[w]()
