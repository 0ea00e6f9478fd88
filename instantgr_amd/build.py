"""Build recipe for the native libraries of instantgr_amd.

Outputs, all in-tree under instantgr_amd/lib/:
  libigr_host.so  host C++ (design I/O, trees, writer, evaluator), g++
  libigr.so       HIP kernels for gfx950 + the C ABI, hipcc
  igr_route       the -cap/-net/-out command-line driver

Every HIP translation unit is compiled with -ffp-contract=off (the routing
arithmetic must round exactly as written) and without packed-FP32 ops.
"""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(ROOT, "csrc")
LIB = os.path.join(ROOT, "lib")
INCLUDE = os.path.join(os.path.dirname(ROOT), "include")

HOST_SRCS = ["io.cpp", "netinit.cpp", "gen.cpp", "steiner.cpp", "prep.cpp", "dag.cpp", "evaluator.cpp", "igr_host.cpp"]
HIP_SRCS = ["grid_kernels.hip", "lshape_kernels.hip", "sched_kernels.hip", "sched_batched.hip", "dag_kernels.hip", "dag_layout.hip", "track_kernels.hip",
            "igr_ctx.hip", "serve.hip", "eval_kernels.hip",
            "pipeline.hip", "noalloc.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def _newer(out, deps):
    if not os.path.exists(out):
        return True
    t = os.path.getmtime(out)
    return any(os.path.getmtime(d) > t for d in deps)


def _objs(srcs, subdir, compiler, flags, force=False, ext_out=".o"):
    objdir = os.path.join(LIB, "obj", subdir)
    os.makedirs(objdir, exist_ok=True)
    headers = []
    for d in (os.path.join(CSRC, "host"), os.path.join(CSRC, "hip"), INCLUDE):
        headers += [os.path.join(d, f) for f in os.listdir(d) if f.endswith((".h", ".hpp"))]
    outs = []
    jobs = []
    for s in srcs:
        src = os.path.join(CSRC, subdir, s)
        out = os.path.join(objdir, s + ext_out)
        outs.append(out)
        if force or _newer(out, [src] + headers):
            jobs.append([compiler] + flags + ["-c", src, "-o", out])
    procs = []
    for j in jobs:
        print("+", " ".join(j), flush=True)
        procs.append(subprocess.Popen(j))
        if len(procs) >= 8:
            if procs.pop(0).wait() != 0:
                raise RuntimeError("compile failed")
    for p in procs:
        if p.wait() != 0:
            raise RuntimeError("compile failed")
    return outs


def build_host(force=False):
    os.makedirs(LIB, exist_ok=True)
    flags = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-result",
             "-I" + INCLUDE, "-I" + os.path.join(CSRC, "host")]
    objs = _objs(HOST_SRCS, "host", "g++", flags, force)
    out = os.path.join(LIB, "libigr_host.so")
    if force or _newer(out, objs):
        _run(["g++", "-shared", "-o", out] + objs + ["-lpthread"])
    return out


def build_hip(force=False):
    host = build_host(force)
    flags = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-munsafe-fp-atomics", "-Wno-unused-result",
             "-I" + INCLUDE, "-I" + os.path.join(CSRC, "host"), "-I" + os.path.join(CSRC, "hip")]
    srcs = [s for s in HIP_SRCS if os.path.exists(os.path.join(CSRC, "hip", s))]
    objs = _objs(srcs, "hip", HIPCC, flags, force)
    out = os.path.join(LIB, "libigr.so")
    if force or _newer(out, objs + [host]):
        _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", out] + objs +
             ["-L" + LIB, "-ligr_host", "-Wl,-rpath,$ORIGIN", "-lpthread"])
    drv = os.path.join(CSRC, "host", "igr_route.cpp")
    exe = os.path.join(LIB, "igr_route")
    if os.path.exists(drv) and (force or _newer(exe, [drv, out])):
        _run(["g++", "-O2", "-std=c++17", "-I" + INCLUDE, "-I" + os.path.join(CSRC, "host"), drv, "-o", exe,
              "-L" + LIB, "-ligr", "-ligr_host", "-Wl,-rpath,$ORIGIN", "-lpthread"])
    return out


def build_all(force=True):
    build_hip(force)


if __name__ == "__main__":
    build_all(force="--force" in sys.argv)
