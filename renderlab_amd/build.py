"""Builds librq.so (host C++ with g++, device HIP for gfx950 with hipcc).

Usage: python -m renderlab_amd.build   (or build() from Python)
       python -m renderlab_amd.build --sanitize
           builds renderlab_amd/_san/librq_san.so: the host C++ sources (the
           builders, packers, relayouts and the parsers of external words and
           cooked files) compiled by g++ with AddressSanitizer and
           UndefinedBehaviorSanitizer, linked with the usual device objects. Run
           the CPU tests against it with
             RQ_LIBRARY=renderlab_amd/_san/librq_san.so \
             LD_PRELOAD=$(gcc -print-file-name=libasan.so) ASAN_OPTIONS=detect_leaks=0 \
             python -m pytest tests -m "not gpu"

The kernels are compiled without floating-point contraction and with IEEE
division and square root, so they compute the same f32 arithmetic as the
host-side records they read. Packed FP32 instructions are disabled
(-packed-fp32-ops): the code generator would otherwise form them from plain
float code.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "librq.so")
BUILD = os.path.join(HERE, "_build")

HOST_SOURCES = ["api.cpp", "builder.cpp", "toplevel.cpp", "pack.cpp", "layouts.cpp", "scenegen.cpp", "reflect.cpp", "assets.cpp"]
DEVICE_SOURCES = ["trace.hip", "serve.hip", "sort.hip", "refit.hip", "gpubuild.hip", "gicache.hip", "hits.hip", "noalloc.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("RQ_OFFLOAD_ARCH", "gfx950")

HOST_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-Wall", "-Wno-unused-function",
              "-pthread"]
DEVICE_FLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-ffp-contract=off",
                "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-gpu-flush-denormals-to-zero",
                "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("command failed: " + " ".join(cmd))
    return r.stdout


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    jobs = []
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        jobs.append(["g++", *HOST_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in DEVICE_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        jobs.append([HIPCC, *DEVICE_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    procs = [(j, subprocess.Popen(j, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)) for j in jobs]
    failed = False
    for j, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed = True
            sys.stderr.write(out)
            sys.stderr.write("\nfailed: " + " ".join(j) + "\n")
        elif verbose and out.strip():
            sys.stderr.write(out)
    if failed:
        raise RuntimeError("librq build failed")
    tmp = OUT + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs, "-pthread"])
    os.replace(tmp, OUT)
    return OUT


SAN = os.path.join(HERE, "_san")
SAN_FLAGS = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-g"]


def build_sanitized():
    """librq_san.so: sanitized host objects (g++) with the ordinary device objects, linked by g++."""
    build()  # the device objects
    os.makedirs(SAN, exist_ok=True)
    objs, procs = [], []
    for src in HOST_SOURCES:
        obj = os.path.join(SAN, src + ".o")
        cmd = ["g++", *HOST_FLAGS, *SAN_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
    for cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out)
            raise RuntimeError("command failed: " + " ".join(cmd))
    objs += [os.path.join(BUILD, src + ".o") for src in DEVICE_SOURCES]
    out = os.path.join(SAN, "librq_san.so")
    rocm = os.path.dirname(os.path.dirname(os.path.realpath(HIPCC)))
    _run(["g++", "-shared", *SAN_FLAGS, "-o", out, *objs, "-pthread", f"-L{rocm}/lib", f"-Wl,-rpath,{rocm}/lib",
          "-lamdhip64"])
    return out


def build_oracle():
    oracle = os.path.join(os.path.dirname(HERE), "oracle")
    _run(["make", "-s", "-C", oracle])
    return os.path.join(oracle, "liboracle.so")


if __name__ == "__main__":
    if "--sanitize" in sys.argv:
        print(build_sanitized())
    else:
        print(build(verbose=True))
        print(build_oracle())
