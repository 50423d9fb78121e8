"""Builds libtay_amd.so (HIP, gfx950) in-tree, and the CPU oracle used by the
tests.  Run `python -m tay_amd.build` or call build() / build_oracle().

The library is compiled with the float semantics its parity depends on: no
contraction into fused multiply-add (-ffp-contract=off), correctly rounded
division and square root (hipcc's default), and no packed FP32 instructions.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libtay_amd.so")
ORACLE = os.path.join(ROOT, "oracle")
REFERENCE = "/root/reference"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC, "-I" + BUILD,
]

HOST_FLAGS = [
    "-O2", "-std=c++17", "-fPIC", "-fvisibility=hidden", "-ffp-contract=off", "-Wall",
    "-I" + os.path.join(ROOT, "include"),
]

SOURCES = ["gpu_space.hip", "gpu_batch.hip", "dflock.hip", "gpu_static.hip", "jit.cpp", "tay_api.cpp"]
HOST_ONLY = {"tay_api.cpp"}
EMBEDDED = {"device_abi_text": "device_abi.h", "device_math_text": "device_math.h", "jit_device_text": "jit_device.h"}


def _write_embedded(overrides=None, build_dir=BUILD):
    """Embed device_abi.h, device_math.h and jit_device.h as string constants (overrides
    maps a header name to another file to embed in its place)."""
    os.makedirs(build_dir, exist_ok=True)
    out = ['// generated by tay_amd/build.py from csrc/device_abi.h, csrc/device_math.h and csrc/jit_device.h',
           '#include "internal.h"', 'namespace tay {']
    for name, header in EMBEDDED.items():
        with open((overrides or {}).get(header, os.path.join(CSRC, header))) as f:
            text = f.read()
        if ")TAYEMBED\"" in text:
            raise RuntimeError(header + " contains the raw-string delimiter")
        out.append('extern const char *const %s = R"TAYEMBED(%s)TAYEMBED";' % (name, text))
    out.append("}  // namespace tay")
    path = os.path.join(build_dir, "embedded.cpp")
    content = "\n".join(out) + "\n"
    if not os.path.exists(path) or open(path).read() != content:
        with open(path, "w") as f:
            f.write(content)
    return path


def _run(cmd, cwd=None):
    proc = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        raise RuntimeError("command failed (%d): %s\n%s" % (proc.returncode, " ".join(cmd), proc.stdout))
    return proc.stdout


def build_library(verbose=False, out=LIB, overrides=None, build_dir=BUILD):
    """Compile every HIP/C++ source for gfx950 and link libtay_amd.so (or
    `out`, for experiments, with the embedded headers or sources named in
    `overrides` taken from other files)."""
    embedded = _write_embedded(overrides, build_dir)
    objs = []
    jobs = []
    for src in SOURCES + [embedded]:
        path = src if os.path.isabs(src) else (overrides or {}).get(src, os.path.join(CSRC, src))
        obj = os.path.join(build_dir, os.path.basename(path) + ".o")
        objs.append(obj)
        if src in HOST_ONLY:
            # host C++ without HIP headers (sees Tay's float4, not HIP's)
            jobs.append(["g++", "-x", "c++"] + HOST_FLAGS + ["-c", path, "-o", obj])
        else:
            jobs.append([HIPCC, "-x", "hip", "--offload-arch=" + ARCH] + COMMON_FLAGS + ["-c", path, "-o", obj])
    with ThreadPoolExecutor(max_workers=min(4, len(jobs))) as pool:
        outs = list(pool.map(_run, jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    tmp = out + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs +
         ["-Wl,-Bsymbolic", "-lhiprtc", "-lamdhip64"])
    os.replace(tmp, out)
    return out


def build_oracle(with_reference=True):
    """Build the CPU oracle (test infrastructure) and, where the reference
    tree exists, the reference's own grid code into oracle/_ref/."""
    _run(["make", "-s", "-C", ORACLE, "all"])
    if with_reference and os.path.isdir(os.path.join(REFERENCE, "tay")):
        _run(["make", "-s", "-C", ORACLE, "ref", "REF=" + REFERENCE])


C_HOST = os.path.join(ROOT, "tests", "c_host")


def build_c_host():
    """Build the plain C host test program (tests/c_host/flock_host.c)
    against include/tay.h and libtay_amd.so, as a reference host links Tay."""
    out_dir = os.path.join(C_HOST, "_build")
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, "flock_host")
    _run(["gcc", "-O2", "-std=gnu11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
          "-I", os.path.join(PKG, "agent_src"), os.path.join(C_HOST, "flock_host.c"), "-o", out,
          "-L", PKG, "-ltay_amd", "-Wl,-rpath,$ORIGIN/../../../tay_amd", "-Wl,-rpath-link,/opt/rocm/lib"])
    return out


def build(verbose=False):
    lib = build_library(verbose)
    build_oracle()
    build_c_host()
    return lib


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
