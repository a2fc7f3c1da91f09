"""Builds the gfx950 HIP library libemerald_sph.so in-tree with hipcc.

Flags that the bit-exact parity of the path depends on:
  -ffp-contract=off       no a*b+c fused into FMA (the reference's x86-64 build has none)
  -fno-fast-math          IEEE semantics; HIP's f32 divide and sqrt stay correctly rounded
  -Xclang -target-feature -Xclang -packed-fp32-ops
                          no packed FP32 instructions are generated
"""

from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libemerald_sph.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SOURCES = ["engine.hip", "sph_kernels.hip", "sort.hip", "host_scene.cpp", "snapshot.cpp"]
HEADERS = ["device_math.hpp", "kernels.hpp", "sort_networks.hpp"]

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # Loop heads on 128-byte boundaries: the neighbour build's time depends on
    # its code layout, and this layout measured 3-4 % faster than the default.
    "-falign-loops=128",
    "-Wall",
    "-Wno-unused-result",
]


def _objects_dir() -> str:
    d = os.path.join(HERE, "_build")
    os.makedirs(d, exist_ok=True)
    return d


def build(verbose: bool = False, jobs: int = 4) -> str:
    """Compile every source for gfx950 and link libemerald_sph.so."""
    out = _objects_dir()
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(out, src + ".o")
        objs.append(obj)
        cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd = [HIPCC, "-x", "hip", *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                            stderr=subprocess.STDOUT)))
        if len(procs) >= jobs:
            _wait(procs.pop(0))
    for p in procs:
        _wait(p)
    _check_build_layout(os.path.join(out, "sph_kernels.hip.o"))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_headless(verbose)
    build_api_check(verbose)
    return LIB


# The 16-entry IISPH neighbour build and the byte of a 4 KB page its entry
# should sit at in sph_kernels.hip's code object (k_layout_pad there).
BUILD_KERNEL = "_ZN5emsph7k_buildILi1ELi16ELi64ELb0EEEvNS_5BatchE"
BUILD_KERNEL_PAGE_OFFSET = 3840


def _check_build_layout(obj: str) -> None:
    """Warn when the neighbour build's entry is not at BUILD_KERNEL_PAGE_OFFSET
    of a page (an edit before it in sph_kernels.hip moved it; the size of
    k_layout_pad there sets it). Diagnostics only: never fails the build."""
    llvm = os.path.join(os.path.dirname(os.path.realpath(HIPCC)), "..", "lib", "llvm", "bin")
    if not os.path.isdir(llvm):
        llvm = "/opt/rocm/lib/llvm/bin"
    tmp = obj + ".layout"
    try:
        subprocess.run([os.path.join(llvm, "llvm-objcopy"), f"--dump-section=.hip_fatbin={tmp}.fb",
                        obj, os.devnull], check=True, capture_output=True)
        subprocess.run([os.path.join(llvm, "clang-offload-bundler"), "--unbundle", "--type=o",
                        f"--input={tmp}.fb", f"--targets=hipv4-amdgcn-amd-amdhsa--{ARCH}",
                        f"--output={tmp}.co"], check=True, capture_output=True)
        syms = subprocess.run([os.path.join(llvm, "llvm-readelf"), "-s", "--wide", f"{tmp}.co"],
                              check=True, capture_output=True, text=True).stdout
    except (OSError, subprocess.CalledProcessError):
        return
    finally:
        for ext in (".fb", ".co"):
            if os.path.exists(tmp + ext):
                os.remove(tmp + ext)
    for line in syms.splitlines():
        f = line.split()
        if len(f) >= 8 and f[7] == BUILD_KERNEL and f[3] == "FUNC":
            off = int(f[1], 16) % 4096
            if off != BUILD_KERNEL_PAGE_OFFSET:
                print(f"build: the 16-entry IISPH neighbour build starts at byte {off} of a "
                      f"4 KB page, not {BUILD_KERNEL_PAGE_OFFSET} (resize k_layout_pad in "
                      f"sph_kernels.hip; the build's time depends on it)", file=sys.stderr,
                      flush=True)
            return


HEADLESS = os.path.join(HERE, "sph2d_box_headless")


def build_headless(verbose: bool = False) -> str:
    """The C++ headless driver, linked against the C ABI only (plain g++)."""
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-o", HEADLESS + ".tmp",
           os.path.join(CSRC, "headless_main.cpp"), "-L" + HERE, "-lemerald_sph",
           "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(HEADLESS + ".tmp", HEADLESS)
    return HEADLESS


API_CHECK = os.path.join(os.path.dirname(HERE), "tests", "cpp", "api_check")


def build_api_check(verbose: bool = False) -> str:
    """The C++ mirror header's check program (tests/cpp/api_check.cpp), linked
    against the C ABI only (plain g++), run by the GPU tests."""
    src = API_CHECK + ".cpp"
    if not os.path.exists(src):
        return ""
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-o", API_CHECK + ".tmp", src, "-L" + HERE,
           "-lemerald_sph", "-Wl,-rpath,$ORIGIN/../../emerald_amd"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(API_CHECK + ".tmp", API_CHECK)
    return API_CHECK


def _wait(item) -> None:
    src, p = item
    text, _ = p.communicate()
    if p.returncode != 0:
        sys.stderr.write(text.decode(errors="replace"))
        raise RuntimeError(f"hipcc failed on {src}")
    msg = text.decode(errors="replace").strip()
    if msg:
        # hipcc's host pass warns that it does not know the packed-fp32 feature.
        lines = [ln for ln in msg.splitlines() if "packed-fp32-ops" not in ln]
        if lines:
            sys.stderr.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
