"""In-tree build of the gfx950 prover library (no cmake, no JIT cache).

Pipeline for the round-down kernels (every csrc/*.hip, compiled in parallel):
  1. hipcc --cuda-device-only -S            -> gfx950 assembly
  2. set `.amdhsa_float_round_mode_32/_16_64 2` in every ccrd_* kernel
     descriptor, so each wave starts in round-toward-negative-infinity mode
  3. clang -cc1as per unit, ld.lld of all  -> one code object (.hsaco)
  4. the code object is embedded (.incbin) into libcirclecover_amd.so, which
     loads it with hipModuleLoadData.
Host code is plain C++ over the HIP runtime API.
"""
from __future__ import annotations

import os
import re
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(ROOT, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
LLVM = os.path.join(ROCM, "llvm", "bin")
LIB = os.path.join(PKG, "libcirclecover_amd.so")
PROVER_EXE = os.path.join(PKG, "circlecover_prover")

DEVICE_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fno-fast-math", "-ffp-contract=off",
]
HOST_FLAGS = [
    "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
    "-Wall", "-Wextra", "-Wno-unused-parameter",
    "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include",
]


def _run(cmd: list[str], **kw) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True, **kw)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]}")


def patch_round_mode(asm: str) -> tuple[str, int]:
    """Sets round-down in the descriptor of every ccrd_* kernel; returns the
    patched text and the number of kernels patched."""
    out, n = [], 0
    in_rd = False
    for line in asm.splitlines(keepends=True):
        m = re.match(r"\s*\.amdhsa_kernel\s+(\S+)", line)
        if m:
            in_rd = m.group(1).startswith("ccrd_")
            n += in_rd
        elif re.match(r"\s*\.end_amdhsa_kernel", line):
            in_rd = False
        if in_rd:
            line = re.sub(r"(\.amdhsa_float_round_mode_(?:32|16_64))\s+0\b", r"\1 2", line)
        out.append(line)
    return "".join(out), n


def _compile_rd(src: str) -> subprocess.Popen:
    name = os.path.splitext(os.path.basename(src))[0]
    asm = os.path.join(BUILD, name + ".s")
    return subprocess.Popen(["hipcc", *DEVICE_FLAGS, f"-I{CSRC}", f"-I{ROOT}/include", "--cuda-device-only", "-S",
                             "-o", asm, src], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)


def build_rd_module() -> str:
    """Every csrc/*.hip translation unit -> patched assembly -> one object
    each -> a single code object holding all kernels."""
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    procs = [(src, _compile_rd(src)) for src in srcs]
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append(src)
            sys.stderr.write(out)
    if failed:
        raise RuntimeError(f"device compile failed: {failed}")
    objs = []
    for src in srcs:
        name = os.path.splitext(os.path.basename(src))[0]
        text = open(os.path.join(BUILD, name + ".s")).read()
        patched, n = patch_round_mode(text)
        kernels = set(re.findall(r"\.amdhsa_kernel\s+(ccrd_\S+)", text))
        if n != len(kernels):
            raise RuntimeError(f"round-mode patch did not reach every ccrd_ kernel of {src}")
        if patched.count(".amdhsa_float_round_mode_16_64 2") != n:
            raise RuntimeError("unexpected descriptor layout in compiler output")
        pasm = os.path.join(BUILD, name + ".patched.s")
        open(pasm, "w").write(patched)
        obj = os.path.join(BUILD, name + ".o")
        _run([os.path.join(LLVM, "clang"), "-cc1as", "-triple", "amdgcn-amd-amdhsa", "-target-cpu", "gfx950",
              "-filetype", "obj", "-o", obj, pasm])
        objs.append(obj)
    hsaco = os.path.join(BUILD, "prover_rd.hsaco")
    _run([os.path.join(LLVM, "ld.lld"), "-shared", *objs, "-o", hsaco])
    return hsaco


def build_library(hsaco: str) -> None:
    blob = os.path.join(BUILD, "prover_rd_blob.S")
    with open(blob, "w") as f:
        f.write(
            "\t.section .rodata\n\t.p2align 12\n"
            "\t.globl cc_prover_rd_hsaco\ncc_prover_rd_hsaco:\n"
            f'\t.incbin "{hsaco}"\n'
            "\t.globl cc_prover_rd_hsaco_end\ncc_prover_rd_hsaco_end:\n"
            '\t.section .note.GNU-stack,"",@progbits\n'
        )
    host_dir = os.path.join(CSRC, "host")
    host_srcs = [os.path.join(host_dir, f) for f in sorted(os.listdir(host_dir)) if f.endswith(".cpp")]
    # Host units with device globals (noalloc.hip) are ordinary HIP code:
    # hipcc embeds their device part and registers it with the runtime.
    hip_objs = []
    for f in sorted(os.listdir(host_dir)):
        if f.endswith(".hip"):
            obj = os.path.join(BUILD, "host_" + os.path.splitext(f)[0] + ".o")
            _run(["hipcc", *DEVICE_FLAGS, "-fPIC", f"-I{ROOT}/include", f"-I{host_dir}", "-c", "-o", obj,
                  os.path.join(host_dir, f)])
            hip_objs.append(obj)
    tmp = LIB + ".tmp"
    _run(["g++", *HOST_FLAGS, f"-I{ROOT}/include", f"-I{CSRC}/host", "-shared", "-o", tmp, *host_srcs, *hip_objs,
          blob, f"-L{ROCM}/lib", "-lamdhip64", f"-Wl,-rpath,{ROCM}/lib", "-lpthread"])
    os.replace(tmp, LIB)


def build_prover_exe() -> None:
    src = os.path.join(CSRC, "prover_main.cpp")
    if not os.path.exists(src):
        return
    _run(["g++", *HOST_FLAGS, f"-I{ROOT}/include", f"-I{CSRC}/host", "-o", PROVER_EXE, src,
          f"-L{PKG}", "-lcirclecover_amd", "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{ROCM}/lib",
          f"-L{ROCM}/lib", "-lamdhip64"])


def build_oracle() -> None:
    """The CPU checker used by tests and the CPU baseline (never by the product)."""
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


def build_all() -> None:
    if shutil.which("hipcc") is None:
        raise RuntimeError("hipcc not found")
    hsaco = build_rd_module()
    build_library(hsaco)
    build_prover_exe()
    build_oracle()


if __name__ == "__main__":
    build_all()
    print("built", LIB)
