"""Build liblvm_gpu.so (the HIP kernels + C ABI) in-tree for gfx950.

Run ``python -m lambda_vm_amd.build`` (or ``__graft_entry__.build()``).
hipcc cross-compiles for gfx950 without a GPU present.
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "liblvm_gpu.so")

SOURCES = ["ntt.hip", "merkle.hip", "misc.hip", "fri.hip", "ood.hip", "logup.hip", "compose.hip", "grinding.hip", "context.cpp", "compose_jit.cpp", "transcript.cpp", "capi.cpp", "batch.cpp", "rounds.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("LVM_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    # packed fp32 instructions are kept out of every code object
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]


def _embed_field_source(objdir: str) -> str:
    """C++ file holding the text of field.hpp and compose_common.hpp for the
    run-time compiled constraint kernels (include lines dropped: the run-time
    compiler provides the HIP device environment)."""
    body = ""
    for name in ("field.hpp", "compose_common.hpp"):
        text = open(os.path.join(CSRC, name)).read()
        body += "\n".join(line for line in text.splitlines()
                          if not line.startswith("#include") and not line.startswith("#pragma once")) + "\n"
    out = os.path.join(objdir, "jit_field_src.cpp")
    src = ('// generated from csrc/field.hpp and csrc/compose_common.hpp by build.py\n'
           'extern const char *const lvm_jit_field_src = R"LVMFIELD(' + body + ')LVMFIELD";\n')
    if not os.path.exists(out) or open(out).read() != src:
        with open(out, "w") as f:
            f.write(src)
    return out


def build(verbose: bool = False) -> str:
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    procs = []
    sources = [os.path.join(CSRC, s) for s in SOURCES] + [_embed_field_source(objdir)]
    for src in sources:
        s = os.path.basename(src)
        obj = os.path.join(objdir, s + ".o")
        lang = ["-x", "hip"] if s.endswith(".cpp") else []
        cmd = [HIPCC, *FLAGS, *lang, "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((s, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
    failed = []
    for s, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append((s, out.decode(errors="replace")))
        elif verbose and out:
            print(out.decode(errors="replace"))
    if failed:
        for s, msg in failed:
            sys.stderr.write(f"--- {s} ---\n{msg}\n")
        raise RuntimeError("hipcc failed for: " + ", ".join(s for s, _ in failed))
    tmp = LIB + ".tmp"
    link = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lhiprtc"]
    if verbose:
        print(" ".join(link), flush=True)
    subprocess.run(link, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
