"""Builds libmdn.so (HIP for gfx950 + host C++) in-tree.

    python -m md_nucleation_amd.build
"""
from __future__ import annotations

import json
import os
import re
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
ROOT = os.path.dirname(PKG)
LIB = os.path.join(PKG, "libmdn.so")
SOURCES = ["mdn_engine.hip", "mdn_adjoint.hip", "host_sysdef.cpp", "host_run.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", "--offload-arch=gfx950",
         # packed FP32 is kept off for every kernel of this library
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result", "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]


RESOURCES = os.path.join(PKG, "libmdn.resources.json")


def kernel_resources(remarks: str) -> dict:
    """{kernel: {"vgprs", "vgpr_spill", "sgpr_spill", "scratch"}} from the
    compiler's kernel-resource-usage remarks"""
    out, name = {}, None
    for line in remarks.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            name = m.group(1)
            out[name] = {}
            continue
        m = re.search(r"remark: +(VGPRs|VGPRs Spill|SGPRs Spill|ScratchSize \[bytes/lane\]): (\d+)", line)
        if m and name:
            key = {"VGPRs": "vgprs", "VGPRs Spill": "vgpr_spill", "SGPRs Spill": "sgpr_spill",
                   "ScratchSize [bytes/lane]": "scratch"}[m.group(1)]
            out[name][key] = int(m.group(2))
    return out


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-Rpass-analysis=kernel-resource-usage"] + srcs + ["-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stderr[-6000:])
    if verbose and r.stderr:
        print(r.stderr[-3000:])
    os.replace(tmp, LIB)
    # the register budget of every kernel, for the build checks in tests/
    with open(RESOURCES, "w") as f:
        json.dump(kernel_resources(r.stderr), f, indent=1, sort_keys=True)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
