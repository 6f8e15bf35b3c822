"""Builds libintpointsfem.so (host C++ + gfx950 HIP kernels) in-tree with hipcc.

Every translation unit is compiled with -ffp-contract=off (no fused
multiply-add anywhere: strains and branch flags must be bit-identical to the
CPU path) and without packed FP32 instructions.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libintpointsfem.so")
OBJDIR = os.path.join(HERE, "_build")

SOURCES = [
    "ipf_kernels.hip",
    "ipf_solver.hip",
    "ipf_kernels_fmt.hip",
    "ipf_assemble.hip",
    "ipf_pattern.hip",
    "ipf_learned.hip",
    "ipf_patch.hip",
    "ipf_mlp.hip",
    "ipf_device.hip",
    "TPZCompMesh.cpp",
    "TPZIntPoints.cpp",
    "ipf_frontend.cpp",
]

ARCH = os.environ.get("IPF_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]
DEVICE = [
    "--offload-arch=" + ARCH,
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
# Per-source device flags.  ipf_learned.hip: MFMA results in VGPRs (the
# inference loop applies the activation to every MFMA result; with the
# results in AGPRs each one is first copied out).
EXTRA = {
    "ipf_learned.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form=1"],
}


def _hipcc():
    for c in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.sep not in c or os.path.exists(c):
            return c
    return "hipcc"


def build(verbose=False, jobs=None):
    os.makedirs(OBJDIR, exist_ok=True)
    hipcc = _hipcc()
    objs = []
    procs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmd = [hipcc] + COMMON
        if src.endswith(".hip"):
            cmd += DEVICE + EXTRA.get(src, []) + ["-x", "hip"]
        cmd += ["-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        # hipcc's host pass does not know the device-only target feature; harmless
        text = "\n".join(l for l in text.splitlines() if "packed-fp32-ops" not in l)
        if p.returncode != 0:
            failed.append(src)
            sys.stderr.write(text + "\n")
        elif verbose and text.strip():
            print(text)
    if failed:
        raise RuntimeError("compilation failed: " + ", ".join(failed))
    tmp = LIB + ".tmp"
    cmd = [hipcc, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
