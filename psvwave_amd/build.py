"""Build ``libpsvwave_amd.so`` (the HIP kernels plus the C-ABI) for gfx950.

Run as ``python -m psvwave_amd.build``.  The library is written next to this
file so that it travels with the source tree.
"""
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpsvwave_amd.so")
ARCH = os.environ.get("PSV_OFFLOAD_ARCH", "gfx950")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

SOURCES = ["psv_config.cpp", "psv_kernels.hip", "psv_model.hip",
           "psv_batch.hip", "psv_conv.hip", "psv_capi.cpp"]

# -ffp-contract=off: every multiply and add rounds separately, as in the
# reference's x86-64 build, so stencil results match it bit for bit.
# -packed-fp32-ops is disabled for all device code.
COMMON = [
    "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
    "-Wall", "-Wno-unused-function",
]
DEVICE = [
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc():
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found; the ROCm toolchain is required")
    return path


def build(verbose=False):
    """Compile every source (in parallel) and link the shared library
    (always rebuilds)."""
    from concurrent.futures import ThreadPoolExecutor
    cc = hipcc()
    with tempfile.TemporaryDirectory(prefix="psv_build_") as tmp:
        cmds, objs = [], []
        for src in SOURCES:
            obj = os.path.join(tmp, os.path.splitext(src)[0] + ".o")
            lang = (["-x", "hip"] if src.endswith(".hip") else
                    ["-x", "c++", "-D__HIP_PLATFORM_AMD__",
                     f"-I{os.path.join(ROCM, 'include')}"])
            cmds.append([cc, *COMMON, *DEVICE, *lang, "-c",
                         os.path.join(CSRC, src), "-o", obj])
            objs.append(obj)
        if verbose:
            for cmd in cmds:
                print(" ".join(cmd), flush=True)
        jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", "0") or 0)
                          or os.cpu_count() or 1))
        with ThreadPoolExecutor(jobs) as pool:
            for r in pool.map(lambda c: subprocess.run(c), cmds):
                if r.returncode:
                    raise subprocess.CalledProcessError(r.returncode, r.args)
        tmp_lib = LIB + ".tmp"
        cmd = [cc, *DEVICE, "-shared", "-fPIC", "-o", tmp_lib, *objs]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(tmp_lib, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
