"""Build the MI355X (gfx950) shared library ``librpa_amd/librpa_mi355x.so``.

Every HIP and C++ source under ``librpa_amd/csrc`` is compiled with hipcc for
gfx950 and linked into one C-ABI library exporting LibRPA's API
(include/librpa_*.h) and this build's extensions (include/librpa_amd.h).
The minimax coefficient table (librpa_amd/data/minimax_coeffs.txt) is
embedded as a string constant.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "librpa_mi355x.so")
ARCH = os.environ.get("LIBRPA_AMD_ARCH", "gfx950")
# -amdgpu-mfma-vgpr-form: MFMA results in ordinary VGPRs.  Without it the 4-wave GEMM tiles keep
# their accumulators in VGPRs across the K loop and copy all of them to AGPRs and back around the
# MFMAs of every K slice (64 + 64 v_accvgpr moves per 16 MFMAs in the 64x128 T_A tile).
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-mllvm", "-amdgpu-mfma-vgpr-form",
         "-Wno-unused-command-line-argument"]


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _gen_table():
    src = os.path.join(HERE, "data", "minimax_coeffs.txt")
    text = open(src).read()
    out = os.path.join(BUILD, "minimax_table.cpp")
    chunks = [text[i:i + 60000] for i in range(0, len(text), 60000)]
    with open(out, "w") as f:
        f.write("// generated from librpa_amd/data/minimax_coeffs.txt\n")
        f.write("extern const char kMinimaxTable[];\nconst char kMinimaxTable[] =\n")
        for c in chunks:
            f.write('R"TAB(' + c + ')TAB"\n')
        f.write(";\n")
    return out


def _compile(hipcc, src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [hipcc] + FLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    hipcc = _hipcc()
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".cpp", ".hip")))
    srcs.append(_gen_table())
    jobs = min(len(srcs), int(os.environ.get("MAX_JOBS", "8")))
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda s: _compile(hipcc, s), srcs))
    tmp = LIB + ".tmp"
    cmd = [hipcc, "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-lrocsolver", "-lrocblas"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
