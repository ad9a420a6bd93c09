"""Build libcabanamd_amd.so (HIP, gfx950) and the cabanamd command-line driver.

    python -m cabanamd_amd.build          # or: python cabanamd_amd/build.py

Every HIP source is compiled with hipcc for gfx950 only.  FP64 expressions
are evaluated as written (-ffp-contract=off); kernels that want fused
multiply-adds spell them out.  Packed FP32 instructions are disabled, as
required on these machines.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcabanamd_amd.so")
EXE = os.path.join(HERE, "cabanamd")

HIP_SOURCES = ["core.hip", "lj.hip", "neighbor.hip", "tiles.hip", "md.hip", "nnp.hip",
               "batch.hip"]
CXX_SOURCES = ["host_input.cpp", "serve.cpp"]
ARCH = os.environ.get("CMD_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
          "-I" + os.path.join(ROOT, "include")]
HIP_FLAGS = COMMON + [f"--offload-arch={ARCH}", "-munsafe-fp-atomics",
                      "-Xclang", "-target-feature", "-Xclang",
                      "-packed-fp32-ops", "-Wno-unused-result"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True)
    return r.returncode, r.stdout


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC] + HIP_FLAGS + ["-c", path, "-o", obj]
    else:
        cmd = ["g++"] + COMMON + ["-Wall", "-Wextra", "-c", path, "-o", obj]
    rc, out = _run(cmd)
    return src, obj, rc, out


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    sources = [s for s in HIP_SOURCES + CXX_SOURCES
               if os.path.exists(os.path.join(CSRC, s))]
    objs = []
    failed = []
    with cf.ThreadPoolExecutor(max_workers=min(8, len(sources))) as ex:
        for src, obj, rc, out in ex.map(_compile, sources):
            if rc != 0:
                failed.append((src, out))
            else:
                if verbose and out.strip():
                    print(out)
                objs.append(obj)
    if failed:
        for src, out in failed:
            sys.stderr.write(f"--- {src} failed ---\n{out}\n")
        raise RuntimeError("HIP build failed: " +
                           ", ".join(s for s, _ in failed))
    tmp = LIB + ".tmp"
    rc, out = _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}",
                    "-o", tmp] + objs)
    if rc != 0:
        raise RuntimeError("link failed:\n" + out)
    os.replace(tmp, LIB)
    main = os.path.join(CSRC, "main.cpp")
    if os.path.exists(main):
        rc, out = _run(["g++"] + COMMON + ["-Wall", "-Wextra", main, "-o", EXE,
                                           "-L" + HERE, "-lcabanamd_amd",
                                           "-Wl,-rpath,$ORIGIN"])
        if rc != 0:
            raise RuntimeError("cabanamd driver build failed:\n" + out)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
