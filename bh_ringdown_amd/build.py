"""Build libteuk (HIP, gfx950) and the teuk_run driver in-tree.

    python -m bh_ringdown_amd.build

Outputs bh_ringdown_amd/lib/libteuk.so and bh_ringdown_amd/bin/teuk_run.
"""
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

# -packed-fp32-ops: no packed FP32 instructions in any kernel.
# -ffp-contract=off: every sum and product is rounded as written, so the
# device reproduces the host FP64 evaluation order bit for bit.
HIPFLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-result", f"-I{INCLUDE}", f"-I{CSRC}"]

# The stage kernels are built once per theta-chunk length, one object each
# (the longest compiles first).
LIB_SOURCES = ["teuk_stage_r8.hip", "teuk_stage_r4.hip", "teuk_stage_r2.hip", "teuk_kernels.hip", "teuk_so2.hip",
               "teuk_api.hip", "teuk_host.cpp", "teuk_traj.cpp"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        print(out, flush=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}")


def build_lib():
    """Each source compiled on its own (in parallel), then one link."""
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(os.path.join(PKG, "lib"), exist_ok=True)
    objdir = os.path.join(PKG, "lib", "obj")
    os.makedirs(objdir, exist_ok=True)
    out = os.path.join(PKG, "lib", "libteuk.so")
    tmp = out + ".tmp"
    objs = [os.path.join(objdir, os.path.splitext(src)[0] + ".o") for src in LIB_SOURCES]
    with ThreadPoolExecutor(max_workers=min(len(LIB_SOURCES), os.cpu_count() or 1)) as ex:
        list(ex.map(lambda so: _run([HIPCC, *HIPFLAGS, "-c", "-o", so[1], os.path.join(CSRC, so[0])]),
                    zip(LIB_SOURCES, objs)))
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, out)
    return out


def build_driver():
    os.makedirs(os.path.join(PKG, "bin"), exist_ok=True)
    src = os.path.join(PKG, "driver", "teuk_main.cpp")
    if not os.path.exists(src):
        return None
    out = os.path.join(PKG, "bin", "teuk_run")
    _run(["g++", "-O2", "-std=c++17", f"-I{INCLUDE}", "-o", out, src,
          f"-L{os.path.join(PKG, 'lib')}", "-lteuk", "-Wl,-rpath,$ORIGIN/../lib"])
    return out


def main():
    build_lib()
    build_driver()


if __name__ == "__main__":
    sys.exit(main())
