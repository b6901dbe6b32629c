"""Build recipe of libpwdft_hip.so (gfx950) -- run by __graft_entry__.build().

    python -m pwdft_develop_amd.build

compiles every csrc/*.hip and csrc/*.cpp with hipcc into
pwdft_develop_amd/libpwdft_hip.so (in-tree, so the built library travels with
the source tree).  Objects are cached under pwdft_develop_amd/_build/ and
rebuilt when the source or a header is newer.
"""
import glob
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpwdft_hip.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

ARCH = "gfx950"
CFLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 VALU ops off (v_pk_{add,mul,fma}_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-command-line-argument",
    f"-I{os.path.join(ROOT, 'include')}", f"-I{ROCM}/include",
]
LDFLAGS = ["-shared", f"--offload-arch={ARCH}", f"-L{ROCM}/lib", "-lrocblas", "-lrocsolver",
           f"-Wl,-rpath,{ROCM}/lib"]


def _headers():
    return glob.glob(os.path.join(CSRC, "*.h")) + glob.glob(os.path.join(ROOT, "include", "*.h"))


def _compile(src, force):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    newest = max([os.path.getmtime(src)] + [os.path.getmtime(h) for h in _headers()])
    if not force and os.path.exists(obj) and os.path.getmtime(obj) >= newest:
        return obj
    cmd = [HIPCC] + CFLAGS + ["-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC, "-x", "hip"] + CFLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed on {os.path.basename(src)}")
    return obj


def build(force=False, jobs=4):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, force), srcs))
    tmp = LIB + ".tmp"
    r = subprocess.run([HIPCC] + objs + LDFLAGS + ["-o", tmp], capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link of libpwdft_hip.so failed")
    os.replace(tmp, LIB)
    # C++ adapter driver used by the adapter tests (links the library above)
    drv_src = os.path.join(ROOT, "tests", "cpp", "adapter_driver.cpp")
    if os.path.exists(drv_src):
        drv = os.path.join(ROOT, "tests", "cpp", "adapter_driver")
        cmd = [HIPCC, "-O2", "-std=c++17", f"-I{ROCM}/include", drv_src, "-o", drv, f"-L{HERE}", "-lpwdft_hip",
               f"-L{ROCM}/lib", "-lamdhip64", "-Wl,-rpath,$ORIGIN/../../pwdft_develop_amd",
               f"-Wl,-rpath,{ROCM}/lib"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("adapter driver build failed")
    return LIB


if __name__ == "__main__":
    print(build(force="--force" in sys.argv))
