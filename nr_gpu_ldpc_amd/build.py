#!/usr/bin/env python3
"""Build recipe for libnrldpc.so (HIP for gfx950) and the C++ driver executable.

    python nr_gpu_ldpc_amd/build.py        # (a script: importing the package needs the library)

Every source under csrc/ is compiled with hipcc for gfx950 into _build/, then linked into
nr_gpu_ldpc_amd/lib/libnrldpc.so (kept in-tree so it travels with the repository snapshot).
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.environ.get("NRLDPC_BUILD_DIR") or os.path.join(HERE, "_build")
LIBDIR = os.environ.get("NRLDPC_LIB_DIR") or os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libnrldpc.so")
ROOT = os.path.dirname(HERE)

ARCH = os.environ.get("NRLDPC_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",              # every float expression rounds exactly as written
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-command-line-argument",
    f"-I{os.path.join(ROOT, 'include')}",
    f"-I{CSRC}",
] + os.environ.get("NRLDPC_EXTRA_FLAGS", "").split()

SOURCES = ["mixed_k_dup_learned.hip", "mixed_k_slim_learned.hip", "decoder_pair.hip", "mixed_k_dup_plain.hip",
           "mixed_k_slim_plain.hip", "decoder.hip", "decoder_bg1.hip", "decoder_bg2.hip", "decoder_qc.hip", "decoder_slim_bg1.hip", "decoder_slim_bg2.hip",
           "decoder_slim_qc.hip", "decode_tuning.cpp", "tune.cpp", "channel.hip", "learned.hip", "api.cpp", "decode.cpp", "code.cpp",
           "encoder.cpp", "noms_api.cpp", "compat.cpp", "mixed_plan.cpp", "mixed.hip",
           "tb_host.cpp", "tb.hip"]
DRIVER_SRC = os.path.join(HERE, "driver", "ldpc_ber.cpp")
DRIVER = os.path.join(HERE, "bin", "ldpc_ber")
# host programs the GPU tests run (built here, so that no test compiles on the GPU machine)
TEST_PROGRAMS = {os.path.join(ROOT, "tests", "compat", "c_entry.cpp"): os.path.join(ROOT, "tests", "compat", "_bin",
                                                                                      "c_entry")}


def _newest_header():
    hs = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".h", ".inc"))]
    hs += [os.path.join(ROOT, "include", f) for f in os.listdir(os.path.join(ROOT, "include"))]
    return max(os.path.getmtime(h) for h in hs)


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    # NRLDPC_INCREMENTAL=1 (development only) keeps objects newer than their source and every header
    if os.environ.get("NRLDPC_INCREMENTAL") == "1" and os.path.exists(obj) and \
            os.path.getmtime(obj) > max(os.path.getmtime(path), _newest_header()):
        return obj, ""
    cmd = [HIPCC] + COMMON + ["-c", path, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC] + COMMON + ["-x", "hip", "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{out}")
    return obj, out


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    sources = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objs = []
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(6, os.cpu_count() or 2)) as ex:
        for obj, out in ex.map(_compile, sources):
            objs.append(obj)
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    if verbose:
        print("built", os.path.relpath(LIB, ROOT))
    build_driver(verbose)
    build_test_programs(verbose)
    return LIB


def build_driver(verbose=True):
    """The BER/FER driver executable, linked against the library with an rpath to lib/."""
    os.makedirs(os.path.dirname(DRIVER), exist_ok=True)
    tmp = DRIVER + ".tmp"
    cmd = [HIPCC, "-O2", "-std=c++17", "-Wall", f"-I{os.path.join(ROOT, 'include')}", DRIVER_SRC,
           f"-L{LIBDIR}", "-lnrldpc", "-Wl,-rpath,$ORIGIN/../lib", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("driver build failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, DRIVER)
    if verbose:
        print("built", os.path.relpath(DRIVER, ROOT))
    return DRIVER


def build_test_programs(verbose=True):
    """Host programs of the GPU tests, linked against the library (rpath to lib/)."""
    for src, exe in TEST_PROGRAMS.items():
        if not os.path.exists(src):
            continue
        os.makedirs(os.path.dirname(exe), exist_ok=True)
        tmp = exe + ".tmp"
        rel = os.path.relpath(LIBDIR, os.path.dirname(exe))
        cmd = [HIPCC, "-O1", "-std=c++17", f"-I{os.path.join(ROOT, 'include')}", src, f"-L{LIBDIR}", "-lnrldpc",
               f"-Wl,-rpath,$ORIGIN/{rel}", "-o", tmp]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"test program build failed for {src}:\n" + r.stdout + r.stderr)
        os.replace(tmp, exe)
        if verbose:
            print("built", os.path.relpath(exe, ROOT))


if __name__ == "__main__":
    try:
        build()
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
