"""Builds lible_amd/liblbx.so (HIP, gfx950) in-tree with hipcc.

Usage: python -m lible_amd.build [-j N]
The quartet kernels are instantiated by one source (csrc/instances.hip) compiled once per
instantiation unit, so the units build in parallel.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "liblbx.so")
INCLUDE = os.path.join(ROOT, "include")
SELFTEST_SRC = os.path.join(ROOT, "tests", "cpp", "ints_selftest.cpp")
SELFTEST = os.path.join(HERE, "bin", "ints_selftest")

N_UNITS = 24   # csrc/instances.hip: 0..11 FP64 forms and dense outputs, 12..23 FP32 forms
N_GRAD_UNITS = 10   # csrc/grad_instances.hip: one per lane-side main pair class
N_BLOCK_UNITS = 20  # csrc/block_instances.hip: 0..9 FP64 block J/K forms per bra class, 10..19 FP32

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-mcode-object-version=5",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-ffp-contract=fast", f"-I{INCLUDE}",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def runtime_link() -> list[str]:
    """Which HIP runtime the library binds to.  One process must hold one HIP runtime: when
    PyTorch-ROCm is installed it ships its own libamdhip64.so (and HSA runtime), and a second
    runtime from /opt/rocm in the same process cannot open the device.  The library is therefore
    linked against the runtime torch carries when torch is present (recorded as the unversioned
    libamdhip64.so, which the loader shares with torch, or resolves to /opt/rocm/lib when torch is
    absent at run time), and against /opt/rocm's runtime otherwise."""
    try:
        import importlib.util
        spec = importlib.util.find_spec("torch")
        tlib = os.path.join(os.path.dirname(spec.origin), "lib") if spec and spec.origin else None
    except Exception:
        tlib = None
    if tlib and os.path.exists(os.path.join(tlib, "libamdhip64.so")):
        # rocBLAS / rocSOLVER (density-fitted J/K) from the same directory, built against the
        # same runtime
        return ["-no-hip-rt", f"-L{tlib}", "-l:libamdhip64.so", "-l:librocblas.so", "-l:librocsolver.so",
                f"-Wl,-rpath,{tlib}"]
    return ["-L/opt/rocm/lib", "-lrocblas", "-lrocsolver"]


def _big_stack():
    """The optimiser recurses over the control-flow graph of the large, fully unrolled quartet
    kernels deeply enough to overflow the default 8 MiB stack; give the compiler an unlimited
    one."""
    import resource
    try:
        resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))
    except (ValueError, OSError):
        soft, hard = resource.getrlimit(resource.RLIMIT_STACK)
        resource.setrlimit(resource.RLIMIT_STACK, (hard, hard))


def _compile(src: str, obj: str, defines: list[str]) -> tuple[str, float]:
    t = time.time()
    cmd = [hipcc(), "-x", "hip", "-c", src, "-o", obj] + FLAGS + defines
    r = subprocess.run(cmd, capture_output=True, text=True, preexec_fn=_big_stack)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj, time.time() - t


def _stale(src: str, obj: str) -> bool:
    """True unless obj is newer than src, every header and this script."""
    if not os.path.exists(obj):
        return True
    deps = [src, os.path.abspath(__file__)]
    for d in (CSRC, INCLUDE):
        for base, _, files in os.walk(d):
            deps += [os.path.join(base, f) for f in files if f.endswith((".h", ".hpp"))]
    return os.path.getmtime(obj) < max(os.path.getmtime(p) for p in deps)


def build(jobs: int | None = None, verbose: bool = True, incremental: bool = False) -> str:
    """Compiles every unit and links the library (incremental: only units older than their
    sources or any header)."""
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or max(1, min(os.cpu_count() or 4, 8))
    work = [
        (os.path.join(CSRC, "host_setup.cpp"), os.path.join(BUILD, "host_setup.o"), []),
        (os.path.join(CSRC, "ecoef.hip"), os.path.join(BUILD, "ecoef.o"), []),
        (os.path.join(CSRC, "capi.hip"), os.path.join(BUILD, "capi.o"), []),
        (os.path.join(CSRC, "ints_api.cpp"), os.path.join(BUILD, "ints_api.o"), []),
        (os.path.join(CSRC, "d1.hip"), os.path.join(BUILD, "d1.o"), []),
        (os.path.join(CSRC, "ri.hip"), os.path.join(BUILD, "ri.o"), []),
        (os.path.join(CSRC, "generic.hip"), os.path.join(BUILD, "generic.o"), []),
    ]
    for u in range(N_UNITS):
        work.append((os.path.join(CSRC, "instances.hip"), os.path.join(BUILD, f"inst_{u}.o"),
                     [f"-DLBX_UNIT={u}"]))
    for u in range(N_BLOCK_UNITS):
        work.append((os.path.join(CSRC, "block_instances.hip"), os.path.join(BUILD, f"block_{u}.o"),
                     [f"-DLBX_BUNIT={u}"]))
    for u in range(N_GRAD_UNITS):
        work.append((os.path.join(CSRC, "grad_instances.hip"), os.path.join(BUILD, f"grad_{u}.o"),
                     [f"-DLBX_GUNIT={u}"]))
    # heaviest units (highest bra class) first
    work.sort(key=lambda w: -(int(w[2][0].split("=")[1]) % 12) if w[2] else 1)
    t0 = time.time()
    todo = [w for w in work if not incremental or _stale(w[0], w[1])]
    with cf.ThreadPoolExecutor(jobs) as ex:
        futs = [ex.submit(_compile, *w) for w in todo]
        for f in cf.as_completed(futs):
            obj, dt = f.result()
            if verbose:
                print(f"  {os.path.basename(obj)}  {dt:.1f}s", flush=True)
    objs = [w[1] for w in work]
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "-shared", "-o", tmp, "--offload-arch=gfx950", "-fPIC"] + objs + runtime_link()
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    build_selftest()
    if verbose:
        print(f"built {OUT} and {SELFTEST} in {time.time() - t0:.1f}s", flush=True)
    return OUT


def build_selftest() -> str:
    """The C++ known-answer program of the lible::ints interface (tests/cpp/ints_selftest.cpp),
    a plain host program linked against liblbx.so."""
    os.makedirs(os.path.dirname(SELFTEST), exist_ok=True)
    cmd = ["g++", "-O2", "-std=c++17", f"-I{INCLUDE}", SELFTEST_SRC, "-o", SELFTEST,
           f"-L{HERE}", "-l:liblbx.so", "-Wl,-rpath,$ORIGIN/.."]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"selftest build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return SELFTEST


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("--incremental", action="store_true",
                    help="recompile only units older than their sources or any header")
    a = ap.parse_args(argv)
    build(a.j, incremental=a.incremental)


if __name__ == "__main__":
    sys.exit(main())
