"""In-tree build of the native libraries.

- ``althea_amd/libalthea_amd.so``: the HIP kernels and C ABI, gfx950 only.
- ``althea_amd/libalthea_synth.so``: host generators of the synthetic inputs.
- ``oracle/liboracle.so``: the FP64 CPU oracle (test infrastructure).

Every call recompiles from source (no up-to-date checks).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "althea_amd"
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
ORACLE = ROOT / "oracle"

HIP_SOURCES = ["ibl_kernels.hip", "resolve_kernels.hip", "ssr_kernels.hip", "shadow_kernels.hip", "gbuffer_kernels.hip",
               "mlp_kernels.hip",
               "batched_kernels.hip", "althea_capi.cpp", "hdr_io.cpp"]
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", str(ROOT / "include"), "-Wall",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    lines = [ln for ln in proc.stdout.splitlines() if "packed-fp32-ops" not in ln]
    if proc.returncode != 0:
        raise RuntimeError("build step failed:\n  " + " ".join(cmd) + "\n" + "\n".join(lines[-60:]))
    if lines and os.environ.get("ALTHEA_BUILD_VERBOSE"):
        print("\n".join(lines), file=sys.stderr)


def build_hip() -> Path:
    hipcc = _hipcc()
    OBJ.mkdir(exist_ok=True)
    sources = [CSRC / s for s in HIP_SOURCES if (CSRC / s).exists()]
    objs = [OBJ / (s.stem + ".o") for s in sources]

    def compile_one(args):
        src, obj = args
        _run([hipcc, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])

    with cf.ThreadPoolExecutor(max_workers=min(8, len(sources))) as ex:
        list(ex.map(compile_one, zip(sources, objs)))
    out = PKG / "libalthea_amd.so"
    tmp = PKG / "libalthea_amd.so.tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def build_synth() -> Path:
    out = PKG / "libalthea_synth.so"
    tmp = PKG / "libalthea_synth.so.tmp"
    _run(["g++", "-O2", "-std=c++17", "-fopenmp", "-fPIC", "-shared", "-ffp-contract=off",
          str(CSRC / "synth.cpp"), "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def build_oracle() -> Path:
    """The FP64 oracle (no contraction, so every product and sum rounds as
    written) and the FP32 port of the resolve (the CPU baseline: -O3 with AVX2
    and FMA, which every x86-64 host of this project has, and no contraction
    beyond the explicit fmaf of the kernel's formulation) in one library."""
    OBJ.mkdir(exist_ok=True)
    common = ["g++", "-std=c++17", "-fopenmp", "-fPIC", "-ffp-contract=off"]
    o1, o2 = OBJ / "oracle_capi.o", OBJ / "port_resolve.o"
    _run([*common, "-O2", "-c", str(ORACLE / "oracle_capi.cpp"), "-o", str(o1)])
    _run([*common, "-O3", "-mavx2", "-mfma", "-c", str(ORACLE / "port_resolve.cpp"), "-o", str(o2)])
    out = ORACLE / "liboracle.so"
    tmp = ORACLE / "liboracle.so.tmp"
    _run(["g++", "-fopenmp", "-shared", str(o1), str(o2), "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def build_facade_demo() -> Path:
    """C++ consumer of include/althea/ImageBasedLighting.hpp (run by the GPU tests)."""
    out_dir = ROOT / "tests" / "cpp" / "build"
    out_dir.mkdir(parents=True, exist_ok=True)
    out = out_dir / "facade_demo"
    _run([_hipcc(), "-std=c++17", "-O2", "-I", str(ROOT / "include"), str(ROOT / "tests" / "cpp" / "facade_demo.cpp"),
          "-L", str(PKG), "-lalthea_amd", "-Wl,-rpath,$ORIGIN/../../../althea_amd", "-o", str(out)])
    return out


def build_all() -> None:
    with cf.ThreadPoolExecutor(max_workers=3) as ex:
        futs = [ex.submit(build_hip), ex.submit(build_synth), ex.submit(build_oracle)]
        for f in futs:
            f.result()
    build_facade_demo()


if __name__ == "__main__":
    build_all()
    print("built:", PKG / "libalthea_amd.so", PKG / "libalthea_synth.so", ORACLE / "liboracle.so")
