"""Build libsplatt_adatm.so (host C + HIP kernels for gfx950) and the
`splatt` command-line tool, in-tree under adatm_amd/lib/ and adatm_amd/bin/.

Host C is compiled by gcc (C11, OpenMP), kernels by hipcc for gfx950, and the
shared library is linked by hipcc so that the HIP runtime comes with it.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_obj"
LIBDIR = PKG / "lib"
BINDIR = PKG / "bin"
LIBNAME = "libsplatt_adatm.so"

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = os.environ.get("ADATM_OFFLOAD_ARCH", "gfx950")

CFLAGS = [
    "-O3", "-std=gnu11", "-fPIC", "-fopenmp", "-Wall", "-Wno-unused-function",
    "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include", f"-I{ROOT / 'include'}",
]
HIPFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 math stays off: compile every kernel without v_pk_*_f32
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT / 'include'}",
]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip():
        print(out, file=sys.stderr)


def _hipcc() -> str:
    p = shutil.which("hipcc") or f"{ROCM}/bin/hipcc"
    if not Path(p).exists():
        raise RuntimeError("hipcc not found: the HIP toolchain is required to build")
    return p


def sources() -> tuple[list[Path], list[Path]]:
    c = sorted((CSRC / "host").glob("*.c"))
    c = [p for p in c if p.name != "cli.c"]
    h = sorted((CSRC / "hip").glob("*.hip"))
    return c, h


def build(verbose: bool = False, jobs: int = 8) -> Path:
    """Compile everything; always rebuilds (a fresh compile is the build check)."""
    hipcc = _hipcc()
    OBJ.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    BINDIR.mkdir(exist_ok=True)
    csrcs, hsrcs = sources()
    steps = []
    objs = []
    for s in csrcs:
        o = OBJ / (s.stem + ".c.o")
        steps.append(["gcc", *CFLAGS, "-c", str(s), "-o", str(o)])
        objs.append(o)
    for s in hsrcs:
        o = OBJ / (s.stem + ".hip.o")
        steps.append([hipcc, *HIPFLAGS, "-c", str(s), "-o", str(o)])
        objs.append(o)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, steps))
    lib = LIBDIR / LIBNAME
    tmp = LIBDIR / (LIBNAME + ".tmp")
    _run([hipcc, "-shared", "-o", str(tmp), *map(str, objs), "-lgomp", "-lm", "-ldl",
          f"--offload-arch={ARCH}"])
    os.replace(tmp, lib)
    cli = CSRC / "host" / "cli.c"
    if cli.exists():
        exe = BINDIR / "splatt"
        _run(["gcc", *CFLAGS, str(cli), "-o", str(exe), f"-L{LIBDIR}",
              "-lsplatt_adatm", "-Wl,-rpath,$ORIGIN/../lib", "-lm"])
    # the C program replaying the reference drivers' call sequences (tests)
    replay = ROOT / "tests" / "boundary_replay.c"
    if replay.exists():
        _run(["gcc", *CFLAGS, str(replay), "-o", str(BINDIR / "boundary_replay"),
              f"-L{LIBDIR}", "-lsplatt_adatm", "-Wl,-rpath,$ORIGIN/../lib", "-lm"])
    if verbose:
        print(f"built {lib}")
    return lib


if __name__ == "__main__":
    build(verbose=True)
