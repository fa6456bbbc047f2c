"""Build the C++ drop-in test programs into tests/cpp/bin/.

dropin_host  host-only checks of the C++ API layer and RHS evaluation for the
             oracle comparison (runs without a GPU);
dropin_gpu   TreESaP-style programs run through the drop-in integrators on the
             GPU, their right-hand sides mapped onto the library's kernels;
dropin_generic  programs whose right-hand sides the library does not carry,
             run on the stage kernel compiled from the functor;
dropin_bc    programs whose boundary functor the library does not carry, run
             on the boundary kernel compiled from the functor;
generic_timing  the throughput of that stage kernel on a (phi, Pi) functor
             (bench.py's generic leg).

The first two are ordinary g++ C++20 programs, the last three are compiled by hipcc
for gfx950; all include include/treesap/ and link
treesap_amd/libtreesap_hip.so (found at run time through an $ORIGIN rpath).
Usage: python tests/cpp/build.py
"""
from __future__ import annotations

import concurrent.futures as cf
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
BIN = HERE / "bin"
PROGRAMS = ("dropin_host", "dropin_gpu", "dropin_generic", "dropin_bc", "generic_timing")
# programs whose right-hand sides run on the stage kernel compiled from the functor
HIP_PROGRAMS = ("dropin_generic", "dropin_bc", "generic_timing")


def _hipcc() -> str:
    for cand in ("/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def _compile(name: str) -> Path:
    out = BIN / name
    if name in HIP_PROGRAMS:
        head = [_hipcc(), "--offload-arch=gfx950", "-std=c++20", "-O2", "-ffp-contract=off",
                "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall"]
    else:
        head = ["g++", "-std=c++20", "-O2", "-ffp-contract=off", "-Wall", "-Wextra"]
    cmd = [
        *head,
        f"-I{ROOT / 'include' / 'treesap'}", f"-I{ROOT / 'include'}", f"-I{HERE}",
        str(HERE / f"{name}.cpp"), "-o", str(out),
        f"-L{ROOT / 'treesap_amd'}", "-ltreesap_hip", "-Wl,-rpath,$ORIGIN/../../../treesap_amd",
    ]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"g++ failed on {name}.cpp:\n{res.stdout}\n{res.stderr}")
    return out


def build(verbose: bool = True) -> list[Path]:
    if not (ROOT / "treesap_amd" / "libtreesap_hip.so").exists():
        raise RuntimeError("build the HIP library first (treesap_amd/build.py)")
    BIN.mkdir(exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=len(PROGRAMS)) as ex:
        outs = list(ex.map(_compile, PROGRAMS))
    if verbose:
        print(f"built {', '.join(p.name for p in outs)} in {BIN}", flush=True)
    return outs


def binary(name: str) -> Path:
    """Path of a built program; builds the set if it is missing."""
    p = BIN / name
    if not p.exists():
        build(verbose=False)
    return p


if __name__ == "__main__":
    build()
    sys.exit(0)
