"""Build the C++ drop-in test programs into tests/cpp/bin/.

dropin_host  host-only checks of the C++ API layer and RHS evaluation for the
             oracle comparison (runs without a GPU);
dropin_gpu   TreESaP-style programs run through the drop-in integrators on the
             GPU.

Both are ordinary g++ C++20 programs that include include/treesap/ and link
treesap_amd/libtreesap_hip.so (found at run time through an $ORIGIN rpath).
Usage: python tests/cpp/build.py
"""
from __future__ import annotations

import concurrent.futures as cf
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
BIN = HERE / "bin"
PROGRAMS = ("dropin_host", "dropin_gpu")


def _compile(name: str) -> Path:
    out = BIN / name
    cmd = [
        "g++", "-std=c++20", "-O2", "-ffp-contract=off", "-Wall", "-Wextra",
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
