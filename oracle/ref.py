"""Build the reference binaries into oracle/_ref/ when the reference tree is
present (TEST INFRASTRUCTURE ONLY; see oracle/Makefile)."""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
REF = Path(os.environ.get("LLAMA_REFERENCE", "/root/reference"))


def build_if_available() -> bool:
    if not (REF / "benchmark" / "benchmark.cc").exists():
        return False
    subprocess.run(["make", "-C", str(HERE), f"REF={REF}", "-j4"], check=True,
                   stdout=subprocess.DEVNULL)
    return True


def driver() -> Path | None:
    p = HERE / "_ref" / "ref_driver"
    return p if p.exists() else None
