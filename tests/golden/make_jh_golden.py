"""Writes tests/golden/jh_rec709_16.coeff: the Jakob-Hanika RGB-to-spectrum
table (Rec.709 target, 16 cells per axis) produced by the reference's own
generator, tools/JakobHanikaLUTGen.cpp, compiled from the reference sources by
oracle/ref/Makefile into oracle/_ref/jh_lut_gen.  Run in a checkout that has
the reference tree and the built oracle/_ref:

    python tests/golden/make_jh_golden.py
"""
from __future__ import annotations

import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
OUT = Path(__file__).resolve().parent / "jh_rec709_16.coeff"


def main():
    gen = ROOT / "oracle" / "_ref" / "jh_lut_gen"
    r = subprocess.run([str(gen), "--target", "rec709", "--resolution", "16", "--output", str(OUT)],
                       capture_output=True, text=True)
    # The generator exits 2 when more than 1% of cells stay above its
    # tolerance, which the reference's own table also does; the file is written.
    if r.returncode not in (0, 2) or not OUT.exists():
        sys.exit(f"generator failed: {r.returncode}\n{r.stdout}\n{r.stderr}")
    print(r.stdout.strip().splitlines()[-2])


if __name__ == "__main__":
    main()
