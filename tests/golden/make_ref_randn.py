"""Writes tests/golden/ref_randn.json: the first draws of the reference's own
Util::randn (src/util.cpp:10-18) after srand(seed), as IEEE-754 bit patterns.

They come from oracle/_ref/ref_randn, which oracle/Makefile compiles from the
reference's src/util.cpp (run `make -C oracle` first, in the container that
holds the reference sources).  Usage: python tests/golden/make_ref_randn.py
"""
import json
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "..", "..", "oracle", "_ref", "ref_randn")
SEEDS = (20000905, 1)
COUNT = 1000


def main():
    out = {"source": "src/util.cpp:10-18 Util::randn after srand(seed); 16 hex digits of the double",
           "count": COUNT, "draws": {}}
    for s in SEEDS:
        txt = subprocess.run([BIN, str(s), str(COUNT)], check=True, capture_output=True, text=True).stdout
        out["draws"][str(s)] = txt.split()
    with open(os.path.join(HERE, "ref_randn.json"), "w") as f:
        json.dump(out, f, indent=0)


if __name__ == "__main__":
    main()
