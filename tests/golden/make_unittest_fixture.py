"""Regenerates tests/golden/unittest_inputs.npz.

Compiles gen_unittest_inputs.cpp with g++ (libstdc++, -O0 so the key-slice shift is
the scalar x86 shift) and stores its random draws: the key/value lengths of the
reference unit-test generator for the four instantiated length ranges, the two
std::shuffle permutations, and a checksum of every generated key slice.

Usage:  python tests/golden/make_unittest_fixture.py
"""
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    with tempfile.TemporaryDirectory() as td:
        exe = os.path.join(td, "gen")
        subprocess.check_call(["g++", "-O0", "-std=c++17", "-o", exe,
                               os.path.join(HERE, "gen_unittest_inputs.cpp")])
        out = subprocess.check_output([exe]).decode()
    arrays = {}
    for line in out.strip().splitlines():
        parts = line.split()
        name, n = parts[0], int(parts[1])
        vals = np.array([int(x) for x in parts[2:]], dtype=np.uint32)
        assert vals.size == n, name
        arrays[name] = vals
    np.savez_compressed(os.path.join(HERE, "unittest_inputs.npz"), **arrays)
    print("wrote", sorted(arrays))


if __name__ == "__main__":
    main()
