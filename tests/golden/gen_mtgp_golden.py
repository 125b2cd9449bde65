"""Writes tests/golden/mtgp_ref_words.npz from the reference's own vendored
MTGP32 C library (src/partfilt/opencl/mtgp/mtgp32-fast.c and the 200-set
parameter table), compiled where it lies by `make -C oracle ref` into
oracle/_ref/mtgp_ref.  Run in a container that has the reference tree:

    make -C oracle ref && python tests/golden/gen_mtgp_golden.py

Contents (all uint32):
  init_<seed>      [6, 351]   mtgp32_init_state for sets SETS_INIT, seed+set+1
  f32_<seed>       [200, 16]  first 16 close1_open2 float bit patterns, set i, seed+i+1
  u32_1234         [4, 4096]  first 4096 tempered uint32 words, sets 0..3, seed 1234+i+1
  f32long_1234     [4, 4096]  first 4096 float words, sets 0..3, seed 1234+i+1
  buf16x256_step1         [96, 768]  the first-step buffer of the 16x256 robot-arm /
                              unicycle small configuration, seed 1234 (each
                              stream's first 768 float words)
"""
import os
import subprocess
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(HERE, "..", "..", "oracle", "_ref", "mtgp_ref")
SEEDS = [1234, 42, 7, 99, 0xF000000D]
SETS_INIT = [0, 1, 2, 95, 96, 199]


def run(*args):
    out = subprocess.run([REF, *map(str, args)], check=True, capture_output=True, text=True).stdout
    return np.array([int(x, 16) for x in out.split()], dtype=np.uint32)


def main():
    data = {}
    for s in SEEDS:
        data[f"init_{s}"] = np.stack([run("init", (s + i + 1) & 0xFFFFFFFF, i) for i in SETS_INIT])
        data[f"f32_{s}"] = np.stack([run("f32", (s + i + 1) & 0xFFFFFFFF, i, 16) for i in range(200)])
    data["u32_1234"] = np.stack([run("u32", 1234 + i + 1, i, 4096) for i in range(4)])
    data["f32long_1234"] = np.stack([run("f32", 1234 + i + 1, i, 4096) for i in range(4)])
    data["buf16x256_step1"] = np.stack([run("f32", 1234 + i + 1, i, 768) for i in range(96)])
    data["sets_init"] = np.array(SETS_INIT, dtype=np.uint32)
    data["seeds"] = np.array(SEEDS, dtype=np.uint32)
    np.savez_compressed(os.path.join(HERE, "mtgp_ref_words.npz"), **data)


if __name__ == "__main__":
    main()
