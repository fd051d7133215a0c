#!/usr/bin/env python3
"""Writes tests/golden/blue_noise_64x64.npy: the reference's 64x64 blue-noise
threshold map (depth/blue.cpp), read through the reference bridge built from
the reference sources (oracle/_ref/libzimgref.so, `make -C oracle ref`).

The fixture is data: the oracle takes it as an input for RANDOM dither and the
CPU tests compare the product's copy of the table with it."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from oracle_lib import ptr, c_u8_p, ref  # noqa: E402


def main():
    table = np.zeros((64, 64), np.uint8)
    ref().ref_blue_noise(ptr(table, c_u8_p))
    np.save(os.path.join(HERE, "blue_noise_64x64.npy"), table, allow_pickle=False)
    print("wrote", table.shape, int(table.sum()))


if __name__ == "__main__":
    main()
