#!/usr/bin/env python3
"""Writes KCs/gradient_spread.json: the rounding spread of the reference formulation on the KCs
gradient matrices kept as goldens, i.e. max |oracle - second-rounding oracle| per matrix
(oracle/liborc.so vs oracle/liborc_nofma.so; about 4 minutes on 8 cores).  The GPU tests bound the
engine's distance to these goldens by a multiple of it.

    python tests/golden/make_kcs_spread.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from libgrpp_amd import io  # noqa: E402
import oracle  # noqa: E402

MATRICES = {"arep_grad_1x": (0, 1, 0), "so_z_grad_1y": (3, 1, 1)}   # (operator, atom, coordinate)


def main():
    sysm = io.load_system(os.path.join(HERE, "KCs"))
    a, _ = oracle.oracle().gradients(sysm)
    b, _ = oracle.oracle_nofma().gradients(sysm)
    out = {}
    for name, (op, at, c) in MATRICES.items():
        k = 3 * at + c
        out[name] = float(np.abs(np.tril(a[op, k]) - np.tril(b[op, k])).max())
    with open(os.path.join(HERE, "KCs", "gradient_spread.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(out)


if __name__ == "__main__":
    main()
