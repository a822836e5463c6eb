"""Writes the small golden fixtures used by the CPU tests (run: python tests/golden/make_golden.py).

Every vector here comes from an independent library (scipy / numpy), not
from the oracle:
  * shortrange_tables.npz: AREPO's short-range tables (forcetree_walk.c:72-85)
    evaluated with scipy.special.erfc in float64 and rounded to float32.
"""
import os

import numpy as np
from scipy.special import erfc

HERE = os.path.dirname(os.path.abspath(__file__))


def shortrange(rcut=4.5, ntab=127):
    u = (rcut / 2.0) / ntab * np.arange(ntab + 1)
    tp = (-erfc(u)).astype(np.float32)
    with np.errstate(divide="ignore", invalid="ignore"):
        tf = (erfc(u) + 2.0 * u / np.sqrt(np.pi) * np.exp(-u * u) - 1.0) / (u * u)
    tf[0] = 0.0
    return tf.astype(np.float32), tp


def main():
    tf, tp = shortrange()
    np.savez(os.path.join(HERE, "shortrange_tables.npz"), tab_force=tf, tab_pot=tp, rcut=4.5, ntab=127)
    print("wrote shortrange_tables.npz")


if __name__ == "__main__":
    main()
