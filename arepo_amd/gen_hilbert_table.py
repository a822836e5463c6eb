"""Derive the 24-state table of AREPO's 3-D Peano-Hilbert curve (run: python -m arepo_amd.gen_hilbert_table).

AREPO describes its curve twice in peano.c: by orientation/digit tables for the key (peano_hilbert_key,
:334-390) and by the octant order of 24 rotations, two rotation maps and a sense flip for the inverse
(peano_hilbert_key_inverse, :393-496).  This script starts from the second description, restated below:
an orientation state is a (rotation, sense) pair; at each level the octant o = x bit << 2 | y bit << 1 | z bit
falls in quadrant q = QUADRANT[rotation][o], the key digit is q (sense +1) or 7 - q (sense -1), and the
rotation is stepped ROTX[q] times through ROTX_MAP and ROTY[q] times through ROTY_MAP while the sense takes
the factor SENSE[q].  The states reachable from (0, +1) are numbered in breadth-first order (start state 0),
and the table the library uses is printed: entry [8 s + o] = (next state << 3) | digit, one byte each.

The library's keys are checked against the oracle's restatement of peano_hilbert_key (the first
description) by tests/test_abi.py; the oracle's key and inverse are checked against each other by
tests/test_oracle.py.
"""
import random

# peano.c:393-424, quadrants[24][2][2][2], flattened as [rotation][x bit][y bit][z bit]
QUADRANT = [
    0, 7, 1, 6, 3, 4, 2, 5, 7, 4, 6, 5, 0, 3, 1, 2, 4, 3, 5, 2, 7, 0, 6, 1, 3, 0, 2, 1, 4, 7, 5, 6,
    1, 0, 6, 7, 2, 3, 5, 4, 0, 3, 7, 4, 1, 2, 6, 5, 3, 2, 4, 5, 0, 1, 7, 6, 2, 1, 5, 6, 3, 0, 4, 7,
    6, 1, 7, 0, 5, 2, 4, 3, 1, 2, 0, 3, 6, 5, 7, 4, 2, 5, 3, 4, 1, 6, 0, 7, 5, 6, 4, 7, 2, 1, 3, 0,
    7, 6, 0, 1, 4, 5, 3, 2, 6, 5, 1, 2, 7, 4, 0, 3, 5, 4, 2, 3, 6, 7, 1, 0, 4, 7, 3, 0, 5, 6, 2, 1,
    6, 7, 5, 4, 1, 0, 2, 3, 7, 0, 4, 3, 6, 1, 5, 2, 0, 1, 3, 2, 7, 6, 4, 5, 1, 6, 2, 5, 0, 7, 3, 4,
    2, 3, 1, 0, 5, 4, 6, 7, 3, 4, 0, 7, 2, 5, 1, 6, 4, 5, 7, 6, 3, 2, 0, 1, 5, 2, 6, 1, 4, 3, 7, 0,
]
# peano.c:426-433: rotation maps, rotation steps per quadrant, sense factor per quadrant
ROTX_MAP = [4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 0, 1, 2, 3, 17, 18, 19, 16, 23, 20, 21, 22]
ROTY_MAP = [1, 2, 3, 0, 16, 17, 18, 19, 11, 8, 9, 10, 22, 23, 20, 21, 14, 15, 12, 13, 4, 5, 6, 7]
ROTX = [3, 0, 0, 2, 2, 0, 0, 1]
ROTY = [0, 1, 1, 2, 2, 3, 3, 0]
SENSE = [-1, -1, -1, +1, +1, -1, -1, -1]


def step(state, octant):
    """One level of the curve: (digit, next state) of (rotation, sense) for octant x<<2 | y<<1 | z."""
    rotation, sense = state
    q = QUADRANT[8 * rotation + octant]
    digit = q if sense == 1 else 7 - q
    for _ in range(ROTX[q]):
        rotation = ROTX_MAP[rotation]
    for _ in range(ROTY[q]):
        rotation = ROTY_MAP[rotation]
    return digit, (rotation, sense * SENSE[q])


def quadrant_key(x, y, z, b):
    """The b-level key walked directly through the (rotation, sense) states."""
    state, k = (0, 1), 0
    for l in range(b - 1, -1, -1):
        d, state = step(state, (((x >> l) & 1) << 2) | (((y >> l) & 1) << 1) | ((z >> l) & 1))
        k = (k << 3) | d
    return k


def derive():
    ids, order = {(0, 1): 0}, [(0, 1)]
    i = 0
    while i < len(order):
        for o in range(8):
            _, nxt = step(order[i], o)
            if nxt not in ids:
                ids[nxt] = len(order)
                order.append(nxt)
        i += 1
    assert len(order) <= 32, "states must fit 5 bits"
    tab = []
    for s in order:
        for o in range(8):
            d, nxt = step(s, o)
            tab.append((ids[nxt] << 3) | d)
    return tab


def table_key(tab, x, y, z, b):
    s, k = 0, 0
    for l in range(b - 1, -1, -1):
        e = tab[8 * s + ((((x >> l) & 1) << 2) | (((y >> l) & 1) << 1) | ((z >> l) & 1))]
        k = (k << 3) | (e & 7)
        s = e >> 3
    return k


if __name__ == "__main__":
    tab = derive()
    rng = random.Random(1)
    for b in range(1, 43):
        for _ in range(200):
            x, y, z = (rng.getrandbits(b) for _ in range(3))
            assert table_key(tab, x, y, z, b) == quadrant_key(x, y, z, b)
    print("%d states reachable; the packed table equals the state walk for b = 1..42" % (len(tab) // 8))
    print(",\n".join("    " + ", ".join("%3d" % v for v in tab[i:i + 16]) for i in range(0, len(tab), 16)))
