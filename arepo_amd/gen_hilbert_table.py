"""Derive the 24-state table of the library's 3-D Hilbert curve (run: python -m arepo_amd.gen_hilbert_table).

The curve is the one of Skilling's transpose method ("Programming the Hilbert curve", AIP Conf. Proc. 707, 381
(2004)) on b-bit coordinates: undo excess work, Gray-encode, interleave MSB first.  Each level's key digit
depends only on that level's octant and on an orientation state carried down from the levels above; this
script finds the states by evaluating the transpose method on cells of increasing depth, checks that the
transitions do not depend on the representative cell or on b, and prints the table the library uses:
entry [8 s + o] = (next state << 3) | digit, start state 0.
"""
import itertools
import random


def transpose_key(x, y, z, b):
    X = [x, y, z]
    Q = 1 << (b - 1)
    while Q > 1:
        P = Q - 1
        for i in range(3):
            if X[i] & Q:
                X[0] ^= P
            else:
                t = (X[0] ^ X[i]) & P
                X[0] ^= t
                X[i] ^= t
        Q >>= 1
    X[1] ^= X[0]
    X[2] ^= X[1]
    t = 0
    Q = 1 << (b - 1)
    while Q > 1:
        if X[2] & Q:
            t ^= Q - 1
        Q >>= 1
    X = [v ^ t for v in X]
    k = 0
    for l in range(b - 1, -1, -1):
        k = (k << 3) | (((X[0] >> l) & 1) << 2) | (((X[1] >> l) & 1) << 1) | ((X[2] >> l) & 1)
    return k


def _digit_map(prefix, b):
    """Key digit at level len(prefix)+1 of each octant inside the cell reached by `prefix` (octants from the top)."""
    out = []
    for o in range(8):
        x = y = z = 0
        for j, q in enumerate(list(prefix) + [o]):
            sh = b - 1 - j
            x |= ((q >> 2) & 1) << sh
            y |= ((q >> 1) & 1) << sh
            z |= (q & 1) << sh
        out.append((transpose_key(x, y, z, b) >> (3 * (b - 1 - len(prefix)))) & 7)
    return tuple(out)


def derive(b=30, depth=4):
    ids, rep = {}, {}
    for d in range(depth + 1):
        for pre in itertools.product(range(8), repeat=d):
            m = _digit_map(pre, b)
            if m not in ids:
                ids[m] = len(ids)
                rep[ids[m]] = list(pre)
    digit = {s: m for m, s in ids.items()}
    nxt = {s: [ids[_digit_map(rep[s] + [o], b)] for o in range(8)] for s in digit}
    for d in range(depth):  # every representative agrees
        for pre in itertools.product(range(8), repeat=d):
            s = ids[_digit_map(pre, b)]
            assert all(ids[_digit_map(list(pre) + [o], b)] == nxt[s][o] for o in range(8))
    assert ids[_digit_map([], b)] == 0
    return [(nxt[s][o] << 3) | digit[s][o] for s in range(len(digit)) for o in range(8)]


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
    for b in range(1, 31):
        for _ in range(200):
            x, y, z = (rng.getrandbits(b) for _ in range(3))
            assert table_key(tab, x, y, z, b) == transpose_key(x, y, z, b)
    print("%d states, identical to the transpose method for b = 1..30" % (len(tab) // 8))
    print(",\n".join("    " + ", ".join("%3d" % v for v in tab[i:i + 16]) for i in range(0, len(tab), 16)))
