"""Writes the initial-condition fixtures of the reference's example runs (data only, no code).

* ``L50n32_ics``: the cosmological box of examples/cosmo_box_gravity_only_3d/L50n32/ics (Gadget format 1,
  32,768 particles), copied as is.
* ``galaxy_ics_every12``: every 12th particle of each type of examples/isolated_galaxy_collisionless_3d/ICs/ICs
  (Gadget format 1, 300,000 particles: halo, disk, bulge), masses times 12 so that each component keeps its
  mass, written in Gadget format 1 by arepo_amd.gadget (25,000 particles).

Run from the repository root in a container where the reference is mounted at /root/reference.
"""
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from arepo_amd import gadget as G  # noqa: E402

REF = "/root/reference/examples"


def main():
    shutil.copyfile(os.path.join(REF, "cosmo_box_gravity_only_3d", "L50n32", "ics"), os.path.join(HERE, "L50n32_ics"))
    g = G.read_gadget1(os.path.join(REF, "isolated_galaxy_collisionless_3d", "ICs", "ICs"))
    keep = np.concatenate([np.flatnonzero(g.ptype == t)[::12] for t in range(G.NTYPES)])
    keep.sort()
    npart = [int((g.ptype[keep] == t).sum()) for t in range(G.NTYPES)]
    h = G.Header(**{**g.header.__dict__, "npart": npart, "npart_total": npart})
    sub = G.Snapshot(h, g.pos[keep], g.vel[keep], g.ids[keep], g.mass[keep] * 12.0, g.ptype[keep], None, b"",
                     g.float_bytes)
    G.write_gadget1(os.path.join(HERE, "galaxy_ics_every12"), sub)


if __name__ == "__main__":
    main()
