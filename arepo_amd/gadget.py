"""Gadget format-1 initial conditions: the particle load of AREPO's read_ic.c for ICFormat 1.

A format-1 file is a sequence of Fortran-style records, each framed by its
byte length before and after (the SKIP/SKIP2 pairs of read_ic.c:66-73): the
256-byte header (read_ic.c:1017-1024, struct io_header), then POS, VEL and ID
for all particles with npart > 0, MASS for the types whose mass-table entry
is zero, and U for gas (the block order of the iofields enumeration that the
reader walks, read_ic.c:1123-1250).  A first record length other than 256
means the file has the other byte order (read_ic.c:1019-1021).  Particles of
a type with a non-zero mass-table entry get that mass (read_ic.c:254-255).

Velocities are returned as stored (Gadget's u = v_pec / sqrt(a) in comoving
runs); `internal_velocities` applies init.c:208's conversion to AREPO's
internal p = a^2 dx/dt.  Only the fields the gravity step and the unbinding
loop read are kept.  No reference code is used: this is a reader of the
documented file layout, written for this package.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field

import numpy as np

NTYPES = 6
# struct io_header: npart[6] int, mass[6] double, time, redshift double, flag_sfr, flag_feedback int,
# npartTotal[6] unsigned int, flag_cooling, num_files int, BoxSize, Omega0, OmegaLambda, HubbleParam double,
# flag_stellarage, flag_metals int, npartTotalHighWord[6] unsigned int, flag_entropy_instead_u,
# flag_doubleprecision int, then padding to 256 bytes
_HEADER = "6i6d2d2i6I2i4d2i6I2i"
_HEADER_BYTES = struct.calcsize("<" + _HEADER)
assert _HEADER_BYTES <= 256


@dataclass
class Header:
    npart: list
    mass: list
    time: float = 0.0
    redshift: float = 0.0
    flag_sfr: int = 0
    flag_feedback: int = 0
    npart_total: list = field(default_factory=lambda: [0] * NTYPES)
    flag_cooling: int = 0
    num_files: int = 1
    box_size: float = 0.0
    omega0: float = 0.0
    omega_lambda: float = 0.0
    hubble_param: float = 0.0
    flag_stellarage: int = 0
    flag_metals: int = 0
    npart_total_high: list = field(default_factory=lambda: [0] * NTYPES)
    flag_entropy_instead_u: int = 0
    flag_doubleprecision: int = 0
    pad: bytes = b""

    def pack(self, e: str) -> bytes:
        vals = (list(self.npart) + list(self.mass) + [self.time, self.redshift, self.flag_sfr, self.flag_feedback] +
                list(self.npart_total) + [self.flag_cooling, self.num_files, self.box_size, self.omega0,
                                          self.omega_lambda, self.hubble_param, self.flag_stellarage,
                                          self.flag_metals] + list(self.npart_total_high) +
                [self.flag_entropy_instead_u, self.flag_doubleprecision])
        raw = struct.pack(e + _HEADER, *vals)
        pad = self.pad if len(self.pad) == 256 - len(raw) else bytes(256 - len(raw))
        return raw + pad

    @staticmethod
    def unpack(raw: bytes, e: str) -> "Header":
        v = struct.unpack(e + _HEADER, raw[:_HEADER_BYTES])
        return Header(npart=list(v[0:6]), mass=list(v[6:12]), time=v[12], redshift=v[13], flag_sfr=v[14],
                      flag_feedback=v[15], npart_total=list(v[16:22]), flag_cooling=v[22], num_files=v[23],
                      box_size=v[24], omega0=v[25], omega_lambda=v[26], hubble_param=v[27], flag_stellarage=v[28],
                      flag_metals=v[29], npart_total_high=list(v[30:36]), flag_entropy_instead_u=v[36],
                      flag_doubleprecision=v[37], pad=raw[_HEADER_BYTES:256])


@dataclass
class Snapshot:
    header: Header
    pos: np.ndarray       # (N, 3) float64, particles ordered by type as in the file
    vel: np.ndarray       # (N, 3) float64, as stored
    ids: np.ndarray       # (N,) uint32 or uint64
    mass: np.ndarray      # (N,) float64
    ptype: np.ndarray     # (N,) uint8
    utherm: np.ndarray | None = None  # gas only (N_gas,), float64
    extra: bytes = b""    # records after the ones read (kept for a byte-exact rewrite)
    float_bytes: int = 4


class _Records:
    def __init__(self, data: bytes):
        self.data = data
        self.o = 0
        first = struct.unpack("<i", data[:4])[0]
        # read_ic.c:1019-1021: a first record of other length than 256 means swapped bytes
        self.e = "<" if first == 256 else ">"
        if struct.unpack(self.e + "i", data[:4])[0] != 256:
            raise ValueError("not a Gadget format-1 file: the first record is not the 256-byte header")

    def record(self) -> bytes:
        if self.o + 4 > len(self.data):
            raise ValueError("truncated file: record expected at byte %d" % self.o)
        n = struct.unpack(self.e + "i", self.data[self.o:self.o + 4])[0]
        body = self.data[self.o + 4:self.o + 4 + n]
        if len(body) != n or self.data[self.o + 4 + n:self.o + 8 + n] != self.data[self.o:self.o + 4]:
            raise ValueError("broken record framing at byte %d" % self.o)
        self.o += n + 8
        return body


def read_gadget1(path: str) -> Snapshot:
    with open(path, "rb") as f:
        data = f.read()
    rec = _Records(data)
    e = rec.e
    h = Header.unpack(rec.record(), e)
    npart = np.array(h.npart, dtype=np.int64)
    n = int(npart.sum())
    ptype = np.repeat(np.arange(NTYPES, dtype=np.uint8), npart)

    body = rec.record()
    fb = len(body) // (3 * n) if n else 4  # POS: float or double (flag_doubleprecision)
    if fb not in (4, 8) or len(body) != 3 * n * fb:
        raise ValueError("POS record has %d bytes for %d particles" % (len(body), n))
    ft = np.dtype(e + ("f4" if fb == 4 else "f8"))
    pos = np.frombuffer(body, ft).reshape(n, 3).astype(np.float64)
    body = rec.record()
    if len(body) != 3 * n * fb:
        raise ValueError("VEL record has %d bytes for %d particles" % (len(body), n))
    vel = np.frombuffer(body, ft).reshape(n, 3).astype(np.float64)
    body = rec.record()
    ib = len(body) // n if n else 4  # ID: 32-bit, or 64-bit with LONGIDS (read_ic.c:59-63)
    if ib not in (4, 8) or len(body) != n * ib:
        raise ValueError("ID record has %d bytes for %d particles" % (len(body), n))
    ids = np.frombuffer(body, np.dtype(e + ("u4" if ib == 4 else "u8"))).copy()

    # MASS for the types present with a zero mass-table entry; the table otherwise
    mass = np.empty(n)
    var = [t for t in range(NTYPES) if npart[t] > 0 and h.mass[t] == 0]
    nvar = int(sum(npart[t] for t in var))
    mvar = None
    if nvar > 0:
        body = rec.record()
        if len(body) != nvar * fb:
            raise ValueError("MASS record has %d bytes for %d particles" % (len(body), nvar))
        mvar = np.frombuffer(body, ft).astype(np.float64)
    first = np.concatenate([[0], np.cumsum(npart)])
    q = 0
    for t in range(NTYPES):
        a, b = int(first[t]), int(first[t + 1])
        if b == a:
            continue
        if h.mass[t] != 0:
            mass[a:b] = h.mass[t]
        else:
            mass[a:b] = mvar[q:q + b - a]
            q += b - a
    utherm = None
    if npart[0] > 0 and rec.o < len(data):
        body = rec.record()
        if len(body) == int(npart[0]) * fb:
            utherm = np.frombuffer(body, ft).astype(np.float64)
        else:
            rec.o -= len(body) + 8
    return Snapshot(h, pos, vel, ids, mass, ptype, utherm, data[rec.o:], fb)


def write_gadget1(path: str, s: Snapshot, byteorder: str = "<") -> None:
    """Write the records read_gadget1 reads (header, POS, VEL, ID, MASS, U), then any extra records kept."""
    e = byteorder
    ft = np.dtype(e + ("f4" if s.float_bytes == 4 else "f8"))
    it = np.dtype(e + ("u4" if s.ids.dtype.itemsize == 4 else "u8"))

    def rec(b: bytes) -> bytes:
        m = struct.pack(e + "i", len(b))
        return m + b + m

    out = [rec(s.header.pack(e)), rec(s.pos.astype(ft).tobytes()), rec(s.vel.astype(ft).tobytes()),
           rec(s.ids.astype(it).tobytes())]
    var = np.isin(s.ptype, [t for t in range(NTYPES) if s.header.npart[t] > 0 and s.header.mass[t] == 0])
    if var.any():
        out.append(rec(s.mass[var].astype(ft).tobytes()))
    if s.utherm is not None:
        out.append(rec(s.utherm.astype(ft).tobytes()))
    with open(path, "wb") as f:
        f.write(b"".join(out))
        if byteorder == "<" or not s.extra:
            f.write(s.extra)


def internal_velocities(vel: np.ndarray, time_a: float, comoving: bool) -> np.ndarray:
    """init.c:208: in comoving runs the stored u = v/sqrt(a) become p = a^2 dx/dt (u sqrt(a) a)."""
    return vel * (np.sqrt(time_a) * time_a) if comoving else vel.copy()


def gravity_constant(unit_length_cm: float, unit_mass_g: float, unit_velocity_cm_s: float) -> float:
    """begrun.c:273 with GRAVITY = 6.6738e-8 cgs (allvars.h:419); UnitTime = UnitLength / UnitVelocity."""
    unit_time = unit_length_cm / unit_velocity_cm_s
    return 6.6738e-8 / unit_length_cm ** 3 * unit_mass_g * unit_time ** 2


def softening_comoving(eps_comoving: float, eps_max_phys: float, time_a: float, comoving: bool) -> float:
    """SofteningTable of set_softenings (grav_softening.c:59-71): the comoving value, or max_phys / a
    where the comoving value would exceed max_phys in physical units."""
    if comoving and eps_comoving * time_a > eps_max_phys:
        return eps_max_phys / time_a
    return eps_comoving
