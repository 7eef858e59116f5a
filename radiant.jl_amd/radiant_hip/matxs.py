"""MATXS multigroup libraries (the NJOY/CCCC interface format, in the form Radiant writes and reads):
the reader of cross_sections/read_matxs.jl:28-211 and the writer of cross_sections/write_matxs.jl:44-197.

A file is a sequence of records:
- " 0v" file identification, " 1d" file control (npart, ntype, nholl, nmat, maxw, length),
  " 2d" set identification (nholl words), " 3d" file data (particle, data-type and material
  names; groups per particle, jinp, joutp, submaterials and locations per material);
- " 4d" per particle: the Ng + 1 group boundaries in eV, descending;
- per material: " 5d" material control (name, amass, then per submaterial temp, sigz and the
  integers itype, n1d, n2d, locs), and per submaterial " 6d"/" 7d" vector control and data when it
  holds vectors, " 8d"/" 9d" matrix control and banded data.
- One data type per ordered (incident, outgoing) particle pair (ntype = npart^2, type
  (i - 1) npart + o); the self pairs carry the incident particle's vectors, every pair its
  transfer matrix.
- Vectors: gtot0 gheat gchar gabs gedep (photon), btot0 bheat bchar babs EMOMTR BSTC bedep
  (electron), ptot0 pheat pchar pabs PSTC pedep (positron), first group 1, last group Ng (Ng + 1
  for the stopping powers and edep).  EMOMTR is the momentum transfer, *STC the boundary stopping
  powers; *heat is not read back (edep holds the deposition); *char (Ng values) stands for the
  charge deposition with 0 at the cut-off, or a *cdep of Ng + 1 values.  *tot0, *abs, *edep and a
  charge vector are required, *momt/EMOMTR and the stopping powers default to zero.
- Matrix: Legendre count lord, then per outgoing group the band length jband and the highest
  incident group ijj; the data run over outgoing groups, Legendre moments, then incident groups
  ijj down to ijj - jband + 1.
- Values in the file are per unit density: written divided by the material density, read back
  multiplied by it.

Encodings.  Text ("BCD"): each record starts with its 4-character tag; reals are 12-character
"%12.5E" fields, integers 6-character fields, names 8-character words.  The line layout after the
tag: " 0v" two names, a blank word and an integer; " 1d" two blanks then the integers (6 on the
first line, 12 per line after); " 2d", " 3d", " 6d" four blanks then names (8 on the first line,
9 per line after) and for 3d/6d the integers on new lines (12 per line); " 4d", " 7d", " 9d" eight
blanks then reals (5 on the first line, 6 per line after); " 5d" the name and amass, then one line
per submaterial (two reals, four integers); " 8d" four blanks and the name, then the integers.
Binary: each record is one Fortran-style unformatted record, a 4-byte length, the payload
(int32 integers, float64 reals, 8-byte names, little-endian) and the length again.  The reader
tells the two apart by the first four bytes (" 0v ").
"""
import struct

import numpy as np

from .fmac import Library

CODES = {"photon": "g", "electron": "b", "positron": "p"}
PARTICLES = {v: k for k, v in CODES.items()}
VECTORS = {"g": ["gtot0", "gheat", "gchar", "gabs", "gedep"],
           "b": ["btot0", "bheat", "bchar", "babs", "EMOMTR", "BSTC", "bedep"],
           "p": ["ptot0", "pheat", "pchar", "pabs", "PSTC", "pedep"]}
TAGS = ("0v", "1d", "2d", "3d", "4d", "5d", "6d", "7d", "8d", "9d", "10d")
EV_PER_MEV = 1.0e6


def _key(name):
    """Quantity of a vector name: its suffix after the particle code, with the charged-particle
    names mapped (read_matxs.jl:424-434)."""
    name = name.strip()
    if name == "EMOMTR":
        return "momt"
    if name in ("BSTC", "CSTC", "PSTC"):
        return "stpw"
    return name[1:] if len(name) >= 2 else name


# ----------------------------------------------------------------------------------- reading
class _Records:
    """The records of a file, each read as a stream of names, reals and integers."""

    def __init__(self, raw):
        self.binary = raw[:4] != b" 0v "
        if self.binary:
            self.recs, pos = [], 0
            while pos + 4 <= len(raw):
                (m,) = struct.unpack_from("<i", raw, pos)
                self.recs.append(raw[pos + 4:pos + 4 + m])
                pos += 8 + m
        else:
            lines = raw.decode("latin-1").split("\n")
            self.lines = [l[:-1] if l.endswith("\r") else l for l in lines]
            self.li = 0
        self.ri = 0

    @staticmethod
    def _is_tag(line):
        return len(line) >= 4 and line[:4].strip() in TAGS

    def next(self, tag):
        if self.binary:
            if self.ri >= len(self.recs):
                raise ValueError(f"Expected MATXS record '{tag}' at end of file.")
            self.buf, self.pos = self.recs[self.ri], 0
            self.ri += 1
            return
        if self.li >= len(self.lines) or self.lines[self.li][:4].strip() != tag.strip():
            where = self.lines[self.li] if self.li < len(self.lines) else "end of file"
            raise ValueError(f"Expected MATXS record '{tag}' at line {self.li + 1}: {where}")
        first = self.lines[self.li].ljust(4)[4:]
        self.li += 1
        cont = []
        while self.li < len(self.lines) and not self._is_tag(self.lines[self.li]):
            cont.append(self.lines[self.li])
            self.li += 1
        t = tag.strip()
        if first.strip() == "":
            payload = cont
        elif t == "0v":
            r = first.ljust(35)
            payload = [r[0:8] + r[9:17], r[-6:]]
        elif t == "1d":
            payload = [first[2:]] + cont
        elif t in ("2d", "3d", "6d"):
            payload = [first[4:]] + cont
        elif t in ("4d", "7d", "9d", "10d"):
            payload = [first[8:]] + cont
        elif t == "5d":
            r = first.ljust(20)
            reals, ints = [r[8:]], []
            for c in cont:
                c = c.ljust(48)
                reals.append(c[:24])
                ints.append(c[24:48])
            payload = [r[:8]] + reals + ints
        elif t == "8d":
            payload = [first.ljust(12)[4:12]] + cont
        else:
            raise ValueError(f"Unsupported MATXS record tag '{tag}'.")
        self.payload, self.pi = payload, 0

    def _fields(self, n, width, minimum=0):
        out = []
        while len(out) < n:
            line = self.payload[self.pi]
            self.pi += 1
            k = min(max(len(line) // width, minimum), n - len(out))
            line = line.ljust(k * width)
            out += [line[j * width:(j + 1) * width] for j in range(k)]
        return out

    def reals(self, n):
        if self.binary:
            v = struct.unpack_from(f"<{n}d", self.buf, self.pos)
            self.pos += 8 * n
            return np.array(v, dtype=np.float64)
        return np.array([float(f) for f in self._fields(n, 12)], dtype=np.float64)

    def ints(self, n):
        if self.binary:
            v = struct.unpack_from(f"<{n}i", self.buf, self.pos)
            self.pos += 4 * n
            return [int(x) for x in v]
        return [int(f) for f in self._fields(n, 6)]

    def names(self, n):
        if self.binary:
            v = [self.buf[self.pos + 8 * k:self.pos + 8 * k + 8].decode("latin-1") for k in range(n)]
            self.pos += 8 * n
            return v
        return self._fields(n, 8, minimum=1)  # an all-blank line is one blank word


def read_matxs(path, densities, particles=None):
    """Library of a MATXS file.  densities: the materials' densities in file order (the values are
    multiplied by them); particles: the particle of each library slot ("photon", "electron",
    "positron"; default: the file's, in file order).  Follows read_matxs.jl:28-211."""
    with open(path, "rb") as f:
        rd = _Records(f.read())
    rd.next(" 0v ")
    rd.names(2)
    rd.ints(1)
    rd.next(" 1d ")
    npart, ntype, nholl, nmat = rd.ints(6)[:4]
    densities = np.asarray(densities, dtype=np.float64)
    if densities.size != nmat:
        raise ValueError("MATXS file material count mismatch.")
    if np.any(densities <= 0.0):
        raise ValueError("MATXS input requires strictly positive material densities.")
    rd.next(" 2d ")
    rd.names(nholl)
    rd.next(" 3d ")
    holl = rd.names(npart + ntype + nmat)
    codes = [h.strip() for h in holl[:npart]]
    ints = rd.ints(npart + 2 * ntype + 2 * nmat)
    ng_file, jinp, joutp = ints[:npart], ints[npart:npart + ntype], ints[npart + ntype:npart + 2 * ntype]
    if particles is None:
        particles = [PARTICLES[c] for c in codes]
    if len(particles) != npart:
        raise ValueError(f"MATXS file holds {npart} particles but {len(particles)} were provided.")
    slot = []
    for c in codes:
        if c not in PARTICLES or PARTICLES[c] not in particles:
            raise ValueError(f"MATXS particle code '{c}' has no matching provided particle.")
        slot.append(particles.index(PARTICLES[c]))
    groups, bounds = [0] * npart, [None] * npart
    for j in range(npart):
        rd.next(" 4d ")
        groups[slot[j]] = ng_file[j]
        bounds[slot[j]] = rd.reals(ng_file[j] + 1) / EV_PER_MEV
    vec = [[dict() for _ in range(nmat)] for _ in range(npart)]
    mats = {}
    lord_max = 1
    for m in range(nmat):
        rd.next(" 5d ")
        rd.names(1)
        rd.reals(1 + 2 * ntype)
        ctl = rd.ints(4 * ntype)
        for s in range(ntype):
            t, n1d, n2d = ctl[4 * s], ctl[4 * s + 1], ctl[4 * s + 2]
            i, o = slot[jinp[t - 1] - 1], slot[joutp[t - 1] - 1]
            if n1d > 0:
                rd.next(" 6d ")
                names = rd.names(n1d)
                band = rd.ints(2 * n1d)
                lengths = [band[n1d + k] - band[k] + 1 for k in range(n1d)]
                rd.next(" 7d ")
                data = rd.reals(sum(lengths))
                pos = 0
                for k in range(n1d):
                    vec[i][m][_key(names[k])] = data[pos:pos + lengths[k]]
                    pos += lengths[k]
            if n2d > 0:
                rd.next(" 8d ")
                rd.names(1)
                go, gi = groups[o], groups[i]
                hdr = rd.ints(2 + 2 * go)
                lord, jband, ijj = hdr[0], hdr[2:2 + go], hdr[2 + go:2 + 2 * go]
                lord_max = max(lord_max, lord)
                rd.next(" 9d ")
                data = rd.reals(lord * sum(jband))
                a = np.zeros((gi, go, lord))
                pos = 0
                for gf in range(go):
                    for l in range(lord):
                        for g in range(ijj[gf], ijj[gf] - jband[gf], -1):
                            a[g - 1, gf, l] = data[pos]
                            pos += 1
                mats[(i, o, m)] = a
    L = lord_max - 1
    lib = Library(particles=list(particles), groups=groups, legendre_order=L, energy_boundaries=bounds, total=[],
                  absorption=[], stopping_powers=[], momentum_transfer=[], energy_deposition=[],
                  charge_deposition=[], title="")

    def take(n, m, name, length, required=True):
        d = vec[n][m]
        if name not in d:
            if required:
                raise ValueError(f"MATXS file is missing required vector '*{name}' for incident particle "
                                 f"'{CODES[particles[n]]}'.")
            return np.zeros(length)
        v = d[name]
        if len(v) != length:
            raise ValueError(f"MATXS vector '*{name}' for incident particle '{CODES[particles[n]]}' has length "
                             f"{len(v)}; expected {length}.")
        return v

    for n in range(npart):
        G = groups[n]
        cols = {k: [] for k in ("tot", "abs", "edep", "momt", "stpw", "cdep")}
        for m in range(nmat):
            rho = densities[m]
            cols["tot"].append(take(n, m, "tot0", G) * rho)
            cols["abs"].append(take(n, m, "abs", G) * rho)
            cols["edep"].append(take(n, m, "edep", G + 1) * rho)
            cols["momt"].append(take(n, m, "momt", G, required=False) * rho)
            cols["stpw"].append(take(n, m, "stpw", G + 1, required=False) * rho)
            if "cdep" in vec[n][m]:
                c = take(n, m, "cdep", G + 1)
            elif "char" in vec[n][m]:
                c = np.append(take(n, m, "char", G), 0.0)
            else:
                raise ValueError("MATXS file is missing required charge-deposition vector 'cdep' or 'char' for "
                                 f"incident particle '{CODES[particles[n]]}'.")
            cols["cdep"].append(c * rho)
        stack = {k: np.stack(v, axis=1) for k, v in cols.items()}  # [length, Nmat]
        lib.total.append(stack["tot"])
        lib.absorption.append(stack["abs"])
        lib.momentum_transfer.append(stack["momt"])
        lib.stopping_powers.append(stack["stpw"])
        lib.energy_deposition.append(stack["edep"][:G])
        lib.energy_deposition_cutoff.append(stack["edep"][G])
        lib.charge_deposition.append(stack["cdep"][:G])
        lib.charge_deposition_cutoff.append(stack["cdep"][G])
    for i in range(npart):
        for o in range(npart):
            s = np.zeros((nmat, groups[i], groups[o], L + 1))
            for m in range(nmat):
                a = mats.get((i, o, m))
                if a is not None:
                    s[m, :, :, :a.shape[2]] = a * densities[m]
            lib.scattering[(i, o)] = s
    return lib


# ----------------------------------------------------------------------------------- writing
def _lines(fields, first, cont):
    out, i = [fields[:first]], first
    while i < len(fields):
        out.append(fields[i:i + cont])
        i += cont
    return ["".join(f) for f in out]


def _word(s):
    return str(s)[:8].ljust(8)


class _Writer:
    def __init__(self, f, binary):
        self.f, self.binary = f, binary

    def record(self, tag, *segments):
        """segments: ("i" | "r" | "h", values) in record order."""
        if self.binary:
            buf = b""
            for kind, vals in segments:
                if kind == "i":
                    buf += struct.pack(f"<{len(vals)}i", *[int(v) for v in vals])
                elif kind == "r":
                    buf += struct.pack(f"<{len(vals)}d", *[float(v) for v in vals])
                else:
                    buf += b"".join(_word(v).encode("latin-1") for v in vals)
            self.f.write(struct.pack("<i", len(buf)) + buf + struct.pack("<i", len(buf)))
            return
        ri = lambda v: [f"{int(x):6d}" for x in v]
        rr = lambda v: [f"{float(x):12.5E}" for x in v]
        rh = lambda v: [_word(x) for x in v]
        t = tag.strip()
        seg = dict(segments) if t != "5d" else None
        out = []
        if t == "0v":
            h, i = segments[0][1], segments[1][1]
            out = [tag + _word(h[0]) + " " + _word(h[1]) + " " * 8 + " " + f"{int(i[0]):6d}"]
        elif t == "1d":
            out = _lines(ri(seg["i"]), 6, 12)
            out[0] = tag + "  " + out[0]
        elif t in ("2d", "3d", "6d"):
            out = _lines(rh(seg["h"]), 8, 9)
            out[0] = tag + "    " + out[0]
            if "i" in seg:
                out += _lines(ri(seg["i"]), 12, 12)
        elif t in ("4d", "7d", "9d"):
            out = _lines(rr(seg["r"]), 5, 6)
            out[0] = tag + " " * 8 + out[0]
        elif t == "5d":
            h, r, i = segments[0][1], segments[1][1], segments[2][1]
            out = [tag + _word(h[0]) + f"{float(r[0]):12.5E}"]
            for s in range(len(i) // 4):
                out.append(f"{float(r[1 + 2 * s]):12.5E}{float(r[2 + 2 * s]):12.5E}" +
                           "".join(f"{int(x):6d}" for x in i[4 * s:4 * s + 4]))
        elif t == "8d":
            out = [tag + "    " + _word(segments[0][1][0])] + _lines(ri(segments[1][1]), 12, 12)
        else:
            raise ValueError(f"Unsupported MATXS record tag '{tag}'.")
        self.f.write(("\n".join(out) + "\n").encode("latin-1"))


def _band(a):
    """jband, ijj of a [G_in, G_out, lord] matrix: per outgoing group the span of incident groups
    with a non-zero moment and the highest of them (1-based; 0 and 1 when none)."""
    gi, go = a.shape[:2]
    jband, ijj = [0] * go, [1] * go
    for gf in range(go):
        nz = np.nonzero(np.any(a[:, gf, :] != 0.0, axis=1))[0]
        if nz.size:
            jband[gf] = int(nz[-1] - nz[0] + 1)
            ijj[gf] = int(nz[-1] + 1)
    return jband, ijj


def write_matxs(path, lib, densities, material_names=None, binary=False):
    """Writes a Library as a MATXS file (write_matxs.jl:44-197): values divided by the material
    densities, group boundaries in eV, BCD text unless binary."""
    densities = np.asarray(densities, dtype=np.float64)
    nmat, npart = lib.Nmat, len(lib.particles)
    if densities.size != nmat or np.any(densities <= 0.0):
        raise ValueError("MATXS output requires strictly positive material densities.")
    names = material_names or [f"mat{m + 1}" for m in range(nmat)]
    codes = [CODES[p] for p in lib.particles]
    G, lord = lib.groups, lib.legendre_order + 1
    ntype = npart * npart
    jinp = [i + 1 for i in range(npart) for _ in range(npart)]
    joutp = [o + 1 for _ in range(npart) for o in range(npart)]
    htype = [codes[i] + codes[o] for i in range(npart) for o in range(npart)]
    n1d = [len(VECTORS[codes[jinp[t] - 1]]) if jinp[t] == joutp[t] else 0 for t in range(ntype)]
    counts = [(2 if n1d[t] else 0) + 2 for t in range(ntype)]
    locs = [sum(counts[:t]) for t in range(ntype)]
    per_mat = 1 + sum(counts)

    def cutoff(name, n, m):
        v = getattr(lib, name)
        return float(v[n][m]) if len(v) > n else 0.0

    with open(path, "wb") as f:
        w = _Writer(f, binary)
        w.record(" 0v ", ("h", ["matxs", "Radiant"]), ("i", [5]))
        w.record(" 1d ", ("i", [npart, ntype, 9, nmat, 99999, 0]))
        title = (lib.title or "Coupled multigroup cross-sections produced with Radiant.jl").ljust(72)[:72]
        w.record(" 2d ", ("h", [title[8 * k:8 * k + 8] for k in range(9)]))
        w.record(" 3d ", ("h", codes + htype + [n[:8] for n in names]),
                 ("i", list(G) + jinp + joutp + [ntype] * nmat + [m * per_mat for m in range(nmat)]))
        for n in range(npart):
            w.record(" 4d ", ("r", list(np.asarray(lib.energy_boundaries[n], dtype=np.float64) * EV_PER_MEV)))
        for m in range(nmat):
            rho = densities[m]
            reals, ints = [0.0], []
            for t in range(ntype):
                reals += [0.0, 0.0]
                ints += [t + 1, n1d[t], 1, locs[t]]
            w.record(" 5d ", ("h", [names[m]]), ("r", reals), ("i", ints))
            for t in range(ntype):
                i, o = jinp[t] - 1, joutp[t] - 1
                if i == o:
                    hv = VECTORS[codes[i]]
                    Gi = G[i]
                    nlg = [Gi + 1 if (v in ("BSTC", "CSTC", "PSTC") or v.endswith("edep")) else Gi for v in hv]
                    w.record(" 6d ", ("h", hv), ("i", [1] * len(hv) + nlg))
                    edep = np.append(lib.energy_deposition[i][:, m], cutoff("energy_deposition_cutoff", i, m))
                    data = []
                    for v in hv:
                        if v.endswith("tot0"):
                            x = lib.total[i][:, m]
                        elif v.endswith("heat"):
                            x = edep[:Gi]
                        elif v.endswith("char"):
                            x = lib.charge_deposition[i][:, m]
                        elif v.endswith("abs"):
                            x = lib.absorption[i][:, m]
                        elif v == "EMOMTR":
                            x = lib.momentum_transfer[i][:, m]
                        elif v in ("BSTC", "CSTC", "PSTC"):
                            x = lib.stopping_powers[i][:, m]
                        else:  # *edep
                            x = edep
                        data += list(np.asarray(x, dtype=np.float64) / rho)
                    w.record(" 7d ", ("r", data))
                s = lib.scattering.get((i, o))
                a = np.zeros((G[i], G[o], lord))
                if s is not None:
                    k = min(lord, s.shape[3])
                    a[:, :, :k] = s[m, :, :, :k]
                jband, ijj = _band(a)
                w.record(" 8d ", ("h", [codes[i] + "scat"]), ("i", [lord, 0] + jband + ijj))
                data = []
                for gf in range(G[o]):
                    for l in range(lord):
                        for g in range(ijj[gf], ijj[gf] - jband[gf], -1):
                            data.append(a[g - 1, gf, l] / rho)
                w.record(" 9d ", ("r", data))
