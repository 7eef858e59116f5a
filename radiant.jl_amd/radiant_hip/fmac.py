"""FMAC-M multigroup cross-section libraries: the reader of cross_sections/read_fmac_m.jl:16-461
and the writer of cross_sections/write_fmac_m.jl:16-377.

The file is text.
- Line 1 starts with "FMAC-M"; lines 2-6 are not read.
- Lines 7-13 hold the 40 integer parameters of record II+1:
  - P[1] cumulative number of groups;
  - P[2] number of particles;
  - P[15] number of materials;
  - P[17] length of record II+2;
  - P[20] Legendre order + 1.
- Record II+2 holds the lengths L[i] of the records III+i.
- The records III+i with L[i] != 0 follow in order of i, each over ceil(L[i]/6) lines. The title
  and the particle names take one line each.
- Then come the scattering records III+39, one pair per (from-group, to-group):
  - an integer line: from-group, to-group, one entry per material (0: no scattering; -(n): n
    Legendre moments follow), and the count of values that follow;
  - the values: one 0.0 for a material without scattering, else its moments.
- Lines are read after collapsing runs of spaces into one (tools/one_space.jl). Line numbers count
  empty lines too.
- Integers are written in 12-character fields, six per line.
- Reals are written as 12-character fields "sd.dddddEsxx", six per line. Reading takes the
  mantissa from characters 2-8 and the exponent from 10-12 of each field.

Group indices run over all particles (cumulative), as in the reference.
"""
import math
from dataclasses import dataclass, field

import numpy as np

NAMES = {"GAMA": "photon", "BETA": "electron", "POSITR": "positron"}
FMAC = {v: k for k, v in NAMES.items()}


@dataclass
class Library:
    """A multigroup library as Radiant's Cross_Sections holds it after read_fmac_m.jl:318-373."""
    particles: list                       # "photon" / "electron" / "positron", file order
    groups: list                          # number of groups per particle
    legendre_order: int
    energy_boundaries: list               # per particle: [G+1]
    total: list                           # per particle: [G, Nmat]
    absorption: list                      # per particle: [G, Nmat]
    stopping_powers: list                 # per particle: [G+1, Nmat] at the group boundaries
    momentum_transfer: list               # per particle: [G, Nmat]
    energy_deposition: list               # per particle: [G, Nmat] (the cut-off term is not included)
    charge_deposition: list               # per particle: [G, Nmat]
    scattering: dict = field(default_factory=dict)  # (i, j) -> [Nmat, G_i, G_j, L+1] from i to j
    title: str = ""
    # energy and charge deposited at the cut-off (the value at the lowest group boundary) per
    # particle: [Nmat] each; MATXS files carry them (edep, cdep of length G + 1), FMAC-M files do not
    energy_deposition_cutoff: list = field(default_factory=list)
    charge_deposition_cutoff: list = field(default_factory=list)

    @property
    def Nmat(self):
        return self.total[0].shape[1]

    def index(self, particle):
        particle = NAMES.get(particle, particle)
        if particle not in self.particles:
            raise ValueError(f"Cross-sections don't contain data for the {particle} particle.")
        return self.particles.index(particle)

    def transport(self, particle="photon", L=None):
        """(sigt [G, Nmat], sigs [Nmat, G, G, L+1]) of one particle, in the layouts the solve takes
        (sn_flux.jl:69-83: get_total and get_scattering with zero padding above the file's order)."""
        i = self.index(particle)
        L = self.legendre_order if L is None else L
        s = self.scattering[(i, i)]
        out = np.zeros(s.shape[:3] + (L + 1,))
        n = min(L + 1, s.shape[3])
        out[..., :n] = s[..., :n]
        return self.total[i].copy(), out


# ----------------------------------------------------------------------------------- reading
def _one_space(line):
    """tools/one_space.jl: leading blanks removed, every other run of blanks kept as one."""
    out = []
    prev_space = True
    for ch in line:
        if ch.isspace():
            if not prev_space:
                out.append(ch)
            prev_space = True
        else:
            out.append(ch)
            prev_space = False
    return "".join(out)


def _floats(line, count):
    """The real fields of one line (read_fmac_m.jl:444-461)."""
    if line[0] != "-":
        line = "+" + line
    vals = []
    for i in range(count):
        f = line[12 * i:12 * i + 12]
        if len(f) < 12:
            raise ValueError(f"FMAC-M: short real field in line {line!r}")
        mant = (-1.0 if f[0] == "-" else 1.0) * float(f[1:8])
        ex = (-1.0 if f[9] == "-" else 1.0) * float(f[10:12])
        vals.append(mant * 10.0 ** ex)
    return vals


def read_fmac_m(path):
    """Reads an FMAC-M file into a Library (read_fmac_m.jl:16-373)."""
    with open(path) as f:
        raw = f.read().split("\n")
    if raw and raw[-1] == "":
        raw = raw[:-1]
    lines = [_one_space(x) for x in raw]
    if not lines or lines[0][:6] != "FMAC-M":
        raise ValueError("The file is not a FMAC-M format file.")
    k = 6  # 0-based index of the first line of record II+1 (line 7)

    def ints(nvals):
        nonlocal k
        v = []
        while len(v) < nvals:
            s = lines[k]
            k += 1
            if s == "":
                continue
            v += [int(t) for t in s.split(" ") if t != ""]
        return v

    def reals(nvals):
        nonlocal k
        v = []
        nl = math.ceil(nvals / 6)
        j = 0
        while j < nl:
            if lines[k] == "":
                k += 1
                continue
            cnt = 6 if j < nl - 1 else (nvals % 6 or 6)
            v += _floats(lines[k], cnt)
            k += 1
            j += 1
        return v

    P = ints(40)
    Gtot, Npart, Nmat, n2, Lp1 = P[0], P[1], P[14], P[16], P[19]
    L = Lp1 - 1
    Lrec = ints(n2)
    rec = {i + 1: n for i, n in enumerate(Lrec) if n != 0}
    known = {1, 2, 3, 4, 5, 6, 11, 12, 14, 15, 16, 35, 36, 37, 38}
    for i in rec:
        if i not in known:
            raise ValueError(f"FMAC-M: record III+{i} is not read (the reference reader does not read it either)")
    groups, names, eb_raw = [], [], []
    tot = ab = sp = mt = ed = cd = None
    title = ""
    for i in sorted(rec):
        n = rec[i]
        if i == 1:
            title = lines[k]
            k += 1
        elif i == 2:
            groups = ints(n)
        elif i == 3:
            names = [t for t in lines[k].split(" ") if t != ""]
            k += 1
        elif i in (4, 6):
            reals(n)  # rest energies, group velocities: not used by the transport solve
        elif i == 5:
            eb_raw = reals(n)
        elif i in (11, 12, 14):
            ints(n)
        elif i == 15:
            tot = np.array(reals(n)).reshape(Gtot, Nmat)
        elif i == 16:
            ab = np.array(reals(n)).reshape(Gtot, Nmat)
        elif i == 35:
            sp = reals(n)
        elif i == 36:
            mt = np.array(reals(n)).reshape(Gtot, Nmat)
        elif i == 37:
            ed = np.array(reals(n)).reshape(Gtot, Nmat)
        elif i == 38:
            cd = np.array(reals(n)).reshape(Gtot, Nmat)
    if len(groups) != Npart or sum(groups) != Gtot:
        raise ValueError("FMAC-M: group counts disagree with the cumulative number of groups")
    # scattering records III+39
    scat = {}  # (material, to-group, from-group) -> moments (0-based cumulative groups)
    while k < len(lines):
        if lines[k] == "":
            k += 1
            continue
        head = ints(3 + Nmat)
        gi, gf, ns, cnt = head[0] - 1, head[1] - 1, [abs(x) for x in head[2:2 + Nmat]], head[2 + Nmat]
        vals = reals(cnt)
        pos = 0
        for m in range(Nmat):
            scat[(m, gf, gi)] = vals[pos:pos + ns[m]]
            pos += ns[m] if ns[m] != 0 else 1
    off = np.concatenate([[0], np.cumsum(groups)]).astype(int)

    def part(a, n):
        if a is None:
            return np.zeros((groups[n], Nmat))
        return a[off[n]:off[n] + groups[n]].copy()

    lib = Library(particles=[NAMES.get(s, s) for s in names[:Npart]], groups=list(groups), legendre_order=L,
                  energy_boundaries=[], total=[], absorption=[], stopping_powers=[], momentum_transfer=[],
                  energy_deposition=[], charge_deposition=[], title=title)
    # energy boundaries: each particle's lower boundaries in file order, then the last boundaries
    # of the particles in reverse order (write_fmac_m.jl:102-115, read_fmac_m.jl:166-178)
    for n in range(Npart):
        if eb_raw:
            lo = eb_raw[off[n]:off[n] + groups[n]]
            last = eb_raw[Gtot + (Npart - 1 - n)]
            lib.energy_boundaries.append(np.array(lo + [last]))
        else:
            lib.energy_boundaries.append(np.zeros(groups[n] + 1))
        lib.total.append(part(tot, n))
        lib.absorption.append(part(ab, n))
        lib.momentum_transfer.append(part(mt, n))
        lib.energy_deposition.append(part(ed, n))
        lib.charge_deposition.append(part(cd, n))
        s = np.zeros((groups[n] + 1, Nmat))
        if sp is not None:
            spa = np.array(sp)
            s[:groups[n]] = spa[off[n] * Nmat:(off[n] + groups[n]) * Nmat].reshape(groups[n], Nmat)
            s[groups[n]] = spa[Gtot * Nmat + (Npart - 1 - n) * Nmat:Gtot * Nmat + (Npart - n) * Nmat]
        lib.stopping_powers.append(s)
    for n in range(Npart):
        for m in range(Npart):
            a = np.zeros((Nmat, groups[n], groups[m], L + 1))
            for mat in range(Nmat):
                for g_i in range(groups[n]):
                    for g_f in range(groups[m]):
                        v = scat.get((mat, off[m] + g_f, off[n] + g_i))
                        if v:
                            a[mat, g_i, g_f, :len(v)] = v[:L + 1]
            lib.scattering[(n, m)] = a
    return lib


# ----------------------------------------------------------------------------------- writing
def _int_field(i):
    s = str(int(i))
    return " " * (12 - len(s)) + s


def _real_field(x):
    """write_fmac_m.jl:283-313: sign, mantissa rounded to 5 decimals (shortest decimal form) and
    padded with zeros to 7 characters, 'E', exponent sign and two digits.  A decimal exponent of 99
    or more in magnitude is written as zero (the reference's test of the exponent sign never holds,
    so large values are zeroed as well as small ones)."""
    x = float(x)
    sg = "-" if x < 0 else " "
    e = math.floor(math.log10(abs(x))) if x != 0 else 0
    se = "-" if e < 0 else "+"
    mant = abs(x) / 10.0 ** e
    e = abs(e)
    if e < 10:
        es = "0" + str(e)
    elif e < 99:
        es = str(e)
    else:
        es, mant, sg = "00", 0.0, " "
    ms = repr(round(mant, 5))
    ms = ms + "0" * max(0, 7 - len(ms))
    return f"{sg}{ms}E{se}{es}"


def _write_ints(f, v):
    for j, i in enumerate(v):
        f.write(_int_field(i))
        if (j + 1) % 6 == 0:
            f.write("\n")
    if len(v) % 6 != 0:
        f.write("\n")


def _write_reals(f, v):
    for j, x in enumerate(v):
        f.write(_real_field(x))
        if (j + 1) % 6 == 0:
            f.write("\n")
    if len(v) % 6 != 0:
        f.write("\n")


def write_fmac_m(path, lib):
    """Writes a Library in the layout of write_fmac_m.jl:16-235."""
    Npart, Nmat, L = len(lib.particles), lib.Nmat, lib.legendre_order
    Ng = list(lib.groups)
    Gtot = sum(Ng)
    with open(path, "w") as f:
        f.write("FMAC-M                       5\n")
        f.write("\n\n")
        v0 = [0] * 10
        v0[0] = 1
        _write_ints(f, v0)
        f.write("NGCAP\n")
        v1 = [0] * 40
        v1[0], v1[1], v1[14], v1[16], v1[19] = Gtot, Npart, Nmat, 38, L + 1
        _write_ints(f, v1)
        v2 = [0] * 38
        v2[0], v2[1], v2[2], v2[4] = 12, Npart, Npart, Gtot + Npart
        v2[14] = Gtot * Nmat
        v2[34] = (Gtot + Npart) * Nmat
        v2[35] = v2[36] = v2[37] = Gtot * Nmat
        _write_ints(f, v2)
        f.write("Coupled multigroup cross-sections produced with Radiant.jl\n")
        _write_ints(f, Ng)
        f.write("".join(FMAC[p] + "  " for p in lib.particles) + "\n")
        eb = []
        for n in range(Npart):
            eb += list(lib.energy_boundaries[n][:-1])
        for n in range(Npart):
            eb.append(lib.energy_boundaries[Npart - 1 - n][-1])
        _write_reals(f, eb)
        _write_reals(f, [lib.total[n][g, m] for n in range(Npart) for g in range(Ng[n]) for m in range(Nmat)])
        sp = [lib.stopping_powers[n][g, m] for n in range(Npart) for g in range(Ng[n]) for m in range(Nmat)]
        sp += [lib.stopping_powers[Npart - 1 - n][-1, m] for n in range(Npart) for m in range(Nmat)]
        _write_reals(f, sp)
        for arr in (lib.momentum_transfer, lib.energy_deposition, lib.charge_deposition):
            _write_reals(f, [arr[n][g, m] for n in range(Npart) for g in range(Ng[n]) for m in range(Nmat)])
        off = np.concatenate([[0], np.cumsum(Ng)]).astype(int)
        for j in range(Npart):
            for gf in range(Ng[j]):
                for i in range(Npart):
                    for gi in range(Ng[i]):
                        d = lib.scattering[(i, j)][:, gi, gf, :]
                        p = [0] * (3 + Nmat)
                        p[0], p[1] = off[i] + gi + 1, off[j] + gf + 1
                        vals = []
                        for m in range(Nmat):
                            if np.all(d[m] == 0.0):
                                p[2 + m] = 0
                                p[-1] += 1
                                vals.append(0.0)
                            else:
                                nz = np.nonzero(d[m])[0]
                                le = int(nz[-1])
                                p[2 + m] = -(le + 1)
                                p[-1] += le + 1
                                vals += list(d[m, :le + 1])
                        _write_ints(f, p)
                        _write_reals(f, vals)
