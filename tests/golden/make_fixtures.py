"""Extract golden vectors from the reference's own test logs into small JSON
fixtures under tests/golden/.  Run in a container where the reference
repository is mounted (default /root/reference); the fixtures it writes are
data (inputs and expected outputs) and are committed.

Sources:
  tests/scripts/nosg.log                 Al hkl families, no space group, dcutoff 0.4
  tests/src/app_debyemsd/main.cc,test.log  sqrt(MSD) table (Debye model)
  tests/src/app_sabutils/main.cc:88-135   mpmath known answers of the alpha-bin integral
  tests/scripts/ncmat2hkl.log + data/*.ncmat  hkl tables of crystals with a space
                                         group (10 significant digits) with their inputs
  tests/src/app_pointwisedist/main.cc,test.log  percentiles of a uniform density
                                         given as a piecewise-linear one
  tests/src/app_sabsliceinteg/main.cc,test.log  tailed breakdowns of S = exp(alpha)
                                         over alpha ranges (front/middle/back integrals)
"""
import json
import pathlib
import re
import sys

REF = pathlib.Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
OUT = pathlib.Path(__file__).resolve().parent


def nosg():
    lines = (REF / "tests/scripts/nosg.log").read_text().splitlines()
    # the verbosity-3 Info dump of Al_nosg.ncmat lists every family with its
    # expanded hkl list
    start = next(i for i, l in enumerate(lines)
                 if 'Info-dump of "Al_nosg.ncmat;dcutoff=0.4" with verbosity lvl 3' in l)
    fams = []
    started = False
    for l in lines[start:]:
        if l.startswith("HKL planes"):
            started = True
            continue
        if not started:
            continue
        m = re.match(r"^\s+(-?\d+)\s+(-?\d+)\s+(-?\d+)\s+([0-9.eE+-]+)\s+(\d+)\s+([0-9.eE+-]+)\s+(.*)$", l)
        if not m:
            if fams:
                break
            continue
        members = [tuple(int(x) for x in t.split(","))
                   for t in re.findall(r"\((-?\d+,-?\d+,-?\d+)\)", m.group(7))]
        fams.append(dict(hkl=[int(m.group(1)), int(m.group(2)), int(m.group(3))], d=float(m.group(4)),
                         mult=int(m.group(5)), fsq=float(m.group(6)), members=members))
    data = dict(
        source="tests/scripts/nosg.log (Info-dump of Al_nosg.ncmat;dcutoff=0.4, verbosity 3) and "
               "tests/scripts/nosg.py:21-33 (material)",
        material=dict(a=4.04958, positions=[[0, .5, .5], [0, 0, 0], [.5, .5, 0], [.5, 0, .5]],
                      debye_temp=410.3542, temperature=293.15, mass=26.981538408, csl=0.3449, dcutoff=0.4,
                      msd_printed=0.00989097),
        families=fams)
    (OUT / "nosg_al_hkl.json").write_text(json.dumps(data, indent=1))
    return len(fams)


def debyemsd():
    vals = []
    for l in (REF / "tests/src/app_debyemsd/test.log").read_text().splitlines():
        m = re.match(r"sqrt\(MSD\[m=([0-9.]+)u, Tdebye=([0-9.]+)K, T=([0-9.]+)K\]\) = ([0-9.eE+-]+)Aa", l)
        if m:
            vals.append([float(m.group(1)), float(m.group(2)), float(m.group(3)), float(m.group(4))])
    # the log prints rounded masses/temperatures; the program's exact inputs
    # (tests/src/app_debyemsd/main.cc) are:
    masses = {"1.00794": 1.00794, "26.9815": 26.981539, "238.029": 238.02891}
    tds = {"100": 100.0, "410.354": 410.3542, "2000": 2000.0}
    exact = []
    for m_, td, t, v in vals:
        mm = min(masses.values(), key=lambda x: abs(x - m_))
        tt = min(tds.values(), key=lambda x: abs(x - td))
        exact.append(dict(mass=mm, debye_temp=tt, temperature=t, sqrt_msd=v))
    (OUT / "debye_msd.json").write_text(json.dumps(dict(
        source="tests/src/app_debyemsd/test.log with inputs from tests/src/app_debyemsd/main.cc",
        printed_digits=6, values=exact), indent=1))
    return len(exact)


def sabutils():
    src = (REF / "tests/src/app_sabutils/main.cc").read_text()
    pairs = re.findall(r"ref\.emplace_back\(\s*([-0-9.eE+]+)\s*,\s*([-0-9.eE+]+)\s*\);", src)
    (OUT / "sab_alpha_integrals.json").write_text(json.dumps(dict(
        source="tests/src/app_sabutils/main.cc:88-135 (mpmath, 30 digits): y -> exact "
               "integral of the log-linear interpolant over [0,1] with s1=(1-y)/2, s2=(1+y)/2",
        tolerance_fast=1e-14, values=[[float(a), float(b)] for a, b in pairs]), indent=1))
    return len(pairs)


def _frac(t):
    if "/" in t:
        num, den = t.split("/")
        return float(num) / float(den)
    return float(t)


def _ncmat_positions(path):
    """@ATOMPOSITIONS of an NCMAT file: {element: [[x, y, z], ...]} in file order"""
    out, on = {}, False
    for raw in path.read_text().splitlines():
        line = raw.split("#")[0].strip()
        if not line:
            continue
        if line.startswith("@"):
            on = line.startswith("@ATOMPOSITIONS")
            continue
        if on:
            t = line.split()
            out.setdefault(t[0], []).append([_frac(x) for x in t[1:4]])
    return out


def spacegroup_hkl():
    """laz tables of tests/scripts/ncmat2hkl.log: per material the header
    values (space group, cell, temperature, effective Debye temperature),
    the per-element coherent scattering lengths and masses of the log's info
    dump, the NCMAT file's atom positions and the table (h k l mult d F2)."""
    lines = (REF / "tests/scripts/ncmat2hkl.log").read_text().splitlines()
    heads = [i for i, l in enumerate(lines) if l.startswith("====> Conversion of stdlib::") and " to laz " in l]
    mats = {}
    for i0 in heads:
        m = re.match(r"====> Conversion of stdlib::(\S+?\.ncmat)(?:;dcutoff=([0-9.]+))? to laz", lines[i0])
        fname, dcut = m.group(1), m.group(2)
        if dcut is None:
            continue  # automatic dcutoff: not a fixed input
        hdr, table, atoms, snipped, tail = {}, [], {}, 0, []
        j = i0 + 1
        while not lines[j].startswith("====> NCrystal dump of converted laz"):
            l = lines[j]
            mh = re.match(r"^(?:STDOUT|res|OUTFILE)># (\w+) ([-0-9.eE+]+)", l)
            if mh and mh.group(1) in ("temperature", "spacegroup", "lattice_a", "lattice_b", "lattice_c",
                                      "lattice_aa", "lattice_bb", "lattice_cc", "debye_temperature",
                                      "average_mass"):
                hdr[mh.group(1)] = float(mh.group(2))
            mt = re.match(r"^(?:STDOUT|res|OUTFILE)>\s+(-?\d+)\s+(-?\d+)\s+(-?\d+)\s+(\d+)\s+([0-9.eE+-]+)\s+([0-9.eE+-]+)\s*$", l)
            if mt:
                row = [int(mt.group(1)), int(mt.group(2)), int(mt.group(3)), int(mt.group(4)),
                       float(mt.group(5)), float(mt.group(6))]
                (tail if snipped else table).append(row)
            ms = re.match(r"^(?:STDOUT|res|OUTFILE)>\s+<<<SNIPPED (\d+) LINES", l)
            if ms:
                snipped = int(ms.group(1))
            j += 1
        while not lines[j].startswith("Atom data:"):
            j += 1
        j += 1
        while lines[j].startswith("   "):
            ma = re.match(r"^\s+(\w+)\s+= \w+\(cohSL=([-0-9.eE+]+)fm .* mass=([0-9.eE+-]+)u", lines[j])
            atoms[ma.group(1)] = dict(csl_fm=float(ma.group(2)), mass=float(ma.group(3)))
            j += 1
        pos = _ncmat_positions(REF / "data" / fname)
        species = [dict(element=e, csl_fm=atoms[e]["csl_fm"], mass=atoms[e]["mass"], positions=p)
                   for e, p in pos.items()]
        mats[fname] = dict(spacegroup=int(hdr["spacegroup"]), dcutoff=float(dcut),
                           cell=[hdr["lattice_a"], hdr["lattice_b"], hdr["lattice_c"], hdr["lattice_aa"],
                                 hdr["lattice_bb"], hdr["lattice_cc"]],
                           temperature=hdr["temperature"], debye_temperature=hdr["debye_temperature"],
                           average_mass=hdr["average_mass"], species=species, table=table,
                           snipped=snipped, tail=tail)
    (OUT / "hkl_spacegroup.json").write_text(json.dumps(dict(
        source="tests/scripts/ncmat2hkl.log (laz tables at 10 significant digits, header values, info-dump "
               "atom data) with atom positions from data/<file>.ncmat; F2 in barn, d in Aa, csl in fm. The "
               "reference's MSDs come from each element's VDOS; debye_temperature is the effective, "
               "sigma-weighted value the log prints. The log shows long tables as head rows ('table'), "
               "a count of snipped rows and tail rows ('tail').", materials=mats), indent=1))
    return {k: len(v["table"]) + v["snipped"] + len(v["tail"]) for k, v in mats.items()}


def pointwisedist():
    # main.cc: x = {0,1,2,3,4}, weights all 4, 117 percentiles r = linspace(0,1,117)
    rows = []
    for l in (REF / "tests/src/app_pointwisedist/test.log").read_text().splitlines():
        m = re.match(r"percentile\( ([0-9.eE+-]+) \) = ([0-9.eE+-]+)$", l)
        if m:
            rows.append([float(m.group(1)), float(m.group(2))])
    (OUT / "pointwisedist_uniform.json").write_text(json.dumps(dict(
        source="tests/src/app_pointwisedist/main.cc (inputs) and test.log (printed %g values)",
        x=[0, 1, 2, 3, 4], w=[4, 4, 4, 4, 4], npoints=117, printed=rows), indent=1))
    return len(rows)


def sabsliceinteg():
    # main.cc: alpha grid 0.1..2.0 step 0.1, S = exp(alpha); checkXS(a0, a1) cases in order
    ranges = [(0.1, 2.0), (0.2, 1.8), (0.1123, 1.8), (0.112, 1.8123), (0.123, 1.8123), (0.15, 1.85),
              (0.1234, 0.1567), (0.20000001, 0.2567), (0.2, 0.2567), (0.2, 0.5567), (0.2, 1.4),
              (100.0, 101.0), (1e-10, 1e-9), (0.05, 1.123), (0.05, 2.3), (1.05, 2.3)]
    cases = []
    lines = (REF / "tests/src/app_sabsliceinteg/test.log").read_text().splitlines()
    idx = None
    for l in lines:
        m = re.match(r"makeTailedBreakdown DBG: aidx_low=(\d+), aidx_upp=(\d+)", l)
        if m:
            idx = [int(m.group(1)), int(m.group(2))]
            continue
        m = re.match(r" checkXS\((\d+)\): xs_front = (\S+) xs_middle = (\S+) xs_back = (\S+) xs_sum = (\S+)"
                     r" expected = (\S+) reldiff<1e-15\? = (\w+)", l)
        if m:
            k = int(m.group(1))
            cases.append(dict(alow=ranges[k - 1][0], aupp=ranges[k - 1][1], aidx=idx,
                              front=float(m.group(2)), middle=float(m.group(3)), back=float(m.group(4)),
                              total=float(m.group(5)), expected=float(m.group(6)), ok=m.group(7) == "yes"))
    (OUT / "sabsliceinteg.json").write_text(json.dumps(dict(
        source="tests/src/app_sabsliceinteg/main.cc:128-177 (inputs) and test.log (printed values, 6 digits)",
        alpha=[round(0.1 * i, 1) for i in range(1, 21)], cases=cases), indent=1))
    return len(cases)


if __name__ == "__main__":
    print("pointwisedist percentiles:", pointwisedist())
    print("sabsliceinteg cases:", sabsliceinteg())
    print("space-group hkl tables:", spacegroup_hkl())
    print("nosg families:", nosg())
    print("debye msd values:", debyemsd())
    print("alpha integral known answers:", sabutils())
