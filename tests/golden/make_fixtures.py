"""Extract golden vectors from the reference's own test logs into small JSON
fixtures under tests/golden/.  Run in a container where the reference
repository is mounted (default /root/reference); the fixtures it writes are
data (inputs and expected outputs) and are committed.

Sources:
  tests/scripts/nosg.log                 Al hkl families, no space group, dcutoff 0.4
  tests/src/app_debyemsd/main.cc,test.log  sqrt(MSD) table (Debye model)
  tests/src/app_sabutils/main.cc:88-135   mpmath known answers of the alpha-bin integral
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


if __name__ == "__main__":
    print("nosg families:", nosg())
    print("debye msd values:", debyemsd())
    print("alpha integral known answers:", sabutils())
