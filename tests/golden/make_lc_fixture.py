"""Writes tests/golden/lc_graphite.json from the reference's own test
outputs (data only):

* tests/src/app_lc/test.log: cross sections of the layered crystal
  "C_sg194_pyrolytic_graphite.ncmat;dcutoff=0.5;mos=4deg;
  dir1=@crys_hkl:0,0,1@lab:0,0,1;dir2=@crys_hkl:0,1,0@lab:0,1,0;lcaxis=0,0,1"
  at 43 wavelengths for the neutron direction (0, 1/sqrt 2, -1/sqrt 2), with
  lcmode -1, 0 and 20 (tests/src/app_lc/main.cc);
* tests/scripts/ncmat2hkl.log: the same material's cell, volume, sigma_coh,
  average mass and hkl family table (laz format; head and tail rows).

Run in a checkout next to the reference: python tests/golden/make_lc_fixture.py
"""
import json
import pathlib
import re

REF = pathlib.Path("/root/reference/tests")
OUT = pathlib.Path(__file__).parent / "lc_graphite.json"


def app_lc():
    txt = (REF / "src" / "app_lc" / "test.log").read_text()
    out = {}
    for block in txt.split("----------------- Testing cfg ")[1:]:
        mode = int(re.search(r"lcmode=(-?\d+)", block).group(1))
        wl, xs = [], []
        for m in re.finditer(r"xs@[0-9.e+-]+eV \(([0-9.e+-]+) Aa\): ([0-9.e+-]+)", block):
            wl.append(float(m.group(1)))
            xs.append(float(m.group(2)))
        out[str(mode)] = dict(wavelength=wl, xs=xs)
    return out


def laz():
    txt = (REF / "scripts" / "ncmat2hkl.log").read_text()
    start = txt.index("====> Conversion of stdlib::C_sg194_pyrolytic_graphite.ncmat to laz <=====")
    block = txt[start:txt.index("====> NCrystal dump of converted laz <=====", start)]
    hdr = {}
    for key in ("lattice_a", "lattice_c", "Vc", "sigma_coh", "average_mass", "temperature"):
        hdr[key] = float(re.search(r"# %s ([0-9.e+-]+)" % key, block).group(1))
    head, tail, seen_snip = [], [], False
    for line in block.splitlines():
        if "SNIPPED" in line:
            seen_snip = True
            snipped = int(re.search(r"SNIPPED (\d+) LINES", line).group(1))
            continue
        m = re.match(r"OUTFILE>\s+(-?\d+)\s+(-?\d+)\s+(-?\d+)\s+(\d+)\s+([0-9.e+-]+)\s+([0-9.e+-]+)\s*$", line)
        if m:
            row = [int(m.group(i)) for i in range(1, 5)] + [float(m.group(5)), float(m.group(6))]
            (tail if seen_snip else head).append(row)
    return dict(header=hdr, table=head, tail=tail, snipped=snipped)


if __name__ == "__main__":
    data = dict(source="reference tests/src/app_lc/test.log and tests/scripts/ncmat2hkl.log",
                direction=[0.0, 0.7071067811865475, -0.7071067811865475], mosaicity_deg=4.0, dcutoff=0.5,
                positions=[[0, 0, 0.25], [0, 0, 0.75], [1 / 3, 2 / 3, 0.25], [2 / 3, 1 / 3, 0.75]],
                spacegroup=194, lcaxis=[0.0, 0.0, 1.0], app_lc=app_lc(), laz=laz())
    OUT.write_text(json.dumps(data, indent=1))
    print("wrote", OUT)
