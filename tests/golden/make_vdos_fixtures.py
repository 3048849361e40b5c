"""Writes tests/golden/vdos_reference.json: phonon densities of states from
the reference's data library with the values its own test logs record for
them.  Run in a checkout that holds the reference under /root/reference:

    python tests/golden/make_vdos_fixtures.py

Read (as data, nothing is executed):
  * data/<material>.ncmat: the @DYNINFO sections (element, type vdos with
    vdos_egrid / vdos_density, or type vdosdebye with debye_temp);
  * ncrystal_core/src/atomdb/NCAtomDB.cc: natural-element masses
    (addNatElem lines);
  * tests/scripts/analysevdos.log: gamma0, integral, msd and effective
    temperature per material, element and temperature;
  * tests/scripts/vdosgn.log: moments of G1, G2, G5, G8, G20 and G50;
  * tests/scripts/vdosregularise.log: the regularised grid (emin, emax, point
    count) and first/last density values of every data-library VDOS;
  * tests/scripts/vdos2knl.log: kernel grid sizes and suggested Emax for the
    Debye VDOS of 400 K;
  * tests/scripts/basic.log: aluminium cross sections (PowderBragg at 1-4 Aa,
    absorption, incoherent elastic + inelastic at 4 Aa) with the cell of
    data/Al_sg225.ncmat.
"""
import json
import pathlib
import re

REF = pathlib.Path("/root/reference")
OUT = pathlib.Path(__file__).resolve().parent / "vdos_reference.json"


def masses():
    txt = (REF / "ncrystal_core/src/atomdb/NCAtomDB.cc").read_text()
    out = {}
    for m in re.finditer(r"addNatElem\(\s*(\d+)\s*,\s*([0-9.eE+-]+)\s*,[^;]*;\s*//\s*([A-Za-z]+)", txt):
        out[m.group(3)] = float(m.group(2))
    return out


def _num(tok):
    if "/" in tok:
        a, b = tok.split("/")
        return float(a) / float(b)
    return float(tok)


def _nums(toks):
    """numbers of an NCMAT field; "XrN" is the value X repeated N times"""
    out = []
    for t in toks:
        if "r" in t:
            v, n = t.split("r")
            out.extend([_num(v)] * int(n))
        else:
            out.append(_num(t))
    return out


def dyninfos(path):
    """[(element, dict of fields)] of the @DYNINFO sections, and the
    @DEBYETEMPERATURE values"""
    sections, cur, name = [], None, None
    for raw in path.read_text().splitlines():
        line = raw.split("#", 1)[0].strip()
        if not line:
            continue
        if line.startswith("@"):
            name = line[1:].split()[0]
            cur = {"_section": name, "_tokens": []}
            sections.append(cur)
            continue
        if cur is not None:
            cur["_tokens"].append(line.split())
    out, debye = [], {}
    for s in sections:
        lines = s["_tokens"]
        if s["_section"] == "DEBYETEMPERATURE":
            for toks in lines:
                for i in range(0, len(toks) - 1, 2):
                    debye[toks[i]] = float(toks[i + 1])
        if s["_section"] != "DYNINFO":
            continue
        # a field name opens a line; lines of numbers continue the field
        fields, key = {}, None
        for toks in lines:
            if re.match(r"^[a-z_][a-z_0-9]*$", toks[0]):
                key = toks[0]
                fields[key] = list(toks[1:])
            elif key is not None:
                fields[key].extend(toks)
        out.append((fields["element"][0], fields))
    return out, debye


def vdos_input(material, element, mass_table):
    path = REF / "data" / f"{material}.ncmat"
    dis, debye = dyninfos(path)
    for el, f in dis:
        if el != element:
            continue
        kind = f["type"][0]
        if element not in mass_table:  # isotopes (D, ...) are not natural elements
            return None
        rec = {"mass": mass_table[element]}
        if kind == "vdos":
            rec.update(kind="vdos", egrid=_nums(f["vdos_egrid"]), density=_nums(f["vdos_density"]))
            return rec
        if kind == "vdosdebye":
            if "debye_temp" in f:
                rec.update(kind="debye", debye_temp=_num(f["debye_temp"][0]))
                return rec
            if element in debye:
                rec.update(kind="debye", debye_temp=debye[element])
                return rec
        return None
    return None


def main():
    mass_table = masses()
    vdos = {}

    def key_for(material, element):
        k = f"{material}/{element}"
        if k not in vdos:
            rec = vdos_input(material, element, mass_table)
            if rec is None:
                return None
            vdos[k] = rec
        return k

    analyse = []
    for line in (REF / "tests/scripts/analysevdos.log").read_text().splitlines():
        m = re.match(r"==> (\S+?)\.ncmat;temp=([^;]+);vdoslux=(\d)//(\w+) => (.*)", line)
        if not m:
            continue
        k = key_for(m.group(1), m.group(4))
        if k is None:
            continue
        vals = dict(kv.split("=") for kv in m.group(5).split("; "))
        analyse.append(dict(vdos=k, temperature=float(m.group(2)), vdoslux=int(m.group(3)),
                            expected={kk: float(v) for kk, v in vals.items()}))

    gn = []
    cur = None
    for line in (REF / "tests/scripts/vdosgn.log").read_text().splitlines():
        m = re.match(r'Investigating "stdlib::(\S+?)\.ncmat;temp=([0-9.]+)K;comp=inelas;vdoslux=(\d)" / (\w+)', line)
        if m:
            k = key_for(m.group(1), m.group(4))
            cur = None if k is None else dict(vdos=k, temperature=float(m.group(2)), vdoslux=int(m.group(3)), orders={})
            if cur is not None:
                gn.append(cur)
            continue
        m = re.match(r'Distribution "G(\d+)" with:', line)
        if m and cur is not None:
            order = m.group(1)
            cur["orders"][order] = {}
            continue
        m = re.match(r"\s+(integral|mean|rms|skewness|kurtosis|\dth moment \(std\))\s*: (\S+)", line)
        if m and cur is not None and cur["orders"]:
            last = list(cur["orders"])[-1]
            cur["orders"][last][m.group(1)] = float(m.group(2))

    regul = []
    lines = (REF / "tests/scripts/vdosregularise.log").read_text().splitlines()
    for i, line in enumerate(lines):
        m = re.match(r"Validating VDOS regularisation of stdlib::(\S+?)\.ncmat / (\w+):", line)
        if not m:
            continue
        m2 = re.match(r"\s+egrid \[meV\]: (\S+) \.\. (\S+) \((\d+) points\)", lines[i + 1])
        m3 = re.match(r"\s+density: \[(.*)\]", lines[i + 2])
        vals = [v.strip() for v in m3.group(1).split(",")]
        regul.append(dict(material=m.group(1), element=m.group(2), emin_mev=float(m2.group(1)),
                          emax_mev=float(m2.group(2)), npts=int(m2.group(3)),
                          first3=[float(v) for v in vals[:3]], last3=[float(v) for v in vals[4:]]))
    # every 6th data-library entry, and all Debye ones with a given temperature
    keep = []
    for j, r in enumerate(regul):
        k = key_for(r["material"], r["element"])
        if k is None:
            continue
        if j % 6 == 0 or vdos[k]["kind"] == "debye":
            r["vdos"] = k
            keep.append(r)

    knl = []
    for block in (REF / "tests/scripts/vdos2knl.log").read_text().split("Calling ")[1:]:
        m = re.match(r"nc_vdos\.extractKnl\( m=(\S+), T=(\S+), vdoslux=(\d), target_emax=(\S+) \)", block)
        if not m:
            continue
        rec = dict(mass=float(m.group(1)), temperature=float(m.group(2)), vdoslux=int(m.group(3)),
                   target_emax=float(m.group(4)), debye_temp=400.0)
        sa = re.search(r"Got alpha : NumpyArray\( shape=(\d+) \)", block)
        sb = re.search(r"Got beta : NumpyArray\( shape=(\d+) \)", block)
        se = re.search(r"Got suggested_emax : (\S+)", block)
        dis = re.search(r"Discarding (\d+) edges", block)
        err = re.search(r"Caught expected NCCalcError\('(.*)'\)", block)
        if sa:
            rec.update(nalpha=int(sa.group(1)), nbeta=int(sb.group(1)), suggested_emax=float(se.group(1)),
                       discarded_edges=int(dis.group(1)) if dis else 0)
        elif err:
            rec.update(error=err.group(1))
        knl.append(rec)

    # tests/scripts/basic.log: aluminium from data/Al_sg225.ncmat at 293.15 K
    # (PowderBragg at 1-4 Aa, absorption and the "Bkgd" = incoherent elastic +
    # inelastic at 4 Aa); cell and atom data from that file and NCAtomDB.cc
    blog = (REF / "tests/scripts/basic.log").read_text()
    mpb = re.search(r"PowderBragg diffraction cross-section @ \[1, 2, 3, 4\]Aa: \[(.*?)\] barn", blog)
    mabs = re.search(r"absorption cross-section @ 4.0Aa: (\S+) barn", blog)
    mbk = re.search(r"Bkgd cross-section @ 4.0Aa: (\S+) barn", blog)
    mbp = re.search(r"Bkgd\+PowderBragg cross-section @ 4.0Aa: (\S+) barn", blog)
    mT = re.search(r"Temperature : (\S+)", blog)
    atom = re.search(r"addNatElem\(13, ([0-9.]+), ([0-9.]+), ([0-9.]+), ([0-9.]+)\);//Al",
                     (REF / "ncrystal_core/src/atomdb/NCAtomDB.cc").read_text())
    kal = key_for("Al_sg225", "Al")
    basic = dict(vdos=kal, temperature=float(mT.group(1)), spacegroup=225, a=4.04958,
                 positions=[[0, 0.5, 0.5], [0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5]],
                 mass=float(atom.group(1)), coh_sl=float(atom.group(2)), sigma_inc=float(atom.group(3)),
                 sigma_abs=float(atom.group(4)), dcutoff=0.1,
                 powderbragg_1to4=[float(v) for v in mpb.group(1).split(",")], absorption_4=float(mabs.group(1)),
                 bkgd_4=float(mbk.group(1)), bkgd_plus_bragg_4=float(mbp.group(1)))

    used = {r["vdos"] for r in analyse} | {r["vdos"] for r in gn} | {r["vdos"] for r in keep} | {kal}
    out = dict(vdos={k: v for k, v in vdos.items() if k in used}, analyse=analyse, gn=gn, regularise=keep,
               kernel=knl, basic=basic)
    OUT.write_text(json.dumps(out, indent=0))
    print(f"wrote {OUT}: {len(out['vdos'])} densities, {len(analyse)} analyse, {len(gn)} Gn, "
          f"{len(keep)} regularise, {len(knl)} kernel records")


if __name__ == "__main__":
    main()
