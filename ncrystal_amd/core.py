"""Python mirror of NCrystal's Process/Scatter interface for the GPU path.

Method names, argument meaning and output shapes follow
ncrystal_python/src/NCrystal/core.py (Process.crossSectionIsotropic /
xsect :1277-1310, Scatter.sampleScatterIsotropic :1492-1504,
Scatter.sampleScatter :1478-1489); the heavy lifting is one batched call
into libncrystal_amd per method call.
"""
import ctypes

import numpy as np

from . import _capi as C
from .synth import WL2EKIN


def wl2ekin(wl):
    wl = np.asarray(wl, dtype=np.float64)
    return WL2EKIN / (wl * wl)


class Process:
    """Owns one ncamd_process_t handle (reference counted on the C side)."""

    def __init__(self, handle, name="process", keep=()):
        if not handle.internal:
            C.check()
            raise C.NCAMDError("process creation failed")
        self._h = handle
        self.name = name
        self._keep = keep

    @property
    def handle(self):
        return self._h

    def __del__(self):
        try:
            if self._h is not None and self._h.internal:
                C.lib().ncamd_unref(ctypes.byref(self._h))
        except Exception:
            pass

    def refCount(self):
        return C.call("ncamd_refcount", self._h)

    # --- cross sections
    def crossSectionIsotropic(self, ekin, repeat=None):
        scalar = np.ndim(ekin) == 0 and repeat is None
        e = C.as_f64(np.atleast_1d(ekin))
        rep = 1 if repeat is None else int(repeat)
        out = np.empty(e.size * rep)
        C.call("ncamd_crosssection_nonoriented_many", self._h, C.dptr(e), e.size, rep, C.dptr(out))
        if scalar:
            return float(out[0])
        return out if repeat is not None else out.reshape(np.shape(ekin))

    def crossSectionIsotropic_fp32(self, ekin):
        e = C.as_f64(np.atleast_1d(ekin))
        out = np.empty(e.size)
        C.call("ncamd_crosssection_nonoriented_many_fp32", self._h, C.dptr(e), e.size, 1, C.dptr(out))
        return out

    def crossSection(self, ekin, direction):
        e = C.as_f64(np.atleast_1d(ekin))
        d = C.as_f64(np.broadcast_to(np.asarray(direction, dtype=np.float64), (e.size, 3)))
        out = np.empty(e.size)
        C.call("ncamd_crosssection_many_dirs", self._h, C.dptr(e), C.dptr(d), e.size, C.dptr(out))
        return out

    def xsect(self, ekin=None, wl=None, direction=None, repeat=None):
        if (ekin is None) == (wl is None):
            raise C.NCBadInput("specify exactly one of ekin and wl")
        e = wl2ekin(wl) if wl is not None else np.asarray(ekin, dtype=np.float64)
        if direction is None:
            return self.crossSectionIsotropic(e, repeat=repeat)
        return self.crossSection(e, direction)

    # --- device-resident batched calls (pointers to device memory, stream)
    def dev_crossSectionIsotropic(self, d_ekin, n, d_out, stream=0, repeat=1, fp32=False):
        fn = "ncamd_dev_crosssection_nonoriented_many_fp32" if fp32 else "ncamd_dev_crosssection_nonoriented_many"
        C.call(fn, self._h, d_ekin, n, repeat, d_out, stream)


class Scatter(Process):
    def setRandomSeed(self, seed):
        C.call("ncamd_setrngseed", self._h, int(seed))

    def getRNGState(self):
        s = ctypes.c_uint64()
        c = ctypes.c_uint64()
        C.call("ncamd_getrngstate", self._h, ctypes.byref(s), ctypes.byref(c))
        return s.value, c.value

    def setRNGState(self, seed, callcounter):
        C.call("ncamd_setrngstate", self._h, int(seed), int(callcounter))

    def sampleScatterIsotropic(self, ekin, repeat=None):
        e = C.as_f64(np.atleast_1d(ekin))
        rep = 1 if repeat is None else int(repeat)
        oe = np.empty(e.size * rep)
        om = np.empty(e.size * rep)
        C.call("ncamd_samplescatterisotropic_many", self._h, C.dptr(e), e.size, rep, C.dptr(oe), C.dptr(om))
        if np.ndim(ekin) == 0 and repeat is None:
            return float(oe[0]), float(om[0])
        return oe, om

    def sampleScatter(self, ekin, direction, repeat=None):
        if np.ndim(ekin) == 0 and np.ndim(direction) == 1:
            rep = 1 if repeat is None else int(repeat)
            oe, ox, oy, oz = (np.empty(rep) for _ in range(4))
            d = (ctypes.c_double * 3)(*[float(x) for x in direction])
            C.call("ncamd_samplescatter_many", self._h, float(ekin), ctypes.byref(d), rep, C.dptr(oe),
                   C.dptr(ox), C.dptr(oy), C.dptr(oz))
            if repeat is None:
                return float(oe[0]), (float(ox[0]), float(oy[0]), float(oz[0]))
            return oe, np.stack([ox, oy, oz], axis=1)
        e = C.as_f64(np.atleast_1d(ekin))
        d = C.as_f64(np.broadcast_to(np.asarray(direction, dtype=np.float64), (e.size, 3)))
        oe = np.empty(e.size)
        od = np.empty((e.size, 3))
        C.call("ncamd_samplescatter_many_dirs", self._h, C.dptr(e), C.dptr(d), e.size, C.dptr(oe), C.dptr(od))
        return oe, od

    def dev_sampleScatterIsotropic(self, d_ekin, n, d_oe, d_omu, stream=0, repeat=1):
        C.call("ncamd_dev_samplescatterisotropic_many", self._h, d_ekin, n, repeat, d_oe, d_omu, stream)

    def dev_sampleScatter(self, d_ekin, d_dirs, n, d_oe, d_od, stream=0):
        """per-neutron directions in and out (3 doubles per neutron)"""
        C.call("ncamd_dev_samplescatter_many_dirs", self._h, d_ekin, d_dirs, n, d_oe, d_od, stream)

    def dev_check_errors(self, stream=0):
        return C.call("ncamd_dev_check_errors", self._h, stream)


# ---------------------------------------------------------------- factories

def debye_msd(debye_temp, temperature, mass):
    return C.call("ncamd_debye_msd", float(debye_temp), float(temperature), float(mass))


def hkl_orthogonal(a, b, c, species, dcutoff, fsquarecut=1e-5, merge_tolerance=1e-6):
    """species: list of (csl [sqrt(barn)], msd [Aa^2], positions (n,3)).
    Returns (hkl[n,3], mult[n], dspacing[n], fsquared[n]) sorted like
    NCrystal's InfoBuilder."""
    natoms = np.array([len(s[2]) for s in species], dtype=np.int32)
    pos = C.as_f64(np.concatenate([np.asarray(s[2], dtype=np.float64).reshape(-1, 3) for s in species]).ravel())
    csl = C.as_f64([s[0] for s in species])
    msd = C.as_f64([s[1] for s in species])
    maxout = 4096
    req = ctypes.c_int64()
    while True:
        hkl = np.empty(3 * maxout, dtype=np.int32)
        mult = np.empty(maxout, dtype=np.int32)
        dsp = np.empty(maxout)
        fsq = np.empty(maxout)
        n = C.call("ncamd_hkl_orthogonal", a, b, c, len(species), C.iptr(natoms), C.dptr(pos), C.dptr(csl),
                   C.dptr(msd), dcutoff, fsquarecut, merge_tolerance, maxout, C.iptr(hkl), C.iptr(mult),
                   C.dptr(dsp), C.dptr(fsq), ctypes.byref(req))
        if n == -1:
            maxout = int(req.value)
            continue
        if n == -2:
            raise C.NCCalcError(f"Combinatorics too great to reach dcutoff = {dcutoff} Aa")
        if n < 0:
            raise C.NCAMDError("hkl enumeration failed")
        return hkl[:3 * n].reshape(n, 3).copy(), mult[:n].copy(), dsp[:n].copy(), fsq[:n].copy()


def hkl_cell_members(cell, species, dcutoff, fsquarecut=1e-5, merge_tolerance=1e-6):
    """hkl_orthogonal_members for any cell = (a, b, c, alpha, beta, gamma)
    (enumeration without symmetry over the reference's reciprocal lattice)."""
    return hkl_orthogonal_members(*cell[:3], species, dcutoff, fsquarecut, merge_tolerance, _angles=tuple(cell[3:]))


def hkl_orthogonal_members(a, b, c, species, dcutoff, fsquarecut=1e-5, merge_tolerance=1e-6, _angles=None):
    """As hkl_orthogonal, plus the explicit hkl members of every family:
    returns (hkl, mult, dspacing, fsquared, mem_off[n+1], mem_hkl[m,3])."""
    natoms = np.array([len(s[2]) for s in species], dtype=np.int32)
    pos = C.as_f64(np.concatenate([np.asarray(s[2], dtype=np.float64).reshape(-1, 3) for s in species]).ravel())
    csl = C.as_f64([s[0] for s in species])
    msd = C.as_f64([s[1] for s in species])
    maxout, maxmem = 4096, 65536
    req, mreq = ctypes.c_int64(), ctypes.c_int64()
    while True:
        hkl = np.empty(3 * maxout, dtype=np.int32)
        mult = np.empty(maxout, dtype=np.int32)
        dsp = np.empty(maxout)
        fsq = np.empty(maxout)
        moff = np.empty(maxout + 1, dtype=np.int64)
        mhkl = np.empty(3 * maxmem, dtype=np.int32)
        fn, cellargs = (("ncamd_hkl_orthogonal_members", (a, b, c)) if _angles is None else
                        ("ncamd_hkl_cell_members", (a, b, c) + tuple(float(x) for x in _angles)))
        n = C.call(fn, *cellargs, len(species), C.iptr(natoms), C.dptr(pos),
                   C.dptr(csl), C.dptr(msd), dcutoff, fsquarecut, merge_tolerance, maxout, C.iptr(hkl),
                   C.iptr(mult), C.dptr(dsp), C.dptr(fsq), ctypes.byref(req), maxmem, C.i64ptr(moff),
                   C.iptr(mhkl), ctypes.byref(mreq))
        if n == -1:
            maxout = max(maxout, int(req.value))
            maxmem = max(maxmem, int(mreq.value))
            continue
        if n == -2:
            raise C.NCCalcError(f"Combinatorics too great to reach dcutoff = {dcutoff} Aa")
        if n < 0:
            raise C.NCAMDError("hkl enumeration failed")
        nm = int(moff[n])
        return (hkl[:3 * n].reshape(n, 3).copy(), mult[:n].copy(), dsp[:n].copy(), fsq[:n].copy(),
                moff[:n + 1].copy(), mhkl[:3 * nm].reshape(nm, 3).copy())


def hkl_spacegroup_members(spacegroup, cell, species, dcutoff, fsquarecut=1e-5):
    """hkl families of a crystal with a space group (NCrystal's
    calculateHKLPlanesWithSymEqRefl, ncrystal_core/src/extd_utils/NCFillHKL.cc:545-709):
    cell = (a, b, c, alpha, beta, gamma) in Aa and degrees, species as for
    hkl_orthogonal.  Returns (hkl, mult, dspacing, fsquared, mem_off[n+1],
    mem_hkl[m,3]); the members of a family are its equivalent reflections, one
    per +-pair."""
    natoms = np.array([len(s[2]) for s in species], dtype=np.int32)
    pos = C.as_f64(np.concatenate([np.asarray(s[2], dtype=np.float64).reshape(-1, 3) for s in species]).ravel())
    csl = C.as_f64([s[0] for s in species])
    msd = C.as_f64([s[1] for s in species])
    a, b, c, al, be, ga = (float(x) for x in cell)
    maxout, maxmem = 4096, 65536
    req, mreq = ctypes.c_int64(), ctypes.c_int64()
    while True:
        hkl = np.empty(3 * maxout, dtype=np.int32)
        mult = np.empty(maxout, dtype=np.int32)
        dsp = np.empty(maxout)
        fsq = np.empty(maxout)
        moff = np.empty(maxout + 1, dtype=np.int64)
        mhkl = np.empty(3 * maxmem, dtype=np.int32)
        n = C.call("ncamd_hkl_spacegroup_members", int(spacegroup), a, b, c, al, be, ga, len(species),
                   C.iptr(natoms), C.dptr(pos), C.dptr(csl), C.dptr(msd), dcutoff, fsquarecut, maxout, C.iptr(hkl),
                   C.iptr(mult), C.dptr(dsp), C.dptr(fsq), ctypes.byref(req), maxmem, C.i64ptr(moff),
                   C.iptr(mhkl), ctypes.byref(mreq))
        if n == -1:
            maxout = max(maxout, int(req.value))
            maxmem = max(maxmem, int(mreq.value))
            continue
        if n == -2:
            raise C.NCCalcError(f"Combinatorics too great to reach dcutoff = {dcutoff} Aa")
        if n < 0:
            raise C.NCAMDError("hkl enumeration failed")
        nm = int(moff[n])
        return (hkl[:3 * n].reshape(n, 3).copy(), mult[:n].copy(), dsp[:n].copy(), fsq[:n].copy(),
                moff[:n + 1].copy(), mhkl[:3 * nm].reshape(nm, 3).copy())


def reci_lattice(cell):
    """Reciprocal lattice (row-major 3x3) of cell = (a, b, c, alpha, beta,
    gamma) as ncrystal_core/src/utils/NCLatticeUtils.cc:95-149 builds it:
    closed forms for 90/90/90, 90/90/120 and 90/beta/90 cells, 2 pi times the
    inverse of the reference's lattice matrix for any other cell."""
    out = np.empty(9)
    C.call("ncamd_reci_lattice", *[float(x) for x in cell], C.dptr(out))
    return out


def reci_lattice_orthogonal(a, b, c):
    """Reciprocal lattice of a cell with 90-degree angles, row-major 3x3
    (diag(2pi/a, 2pi/b, 2pi/c), ncrystal_core/src/utils/NCLatticeUtils.cc:95-117)."""
    k = 2.0 * np.pi
    return np.array([k / a, 0.0, 0.0, 0.0, k / b, 0.0, 0.0, 0.0, k / c])


class SCBragg(Scatter):
    """Oriented mosaic single crystal (NCrystal's SCBragg).  Use
    crossSection / sampleScatter / xsect(direction=...)."""

    def info(self):
        nf, nn, thr = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_double()
        gos = np.empty(15)
        C.call("ncamd_scbragg_info", self._h, ctypes.byref(nf), ctypes.byref(nn), ctypes.byref(thr), C.dptr(gos))
        return dict(nfam=nf.value, nnormals=nn.value, threshold=thr.value, gos=gos)

    def tables(self):
        inf = self.info()
        nf, nn, nlt = inf["nfam"], inf["nnormals"], int(inf["gos"][10])
        inv2d, xsf = np.empty(nf), np.empty(nf)
        off = np.empty(nf + 1, dtype=np.int64)
        nrm = np.empty((nn, 3))
        lt1, lt2 = np.empty(2 * nlt), np.empty(2 * nlt)
        C.call("ncamd_scbragg_tables", self._h, C.dptr(inv2d), C.dptr(xsf), C.i64ptr(off), C.dptr(nrm),
               C.dptr(lt1), C.dptr(lt2))
        return dict(inv2d=inv2d, xsfact=xsf, off=off, normals=nrm, lt_evalcosx=lt1.reshape(-1, 2),
                    lt_sofcosd=lt2.reshape(-1, 2))

    def dev_crossSection(self, d_ekin, d_dirs, n, d_out, stream=0):
        C.call("ncamd_dev_crosssection_many_dirs", self._h, d_ekin, d_dirs, n, d_out, stream)

    def dev_sampleScatter(self, d_ekin, d_dirs, n, d_oe, d_od, stream=0):
        C.call("ncamd_dev_samplescatter_many_dirs", self._h, d_ekin, d_dirs, n, d_oe, d_od, stream)


def create_scbragg(v0_times_natoms, dspacing, fsquared, mem_off, mem_hkl, reci_lattice, cry2lab,
                   mosaicity_fwhm, prec=1e-3, ntrunc=0.0):
    """SCBragg from hkl families with explicit members (see
    hkl_orthogonal_members); cry2lab: 3x3 crystal-to-lab rotation;
    mosaicity_fwhm in radians."""
    d, f = C.as_f64(dspacing), C.as_f64(fsquared)
    mo = np.ascontiguousarray(mem_off, dtype=np.int64)
    mh = np.ascontiguousarray(mem_hkl, dtype=np.int32).ravel()
    L = C.as_f64(np.asarray(reci_lattice, dtype=np.float64).ravel())
    R = C.as_f64(np.asarray(cry2lab, dtype=np.float64).ravel())
    h = C.call("ncamd_create_scbragg", float(v0_times_natoms), d.size, C.dptr(d), C.dptr(f), C.i64ptr(mo),
               C.iptr(mh), C.dptr(L), C.dptr(R), float(mosaicity_fwhm), float(prec), float(ntrunc))
    return SCBragg(h, name="SCBragg")


class LCBragg(Scatter):
    """Layered crystal by one of NCrystal's reference models
    (ncrystal_core/src/lcbragg/NCLCBragg.cc:62-73): a single crystal
    rotated around the layer axis.  Use crossSection / sampleScatter."""

    def __init__(self, h, crystal, name="LCBragg"):
        super().__init__(h, name=name)
        self.crystal = crystal

    def dev_crossSection(self, d_ekin, d_dirs, n, d_out, stream=0):
        C.call("ncamd_dev_crosssection_many_dirs", self._h, d_ekin, d_dirs, n, d_out, stream)

    def dev_sampleScatter(self, d_ekin, d_dirs, n, d_oe, d_od, stream=0):
        C.call("ncamd_dev_samplescatter_many_dirs", self._h, d_ekin, d_dirs, n, d_oe, d_od, stream)


def create_lcbragg_ref(crystal, lcaxis_lab, mode):
    """LCBragg with lcmode = mode != 0 over an SCBragg process `crystal`
    (NCLCRefModels.cc): mode > 0 averages cross sections over the smallest
    prime >= mode of equally spaced rotations around lcaxis_lab and scatters
    over `mode` random ones picked by cross section; mode < 0 uses -mode
    random rotations for both."""
    ax = C.as_f64(np.asarray(lcaxis_lab, dtype=np.float64).ravel())
    h = C.call("ncamd_create_lcbragg_ref", crystal._h, C.dptr(ax), int(mode))
    return LCBragg(h, crystal)


def create_lcbragg(v0_times_natoms, dspacing, fsquared, mem_off, mem_hkl, reci_lattice, cry2lab, lcaxis_crystal,
                   mosaicity_fwhm, prec=1e-3, ntrunc=0.0, mode=0):
    """LCBragg from the inputs of create_scbragg and the layer axis in the
    crystal frame (NCLCBragg.cc:36-73): mode 0 the layered-crystal model
    (cross sections), else a reference model (see create_lcbragg_ref)."""
    d, f = C.as_f64(dspacing), C.as_f64(fsquared)
    mo = np.ascontiguousarray(mem_off, dtype=np.int64)
    mh = np.ascontiguousarray(mem_hkl, dtype=np.int32).ravel()
    L = C.as_f64(np.asarray(reci_lattice, dtype=np.float64).ravel())
    R = C.as_f64(np.asarray(cry2lab, dtype=np.float64).ravel())
    ax = C.as_f64(np.asarray(lcaxis_crystal, dtype=np.float64).ravel())
    h = C.call("ncamd_create_lcbragg", float(v0_times_natoms), d.size, C.dptr(d), C.dptr(f), C.i64ptr(mo),
               C.iptr(mh), C.dptr(L), C.dptr(R), C.dptr(ax), float(mosaicity_fwhm), float(prec), float(ntrunc),
               int(mode))
    return LCBragg(h, None)


def create_powderbragg(v0_times_natoms, dspacings, fsq_x_mult, fp32_tables=True):
    d = C.as_f64(dspacings)
    f = C.as_f64(fsq_x_mult)
    h = C.call("ncamd_create_powderbragg", float(v0_times_natoms), C.dptr(d), C.dptr(f), d.size,
               1 if fp32_tables else 0)
    return Scatter(h, name="PowderBragg")


def powderbragg_tables(proc):
    n = C.call("ncamd_powderbragg_tables", proc.handle, None, None, None, 0)
    e2d, fdm, f32 = np.empty(n), np.empty(n), np.empty(n, dtype=np.float32)
    C.call("ncamd_powderbragg_tables", proc.handle, C.dptr(e2d), C.dptr(fdm), C.fptr(f32), n)
    return e2d, fdm, f32


def create_elincscatter(msd, bound_incoh_xs, scale):
    m, b, s = C.as_f64(msd), C.as_f64(bound_incoh_xs), C.as_f64(scale)
    h = C.call("ncamd_create_elincscatter", m.size, C.dptr(m), C.dptr(b), C.dptr(s))
    return Scatter(h, name="ElIncScatter")


def create_freegas(temperature, mass_amu, sigma_free):
    """FreeGas scatter process (NCrystal's FreeGas(Temperature, AtomMass,
    SigmaFree)); sigma_free in barn."""
    h = C.call("ncamd_create_freegas", float(temperature), float(mass_amu), float(sigma_free))
    return Scatter(h, name="FreeGas")


def create_absoov(sigma_abs):
    """1/v absorption process (NCrystal's AbsOOV); cross sections only."""
    h = C.call("ncamd_create_absoov", float(sigma_abs))
    return Process(h, name="AbsOOV")


class SABScatter(Scatter):
    def egrid(self):
        n = C.call("ncamd_sab_egrid", self._h, None, None, 0)
        eg, xs = np.empty(n), np.empty(n)
        C.call("ncamd_sab_egrid", self._h, C.dptr(eg), C.dptr(xs), n)
        return eg, xs

    def derived(self, na, nb):
        ls = np.empty(na * nb)
        cu = np.empty(na * nb)
        C.call("ncamd_sab_derived", self._h, C.dptr(ls), C.dptr(cu))
        return ls.reshape(nb, na), cu.reshape(nb, na)

    def sampler(self, ie, nb):
        kind, ibl, nv = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        fb = ctypes.c_double()
        n = nb + 1
        bx, by, bc = np.empty(n), np.empty(n), np.empty(n)
        idbl = np.empty(8 * nb)
        iidx = np.empty(2 * nb, dtype=np.int32)
        C.call("ncamd_sab_sampler", self._h, ie, ctypes.byref(kind), ctypes.byref(ibl), ctypes.byref(nv),
               ctypes.byref(fb), C.dptr(bx), C.dptr(by), C.dptr(bc), C.dptr(idbl), C.iptr(iidx))
        res = dict(kind=kind.value, ibeta_low=ibl.value, nvals=nv.value, firstbin=fb.value)
        if kind.value == 1:
            k = nv.value
            res.update(bx=bx[:k], by=by[:k], bcdf=bc[:k], info=idbl.reshape(-1, 8)[:k - 1],
                       infoidx=iidx.reshape(-1, 2)[:k - 1])
        return res

    def dev_sample_alphabeta(self, d_ekin, n, d_a, d_b, stream=0):
        C.call("ncamd_dev_sab_sample_alphabeta", self._h, d_ekin, n, d_a, d_b, stream)

    # exact envelope sampler (box rejection in kinematic coordinates) as the sampling mode
    def use_envelope_sampler(self, nintervals=256, emin=0.0, emax=0.0, nbox_w=128, nbox_k=32, maxtries=64):
        C.call("ncamd_sab_use_envelope_sampler", self._h, int(nintervals), float(emin), float(emax), int(nbox_w),
               int(nbox_k), int(maxtries))
        self._kb = (int(nintervals), int(nbox_w) * int(nbox_k))

    def use_learned_sampler(self, model, nintervals=256, emin=0.0, emax=0.0, nbox_w=128, nbox_k=32, maxtries=64):
        """the envelope sampler on box grids placed by the trained model"""
        C.call("ncamd_sab_use_learned_sampler", self._h, model.handle, int(nintervals), float(emin), float(emax),
               int(nbox_w), int(nbox_k), int(maxtries))
        self._kb = (int(nintervals), int(nbox_w) * int(nbox_k))

    def use_plain_sampler(self):
        C.call("ncamd_sab_use_plain_sampler", self._h)

    def envelope_tables(self):
        """interval edges [nn+1] and boxes [nn, nbox, 6]: w0, w1, kappa0, kappa1, upper, lower bound"""
        nn, nbox = self._kb
        node = np.empty(nn + 1)
        boxes = np.empty(nn * nbox * 6)
        C.call("ncamd_sab_envelope_tables", self._h, C.dptr(node), C.dptr(boxes))
        return node, boxes.reshape(nn, nbox, 6)

    def envelope_stats(self):
        """per interval: envelope mass in use, of the uniform grid, uniform blend weight, squeeze mass"""
        out = np.empty(4 * self._kb[0])
        C.call("ncamd_sab_envelope_stats", self._h, C.dptr(out))
        return out.reshape(-1, 4)

    def set_envelope_squeeze(self, on):
        C.call("ncamd_sab_envelope_squeeze", self._h, int(bool(on)))

    def dev_sample_alphabeta_envelope(self, d_ekin, n, d_a, d_b, stream=0):
        return C.call("ncamd_dev_sab_sample_alphabeta_envelope", self._h, d_ekin, n, d_a, d_b, stream)


def sab_std_format(knltype, alpha, beta, sab, beta_grid_optimised=False):
    """Scattering-kernel data to the standard S(alpha,beta) table
    (transformKernelToStdFormat, ncrystal_core/src/sab/NCSABUtils.cc:134-325):
    knltype 0 S, 1 scaled S, 2 scaled symmetric S (beta grid from 0).
    Returns dict(alpha, beta, sab[nbeta][nalpha], ntrimmed)."""
    a = np.ascontiguousarray(alpha, dtype=np.float64)
    b = np.ascontiguousarray(beta, dtype=np.float64)
    s = np.ascontiguousarray(sab, dtype=np.float64).reshape(b.size, a.size)
    na, nb, nt = ctypes.c_int(), ctypes.c_int(), ctypes.c_int64()
    args = (int(knltype), int(bool(beta_grid_optimised)), a.size, b.size, C.dptr(a), C.dptr(b), C.dptr(s))
    C.call("ncamd_sab_std_format", *args, ctypes.byref(na), ctypes.byref(nb), None, None, None, ctypes.byref(nt))
    oa, ob, os_ = np.empty(na.value), np.empty(nb.value), np.empty((nb.value, na.value))
    C.call("ncamd_sab_std_format", *args, ctypes.byref(na), ctypes.byref(nb), C.dptr(oa), C.dptr(ob), C.dptr(os_),
           ctypes.byref(nt))
    return dict(alpha=oa, beta=ob, sab=os_, ntrimmed=nt.value)


def create_sabscatter(alpha, beta, sab, temperature, mass_amu, sigma_bound, suggested_emax=0.0, egrid=None):
    a, b = C.as_f64(alpha), C.as_f64(beta)
    s = C.as_f64(sab).reshape(b.size, a.size)
    eg = C.as_f64(egrid if egrid is not None else np.zeros(0))
    h = C.call("ncamd_create_sabscatter", a.size, b.size, C.dptr(a), C.dptr(b), C.dptr(s), float(temperature),
               float(mass_amu), float(sigma_bound), float(suggested_emax), eg.size,
               C.dptr(eg) if eg.size else None)
    return SABScatter(h, name="SABScatter")


def _vdos_args(vdos):
    """(n_egrid, egrid, n_density, density) of a VDOS given as (egrid,
    density) or as a Debye temperature (createVDOSDebye)"""
    if np.isscalar(vdos):
        eg = C.as_f64([float(vdos)])
        return 1, eg, 0, None, eg
    egrid, density = vdos
    eg, d = C.as_f64(egrid), C.as_f64(density)
    return eg.size, eg, d.size, d, (eg, d)


def vdos_regularise(vdos):
    """regulariseVDOSGrid: (emin, emax, density) of the regular grid"""
    ne, eg, nd, d, keep = _vdos_args(vdos)
    dp = C.dptr(d) if d is not None else None
    lo, hi = ctypes.c_double(), ctypes.c_double()
    n = C.call("ncamd_vdos_regularise", ne, C.dptr(eg), nd, dp, ctypes.byref(lo), ctypes.byref(hi), None, 0)
    out = np.empty(n)
    C.call("ncamd_vdos_regularise", ne, C.dptr(eg), nd, dp, ctypes.byref(lo), ctypes.byref(hi), C.dptr(out), n)
    return lo.value, hi.value, out


def vdos_analyse(vdos, temperature, mass_amu):
    """VDOSEval values: dict(gamma0, msd [Aa^2], teff [K], integral)"""
    ne, eg, nd, d, keep = _vdos_args(vdos)
    out = np.empty(4)
    C.call("ncamd_vdos_analyse", ne, C.dptr(eg), nd, C.dptr(d) if d is not None else None, float(temperature),
           float(mass_amu), C.dptr(out))
    return dict(gamma0=out[0], msd=out[1], teff=out[2], integral=out[3])


MINIMC_TALLIES = {"theta": 0, "mu": 1, "nscat": 2, "w": 3, "nscat_uw": 4, "e": 5, "l": 6, "de": 7, "q": 8}
MINIMC_CATEGORIES = ("NOSCAT", "SINGLESCAT_ELAS", "SINGLESCAT_INELAS", "MULTISCAT_PUREELAS", "MULTISCAT_OTHER")


def minimc_sphere(scatter, absorption, numdens, radius, ekin, n, source="constant", pos=(0.0, 0.0, 0.0),
                  direction=(0.0, 0.0, 1.0), beam_radius=0.0, tally="theta", nbins=90, xrange=(0.0, 180.0),
                  breakdown=True, roulette=(0.1, 1e-2, 2), nscatlimit=None, ignoremiss=False,
                  include_absorption=True, seed=0):
    """MiniMC through a sphere on the GPU (ncamd_minimc_sphere).  Returns a
    dict: "total" and, with breakdown, each scattering-history category, as
    dict(content, errorsq, underflow, underflow_errorsq, overflow,
    overflow_errorsq) over nbins bins of xrange; plus "info"."""
    kinds = {"constant": 0, "circular": 1}
    src = np.array([ekin, *pos, *direction, beam_radius], dtype=np.float64)
    eng = np.array([roulette[0], roulette[1], roulette[2], -1 if nscatlimit is None else nscatlimit,
                    1 if ignoremiss else 0, 1 if include_absorption else 0], dtype=np.float64)
    hist = np.zeros((6, 2, nbins + 2))
    info = np.zeros(5)
    null = C.Handle()
    C.call("ncamd_minimc_sphere", scatter.handle if scatter is not None else null,
           absorption.handle if absorption is not None else null, float(numdens), float(radius), kinds[source],
           C.dptr(src), int(n), MINIMC_TALLIES[tally], int(nbins), float(xrange[0]), float(xrange[1]),
           1 if breakdown else 0, C.dptr(eng), int(seed), C.dptr(hist), C.dptr(info))

    def h(k):
        return dict(content=hist[k, 0, 1:-1].copy(), errorsq=hist[k, 1, 1:-1].copy(), underflow=hist[k, 0, 0],
                    underflow_errorsq=hist[k, 1, 0], overflow=hist[k, 0, -1], overflow_errorsq=hist[k, 1, -1])
    out = {"total": h(0), "info": dict(max_steps=int(info[0]), neutron_steps=int(info[1]), scattered=int(info[3]),
                                       step_kernel_ms=float(info[4]))}
    if breakdown:
        for i, name in enumerate(MINIMC_CATEGORIES):
            out[name] = h(1 + i)
    return out


def vdos_gn(vdos, temperature, mass_amu, order):
    """(egrid, spectrum) of the phonon-order-n function G_n of the expansion"""
    ne, eg, nd, d, keep = _vdos_args(vdos)
    dp = C.dptr(d) if d is not None else None
    lo, bw = ctypes.c_double(), ctypes.c_double()
    args = (ne, C.dptr(eg), nd, dp, float(temperature), float(mass_amu), int(order), ctypes.byref(lo),
            ctypes.byref(bw))
    n = C.call("ncamd_vdos_gn", *args, None, 0)
    spec = np.empty(n)
    C.call("ncamd_vdos_gn", *args, C.dptr(spec), n)
    return lo.value + bw.value * np.arange(n), spec


def vdos_kernel(vdos, temperature, mass_amu, vdoslux=3, target_emax=0.0):
    """createScatteringKernel on the GPU: dict(alpha, beta, sab[nb][na],
    suggested_emax, msd, gamma0, norders, orders_summed, rows_summed, gpu_ms)"""
    ne, eg, nd, d, keep = _vdos_args(vdos)
    dp = C.dptr(d) if d is not None else None
    args = (ne, C.dptr(eg), nd, dp, float(temperature), float(mass_amu), int(vdoslux), float(target_emax))
    na, nb, info = ctypes.c_int(), ctypes.c_int(), np.empty(7)
    C.call("ncamd_vdos_kernel", *args, ctypes.byref(na), ctypes.byref(nb), None, None, None, C.dptr(info))
    a, b, s = np.empty(na.value), np.empty(nb.value), np.empty((nb.value, na.value))
    C.call("ncamd_vdos_kernel", *args, ctypes.byref(na), ctypes.byref(nb), C.dptr(a), C.dptr(b), C.dptr(s), None)
    return dict(alpha=a, beta=b, sab=s, suggested_emax=info[0], msd=info[1], gamma0=info[2], norders=int(info[3]),
                orders_summed=int(info[4]), rows_summed=int(info[5]), gpu_ms=info[6])


def create_sabscatter_from_vdos(vdos, temperature, mass_amu, sigma_bound, vdoslux=3, target_emax=0.0):
    """SABScatter of an element with VDOS dynamics (the kernel expanded on the
    GPU, its suggested Emax as the top of the energy grid)"""
    ne, eg, nd, d, keep = _vdos_args(vdos)
    h = C.call("ncamd_create_sabscatter_from_vdos", ne, C.dptr(eg), nd, C.dptr(d) if d is not None else None,
               float(temperature), float(mass_amu), float(sigma_bound), int(vdoslux), float(target_emax))
    return SABScatter(h, name="SABScatter")


def create_composition(procs, scales):
    """ProcComposition of the processes (NCProcImpl.cc:255-338): nested
    compositions are flattened, repeated processes add their scales,
    PowderBragg and ElInc components of several phases are merged."""
    arr = (C.Handle * len(procs))(*[p.handle for p in procs])
    sc = C.as_f64(scales)
    h = C.call("ncamd_create_composition", len(procs), arr, C.dptr(sc))
    return Scatter(h, name="ProcComposition", keep=tuple(procs))


def composition_components(comp):
    """[(component, scale)] of a composition after merging
    (ProcComposition::components(), NCProcImpl.hh:328)."""
    n = C.call("ncamd_composition_component", comp.handle, -1, None, None)
    out = []
    for i in range(n):
        h = C.Handle()
        s = np.zeros(1)
        C.call("ncamd_composition_component", comp.handle, i, ctypes.byref(h), C.dptr(s))
        out.append((Scatter(h, name="component", keep=(comp,)), float(s[0])))
    return out


class Batch:
    """Many materials evaluated in one persistent launch (device pointers).
    Inputs and outputs are concatenated per material in the given order."""

    def __init__(self, procs, counts):
        arr = (C.Handle * len(procs))(*[p.handle for p in procs])
        cnt = np.ascontiguousarray(counts, dtype=np.int64)
        h = C.call("ncamd_batch_create", len(procs), arr, C.i64ptr(cnt))
        if not h.internal:
            C.check()
            raise C.NCAMDError("batch creation failed")
        self._h = h
        self._procs = tuple(procs)
        self.counts = cnt.copy()
        self.offsets = np.concatenate([[0], np.cumsum(cnt)])

    def __del__(self):
        try:
            if self._h is not None and self._h.internal:
                C.lib().ncamd_batch_destroy(ctypes.byref(self._h))
        except Exception:
            pass

    @property
    def total(self):
        return C.call("ncamd_batch_total", self._h)

    def dev_crossSection(self, d_ekin, d_xs, stream=0):
        C.call("ncamd_dev_batch_crosssection", self._h, d_ekin, d_xs, stream)

    def dev_sampleScatterIsotropic(self, d_ekin, d_oe, d_omu, stream=0):
        C.call("ncamd_dev_batch_samplescatterisotropic", self._h, d_ekin, d_oe, d_omu, stream)

    def dev_check_errors(self, stream=0):
        return C.call("ncamd_dev_batch_check_errors", self._h, stream)

    def tuning(self):
        t, b, o = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        C.call("ncamd_batch_get_tuning", self._h, ctypes.byref(t), ctypes.byref(b), ctypes.byref(o))
        return dict(tile=t.value, nblocks=b.value, ordering=o.value)

    def graph_stats(self, mode):
        """(captures, replays) of the graph of mode 0 (xs) or 1 (sampling)"""
        c, r = ctypes.c_int64(), ctypes.c_int64()
        C.call("ncamd_batch_graph_stats", self._h, int(mode), ctypes.byref(c), ctypes.byref(r))
        return c.value, r.value

    def set_tuning(self, tile, nblocks, ordering):
        C.call("ncamd_batch_set_tuning", self._h, int(tile), int(nblocks), int(ordering))

    def dev_autotune(self, d_ekin, d_o1, d_o2, stream=0):
        return C.call("ncamd_dev_batch_autotune", self._h, d_ekin, d_o1, d_o2, stream)


class MLPProposal:
    """The learned proposal model (8 energy features -> 3 x 64 -> 8-component
    bivariate Gaussian mixture over (w, kappa)), trained on the GPU."""

    def __init__(self, emin, emax, seed=1):
        h = C.call("ncamd_mlp_create", int(seed), float(emin), float(emax))
        if not h.internal:
            C.check()
            raise C.NCAMDError("model creation failed")
        self._h = h
        self.emin, self.emax = float(emin), float(emax)

    def __del__(self):
        try:
            if self._h is not None and self._h.internal:
                C.lib().ncamd_mlp_destroy(ctypes.byref(self._h))
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def nparams(self):
        return C.call("ncamd_mlp_nparams")

    def params(self):
        out = np.empty(self.nparams(), dtype=np.float32)
        C.call("ncamd_mlp_get_params", self._h, C.fptr(out))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        C.call("ncamd_mlp_set_params", self._h, C.fptr(p))

    def forward(self, ekin):
        e = C.as_f64(np.atleast_1d(ekin))
        out = np.empty((e.size, 48), dtype=np.float32)
        C.call("ncamd_mlp_forward", self._h, C.dptr(e), e.size, C.fptr(out))
        return out

    def loss_grad(self, ekin, w, kappa):
        e, ww, kk = C.as_f64(ekin), C.as_f64(w), C.as_f64(kappa)
        g = np.empty(self.nparams(), dtype=np.float32)
        loss = C.call("ncamd_mlp_loss_grad", self._h, C.dptr(e), C.dptr(ww), C.dptr(kk), e.size, C.fptr(g))
        return loss, g

    def train_step(self, ekin, w, kappa, lr):
        e, ww, kk = C.as_f64(ekin), C.as_f64(w), C.as_f64(kappa)
        return C.call("ncamd_mlp_train_step", self._h, C.dptr(e), C.dptr(ww), C.dptr(kk), e.size, float(lr))

    def train_sab(self, sab, steps=2000, batch=8192, lr=3e-3, seed=1):
        return C.call("ncamd_mlp_train_sab", self._h, sab.handle, int(steps), int(batch), float(lr), int(seed))


def debug_gemm_bf16(A, B, C0=None, ta=False, tb=False, accum=False):
    A = np.ascontiguousarray(A, dtype=np.float32)
    B = np.ascontiguousarray(B, dtype=np.float32)
    M = A.shape[1] if ta else A.shape[0]
    K = A.shape[0] if ta else A.shape[1]
    N = B.shape[0] if tb else B.shape[1]
    out = np.zeros((M, N), dtype=np.float32) if C0 is None else np.ascontiguousarray(C0, dtype=np.float32).copy()
    C.call("ncamd_debug_gemm_bf16", C.fptr(A), C.fptr(B), C.fptr(out), M, N, K, int(ta), int(tb), int(accum))
    return out
