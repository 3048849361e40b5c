"""Benchmark: the per-neutron physics hot path on one MI355X.

Default run (`python bench.py`): the headline workload and, in the same JSON
line under "subresults", every other configuration of the workload set, each
with its own roofline figure and CPU baseline:

  headline (C3): S(alpha,beta) scatter sampling.  Table 200 (alpha) x 800
    (beta) FP64 from a three-Gaussian VDOS (centres 0.01-0.12 eV, widths
    0.003-0.02 eV, seeded) on 500 points up to 0.15 eV, T = 293 K; batch of
    1e8 neutrons with E log-uniform in [1e-3, 1] eV resident in HBM; outputs
    E' and mu (FP64) in HBM.  Three samplers are timed: the plain (Alg1)
    sampler, the box-envelope sampler on a uniform grid and the learned
    sampler (the same envelope sampler on grids placed by the MLP trained on
    the plain sampler's samples; its training time is reported); `value` is
    the learned sampler.  KS distances of E' and mu against the CPU checker's
    plain sampler at fixed energies are reported.
  C1: fcc powder (a = 4.05 Aa, b_coh 3.449 fm, T_Debye 410 K, T 293 K,
    dcutoff 0.5 Aa): ElInc + PowderBragg + S(alpha,beta) from the Debye VDOS;
    1e6 neutrons, E log-uniform 1e-5..10 eV, isotropic directions; one step =
    cross sections + oriented scatter sampling.
  C2: PowderBragg of the seeded orthorhombic 24-atom cell at dcutoff 0.1 Aa
    (~1e5 planes), 1e8 neutrons, E log-uniform 1e-4..1 eV.
  C4: 512 materials (ElInc + PowderBragg + S(alpha,beta) each) of uneven
    sizes in one persistent launch, ~1.1e8 neutrons per step.
  C5: mosaic single crystal, the seeded orthorhombic cell, dcutoff 0.2 Aa,
    mosaicity FWHM 30 arcmin, seeded random orientation; 1e7 neutrons with
    wavelength ~ U(1,6) Aa and directions in a 2-degree cone about a seeded
    axis; one step = cross sections + scatter sampling.
  F2: VDOS -> S(alpha,beta) expansion of the data library's aluminium VDOS at
    vdoslux 5; one step = one whole expansion.
  F3: MiniMC: 1e7 neutrons of a 4 Aa pencil beam into a sphere of Al_sg225
    one scattering length across (forced collisions, absorption, roulette,
    theta tally); one step = one whole engine call.
  F4: LCBragg, the layered-crystal model (lcmode=0) of pyrolytic graphite:
    cross sections and scattering of 2^20 neutrons (0.5-7 Aa, isotropic).

`--workload` runs one of them alone.  Roofline figures come from HIP events
on the launch stream; the CPU baseline is the CPU checker timed on a bounded
sample of the same workload on this host's cores.
"""
import argparse
import json
import math
import os
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# HBM traffic of one launch of the learned sampler on this workload (256
# intervals x 128 x 32 boxes): 20.86 GB fetched + 2.21 GB written, from the
# FETCH_SIZE / WRITE_SIZE counter passes scaled by a plain 8e8-byte copy
# (profiles/k_kb_sample_sab200x800_learned_grid.md); counters cannot be read
# inside this process, so the figure is carried here
LEARNED_TRAFFIC_GB = 23.07
BYTES_PER_SAMPLE = 24  # E in (8) + E' out (8) + mu out (8)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=100_000_000)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-batch", type=int, default=2_000_000)
    ap.add_argument("--copy-calib", action="store_true",
                    help="also run one device-to-device copy of the output size (counter calibration)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=["all", "sab", "c1", "scbragg", "powderbragg", "batch", "vdos", "minimc", "lcbragg"],
                    default="all")
    ap.add_argument("--sub-cpu-seconds", type=float, default=8.0,
                    help="CPU-baseline sample length of each sub-result")
    ap.add_argument("--train-steps", type=int, default=2000)
    ap.add_argument("--ks-samples", type=int, default=10_000_000,
                    help="samples per energy for the KS check against the CPU reference sampler (0: skip)")
    return ap.parse_args()


# the host's share of CPU threads on a one-GPU node: the CPU baseline uses all
# threads of this process's affinity mask, at most this many
CPU_SHARE = 16


def progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_threads(args):
    if args.cpu_threads:
        return args.cpu_threads
    return max(1, min(len(os.sched_getaffinity(0)), CPU_SHARE))


def host_info():
    model = None
    try:
        for line in pathlib.Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity_threads": len(os.sched_getaffinity(0)),
            "baseline_threads_cap": CPU_SHARE}


def library_info():
    """which native library ran: path and content hash (the pushed build is
    run as is; nothing is rebuilt on the GPU host)"""
    import hashlib
    lib = ROOT / "ncrystal_amd" / "libncrystal_amd.so"
    h = hashlib.sha256(lib.read_bytes()).hexdigest()[:16] if lib.exists() else None
    return {"path": str(lib.relative_to(ROOT)), "sha256_16": h}


def cpu_baseline(table, batch, seconds, threads):
    """CPU checker (oracle/) timed on a bounded sample of the same workload:
    repeated batches of the same energy distribution for about `seconds`."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import numpy as np
    import pyoracle
    from ncrystal_amd import synth
    pyoracle.build()
    o = pyoracle.sab(table["alpha"], table["beta"], table["sab"], table["temperature"], table["mass"],
                     table["sigma_bound"])
    e = synth.energies_logu(batch, 1e-3, 1.0, seed=2)
    o.sample_iso(e[:10000], seed=5, nthreads=threads)  # warm caches
    done = 0
    t0 = time.perf_counter()
    k = 0
    while True:
        _, _, nerr = o.sample_iso(e, seed=5, stream=k, nthreads=threads)
        done += batch
        k += 1
        dt = time.perf_counter() - t0
        if dt >= seconds:
            break
    return dict(value=done / dt, unit="samples/s", cores=threads, kind="port",
                sample=f"{done} neutrons ({k} batches of {batch}) of the same workload "
                       f"(E log-uniform 1e-3..1 eV, same table), {threads} threads, {dt:.1f} s")


FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6
# FP32 pre-filter of the scan per (neutron, demi-normal): dot product (3 mul + 2 add),
# sin(theta) = w * 1/(2d) (1 mul), | |dot| - sin(theta) | (1 sub), compare (1)
SCB_FLOPS_PER_NORMAL = 8


def scb_setup(n):
    import numpy as np
    from ncrystal_amd import core, synth
    cr = synth.ortho_random(seed=1, dcutoff=0.2)
    sp = [(s.csl, core.debye_msd(s.debye_temp, cr.temperature, s.mass), s.positions) for s in cr.species]
    hkl, mult, dsp, fsq, moff, mh = core.hkl_orthogonal_members(cr.a, cr.b, cr.c, sp, cr.dcutoff)
    R = synth.random_rotation(4)
    args = (cr.volume * cr.n_atoms, dsp, fsq, moff, mh, core.reci_lattice_orthogonal(cr.a, cr.b, cr.c), R,
            30 * math.pi / (180 * 60))
    wl = synth.wavelengths_uniform(n, 1.0, 6.0, seed=6)
    e = synth.WL2EKIN / (wl * wl)
    dirs = synth.cone_dirs(n, synth.random_unit_vector(5), 2 * math.pi / 180, seed=3)
    return cr, args, e, np.ascontiguousarray(dirs)


def cpu_baseline_scb(args_scb, e, dirs, seconds, threads):
    """CPU checker timed on a bounded sample of the single-crystal workload:
    cross section + sampling of the first neutrons of the batch, repeated
    for about `seconds`."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle
    pyoracle.build()
    v0n, dsp, fsq, moff, mh, L, R, mos = args_scb
    o = pyoracle.scbragg(0, 0, 0, v0n, dsp, fsq, moff, mh, R, mos, reci=L)
    batch = 20000
    eb, db = e[:batch], dirs[:batch]
    done, k = 0, 0
    t0 = time.perf_counter()
    while True:
        o.xs_dir(eb, db, nthreads=threads)
        o.sample_dir(eb, db, seed=5, stream=k, nthreads=threads)
        done += batch
        k += 1
        dt = time.perf_counter() - t0
        if dt >= seconds:
            break
    return dict(value=done / dt, unit="neutrons/s (xs + sample)", cores=threads, kind="port",
                sample=f"{done} neutrons ({k} passes over the first {batch} neutrons of the batch, "
                       f"cross section + sampling), {threads} threads, {dt:.1f} s")


def c1_setup():
    """the fcc powder: ElInc, PowderBragg and the S(alpha,beta) expanded on the
    GPU from the Debye VDOS of T_Debye 410 K (createVDOSDebye; NCrystal's
    default inelastic model for a crystal with a Debye temperature), vdoslux 3"""
    import math
    from ncrystal_amd import core, materials, synth
    cr = synth.fcc_powder()
    sp = cr.species[0]
    sigma = 4 * math.pi * sp.csl ** 2 + sp.sigma_inc
    dyn = {"vdos": sp.debye_temp, "mass": sp.mass, "sigma_bound": sigma, "vdoslux": 3}
    m = materials.Material(cr, dynamics=[(1.0, dyn)])
    k = core.vdos_kernel(sp.debye_temp, cr.temperature, sp.mass, vdoslux=3)
    t = core.sab_std_format(0, k["alpha"], k["beta"], k["sab"], beta_grid_optimised=True)
    sab = dict(alpha=t["alpha"], beta=t["beta"], sab=t["sab"], temperature=cr.temperature, mass=sp.mass,
               sigma_bound=sigma, suggested_emax=k["suggested_emax"])
    m.components["inelas"] = m.inelastic[0]
    return cr, sab, m


def cpu_baseline_c1(m, sab, e, dirs, seconds, threads):
    """CPU checker timed on the C1 composition: cross sections + oriented
    sampling of the whole 1e6-neutron batch, repeated for about `seconds`."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle
    from ncrystal_amd import materials
    pyoracle.build()
    mm, bb, ss = materials.elinc_inputs(m.crystal, m.msd)
    v0n, dsp, fm = m.pb_inputs
    o = pyoracle.composition([pyoracle.elinc(mm, bb, ss), pyoracle.powderbragg(v0n, dsp, fm),
                              pyoracle.sab(sab["alpha"], sab["beta"], sab["sab"], sab["temperature"], sab["mass"],
                                           sab["sigma_bound"], suggested_emax=sab.get("suggested_emax", 0.0))],
                             [1.0, 1.0, 1.0])
    done, k = 0, 0
    t0 = time.perf_counter()
    while True:
        o.xs(e, nthreads=threads)
        o.sample_dir(e, dirs, seed=5, stream=k, nthreads=threads)
        done += e.size
        k += 1
        dt = time.perf_counter() - t0
        if dt >= seconds:
            break
    return dict(value=done / dt, unit="neutrons/s (xs + oriented sample)", cores=threads, kind="port",
                sample=f"{done} neutrons ({k} passes over the 1e6-neutron batch), cross section + oriented "
                       f"sampling, {threads} threads, {dt:.1f} s")


def run_c1(args):
    import numpy as np
    from ncrystal_amd import _capi as C
    from ncrystal_amd import synth
    n = 1_000_000
    t0 = time.perf_counter()
    cr, sab, m = c1_setup()
    t_setup = time.perf_counter() - t0
    g = m.scatter
    e = synth.energies_logu(n, 1e-5, 10.0, seed=1)
    dirs = np.ascontiguousarray(synth.isotropic_dirs(n, seed=1))
    stream = C.Stream()
    de, dd = C.DeviceArray.from_numpy(e), C.DeviceArray.from_numpy(dirs.ravel())
    xs, oe, od = C.DeviceArray(n), C.DeviceArray(n), C.DeviceArray(3 * n)
    g.setRandomSeed(1)
    ev = [C.Event() for _ in range(3)]
    for _ in range(args.warmup):
        g.dev_crossSectionIsotropic(de.ptr, n, xs.ptr, stream=stream.ptr)
        g.dev_sampleScatter(de.ptr, dd.ptr, n, oe.ptr, od.ptr, stream=stream.ptr)
    stream.sync()
    xs_ms = samp_ms = 0.0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ev[0].record(stream)
        g.dev_crossSectionIsotropic(de.ptr, n, xs.ptr, stream=stream.ptr)
        ev[1].record(stream)
        g.dev_sampleScatter(de.ptr, dd.ptr, n, oe.ptr, od.ptr, stream=stream.ptr)
        ev[2].record(stream)
        stream.sync()
        xs_ms += ev[0].elapsed_ms(ev[1])
        samp_ms += ev[1].elapsed_ms(ev[2])
    wall = time.perf_counter() - t0
    xs_ms /= args.steps
    samp_ms /= args.steps
    assert g.dev_check_errors(stream=stream.ptr) == 0
    o = od.numpy().reshape(n, 3)
    assert np.all(xs.numpy() >= 0.0) and np.abs(np.linalg.norm(o, axis=1) - 1).max() < 1e-12, "invalid output"
    # E + dir in (32 B), E' + dir out (32 B); xs: E in, xs out (16 B)
    ach = 64 * n / (samp_ms * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_c1(m, sab, e, dirs, args.cpu_seconds, cpu_threads(args))
    return {
        "metric": "fcc powder neutrons/s on 1 MI355X (cross section + oriented scatter sampling)",
        "value": n * args.steps / wall, "unit": "neutrons/s",
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "dtype": "f64",
        "config": {"workload": "fcc powder a=4.05 Aa, b_coh 3.449 fm, sigma_inc 0.0082 b, T_Debye 410 K, T 293 K, "
                               "dcutoff 0.5 Aa: ElInc + PowderBragg + S(alpha,beta) expanded on the GPU from the "
                               "Debye VDOS (vdoslux 3: %dx%d, Emax %.3f eV, free-gas extender above); E "
                               "log-uniform 1e-5..10 eV, isotropic directions, seed 1"
                               % (sab["alpha"].size, sab["beta"].size, m.components["inelas"].egrid()[0][-1]),
                   "neutrons_per_step": n, "families": int(len(m.pb_inputs[1])), "setup_s": t_setup},
        "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "oriented sampling call (k_comp_pick + k_comp_sab + high-energy queue)",
                     "kernel_ms": samp_ms, "bytes_per_neutron": 64, "xs_call_ms": xs_ms,
                     "xs_GBs": 16 * n / (xs_ms * 1e-3) / 1e9},
        "cpu_baseline": cpu,
    }


def run_scbragg(args):
    import numpy as np
    from ncrystal_amd import _capi as C
    from ncrystal_amd import core
    n = args.n if args.n != 100_000_000 else 10_000_000
    cr, a, e, dirs = scb_setup(n)
    sc = core.create_scbragg(*a)
    t = sc.tables()
    # demi-normals with 2d > wl scanned per (E, dir) evaluation
    wl = np.sqrt(0.081804209605330899 / e)
    nf = np.searchsorted(t["inv2d"], (1.0 - 2 * 2.220446049250313e-16) / wl, side="left")
    scanned = float(t["off"][nf].sum())
    stream = C.Stream()
    de, dd = C.DeviceArray.from_numpy(e), C.DeviceArray.from_numpy(dirs.ravel())
    xs, oe, od = C.DeviceArray(n), C.DeviceArray(n), C.DeviceArray(3 * n)
    sc.setRandomSeed(77)
    ev = [C.Event() for _ in range(3)]

    def step():
        sc.dev_crossSection(de.ptr, dd.ptr, n, xs.ptr, stream=stream.ptr)
        sc.dev_sampleScatter(de.ptr, dd.ptr, n, oe.ptr, od.ptr, stream=stream.ptr)

    for _ in range(args.warmup):
        step()
    stream.sync()
    xs_ms = samp_ms = 0.0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ev[0].record(stream)
        sc.dev_crossSection(de.ptr, dd.ptr, n, xs.ptr, stream=stream.ptr)
        ev[1].record(stream)
        sc.dev_sampleScatter(de.ptr, dd.ptr, n, oe.ptr, od.ptr, stream=stream.ptr)
        ev[2].record(stream)
        stream.sync()
        xs_ms += ev[0].elapsed_ms(ev[1])
        samp_ms += ev[1].elapsed_ms(ev[2])
    t1 = time.perf_counter()
    wall = t1 - t0
    xs_ms /= args.steps
    samp_ms /= args.steps
    x = xs.numpy()
    o = od.numpy().reshape(n, 3)
    assert np.all(x >= 0.0) and np.abs(np.linalg.norm(o[:100000], axis=1) - 1).max() < 1e-12, "invalid output"
    # dominant work: the scans of the sampling call, counted as one full scan
    # of the reference's algorithm per neutron (the second, partial scan of
    # the selection is left out) over the whole call's time (sorting, the
    # count, pick and scatter kernels and the Romberg queue included), so the
    # figure is a lower bound
    flops = SCB_FLOPS_PER_NORMAL * scanned
    achieved = flops / (samp_ms * 1e-3) / 1e12
    cpu = None
    if not args.no_cpu:
        threads = cpu_threads(args)
        cpu = cpu_baseline_scb(a, e, dirs, args.cpu_seconds, threads)
    out = {
        "metric": "SCBragg oriented neutrons/s on 1 MI355X (cross section + scatter sampling)",
        "value": n * args.steps / wall,
        "unit": "neutrons/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "SCBragg, orthorhombic cell seed 1 (a,b,c=%.3f,%.3f,%.3f Aa, 24 atoms), dcutoff 0.2 Aa, "
                               "mosaicity 30 arcmin, random orientation, wavelength U(1,6) Aa, 2-degree cone"
                               % (cr.a, cr.b, cr.c),
                   "neutrons_per_step": n, "families": int(len(t["inv2d"])),
                   "demi_normals": int(t["off"][-1]), "parallelism": "single GPU"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS, "frac_vs_fp64_peak": achieved / FP64_PEAK_TFLOPS,
                     "traffic": None, "kernel": "sampling call: k_scb_sample_count + k_scb_sample_pick + k_scb_sample_scatter "
                                                               "+ sort + Romberg queue",
                     "kernel_ms": samp_ms, "xs_call_ms": xs_ms,
                     "xs_achieved_TFLOPs": flops / (xs_ms * 1e-3) / 1e12,
                     "flops_per_launch": flops, "normals_scanned_per_neutron": scanned / n},
        "cpu_baseline": cpu,
    }
    return out


def cpu_baseline_pb(v0n, dsp, fm, seconds, threads):
    """CPU checker timed on a bounded sample of the PowderBragg workload:
    cross section + sampling of 2e6 neutrons, repeated for about `seconds`."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle
    from ncrystal_amd import synth
    pyoracle.build()
    o = pyoracle.powderbragg(v0n, dsp, fm)
    batch = 2_000_000
    e = synth.energies_logu(batch, 1e-4, 1.0, seed=2)
    done, k = 0, 0
    t0 = time.perf_counter()
    while True:
        o.xs(e, nthreads=threads)
        o.sample_iso(e, seed=5, stream=k, nthreads=threads)
        done += batch
        k += 1
        dt = time.perf_counter() - t0
        if dt >= seconds:
            break
    return dict(value=done / dt, unit="neutrons/s (xs + sample)", cores=threads, kind="port",
                sample=f"{done} neutrons ({k} batches of {batch}, E log-uniform 1e-4..1 eV), "
                       f"cross section + sampling, {threads} threads, {dt:.1f} s")


def run_powderbragg(args):
    import numpy as np
    from ncrystal_amd import _capi as C
    from ncrystal_amd import core, synth
    n = args.n
    cr = synth.ortho_random(seed=1, dcutoff=0.1)
    sp = [(s.csl, core.debye_msd(s.debye_temp, cr.temperature, s.mass), s.positions) for s in cr.species]
    hkl, mult, dsp, fsq = core.hkl_orthogonal(cr.a, cr.b, cr.c, sp, cr.dcutoff)
    v0n = cr.volume * cr.n_atoms
    pb = core.create_powderbragg(v0n, dsp, fsq * mult)
    nplanes = len(core.powderbragg_tables(pb)[0])
    stream = C.Stream()
    ekin, xs, oe, om = C.DeviceArray(n), C.DeviceArray(n), C.DeviceArray(n), C.DeviceArray(n)
    C.call("ncamd_dev_fill_loguniform", ekin.ptr, n, 1e-4, 1.0, 1, stream.ptr)
    stream.sync()
    pb.setRandomSeed(9)
    ev = [C.Event() for _ in range(4)]
    for _ in range(args.warmup):
        pb.dev_crossSectionIsotropic(ekin.ptr, n, xs.ptr, stream=stream.ptr)
        pb.dev_sampleScatterIsotropic(ekin.ptr, n, oe.ptr, om.ptr, stream=stream.ptr)
    stream.sync()
    t64 = t32 = ts = 0.0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ev[0].record(stream)
        pb.dev_crossSectionIsotropic(ekin.ptr, n, xs.ptr, stream=stream.ptr)
        ev[1].record(stream)
        pb.dev_crossSectionIsotropic(ekin.ptr, n, xs.ptr, stream=stream.ptr, fp32=True)
        ev[2].record(stream)
        pb.dev_sampleScatterIsotropic(ekin.ptr, n, oe.ptr, om.ptr, stream=stream.ptr)
        ev[3].record(stream)
        stream.sync()
        t64 += ev[0].elapsed_ms(ev[1])
        t32 += ev[1].elapsed_ms(ev[2])
        ts += ev[2].elapsed_ms(ev[3])
    wall = time.perf_counter() - t0
    t64, t32, ts = t64 / args.steps, t32 / args.steps, ts / args.steps
    assert pb.dev_check_errors(stream=stream.ptr) == 0
    mu = om.numpy(1000000)
    assert np.all(np.abs(mu) <= 1.0), "invalid output"
    ach = 24 * n / (ts * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu:
        threads = cpu_threads(args)
        cpu = cpu_baseline_pb(v0n, dsp, fsq * mult, args.cpu_seconds, threads)
    out = {
        "metric": "PowderBragg neutrons/s on 1 MI355X (FP64 xs + FP32-table xs + scatter sampling)",
        "value": n * args.steps / wall, "unit": "neutrons/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "PowderBragg, orthorhombic cell seed 1 (a,b,c=%.3f,%.3f,%.3f Aa, 24 atoms), "
                               "dcutoff 0.1 Aa, E log-uniform 1e-4..1 eV" % (cr.a, cr.b, cr.c),
                   "neutrons_per_step": n, "families": int(len(dsp)), "planes": nplanes,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBS, "traffic": None, "kernel": "k_sample_iso<kCompPBOnly>",
                     "kernel_ms": ts, "xs_fp64_ms": t64, "xs_fp32_ms": t32,
                     "xs_fp64_GBs": 16 * n / (t64 * 1e-3) / 1e9, "xs_fp32_GBs": 16 * n / (t32 * 1e-3) / 1e9},
        "cpu_baseline": cpu,
    }
    return out


def batch_setup(core, mats):
    procs = []
    for i, m in enumerate(mats):
        v0n, d, fm = m["pb"]
        ei = core.create_elincscatter(m["elinc"]["msd"], m["elinc"]["bixs"], m["elinc"]["scale"])
        pb = core.create_powderbragg(v0n, d, fm)
        t = m["sab"]
        sab = core.create_sabscatter(t["alpha"], t["beta"], t["sab"], t["temperature"], t["mass"], t["sigma_bound"])
        p = core.create_composition([ei, pb, sab], [1.0, 1.0, 1.0])
        p.setRandomSeed(1000 + i)
        procs.append(p)
    return procs


def cpu_baseline_batch(mats, seconds, threads, nmat=8):
    """CPU checker timed on a bounded sample of the many-materials workload:
    cross section + sampling for the neutrons of the first `nmat` materials,
    repeated for about `seconds`."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle
    from ncrystal_amd import synth
    pyoracle.build()
    objs = []
    for m in mats[:nmat]:
        v0n, d, fm = m["pb"]
        t = m["sab"]
        parts = [pyoracle.elinc(m["elinc"]["msd"], m["elinc"]["bixs"], m["elinc"]["scale"]),
                 pyoracle.powderbragg(v0n, d, fm),
                 pyoracle.sab(t["alpha"], t["beta"], t["sab"], t["temperature"], t["mass"], t["sigma_bound"])]
        objs.append((pyoracle.composition(parts, [1.0, 1.0, 1.0]), parts,
                     synth.energies_logu(m["count"], 1e-4, 1.0, seed=3)))
    done, k = 0, 0
    t0 = time.perf_counter()
    while True:
        for o, _, e in objs:
            o.xs(e, nthreads=threads)
            o.sample_iso(e, seed=5, stream=k, nthreads=threads)
            done += e.size
        k += 1
        dt = time.perf_counter() - t0
        if dt >= seconds:
            break
    return dict(value=done / dt, unit="neutrons/s (xs + sample)", cores=threads, kind="port",
                sample=f"{done} neutrons: {k} passes over the first {nmat} materials "
                       f"(ElInc + PowderBragg + SAB each, {sum(x[2].size for x in objs)} neutrons per pass), "
                       f"cross section + sampling, {threads} threads, {dt:.1f} s")


def run_batch(args):
    import numpy as np
    from ncrystal_amd import _capi as C
    from ncrystal_amd import core, synth
    t0 = time.perf_counter()
    mats = synth.batch_materials(nmat=512, seed=11)
    procs = batch_setup(core, mats)
    counts = [m["count"] for m in mats]
    batch = core.Batch(procs, counts)
    t_setup = time.perf_counter() - t0
    tot = batch.total
    stream = C.Stream()
    ekin, xs, oe, om = C.DeviceArray(tot), C.DeviceArray(tot), C.DeviceArray(tot), C.DeviceArray(tot)
    C.call("ncamd_dev_fill_loguniform", ekin.ptr, tot, 1e-4, 1.0, 1, stream.ptr)
    stream.sync()
    t0 = time.perf_counter()
    batch.dev_autotune(ekin.ptr, oe.ptr, om.ptr, stream=stream.ptr)
    stream.sync()
    t_tune = time.perf_counter() - t0
    tuning = batch.tuning()

    def step_batch():
        batch.dev_crossSection(ekin.ptr, xs.ptr, stream=stream.ptr)
        batch.dev_sampleScatterIsotropic(ekin.ptr, oe.ptr, om.ptr, stream=stream.ptr)

    off = batch.offsets

    def step_loop():
        for i, p in enumerate(procs):
            a, n = 8 * int(off[i]), int(counts[i])
            p.dev_crossSectionIsotropic(ekin.ptr + a, n, xs.ptr + a, stream=stream.ptr)
            p.dev_sampleScatterIsotropic(ekin.ptr + a, n, oe.ptr + a, om.ptr + a, stream=stream.ptr)

    def timed(step, steps):
        for _ in range(args.warmup):
            step()
        stream.sync()
        ev0, ev1 = C.Event(), C.Event()
        t0 = time.perf_counter()
        ev0.record(stream)
        for _ in range(steps):
            step()
        ev1.record(stream)
        stream.sync()
        return time.perf_counter() - t0, ev0.elapsed_ms(ev1) / steps

    wall, ms = timed(step_batch, args.steps)
    nerr = batch.dev_check_errors(stream=stream.ptr)
    x = xs.numpy(2000000)
    mu = om.numpy(2000000)
    assert nerr == 0 and np.all(x >= 0.0) and np.all(np.abs(mu) <= 1.0), "invalid output"
    loop_wall, loop_ms = timed(step_loop, max(1, args.steps // 2))
    value = tot * args.steps / wall
    cpu = None
    if not args.no_cpu:
        threads = cpu_threads(args)
        cpu = cpu_baseline_batch(mats, args.cpu_seconds, threads)
    # the whole step (cross-section launch + the three sampling stages, timed together):
    # E in, xs out (16 B) + E in, E' and mu out (24 B) per neutron
    achieved = 40.0 * tot / (ms * 1e-3) / 1e9
    nplanes = [len(m["pb"][1]) for m in mats]
    grids = [m["sab"]["sab"].size for m in mats]
    out = {
        "metric": "neutrons/s (cross section + isotropic scatter) over 512 uneven materials in one persistent launch",
        "value": value,
        "unit": "neutrons/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "512 materials (ElInc + PowderBragg + SAB each), N_planes log-U 1e2..1e5, "
                               "SAB grids 50x100..300x1200, neutrons per material log-U 1e4..1e6, "
                               "E log-U 1e-4..1 eV; one persistent launch per call replayed from a HIP graph",
                   "materials": len(procs), "neutrons_per_step": int(tot),
                   "planes_min_max": [int(min(nplanes)), int(max(nplanes))],
                   "sab_cells_min_max": [int(min(grids)), int(max(grids))],
                   "tuning": tuning, "autotune_s": t_tune, "setup_s": t_setup,
                   "per_material_calls_neutrons_per_s": tot * max(1, args.steps // 2) / loop_wall,
                   "batched_over_per_material_calls": (tot * args.steps / wall) /
                                                      (tot * max(1, args.steps // 2) / loop_wall),
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "k_batch_xs + k_batch_pick + k_batch_sab + k_batch_highe",
                     "kernel_ms": ms, "bytes_per_neutron": 40},
        "cpu_baseline": cpu,
    }
    return out


def ks_sorted(a_sorted, b):
    """two-sample Kolmogorov-Smirnov statistic; `a_sorted` already sorted"""
    import numpy as np
    b = np.sort(b)
    allv = np.concatenate([a_sorted, b])
    ca = np.searchsorted(a_sorted, allv, side="right") / a_sorted.size
    cb = np.searchsorted(b, allv, side="right") / b.size
    return float(np.abs(ca - cb).max())


def ks_check(procs, table, energies, nks, threads):
    """two-sample KS of E' and of mu at fixed energies: each GPU sampler
    against the plain sampler of the CPU checker (the reference algorithm)"""
    sys.path.insert(0, str(ROOT / "oracle"))
    import numpy as np
    import pyoracle
    pyoracle.build()
    o = pyoracle.sab(table["alpha"], table["beta"], table["sab"], table["temperature"], table["mass"],
                     table["sigma_bound"])
    out = {k: {} for k in procs}
    for E in energies:
        e = np.full(nks, E)
        re_, rm, _ = o.sample_iso(e, seed=77, nthreads=threads)
        re_, rm = np.sort(re_), np.sort(rm)
        for k, p in procs.items():
            p.setRandomSeed(501)
            ge, gm = p.sampleScatterIsotropic(e)
            out[k][str(E)] = max(ks_sorted(re_, ge), ks_sorted(rm, gm))
    return out


def run_sab(args):
    import numpy as np
    from ncrystal_amd import _capi as C
    from ncrystal_amd import core, synth

    table = synth.sab_table_c3(seed=1)
    targs = (table["alpha"], table["beta"], table["sab"], table["temperature"], table["mass"], table["sigma_bound"])
    n = args.n
    stream = C.Stream()
    ekin = C.DeviceArray(n)
    oe = C.DeviceArray(n)
    om = C.DeviceArray(n)
    # inputs resident in HBM before timing: E log-uniform in [1e-3, 1] eV
    C.call("ncamd_dev_fill_loguniform", ekin.ptr, n, 1e-3, 1.0, 1, stream.ptr)
    stream.sync()

    def measure(proc):
        proc.setRandomSeed(1234)
        for _ in range(args.warmup):
            proc.dev_sampleScatterIsotropic(ekin.ptr, n, oe.ptr, om.ptr, stream=stream.ptr)
        stream.sync()
        nerr = proc.dev_check_errors(stream=stream.ptr)
        ev0, ev1 = C.Event(), C.Event()
        t0 = time.perf_counter()
        ev0.record(stream)
        for _ in range(args.steps):
            proc.dev_sampleScatterIsotropic(ekin.ptr, n, oe.ptr, om.ptr, stream=stream.ptr)
        ev1.record(stream)
        stream.sync()
        wall = time.perf_counter() - t0
        nerr += proc.dev_check_errors(stream=stream.ptr)
        mu = om.numpy(1000000)
        ef = oe.numpy(1000000)
        assert np.all(np.abs(mu) <= 1.0) and np.all(ef >= 0.0) and nerr == 0, "invalid samples"
        return dict(wall=wall, ms=ev0.elapsed_ms(ev1) / args.steps, value=n * args.steps / wall)

    plain = core.create_sabscatter(*targs)
    r_plain = measure(plain)
    progress(f"C3 plain sampler: {r_plain['ms']:.3f} ms per step")
    # learned proposal: trained on the plain sampler's samples with the GPU kernels
    t0 = time.perf_counter()
    model = core.MLPProposal(1e-3, 1.0, seed=1)
    loss = model.train_sab(plain, steps=args.train_steps, batch=8192, lr=3e-3, seed=1)
    learned = core.create_sabscatter(*targs)
    learned.use_learned_sampler(model, nintervals=256, emin=0.999e-3, emax=1.001, nbox_w=128, nbox_k=32)
    t_setup = time.perf_counter() - t0
    mass, uni, lam, sq = learned.envelope_stats().T
    r_learned = measure(learned)
    progress(f"C3 learned sampler: {r_learned['ms']:.3f} ms per step (training and setup {t_setup:.1f} s)")
    # the same sampler on the uniform grid: what the network gains
    envelope = core.create_sabscatter(*targs)
    envelope.use_envelope_sampler(nintervals=256, emin=0.999e-3, emax=1.001, nbox_w=128, nbox_k=32)
    emass, _, _, esq = envelope.envelope_stats().T
    r_env = measure(envelope)
    progress(f"C3 uniform-grid envelope sampler: {r_env['ms']:.3f} ms per step")

    threads = cpu_threads(args)
    ks = None
    if args.ks_samples > 0:
        ks = ks_check({"plain": plain, "learned": learned, "envelope": envelope}, table,
                      [0.001, 0.01, 0.025, 0.1, 0.5, 1.0], args.ks_samples, threads)
    progress("C3 KS checks done")
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(table, args.cpu_batch, args.cpu_seconds, threads)

    achieved = BYTES_PER_SAMPLE * n / (r_learned["ms"] * 1e-3) / 1e9
    out = {
        "metric": "scatter samples/s on 1 MI355X: learned vs plain SAB sampler at equal KS error",
        "value": r_learned["value"],
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r_learned["wall"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "S(alpha,beta) 200x800 from 3-Gaussian VDOS, T=293K, "
                               "E log-uniform 1e-3..1 eV, learned-proposal sampler: box envelope on grids "
                               "placed by the MLP (3x64, BF16 matrix cores) with exact accept/reject",
                   "neutrons_per_step": n, "sampler": "learned", "parallelism": "single GPU",
                   "plain_samples_per_s": r_plain["value"], "learned_over_plain": r_learned["value"] / r_plain["value"],
                   "uniform_grid_samples_per_s": r_env["value"],
                   "learned_over_uniform_grid": r_learned["value"] / r_env["value"],
                   "envelope_mass_uniform_over_learned_median": float(np.median(uni / mass)),
                   "squeeze_fraction_median_learned": float(np.median(sq / mass)),
                   "squeeze_fraction_median_uniform": float(np.median(esq / emass)),
                   "train_steps": args.train_steps, "train_and_setup_s": t_setup, "train_final_loss": loss,
                   "ks_vs_reference": ks, "ks_samples_per_energy": args.ks_samples},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": LEARNED_TRAFFIC_GB,
                     "traffic_unit": "GB per launch (FETCH_SIZE + WRITE_SIZE, each scaled by a plain copy)",
                     "kernel": "k_kb_sample<0>", "kernel_ms": r_learned["ms"],
                     "plain_kernel": "k_sab_sample_iso_main", "plain_kernel_ms": r_plain["ms"],
                     "bytes_per_sample": BYTES_PER_SAMPLE},
        "cpu_baseline": cpu,
    }
    return out


def run_vdos(args):
    """VDOS -> S(alpha,beta) expansion of the aluminium VDOS of the data
    library (tests/golden/vdos_reference.json) at 293.15 K, vdoslux 5 (the
    largest standard table, 1600 x 3200): whole calls (host chain + GPU
    operand tables and order sum) per second; roofline of the order sum on
    the FP64 matrix cores"""
    import json as _json
    import numpy as np
    from ncrystal_amd import core
    fix = _json.loads((ROOT / "tests" / "golden" / "vdos_reference.json").read_text())
    v = fix["vdos"]["Al_sg225/Al"]
    vin = (np.array(v["egrid"]), np.array(v["density"]))
    T0, mass, lux = 293.15, v["mass"], 5
    steps = max(1, min(args.steps, 5))
    for w in range(max(1, min(args.warmup, 2))):
        core.vdos_kernel(vin, T0 + 1e-9 * (w + 1), mass, vdoslux=lux)
    # each call at a temperature 1e-12 apart, so that every call expands anew
    t0 = time.perf_counter()
    gms = 0.0
    for i in range(steps):
        k = core.vdos_kernel(vin, T0 * (1.0 + 1e-12 * i), mass, vdoslux=lux)
        gms += k["gpu_ms"]
    wall = time.perf_counter() - t0
    gms /= steps
    M, N, K = k["rows_summed"], k["alpha"].size, k["orders_summed"]
    flops = 2.0 * M * N * ((K + 3) // 4 * 4)
    ach = flops / (gms * 1e-3) / 1e12
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import pyoracle
        pyoracle.build()
        lo, hi, d = pyoracle.vdos_regularise(*vin)
        tc = time.perf_counter()
        pyoracle.vdos_kernel(lo, hi, d, T0, mass, vdoslux=lux)
        dt = time.perf_counter() - tc
        cpu = dict(value=1.0 / dt, unit="expansions/s", cores=1, kind="port",
                   sample=f"one expansion of the same VDOS at vdoslux {lux} by the checker (host chain and "
                          f"order sum on one core), {dt:.1f} s")
    return {
        "metric": "VDOS -> S(alpha,beta) expansions per second on 1 MI355X",
        "value": steps / wall, "unit": "expansions/s", "steps": steps, "warmup": 1,
        "ms_per_step": wall / steps * 1e3, "dtype": "f64",
        "config": {"workload": "aluminium VDOS of the data library, T 293.15 K, vdoslux 5: %dx%d table, %d phonon "
                               "orders (%d summed)" % (N, k["beta"].size, k["norders"], K),
                   "gpu_sum_ms": gms},
        "roofline": {"bound": "mfma", "achieved": ach, "peak": 78.6, "unit": "TFLOP/s", "frac": ach / 78.6,
                     "traffic": None, "kernel": "k_vdos_gtab + k_vdos_afac + k_vdos_fill (order sum, "
                     "v_mfma_f64_16x16x4f64)", "kernel_ms": gms, "flops_per_launch": flops},
        "cpu_baseline": cpu,
    }


def al_sg225_inputs():
    """NCrystal's Al_sg225 (structure, scattering lengths, absorption, VDOS of
    the data library) from tests/golden/vdos_reference.json"""
    import json as _json
    fix = _json.loads((ROOT / "tests" / "golden" / "vdos_reference.json").read_text())
    return fix["basic"], fix["vdos"][fix["basic"]["vdos"]]


def run_minimc(args):
    """MiniMC: a pencil beam of 4 Aa neutrons into a sphere of NCrystal's
    Al_sg225 one scattering length across (the reference's mmc_al test
    geometry), forced-collision engine with absorption and roulette, theta
    tally with the breakdown by scattering history; source neutrons per
    second of whole engine calls; roofline of the engine's step kernel"""
    import numpy as np
    from ncrystal_amd import core
    b, v = al_sg225_inputs()
    T, mass, a = b["temperature"], b["mass"], b["a"]
    vin = (np.array(v["egrid"]), np.array(v["density"]))
    sigma_bound = 4 * math.pi * b["coh_sl"] ** 2 + b["sigma_inc"]
    msd = core.vdos_analyse(vin, T, mass)["msd"]
    hkl, mult, dsp, fsq, _, _ = core.hkl_spacegroup_members(b["spacegroup"], (a, a, a, 90, 90, 90),
                                                             [(b["coh_sl"], msd, np.array(b["positions"]))],
                                                             b["dcutoff"])
    ei = core.create_elincscatter([msd], [b["sigma_inc"]], [1.0])
    pb = core.create_powderbragg(a ** 3 * len(b["positions"]), dsp, fsq * mult)
    inel = core.create_sabscatter_from_vdos(vin, T, mass, sigma_bound, vdoslux=3)
    scatter = core.create_composition([ei, pb, inel], [1.0, 1.0, 1.0])
    ab = core.create_absoov(b["sigma_abs"])
    nd = len(b["positions"]) / a ** 3
    ekin = 0.081804209605330899 / 4.0 ** 2
    r = 0.5 / (nd * 100.0 * scatter.crossSectionIsotropic(np.array([ekin]))[0])
    z = -r * (1 - 1e-13)
    n = 10_000_000
    steps = max(1, min(args.steps, 5))
    for w in range(max(1, min(args.warmup, 2))):
        core.minimc_sphere(scatter, ab, nd, r, ekin, n, pos=(0.0, 0.0, z), seed=100 + w)
    t0 = time.perf_counter()
    ms = nst = nsc = 0.0
    for i in range(steps):
        res = core.minimc_sphere(scatter, ab, nd, r, ekin, n, pos=(0.0, 0.0, z), seed=i)
        ms += res["info"]["step_kernel_ms"]
        nst += res["info"]["neutron_steps"]
        nsc += res["info"]["scattered"]
    wall = time.perf_counter() - t0
    tot = res["total"]["content"].sum()
    assert 0.5 * n < tot < n, "tallied weight out of range"
    # step kernel: reads 80 B of neutron state + 16 B of cross sections per
    # neutron and step, writes 80 B per neutron passed on to scatter
    bytes_ = 96.0 * nst + 80.0 * nsc
    ach = bytes_ / (ms * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import pyoracle as po
        import pymmc
        po.build()
        threads = cpu_threads(args)
        lo, hi, d = po.vdos_regularise(*vin)
        k = po.vdos_kernel(lo, hi, d, T, mass, vdoslux=3)
        sf = po.sab_std_format(0, k["alpha"], k["beta"], k["sab"], beta_grid_optimised=True)
        cms = po.vdos_analyse(lo, hi, d, T, mass)["msd"]
        h2, m2, d2, f2, _, _ = po.hkl_spacegroup_members(b["spacegroup"], (a, a, a, 90, 90, 90),
                                                         [(b["coh_sl"], cms, np.array(b["positions"]))], b["dcutoff"])
        cs = po.composition([po.elinc(np.array([cms]), np.array([b["sigma_inc"]]), np.array([1.0])),
                             po.powderbragg(a ** 3 * len(b["positions"]), d2, f2 * m2),
                             po.sab(sf["alpha"], sf["beta"], sf["sab"], T, mass, sigma_bound,
                                    suggested_emax=k["suggested_emax"])], [1.0, 1.0, 1.0])
        cab = po.absoov(b["sigma_abs"])
        nc = 200_000
        tc = time.perf_counter()
        pymmc.minimc_sphere(cs, cab, nd, r, ekin, nc, pos=(0.0, 0.0, z), seed=1, nthreads=threads)
        dt = time.perf_counter() - tc
        cpu = dict(value=nc / dt, unit="neutrons/s", cores=threads, kind="port",
                   sample=f"{nc} source neutrons of the same simulation by the checker's engine (numpy steps over "
                          f"the checker's composition, {threads} threads), {dt:.1f} s")
    return {
        "metric": "MiniMC source neutrons/s on 1 MI355X (Al_sg225 sphere, 4 Aa pencil beam)",
        "value": n * steps / wall, "unit": "neutrons/s", "steps": steps, "warmup": 1,
        "ms_per_step": wall / steps * 1e3, "dtype": "f64",
        "config": {"workload": "pencil beam of 4 Aa neutrons into a sphere of Al_sg225 (ElInc + PowderBragg + "
                               "VDOS S(alpha,beta), 1/v absorption) one scattering length across (r = %.4f m), "
                               "forced collisions with roulette, theta tally with breakdown" % r,
                   "neutrons_per_step": n, "neutron_steps_per_call": nst / steps,
                   "engine_steps_per_call": res["info"]["max_steps"], "tallied_weight_fraction": tot / n},
        "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                     "traffic": None, "kernel": "k_mmc_step (engine step: transmission, tally, collision point)",
                     "kernel_ms": ms / steps, "bytes_per_call": bytes_ / steps},
        "cpu_baseline": cpu,
    }


def run_lcbragg(args):
    """LCBragg, the layered-crystal model (lcmode=0) of pyrolytic graphite
    (the reference's app_lc material: sg 194, dcutoff 0.5 Aa, 4 deg mosaicity,
    layer axis (0,0,1)): cross sections and scattering of 2^20 neutrons with
    wavelengths uniform in 0.5-7 Aa and isotropic directions, device-resident"""
    import json as _json
    import numpy as np
    from ncrystal_amd import _capi as C, core
    g = _json.loads((ROOT / "tests" / "golden" / "lc_graphite.json").read_text())
    v = _json.loads((ROOT / "tests" / "golden" / "vdos_reference.json").read_text())["vdos"][
        "C_sg194_pyrolytic_graphite/C"]
    h = g["laz"]["header"]
    a, c = h["lattice_a"], h["lattice_c"]
    cell = (a, a, c, 90, 90, 120)
    msd = core.vdos_analyse((np.array(v["egrid"]), np.array(v["density"])), h["temperature"], h["average_mass"])["msd"]
    csl = math.sqrt(h["sigma_coh"] / (4 * math.pi))
    pos = np.array(g["positions"])
    _, _, dsp, fsq, moff, mh = core.hkl_spacegroup_members(g["spacegroup"], cell, [(csl, msd, pos)], g["dcutoff"])
    L = core.reci_lattice(cell)
    lc = core.create_lcbragg(h["Vc"] * len(pos), dsp, fsq, moff, mh, L, np.eye(3), g["lcaxis"],
                             math.radians(g["mosaicity_deg"]))
    n = 1 << 20
    rng = np.random.default_rng(12)
    wl = rng.uniform(0.5, 7.0, n)
    e = 0.081804209605330899 / wl ** 2
    dirs = rng.normal(size=(n, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    stream = C.Stream()
    de, dd = C.DeviceArray.from_numpy(e), C.DeviceArray.from_numpy(dirs.ravel())
    xs, oe, od = C.DeviceArray(n), C.DeviceArray(n), C.DeviceArray(3 * n)
    lc.setRandomSeed(5)
    ev = [C.Event() for _ in range(3)]
    steps = max(1, min(args.steps, 5))
    for _ in range(max(1, min(args.warmup, 2))):
        lc.dev_crossSection(de.ptr, dd.ptr, n, xs.ptr, stream=stream.ptr)
        lc.dev_sampleScatter(de.ptr, dd.ptr, n, oe.ptr, od.ptr, stream=stream.ptr)
    stream.sync()
    xs_ms = samp_ms = 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        ev[0].record(stream)
        lc.dev_crossSection(de.ptr, dd.ptr, n, xs.ptr, stream=stream.ptr)
        ev[1].record(stream)
        lc.dev_sampleScatter(de.ptr, dd.ptr, n, oe.ptr, od.ptr, stream=stream.ptr)
        ev[2].record(stream)
        stream.sync()
        xs_ms += ev[0].elapsed_ms(ev[1])
        samp_ms += ev[1].elapsed_ms(ev[2])
    wall = time.perf_counter() - t0
    xs_ms /= steps
    samp_ms /= steps
    x = xs.numpy()
    assert np.all(np.isfinite(x)) and np.mean(x > 0) > 0.3, "LCBragg cross sections out of range"
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import pyoracle as po
        po.build()
        lo, hi, dv = po.vdos_regularise(np.array(v["egrid"]), np.array(v["density"]))
        cms = po.vdos_analyse(lo, hi, dv, h["temperature"], h["average_mass"])["msd"]
        fam = po.hkl_spacegroup_members(g["spacegroup"], cell, [(csl, cms, pos)], g["dcutoff"])
        ol = po.reci_lattice(cell)
        osc = po.scbragg(a, a, c, h["Vc"] * len(pos), fam[2], fam[3], fam[4], fam[5], np.eye(3),
                         math.radians(g["mosaicity_deg"]), reci=ol)
        nrm = fam[5] @ ol.reshape(3, 3).T
        nrm /= np.linalg.norm(nrm, axis=1)[:, None]
        of = np.repeat(np.arange(len(fam[2])), np.diff(fam[4]))
        ochk = po.LCHelperXS(osc, nrm, fam[3][of], fam[2][of], g["lcaxis"], g["lcaxis"], h["Vc"] * len(pos))
        m = 0
        tc = time.perf_counter()
        while time.perf_counter() - tc < min(args.sub_cpu_seconds, 10.0) and m < n:
            ochk.xs_dir(e[m:m + 200], dirs[m:m + 200])
            m += 200
        dt = time.perf_counter() - tc
        cpu = dict(value=m / dt, unit="neutrons/s (cross sections)", cores=1, kind="port",
                   sample=f"{m} neutrons of the same batch through the checker's lcmode=0 cross section "
                          f"(one thread), {dt:.1f} s")
    return {
        "metric": "LCBragg lcmode=0 neutrons/s on 1 MI355X (pyrolytic graphite, cross section + scattering)",
        "value": n / ((xs_ms + samp_ms) * 1e-3), "unit": "neutrons/s", "steps": steps, "warmup": 1,
        "ms_per_step": wall / steps * 1e3, "dtype": "f64",
        "config": {"workload": "pyrolytic graphite layered crystal (sg 194, dcutoff 0.5 Aa, 4 deg mosaicity, "
                               "layer axis 001), 2^20 neutrons, 0.5-7 Aa, isotropic directions",
                   "neutrons": n, "xs_ms": xs_ms, "sample_ms": samp_ms,
                   "xs_neutrons_per_s": n / (xs_ms * 1e-3)},
        "roofline": {"bound": "fp64-valu", "achieved": None, "peak": None, "unit": None, "frac": None,
                     "traffic": None, "kernel": "k_lc0_xs (per-neutron Romberg averages over rotation ranges)",
                     "note": "40 B of HBM traffic per neutron; the FP64 Romberg arithmetic of each range bounds it"},
        "cpu_baseline": cpu,
    }


def copy_calibration(nbytes=800_000_000):
    """one plain device-to-device copy of nbytes (read nbytes, write nbytes):
    the reference launch that scales the FETCH_SIZE / WRITE_SIZE counters of
    a profiling pass of this process"""
    from ncrystal_amd import _capi as C
    stream = C.Stream()
    src, dst = C.DeviceArray(nbytes // 8), C.DeviceArray(nbytes // 8)
    C.call("ncamd_memcpy_dtod", dst.ptr, src.ptr, nbytes, stream.ptr)
    stream.sync()
    progress(f"copy calibration: {nbytes} bytes device to device")


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only (a one-GPU node); run with --gpus 1"}))
        sys.exit(2)
    runners = {"sab": run_sab, "c1": run_c1, "powderbragg": run_powderbragg, "batch": run_batch,
               "scbragg": run_scbragg, "vdos": run_vdos, "minimc": run_minimc, "lcbragg": run_lcbragg}
    if args.copy_calib:
        copy_calibration()
    if args.workload != "all":
        out = runners[args.workload](args)
        out.update(n_gpus=1, host=host_info(), library=library_info())
        print(json.dumps(out))
        return
    out = run_sab(args)
    sub_args = argparse.Namespace(**vars(args))
    sub_args.cpu_seconds = args.sub_cpu_seconds
    sub = {}
    for key, name in (("C1", "c1"), ("C2", "powderbragg"), ("C4", "batch"), ("C5", "scbragg"), ("F2", "vdos"),
                      ("F3", "minimc"), ("F4", "lcbragg")):
        t0 = time.perf_counter()
        r = runners[name](sub_args)
        for k in ("n_gpus", "higher_is_better", "scaling", "vs_baseline", "data"):
            r.pop(k, None)
        r["wall_s"] = time.perf_counter() - t0
        sub[key] = r
        print(f"[bench] {key} done: {r['value']:.4g} {r['unit']}", file=sys.stderr, flush=True)
    out["config"]["workload_key"] = "C3"
    out["subresults"] = sub
    out["host"] = host_info()
    out["library"] = library_info()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
