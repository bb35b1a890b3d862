"""Benchmark of the EDIPIC-2D particle hot path on one MI355X.

Workload (the headline): 2048x1024 cells, X-periodic with grounded walls at y=0 and
y=Ly, 200 particles/cell per species (electrons + Ar+), uniform Bz = 0.01 T,
Te = 2 eV, n = 1e15 m^-3, Ar gas 1e20 m^-3 at 300 K with e-n MCC (elastic,
excitation, ionisation) and ion charge exchange, FFT_X Poisson, FP64 particles,
N_subcycles = 1.  One step = the hot-path part of one iteration of the reference
main loop (gather rho -> Poisson -> E -> push e + MCC -> push i + RCX -> appends).
Synthetic initial particles (uniform positions, Maxwellian velocities) are
generated on the device.

Prints one JSON line.  `value` is particle-steps/s over the timed steps (electrons +
ions pushed per step / wall time).  Extra objects: the same workload with mixed
precision particles, the Poisson solve on the inner-object geometry (1024x1024,
two biased boxes), the roofline (the whole step's particle-steps/s times the algorithmic
bytes per particle-step against the HBM peak, with the electron push kernel's own figure
beside it) and the CPU baseline (the C oracle on the host cores, on a bounded sample of the
same workload).
"""
import argparse
import concurrent.futures as cf
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "edipic-2d_amd"))

import numpy as np  # noqa: E402

import edgpu  # noqa: E402
from edgpu import config  # noqa: E402

HBM_PEAK_GBS = 8000.0


# Algorithmic bytes of one particle push: the particle state read once and written once,
# FP64 x, y, vx, vy, vz (80 B) or int32 cell + 5 x FP32 (48 B).
BYTES_PER_PARTICLE_STEP = {edgpu.FP64: 80, edgpu.MIXED: 48}


def push_bytes_moved(prob, precision):
    """Bytes the electron push actually moves per particle: with a magnetic field along z only
    and no field along z the push cannot change vz, so vz is read (collision test) but not
    written back: FP64 40 + 32 B, mixed 24 + 20 B."""
    vz_fixed = prob.Bx == 0.0 and prob.By == 0.0 and prob.Ez == 0.0
    full = BYTES_PER_PARTICLE_STEP[precision]
    return full - (8 if precision == edgpu.FP64 else 4) if vz_fixed else full


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_gpu(prob, precision, steps, warmup, sort_interval, ion_sort_interval, sorting_push=True):
    rt = edgpu.Runtime(prob, precision=precision, sort_interval=sort_interval, ion_sort_interval=ion_sort_interval,
                       capacity_factor=1.1)
    try:
        rt.set_sorting_push(sorting_push)
        rt.initiate_electron_neutral_collisions()
        rt.initiate_ion_neutral_collisions()
        rt.load_initial_particles()
        log("  loaded %d + %d particles" % (rt.count(0), rt.count(1)))
        for _ in range(warmup):
            rt.step()
        ps = 0
        push_e_ms, push_i_ms, n_e, sorted_step = [], [], [], []
        phases = {}
        t0 = time.perf_counter()
        for _ in range(steps):
            se, si = rt.step()
            ps += se["n_before"] + si["n_before"]
            t = rt.last_step_timing()
            push_e_ms.append(t["push_e"])
            push_i_ms.append(t["push_i"])
            sorted_step.append(t["sort"] > 0.05)  # a sort or a tile count ran (empty: a few microseconds)
            n_e.append(se["n_before"])
            for k, v in t.items():
                phases[k] = phases.get(k, 0.0) + v
        t1 = time.perf_counter()
        coll = se["collisions"]
        # the electron push of the steps without a sort (the plain push kernel), and that of the
        # steps with one (the sorting push, or the plain push after the separate sort)
        pe, srt = np.array(push_e_ms), np.array(sorted_step)
        plain_ms = float(np.mean(pe[~srt])) if (~srt).any() else float(np.mean(pe))
        sort_step_ms = float(np.mean(pe[srt])) if srt.any() else None
        return dict(value=ps / (t1 - t0), ms_per_step=1e3 * (t1 - t0) / steps, push_e_ms=float(np.mean(pe)),
                    push_e_plain_ms=plain_ms, push_e_sort_step_ms=sort_step_ms,
                    push_i_ms=float(np.mean(push_i_ms)), n_e=float(np.mean(n_e)),
                    phases_ms={k: v / steps for k, v in phases.items()}, n_final=(rt.count(0), rt.count(1)),
                    last_collisions=coll[:3], last_rcx=si["collisions"][0])
    finally:
        rt.close()


def run_poisson(pairs=24, train_steps=400, eval_steps=4):
    """Krylov Poisson on the inner-object geometry (1024x1024 cells, two boxes at +-50 V,
    100 ppc, PCG to relative residual 1e-8): the warm-up records (rho, phi) pairs from the
    run's own solves, the network trains on the device, then consecutive blocks of steps use
    the zero guess (the reference's KSP default), the previous potential and the learned
    guess, first with the multigrid preconditioner (the default), then with Jacobi.  The warm-up and evaluation lengths are scaled down from 2000 + 2000 steps so
    that the benchmark finishes in minutes."""
    prob = config.workload(4)
    rt = edgpu.Runtime(prob, precision=edgpu.FP64, capacity_factor=1.05, fuse_collisions=0)
    try:
        rt.load_initial_particles()
        rt.cnn_init(seed=11, max_pairs=pairs)

        def step(warm):
            rt.gather_ion_charge_density(host=False)
            rt.gather_electron_charge_density(host=False)
            t0 = time.perf_counter()
            _, it, rr = rt.solve_potential_krylov(rtol=1e-8, warm_start=warm, host=False)
            dt = time.perf_counter() - t0
            rt.clear_accumulated_fields()
            rt.calculate_electric_field(host=False)
            rt.advance_electrons()
            rt.advance_ions()
            return dt, it

        step(0)
        for _ in range(pairs):
            step(1)
            rt.cnn_record_pair()
        t0 = time.perf_counter()
        loss0 = rt.cnn_train(1, lr=1e-3)
        loss = rt.cnn_train(train_steps, lr=1e-3)
        t_train = time.perf_counter() - t0
        out = {"training": dict(pairs=pairs, adam_steps=train_steps + 1, seconds=t_train, loss_first=loss0,
                                loss_last16=loss)}
        names = {0: "zero_guess", 1: "previous_potential", 2: "learned"}
        for pc, pname in ((1, "multigrid"), (0, "jacobi")):
            rt.set_krylov_preconditioner(pc)
            blk = {}
            for warm in (0, 1, 2):
                r = [step(warm) for _ in range(eval_steps)]
                blk[names[warm]] = dict(ms=1e3 * float(np.mean([x[0] for x in r])),
                                        iterations=float(np.mean([x[1] for x in r])))
            blk["learned_speedup_vs_zero_guess"] = blk["zero_guess"]["ms"] / blk["learned"]["ms"]
            blk["learned_speedup_vs_previous"] = blk["previous_potential"]["ms"] / blk["learned"]["ms"]
            out[pname] = blk
        return out
    finally:
        rt.close()


def run_combined(steps, warmup, pairs=12, train_steps=300):
    """The combined largest single-GPU case (4096x2048 cells, 100 ppc per species, one grounded
    256x128 inner box, so the Krylov solve to relative residual 1e-8; y walls with secondary
    yield 0.5, a 100 eV beam at y = 0, e-n MCC + RCX): the plain FP64 path with the
    reference's zero initial guess, then mixed precision with the learned warm start, whose
    network trains on the run's own solves during the warm-up, and, for comparison, mixed
    precision starting each solve from the previous potential; all on the same seed."""
    prob = config.workload(5)
    out = {}
    for name, precision, warm in (("fp64_zero_guess", edgpu.FP64, 0), ("mixed_learned", edgpu.MIXED, 2),
                                  ("mixed_previous_potential", edgpu.MIXED, 1)):
        rt = edgpu.Runtime(prob, precision=precision, sort_interval=12, ion_sort_interval=512, capacity_factor=1.1)
        try:
            rt.setup_wall_emission()
            rt.initiate_electron_neutral_collisions()
            rt.initiate_ion_neutral_collisions()
            rt.load_initial_particles()
            log("  %s: loaded %d + %d particles" % (name, rt.count(0), rt.count(1)))
            t_train = None
            if warm == 2:
                rt.cnn_init(seed=11, max_pairs=pairs)
                rt.step(krylov_warm_start=1)
                for _ in range(pairs):
                    rt.step(krylov_warm_start=1)
                    rt.cnn_record_pair()
                t0 = time.perf_counter()
                rt.cnn_train(train_steps, lr=1e-3)
                t_train = time.perf_counter() - t0
            for _ in range(warmup):
                rt.step(krylov_warm_start=warm)
            ps, solve, emitted, hits = 0, 0.0, 0, 0
            t0 = time.perf_counter()
            for _ in range(steps):
                se, si = rt.step(krylov_warm_start=warm)
                ps += se["n_before"] + si["n_before"]
                solve += rt.last_step_timing()["solve"]
                emitted += sum(se["emit_count"])
                hits += sum(se["hit_count"])
            t1 = time.perf_counter()
            out[name] = dict(value=ps / (t1 - t0), ms_per_step=1e3 * (t1 - t0) / steps, solve_ms=solve / steps,
                             electrons_emitted_per_step=emitted / steps, electron_wall_hits_per_step=hits / steps,
                             train_seconds=t_train)
            log("  ", name, out[name])
        finally:
            rt.close()
    out["speedup"] = out["mixed_learned"]["value"] / out["fp64_zero_guess"]["value"]
    out["solve_speedup"] = out["fp64_zero_guess"]["solve_ms"] / out["mixed_learned"]["solve_ms"]
    return out


def host_cores():
    """Cores this process may run on, and the thread budget: the affinity mask, capped by
    OMP_NUM_THREADS where the host sets it (a GPU's share of a shared host)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    cap = os.environ.get("OMP_NUM_THREADS")
    threads = min(avail, int(cap)) if cap and cap.isdigit() and int(cap) > 0 else avail
    return avail, threads


def cpu_baseline(sample_ppc=25, steps=4, threads=None):
    """The C oracle (the project's FP64 C restatement of the reference routines, not the
    reference's Fortran) on the host cores: the headline grid with sample_ppc particles/cell,
    per step: deposit e+i, fold, FFT_X solve, E field, push e+i, e-n MCC (the reference's
    sampling without replacement, per chunk) and RCX.  Particles are split into one chunk per
    thread (as MPI ranks share a cluster in the reference)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ctypes as C

    import oracle_ffi as O
    avail, budget = host_cores()
    threads = threads or budget
    prob = config.workload(2)
    full_ppc = prob.N_of_particles_cell
    n = prob.n_particles * sample_ppc // full_ppc
    g = O.make_geom(prob.geometry())
    v, _ = O.maxwell_tables()
    chunks = np.array_split(np.arange(n), threads)

    def gen(k):
        w = O.well(1000 + k)
        m = chunks[k].size
        e = [np.empty(m) for _ in range(5)]
        O.lib().orc_distribute(C.byref(w), O.dp(v), m, prob.x_limits[0], prob.x_limits[1], prob.y_limits[0],
                               prob.y_limits[1], 0.0, float(prob.gmi), 0.0, float(prob.gmj), prob.factor_convert_e,
                               *[O.dp(a) for a in e])
        i = [np.empty(m) for _ in range(5)]
        O.lib().orc_distribute_ions(C.byref(w), O.dp(v), m, O.dp(e[0]), O.dp(e[1]), prob.factor_convert_i,
                                    *[O.dp(a) for a in i])
        return e, i, w

    with cf.ThreadPoolExecutor(threads) as ex:
        parts = list(ex.map(gen, range(threads)))
    tab = O.MccTable([config.ENERGY_SEGMENTS[0]] + [s[0] for s in config.ENERGY_SEGMENTS[1]],
                     [s[1] for s in config.ENERGY_SEGMENTS[1]],
                     [dict(p, ion_velocity_factor=prob.ion_velocity_factor()) for p in prob.cross_sections()],
                     prob.M_i_amu, prob.neutral_N_m3, prob.N_subcycles, prob.dt, prob.V_scale)
    rcx = O.Rcx()
    s1, alpha = prob.rcx_sigma()
    O.lib().orc_rcx_setup(C.byref(rcx), s1, alpha, prob.neutral_N_m3, prob.neutral_T_K, prob.dt, prob.N_subcycles,
                          prob.Ms, prob.energy_factor_eV, prob.V_scale, prob.T_e_eV, prob.N_max_vel)
    a_eq = O.prepare_sys_y(prob.gmi - 1, prob.gmj)
    sp_e = O.make_species(True)
    sp_i = O.make_species(False, Qs=prob.Qs, QM2s=prob.QM2s, N_subcycles=1, sense_B=1)
    B = (prob.Bx, prob.By, prob.Bz)
    nn = prob.nx * prob.ny
    accEX, accEY = np.zeros(nn), np.zeros(nn)

    def deposit(k, which):
        e, i, _ = parts[k]
        a = e if which == 0 else i
        return O.deposit(g, 1, a[0], a[1])

    def push(k, EX, EY, aEX, aEY):
        e, i, w = parts[k]
        out = []
        for arrs, sp, fx, fy in ((e, sp_e, EX, EY), (i, sp_i, aEX, aEY)):
            rc, rem, _, _ = O.advance(g, sp, fx, fy, B, prob.Ez, *arrs, 4)
            keep = rem == 0
            for q in range(5):
                arrs[q] = np.ascontiguousarray(arrs[q][keep])
            out.append(keep.size)
        m = e[0].size
        kind = np.zeros(m, np.int32)
        cnt = np.zeros(8, np.int64)
        nnew = C.c_int64(0)
        cap = max(1024, m // 20)
        enew, inew = np.zeros(5 * cap), np.zeros(5 * cap)
        O.lib().orc_mcc_en_reference(C.byref(tab.t), O.dp(v), C.byref(w), m, O.dp(e[0]), O.dp(e[1]), O.dp(e[2]),
                                     O.dp(e[3]), O.dp(e[4]), O.ip(kind), O.lp(cnt), cap, C.byref(nnew), O.dp(enew),
                                     O.dp(inew))
        O.lib().orc_rcx_reference(C.byref(rcx), O.dp(v), C.byref(w), i[0].size, O.dp(i[2]), O.dp(i[3]), O.dp(i[4]))
        return sum(out)

    total = 0
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(threads) as ex:
        for _ in range(steps):
            rho_e = O.fold_walls(g, sum(ex.map(lambda k: deposit(k, 0), range(threads))))
            rho_i = O.fold_walls(g, sum(ex.map(lambda k: deposit(k, 1), range(threads))))
            phi = O.solve_poisson_fftx(g, prob.N_of_particles_cell, rho_e, rho_i, a_eq)
            EX, EY = O.efield(g, phi, prob.factor1_E_from_F, prob.factor2_E_from_F, (accEX, accEY))
            total += sum(ex.map(lambda k: push(k, EX, EY, accEX, accEY), range(threads)))
            accEX[:] = 0.0
            accEY[:] = 0.0
    t1 = time.perf_counter()
    ref_push = None
    if O.ref_push_lib() is not None:
        # the reference's own ADVANCE_ELECTRONS (compiled from its sources) beside the C
        # restatement's push, one thread each, on the same electrons in the last step's field
        e = [a.copy() for a in parts[0][0]]
        m = e[0].size
        t2 = time.perf_counter()
        O.ref_advance_electrons(prob.geometry(), B, prob.Ez, *e, EX, EY)
        t3 = time.perf_counter()
        O.advance(g, sp_e, EX, EY, B, prob.Ez, *[a.copy() for a in e], 4)
        t4 = time.perf_counter()
        ref_push = dict(value=m / (t3 - t2), port_value=m / (t4 - t3), unit="electron pushes/s", cores=1,
                        sample="%d electrons of the sample, one ADVANCE_ELECTRONS of the reference's own Fortran "
                               "(its copy into and out of the particle array included) and one push of the C "
                               "restatement, one thread each" % m)
    return dict(value=total / (t1 - t0), unit="particle-steps/s", cores=threads, kind="port",
                reference_push=ref_push, host_cores_available=avail, host_cpu_count=os.cpu_count(),
                sample="2048x1024 cells, %d particles/cell per species (%d of each, 1/%d of the workload), %d steps, "
                       "the C restatement of the reference routines on %d host threads (%d cores in the affinity "
                       "mask, OMP_NUM_THREADS=%s)" % (sample_ppc, n, full_ppc // sample_ppc, steps, threads, avail,
                                                      os.environ.get("OMP_NUM_THREADS")))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--sort-interval", type=int, default=12, help="electron steps between per-cell sorts")
    ap.add_argument("--ion-sort-interval", type=int, default=512,
                    help="steps between ion sorts (ions drift ~1e-4 cell per step: their push does not slow down "
                         "between sorts)")
    ap.add_argument("--sorting-push", type=int, default=1, choices=[0, 1],
                    help="sort the electrons of a sort step with their push (1) or with a separate sort (0)")
    ap.add_argument("--skip-mixed", action="store_true")
    ap.add_argument("--skip-poisson", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--only-precision", choices=["fp64", "mixed"], default=None)
    ap.add_argument("--combined", action="store_true",
                    help="run the combined 4096x2048 case (inner box, wall emission, Krylov) instead")
    ap.add_argument("--copy-probe", type=float, default=0.0,
                    help="only run the plain copy kernel on this many GiB and print its bandwidth")
    args = ap.parse_args()
    if args.copy_probe > 0:
        nbytes = int(args.copy_probe * (1 << 30))
        ms = edgpu.copy_probe(nbytes, 10)
        print(json.dumps({"copy_probe_bytes": nbytes, "ms": ms, "GBps_read_plus_write": 2 * nbytes / ms / 1e6}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)
    if args.combined:
        log("combined 4096x2048 case: %d steps after %d warm-up" % (args.steps, args.warmup))
        c = run_combined(args.steps, args.warmup)
        print(json.dumps({"metric": "particle-steps/s (push+deposit+MCC+emission), Krylov Poisson",
                          "value": c["mixed_learned"]["value"], "unit": "particle-steps/s", "n_gpus": 1,
                          "steps": args.steps, "warmup": args.warmup,
                          "ms_per_step": c["mixed_learned"]["ms_per_step"], "higher_is_better": True,
                          "scaling": "weak", "vs_baseline": None, "dtype": "f32+i32 (f64 grid)",
                          "data": "synthetic (device-generated uniform positions, Maxwellian velocities)",
                          "config": {"workload": "4096x2048 cells, 100 ppc/species, grounded 256x128 inner box, "
                                                 "periodic x, SEE yield 0.5, 100 eV beam, PCG to 1e-8"},
                          "combined": c}), flush=True)
        return
    prob = config.workload(2)
    res = {}
    if args.only_precision != "mixed":
        log("fp64: %d steps after %d warm-up" % (args.steps, args.warmup))
        res["fp64"] = run_gpu(prob, edgpu.FP64, args.steps, args.warmup, args.sort_interval, args.ion_sort_interval,
                              bool(args.sorting_push))
        log("  ", res["fp64"])
    if not args.skip_mixed and args.only_precision != "fp64":
        log("mixed: %d steps after %d warm-up" % (args.steps, args.warmup))
        res["mixed"] = run_gpu(prob, edgpu.MIXED, args.steps, args.warmup, args.sort_interval, args.ion_sort_interval,
                               bool(args.sorting_push))
        log("  ", res["mixed"])
    head = res.get("fp64") or res.get("mixed")
    prec = edgpu.FP64 if "fp64" in res else edgpu.MIXED
    bpp = BYTES_PER_PARTICLE_STEP[prec]
    # whole step: particle-steps/s x algorithmic bytes per particle-step against the HBM peak
    # (sort, compaction, grid work and host gaps are charged against it, not added)
    achieved = bpp * head["value"] / 1e9
    # the dominant kernel alone: the electron push of the steps without a sort, its algorithmic
    # bytes over its mean duration measured with HIP events on the stream it runs on
    kernel_achieved = bpp * head["n_e"] / (head["push_e_plain_ms"] * 1e-3) / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "push_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("fp64" if prec == edgpu.FP64 else "mixed")
    line = {
        "metric": "particle-steps/s (push+deposit+MCC), % HBM roofline; Poisson ms/step learned vs plain",
        "value": head["value"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == edgpu.FP64 else "f32+i32 (f64 grid)",
        "data": "synthetic (device-generated uniform positions, Maxwellian velocities)",
        "config": {"workload": "2048x1024 cells, 200 ppc/species, Bz=0.01 T, periodic x, FFT_X, e-n MCC + RCX, "
                               "FP64 particles, N_subcycles=1",
                   "particles_per_species": prob.n_particles, "grid": [prob.gmi - 1, prob.gmj],
                   "sort_interval": {"electrons": args.sort_interval, "ions": args.ion_sort_interval},
                   "electrons_sorted_by_push": bool(args.sorting_push)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "scope": "whole step: particle-steps/s x %d B per particle-step" % bpp,
                     "algorithmic_bytes_per_particle": bpp,
                     "kernel": "k_advance_%s<electron, MCC candidacy> (push+gather+walls+deposit)" %
                               ("fp64" if prec == edgpu.FP64 else "mixed"),
                     "kernel_ms": head["push_e_plain_ms"], "kernel_achieved": kernel_achieved,
                     "sort_step_push_ms": head["push_e_sort_step_ms"],
                     "kernel_frac": kernel_achieved / HBM_PEAK_GBS,
                     "bytes_moved_per_particle": push_bytes_moved(prob, prec)},
        "phases_ms": head["phases_ms"],
    }
    if "mixed" in res:
        m = res["mixed"]
        line["mixed"] = {"value": m["value"], "ms_per_step": m["ms_per_step"], "push_e_ms": m["push_e_plain_ms"],
                         "sort_step_push_ms": m["push_e_sort_step_ms"],
                         "roofline_frac": 48 * m["value"] / 1e9 / HBM_PEAK_GBS,
                         "kernel_frac": 48 * m["n_e"] / (m["push_e_plain_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                         "phases_ms": m["phases_ms"],
                         "speedup_vs_fp64": (m["value"] / res["fp64"]["value"]) if "fp64" in res else None}
    if not args.skip_poisson:
        log("poisson (1024x1024, two biased boxes, PCG to 1e-8, learned warm start)")
        line["poisson"] = run_poisson()
        log("  ", line["poisson"])
    if not args.skip_cpu:
        log("cpu baseline")
        line["cpu_baseline"] = cpu_baseline()
        log("  ", line["cpu_baseline"])
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
