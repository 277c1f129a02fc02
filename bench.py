#!/usr/bin/env python3
"""Benchmark of the plane-wave inner loop on one MI355X.

A step is one Car-Parrinello MD step of the 512-state workload on the 192^3
mesh (198,204 Gamma plane waves, 1.6M density G vectors, cubic a = 32 bohr,
256 atoms, ecut 80 Ry) with every input resident in HBM, in CPMD's velocity
Verlet order (src/mdmain_utils.mod.F90:710-966):

  velupi, velupa + posupi + phfac + posupa (rortog),
  forces: rnlsm1 + rnlsm2 -> rhoofr -> vofrho (+ potfor) -> Ewald
          -> vpsi + fnonloc (c2 = -f H c) -> rnlfor,
  velupi, velupa + rortv, rekine.

The default start is the ODIIS ground state of the seeded coefficients (30
FP64 steepest-descent steps, then DIIS to gemax < 1e-5, untimed), dt = 4 a.u.,
fictitious mass 400 a.u.  --step wfopt times a steepest-descent wavefunction
optimisation step instead (Lagrange projection, update, Loewdin).
Wavefunctions: seeded complex Gaussian ~exp(-|G|^2/8), Loewdin-orthonormalised;
GTH pseudopotential with one s projector per atom; LDA (Pade).

Two precision paths are timed from the same start, the first one being the
headline `value`:
  mixed  rhoofr/vpsi transforms and their workspaces in FP32, the wavefunction
         contractions as 3-term bf16 split MFMA products (FP32 accumulation);
         vofrho, Ewald, n x n algebra and every energy sum in FP64;
  fp64   every operation in FP64.
Reported errors: E_KS and F_ion of the headline path against FP64 on the same
coefficients and positions, the same after the warmup + timed steps from the
same start (both evaluated in FP64), ||C^T C - I||, and the E_ham = E_KS +
ekinp + ekinc drift of each path.

Prints one JSON line: metric, value (ms per CP step of the mixed path), the
roofline of the step's dominant kernel (the largest time per step among the
transform kernels and the A^T B contraction kernel, with the measured issue
rate of its MFMA instruction), the HBM rate of every transform stage, MFMA
summaries, the error fields and the CPU baseline.  --pmc-copy only runs the
known-size copy that scales rocprofv3's byte counters.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "CP step ms at 512 states/192³: mixed vs FP64 GPU, |ΔE|<1e-6 Ha/atom"
HBM_PEAK_GBS = 8000.0


def log(msg: str):
    print(msg, file=sys.stderr, flush=True)


def kernel_bytes(ctx_geom, kid: str, np_pairs: float, cw: float = 16.0) -> float:
    """Algorithmic HBM bytes of one launch of kernel `kid` processing np_pairs
    state pairs.  cw = bytes of a workspace complex (16 FP64, 8 FP32);
    wavefunctions are FP64 complex (16 B), V and rho FP64 (8 B)."""
    ngw, S, Zw, n1, n2, n3 = ctx_geom
    c, r = 16.0, 8.0
    nnn = n1 * n2 * n3
    col = n3 * Zw * n2          # column workspace per pair (complex)
    stick = S * n3              # stick workspace per pair (complex)
    if kid == "zinv":
        return np_pairs * (c * 2 * ngw + cw * stick)
    if kid == "yinv" or kid == "yfwd":
        return np_pairs * cw * (stick + col)
    if kid == "xrho":
        return np_pairs * cw * col + 2 * r * nnn
    if kid == "xvpsi":
        return np_pairs * cw * 2 * col + r * nnn
    if kid == "zfwd":
        # c0 read, c2 written (mc_forces_c2_dev overwrites c2: no read of it)
        return np_pairs * (cw * stick + c * 4 * ngw)
    raise KeyError(kid)


def step_model_bytes(ctx_geom, nstate: int, cw: float = 16.0) -> float:
    """Per-step algorithmic bytes of rhoofr + vpsi (the stick-sparse transform model with the
    transform workspaces at cw bytes per complex): per state pair
    vpsi = cw(4 S N + 4 N^2 Zw) + 16*6 NGW + 8 N^3, rhoofr = cw(2 S N + 2 N^2 Zw) + 16*2 NGW,
    plus rho written once (8 N^3).  This per-pair model reads V once per pair and runs vpsi's
    inverse passes; the kernels read V once per launch and, with column reuse, vpsi starts
    from rhoofr's columns, so the transform roofline uses the bytes of the launches made
    (kernel_bytes) and reports this model beside it."""
    ngw, S, Zw, n1, n2, n3 = ctx_geom
    r = 8.0
    npairs = (nstate + 1) // 2
    vp = cw * (4 * S * n3 + 4 * n1 * n2 * Zw) + 16.0 * 6 * ngw + r * n1 * n2 * n3
    rh = cw * (2 * S * n3 + 2 * n1 * n2 * Zw) + 16.0 * 2 * ngw
    return npairs * (vp + rh) + 2 * r * n1 * n2 * n3


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(sysm, c0, v, sample_pairs: int = 64):
    """CPMD's CPU algorithm for the SCF-step content of the GPU measurement,
    on the host's cores: rhoofr + vpsi over Gamma state pairs with stick-sparse
    3D transforms (oracle/cpu_step.c: one state pair per OpenMP thread, 1D
    transforms through the FFTW3 interface), the projector contractions
    rnlsm1 + fnonloc and the Loewdin orthonormalisation (threaded BLAS /
    LAPACK).  The transforms run on `sample_pairs` of the state pairs and are
    scaled to all of them; the contractions and the orthonormalisation run in
    full."""
    from oracle.cpu_step import CpuStep
    from cpmd_amd import system as S
    nstate = c0.shape[1]
    npairs = (nstate + 1) // 2
    ps = min(sample_pairs, npairs)
    cs = CpuStep(sysm.cell, sysm.gset)
    threads = cs.threads()
    gs = sysm.gset
    ngw = gs.ngw
    f = sysm.f
    na, nhxs, ngh, lang, hmat, twnl, eigr = S.projector_set(sysm, gs.miller[:ngw], gs.hg[:ngw])
    beta = np.asfortranarray(twnl[:, 0, 0][:, None] * eigr)           # s projector: (-i)^0 twnl eigr
    cs.rhoofr(c0[:, :2], f[:2])                                         # plans, first touch
    t = {}
    t0 = time.perf_counter()
    cs.rhoofr(c0, f, pairs=ps)
    t["rhoofr"] = (time.perf_counter() - t0) * npairs / ps
    t0 = time.perf_counter()
    cs.vpsi(c0, np.zeros_like(c0), f, v, pairs=ps)
    t["vpsi"] = (time.perf_counter() - t0) * npairs / ps
    t0 = time.perf_counter()
    fnl = cs.rnlsm1(c0, beta)
    dd = -(f[None, :] * hmat[0, 0, 0]) * fnl
    cs.fnonloc(np.zeros_like(c0), beta, dd)
    t["rnlsm1+fnonloc"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    cs.lowdin(c0)
    t["lowdin"] = time.perf_counter() - t0
    total = sum(t.values())
    return {
        "value": round(total * 1e3, 1),
        "unit": "ms/step",
        "cores": threads,
        "cpu": cpu_model(),
        "kind": "port",
        "components_ms": {k: round(x * 1e3, 1) for k, x in t.items()},
        "sample": (f"rhoofr + vpsi with stick-sparse transforms (FFTW3 interface, one state pair per thread on "
                   f"{threads} threads) on {ps} of {npairs} state pairs scaled x{npairs / ps:g}, plus rnlsm1 + "
                   f"fnonloc + Loewdin in full (threaded BLAS / LAPACK): the SCF-step content of CPMD's CPU path "
                   f"(oracle/cpu_step.py)"),
    }


def gpu_same_content_ms(ctx, ks, buf, f, nstate, reps: int = 5) -> float:
    """The GPU's FP64 time for the content of cpu_baseline: rnlsm1, rhoofr,
    vpsi + fnonloc, Loewdin (ms per step, median of reps), with column reuse
    off so that vpsi transforms from C as the CPU path does."""
    ctx.set_precision("fp64")
    ctx.set_column_reuse(False)
    ts = []
    for _ in range(reps + 1):
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.rnlsm_dev(buf["c0"], nstate, tfor=False, want_fnl=False)
        ctx.rhoofr_dev(buf["c0"], f, nstate, ks.d_rho)
        ctx.forces_c2_dev(buf["c0"], buf["c2"], f, ks.d_v, nstate)
        ctx.lowdin_dev(buf["cp"], nstate)
        ctx.synchronize()
        ts.append(time.perf_counter() - t0)
    ctx.set_column_reuse(True)
    return float(np.median(ts[1:]) * 1e3)


FFT_KIDS = ("zinv", "yinv", "xrho", "xvpsi", "yfwd", "zfwd")
AMU = 1822.888486                  # electron masses per atomic mass unit
ION_MASS = 12.0 * AMU             # the synthetic species (Z = 6)


def pmc_copy(mib: int, repeat: int = 3) -> int:
    """y = 1 x + 0 y over whole wavefunction columns (mc_wfn_axpby): reads and
    writes about `mib` MiB per call, for scaling the PMC byte counters."""
    from cpmd_amd import ops
    from cpmd_amd import system as S
    sysm = S.make_system(12.0, 8, 40.0, nstate=16)
    ctx = ops.Context(sysm.cell, sysm.gset)
    per_state = 16 * ctx.ld
    nstate = max(1, (mib << 20) // per_state)
    x = ctx.alloc_wfn(nstate)
    y = ctx.alloc_wfn(nstate)
    for _ in range(repeat):
        ctx.axpby_wfn(1.0, x, 0.0, y, nstate)
    ctx.synchronize()
    print(f"copy: {repeat} x {nstate * per_state} bytes read and written by mc::k_axpby", flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="cfg4")
    ap.add_argument("--nstate", type=int, default=0)
    ap.add_argument("--batch-rho", type=int, default=128)
    ap.add_argument("--batch-vpsi", type=int, default=64)
    ap.add_argument("--step", choices=("cp", "wfopt"), default="cp")
    ap.add_argument("--dt", type=float, default=0.0,
                    help="time step (a.u.; cp default 4 from the optimised start, 1 from --prep sd; wfopt 0.02)")
    ap.add_argument("--emass", type=float, default=400.0)
    ap.add_argument("--prep-steps", type=int, default=30,
                    help="untimed FP64 steepest-descent steps before the CP paths (the CP start state)")
    ap.add_argument("--prep", choices=("sd", "opt"), default="opt",
                    help="CP start: sd = --prep-steps plain steepest-descent steps; opt = the wavefunction "
                         "optimisation (mixed then FP64) to gemax < --prep-tol")
    ap.add_argument("--prep-tol", type=float, default=1e-5)
    ap.add_argument("--paths", default="mixed,fp64", help="comma list of paths to time (first = headline)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--opt-policy", action="store_true",
                    help="also optimise the wavefunctions from the seeded start (preconditioned steepest "
                         "descent) mixed-then-FP64 and FP64 only, and report iterations, times and energies")
    ap.add_argument("--opt-tol", type=float, default=1e-5)
    ap.add_argument("--opt-method", choices=("sd", "diis"), default="diis",
                    help="optimiser of --prep opt and --opt-policy (diis: CPMD's ODIIS, hthrs 0.5)")
    ap.add_argument("--opt-dt", type=float, default=10.0,
                    help="optimisation time step (CPMD TIMESTEP; hthrs 2); its bound from the Hartree "
                         "curvature at the smallest G falls as 1/a: 10 for a = 24 bohr, 7 for a = 32")
    ap.add_argument("--opt-maxiter", type=int, default=4000)
    ap.add_argument("--pmc-copy", type=int, default=0, metavar="MIB",
                    help="only run a copy of MIB MiB (read and written) three times through mc::k_axpby: "
                         "the known-size kernel that scales rocprofv3's FETCH_SIZE / WRITE_SIZE")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: the one-GPU build runs the whole workload on a "
              "single MI355X; multi-GPU sharding is not launched by this benchmark.",
              file=sys.stderr)
        sys.exit(2)

    from cpmd_amd import ops
    from cpmd_amd import system as S
    from cpmd_amd.ks import KSDevice

    if args.pmc_copy:
        return pmc_copy(args.pmc_copy)
    cfg = S.CONFIGS[args.config]
    nstate = args.nstate or cfg["nstate"]
    # CP from the optimised ground state runs at dt = 4 a.u.; from a
    # non-converged start (--prep sd) the per-step move dt^2/emass f eps c0
    # along each state must stay small for the constraint iteration (rortho)
    # to converge, and dt = 1 a.u. keeps it at a few 1e-2
    if args.step == "cp":
        dt = args.dt or (4.0 if args.prep == "opt" else 1.0)
    else:
        dt = args.dt or 0.02
    log(f"[bench] building {args.config}: a={cfg['a']} nat={cfg['nat']} ecut={cfg['ecut']} nstate={nstate}")
    sysm = S.make_system(cfg["a"], cfg["nat"], cfg["ecut"], nstate=nstate, with_density=True)
    c0 = S.initial_coefficients(sysm)
    ngw = sysm.gset.ngw
    n1, n2, n3 = sysm.cell.nr
    ks = KSDevice(sysm)
    ctx = ks.ctx
    ctx.set_batch(args.batch_rho, args.batch_vpsi)
    sticks, cols, _ = ops.layout_stats(sysm.cell, sysm.gset)
    geom = (ngw, sticks, cols, n1, n2, n3)
    log(f"[bench] mesh {sysm.cell.nr} ngw {ngw} nhg {sysm.gset.nhg} sticks {sticks} columns {cols}")
    buf = {k: ctx.alloc_wfn(nstate) for k in ("c0", "cm", "cp", "c2")}
    f = sysm.f
    nat = sysm.nat
    npairs = (nstate + 1) // 2
    K = 2.0 * ngw
    n = float(nstate)
    if args.step == "cp":
        tn_gemm_flops = (4 * 2 * nat * n * K      # rnlsm1 + rnlsm2 (3 components): the A^T B products
                         + 2 * n * n * K          # rortog: <c0|cp>
                         + n * n * K              # rortog: <cp|cp> (symmetric)
                         + 2 * n * n * K)         # rortv: <cm|c0>
        gemm_flops = (tn_gemm_flops
                      + 2 * K * n * nat          # fnonloc
                      + 2 * K * n * n            # rortog: cp += c0 X
                      + 2 * K * n * n)           # rortv: cm += c0 gamy
        step_desc = ("CPMD's velocity Verlet CP step: velupi + velupa + posupi + phfac + posupa (rortog), "
                     "forces (rnlsm1 + rnlsm2 + rhoofr + vofrho(+potfor) + Ewald + vpsi + fnonloc + rnlfor), "
                     "velupi + velupa + rortv + rekine")
    else:
        gemm_flops = (2 * nat * n * K + 2 * K * n * nat + 2 * n * n * K + 2 * K * n * n + n * n * K + 2 * K * n * n)
        tn_gemm_flops = 2 * nat * n * K + 2 * n * n * K + n * n * K
        step_desc = "rnlsm1 + rhoofr + vofrho + vpsi + fnonloc + Lagrange projection + SD update + Loewdin"

    def opt_progress(it, mode, etot, gemax):
        if it % 250 == 0:
            log(f"[bench]   optimisation iteration {it} ({mode}): E_KS {etot:.10f} gemax {gemax:.3e}")

    start_desc = "the seeded coefficients"
    if args.step == "cp" and args.prep_steps > 0:
        # CP dynamics starts from a (partially) optimised state: a few FP64
        # steepest-descent steps (Lagrange projection, dt 0.02, Loewdin) bring
        # the band energies of the random start down so that the orthonormality
        # constraint of the first CP steps converges
        ctx.set_precision("fp64")
        ctx.upload(c0, buf["c0"])
        ks.set_ions(sysm.tau)
        t0 = time.perf_counter()
        for _ in range(args.prep_steps):
            ks.energy(buf["c0"], f, gradient=buf["c2"])
            ctx.lagrange_dev(buf["c0"], buf["c2"], f, nstate)
            ctx.axpy_wfn(0.02, buf["c2"], buf["c0"], nstate)
            ctx.lowdin_dev(buf["c0"], nstate)
        e_prep = ks.energy(buf["c0"], f)["etot"]
        c0 = ctx.download(buf["c0"], nstate)
        log(f"[bench] {args.prep_steps} FP64 optimisation steps ({time.perf_counter() - t0:.1f} s): "
            f"E_KS {e_prep:.8f} Ha")
        start_desc = f"{args.prep_steps} FP64 steepest-descent steps from the seeded coefficients"

    c_sd = c0
    if args.step == "cp" and args.prep == "opt":
        ks.set_ions(sysm.tau)
        ctx.upload(c0, buf["c0"])
        t0 = time.perf_counter()
        r = ks.optimize(buf["c0"], f, buf["c2"], tol=args.prep_tol, dt=args.opt_dt, emass=args.emass,
                        maxiter=min(args.opt_maxiter, 1500), callback=opt_progress, method=args.opt_method)
        c0 = ctx.download(buf["c0"], nstate)
        if r["converged"]:
            start_desc = (f"{start_desc}, then optimised ({args.opt_method}) to gemax < {args.prep_tol} "
                          f"({r['iters_mixed']} mixed + {r['iters_fp64']} FP64 iterations)")
        else:
            log(f"[bench] the start optimisation stopped at gemax {r['gemax']:.2e} (> {args.prep_tol}); "
                f"the CP paths start from that state")
            start_desc = (f"{start_desc}, then {r['iters_mixed'] + r['iters_fp64']} {args.opt_method} "
                          f"optimisation iterations (gemax {r['gemax']:.2e})")
        log(f"[bench] start optimisation: {r['iters_mixed']} mixed + {r['iters_fp64']} fp64 iterations "
            f"({time.perf_counter() - t0:.1f} s): E_KS {r['etot']:.8f} Ha")

    TERMS = ("ekin", "ehep", "epseu", "exc", "enl", "esr", "eself")

    def fp64_eval(tau):
        ctx.set_precision("fp64")
        return ks.energy(buf["c0"], f, tau=tau, gradient=buf["c2"], forces=True)

    # fp32fft: FP32 transforms, FP64 contractions; bf16x3: every split product 3-term
    modes = {"mixed": "mixed", "fp64": "fp64", "fp32fft": "fp32", "bf16x3": "mixed_bf16x3"}

    def run_path(path):
        ctx.upload(c0, buf["c0"])                 # every path starts from the same coefficients, at rest
        ctx.axpby_wfn(0.0, buf["c0"], 0.0, buf["cm"], nstate)
        iters = []
        eham = []
        if args.step == "cp":
            # the start in FP64: E_ham = E_KS at rest
            ctx.set_precision("fp64")
            e0_fp64 = ks.cp_init(buf["c0"], buf["cm"], buf["c2"], f, sysm.tau, emass=args.emass,
                                 ion_mass=ION_MASS)["eham"]
            ctx.set_precision(modes[path])
            st = ks.cp_init(buf["c0"], buf["cm"], buf["c2"], f, sysm.tau, emass=args.emass, ion_mass=ION_MASS)
            start = st["terms"]
        else:
            tau = sysm.tau.copy()
            ctx.set_precision(modes[path])
            start = ks.energy(buf["c0"], f, tau=tau, gradient=buf["c2"], forces=True)

        def step_cp():
            # CPMD's velocity Verlet CP step (src/mdmain_utils.mod.F90:710-966)
            ks.cp_step(st, buf, f, dt, epsog=1e-8, maxit=60)
            iters.append(st["iters"])
            eham.append(st["eham"])

        def step_wfopt():
            ks.energy(buf["c0"], f, tau=tau, gradient=buf["c2"])
            ctx.lagrange_dev(buf["c0"], buf["c2"], f, nstate)
            ctx.axpy_wfn(dt, buf["c2"], buf["c0"], nstate)
            iters.append(ctx.lowdin_dev(buf["c0"], nstate))

        step = step_cp if args.step == "cp" else step_wfopt
        all_kids = FFT_KIDS + ("misc", "gemm", "gemm_tn", "tn_planes", "vofrho")

        def read_timers(nsteps):
            out = {}
            for kid in all_kids:
                tot, cnt = ctx.timer_get(kid)
                if cnt:
                    out[kid] = {"total_ms_per_step": tot / nsteps, "launches_per_step": cnt / nsteps,
                                "avg_ms": tot / cnt}
            return out

        # Per-kernel breakdown: every kernel timed (two event records per
        # launch) on the last warmup step; the timed steps then carry events
        # around the dominant kernels only (the A^T B contraction kernel of
        # the path and the largest transform kernel), whose launch averages
        # the rooflines use.
        kernels = None
        for i in range(args.warmup):
            last = i == args.warmup - 1
            if last:
                ctx.synchronize()
                ctx.timer_reset()
                ctx.timer_select(None)
                ctx.timer_enable(True)
            step()
            if last:
                ctx.synchronize()
                ctx.timer_enable(False)
                kernels = read_timers(1)
        ctx.synchronize()
        log(f"[bench] {path}: {args.warmup} warmup steps done")
        if kernels:
            tn_kid = "tn_planes" if modes[path] in ("mixed", "mixed_bf16x3") else "gemm_tn"
            live = {k for k in (tn_kid, max((k for k in FFT_KIDS if k in kernels), default=None,
                                            key=lambda k: kernels[k]["total_ms_per_step"])) if k in kernels}
            ctx.timer_select(sorted(live))
        else:
            live = set(all_kids)
            ctx.timer_select(None)
        ctx.timer_reset()
        ctx.timer_enable(True)
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        ctx.synchronize()
        t1 = time.perf_counter()
        ctx.timer_enable(False)
        ctx.timer_select(None)
        ms = (t1 - t0) * 1e3 / args.steps
        timed = read_timers(args.steps)
        if kernels is None:
            kernels = timed
        for k in live:   # the live launch averages of the timed steps
            if k in timed:
                kernels[k] = dict(timed[k], source="timed steps")
        # charge sums of the path's rhoofr on the end state: csumr before the
        # FP32 rescale against the FP64 G-space csumg (CPMD's check is
        # |csumr - csumg| < 1e-6, rhoofr:625-635)
        ctx.rhoofr_dev(buf["c0"], f, nstate, ks.d_rho)
        cg, cr_raw = ctx.rhoofr_raw_charge()
        charge = {"csumg": cg, "csumr_before_rescale": cr_raw, "diff": cr_raw - cg, "rel": (cr_raw - cg) / cg}
        # the state after warmup + steps, evaluated in FP64
        ctx.set_precision("fp64")
        energy = None
        if args.step == "cp":
            # CPMD's constants of motion (md_driver): E_cons = E_KS + ekinp,
            # E_ham = E_cons + ekinc, with the full-step velocities of the
            # velocity Verlet step (ekinc from rekine after rortv), all in FP64
            st_end = ks.cp_init(buf["c0"], buf["cm"], buf["c2"], f, st["tau"], vion=st["vion"], emass=args.emass,
                                ion_mass=ION_MASS)
            end = st_end["terms"]
            energy = {"E_ham_start_Ha": e0_fp64, "E_KS_end_Ha": end["etot"], "ekinp_end_Ha": st_end["ekinp"],
                      "ekinc_end_Ha": st_end["ekinc"], "E_cons_end_Ha": st_end["econs"],
                      "E_ham_end_Ha": st_end["eham"],
                      "E_ham_drift_per_atom_Ha": (st_end["eham"] - e0_fp64) / nat,
                      "E_cons_drift_per_atom_Ha": (st_end["econs"] - e0_fp64) / nat,
                      "E_ham_path_per_step_Ha": eham,
                      "E_ham_max_excursion_per_atom_Ha": max(abs(e - e0_fp64) for e in eham) / nat,
                      "steps": args.warmup + args.steps}
            log(f"[bench] {path}: E_ham drift {energy['E_ham_drift_per_atom_Ha']:.3e} Ha/atom over "
                f"{energy['steps']} steps (ekinc {st_end['ekinc']:.3e}, ekinp {st_end['ekinp']:.3e}), largest "
                f"excursion {energy['E_ham_max_excursion_per_atom_Ha']:.3e} Ha/atom")
        else:
            end = fp64_eval(tau)
        orth = float(np.abs(ctx.ovlap_dev(buf["c0"], buf["c0"], nstate) - np.eye(nstate)).max())
        log(f"[bench] {path}: {args.steps} steps: {ms:.3f} ms/step, E_KS(start) {start['etot']:.10f}, "
            f"E_KS(end, fp64) {end['etot']:.10f}, |C^T C - I| {orth:.2e}, iters {iters[-args.steps:]}")
        return {"ms": ms, "kernels": kernels, "start": start, "end": end, "orth": orth,
                "iters": iters[-args.steps:], "energy": energy, "charge": charge}

    paths = [p.strip() for p in args.paths.split(",") if p.strip()]
    res = {p: run_path(p) for p in paths}
    head = paths[0]
    R = res[head]
    cw = 8.0 if head in ("mixed", "fp32fft", "bf16x3") else 16.0
    kernels = R["kernels"]
    lr = -(-npairs // min(args.batch_rho, npairs))
    lv = -(-npairs // min(args.batch_vpsi, npairs))

    def pairs_per_launch(k):
        if k == "xrho":
            return npairs / lr
        if k in ("xvpsi", "yfwd", "zfwd"):
            return npairs / lv
        # zinv / yinv: rhoofr's batches, and vpsi's unless it starts from
        # rhoofr's kept columns (column reuse: one launch per rhoofr batch)
        per_step = kernels[k]["launches_per_step"]
        return npairs / lr if round(per_step) == lr else 2 * npairs / (lr + lv)

    def launch_bytes(k):
        ppl = pairs_per_launch(k)
        return kernel_bytes(geom, k, 0, cw) + (kernel_bytes(geom, k, 1, cw) - kernel_bytes(geom, k, 0, cw)) * ppl

    def fft_kernel_roofline(k):
        b = launch_bytes(k)
        a = b / (kernels[k]["avg_ms"] * 1e-3) / 1e9
        return {"bound": "hbm", "kernel": k, "achieved": round(a, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(a / HBM_PEAK_GBS, 4), "bytes_per_launch": b, "avg_launch_ms": kernels[k]["avg_ms"]}

    def tn_kernel_roofline():
        """The A^T B contraction kernel (k_tn_planes in the split-bf16 paths,
        k_gemm_tn in FP64): the algorithm's MFMA flops per step (6 bf16
        products per FP64-equivalent product for the 3-part split, symmetric
        overlaps counted once) over its time per step."""
        split = head in ("mixed", "bf16x3")
        k = kernels["tn_planes" if split else "gemm_tn"]
        mult = (6.0 if head == "mixed" else 3.0) if split else 1.0
        flops = mult * tn_gemm_flops
        a = flops / (k["total_ms_per_step"] * 1e-3) / 1e12
        peak = 2500.0 if split else 78.6
        out = {"bound": "mfma", "kernel": "tn_planes" if split else "gemm_tn", "achieved": round(a, 2),
               "peak": peak, "unit": "TFLOP/s", "frac": round(a / peak, 4),
               "flops_per_launch": flops / k["launches_per_step"], "avg_launch_ms": k["avg_ms"],
               "launches_per_step": k["launches_per_step"],
               "dtype": "bf16 (3-part split)" if head == "mixed" else ("bf16 (2-part split)" if split else "f64"),
               "useful_fp64_equiv_tflops": round(tn_gemm_flops / (k["total_ms_per_step"] * 1e-3) / 1e12, 2)}
        # the rate the kernel's own MFMA instruction reaches on this GPU
        # (mc_probe_mfma: independent chains, no memory traffic), beside the
        # published dense peak
        probe = "bf16_16x16x32" if split else "f64_16x16x4"
        ceil = ops.probe_mfma(probe)
        out["issue_ceiling"] = {"instruction": "v_mfma_f32_16x16x32_bf16" if split else "v_mfma_f64_16x16x4_f64",
                                "tflops": round(ceil, 1), "frac": round(a / ceil, 4)}
        return out

    # the dominant kernel: the largest time per step among the transform
    # kernels and the A^T B contraction kernel
    tn_kid = "tn_planes" if head in ("mixed", "bf16x3") else "gemm_tn"
    cands = [k for k in FFT_KIDS if k in kernels] + ([tn_kid] if tn_kid in kernels else [])
    dom = max(cands, key=lambda k: kernels[k]["total_ms_per_step"])
    roof = tn_kernel_roofline() if dom == tn_kid else fft_kernel_roofline(dom)
    dom_fft = max((k for k in FFT_KIDS if k in kernels), key=lambda k: kernels[k]["total_ms_per_step"])
    step_bytes = step_model_bytes(geom, nstate, cw)
    fft_ms = sum(kernels[k]["total_ms_per_step"] for k in FFT_KIDS if k in kernels)
    # algorithmic bytes of the transform launches the step made
    run_bytes = sum(launch_bytes(k) * kernels[k]["launches_per_step"] for k in FFT_KIDS if k in kernels)

    def mfma_roofline(path, r):
        """FP64 path: FP64 MFMA flops / 78.6 TF.  Mixed paths: bf16 MFMA issue
        (6 products per FP64-equivalent product for the 3-part A^T B products,
        3 for the 2-part ones) against the 2.5 PF dense bf16 peak; 'useful'
        counts each product once."""
        ms = sum(r["kernels"].get(k, {}).get("total_ms_per_step", 0.0)
                 for k in ("gemm", "gemm_tn", "tn_planes")) or float("nan")
        useful = gemm_flops / (ms * 1e-3) / 1e12
        if path in ("mixed", "bf16x3"):
            # 3-part A^T B products issue 6 bf16 products per FP64 product, 2-part 3
            tn_flops = tn_gemm_flops if path == "mixed" else 0.0
            issued = 3 * gemm_flops + 3 * tn_flops
            tf = issued / (ms * 1e-3) / 1e12
            return {"bound": "mfma", "dtype": "bf16 split (3 or 6 products)", "flops_per_step": issued,
                    "gemm_ms_per_step": ms, "achieved": round(tf, 2), "peak": 2500.0, "unit": "TFLOP/s",
                    "frac": round(tf / 2500.0, 4), "useful_fp64_equiv_tflops": round(useful, 2)}
        return {"bound": "mfma", "dtype": "f64", "flops_per_step": gemm_flops, "gemm_ms_per_step": ms,
                "achieved": round(useful, 2), "peak": 78.6, "unit": "TFLOP/s", "frac": round(useful / 78.6, 4)}

    # HBM bytes per launch of the dominant kernel from the rocprofv3 --pmc
    # summary committed under profiles/ (FETCH_SIZE, WRITE_SIZE of the same
    # command, scaled by a copy kernel); a profile constant, not measured here
    traffic = None
    tf = ROOT / "profiles" / "traffic.json"
    if tf.exists():
        try:
            traffic = json.loads(tf.read_text()).get(f"{head}:{roof['kernel']}")
        except Exception:
            traffic = None
    roof["traffic"] = traffic
    roof["traffic_source"] = ("profiles/traffic.json (rocprofv3 --pmc FETCH_SIZE + WRITE_SIZE per dispatch, "
                              "copy-kernel scaled)") if traffic is not None else None

    cpu = None
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline (threaded stick-sparse CPU path) ...")
        cpu = cpu_baseline(sysm, c0, np.ascontiguousarray(ks.d_v.to_host(ks.nnr1)))
        ctx.upload(c0, buf["c0"])
        ctx.upload(c0, buf["cp"])
        cpu["gpu_fp64_same_content_ms"] = round(gpu_same_content_ms(ctx, ks, buf, f, nstate), 3)
        ctx.set_precision(modes[head])
        log(f"[bench] cpu baseline {cpu['value']:.1f} ms/step on {cpu['cores']} threads ({cpu['cpu']}): "
            f"{cpu['components_ms']}; the same content on the GPU in FP64 {cpu['gpu_fp64_same_content_ms']:.2f} ms")

    opt = None
    if args.opt_policy:
        # the precision policy of the wavefunction optimisation: FP32 transforms /
        # split-bf16 contractions while gemax >= 1e-4, FP64 after
        opt = {"tol": args.opt_tol, "switch": 1e-4, "dt": args.opt_dt, "emass": args.emass,
               "method": args.opt_method, "hthrs": 2.0 if args.opt_method == "sd" else 0.5,
               "start": "the CP start before optimisation (the steepest-descent prepared state)"}
        ks.set_ions(sysm.tau)
        for label, mixed in (("mixed_then_fp64", True), ("fp64", False)):
            ctx.upload(c_sd, buf["c0"])
            ctx.synchronize()
            t0 = time.perf_counter()
            r = ks.optimize(buf["c0"], f, buf["c2"], tol=args.opt_tol, switch=1e-4, mixed=mixed, dt=args.opt_dt,
                            emass=args.emass, maxiter=args.opt_maxiter, callback=opt_progress,
                            method=args.opt_method)
            ctx.synchronize()
            el = time.perf_counter() - t0
            h = r["history"]
            k = next((i for i, x in enumerate(h) if x[2] < 1e-4), None)
            opt[label] = {"converged": r["converged"], "iters_mixed": r["iters_mixed"],
                          "iters_fp64": r["iters_fp64"], "seconds": round(el, 3),
                          "ms_per_iter": round(el * 1e3 / max(len(h), 1), 3),
                          "first_iter_below_1e-4": k, "gemax": r["gemax"], "E_KS_Ha": r["etot"]}
            log(f"[bench] optimise {label}: converged {r['converged']} after {r['iters_mixed']} mixed + "
                f"{r['iters_fp64']} fp64 iterations, {el:.2f} s, E_KS {r['etot']:.10f}, gemax {r['gemax']:.2e}")
        if opt["fp64"]["converged"] and opt["mixed_then_fp64"]["converged"]:
            opt["dE_KS_per_atom_Ha"] = abs(opt["mixed_then_fp64"]["E_KS_Ha"] - opt["fp64"]["E_KS_Ha"]) / nat
        ctx.set_precision(modes[head])

    compare = {}
    if "fp64" in res and head != "fp64":
        F = res["fp64"]
        compare = {
            "fp64_ms_per_step": round(F["ms"], 4),
            "speedup_vs_fp64": round(F["ms"] / R["ms"], 3),
            # same C and R, the headline path's arithmetic against FP64
            "dE_KS_per_atom_Ha": abs(R["start"]["etot"] - F["start"]["etot"]) / nat,
            "dF_max_Ha_per_bohr": float(np.abs(R["start"]["fion"] - F["start"]["fion"]).max()),
            # after warmup + steps from the same start, both evaluated in FP64
            "dE_KS_after_steps_per_atom_Ha": abs(R["end"]["etot"] - F["end"]["etot"]) / nat,
            "dF_after_steps_max_Ha_per_bohr": float(np.abs(R["end"]["fion"] - F["end"]["fion"]).max()),
            "orthonormality_err": R["orth"],
            "orthonormality_err_fp64": F["orth"],
            "E_KS_fp64_Ha": F["start"]["etot"],
            "dE_terms_per_atom_Ha": {k: (R["start"][k] - F["start"][k]) / nat for k in TERMS},
        }
        if R["energy"] and F["energy"]:
            compare["E_ham_drift_per_atom_Ha"] = R["energy"]["E_ham_drift_per_atom_Ha"]
            compare["E_ham_drift_per_atom_Ha_fp64"] = F["energy"]["E_ham_drift_per_atom_Ha"]
            compare["dE_ham_end_per_atom_Ha"] = abs(R["energy"]["E_ham_end_Ha"] - F["energy"]["E_ham_end_Ha"]) / nat
        for p in paths:
            if p not in ("fp64", head):
                compare[f"dE_KS_per_atom_Ha_{p}"] = abs(res[p]["start"]["etot"] - F["start"]["etot"]) / nat
                compare[f"dF_max_Ha_per_bohr_{p}"] = float(np.abs(res[p]["start"]["fion"] - F["start"]["fion"]).max())
    out = {
        "metric": METRIC,
        "value": round(R["ms"], 4),
        "unit": "ms/step",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(R["ms"], 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 transforms, bf16-split MFMA contractions, f64 sums" if head == "mixed" else "f64",
        "data": "synthetic (seeded Gaussian wavefunctions, Loewdin-orthonormalised; seeded atom positions, "
                "analytic GTH pseudopotential)",
        "config": {
            "workload": f"{args.config}: {nstate} states, {n1}x{n2}x{n3} mesh, {ngw} Gamma plane waves, "
                        f"{sysm.gset.nhg} density G, {nat} atoms, a={cfg['a']} bohr, ecut {cfg['ecut']} Ry",
            "step": step_desc,
            "step_kind": args.step,
            "start": start_desc,
            "dt": dt,
            "emass": args.emass,
            "precision_path": head,
            "batch_pairs": {"rhoofr": args.batch_rho, "vpsi": args.batch_vpsi},
        },
        **compare,
        "roofline": roof,
        "fft_kernel_roofline": fft_kernel_roofline(dom_fft),
        "fft_stages": {k: fft_kernel_roofline(k) for k in FFT_KIDS if k in kernels},
        "fft_roofline": {"bytes_per_step": run_bytes, "ms_per_step": round(fft_ms, 4),
                         "achieved": round(run_bytes / (fft_ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(run_bytes / (fft_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                         "pair_model_bytes": step_bytes,
                         "pair_model_rate": round(step_bytes / (fft_ms * 1e-3) / 1e9, 1)},
        "mfma_roofline": mfma_roofline(head, R),
        "paths": {p: {"ms_per_step": round(res[p]["ms"], 4), "E_KS_start_Ha": res[p]["start"]["etot"],
                      "E_KS_end_fp64_Ha": res[p]["end"]["etot"], "orthonormality_err": res[p]["orth"],
                      "constraint_iters": res[p]["iters"], "cp_energy": res[p]["energy"],
                      "charge": res[p]["charge"],
                      "mfma_roofline": mfma_roofline(p, res[p]),
                      "kernels": res[p]["kernels"]} for p in paths},
        "cpu_baseline": cpu,
    }
    if opt is not None:
        out["optimize"] = opt
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    sys.exit(main() or 0)
