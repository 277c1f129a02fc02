#!/usr/bin/env python3
"""Benchmark of the plane-wave inner loop on one MI355X.

A step is one wavefunction-optimisation step of the 512-state workload on the
192^3 mesh (198,204 Gamma plane waves, cubic a = 32 bohr, 256 atoms, ecut
80 Ry) with every input resident in HBM:

  rnlsm1 -> rhoofr -> vpsi -> fnonloc -> Lagrange projection
         -> steepest-descent update c0 += dt c2 -> Loewdin orthonormalisation

(CPMD: forces_driver rnlsm / rscpot(rhoofr) / vpsi / fnonloc / addup2, then
updwf + ortho).  The local potential V(r) is a fixed seeded input (vofrho is
not part of the measured path).  Wavefunctions: seeded complex Gaussian
~exp(-|G|^2/8), Loewdin-orthonormalised; one s projector per atom (GTH).

Two precision paths are timed from the same start, the first one being the
headline `value`:
  mixed  rhoofr/vpsi transforms and their workspaces in FP32 (everything
         else FP64);
  fp64   every operation in FP64.
The band energy sum_i f_i <c_i|H|c_i> of the starting coefficients is
evaluated on both paths and its difference per atom is reported.

Prints one JSON line (see the measurement contract in DESIGN.md).
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
        return np_pairs * (cw * stick + c * 6 * ngw)
    raise KeyError(kid)


def step_model_bytes(ctx_geom, nstate: int, cw: float = 16.0) -> float:
    """Per-step algorithmic bytes of rhoofr + vpsi (the survey's model with the
    transform workspaces at cw bytes per complex): per state pair
    vpsi = cw(4 S N + 4 N^2 Zw) + 16*6 NGW + 8 N^3, rhoofr = cw(2 S N + 2 N^2 Zw) + 16*2 NGW,
    plus rho written once (8 N^3)."""
    ngw, S, Zw, n1, n2, n3 = ctx_geom
    r = 8.0
    npairs = (nstate + 1) // 2
    vp = cw * (4 * S * n3 + 4 * n1 * n2 * Zw) + 16.0 * 6 * ngw + r * n1 * n2 * n3
    rh = cw * (2 * S * n3 + 2 * n1 * n2 * Zw) + 16.0 * 2 * ngw
    return npairs * (vp + rh) + 2 * r * n1 * n2 * n3


def cpu_baseline(sysm, c0, v, max_pairs: int = 1):
    """Oracle (single-thread C restatement of CPMD's rhoofr + vpsi) on a bounded
    sample of the same step: `max_pairs` state pairs, scaled to all pairs."""
    import oracle
    nstate = c0.shape[1]
    npairs = (nstate + 1) // 2
    ns = 2 * max_pairs
    L = oracle.Layout(sysm.cell, sysm.gset)
    cs = np.asfortranarray(c0[:, :ns])
    f = sysm.f[:ns]
    t0 = time.perf_counter()
    oracle.rhoofr(L, cs, f)
    oracle.vpsi(L, cs, np.zeros_like(cs), f, v)
    dt = time.perf_counter() - t0
    return {
        "value": dt * 1e3 * npairs / max_pairs,
        "unit": "ms/step",
        "cores": 1,
        "kind": "port",
        "sample": f"oracle rhoofr+vpsi on {max_pairs} of {npairs} state pairs ({dt:.1f} s), scaled x{npairs // max_pairs}",
    }


FFT_KIDS = ("zinv", "yinv", "xrho", "xvpsi", "yfwd", "zfwd")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="cfg4")
    ap.add_argument("--nstate", type=int, default=0)
    ap.add_argument("--batch-rho", type=int, default=16)
    ap.add_argument("--batch-vpsi", type=int, default=8)
    ap.add_argument("--dt", type=float, default=0.02)
    ap.add_argument("--paths", default="mixed,fp64", help="comma list of paths to time (first = headline)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: the one-GPU build runs the whole workload on a "
              "single MI355X; multi-GPU sharding is described in DESIGN.md, not launched.",
              file=sys.stderr)
        sys.exit(2)

    from cpmd_amd import ops
    from cpmd_amd import system as S

    cfg = S.CONFIGS[args.config]
    nstate = args.nstate or cfg["nstate"]
    log(f"[bench] building {args.config}: a={cfg['a']} nat={cfg['nat']} ecut={cfg['ecut']} nstate={nstate}")
    sysm = S.make_system(cfg["a"], cfg["nat"], cfg["ecut"], nstate=nstate)
    c0 = S.initial_coefficients(sysm)
    v = S.real_space_potential(sysm)
    ngw = sysm.gset.ngw
    n1, n2, n3 = sysm.cell.nr
    sticks, cols, _ = ops.layout_stats(sysm.cell, sysm.gset)
    geom = (ngw, sticks, cols, n1, n2, n3)
    log(f"[bench] mesh {sysm.cell.nr} ngw {ngw} sticks {sticks} columns {cols}")

    ctx = ops.Context(sysm.cell, sysm.gset)
    ctx.set_batch(args.batch_rho, args.batch_vpsi)
    na, nhxs, ngh, lang, hmat, twnl, eigr = S.projector_set(sysm, sysm.gset.miller[:ngw], sysm.gset.hg[:ngw])
    ctx.set_projectors(na, nhxs, ngh, lang, hmat, twnl)
    ctx.set_eigr(eigr)
    del eigr, twnl
    d_c0 = ctx.alloc_wfn(nstate)
    d_c2 = ctx.alloc_wfn(nstate)
    d_rho = ops.DeviceArray(8 * ctx.nnr1)
    d_v = ops.DeviceArray(8 * ctx.nnr1).from_host(v)
    f = sysm.f
    nat = int(np.sum(na))
    npairs = (nstate + 1) // 2
    K = 2.0 * ngw
    n = float(nstate)
    gemm_flops = (2 * nat * n * K            # rnlsm1
                  + 2 * K * n * nat          # fnonloc
                  + 2 * n * n * K            # Lagrange: ovlap(c2, c0)
                  + 2 * K * n * n            # Lagrange: rotate
                  + n * n * K                # Loewdin: S = C^T C (symmetric)
                  + 2 * K * n * n)           # Loewdin: C S^-1/2

    def band_energy():
        """E_band = sum_i f_i <c_i|T + V + V_NL|c_i> = -sum_i dotp(c0_i, c2_i) after vpsi + fnonloc."""
        ctx.rnlsm1_dev(d_c0, nstate)
        ctx.zero_wfn(d_c2, nstate)
        ctx.vpsi_dev(d_c0, d_c2, f, d_v, nstate)
        ctx.fnonloc_dev(d_c2, f, nstate)
        return -float(np.sum(ctx.dotp_wfn(d_c0, d_c2, nstate)))

    def run_path(path):
        ctx.set_precision("mixed" if path == "mixed" else "fp64")
        ctx.upload(c0, d_c0)                 # every path starts from the same coefficients
        e_band = band_energy()
        energies = {}
        lowdin_iters = []

        def step():
            ctx.rnlsm1_dev(d_c0, nstate)
            ek, cg, cr = ctx.rhoofr_dev(d_c0, f, nstate, d_rho)
            energies.update(ekin=ek, csumg=cg, csumr=cr)
            ctx.zero_wfn(d_c2, nstate)
            ctx.vpsi_dev(d_c0, d_c2, f, d_v, nstate)
            ctx.fnonloc_dev(d_c2, f, nstate)
            ctx.lagrange_dev(d_c0, d_c2, f, nstate)
            ctx.axpy_wfn(args.dt, d_c2, d_c0, nstate)
            lowdin_iters.append(ctx.lowdin_dev(d_c0, nstate))

        for i in range(args.warmup):
            step()
        ctx.synchronize()
        log(f"[bench] {path}: {args.warmup} warmup steps done")
        ctx.timer_reset()
        ctx.timer_enable(True)
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        ctx.synchronize()
        t1 = time.perf_counter()
        ctx.timer_enable(False)
        ms = (t1 - t0) * 1e3 / args.steps
        kernels = {}
        for kid in FFT_KIDS + ("misc", "gemm"):
            tot, cnt = ctx.timer_get(kid)
            if cnt:
                kernels[kid] = {"total_ms_per_step": tot / args.steps, "launches": cnt, "avg_ms": tot / cnt}
        # after warmup + steps updates: E_band of the final C and ||C^T C - I||, both in FP64
        ctx.set_precision("fp64")
        e_end = band_energy()
        orth = float(np.abs(ctx.ovlap_dev(d_c0, d_c0, nstate) - np.eye(nstate)).max())
        log(f"[bench] {path}: {args.steps} steps: {ms:.3f} ms/step, E_band(start) {e_band:.10f}, "
            f"E_band(end, fp64) {e_end:.10f}, |C^T C - I| {orth:.2e}, "
            f"csumr-csumg {energies['csumr'] - energies['csumg']:.2e}")
        return {"ms": ms, "kernels": kernels, "e_band": e_band, "e_end": e_end, "orth": orth,
                "csum_diff": energies["csumr"] - energies["csumg"], "lowdin_iters": lowdin_iters[-args.steps:]}

    paths = [p.strip() for p in args.paths.split(",") if p.strip()]
    res = {p: run_path(p) for p in paths}
    head = paths[0]
    R = res[head]
    cw = 8.0 if head == "mixed" else 16.0
    kernels = R["kernels"]
    # dominant transform kernel by time per step
    dom = max((k for k in kernels if k in FFT_KIDS), key=lambda k: kernels[k]["total_ms_per_step"])
    # pairs per launch of the dominant kernel (average over the step)
    lr = -(-npairs // min(args.batch_rho, npairs))
    lv = -(-npairs // min(args.batch_vpsi, npairs))
    if dom == "xrho":
        ppl = npairs / lr
    elif dom in ("xvpsi", "yfwd", "zfwd"):
        ppl = npairs / lv
    else:   # zinv / yinv run in both rhoofr and vpsi batches
        ppl = 2 * npairs / (lr + lv)
    b_launch = kernel_bytes(geom, dom, 0, cw) + (kernel_bytes(geom, dom, 1, cw) - kernel_bytes(geom, dom, 0, cw)) * ppl
    achieved = b_launch / (kernels[dom]["avg_ms"] * 1e-3) / 1e9
    step_bytes = step_model_bytes(geom, nstate, cw)
    fft_ms = sum(kernels[k]["total_ms_per_step"] for k in FFT_KIDS if k in kernels)

    def mfma_roofline(path, r):
        """FP64 path: FP64 MFMA flops / 78.6 TF.  Mixed path: bf16 MFMA issue
        (3 products per FP64-equivalent product, 4 on the symmetric overlap)
        against the 2.5 PF dense bf16 peak; 'useful' counts each product once."""
        ms = r["kernels"].get("gemm", {}).get("total_ms_per_step", float("nan"))
        useful = gemm_flops / (ms * 1e-3) / 1e12
        if path == "mixed":
            issued = 3 * gemm_flops + n * n * K      # + the lo*lo term of S = C^T C
            tf = issued / (ms * 1e-3) / 1e12
            return {"bound": "mfma", "dtype": "bf16 (3-term split)", "flops_per_step": issued,
                    "gemm_ms_per_step": ms, "achieved": round(tf, 2), "peak": 2500.0, "unit": "TFLOP/s",
                    "frac": round(tf / 2500.0, 4), "useful_fp64_equiv_tflops": round(useful, 2)}
        return {"bound": "mfma", "dtype": "f64", "flops_per_step": gemm_flops, "gemm_ms_per_step": ms,
                "achieved": round(useful, 2), "peak": 78.6, "unit": "TFLOP/s", "frac": round(useful / 78.6, 4)}

    traffic = None
    tf = ROOT / "profiles" / "traffic.json"
    if tf.exists():
        try:
            traffic = json.loads(tf.read_text()).get(f"{head}:{dom}")
        except Exception:
            traffic = None

    cpu = None
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline (oracle, 1 pair sample) ...")
        cpu = cpu_baseline(sysm, c0, v)
        log(f"[bench] cpu baseline {cpu['value']:.1f} ms/step")

    compare = {}
    if "fp64" in res and head != "fp64":
        compare = {
            "fp64_ms_per_step": round(res["fp64"]["ms"], 4),
            "speedup_vs_fp64": round(res["fp64"]["ms"] / R["ms"], 3),
            "dE_band_per_atom_Ha": abs(R["e_band"] - res["fp64"]["e_band"]) / nat,
            "E_band_fp64_Ha": res["fp64"]["e_band"],
            "dE_band_after_steps_per_atom_Ha": abs(R["e_end"] - res["fp64"]["e_end"]) / nat,
            "orthonormality_err": R["orth"],
            "orthonormality_err_fp64": res["fp64"]["orth"],
        }
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
        "data": "synthetic (seeded Gaussian wavefunctions, Loewdin-orthonormalised; seeded smooth V(r))",
        "config": {
            "workload": f"{args.config}: {nstate} states, {n1}x{n2}x{n3} mesh, {ngw} Gamma plane waves, "
                        f"{nat} atoms, a={cfg['a']} bohr, ecut {cfg['ecut']} Ry",
            "step": "rnlsm1 + rhoofr + vpsi + fnonloc + Lagrange projection + SD update + Loewdin",
            "precision_path": head,
            "batch_pairs": {"rhoofr": args.batch_rho, "vpsi": args.batch_vpsi},
        },
        **compare,
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "bytes_per_launch": b_launch,
            "avg_launch_ms": kernels[dom]["avg_ms"],
        },
        "fft_roofline": {"model_bytes": step_bytes, "ms_per_step": round(fft_ms, 4),
                         "achieved": round(step_bytes / (fft_ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(step_bytes / (fft_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        "mfma_roofline": mfma_roofline(head, R),
        "paths": {p: {"ms_per_step": round(res[p]["ms"], 4), "E_band_Ha": res[p]["e_band"],
                      "E_band_after_steps_Ha": res[p]["e_end"], "orthonormality_err": res[p]["orth"],
                      "csumr_minus_csumg": res[p]["csum_diff"],
                      "lowdin_newton_schulz_iters": res[p]["lowdin_iters"],
                      "mfma_roofline": mfma_roofline(p, res[p]),
                      "kernels": res[p]["kernels"]} for p in paths},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
