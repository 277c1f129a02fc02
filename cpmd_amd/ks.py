"""Kohn-Sham energy, electronic gradient (H psi) and ionic forces on the GPU,
composed from the library's entry points in the order of CPMD's forces driver
for the Gamma, closed-shell, LDA, norm-conserving path
(src/forces_driver.mod.F90:175-384):

    phfac -> rnlsm1 (+ rnlsm2) -> rhoofr -> vofrho (+ potfor) -> Ewald
          -> vpsi + fnonloc (c2 = -f H c) -> rnlfor

E_KS = ekin + ehep + epseu + exc + esr - eself + enl (ener_com%etot).  The
wavefunctions stay on the device; only energies, fnl and forces come back.
KSDevice.optimize is the wavefunction optimisation at fixed ions with the
mixed-then-FP64 precision policy; KSDevice.cp_init / cp_step are CPMD's
velocity Verlet Car-Parrinello step (src/mdmain_utils.mod.F90:529-966).
"""
from __future__ import annotations

import numpy as np

from . import ops


class KSDevice:
    """One system on one GPU: layout, projectors, density set, local PP, ions."""

    def __init__(self, sysm, raggio: float = 1.2, iesr: int = 2, column_reuse: bool = True):
        from . import system as S
        gs = sysm.gset
        if gs.nhg == gs.ngw:
            raise ValueError("KSDevice needs the density G set (make_system(..., with_density=True))")
        self.sysm = sysm
        self.iesr = iesr
        ctx = ops.Context(sysm.cell, gs)
        ngw = gs.ngw
        na, nhxs, ngh, lang, hmat, twnl, _ = S.projector_set(sysm, gs.miller[:ngw], gs.hg[:ngw], with_eigr=False)
        ctx.set_projectors(na, nhxs, ngh, lang, hmat, twnl)
        ctx.set_gk((gs.miller[:ngw] @ sysm.cell.bvec).T)
        zv, rg, vps, rhops = S.local_pp_set(sysm, gs.hg, raggio=raggio)
        nh = np.array([n // 2 + 1 for n in sysm.cell.nr])
        ctx.set_density_gset(gs.miller + nh, gs.hg, gs.miller @ sysm.cell.bvec)
        ctx.set_local_pp(na, zv, rg, vps, rhops)
        # G=0 local potential (ppener's vploc = eivps(G=0), src/ppener_utils.mod.F90:58-60)
        self.vploc = float(np.dot(np.asarray(na, dtype=np.float64), vps[:, 0])) if gs.hg[0] == 0.0 else 0.0
        # energy() calls forces_c2_dev right after rhoofr_dev on the same
        # coefficients: vpsi starts from rhoofr's y-transformed columns
        ctx.set_column_reuse(column_reuse)
        self.ctx = ctx
        self.hmat = np.asarray(hmat, dtype=np.float64).reshape(nhxs, nhxs, -1, order="F")
        self.na = np.asarray(na)
        self.nhxs = nhxs
        self.ngh = np.asarray(ngh)
        self.nnr1 = ctx.nnr1
        self.d_rho = ops.DeviceArray(8 * self.nnr1)
        self.d_v = ops.DeviceArray(8 * self.nnr1)
        self._tau = None

    def set_ions(self, tau: np.ndarray):
        self.ctx.set_ions(tau, self.sysm.cell.avec)
        self.ctx.phfac()
        self._tau = np.array(tau, dtype=np.float64)

    def energy(self, d_c0: ops.DeviceArray, f: np.ndarray, tau: np.ndarray | None = None,
               gradient: ops.DeviceArray | None = None, forces: bool = False) -> dict:
        """E_KS terms of the device coefficients d_c0; with gradient=d_c2,
        d_c2 <- -f H c; with forces, 'fion' (3, nat)."""
        ctx = self.ctx
        f = np.ascontiguousarray(f, dtype=np.float64)
        n = len(f)
        if tau is not None and (self._tau is None or not np.array_equal(tau, self._tau)):
            self.set_ions(tau)
        ctx.rnlsm_dev(d_c0, n, tfor=forces)                 # fnl (and dfnl) stay on the device
        ekin, _, _ = ctx.rhoofr_dev(d_c0, f, n, self.d_rho)
        en, floc = ctx.vofrho_dev(self.d_rho, self.d_v, forces=forces)
        esr, eself, fesr = ctx.ewald(self.iesr, forces=forces)
        if gradient is not None:
            ctx.forces_c2_dev(d_c0, gradient, f, self.d_v, n)
        # E_NL after the gradient's launches: its read-back waits while they run
        out = dict(ekin=ekin, ehep=en["ehep"], epseu=en["epseu"], exc=en["exc"], esr=esr, eself=eself,
                   enl=ctx.rnlrh(f))
        out["etot"] = (out["ekin"] + out["ehep"] + out["epseu"] + out["exc"] + out["esr"] - out["eself"]
                       + out["enl"])
        if forces:
            out["fion"] = ctx.rnlfor(f, fion=floc + fesr)
        return out

    def optimize(self, d_c0: ops.DeviceArray, f: np.ndarray, d_c2: ops.DeviceArray, tol: float = 1e-5,
                 switch: float = 1e-4, mixed: bool = True, dt: float = 10.0, emass: float = 400.0,
                 hthrs: float | None = None, maxiter: int = 4000, callback=None, method: str = "diis",
                 mdiis: int = 10) -> dict:
        """Wavefunction optimisation at fixed ions: rwfopt -> updwf
        (src/updwf_utils.mod.F90:64-160).  Each iteration runs the forces
        driver (E_KS and c2 = -f H c), the Lagrange projection of c2 (lproj),
        csize, the update and Gram-Schmidt (ortho -> rgs).  The update is
        method "diis" (default), CPMD's default ODIIS over mdiis iterations
        with its progress reset at nreset = mdiis (mc_odiis_dev; hesele's
        vpp; empty states move by dt^2 / emass), or "sd", the preconditioned
        steepest descent c0 += vpp pf1 c2 (pf1 = dt^2 hthrs / emass).  hthrs
        defaults to CPMD's 0.5 for "diis" and to 2 for "sd".  The DIIS history
        is kept across the precision switch (its mixed-precision gradients
        are accurate far below the switch threshold).

        Precision policy: with mixed, the transforms run in FP32 and the
        contractions in split bf16 while the gradient size gemax is >= switch;
        from the first iteration below it everything runs in FP64 until
        gemax < tol (the convergence test of tol_chk_cnvgrad; 1e-5 is CPMD's
        default, src/control_test_utils.mod.F90:595-601).  mixed=False is
        the plain FP64 optimisation.  The returned etot is the energy of the
        final (converged) coefficients, evaluated in FP64.

        dt and hthrs are CPMD's TIMESTEP and HTHRS inputs.  For "sd" the step
        on a component with preconditioned curvature 1 is f pf1 (1 for f = 2,
        dt = 10, hthrs = 2: Newton-like at large G); below the clamp it is
        f dt^2 / emass * curvature, which bounds dt by the charge-sloshing
        (Hartree) curvature at the smallest G: dt = 20, hthrs = 0.5 diverges
        for a 24 bohr cell where dt = 10, hthrs = 2 converges.  callback, if
        given, is called after each gradient as callback(it, mode, etot, gemax)."""
        ctx = self.ctx
        f = np.ascontiguousarray(f, dtype=np.float64)
        n = len(f)
        if method not in ("sd", "diis"):
            raise ValueError("method must be 'sd' or 'diis'")
        if hthrs is None:
            hthrs = 2.0 if method == "sd" else 0.5      # 0.5: CPMD's default HTHRS
        ctx.hesele(True, 0.0, self.vploc, hthrs)
        pf1 = dt * dt * hthrs / emass
        reinit = True
        mode = "mixed" if mixed and switch > tol else "fp64"
        ctx.set_precision(mode)
        hist = []
        iters = {"mixed": 0, "fp64": 0}
        converged = False
        e, gemax, cnorm = None, float("inf"), float("inf")
        for _ in range(maxiter):
            e = self.energy(d_c0, f, gradient=d_c2)
            ctx.lagrange_dev(d_c0, d_c2, f, n)
            gemax, cnorm = ctx.csize_dev(d_c2, f, n)
            hist.append((mode, e["etot"], gemax, cnorm))
            iters[mode] += 1
            if callback is not None:
                callback(len(hist) - 1, mode, e["etot"], gemax)
            if not np.isfinite(gemax):
                break
            if mode == "fp64" and gemax < tol:
                converged = True
                break
            if method == "sd":
                ctx.sd_update_dev(d_c0, d_c2, n, pf1)
            else:
                ctx.odiis_dev(d_c0, d_c2, f, n, e["etot"], dt * dt / emass, mdiis=mdiis, reinit=reinit)
                reinit = False
            ctx.rgs_dev(d_c0, n)
            if mode == "mixed" and gemax < switch:
                mode = "fp64"
                ctx.set_precision(mode)
        if method == "diis":
            ctx.odiis_release()      # 2 mdiis n ldg complex doubles of history
        return dict(etot=e["etot"] if e else float("nan"), terms=e, gemax=gemax, cnorm=cnorm, converged=converged,
                    iters_mixed=iters["mixed"], iters_fp64=iters["fp64"], history=hist)

    # ------------------------------------------------------------------ CP MD
    def cp_init(self, d_c0: ops.DeviceArray, d_cm: ops.DeviceArray, d_c2: ops.DeviceArray, f: np.ndarray,
                tau: np.ndarray, vion: np.ndarray | None = None, emass: float = 400.0,
                ion_mass: float | np.ndarray = 1.0) -> dict:
        """Start of a CP run (src/mdmain_utils.mod.F90:529-643): forces at
        the start (forcedr: E_KS, c2, F_ion), the velocity constraint rortv on
        cm and ekinc (rekine).  Returns the state that cp_step advances:
        tau, vion (nat, 3), fion (3, nat), etot, ekinc, ekinp, eham."""
        f = np.ascontiguousarray(f, dtype=np.float64)
        n = len(f)
        tau = np.array(tau, dtype=np.float64)
        vion = np.zeros_like(tau) if vion is None else np.array(vion, dtype=np.float64)
        r = self.energy(d_c0, f, tau=tau, gradient=d_c2, forces=True)
        ekinc, _ = self.ctx.cp_velupa_rortv_dev(d_c0, None, d_cm, n, 0.0, emass)
        st = dict(tau=tau, vion=vion, fion=r["fion"], etot=r["etot"], terms=r, ekinc=ekinc, emass=emass,
                  mass=np.broadcast_to(np.asarray(ion_mass, dtype=np.float64), (tau.shape[0],))[:, None].copy())
        self._cp_constants(st)
        return st

    @staticmethod
    def _cp_constants(st: dict):
        # md_driver's constants of motion: econs = etot + ekinp, eham = econs + ekinc
        st["ekinp"] = 0.5 * float(np.sum(st["mass"] * st["vion"] * st["vion"]))
        st["econs"] = st["etot"] + st["ekinp"]
        st["eham"] = st["econs"] + st["ekinc"]

    def cp_step(self, st: dict, bufs: dict, f: np.ndarray, dt: float, epsog: float = 1e-8, maxit: int = 60) -> dict:
        """One Car-Parrinello step in CPMD's velocity Verlet order
        (src/mdmain_utils.mod.F90:710-966, no thermostats or constraints):

            velupi (vion += dt/2M F), velupa + posupi + phfac + posupa
            (cm += dt/2mu c2, tau += dt vion, cp = c0 + dt cm, rortog, cm),
            forcedr (E_KS, c2, F at the new positions),
            velupi, velupa + rortv + rekine (mc_cp_velupa_rortv_dev).

        bufs holds the device arrays 'c0', 'cm', 'cp', 'c2'; 'c0' and 'cp'
        are swapped (CPMD's final dcopy of posupa).  st is updated in place
        and returned; st['iters'] is the rortog iteration count."""
        ctx = self.ctx
        f = np.ascontiguousarray(f, dtype=np.float64)
        n = len(f)
        emass = st["emass"]
        dtm = 0.5 * dt / emass
        st["vion"] = st["vion"] + 0.5 * dt / st["mass"] * st["fion"].T          # velupi
        st["tau"] = st["tau"] + dt * st["vion"]                                  # posupi
        self.set_ions(st["tau"])                                                 # phfac
        _, it = ctx.posupa_dev(bufs["c0"], bufs["c2"], bufs["cm"], bufs["cp"], n, dtm, dt, epsog=epsog,
                               maxit=maxit)                                      # velupa + posupa
        bufs["c0"], bufs["cp"] = bufs["cp"], bufs["c0"]
        r = self.energy(bufs["c0"], f, tau=st["tau"], gradient=bufs["c2"], forces=True)   # forcedr
        st["fion"], st["etot"], st["terms"] = r["fion"], r["etot"], r
        st["vion"] = st["vion"] + 0.5 * dt / st["mass"] * st["fion"].T          # velupi
        st["ekinc"], _ = ctx.cp_velupa_rortv_dev(bufs["c0"], bufs["c2"], bufs["cm"], n, dtm, emass)
        st["iters"] = it
        self._cp_constants(st)
        return st

