"""The CPU oracle as a backend of the case runners (test infrastructure).

The package's case runners (rheotool_amd.stressfoam.StressCase, the
rheoTestFoam protocol of rheotool_amd.rheotest) drive the device.  The
subclasses here run the same protocols through the oracle, so the tests can
compare the two on the same case directories.  Only tests import this."""
from __future__ import annotations

import numpy as np

from oracle import oracle as O
from rheotool_amd import foamdict, rheotest, stressfoam
from rheotool_amd.cases import Fields


def oracle_scheme(text):
    """The oracle's scheme name for a div(phi,theta) entry."""
    w = text.split()
    bounded = w[0] == "bounded"
    w = [x for x in w if x not in ("bounded", "GaussDefCmpw", "Gauss")]
    return ("bounded " if bounded else "") + w[0]


def oracle_params(e):
    """Oracle parameters and etaS of one model dictionary."""
    t = e.word("type")
    kw = dict(etaS=e.scalar("etaS"), etaP=e.scalar("etaP"), rho=e.scalar("rho", 1.0))
    for k, o in (("lambda", "lam"), ("alpha", "alpha"), ("L2", "L2"), ("epsilon", "epsilon"),
                 ("zeta", "zeta"), ("lambdaR", "lambdaR"), ("lambdaD", "lambdaD"), ("beta", "beta"),
                 ("delta", "delta"), ("chiMax", "chiMax"), ("lambdaS", "lambdaS"),
                 ("lambdaB", "lambdaB"), ("q", "q"), ("m", "m"), ("K", "K"), ("L", "Lc"),
                 ("a", "a"), ("b", "b")):
        if k in e and not isinstance(e[k], foamdict.FoamDict):
            kw[o] = e.scalar(k)
    if t == "BMPLog":  # BMPLog.C:113-119
        kw.update(G0=e.scalar("G0"), k_bmp=e.scalar("k"), Phi0=e.scalar("Phi0"), PhiInf=e.scalar("PhiInf"))
    elif t == "SaramitoLog":  # SaramitoLog.C:117-170
        kw.update(tau0=e.scalar("tau0"), sar_n=e.scalar("n"),
                  dims=tuple(float(x) for x in e["dims"][0]))
        kw.pop("K", None)
        if kw["sar_n"] == 1.0:
            kw["ptt_function"] = e.word("PTTfunction")
        else:
            kw["sar_k"] = e.scalar("k")
    elif "n" in e:
        kw["n" if t == "XPomPomLog" else "wm_n"] = e.scalar("n")
    if t == "Oldroyd-BRootk":  # Oldroyd_BRootk.C:121
        kw["root_k"] = e.scalar("k")
    if t == "PTTLog":
        # PTTLog.C:139 reads destructionFunctionType (no default)
        fn = e.word("destructionFunctionType")
        if fn not in ("linear", "exponential", "generalized"):
            raise NotImplementedError(f"PTT destructionFunctionType {fn}")
        kw["ptt_exp"] = fn == "exponential"
        if fn == "generalized":  # PTTLog.C:135-152
            kw["ptt_gen"] = (kw.pop("alpha"), kw.pop("beta"))
            kw["ml_rtol"] = e.scalar("rTolMittagLeffler", 1e-12)
            kw["ml_max_iter"] = int(e.scalar("maxIterMittagLeffler", 200))
    return O.params(t, **kw), kw.get("etaS", 0.0)


class OracleStressCase(stressfoam.StressCase):
    """StressCase through the oracle."""

    def _init_backend(self):
        self.O = O
        N = self.m.nCells
        for md in self.modes:
            md.oparams, md.etaS = oracle_params(md.params)
            md.ev = np.broadcast_to(np.eye(3), (N, 3, 3)).copy() if md.eigVals is None else md.eigVals
            md.vec = np.broadcast_to(np.eye(3), (N, 3, 3)).copy() if md.eigVecs is None else md.eigVecs
            bc = np.zeros(self.m.nPatches, np.int32)
            bcval = np.zeros((max(self.m.nBoundary, 1), 6))
            for patch, t, vals in md.bc:
                bc[patch] = {stressfoam.BC_SYM: O.BC_SYMMETRY, stressfoam.BC_WEDGE: O.BC_WEDGE}.get(t, t)
                if vals is not None:
                    s = self.m.patchStart[patch]
                    bcval[s:s + len(vals)] = vals
            md.obc, md.obcval = bc, bcval
            md.tau = np.zeros((N, 6)) if md.tau0 is None else md.tau0.copy()
            if md.bmp:
                fb = np.zeros(self.m.nPatches, np.int32)
                fv = np.zeros(max(self.m.nBoundary, 1))
                for patch, t, vals in md.fbc:
                    fb[patch] = t
                    if vals is not None:
                        s = self.m.patchStart[patch]
                        fv[s:s + len(vals)] = vals
                md.ofbc, md.ofbcval = fb, fv
        self.fields = Fields(U=self.U, Ub=self.Ub, phi=self.phi, gradU=None,
                             theta=None, eigVals=None, eigVecs=None)
        self.octl = O.controls()

    def _step_backend(self, dt):
        import dataclasses
        for md in self.modes:
            f = dataclasses.replace(self.fields, theta=md.theta, eigVals=md.ev, eigVecs=md.vec)
            if md.bmp:
                th, ev, vec, tau, _, md.Phi, _ = self.O.bmp_correct(
                    self.m, md.oparams, f, dt, md.Phi, md.tau, ctl=self.octl, bc=md.obc, bcval=md.obcval,
                    relax=md.relax, scheme=oracle_scheme(md.scheme), bc_phi=md.ofbc,
                    bcval_phi=md.ofbcval, relax_phi=md.frelax, scheme_phi=oracle_scheme(md.fscheme),
                    ctl_phi=self.octl)
            else:
                th, ev, vec, tau, _ = self.O.correct(self.m, md.oparams, f, dt, ctl=self.octl,
                                                     bc=md.obc, bcval=md.obcval, relax=md.relax,
                                                     scheme=oracle_scheme(md.scheme), tau_old=md.tau)
            md.theta, md.ev, md.vec, md.tau = th, ev, vec, tau
        return 0

    def _mode_state(self, md):
        out = dict(theta=md.theta, tau=md.tau, eigVals=md.ev, eigVecs=md.vec)
        if md.bmp:
            out["Phi"] = md.Phi
        return out


class OracleRunner(rheotest.Runner):
    """The rheoTestFoam protocol through the oracle."""

    def _init_backend(self, th0):
        self.O = O
        d = foamdict.parse(self.case.model_text).sub("parameters")
        if d.word("type") == "multiMode":
            entries = [sub for _, sub in d["models"][0]]
        else:
            entries = [d]
        self.modes = []
        for e in entries:
            st = dict(theta=th0.copy(), ev=np.eye(3)[None].copy(), vec=np.eye(3)[None].copy(),
                      tau=np.zeros((1, 6)), old=None)
            p, etaS = oracle_params(e)
            self.modes.append((p, etaS, st))
        self.ctl = O.controls(tol=1e-10)

    def _set_flow(self):
        pass

    def step(self, dt, iteration=None):
        return self._step_oracle(dt, iteration)

    def extra_stress(self):
        return self._extra_oracle()

    def _step_oracle(self, dt, iteration):
        import dataclasses
        for p, _, st in self.modes:
            if iteration == 0:
                st["old"] = st["theta"].copy()
            f = dataclasses.replace(self.fields, theta=st["theta"], eigVals=st["ev"], eigVecs=st["vec"])
            th, ev, vec, tau, _ = self.O.correct(self.m, p, f, dt, ctl=self.ctl, relax=self.case.relax,
                                                 theta_old=st["old"] if iteration else None)
            st.update(theta=th, ev=ev, vec=vec, tau=tau)

    def _extra_oracle(self):
        tau = sum(st["tau"] for _, _, st in self.modes)
        etaS = sum(e for _, e, _ in self.modes)
        return self.O.tau_total(self.m, self.fields.U, self.fields.Ub, tau, etaS)[0]


def rheotest_run(case_path, **kw):
    return rheotest.run(case_path, runner=OracleRunner, **kw)
