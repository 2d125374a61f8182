"""rheoTestFoam on the device path: the constitutive equation of one cell
under a prescribed homogeneous velocity gradient.

Follows solvers/rheoTestFoam (rheoTestFoam.C, createFields.H, reStartOrEnd.H,
timeWrite.H):
- the case's constant/constitutiveProperties holds the model (`parameters`,
  single mode or multiMode) and `rheoTestFoamParameters` (the tensor gradU,
  the rates gammaEpsilonDotL, ramp, maxIter, relTol);
- one cell whose U patch values make the Gauss gradient equal gradU*rate
  (createFields.H:95-146);
- transient mode: every step calls correct() nInIter times (fvSolution
  SIMPLE.nInIter) and writes t and extraStress = tauTotal() (timeWrite.H);
- ramp mode: one correct() per step with deltaT = (lambdaMin+lambdaMax)/2/rate
  until extraStress changes by less than relTol (or maxIter steps), then a
  row `rate  extraStress  status  relError` and the next rate
  (createFields.H:208-229, rheoTestFoam.C:73-87, reStartOrEnd.H);
- the report goes to <case>/Report (or --out), in the reference's format.

One deviation: the ramp's automatic time step (lambdaMin+lambdaMax)/2/rate
suits the stress-form models, whose relaxation term is implicit; the log
models treat their source explicitly, and a step of many relaxation times
lands on a non-physical root (for FENE-CR one with tr C > L^2).  The ramp
step is therefore capped at dt_cap * lambdaMin (default 0.1).

Only the log-conformation models run on the device: a stress-form type name of
the case (e.g. "FENE-CR", "XPomPom") is mapped to its log variant, which has
the same steady states.

    python -m rheotool_amd.rheotest CASE [--out FILE] [--rates N]
"""
from __future__ import annotations

import argparse
import os
import re
import sys

import numpy as np

from . import foamdict

SMALL = 1e-15  # OpenFOAM's SMALL for double

LOG_NAMES = {"Oldroyd-B": "Oldroyd-BLog", "Giesekus": "GiesekusLog", "FENE-P": "FENE-PLog",
             "PTT": "PTTLog", "FENE-CR": "FENE-CRLog", "WhiteMetznerCY": "WhiteMetznerCYLog",
             "Rolie-Poly": "Rolie-PolyLog", "XPomPom": "XPomPomLog"}


def log_variant_text(text: str) -> str:
    """Maps `type <stress-form>;` entries to their log variants."""
    def rep(mo):
        t = mo.group(1)
        return f"type {LOG_NAMES.get(t, t)} ;"
    return re.sub(r"type\s+([^\s;]+)\s*;", rep, text)


def mag_symm(s):
    """mag of a symmTensor (xx xy xz yy yz zz): off-diagonals count twice."""
    return float(np.sqrt(s[0] ** 2 + 2 * s[1] ** 2 + 2 * s[2] ** 2 + s[3] ** 2 + 2 * s[4] ** 2 + s[5] ** 2))


class Case:
    def __init__(self, path):
        self.path = path
        cp = foamdict.read(os.path.join(path, "constant", "constitutiveProperties"))
        self.params = cp.sub("parameters")
        rt = cp.sub("rheoTestFoamParameters")
        self.gradU = np.array(rt["gradU"][0], dtype=float).reshape(3, 3)
        self.rates = [float(x) for x in rt["gammaEpsilonDotL"][0]]
        self.ramp = str(rt.word("ramp", "false")).lower() in ("true", "on", "yes", "1")
        self.maxIter = int(rt.scalar("maxIter", 5000.0))
        self.relTol = rt.scalar("relTol", 1e-8)
        cd = foamdict.read(os.path.join(path, "system", "controlDict"))
        self.deltaT = cd.scalar("deltaT")
        self.endTime = cd.scalar("endTime")
        fs = foamdict.read(os.path.join(path, "system", "fvSolution"))
        simple = fs.get("SIMPLE")
        self.nInIter = int(simple.scalar("nInIter", 0.0)) if isinstance(simple, foamdict.FoamDict) else 0
        self.solver = self._solver_text(fs)
        self.relax = 0.0
        rf = fs.get("relaxationFactors")
        if isinstance(rf, foamdict.FoamDict):
            eq = rf.get("equations")
            src = eq if isinstance(eq, foamdict.FoamDict) else rf
            for key in ("theta", "tau"):
                v = src.lookup_regex(key)
                if v is not None and not isinstance(v, foamdict.FoamDict):
                    self.relax = float(v[-1])
                    break
        self.model_text = "parameters\n{\n" + log_variant_text(self.params.text) + "\n}\n"
        # lambdaMin/Max of the ramp's time step (createFields.H:186-206)
        if self.params.word("type") == "multiMode":
            lams = [d.scalar("lambda") for _, d in self.params["models"][0]]
        else:
            lams = [self.params.scalar("lambda")]
        self.lambdaMin, self.lambdaMax = min(lams), max(lams)

    @staticmethod
    def _solver_text(fs):
        """solvers entry for theta (falling back to the case's tau entry)."""
        sv = fs.get("solvers")
        ent = None
        if isinstance(sv, foamdict.FoamDict):
            for name in ("theta", "thetaM1", "tau", "tauM1"):
                ent = sv.lookup_regex(name)
                if isinstance(ent, foamdict.FoamDict):
                    break
        if not isinstance(ent, foamdict.FoamDict):
            return None
        out = []
        for k, v in ent.items():
            if isinstance(v, foamdict.FoamDict):  # preconditioner { preconditioner DILU; }
                v = v.get("preconditioner", ["DILU"])
            out.append(f"{k} {' '.join(_tok(x) for x in v)};")
        return " ".join(out)


def _tok(x):
    """A dictionary token back as text (integral numbers without '.0')."""
    if isinstance(x, float):
        return str(int(x)) if x.is_integer() and abs(x) < 2 ** 53 else repr(x)
    return str(x)


def one_cell():
    from . import cases
    return cases.cartesian(1, 1, 1)


def patch_velocity(m, G):
    """U patch values of createFields.H:95-146: +-(G row of the face normal) * h/2."""
    Ub = np.zeros((m.nBoundary, 3))
    for b in range(m.nBoundary):
        f = m.nInternal + b
        n = m.Sf[f] / np.linalg.norm(m.Sf[f])
        ax = int(np.argmax(np.abs(n)))
        h = 2.0 * abs(m.Cf[f, ax] - m.C[m.owner[f], ax])
        Ub[b] = np.sign(n[ax]) * G[ax] * 0.5 * h
    return Ub


class Runner:
    """The device path on the one-cell case.  Subclasses may replace the
    backend (_init_backend, _set_flow, step, extra_stress)."""

    def __init__(self, case: Case):
        from . import cases
        self.case = case
        self.m = one_cell()
        th0 = np.zeros((1, 6))
        self.fields = cases.make_fields(self.m, cases.FIELDS_TG, seed=1)
        self.fields.theta[:] = 0.0
        I9 = np.eye(3)[None]
        self.fields.eigVals[:] = I9
        self.fields.eigVecs[:] = I9
        self.fields.U[:] = 0.0
        self._init_backend(th0)

    # ---------------------------------------------------------------- device
    def _init_backend(self, th0):
        from . import abi
        self.abi = abi
        self.ctx = abi.Context(0)
        self.dm = self.ctx.mesh(self.m)
        self.eq = self.ctx.constitutiveEq("", self.dm, self.case.model_text)
        self.units = self.eq.modes() or [self.eq]
        for u in self.units:
            u.set_state(th0)
        if self.case.solver:
            self.eq.set_solver(self.case.solver)
        self.eq.set_relax(self.case.relax)

    def set_rate(self, rate):
        G = self.case.gradU * rate
        self.fields.Ub[:] = patch_velocity(self.m, G)
        self.fields.phi[:] = 0.0
        nb = self.m.nInternal
        self.fields.phi[nb:] = np.einsum("bi,bi->b", self.fields.Ub, self.m.Sf[nb:])
        self._set_flow()

    def _set_flow(self):
        self.eq.set_flow(self.fields.U, self.fields.Ub, self.fields.phi)

    def step(self, dt, iteration=None):
        self.eq.correct(dt, iteration=iteration)

    def extra_stress(self):
        return self.eq.tau_total()[0]

def run(case_path, out=None, rates=None, max_steps=None, log=None, dt_cap=0.1, runner=Runner):
    case = Case(case_path)
    r = runner(case)
    lines = []
    if case.ramp:
        lines.append("******* Ramp mode: variable ε̇|γ̇ **********")
        lines.append("\t".join(["ε̇|γ̇", "extStressXX", "extStressXY", "extStressXZ",
                                "extStressYY", "extStressYZ", "extStressZZ", "Status", "Relative_Error"]))
        todo = case.rates[:rates] if rates else case.rates
        for rate in todo:
            r.set_rate(rate)
            dt = min((case.lambdaMin + case.lambdaMax) * 0.5 / rate, dt_cap * case.lambdaMin)
            old = np.zeros(6)
            cnt = 0
            while True:
                r.step(dt)
                es = r.extra_stress()
                rel = mag_symm(es - old) / (mag_symm(es) + SMALL)
                if rel < case.relTol or cnt > case.maxIter:
                    status = "Converged" if rel < 1e-8 else "Exceed_Niter"
                    lines.append("\t".join([repr(rate)] + [repr(float(x)) for x in es] + [status, repr(rel)]))
                    if log:
                        log(lines[-1])
                    break
                old = es
                cnt += 1
    else:
        rate = case.rates[0]
        lines.append(f"******* Transient mode: fixed ε̇|γ̇ = {rate!r} **********")
        lines.append("\t".join(["t", "extStressXX", "extStressXY", "extStressXZ", "extStressYY",
                                "extStressYZ", "extStressZZ"]))
        r.set_rate(rate)
        nsteps = int(round(case.endTime / case.deltaT))
        if max_steps:
            nsteps = min(nsteps, max_steps)
        for k in range(1, nsteps + 1):
            for i in range(case.nInIter):
                r.step(case.deltaT, iteration=i if case.nInIter > 1 else None)
            es = r.extra_stress()
            lines.append("\t".join([repr(k * case.deltaT)] + [repr(float(x)) for x in es]))
    text = "\n".join(lines) + "\n"
    if out is not None:
        with open(out, "w") as fh:
            fh.write(text)
    return text


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m rheotool_amd.rheotest")
    ap.add_argument("case")
    ap.add_argument("--out", default=None, help="report file (default <case>/Report)")
    ap.add_argument("--rates", type=int, default=0, help="ramp mode: only the first N rates")
    ap.add_argument("--steps", type=int, default=0, help="transient mode: at most N steps")
    ap.add_argument("--dt-cap", type=float, default=0.1, help="ramp step cap in units of lambdaMin")
    a = ap.parse_args(argv)
    out = a.out or os.path.join(a.case, "Report")
    run(a.case, out=out, rates=a.rates or None, max_steps=a.steps or None,
        log=lambda s: print(s, flush=True), dt_cap=a.dt_cap)
    print(f"report written to {out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
