"""The stress step on an OpenFOAM case with the velocity held fixed.

What rheoFoam does for the stress each time step (solvers/rheoFoam/rheoFoam.C:153
`constEq.correct()`, the fields of createFields.H) is run here, through the device
path, on a case directory:
- the mesh comes from constant/polyMesh (ASCII), or from system/blockMeshDict
  when there is none (the blockMesh subset of polymesh.py);
- constant/constitutiveProperties holds the model (`parameters`, a *Log type or
  multiMode of *Log types);
- system/fvSchemes gives `div(phi,theta<mode>)` (regex keys allowed, as in
  OpenFOAM), system/fvSolution the theta solver entry and the relaxation
  factor, system/controlDict the time controls (startFrom, deltaT, endTime,
  writeControl/writeInterval, timePrecision, writeFormat ascii or binary);
- <start>/theta<mode> (MUST_READ) with zeroGradient, fixedValue and empty
  patches; eigVals<mode>/eigVecs<mode> when present (READ_IF_PRESENT);
  tau<mode> for SaramitoLog and BMPLog (MUST_READ: their sources read it);
  for BMPLog also Phi<mode> (MUST_READ) with its patches, `div(phi,Phi<mode>)`,
  the Phi solver and relaxation entries;
- U and, when present, phi from the velocity time directory (default: the
  start time).  Patch velocities: the patch's `value` entry, the face cell's
  value for zeroGradient, zero for noSlip.  Without a phi file the flux is
  interpolate(U) & Sf with linear weights (fvc::flux).
Every write time gets theta<mode>, tau<mode>, eigVals<mode>, eigVecs<mode>
(Phi<mode> for BMPLog, and the summed tau of multiMode), in OpenFOAM's ASCII
format.

    python -m rheotool_amd.stressfoam CASE [--U-time T] [--steps N] [--no-write]

The velocity is not solved: the momentum and pressure equations of rheoFoam are
outside this library.
"""
from __future__ import annotations

import argparse
import os
import sys

import numpy as np

from . import foamdict, polymesh

BC_ZG, BC_FV, BC_EMPTY, BC_SYM, BC_WEDGE = 0, 1, 2, 3, 4
SYMMETRY_TYPES = ("symmetryPlane", "symmetry")


def _bc_entry(t, vals):
    """The volField patch entry written back for a boundary condition code."""
    return {BC_EMPTY: ("empty", None), BC_ZG: ("zeroGradient", None), BC_SYM: ("symmetryPlane", None),
            BC_WEDGE: ("wedge", None)}.get(t, ("fixedValue", vals))


def wedge_face_t(sf):
    """faceT of a wedge patch from its face area vectors sf [n, 3] ([OF9]
    wedgePolyPatch::calcGeometry): the rotation taking the centre-plane normal
    (each component of the mean unit normal n taken to sign(n_i)(max(|n_i|,
    1/2) - 1/2), normalised) onto n ([OF9] transform.H rotationTensor)."""
    n = np.mean(sf / np.linalg.norm(sf, axis=1)[:, None], axis=0)
    cn = np.sign(n) * (np.maximum(np.abs(n), 0.5) - 0.5)
    if not np.any(cn):
        raise ValueError("wedge patch not aligned with a coordinate axis")
    cn = cn / np.linalg.norm(cn)
    if abs(cn.sum()) < 1 - 1e-15:
        raise ValueError("wedge patch not aligned with a coordinate axis")
    s = float(cn @ n)
    a = np.cross(cn, n)
    ma = float(a @ a)
    if np.sqrt(ma) < 1e-15:
        raise ValueError("wedge patch lies in a coordinate plane (no wedge angle)")
    return s * np.eye(3) + (1 - s) * np.outer(a, a) / ma + (np.outer(n, cn) - np.outer(cn, n))


def _time_name(t, prec=6):
    s = f"{t:.{prec}g}"
    return s


def _time_dirs(case):
    out = []
    for n in os.listdir(case):
        try:
            out.append((float(n), n))
        except ValueError:
            pass
    return sorted(out)


class Controls:
    def __init__(self, case):
        cd = foamdict.read(os.path.join(case, "system", "controlDict"))
        self.deltaT = cd.scalar("deltaT")
        self.endTime = cd.scalar("endTime")
        self.writeControl = cd.word("writeControl", "timeStep")
        self.writeInterval = cd.scalar("writeInterval", 1.0)
        self.timePrecision = int(cd.scalar("timePrecision", 6.0))
        self.binary = cd.word("writeFormat", "ascii") == "binary"
        start = cd.word("startFrom", "startTime")
        times = _time_dirs(case)
        if start == "latestTime" and times:
            self.startTime, self.startName = times[-1]
        elif start == "firstTime" and times:
            self.startTime, self.startName = times[0]
        else:
            self.startTime = cd.scalar("startTime", 0.0)
            self.startName = next((n for t, n in times if abs(t - self.startTime) < 1e-12),
                                  _time_name(self.startTime, self.timePrecision))


def _scheme_text(divs, field):
    ent = divs.lookup_regex(f"div(phi,{field})")
    if ent is None:
        ent = divs.get("default")
        if ent is None or ent == ["none"]:
            raise KeyError(f"keyword div(phi,{field}) is undefined in divSchemes")
    if isinstance(ent, foamdict.FoamDict):
        raise ValueError(f"div(phi,{field}): a scheme name is expected")
    return " ".join(str(x) for x in ent)


def _solver_text(fs, field):
    sv = fs.get("solvers")
    ent = sv.lookup_regex(field) if isinstance(sv, foamdict.FoamDict) else None
    if not isinstance(ent, foamdict.FoamDict):
        return None
    from .rheotest import _tok
    out = []
    for k, v in ent.items():
        if isinstance(v, foamdict.FoamDict):
            v = v.get("preconditioner", ["DILU"])
        out.append(f"{k} {' '.join(_tok(x) for x in v)};")
    return " ".join(out)


def _relax(fs, field):
    rf = fs.get("relaxationFactors")
    if not isinstance(rf, foamdict.FoamDict):
        return 0.0
    eq = rf.get("equations")
    src = eq if isinstance(eq, foamdict.FoamDict) else rf
    v = src.lookup_regex(field)
    if v is None or isinstance(v, foamdict.FoamDict):
        return 0.0
    return float(v[-1])


class Mode:
    """One constitutive equation of the case (the single model or one mode)."""

    def __init__(self, name, params):
        self.name, self.params = name, params
        self.field = "theta" + name


class StressCase:
    """A case directory's stress step on the device.  Subclasses may replace
    the backend (_init_backend, _step_backend, _mode_state)."""

    def __init__(self, path, u_time=None):
        self.path = path
        self.pm = polymesh.case_mesh(path)
        self.m = self.pm.to_mesh()
        self.ctl = Controls(path)
        cp = foamdict.read(os.path.join(path, "constant", "constitutiveProperties"))
        self.params = cp.sub("parameters")
        self.model_text = "parameters\n{\n" + self.params.text + "\n}\n"
        if self.params.word("type") == "multiMode":
            self.modes = [Mode(n, d) for n, d in self.params["models"][0]]
        else:
            self.modes = [Mode("", self.params)]
        for md in self.modes:
            t = md.params.word("type")
            if not t.endswith("Log"):
                raise NotImplementedError(f"model type {t}: only the log-conformation models run on "
                                          "the device path")
        divs = foamdict.read(os.path.join(path, "system", "fvSchemes")).sub("divSchemes")
        fs = foamdict.read(os.path.join(path, "system", "fvSolution"))
        for md in self.modes:
            md.scheme = _scheme_text(divs, md.field)
            md.solver = _solver_text(fs, md.field)
            md.relax = _relax(fs, md.field)
            md.bmp = md.params.word("type") == "BMPLog"
            if md.bmp:  # the fluidity equation's entries
                md.fscheme = _scheme_text(divs, "Phi" + md.name)
                md.fsolver = _solver_text(fs, "Phi" + md.name)
                md.frelax = _relax(fs, "Phi" + md.name)
        t0 = os.path.join(path, self.ctl.startName)
        for md in self.modes:
            fld = polymesh.read_field(os.path.join(t0, md.field), self.pm)
            md.theta = fld.internal.copy()
            md.bc = self._theta_bc(fld)
            md.eigVals = self._opt_tensor(t0, "eigVals" + md.name)
            md.eigVecs = self._opt_tensor(t0, "eigVecs" + md.name)
            # SaramitoLog reads its tau (MUST_READ, SaramitoLog.C:51-62) for the
            # plasticity factor of the first step
            md.tau0 = None
            tp = os.path.join(t0, "tau" + md.name)
            if md.params.word("type") in ("SaramitoLog", "BMPLog"):
                md.tau0 = polymesh.read_field(tp, self.pm).internal.copy()
            if md.bmp:
                pf = polymesh.read_field(os.path.join(t0, "Phi" + md.name), self.pm)
                md.Phi = pf.internal[:, 0].copy()
                # a scalar's symmetry value is its cell value (zeroGradient while the
                # patch flux is zero)
                md.fbc = [(p, BC_ZG if t in (BC_SYM, BC_WEDGE) else t, None if v is None else v[:, 0].copy())
                          for p, t, v in self._theta_bc(pf)]
        ut = os.path.join(path, u_time if u_time is not None else self.ctl.startName)
        self.U, self.Ub = self._velocity(ut)
        self.phi = self._flux(ut)
        self.time = self.ctl.startTime
        self._init_backend()

    # ----------------------------------------------------------------- input
    def _opt_tensor(self, tdir, name):
        p = os.path.join(tdir, name)
        if not os.path.isfile(p):
            return None
        return polymesh.read_field(p, self.pm).internal.reshape(-1, 3, 3).copy()

    def _theta_bc(self, fld):
        """[(patch, type, values or None)] of the theta field."""
        out = []
        for i, (name, ptype, start, n) in enumerate(self.pm.patches):
            t = fld.patch_type(name)
            if t == "empty":
                out.append((i, BC_EMPTY, None))
            elif t == "zeroGradient":
                out.append((i, BC_ZG, None))
            elif t == "fixedValue":
                out.append((i, BC_FV, fld.patch_values(name, n)))
            elif t in SYMMETRY_TYPES:
                out.append((i, BC_SYM, None))
            elif t == "wedge":
                out.append((i, BC_WEDGE, None))
            else:
                raise NotImplementedError(f"{fld.name} patch {name}: boundary type {t} is not supported "
                                          "(zeroGradient, fixedValue, symmetryPlane, symmetry, wedge, empty)")
        return out

    def _velocity(self, tdir):
        fld = polymesh.read_field(os.path.join(tdir, "U"), self.pm)
        U = fld.internal.copy()
        nI = self.pm.nInternal
        Ub = np.zeros((len(self.pm.owner) - nI, 3))
        for name, ptype, start, n in self.pm.patches:
            s = start - nI
            t = fld.patch_type(name)
            if t == "empty":
                continue
            vals = fld.patch_values(name, n)
            if vals is not None:
                Ub[s:s + n] = vals
            elif t == "noSlip":
                Ub[s:s + n] = 0.0
            elif t == "zeroGradient":
                Ub[s:s + n] = U[self.pm.owner[start:start + n]]
            elif t in SYMMETRY_TYPES:
                # basicSymmetry: (U_P + (I - 2 n n) U_P)/2, the cell value without
                # its normal component
                Uc = U[self.pm.owner[start:start + n]]
                sf = self.m.Sf[start:start + n]
                nh = sf / np.linalg.norm(sf, axis=1)[:, None]
                Ub[s:s + n] = (Uc + (Uc - 2.0 * nh * np.sum(nh * Uc, axis=1)[:, None])) / 2.0
            elif t == "wedge":
                # wedgeFvPatchField: transform(faceT, U_P)
                T = wedge_face_t(self.m.Sf[start:start + n])
                Ub[s:s + n] = U[self.pm.owner[start:start + n]] @ T.T
            else:
                raise NotImplementedError(f"U patch {name}: type {t} without a value entry")
        return U, Ub

    def _flux(self, tdir):
        m, nI = self.m, self.m.nInternal
        p = os.path.join(tdir, "phi")
        phi = np.zeros(nI + m.nBoundary)
        if os.path.isfile(p):
            fld = polymesh.read_surface_field(p, self.pm)
            phi[:nI] = fld[0]
            phi[nI:] = fld[1]
        else:
            o, n = m.owner[:nI], m.neighbour
            Uf = m.w[:, None] * (self.U[o] - self.U[n]) + self.U[n]
            phi[:nI] = np.einsum("fi,fi->f", Uf, m.Sf[:nI])
            phi[nI:] = np.einsum("fi,fi->f", self.Ub, m.Sf[nI:])
        for k in range(m.nPatches):
            if m.patchKind[k] == 2:  # empty patches carry no flux
                s = nI + m.patchStart[k]
                phi[s:s + m.patchSize[k]] = 0.0
        return phi

    # ---------------------------------------------------------------- device
    def _init_backend(self):
        from . import abi
        self.abi = abi
        self.ctx = abi.Context(0)
        self.dm = self.ctx.mesh(self.m)
        self.eq = self.ctx.constitutiveEq("", self.dm, self.model_text)
        units = self.eq.modes() or [self.eq]
        for md, u in zip(self.modes, units):
            md.unit = u
            u.set_state(md.theta, md.eigVals, md.eigVecs, md.tau0)
            if md.solver:
                u.set_solver(md.solver)
            u.set_relax(md.relax)
            u.set_scheme(md.scheme)
            for patch, t, vals in md.bc:
                u.set_bc(patch, t, vals)
            if md.bmp:
                fl = u.fluidity()
                md.funit = fl
                fl.set_state(md.Phi)
                if md.fsolver:
                    fl.set_solver(md.fsolver)
                fl.set_relax(md.frelax)
                fl.set_scheme(md.fscheme)
                for patch, t, vals in md.fbc:
                    fl.set_bc(patch, t, vals)
        self.eq.set_flow(self.U, self.Ub, self.phi)

    def _step_backend(self, dt):
        rc = self.eq.correct(dt)
        return rc

    # ------------------------------------------------------------------- run
    def step(self):
        dt = self.ctl.deltaT
        rc = self._step_backend(dt)
        self.time += dt
        return rc

    def _mode_state(self, md):
        u = md.unit
        out = dict(theta=u.theta(), tau=u.tau(), eigVals=u.eigVals(), eigVecs=u.eigVecs())
        if md.bmp:
            out["Phi"] = md.funit.Phi()
        return out

    def state(self):
        """{mode: dict(theta, tau, eigVals, eigVecs[, Phi])} and the summed tau."""
        out = {md.name: self._mode_state(md) for md in self.modes}
        tau = sum(s["tau"] for s in out.values())
        return out, tau

    def _write_field(self, *args, **kw):
        polymesh.write_field(*args, binary=self.ctl.binary, **kw)

    def write(self):
        tname = _time_name(self.time, self.ctl.timePrecision)
        tdir = os.path.join(self.path, tname)
        per, tau = self.state()
        pm = self.pm
        for md in self.modes:
            s = per[md.name]
            th_bc = {}
            for patch, t, vals in md.bc:
                name = pm.patches[patch][0]
                th_bc[name] = _bc_entry(t, vals)
            self._write_field(os.path.join(tdir, md.field), md.field, s["theta"], pm, th_bc, tname)
            calc = {n: ("empty", None) if th_bc[n][0] == "empty" else ("calculated", _face_vals(pm, s["tau"], n))
                    for n in th_bc}
            self._write_field(os.path.join(tdir, "tau" + md.name), "tau" + md.name, s["tau"], pm, calc,
                                 tname, dims="[1 -1 -2 0 0 0 0]")
            for fname in ("eigVals", "eigVecs"):
                v = s[fname].reshape(-1, 9)
                zg = {n: ("empty", None) if th_bc[n][0] == "empty" else ("zeroGradient", None) for n in th_bc}
                self._write_field(os.path.join(tdir, fname + md.name), fname + md.name, v, pm, zg, tname)
            if md.bmp:
                f_bc = {}
                for patch, t, vals in md.fbc:
                    f_bc[pm.patches[patch][0]] = _bc_entry(t, vals)
                self._write_field(os.path.join(tdir, "Phi" + md.name), "Phi" + md.name, s["Phi"], pm, f_bc,
                                     tname, dims="[-1 1 1 0 0 0 0]")
        if len(self.modes) > 1:
            calc = {p[0]: ("empty", None) if p[1] == "empty" else ("calculated", _face_vals(pm, tau, p[0]))
                    for p in pm.patches}
            self._write_field(os.path.join(tdir, "tau"), "tau", tau, pm, calc, tname,
                                 dims="[1 -1 -2 0 0 0 0]")
        return tdir

    def run(self, steps=None, write=True, log=None):
        n = int(round((self.ctl.endTime - self.ctl.startTime) / self.ctl.deltaT))
        if steps is not None:
            n = min(n, steps)
        written = []
        wc, wi = self.ctl.writeControl, self.ctl.writeInterval
        for k in range(1, n + 1):
            rc = self.step()
            if log:
                log(f"Time = {_time_name(self.time, self.ctl.timePrecision)}  correct rc {rc}")
            due = (k % max(1, int(round(wi))) == 0) if wc == "timeStep" else \
                abs((self.time - self.ctl.startTime) / wi - round((self.time - self.ctl.startTime) / wi)) < 1e-6
            if write and (due or k == n):
                written.append(self.write())
        return written


def _face_vals(pm, cellvals, patch):
    """Patch values of a calculated field: the face cells' values."""
    for name, _, start, n in pm.patches:
        if name == patch:
            return cellvals[pm.owner[start:start + n]]
    raise KeyError(patch)


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m rheotool_amd.stressfoam")
    ap.add_argument("case")
    ap.add_argument("--U-time", default=None, help="time directory holding U (and phi)")
    ap.add_argument("--steps", type=int, default=None, help="at most N time steps")
    ap.add_argument("--no-write", action="store_true")
    a = ap.parse_args(argv)
    sc = StressCase(a.case, u_time=a.U_time)
    out = sc.run(steps=a.steps, write=not a.no_write, log=lambda s: print(s, flush=True))
    for d in out:
        print(f"wrote {d}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
