"""ctypes wrapper of the CPU FP64 oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The product path never
imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

OLDROYDB, GIESEKUS, FENEP, PTT_LINEAR, PTT_EXP = 0, 1, 2, 3, 4
BC_ZEROGRADIENT, BC_FIXEDVALUE, BC_EMPTY, BC_COUPLED, BC_SYMMETRY, BC_WEDGE = 0, 1, 2, 3, 4, 5
PBICGSTAB, PBICG = 0, 1
PC_NONE, PC_DIAGONAL, PC_DILU = 0, 1, 2

FENECR, WMCY, ROLIEPOLY, XPP, SARAMITO, BMP, FLUIDITY = 5, 6, 7, 8, 9, 10, 11
ROOTK, PTT_GEN = 12, 13
MODEL_IDS = {"Oldroyd-BLog": OLDROYDB, "GiesekusLog": GIESEKUS, "FENE-PLog": FENEP,
             "PTTLog": PTT_LINEAR, "FENE-CRLog": FENECR, "WhiteMetznerCYLog": WMCY,
             "Rolie-PolyLog": ROLIEPOLY, "XPomPomLog": XPP, "SaramitoLog": SARAMITO, "BMPLog": BMP,
             "Oldroyd-BRootk": ROOTK}
SARAMITO_PTT = {"none": 0, "linear": 1, "exponential": 2}


class Mesh(C.Structure):
    _fields_ = [("nCells", C.c_int32), ("nInternal", C.c_int32), ("nBoundary", C.c_int32),
                ("nPatches", C.c_int32), ("owner", C.c_void_p), ("neighbour", C.c_void_p),
                ("Sf", C.c_void_p), ("w", C.c_void_p), ("V", C.c_void_p),
                ("patchStart", C.c_void_p), ("patchSize", C.c_void_p),
                ("patchEmpty", C.c_void_p)]


class Params(C.Structure):
    _fields_ = [("model", C.c_int32), ("rho", C.c_double), ("etaS", C.c_double),
                ("etaP", C.c_double), ("lambda_", C.c_double), ("alpha", C.c_double),
                ("L2", C.c_double), ("epsilon", C.c_double), ("zeta", C.c_double),
                ("ofInvQuirk", C.c_int32),
                ("lambdaR", C.c_double), ("lambdaD", C.c_double), ("beta", C.c_double),
                ("delta", C.c_double), ("chiMax", C.c_double),
                ("lambdaS", C.c_double), ("lambdaB", C.c_double), ("q", C.c_double),
                ("nXPP", C.c_double),
                ("wmM", C.c_double), ("wmN", C.c_double), ("wmK", C.c_double), ("wmL", C.c_double),
                ("wmA", C.c_double), ("wmB", C.c_double),
                ("tau0", C.c_double), ("sarN", C.c_double), ("sarK", C.c_double),
                ("dimX", C.c_double), ("dimY", C.c_double), ("dimZ", C.c_double),
                ("sarPTT", C.c_int32),
                ("G0", C.c_double), ("kBMP", C.c_double), ("Phi0", C.c_double), ("PhiInf", C.c_double),
                ("rootK", C.c_double), ("mlAlpha", C.c_double), ("mlBeta", C.c_double),
                ("mlRtol", C.c_double), ("mlMaxIter", C.c_int32)]


class Controls(C.Structure):
    _fields_ = [("solver", C.c_int32), ("precond", C.c_int32), ("tolerance", C.c_double),
                ("relTol", C.c_double), ("maxIter", C.c_int32), ("minIter", C.c_int32)]


class Perf(C.Structure):
    _fields_ = [("initialResidual", C.c_double), ("finalResidual", C.c_double),
                ("nIterations", C.c_int32), ("converged", C.c_int32), ("solved", C.c_int32)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        vp = C.c_void_p
        L.orc_grad_gauss.argtypes = [vp] * 4
        L.orc_rhs.argtypes = [vp, C.c_int32] + [vp] * 6
        L.orc_rhs_tau.argtypes = [vp, C.c_int32] + [vp] * 8
        L.orc_assemble.argtypes = [vp, vp, vp, vp, C.c_double] + [vp] * 8
        L.orc_relax.argtypes = [vp, vp, C.c_double] + [vp] * 6
        L.orc_solve_segregated.argtypes = [vp] * 12
        L.orc_ldu_solve.restype = Perf
        L.orc_ldu_solve.argtypes = [C.c_int32, C.c_int32] + [vp] * 8
        L.orc_foam_residual.restype = C.c_double
        L.orc_foam_residual.argtypes = [C.c_int32, C.c_int32] + [vp] * 7
        L.orc_calc_eig.argtypes = [C.c_int32, vp, vp, vp, C.c_int32]
        L.orc_tau.argtypes = [vp, C.c_int32, vp, vp, vp, vp]
        L.orc_correct.argtypes = [vp, vp, vp, vp, vp, vp, C.c_double, vp, vp, vp, vp,
                                  C.c_double, vp, vp, vp, vp, vp]
        L.orc_correct_scheme.argtypes = [vp, vp, vp, vp, vp, vp, C.c_double, vp, vp, vp, vp,
                                         C.c_double, C.c_int32, vp, vp, vp, vp, vp, vp]
        L.orc_correct_iter.argtypes = [vp, vp, vp, vp, vp, vp, C.c_double, vp, vp, vp, vp,
                                       C.c_double, C.c_int32, vp, vp, vp, vp, vp, vp, vp]
        L.orc_correct_aux.argtypes = [vp, vp, vp, vp, vp, vp, C.c_double, vp, vp, vp, vp,
                                      C.c_double, C.c_int32, vp, vp, vp, vp, vp, vp, vp, vp]
        L.orc_grad_gauss_symm.argtypes = [vp] * 4
        L.orc_defc_faces.argtypes = [vp, C.c_int32] + [vp] * 5
        L.orc_defc_sout.argtypes = [vp] * 4
        L.orc_div_tau.argtypes = [vp, vp, vp, C.c_double, vp]
        L.orc_tau_total.argtypes = [vp, vp, vp, vp, C.c_double, vp]
        L.orc_div_tau_stab.argtypes = [vp, vp, vp, vp, vp, vp, vp, C.c_double, C.c_double, C.c_int32,
                                       C.c_int32, vp]
        L.orc_decomp_create.restype = vp
        L.orc_decomp_create.argtypes = [vp, C.c_int32, vp]
        L.orc_decomp_destroy.argtypes = [vp]
        L.orc_decomp_set_fields.argtypes = [vp] * 9
        L.orc_decomp_correct.restype = C.c_int32
        L.orc_decomp_correct.argtypes = [vp, vp, C.c_double, vp, vp, vp, C.c_int32, vp]
        L.orc_decomp_get.argtypes = [vp] * 5
        L.orc_wedge_face_t.argtypes = [vp, C.c_int32, vp]
        L.orc_wedge_face_t.restype = C.c_int
        _LIB = L
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class OracleMesh:
    """Keeps the numpy arrays alive behind an orc_mesh struct."""

    def __init__(self, m):
        self.m = m
        self.empty = np.ascontiguousarray((m.patchKind == 2).astype(np.int32))
        self.s = Mesh(m.nCells, m.nInternal, m.nBoundary, len(m.patchStart), _p(m.owner),
                      _p(m.neighbour), _p(m.Sf), _p(m.w), _p(m.V), _p(m.patchStart),
                      _p(m.patchSize), _p(self.empty))

    @property
    def ref(self):
        return C.byref(self.s)


def params(model="Oldroyd-BLog", etaS=0.59, etaP=0.41, lam=1.0, alpha=0.0, L2=100.0,
           epsilon=0.0, zeta=0.0, rho=1.0, ptt_exp=False, of_inv_quirk=False,
           lambdaR=1.0, lambdaD=1.0, beta=0.0, delta=-0.5, chiMax=0.0,
           lambdaS=1.0, lambdaB=1.0, q=1.0, n=0.0, m=1.0, K=0.0, Lc=0.0, a=2.0, b=2.0,
           wm_n=1.0, tau0=0.0, sar_n=1.0, sar_k=None, dims=(1.0, 1.0, 1.0), ptt_function="none",
           G0=1.0, k_bmp=0.0, Phi0=0.0, PhiInf=0.0, root_k=2.0, ptt_gen=None, ml_rtol=1e-12,
           ml_max_iter=200):
    """Model constants.  Rolie-Poly: lambdaR lambdaD beta delta chiMax; XPP:
    lambdaS lambdaB alpha q n; White-Metzner CY: lam m wm_n K Lc(=L) a b;
    Saramito: tau0 sar_n sar_k (etaP when sar_n = 1) dims ptt_function
    (n = 1 only) epsilon zeta; BMP: lam G0 k_bmp Phi0 PhiInf; Oldroyd-BRootk:
    root_k; PTTLog generalized: ptt_gen = (alpha, beta) of the Mittag-Leffler
    function, ml_rtol, ml_max_iter."""
    mid = MODEL_IDS[model]
    if model == "PTTLog" and ptt_exp:
        mid = PTT_EXP
    ml_a, ml_b = 1.0, 1.0
    if model == "PTTLog" and ptt_gen is not None:
        mid = PTT_GEN
        ml_a, ml_b = (float(x) for x in ptt_gen)
    k = etaP if (sar_k is None or sar_n == 1.0) else sar_k
    ptt = SARAMITO_PTT[ptt_function] if sar_n == 1.0 else 0
    return Params(mid, rho, etaS, etaP, lam, alpha, L2, epsilon, zeta, int(of_inv_quirk),
                  lambdaR, lambdaD, beta, delta, chiMax, lambdaS, lambdaB, q, n,
                  m, wm_n, K, Lc, a, b, tau0, sar_n, k, *[float(x) for x in dims], ptt,
                  G0, k_bmp, Phi0, PhiInf, root_k, ml_a, ml_b, ml_rtol, int(ml_max_iter))


def controls(solver=PBICGSTAB, precond=PC_DILU, tol=1e-10, relTol=0.0, maxIter=1000, minIter=0):
    return Controls(solver, precond, tol, relTol, maxIter, minIter)


def default_bc(m):
    return np.where(m.patchKind == 2, BC_EMPTY, BC_ZEROGRADIENT).astype(np.int32)


def wedge_face_t(m, patch):
    """faceT of a wedge patch ([OF9] wedgePolyPatch), or raise ValueError."""
    om = OracleMesh(m)
    T = np.empty(9)
    rc = lib().orc_wedge_face_t(om.ref, int(patch), _p(T))
    if rc != 0:
        raise ValueError(f"patch {patch} is no valid wedge (code {rc})")
    return T.reshape(3, 3)


def grad(m, U, Ub):
    om = OracleMesh(m)
    g = np.empty((m.nCells, 3, 3))
    Ub = np.ascontiguousarray(Ub) if m.nBoundary else np.zeros((1, 3))
    lib().orc_grad_gauss(om.ref, _p(np.ascontiguousarray(U)), _p(Ub), _p(g))
    return g


# GaussDefCmpw limited schemes (limiters.H:38-100)
SCHEMES = {"upwind": 0, "cubista": 1, "minmod": 2, "smart": 3, "waceb": 4, "superbee": 5, "none": 6}
BOUNDED = 16  # ORC_DC_BOUNDED


def scheme_id(scheme):
    """'cubista' or 'bounded cubista' -> the ORC_DC_* value."""
    w = scheme.split()
    if w[0] == "bounded":
        return BOUNDED | SCHEMES[w[1]]
    return SCHEMES[w[0]]


def grad_symm(m, theta, thetaB):
    """Gauss-linear gradient of each theta component: [N, 6, 3]."""
    om = OracleMesh(m)
    g = np.empty((m.nCells, 6, 3))
    tb = np.ascontiguousarray(thetaB) if m.nBoundary else np.zeros((1, 6))
    lib().orc_grad_gauss_symm(om.ref, _p(np.ascontiguousarray(theta)), _p(tb), _p(g))
    return g


def defc_faces(m, scheme, phi, theta, g, C=None):
    """Deferred-correction face values (limited minus upwind), [nInternal, 6]."""
    om = OracleMesh(m)
    dc = np.empty((m.nInternal, 6))
    Cc = np.ascontiguousarray(m.C if C is None else C)
    lib().orc_defc_faces(om.ref, SCHEMES[scheme], _p(np.ascontiguousarray(phi)),
                         _p(np.ascontiguousarray(theta)), _p(np.ascontiguousarray(g)), _p(Cc), _p(dc))
    return dc


def defc_sout(m, phi, dc):
    om = OracleMesh(m)
    s = np.empty((m.nCells, 6))
    lib().orc_defc_sout(om.ref, _p(np.ascontiguousarray(phi)), _p(np.ascontiguousarray(dc)), _p(s))
    return s


def div_tau(m, tau, tauB=None, rho=1.0):
    """fvc::div(tau/rho), Gauss linear: [N, 3]."""
    om = OracleMesh(m)
    out = np.empty((m.nCells, 3))
    tb = None if tauB is None else np.ascontiguousarray(tauB)
    lib().orc_div_tau(om.ref, _p(np.ascontiguousarray(tau)), _p(tb), rho, _p(out))
    return out


STAB = {"none": 0, "BSD": 1, "coupling": 2}
SNGRAD = {"corrected": 0, "orthogonal": 1, "uncorrected": 2}


def div_tau_stab(m, U, Ub, tau, tauB=None, rho=1.0, etaP=0.0, stab="none", lap="corrected"):
    """The explicit part of divTau(U) for a stabilization option: [N, 3]."""
    om = OracleMesh(m)
    out = np.empty((m.nCells, 3))
    tb = None if tauB is None else np.ascontiguousarray(tauB)
    Cb = np.ascontiguousarray(m.Cf[m.nInternal:]) if m.nBoundary else np.zeros((1, 3))
    Ub = np.ascontiguousarray(Ub) if m.nBoundary else np.zeros((1, 3))
    lib().orc_div_tau_stab(om.ref, _p(np.ascontiguousarray(m.C)), _p(Cb), _p(np.ascontiguousarray(U)), _p(Ub),
                           _p(np.ascontiguousarray(tau)), _p(tb), rho, etaP, STAB[stab], SNGRAD[lap], _p(out))
    return out


def tau_total(m, U, Ub, tau, etaS):
    om = OracleMesh(m)
    out = np.empty((m.nCells, 6))
    Ub = np.ascontiguousarray(Ub) if m.nBoundary else np.zeros((1, 3))
    lib().orc_tau_total(om.ref, _p(np.ascontiguousarray(U)), _p(Ub), _p(np.ascontiguousarray(tau)), etaS,
                        _p(out))
    return out


def rhs(p, gradU, theta, eigVals, eigVecs, tau_old=None, aux=None):
    """The log-form source S and the per-cell factor; tau_old: the model's
    current tau (Saramito's plasticity factor, BMP's fluidity source), None
    reads as zero; aux [n, 6]: BMP's fluidity in column 0."""
    n = theta.shape[0]
    S = np.empty((n, 6))
    f = np.zeros(n)
    t = None if tau_old is None else np.ascontiguousarray(tau_old, dtype=np.float64)
    lib().orc_rhs_tau(C.byref(p), n, _p(np.ascontiguousarray(gradU)), _p(np.ascontiguousarray(theta)),
                      _p(np.ascontiguousarray(eigVals)), _p(np.ascontiguousarray(eigVecs)), _p(t),
                      _p(None if aux is None else np.ascontiguousarray(aux, dtype=np.float64)), _p(S),
                      _p(f))
    return S, f


def assemble(m, phi, theta, S, dt, bc=None, bcval=None):
    om = OracleMesh(m)
    N, F, NB = m.nCells, m.nInternal, m.nBoundary
    bc = default_bc(m) if bc is None else bc
    bcval = np.zeros((max(NB, 1), 6)) if bcval is None else bcval
    out = dict(diag=np.empty(N), lower=np.empty(F), upper=np.empty(F), source=np.empty((N, 6)),
               iC=np.empty((max(NB, 1), 6)), bC=np.empty((max(NB, 1), 6)))
    lib().orc_assemble(om.ref, _p(phi), _p(np.ascontiguousarray(theta)), _p(S), dt, _p(bc),
                       _p(bcval), _p(out["diag"]), _p(out["lower"]), _p(out["upper"]),
                       _p(out["source"]), _p(out["iC"]), _p(out["bC"]))
    return out


def relax(m, sys, alpha, psi, bc=None):
    """[OF9] fvMatrix::relax on an assembled system (copies): returns the
    relaxed diag and source."""
    om = OracleMesh(m)
    bc = default_bc(m) if bc is None else bc
    D = np.array(sys["diag"], dtype=np.float64)
    S = np.array(sys["source"], dtype=np.float64, order="C")
    lib().orc_relax(om.ref, _p(np.ascontiguousarray(bc, dtype=np.int32)), alpha,
                    _p(np.ascontiguousarray(psi, dtype=np.float64)), _p(D), _p(sys["lower"]), _p(sys["upper"]),
                    _p(S), _p(sys["iC"]))
    return D, S


def solve_segregated(m, sys, psi, ctl=None, bc=None, valid=None):
    om = OracleMesh(m)
    bc = default_bc(m) if bc is None else bc
    valid = m.validComponents if valid is None else np.asarray(valid, np.int32)
    ctl = controls() if ctl is None else ctl
    perf = (Perf * 6)()
    psi = np.ascontiguousarray(psi, dtype=np.float64).copy()
    lib().orc_solve_segregated(om.ref, _p(sys["diag"]), _p(sys["lower"]), _p(sys["upper"]),
                               _p(sys["source"]), _p(sys["iC"]), _p(sys["bC"]), _p(bc),
                               _p(valid), C.byref(ctl), _p(psi), perf)
    return psi, [perf[i] for i in range(6)]


def ldu_solve(n, l, u, diag, lower, upper, b, x0, ctl=None):
    ctl = controls() if ctl is None else ctl
    x = np.ascontiguousarray(x0, dtype=np.float64).copy()
    perf = lib().orc_ldu_solve(n, len(l), _p(l), _p(u), _p(diag), _p(lower), _p(upper), _p(b),
                               C.byref(ctl), _p(x))
    return x, perf


def foam_residual(n, l, u, diag, lower, upper, b, x):
    return lib().orc_foam_residual(n, len(l), _p(l), _p(u), _p(diag), _p(lower), _p(upper), _p(b),
                                   _p(np.ascontiguousarray(x)))


def calc_eig(theta, method=0):
    n = theta.shape[0]
    ev = np.zeros((n, 3, 3))
    vec = np.empty((n, 3, 3))
    lib().orc_calc_eig(n, _p(np.ascontiguousarray(theta)), _p(ev), _p(vec), method)
    return ev, vec


def tau(p, eigVals, eigVecs, fOld=None):
    n = eigVals.shape[0]
    t = np.empty((n, 6))
    fOld = np.zeros(n) if fOld is None else fOld
    lib().orc_tau(C.byref(p), n, _p(np.ascontiguousarray(eigVals)),
                  _p(np.ascontiguousarray(eigVecs)), _p(fOld), _p(t))
    return t


def correct(m, p, fields, dt, ctl=None, bc=None, bcval=None, gradU=None, relax=0.0,
            use_grad=False, scheme="upwind", theta_old=None, tau_old=None, aux=None, valid=None):
    """One constitutiveEq::correct(); returns (theta, eigVals, eigVecs, tau, perf).
    scheme: the GaussDefCmpw limiter of div(phi,theta), optionally prefixed by
    "bounded " (OpenFOAM's bounded convection wrapper).  theta_old: the
    step's theta.oldTime() for an inner iteration (None: a new step).
    tau_old: the model's tau before the step (Saramito reads it; None: zero).
    aux [N, 6]: BMP's fluidity of the step in column 0.  valid: the solved
    components (default: the mesh's validComponents)."""
    om = OracleMesh(m)
    bc = default_bc(m) if bc is None else bc
    NB = m.nBoundary
    bcval = np.zeros((max(NB, 1), 6)) if bcval is None else bcval
    ctl = controls() if ctl is None else ctl
    th = fields.theta.copy()
    ev = fields.eigVals.copy()
    vec = fields.eigVecs.copy()
    t = np.zeros_like(th) if tau_old is None else np.array(tau_old, dtype=np.float64, order="C")
    perf = (Perf * 6)()
    g = gradU if gradU is not None else (fields.gradU if use_grad else None)
    Ub = fields.Ub if NB else np.zeros((1, 3))
    Cc = np.ascontiguousarray(m.C)
    to = None if theta_old is None else np.ascontiguousarray(theta_old, dtype=np.float64)
    va = np.ascontiguousarray(m.validComponents if valid is None else valid, dtype=np.int32)
    ax = None if aux is None else np.ascontiguousarray(aux, dtype=np.float64)
    lib().orc_correct_aux(om.ref, C.byref(p), _p(fields.U), _p(Ub), _p(fields.phi),
                          _p(None if g is None else np.ascontiguousarray(g)), dt, _p(bc),
                          _p(bcval), _p(va), C.byref(ctl), relax, scheme_id(scheme),
                          _p(Cc), _p(to), _p(ax), _p(th), _p(ev), _p(vec), _p(t), perf)
    return th, ev, vec, t, [perf[i] for i in range(6)]


def bmp_correct(m, p, fields, dt, Phi, tau_old, ctl=None, bc=None, bcval=None, relax=0.0,
                scheme="upwind", bc_phi=None, bcval_phi=None, relax_phi=0.0, scheme_phi="upwind",
                ctl_phi=None, gradU=None, theta_old=None, Phi_old=None):
    """BMPLog::correct (BMPLog.C:142-202): the fluidity equation, then the
    theta equation with the new fluidity.  Phi [N], bcval_phi [NB] (patch
    values of fixedValue patches).  Returns (theta, eigVals, eigVecs, tau,
    perf, Phi_new, perf_phi)."""
    import dataclasses
    N, NB = m.nCells, m.nBoundary
    pf = Params.from_buffer_copy(p)
    pf.model = FLUIDITY
    P6 = np.zeros((N, 6))
    P6[:, 0] = Phi
    I = np.broadcast_to(np.eye(3), (N, 3, 3)).copy()
    ff = dataclasses.replace(fields, theta=P6, eigVals=I, eigVecs=I)
    bv = np.zeros((max(NB, 1), 6))
    if bcval_phi is not None:
        bv[:NB, 0] = bcval_phi
    po = None
    if Phi_old is not None:
        po = np.zeros((N, 6))
        po[:, 0] = Phi_old
    Pn, _, _, _, perf_phi = correct(m, pf, ff, dt, ctl=ctl_phi, bc=bc_phi, bcval=bv, gradU=gradU,
                                    relax=relax_phi, scheme=scheme_phi, theta_old=po, tau_old=tau_old,
                                    valid=[1, -1, -1, -1, -1, -1])
    th, ev, vec, t, perf = correct(m, p, fields, dt, ctl=ctl, bc=bc, bcval=bcval, gradU=gradU, relax=relax,
                                   scheme=scheme, theta_old=theta_old, tau_old=tau_old, aux=Pn)
    return th, ev, vec, t, perf, Pn[:, 0].copy(), perf_phi


def simple_domains(m, n, axis=2):
    """[OF9] decomposePar's "simple" method with n domains along one axis:
    cells sorted by centre coordinate (ties by cell number), cut into n
    chunks of equal count."""
    order = np.lexsort((np.arange(m.nCells), np.asarray(m.C)[:, axis]))
    dom = np.empty(m.nCells, dtype=np.int32)
    dom[order] = (np.arange(m.nCells, dtype=np.int64) * n // m.nCells).astype(np.int32)
    return dom


class Decomposed:
    """The stress step of a decomposed run, one thread per domain
    (orc_decomp_*): set_fields() distributes the fields (decomposePar),
    correct() runs one step, get() gathers the result (reconstructPar)."""

    def __init__(self, m, domains):
        self.m = m
        self.om = OracleMesh(m)
        self.domains = np.ascontiguousarray(domains, dtype=np.int32)
        self.n = int(self.domains.max()) + 1
        self.h = lib().orc_decomp_create(self.om.ref, self.n, _p(self.domains))
        if not self.h:
            raise ValueError("invalid domain assignment")

    def set_fields(self, fields, tau_old=None, bcval=None):
        NB = self.m.nBoundary
        Ub = fields.Ub if NB else np.zeros((1, 3))
        self._keep = [np.ascontiguousarray(a, dtype=np.float64) for a in
                      (fields.U, Ub, fields.phi, fields.theta, fields.eigVals, fields.eigVecs)]
        t = None if tau_old is None else np.ascontiguousarray(tau_old, dtype=np.float64)
        bv = None if bcval is None else np.ascontiguousarray(bcval, dtype=np.float64)
        lib().orc_decomp_set_fields(self.h, *[_p(a) for a in self._keep], _p(t), _p(bv))

    def correct(self, p, dt, ctl=None, bc=None, valid=None, scheme="upwind"):
        bc = default_bc(self.m) if bc is None else np.ascontiguousarray(bc, dtype=np.int32)
        va = np.ascontiguousarray(self.m.validComponents if valid is None else valid, dtype=np.int32)
        ctl = controls() if ctl is None else ctl
        perf = (Perf * 6)()
        rc = lib().orc_decomp_correct(self.h, C.byref(p), dt, _p(bc), _p(va), C.byref(ctl),
                                      scheme_id(scheme), perf)
        if rc != 0:
            raise ValueError("decomposed step: unsupported scheme, solver or model, "
                             "or fewer threads than domains")
        return [perf[i] for i in range(6)]

    def get(self):
        N = self.m.nCells
        th, t = np.empty((N, 6)), np.empty((N, 6))
        ev, vec = np.empty((N, 3, 3)), np.empty((N, 3, 3))
        lib().orc_decomp_get(self.h, _p(th), _p(ev), _p(vec), _p(t))
        return th, ev, vec, t

    def close(self):
        if self.h:
            lib().orc_decomp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def decomposed_correct(m, p, fields, dt, n_domains, ctl=None, bc=None, bcval=None, tau_old=None,
                       valid=None, scheme="upwind", domains=None):
    """correct() as a run decomposed into n_domains "simple" domains along z
    executes it; returns (theta, eigVals, eigVecs, tau, perf)."""
    d = Decomposed(m, simple_domains(m, n_domains) if domains is None else domains)
    try:
        d.set_fields(fields, tau_old=tau_old, bcval=bcval)
        perf = d.correct(p, dt, ctl=ctl, bc=bc, valid=valid, scheme=scheme)
        return (*d.get(), perf)
    finally:
        d.close()
