"""ctypes binding of librheo_gpu.so (the C-ABI of include/rheo_gpu.h).

Mirrors the reference's call surface for the stress step:

* :class:`Context`                 -- device, stream, HBM budget, kernel timing
* :class:`Mesh`                    -- lduAddressing + geometry (primitiveMesh)
* :meth:`Context.constitutiveEq`   -- constitutiveEq::New(name, U, phi, dict)
  returning a :class:`ConstitutiveEq` with ``correct()``, ``tau()``,
  ``theta()``, ``eigVals()``, ``eigVecs()``
* :meth:`Context.symm_solve`       -- sparseSolver<symmTensor>::solve(fvMatrix&)

There is no fallback: if the HIP library is missing or no gfx950 device is
visible, the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref

import numpy as np

_LIB = None
HERE = os.path.dirname(os.path.abspath(__file__))
# RHEO_GPU_LIB selects another build of the library (e.g. a tuning variant)
LIB_PATH = os.environ.get("RHEO_GPU_LIB") or os.path.join(HERE, "lib", "librheo_gpu.so")

RT_OK, RT_E_UNKNOWN_TYPE, RT_E_BAD_ARG, RT_E_OOM, RT_E_HIP, RT_E_UNSUPPORTED, RT_E_NO_DEVICE = range(7)
RT_W_NOT_CONVERGED = 0x100
RT_THETA, RT_TAU, RT_EIGVALS, RT_EIGVECS = range(4)
RT_RENUMBER_AUTO, RT_RENUMBER_NONE, RT_RENUMBER_RCM, RT_RENUMBER_SFC = range(4)
RT_RENUMBER_NO_COLOUR = 16
RT_BC_ZEROGRADIENT, RT_BC_FIXEDVALUE, RT_BC_EMPTY, RT_BC_SYMMETRY, RT_BC_WEDGE = range(5)


class RheoError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[rt status {code}] {msg}")
        self.code = code
        self.msg = msg


class Patch(C.Structure):
    _fields_ = [("start", C.c_int32), ("size", C.c_int32), ("kind", C.c_int32)]


class MeshDesc(C.Structure):
    _fields_ = [("nCells", C.c_int32), ("nInternalFaces", C.c_int32),
                ("nBoundaryFaces", C.c_int32), ("nPatches", C.c_int32),
                ("owner", C.c_void_p), ("neighbour", C.c_void_p), ("Sf", C.c_void_p),
                ("weights", C.c_void_p), ("V", C.c_void_p), ("patches", C.c_void_p),
                ("C", C.c_void_p), ("Cb", C.c_void_p)]


class SolverPerf(C.Structure):
    _fields_ = [("initialResidual", C.c_double), ("finalResidual", C.c_double),
                ("nIterations", C.c_int32), ("nRefinements", C.c_int32),
                ("converged", C.c_int32), ("solved", C.c_int32)]

    def __repr__(self):
        return (f"SolverPerf(init={self.initialResidual:.3e}, final={self.finalResidual:.3e}, "
                f"it={self.nIterations}, ref={self.nRefinements}, conv={self.converged}, "
                f"solved={self.solved})")


class Request(C.Structure):
    _fields_ = [("mesh", C.c_void_p), ("dict", C.c_char_p), ("controls", C.c_char_p),
                ("theta", C.c_void_p), ("eigVals", C.c_void_p), ("eigVecs", C.c_void_p),
                ("U", C.c_void_p), ("Ub", C.c_void_p), ("phi", C.c_void_p),
                ("deltaT", C.c_double), ("thetaOut", C.c_void_p), ("tauOut", C.c_void_p),
                ("gradU", C.c_void_p)]


class KernelStat(C.Structure):
    _fields_ = [("name", C.c_char * 48), ("launches", C.c_int64), ("total_ms", C.c_double),
                ("bytes", C.c_double)]


class LduSymm(C.Structure):
    _fields_ = [("diag", C.c_void_p), ("lower", C.c_void_p), ("upper", C.c_void_p),
                ("source", C.c_void_p), ("internalCoeffs", C.c_void_p),
                ("boundaryCoeffs", C.c_void_p)]


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with rheotool_amd.build.build_gpu()")
        L = C.CDLL(LIB_PATH)
        vp, i32 = C.c_void_p, C.c_int
        sig = {
            "rt_device_count": (i32, []),
            "rt_ctx_create": (i32, [i32, C.c_size_t, vp]),
            "rt_ctx_destroy": (None, [vp]),
            "rt_last_error": (C.c_char_p, [vp]),
            "rt_ctx_bytes_in_use": (C.c_size_t, [vp]),
            "rt_ctx_set_profiling": (i32, [vp, i32]),
            "rt_ctx_kernel_stats": (i32, [vp, vp, i32]),
            "rt_ctx_reset_stats": (None, [vp]),
            "rt_ctx_copy_bench": (i32, [vp, C.c_size_t, i32, vp]),
            "rt_mesh_create": (i32, [vp, vp, i32, vp]),
            "rt_mesh_destroy": (None, [vp]),
            "rt_mesh_valid_components": (i32, [vp, vp]),
            "rt_model_create": (i32, [vp, C.c_char_p, vp, C.c_char_p, vp]),
            "rt_model_destroy": (None, [vp]),
            "rt_model_type": (C.c_char_p, [vp]),
            "rt_model_set_state": (i32, [vp, vp, vp, vp, vp]),
            "rt_model_set_bc": (i32, [vp, i32, i32, vp]),
            "rt_model_set_flow": (i32, [vp, vp, vp, vp]),
            "rt_model_set_solver": (i32, [vp, C.c_char_p]),
            "rt_model_set_relax": (i32, [vp, C.c_double]),
            "rt_model_set_scheme": (i32, [vp, C.c_char_p]),
            "rt_model_mode_count": (i32, [vp]),
            "rt_model_mode": (vp, [vp, i32]),
            "rt_model_fluidity": (vp, [vp]),
            "rt_model_div_tau": (i32, [vp, vp, vp]),
            "rt_model_tau_total": (i32, [vp, vp]),
            "rt_model_div_tau_coeff": (i32, [vp, vp]),
            "rt_model_set_laplacian_scheme": (i32, [vp, C.c_char_p]),
            "rt_model_correct": (i32, [vp, C.c_double, vp]),
            "rt_model_correct_iter": (i32, [vp, C.c_double, vp, i32]),
            "rt_model_get": (i32, [vp, i32, vp]),
            "rt_model_solver_perf": (i32, [vp, vp]),
            "rt_symm_solve": (i32, [vp, vp, vp, vp, C.c_char_p, vp]),
            "rt_symm_solve_noalloc": (i32, [vp, vp, vp, vp, C.c_char_p, vp]),
            "rt_server_create": (i32, [vp, C.c_size_t, i32, vp]),
            "rt_server_destroy": (None, [vp]),
            "rt_server_submit": (i32, [vp, vp, vp]),
            "rt_server_progress": (i32, [vp]),
            "rt_server_poll": (i32, [vp, C.c_int64, vp, vp, vp, vp]),
            "rt_server_finished": (i32, [vp, vp, i32]),
            "rt_server_set_batching": (i32, [vp, i32]),
            "rt_request_bytes_estimate": (C.c_size_t, [vp, i32, i32]),
            "rt_server_batched_rounds": (C.c_int64, [vp]),
            "rt_host_alloc": (i32, [vp, C.c_size_t, vp]),
            "rt_host_free": (None, [vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def device_count() -> int:
    return lib().rt_device_count()


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f64(a, shape=None):
    """C-contiguous float64 copy or view; with shape, the array must hold
    exactly that many values (the C library reads that many)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        if a.size != int(np.prod(shape)):
            raise ValueError(f"expected {int(np.prod(shape))} values {tuple(shape)}, got shape {a.shape}")
        a = a.reshape(shape)
    return a


class Context:
    def __init__(self, device=0, hbm_budget=0):
        h = C.c_void_p()
        rc = lib().rt_ctx_create(device, hbm_budget, C.byref(h))
        if rc != RT_OK:
            raise RheoError(rc, "rt_ctx_create failed (no gfx950 device?)")
        self.h = h

    def check(self, rc, what=""):
        if rc != RT_OK and rc != RT_W_NOT_CONVERGED:
            raise RheoError(rc, f"{what}: {lib().rt_last_error(self.h).decode()}")
        return rc

    def close(self):
        if getattr(self, "h", None):
            lib().rt_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def bytes_in_use(self):
        return lib().rt_ctx_bytes_in_use(self.h)

    def set_profiling(self, on=True):
        self.check(lib().rt_ctx_set_profiling(self.h, int(on)), "set_profiling")

    def kernel_stats(self):
        buf = (KernelStat * 64)()
        n = lib().rt_ctx_kernel_stats(self.h, buf, 64)
        if n < 0:
            raise RheoError(-n, lib().rt_last_error(self.h).decode())
        return {buf[i].name.decode(): dict(launches=buf[i].launches, ms=buf[i].total_ms,
                                           bytes=buf[i].bytes) for i in range(n)}

    def reset_stats(self):
        lib().rt_ctx_reset_stats(self.h)

    def copy_bench(self, nbytes, reps=20):
        """Mean time (ms) of the plain device copy kernel over nbytes."""
        ms = C.c_double()
        self.check(lib().rt_ctx_copy_bench(self.h, nbytes, reps, C.byref(ms)), "copy_bench")
        return ms.value

    def pinned_empty(self, shape, dtype=np.float64):
        """Uninitialised numpy array in page-locked host memory
        (rt_host_alloc): copies from and to it run asynchronously, so a
        server's uploads and downloads overlap the device's work.  The memory
        is freed when the last array using it is collected."""
        n = max(1, int(np.prod(shape)) * np.dtype(dtype).itemsize)
        p = C.c_void_p()
        self.check(lib().rt_host_alloc(self.h, n, C.byref(p)), "rt_host_alloc")
        buf = (C.c_char * n).from_address(p.value)
        fin = weakref.finalize(buf, lib().rt_host_free, C.c_void_p(p.value))
        fin.atexit = False  # the process's end returns it anyway
        return np.frombuffer(buf, dtype=dtype, count=int(np.prod(shape))).reshape(shape)

    def mesh(self, m, renumber=RT_RENUMBER_AUTO) -> "Mesh":
        return Mesh(self, m, renumber)

    def constitutiveEq(self, name, mesh, dict_text) -> "ConstitutiveEq":
        return ConstitutiveEq(self, name, mesh, dict_text)

    def symm_solve(self, mesh, diag, lower, upper, source, psi, controls, internalCoeffs=None,
                   boundaryCoeffs=None, noalloc=False):
        """sparseSolver<symmTensor>::solve: returns (psi, perf[6]).  noalloc:
        through rt_symm_solve_noalloc (no device allocation; meshes up to the
        library's scratch size)."""
        N, F, NB = mesh.nCells, mesh.nInternal, mesh.nBoundary
        keep = [_f64(diag, (N,)), _f64(lower, (F,)), _f64(upper, (F,)), _f64(source, (N, 6)),
                None if internalCoeffs is None else _f64(internalCoeffs, (NB, 6)),
                None if boundaryCoeffs is None else _f64(boundaryCoeffs, (NB, 6))]
        s = LduSymm(*[_p(k) for k in keep])
        x = _f64(psi, (N, 6)).copy()
        perf = (SolverPerf * 6)()
        fn = lib().rt_symm_solve_noalloc if noalloc else lib().rt_symm_solve
        rc = fn(self.h, mesh.h, C.byref(s), _p(x), controls.encode(), perf)
        self.check(rc, "rt_symm_solve_noalloc" if noalloc else "rt_symm_solve")
        return x, [perf[i] for i in range(6)]


class Mesh:
    """Device mesh built from a rheotool_amd.cases.Mesh (or any object with the
    same arrays)."""

    def __init__(self, ctx: Context, m, renumber=RT_RENUMBER_AUTO):
        self.ctx = ctx
        self.src = m
        d, self._keep = mesh_desc(m)
        h = C.c_void_p()
        ctx.check(lib().rt_mesh_create(ctx.h, C.byref(d), renumber, C.byref(h)), "rt_mesh_create")
        self.h = h
        self.nCells = m.nCells
        self.nInternal = int(m.nInternal)
        self.nBoundary = int(m.nBoundary)
        self.patchSize = [int(x) for x in m.patchSize]

    @property
    def validComponents(self):
        v = np.zeros(6, np.int32)
        lib().rt_mesh_valid_components(self.h, _p(v))
        return v

    def close(self):
        if getattr(self, "h", None):
            lib().rt_mesh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ConstitutiveEq:
    """constitutiveEq selected at run time from a constitutiveProperties dictionary."""

    def __init__(self, ctx: Context, name, mesh: Mesh, dict_text, _handle=None, _owner=None):
        self.ctx, self.mesh = ctx, mesh
        self._owner = _owner  # a mode keeps its multiMode model alive and is not destroyed itself
        if _handle is not None:
            self.h = _handle
        else:
            h = C.c_void_p()
            ctx.check(lib().rt_model_create(ctx.h, name.encode(), mesh.h, dict_text.encode(),
                                            C.byref(h)), "constitutiveEq::New")
            self.h = h
        self.N = mesh.nCells

    def modes(self):
        """The modes of a multiMode model (empty for a single model)."""
        n = lib().rt_model_mode_count(self.h)
        return [ConstitutiveEq(self.ctx, None, self.mesh, None, _handle=C.c_void_p(lib().rt_model_mode(self.h, i)),
                               _owner=self) for i in range(n)]

    def fluidity(self):
        """BMPLog's fluidity equation (None for other models): scalar state
        Phi [N], one fixedValue value per face, its own solver/scheme/relax."""
        h = lib().rt_model_fluidity(self.h)
        return FluidityEq(self.ctx, None, self.mesh, None, _handle=C.c_void_p(h), _owner=self) if h else None

    @property
    def type(self):
        return lib().rt_model_type(self.h).decode()

    def set_state(self, theta, eigVals=None, eigVecs=None, tau=None):
        N = self.N
        keep = [_f64(theta, (N, 6)), None if eigVals is None else _f64(eigVals, (N, 3, 3)),
                None if eigVecs is None else _f64(eigVecs, (N, 3, 3)),
                None if tau is None else _f64(tau, (N, 6))]
        self.ctx.check(lib().rt_model_set_state(self.h, *[_p(k) for k in keep]), "set_state")

    def set_flow(self, U, Ub, phi):
        NB = self.mesh.nBoundary
        keep = [_f64(U, (self.N, 3)), _f64(Ub, (NB, 3)) if Ub is not None and np.size(Ub) else None,
                _f64(phi, (self.mesh.nInternal + NB,))]
        self.ctx.check(lib().rt_model_set_flow(self.h, *[_p(k) for k in keep]), "set_flow")

    _BC_WIDTH = 6  # values per face of a fixedValue patch

    def set_bc(self, patch, bc_type, values=None):
        v = None
        if values is not None:
            if not 0 <= patch < len(self.mesh.patchSize):
                raise ValueError(f"patch {patch} out of range")
            v = _f64(values, (self.mesh.patchSize[patch] * self._BC_WIDTH,))
        self.ctx.check(lib().rt_model_set_bc(self.h, patch, bc_type, _p(v)), "set_bc")

    def set_solver(self, controls: str):
        self.ctx.check(lib().rt_model_set_solver(self.h, controls.encode()), "set_solver")

    def set_relax(self, factor: float):
        self.ctx.check(lib().rt_model_set_relax(self.h, factor), "set_relax")

    def set_scheme(self, div_scheme: str):
        """fvSchemes div(phi,theta): "GaussDefCmpw cubista", "upwind", ..."""
        self.ctx.check(lib().rt_model_set_scheme(self.h, div_scheme.encode()), "set_scheme")

    def div_tau(self, tauB=None):
        """The explicit part of divTau(U) for the model's stabilization option
        (fvc::div(tau/rho), minus the BSD laplacian or the coupling
        div((etaP/rho) grad U)) from the current tau and flow: [N, 3]."""
        out = np.empty((self.N, 3))
        tb = None if tauB is None else _f64(tauB, (self.mesh.nBoundary, 6))
        self.ctx.check(lib().rt_model_div_tau(self.h, _p(tb), _p(out)), "div_tau")
        return out

    def div_tau_coeff(self) -> float:
        """nu of divTau's implicit fvm::laplacian(nu, U)."""
        v = C.c_double()
        self.ctx.check(lib().rt_model_div_tau_coeff(self.h, C.byref(v)), "div_tau_coeff")
        return v.value

    def set_laplacian_scheme(self, text: str):
        """fvSchemes laplacian(eta,U): "Gauss linear corrected|orthogonal|uncorrected"."""
        self.ctx.check(lib().rt_model_set_laplacian_scheme(self.h, text.encode()), "set_laplacian_scheme")

    def tau_total(self):
        out = np.empty((self.N, 6))
        self.ctx.check(lib().rt_model_tau_total(self.h, _p(out)), "tau_total")
        return out

    def correct(self, deltaT, gradU=None, iteration=None):
        """One step; iteration 0, 1, ... makes it an inner iteration of a step
        (the ddt keeps theta.oldTime() of iteration 0)."""
        g = None if gradU is None else _f64(gradU, (self.N, 3, 3))
        if iteration is None:
            return self.ctx.check(lib().rt_model_correct(self.h, deltaT, _p(g)), "correct")
        return self.ctx.check(lib().rt_model_correct_iter(self.h, deltaT, _p(g), int(iteration)), "correct")

    def _get(self, which, shape):
        out = np.empty(shape)
        self.ctx.check(lib().rt_model_get(self.h, which, _p(out)), "get")
        return out

    def get_into(self, which, out):
        """rt_model_get into a caller's C-contiguous float64 array (theta, tau: [N, 6])."""
        if not (isinstance(out, np.ndarray) and out.dtype == np.float64 and out.flags.c_contiguous
                and out.size == (9 if which in (RT_EIGVALS, RT_EIGVECS) else 6) * self.N):
            raise ValueError("get_into: a C-contiguous float64 array of the field's size is needed")
        self.ctx.check(lib().rt_model_get(self.h, which, _p(out)), "get")
        return out

    def theta(self):
        return self._get(RT_THETA, (self.N, 6))

    def tau(self):
        return self._get(RT_TAU, (self.N, 6))

    def eigVals(self):
        return self._get(RT_EIGVALS, (self.N, 3, 3))

    def eigVecs(self):
        return self._get(RT_EIGVECS, (self.N, 3, 3))

    def solver_perf(self):
        perf = (SolverPerf * 6)()
        lib().rt_model_solver_perf(self.h, perf)
        return [perf[i] for i in range(6)]

    def close(self):
        if getattr(self, "h", None):
            if self._owner is None:
                lib().rt_model_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# dimensions of the further models' constants (RoliePolyLog.C:117-121,
# XPomPomLog.C:118-122, WhiteMetznerCYLog.C:119-124)
_DIMS = {"lambdaR": "[0 0 1 0 0 0 0]", "lambdaD": "[0 0 1 0 0 0 0]", "lambdaS": "[0 0 1 0 0 0 0]",
         "lambdaB": "[0 0 1 0 0 0 0]", "K": "[0 0 1 0 0 0 0]", "L": "[0 0 1 0 0 0 0]",
         "tau0": "[1 -1 -2 0 0 0 0]", "k": "[1 -1 -1 0 0 0 0]"}


class FluidityEq(ConstitutiveEq):
    """The fluidity equation of a BMPLog model (owned by it)."""

    _BC_WIDTH = 1

    def set_state(self, Phi, eigVals=None, eigVecs=None, tau=None):
        keep = _f64(Phi, (self.N,))
        self.ctx.check(lib().rt_model_set_state(self.h, _p(keep), None, None, None), "set_state")

    def Phi(self):
        out = np.empty(self.N)
        self.ctx.check(lib().rt_model_get(self.h, RT_THETA, _p(out)), "get")
        return out

    theta = Phi


def model_dict(model="Oldroyd-BLog", etaS=0.59, etaP=0.41, lam=1.0, alpha=None, L2=None,
               epsilon=None, zeta=None, destruction="linear", rho=1.0, stabilization="coupling", **extra):
    """constitutiveProperties `parameters` text in OpenFOAM syntax.  extra:
    further constants by their dictionary keys (lambdaR, lambdaD, beta, delta,
    chiMax, lambdaS, lambdaB, q, n, m, K, L, a, b)."""
    lines = [f"type {model};", f"rho rho [1 -3 0 0 0 0 0] {rho!r};",
             f"etaS etaS [1 -1 -1 0 0 0 0] {etaS!r};", f"etaP etaP [1 -1 -1 0 0 0 0] {etaP!r};"]
    if model not in ("Rolie-PolyLog", "XPomPomLog"):
        lines.append(f"lambda lambda [0 0 1 0 0 0 0] {lam!r};")
    for k, v in extra.items():
        if isinstance(v, str):  # a word, e.g. Saramito's PTTfunction
            lines.append(f"{k} {v};")
        elif isinstance(v, (tuple, list)):  # a vector, e.g. Saramito's dims
            lines.append(f"{k} ({' '.join(repr(float(x)) for x in v)});")
        else:
            lines.append(f"{k} {k} {_DIMS.get(k, '[0 0 0 0 0 0 0]')} {v!r};")
    if alpha is not None:
        lines.append(f"alpha alpha [0 0 0 0 0 0 0] {alpha!r};")
    if L2 is not None:
        lines.append(f"L2 L2 [0 0 0 0 0 0 0] {L2!r};")
    if epsilon is not None:
        lines.append(f"epsilon epsilon [0 0 0 0 0 0 0] {epsilon!r};")
    if zeta is not None:
        lines.append(f"zeta zeta [0 0 0 0 0 0 0] {zeta!r};")
    if model == "PTTLog":
        lines.append(f"destructionFunctionType {destruction};")
    lines.append(f"stabilization {stabilization};")
    return "parameters\n{\n    " + "\n    ".join(lines) + "\n}\n"


def solver_dict(tol=1e-10, precond="DILU", precision="double", maxIter=1000, relTol=0.0,
                innerRelTol=None, solver="PBiCGStab"):
    s = (f"solver {solver}; preconditioner {precond}; tolerance {tol!r}; relTol {relTol!r}; "
         f"maxIter {maxIter}; minIter 0; precision {precision};")
    if innerRelTol is not None:
        s += f" innerRelTol {innerRelTol!r};"
    return s


def mesh_desc(m):
    """rt_mesh_desc for a cases.Mesh; returns (desc, keepalive)."""
    keep = dict(owner=np.ascontiguousarray(m.owner, np.int32),
                nbr=np.ascontiguousarray(m.neighbour, np.int32), Sf=_f64(m.Sf), w=_f64(m.w),
                V=_f64(m.V), C=_f64(m.C) if getattr(m, "C", None) is not None else None,
                Cb=_f64(m.Cf[m.nInternal:]) if getattr(m, "Cf", None) is not None and m.nBoundary else None)
    pt = (Patch * max(1, len(m.patchStart)))()
    for i in range(len(m.patchStart)):
        pt[i] = Patch(int(m.patchStart[i]), int(m.patchSize[i]), int(m.patchKind[i]))
    keep["pt"] = pt
    d = MeshDesc(m.nCells, m.nInternal, m.nBoundary, len(m.patchStart), _p(keep["owner"]),
                 _p(keep["nbr"]), _p(keep["Sf"]), _p(keep["w"]), _p(keep["V"]),
                 C.cast(pt, C.c_void_p), _p(keep["C"]), _p(keep["Cb"]))
    return d, keep


class Server:
    """Continuous-batching request loop (rt_server_*)."""

    def __init__(self, ctx: Context, hbm_budget=0, max_inflight=8, batching=True):
        h = C.c_void_p()
        ctx.check(lib().rt_server_create(ctx.h, hbm_budget, max_inflight, C.byref(h)),
                  "rt_server_create")
        self.ctx, self.h = ctx, h
        self.held = {}  # ticket -> (input arrays, theta out, tau out): the server reads/writes them
        ctx.check(lib().rt_server_set_batching(h, int(batching)), "rt_server_set_batching")

    @property
    def batched_rounds(self) -> int:
        """Batched Krylov rounds launched so far (rt_server_batched_rounds)."""
        return int(lib().rt_server_batched_rounds(self.h))

    def submit(self, mesh: "Mesh", fields, dict_text, deltaT, controls=None, gradU=None, out=None,
               eigen=True) -> int:
        """mesh: an uploaded Mesh (shared by any number of requests).  The
        field arrays are referenced by the server until the ticket finishes.
        gradU: optional prescribed velocity gradient [N, 3, 3].  out: optional
        (theta, tau) result arrays, float64 [N, 6] C-contiguous, written in
        place (page-locked ones, Context.pinned_empty, let the download
        overlap other work).  eigen=False sends no eigen state: the server
        computes it from theta on the device (calcEig)."""
        m = mesh.src
        N, NB = m.nCells, m.nBoundary
        arrs = [_f64(fields.theta, (N, 6)), _f64(fields.eigVals, (N, 3, 3)) if eigen else None,
                _f64(fields.eigVecs, (N, 3, 3)) if eigen else None,
                _f64(fields.U, (N, 3)), _f64(fields.Ub, (NB, 3)) if NB else None,
                _f64(fields.phi, (m.nInternal + NB,))]
        g = None if gradU is None else _f64(gradU, (N, 3, 3))
        if out is None:
            th_out, tau_out = np.empty((N, 6)), np.empty((N, 6))
        else:
            th_out, tau_out = out
            for a in (th_out, tau_out):
                if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.shape == (N, 6)
                        and a.flags.c_contiguous and a.flags.writeable):
                    raise ValueError(f"out arrays must be writable C-contiguous float64 ({N}, 6)")
        q = Request(mesh.h, dict_text.encode(), None if controls is None else controls.encode(),
                    *[_p(a) for a in arrs], deltaT, _p(th_out), _p(tau_out), _p(g))
        arrs.append(g)
        t = C.c_int64()
        self.ctx.check(lib().rt_server_submit(self.h, C.byref(q), C.byref(t)), "rt_server_submit")
        self.held[t.value] = (arrs, th_out, tau_out, mesh)  # the mesh must outlive the request
        return t.value

    def progress(self) -> int:
        n = lib().rt_server_progress(self.h)
        if n < 0:
            raise RheoError(-n, lib().rt_last_error(self.ctx.h).decode())
        return n

    def finished(self, max_n=256):
        """Tickets that finished or failed since the last call."""
        buf = (C.c_int64 * max_n)()
        out = []
        while True:
            n = lib().rt_server_finished(self.h, buf, max_n)
            if n < 0:
                raise RheoError(-n, lib().rt_last_error(self.ctx.h).decode())
            out.extend(buf[i] for i in range(n))
            if n < max_n:
                return out

    def poll(self, ticket):
        """None while pending, else dict(theta, tau, perf, latency_ms)."""
        perf = (SolverPerf * 6)()
        lat = C.c_double()
        rc = lib().rt_server_poll(self.h, ticket, None, None, perf, C.byref(lat))
        if rc == 0:
            return None
        _, th, tau, _ = self.held.pop(ticket)
        if rc < 0:
            raise RheoError(-rc, lib().rt_last_error(self.ctx.h).decode())
        return dict(theta=th, tau=tau, perf=[perf[i] for i in range(6)], latency_ms=lat.value)

    def close(self):
        if getattr(self, "h", None):
            lib().rt_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
