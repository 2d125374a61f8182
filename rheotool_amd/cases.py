"""Synthetic stress-step cases: meshes and seeded fields (ctypes over librheo_case.so).

A :class:`Mesh` holds OpenFOAM-style LDU addressing and geometry as numpy
arrays; :func:`make_fields` fills U, phi, theta and the eigen state.  The
preset workloads are:

* ``channel2d``  -- 128x128x1 Poiseuille channel, Oldroyd-B-Log (CPU case)
* ``box3d``      -- 128^3 Taylor-Green + shear box (FP64 GPU drop-in case)
* ``box3d_perm`` -- 256^3, cells permuted and nodes jittered (precision form)
* ``graph``      -- random face graph (polyhedral stress test)
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_LIB = None
PATCH, WALL, EMPTY = 0, 1, 2
FIELDS_CHANNEL, FIELDS_TG = 0, 1


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "librheo_case.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: run rheotool_amd.build.build_case()")
        L = C.CDLL(path)
        L.rtc_mesh_cartesian.restype = C.c_void_p
        L.rtc_mesh_cartesian.argtypes = [C.c_int] * 3 + [C.c_double] * 3 + [C.c_int, C.c_int,
                                                                               C.c_double,
                                                                               C.c_uint64,
                                                                               C.c_uint64]
        L.rtc_mesh_random_graph.restype = C.c_void_p
        L.rtc_mesh_random_graph.argtypes = [C.c_int, C.c_int, C.c_double, C.c_uint64]
        L.rtc_mesh_from_poly.restype = C.c_void_p
        L.rtc_mesh_from_poly.argtypes = [C.c_int32, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_int32,
                                         C.c_void_p, C.c_void_p, C.c_void_p, C.c_char_p]
        L.rtc_mesh_free.argtypes = [C.c_void_p]
        L.rtc_mesh_sizes.argtypes = [C.c_void_p, C.c_void_p]
        L.rtc_mesh_export.argtypes = [C.c_void_p] + [C.c_void_p] * 11
        L.rtc_mesh_patch_name.restype = C.c_char_p
        L.rtc_mesh_patch_name.argtypes = [C.c_void_p, C.c_int]
        L.rtc_fields.restype = C.c_int
        L.rtc_fields.argtypes = [C.c_void_p, C.c_int, C.c_uint64, C.c_double, C.c_double,
                                 C.c_double] + [C.c_void_p] * 7
        L.rtc_sym_eig.argtypes = [C.c_void_p] * 3
        _LIB = L
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


@dataclass
class Mesh:
    nCells: int
    nInternal: int
    nBoundary: int
    owner: np.ndarray      # int32 [nInternal + nBoundary]
    neighbour: np.ndarray  # int32 [nInternal]
    Sf: np.ndarray         # [nFaces, 3]
    Cf: np.ndarray         # [nFaces, 3]
    w: np.ndarray          # [nInternal]
    V: np.ndarray          # [nCells]
    C: np.ndarray          # [nCells, 3]
    patchStart: np.ndarray
    patchSize: np.ndarray
    patchKind: np.ndarray
    patchNames: list
    origId: np.ndarray
    handle: object = field(default=None, repr=False)

    @property
    def nPatches(self):
        return len(self.patchStart)

    @property
    def emptyDir(self) -> int:
        """Direction normal to the empty patches (-1 for a 3D mesh)."""
        tot = np.zeros(3)
        for p in range(self.nPatches):
            if self.patchKind[p] == EMPTY and self.patchSize[p]:
                s = self.patchStart[p] + self.nInternal
                tot += np.abs(self.Sf[s:s + self.patchSize[p]]).sum(0)
        return int(np.argmax(tot)) if tot.any() else -1

    @property
    def validComponents(self):
        """[OF9] fvMesh::validComponents<symmTensor>() from the solution directions."""
        d = np.ones(3, dtype=np.int32)
        if self.emptyDir >= 0:
            d[self.emptyDir] = -1
        ij = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
        return np.array([d[i] * d[j] for i, j in ij], dtype=np.int32)

    def close(self):
        if self.handle:
            lib().rtc_mesh_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _export(h) -> Mesh:
    L = lib()
    sz = np.zeros(4, dtype=np.int32)
    L.rtc_mesh_sizes(h, _p(sz))
    N, F, NB, NP = (int(x) for x in sz)
    NF = F + NB
    owner = np.empty(NF, np.int32)
    nbr = np.empty(F, np.int32)
    Sf = np.empty((NF, 3))
    Cf = np.empty((NF, 3))
    w = np.empty(F)
    V = np.empty(N)
    Cc = np.empty((N, 3))
    ps = np.empty(NP, np.int32)
    pz = np.empty(NP, np.int32)
    pk = np.empty(NP, np.int32)
    orig = np.empty(N, np.int32)
    L.rtc_mesh_export(h, _p(owner), _p(nbr), _p(Sf), _p(Cf), _p(w), _p(V), _p(Cc), _p(ps),
                      _p(pz), _p(pk), _p(orig))
    names = [L.rtc_mesh_patch_name(h, i).decode() for i in range(NP)]
    return Mesh(N, F, NB, owner, nbr, Sf, Cf, w, V, Cc, ps, pz, pk, names, orig, h)


def cartesian(nx, ny, nz, L=(1.0, 1.0, 1.0), twoD=False, channel=False, jitter=0.0,
              jitter_seed=4, perm_seed=0) -> Mesh:
    Lz = L[2]
    if twoD:
        Lz = L[0] / nx  # one cell thick, dz = dx
    h = lib().rtc_mesh_cartesian(nx, ny, nz, L[0], L[1], Lz, int(twoD), int(channel),
                                 float(jitter), jitter_seed, perm_seed)
    if not h:
        raise ValueError("invalid Cartesian mesh request")
    return _export(h)


def from_poly(points, faces, owner, neighbour, nCells, patches) -> Mesh:
    """A Mesh from polyMesh data: points [nPoints, 3], faces (list of point
    label lists), owner [nFaces], neighbour [nInternal], and patches as
    (name, kind, start, size) with start counted from the first boundary face."""
    pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    fs = np.zeros(len(faces) + 1, np.int32)
    fs[1:] = np.cumsum([len(f) for f in faces])
    fl = np.ascontiguousarray(np.concatenate([np.asarray(f, np.int32) for f in faces]) if faces
                              else np.zeros(0, np.int32))
    own = np.ascontiguousarray(owner, np.int32)
    nbr = np.ascontiguousarray(neighbour, np.int32)
    ps = np.array([p[2] for p in patches], np.int32)
    pz = np.array([p[3] for p in patches], np.int32)
    pk = np.array([p[1] for p in patches], np.int32)
    names = b"".join(p[0].encode() + b"\0" for p in patches)
    h = lib().rtc_mesh_from_poly(len(pts), _p(pts), len(faces), _p(fs), _p(fl), _p(own), len(nbr),
                                 _p(nbr), int(nCells), len(patches), _p(ps), _p(pz), _p(pk), names)
    if not h:
        raise ValueError("inconsistent polyMesh description")
    return _export(h)


def random_graph(nCells, kNear=4, sigma=0.5, seed=1) -> Mesh:
    h = lib().rtc_mesh_random_graph(nCells, kNear, sigma, seed)
    if not h:
        raise ValueError("invalid random graph request")
    return _export(h)


@dataclass
class Fields:
    U: np.ndarray
    Ub: np.ndarray
    phi: np.ndarray
    gradU: np.ndarray
    theta: np.ndarray
    eigVals: np.ndarray
    eigVecs: np.ndarray


def make_fields(m: Mesh, kind=FIELDS_CHANNEL, seed=2, noise=1e-2, lnLo=-20.0, lnHi=5.0,
                want_grad=True) -> Fields:
    N, NB, NF = m.nCells, m.nBoundary, m.nInternal + m.nBoundary
    U = np.empty((N, 3))
    Ub = np.empty((max(NB, 0), 3))
    phi = np.empty(NF)
    g = np.empty((N, 3, 3)) if want_grad else None
    th = np.empty((N, 6))
    ev = np.empty((N, 3, 3))
    evec = np.empty((N, 3, 3))
    rc = lib().rtc_fields(m.handle, kind, seed, noise, lnLo, lnHi, _p(U), _p(Ub) if NB else None,
                          _p(phi), _p(g), _p(th), _p(ev), _p(evec))
    if rc != 0:
        raise RuntimeError("rtc_fields failed")
    return Fields(U, Ub, phi, g, th, ev, evec)


def sym_eig(theta6):
    ev = np.empty(3)
    R = np.empty(9)
    t = np.ascontiguousarray(theta6, dtype=np.float64)
    lib().rtc_sym_eig(_p(t), _p(ev), _p(R))
    return ev, R.reshape(3, 3)


# ---------------------------------------------------------------- presets
def channel2d(n=128, seed=2):
    """128x128x1 Poiseuille channel on [0,1]^2 (Oldroyd-B-Log CPU case)."""
    m = cartesian(n, n, 1, twoD=True, channel=True)
    f = make_fields(m, FIELDS_CHANNEL, seed=seed)
    return m, f


def box3d(n=128, seed=2, lnLo=-20.0, lnHi=5.0, perm_seed=0, jitter=0.0):
    """n^3 Taylor-Green + shear box on [0,1]^3."""
    m = cartesian(n, n, n, jitter=jitter, perm_seed=perm_seed)
    f = make_fields(m, FIELDS_TG, seed=seed, lnLo=lnLo, lnHi=lnHi)
    return m, f
