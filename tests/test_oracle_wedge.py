"""wedge theta patches in the oracle ([OF9] wedgePolyPatch::calcGeometry,
wedgeFvPatchField through gaussDefCmpwConvectionScheme.C:103-117).

Pinned three ways:
- the patch rotation faceT against the closed-form rotation about the x axis
  by the half wedge angle, for both patches of a wedge built with known
  angles, and the case reader's own faceT against the oracle's;
- the convection coefficients on a wedge face with a prescribed flux against
  closed forms: 1 - d_i d_j with d = (0, sin^2 a, sin^2 a), and the patch
  value R(a) theta R(a)^T from trigonometry;
- the whole step against a full ring of sectors: for axisymmetric theta and
  velocity, a wedge of half angle a approximates one sector of the ring to
  O(a^2) (the wedge face values are the rotated cell values where the ring
  interpolates the cell and its rotated neighbour), so halving the angle
  quarters the difference.  A velocity wedge value taken as the cell value,
  or rotated the wrong way, misses the hoop terms and does not converge.
  theta's wedge value moves this step only at O(a^2) (no flux crosses the
  wedge faces, and the deferred correction reads no gradient across them),
  so the coefficient test is what pins it."""
import numpy as np
import pytest

from oracle import oracle as O
from rheotool_amd import foamdict, polymesh, stressfoam

LX, R0, R1 = 1.0, 0.5, 1.5
NXC, NRC = 4, 6


def _dict(a_deg, sectors):
    """blockMeshDict text of `sectors` blocks of half angle a_deg about the x
    axis, sector k centred on the angle 2 a k; one sector has wedge patches
    front (+a) and back (-a), a full ring has none."""
    a = np.radians(a_deg)
    ring = sectors > 1
    nang = sectors if ring else 2
    verts = []
    for m in range(nang):
        ph = -a + 2 * a * m
        for x, r in ((0.0, R0), (LX, R0), (LX, R1), (0.0, R1)):
            verts.append((float(x), float(r * np.cos(ph)), float(r * np.sin(ph))))
    blocks, pats = [], {"inlet": [], "outlet": [], "inner": [], "outer": [], "front": [], "back": []}
    for k in range(sectors):
        lo = 4 * k
        hi = 4 * ((k + 1) % nang)
        v = [lo, lo + 1, lo + 2, lo + 3, hi, hi + 1, hi + 2, hi + 3]
        blocks.append(f"hex ({' '.join(map(str, v))}) ({NXC} {NRC} 1) simpleGrading (1 1 1)")
        pats["inlet"].append((v[0], v[4], v[7], v[3]))
        pats["outlet"].append((v[1], v[2], v[6], v[5]))
        pats["inner"].append((v[0], v[1], v[5], v[4]))
        pats["outer"].append((v[3], v[7], v[6], v[2]))
        if not ring:
            pats["back"].append((v[0], v[3], v[2], v[1]))
            pats["front"].append((v[4], v[5], v[6], v[7]))
    vt = "\n".join(f"({x!r} {y!r} {z!r})" for x, y, z in verts)
    bt = "\n".join(blocks)
    bd = ""
    for name, quads in pats.items():
        if not quads:
            continue
        typ = "wedge" if name in ("front", "back") else "patch"
        qs = "\n".join("(" + " ".join(map(str, q)) + ")" for q in quads)
        bd += f"{name} {{ type {typ}; faces ( {qs} ); }}\n"
    return f"convertToMeters 1;\nvertices ( {vt} );\nblocks ( {bt} );\nboundary ( {bd} );\n"


def mesh(a_deg, sectors):
    return polymesh.block_mesh(foamdict.parse(_dict(a_deg, sectors))).to_mesh()


def rot(ph):
    """Rotation about the x axis by ph (y towards z)."""
    c, s = np.cos(ph), np.sin(ph)
    return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])


def cyl_theta(x, r):
    """theta in the (x, r, phi) frame: xx, xr, xphi, rr, rphi, phiphi."""
    return np.stack([0.4 + 0.3 * np.sin(2 * x) + 0.2 * r, 0.25 * np.cos(x + r), 0.05 * np.sin(r),
                     -0.3 + 0.4 * r * r * np.cos(x), 0.04 * np.cos(2 * x), 0.2 - 0.3 * np.sin(x * r)], 1)


def cyl_u(x, r):
    """Velocity in the (x, r, phi) frame, swirl free."""
    return np.stack([1.0 + 0.5 * np.sin(np.pi * x) * (r - R0) * (R1 - r),
                     0.4 / r + 0.3 * np.cos(np.pi * x) * (r - R0), np.zeros_like(x)], 1)


def _sym(t):
    return np.array([[t[0], t[1], t[2]], [t[1], t[3], t[4]], [t[2], t[4], t[5]]])


def cart_fields(P):
    """Axisymmetric theta and U at the points P, in Cartesian components."""
    x, y, z = P[:, 0], P[:, 1], P[:, 2]
    r, ph = np.hypot(y, z), np.arctan2(z, y)
    tc, uc = cyl_theta(x, r), cyl_u(x, r)
    th = np.empty_like(tc)
    U = np.empty_like(uc)
    for i in range(len(P)):
        R = rot(ph[i])
        A = R @ _sym(tc[i]) @ R.T
        th[i] = A[[0, 0, 0, 1, 1, 2], [0, 1, 2, 1, 2, 2]]
        U[i] = R @ uc[i]
    return th, U


def fields(m, wedge_u="rotate"):
    th, U = cart_fields(m.C)
    w, V = np.linalg.eigh(np.array([_sym(t) for t in th]))
    ev = np.zeros((m.nCells, 3, 3))
    ev[:, [0, 1, 2], [0, 1, 2]] = np.exp(w)
    F, NB = m.nInternal, m.nBoundary
    _, Ub = cart_fields(m.Cf[F:])
    for p, name in enumerate(m.patchNames):
        if name not in ("front", "back"):
            continue
        s, n = int(m.patchStart[p]), int(m.patchSize[p])
        T = O.wedge_face_t(m, p)
        Uc = U[m.owner[F + s:F + s + n]]
        Ub[s:s + n] = {"rotate": Uc @ T.T, "cell": Uc, "backwards": Uc @ T}[wedge_u]
    phi = np.empty(F + NB)
    phi[:F] = np.sum((m.w[:, None] * (U[m.owner[:F]] - U[m.neighbour]) + U[m.neighbour]) * m.Sf[:F], axis=1)
    phi[F:] = np.sum(Ub * m.Sf[F:], axis=1)
    from rheotool_amd import cases
    return cases.Fields(U=U, Ub=Ub, phi=phi, gradU=None, theta=th, eigVals=ev, eigVecs=V)


def step(m, scheme, wedge_bc=O.BC_WEDGE, wedge_u="rotate"):
    bc = O.default_bc(m)
    for p, name in enumerate(m.patchNames):
        if name in ("front", "back"):
            bc[p] = wedge_bc
    p = O.params("GiesekusLog", lam=0.8, alpha=0.2)
    th, _, _, tau, _ = O.correct(m, p, fields(m, wedge_u), 0.02, ctl=O.controls(tol=1e-14, maxIter=500),
                                 bc=bc, scheme=scheme)
    return th


def test_face_t_is_the_rotation_by_the_half_angle():
    for a_deg in (2.5, 7.0):
        m = mesh(a_deg, 1)
        a = np.radians(a_deg)
        for name, sign in (("front", 1.0), ("back", -1.0)):
            p = m.patchNames.index(name)
            T = O.wedge_face_t(m, p)
            assert np.abs(T - rot(sign * a)).max() < 1e-14
            s, n = m.nInternal + m.patchStart[p], m.patchSize[p]
            assert np.abs(stressfoam.wedge_face_t(m.Sf[s:s + n]) - T).max() < 1e-15
    # a patch in a coordinate plane has no angle, and is refused
    m = mesh(2.5, 1)
    with pytest.raises(ValueError):
        O.wedge_face_t(m, m.patchNames.index("inlet"))


def test_wedge_coefficients_closed_form():
    """With a prescribed flux through the wedge faces the upwind coefficients
    show: internalCoeffs phi (1 - d_i d_j), boundaryCoeffs -phi (R theta_P
    R^T - (1 - d_i d_j) theta_P), R = rot(+-a), d = (0, sin^2 a, sin^2 a)."""
    a_deg = 6.0
    a = np.radians(a_deg)
    m = mesh(a_deg, 1)
    f = fields(m)
    bc = O.default_bc(m)
    for p, name in enumerate(m.patchNames):
        if name in ("front", "back"):
            bc[p] = O.BC_WEDGE
    F = m.nInternal
    phi = f.phi.copy()
    rng = np.random.default_rng(5)
    phi[F:] = rng.uniform(-1, 1, m.nBoundary)
    out = O.assemble(m, phi, f.theta, np.zeros((m.nCells, 6)), 0.01, bc=bc)
    d = np.array([0.0, np.sin(a) ** 2, np.sin(a) ** 2])
    ij = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    vic = np.array([1 - d[i] * d[j] for i, j in ij])
    for name, sign in (("front", 1.0), ("back", -1.0)):
        p = m.patchNames.index(name)
        R = rot(sign * a)
        for k in range(m.patchSize[p]):
            b = m.patchStart[p] + k
            tP = f.theta[m.owner[F + b]]
            A = R @ _sym(tP) @ R.T
            tb = A[[0, 0, 0, 1, 1, 2], [0, 1, 2, 1, 2, 2]]
            assert np.allclose(out["iC"][b], phi[F + b] * vic, rtol=0, atol=1e-14)
            assert np.allclose(out["bC"][b], -phi[F + b] * (tb - vic * tP), rtol=0, atol=1e-13)


def _ring_error(a_deg, scheme, **kw):
    w = mesh(a_deg, 1)
    ring = mesh(a_deg, int(round(180.0 / a_deg)))
    tw = step(w, scheme, **kw)
    tr = step(ring, scheme)
    key = lambda C: np.round(C * 1e8).astype(np.int64)
    idx = {tuple(k): i for i, k in enumerate(key(ring.C))}
    match = np.array([idx[tuple(k)] for k in key(w.C)])
    return np.abs(tw - tr[match]).max()


@pytest.mark.parametrize("scheme", ["upwind", "cubista"])
def test_wedge_converges_to_the_ring_at_second_order(scheme):
    e1 = _ring_error(6.0, scheme)
    e2 = _ring_error(3.0, scheme)
    assert e1 < 2e-3
    assert 3.0 < e1 / e2 < 5.0, (e1, e2)


@pytest.mark.parametrize("variant", [dict(wedge_bc=O.BC_ZEROGRADIENT, wedge_u="cell"),
                                     dict(wedge_u="backwards")])
def test_wedge_without_the_rotation_is_caught(variant):
    """The cell value on the wedge faces (zeroGradient), or the rotation the
    wrong way, loses the hoop terms: the difference to the ring stays."""
    e1 = _ring_error(6.0, "upwind", **variant)
    e2 = _ring_error(3.0, "upwind", **variant)
    assert e2 > 0.5 * e1 and e2 > 1e-2, (e1, e2)
