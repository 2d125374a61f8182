/*
 * rtcase.h -- synthetic stress-step cases (meshes + seeded fields).
 *
 * This is the blockMesh / setFields side of the stress step: it builds the
 * LDU-addressed finite-volume meshes and the seeded U, phi, theta and eigen
 * state that both the HIP library and the CPU oracle consume.  Host-only C++.
 *
 * Mesh conventions follow OpenFOAM's primitiveMesh (the layer the reference's
 * constitutiveEq code runs on, e.g. the addressing used by
 * of90/src/libs/sparseMatrixSolvers/segregated/eigenSolver/eigenSolver.C:600-641):
 *   - faces [0, nInternalFaces) are internal, owner < neighbour, ordered by
 *     owner then neighbour (upper-triangular order);
 *   - boundary faces follow, grouped by patch; owner[] of a boundary face is
 *     its faceCell;
 *   - Sf points out of the owner;
 *   - face centres/areas by triangle decomposition about the point average,
 *     cell centres/volumes by face-pyramid decomposition about the average of
 *     the face centres, and linear weights w = |Sf.(C_N-Cf)| / (|Sf.(Cf-C_P)|
 *     + |Sf.(C_N-Cf)|).
 *
 * Randomness comes from a counter-based SplitMix64 keyed by (seed, stream,
 * original cell id), so fields do not depend on any cell permutation.
 */
#ifndef RTCASE_H
#define RTCASE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rtc_mesh rtc_mesh;

enum rtc_patch_kind { RTC_PATCH = 0, RTC_WALL = 1, RTC_EMPTY = 2 };
enum rtc_field_kind {
    RTC_FIELDS_CHANNEL = 0, /* Poiseuille + streamfunction noise; theta = log(I + 0.1 A A^T) */
    RTC_FIELDS_TG = 1       /* Taylor-Green + uniform shear; theta = R diag(ln l) R^T       */
};

/* Cartesian block [0,Lx]x[0,Ly]x[0,Lz], nx*ny*nz hexahedra, cell id
 * i + nx*(j + ny*k).  twoD != 0: nz must be 1 and the z faces form one empty
 * patch ("frontAndBack").  channel != 0: patches inlet(xmin) outlet(xmax)
 * walls(ymin+ymax) [+ frontAndBack]; else xmin xmax ymin ymax zmin zmax.
 * jitter > 0 moves every interior node by U[-jitter,jitter]*spacing per axis.
 * permSeed != 0 renumbers cells by a seeded Fisher-Yates permutation and
 * re-sorts the internal faces into upper-triangular order. */
rtc_mesh* rtc_mesh_cartesian(int nx, int ny, int nz, double Lx, double Ly, double Lz,
                             int twoD, int channel, double jitter, uint64_t jitterSeed,
                             uint64_t permSeed);

/* Random face graph: nCells cells at seeded random positions in [0,1]^3,
 * each joined to its kNear nearest cells (union), |Sf| log-normal with the
 * given sigma about nCells^(-2/3), Sf along C_N - C_P, V = 1/nCells, w = 1/2.
 * No boundary faces. */
rtc_mesh* rtc_mesh_random_graph(int nCells, int kNear, double sigma, uint64_t seed);

/* A mesh from OpenFOAM polyMesh data (constant/polyMesh points, faces,
 * owner, neighbour, boundary).  Face f has the points
 * faceLabels[faceStart[f] .. faceStart[f+1]); faces [0, nInternal) are
 * internal (neighbour[f] given), the rest boundary faces in patch order with
 * patch p covering [patchStart[p], patchStart[p] + patchSize[p]) counted from
 * face nInternal.  patchNames holds nPatches NUL-terminated names back to
 * back.  Geometry as the conventions above.  Returns NULL on an inconsistent
 * description (an index out of range, owner >= neighbour, a face with fewer
 * than 3 points, patches that do not tile the boundary faces). */
rtc_mesh* rtc_mesh_from_poly(int32_t nPoints, const double* points, int32_t nFaces,
                             const int32_t* faceStart, const int32_t* faceLabels,
                             const int32_t* owner, int32_t nInternal, const int32_t* neighbour,
                             int32_t nCells, int32_t nPatches, const int32_t* patchStart,
                             const int32_t* patchSize, const int32_t* patchKind,
                             const char* patchNames);

void rtc_mesh_free(rtc_mesh* m);

/* out[0..3] = nCells, nInternalFaces, nBoundaryFaces, nPatches */
void rtc_mesh_sizes(const rtc_mesh* m, int32_t out[4]);

/* Any pointer may be NULL.  Sizes: owner[nFaces], neighbour[nInternal],
 * Sf/Cf[3*nFaces] (AoS xyz), w[nInternal], V[nCells], C[3*nCells],
 * patchStart/patchSize/patchKind[nPatches] (start counted from the first
 * boundary face), origId[nCells] (original Cartesian id of each cell). */
void rtc_mesh_export(const rtc_mesh* m, int32_t* owner, int32_t* neighbour, double* Sf,
                     double* Cf, double* w, double* V, double* C, int32_t* patchStart,
                     int32_t* patchSize, int32_t* patchKind, int32_t* origId);

const char* rtc_mesh_patch_name(const rtc_mesh* m, int patch);

/* Seeded fields on a mesh.  Outputs (any may be NULL):
 *   U[3*nCells], Ub[3*nBoundary] (analytic at boundary face centres),
 *   phi[nFaces] (U_f.Sf, U_f linear on internal faces, analytic on boundary
 *   faces, 0 on empty faces), gradU[9*nCells] (analytic, row i = d/dx_i),
 *   theta[6*nCells] (xx xy xz yy yz zz), eigVals[9*nCells] and
 *   eigVecs[9*nCells] (row-major tensors; eigVals diagonal holds exp of the
 *   ascending eigenvalues of theta, eigVecs columns the eigenvectors).
 * CHANNEL: U_x = 1.5(1-(2y-1)^2) + noise of max magnitude noiseAmp from
 *   psi = sum_{k,l=1..4} a_kl sin(pi k y) sin(2 pi l x + p_kl).
 *   theta = log(I + 0.1 A A^T), A_ij ~ N(0,1).
 * TG: U = TG(2 pi, seeded phases, amplitude 1) + (y - 1/2, 0, 0);
 *   theta = R diag(ln l) R^T, ln l_i ~ U[lnLo, lnHi], R uniform on SO(3).
 * Returns 0 on success. */
int rtc_fields(const rtc_mesh* m, int kind, uint64_t seed, double noiseAmp, double lnLo,
               double lnHi, double* U, double* Ub, double* phi, double* gradU,
               double* theta, double* eigVals, double* eigVecs);

/* Symmetric 3x3 eigen-decomposition used by the generator (cyclic Jacobi,
 * FP64): theta (6 comps) -> ascending eigenvalues ev[3], column vectors
 * R[9] row-major.  Exposed for tests. */
void rtc_sym_eig(const double theta[6], double ev[3], double R[9]);

#ifdef __cplusplus
}
#endif
#endif
