// oracle.cpp -- CPU FP64 restatement of rheoTool's log-conformation stress step.
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never part of it.
// Build with -O2 -ffp-contract=off so every expression rounds exactly as
// written (no fused multiply-add), matching the reference's plain C++ loops.
//
// Reference files (paths relative to of90/src/libs/):
//   constitutiveEquations/constitutiveEqs/utils/boilerLog.H
//   constitutiveEquations/constitutiveEqs/constitutiveEq/constitutiveEq.C
//   constitutiveEquations/constitutiveEqs/{Oldroyd-B,Giesekus,FENE-P,PTT}/*Log/*.C
//   gaussDefCmpwConvectionScheme/gaussDefCmpwConvectionScheme.C
//   sparseMatrixSolvers/segregated/{sparseSolver.C,eigenSolver/eigenSolver.C}
//   constitutiveEquations/constitutiveEqs/utils/jacobi.H
// OpenFOAM-9 / Eigen 3.2.9 behaviour (third party, absent from the reference
// tree) is restated from their published algorithms and marked [OF9]/[Eigen].
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <omp.h>
#include <vector>

namespace {

typedef double T9[9];

// OpenFOAM tensor inner product t1 & t2 (row-major), sum order k = 0,1,2.
inline void dotTT(const double* a, const double* b, double* c) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            c[3 * i + j] = a[3 * i] * b[j] + a[3 * i + 1] * b[3 + j] + a[3 * i + 2] * b[6 + j];
}
inline void transp(const double* a, double* t) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) t[3 * i + j] = a[3 * j + i];
}
inline void expandSymm(const double* s, double* t) {
    t[0] = s[0]; t[1] = s[1]; t[2] = s[2];
    t[3] = s[1]; t[4] = s[3]; t[5] = s[4];
    t[6] = s[2]; t[7] = s[4]; t[8] = s[5];
}
// [OF9] symm(tensor)
inline void symmT(const double* t, double* s) {
    s[0] = t[0];
    s[1] = 0.5 * (t[1] + t[3]);
    s[2] = 0.5 * (t[2] + t[6]);
    s[3] = t[4];
    s[4] = 0.5 * (t[5] + t[7]);
    s[5] = t[8];
}
// innerP (constitutiveEq.C:471-518): isFirstT -> t1^T & t2 & t1, else t1 & t2 & t1^T
inline void innerP(const double* t1, const double* t2, bool isFirstT, double* r) {
    T9 tt, tmp;
    transp(t1, tt);
    if (isFirstT) {
        dotTT(tt, t2, tmp);
        dotTT(tmp, t1, r);
    } else {
        dotTT(t1, t2, tmp);
        dotTT(tmp, tt, r);
    }
}
// [OF9] det(tensor) and inv(tensor) = cofactor / det
inline double det9(const double* t) {
    return t[0] * t[4] * t[8] + t[1] * t[5] * t[6] + t[2] * t[3] * t[7] - t[0] * t[5] * t[7] -
           t[1] * t[3] * t[8] - t[2] * t[4] * t[6];
}
inline void inv9(const double* t, double* r) {
    const double d = det9(t);
    r[0] = (t[4] * t[8] - t[7] * t[5]) / d;
    r[1] = (t[2] * t[7] - t[1] * t[8]) / d;
    r[2] = (t[1] * t[5] - t[2] * t[4]) / d;
    r[3] = (t[6] * t[5] - t[3] * t[8]) / d;
    r[4] = (t[0] * t[8] - t[2] * t[6]) / d;
    r[5] = (t[3] * t[2] - t[0] * t[5]) / d;
    r[6] = (t[3] * t[7] - t[6] * t[4]) / d;
    r[7] = (t[6] * t[1] - t[0] * t[7]) / d;
    r[8] = (t[0] * t[4] - t[1] * t[3]) / d;
}
const double I9[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};

// [OF9] inv(Field<tensor>) first-cell shift: if the first cell has a diagonal
// component with |T_ii|^2/|T|^2 < small(1e-15), 1 is added to that component
// of every cell before inverting and subtracted after.
void invShiftMask(const double* first, bool mask[3]) {
    double s = 0;
    for (int q = 0; q < 9; ++q) s += first[q] * first[q];
    mask[0] = first[0] * first[0] / s < 1e-15;
    mask[1] = first[4] * first[4] / s < 1e-15;
    mask[2] = first[8] * first[8] / s < 1e-15;
}
inline void invField(const double* t, const bool mask[3], double* r) {
    if (!(mask[0] || mask[1] || mask[2])) {
        inv9(t, r);
        return;
    }
    T9 tp;
    for (int q = 0; q < 9; ++q) tp[q] = t[q];
    if (mask[0]) tp[0] += 1.0;
    if (mask[1]) tp[4] += 1.0;
    if (mask[2]) tp[8] += 1.0;
    inv9(tp, r);
    if (mask[0]) r[0] -= 1.0;
    if (mask[1]) r[4] -= 1.0;
    if (mask[2]) r[8] -= 1.0;
}

// ---------------- eigen-decomposition restatements ----------------------

// [Eigen 3.2.9] SelfAdjointEigenSolver<Matrix3d>::compute restated from the
// algorithm: scale by max|a_ij|, Householder reduction to tridiagonal form,
// implicit symmetric QR steps with Wilkinson shift (Golub & Van Loan 8.3.2),
// deflation when |e_i| <= 2 eps (|d_i| + |d_i+1|), ascending sort.
void eigQR(const double* th, double ev[3], double Q[9]) {
    double a[3][3] = {{th[0], th[1], th[2]}, {th[1], th[3], th[4]}, {th[2], th[4], th[5]}};
    double scale = 0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) scale = std::max(scale, std::fabs(a[i][j]));
    if (scale == 0) scale = 1;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) a[i][j] /= scale;
    double d[3], e[2];
    double Z[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    d[0] = a[0][0];
    if (a[2][0] == 0.0) {
        d[1] = a[1][1];
        d[2] = a[2][2];
        e[0] = a[1][0];
        e[1] = a[2][1];
    } else {
        // reflector H = [[c, s],[s,-c]] on rows/cols 1,2 maps (a10, a20) to (b, 0)
        const double b = std::sqrt(a[1][0] * a[1][0] + a[2][0] * a[2][0]);
        const double c = a[1][0] / b, s = a[2][0] / b;
        const double p = a[1][1], q = a[2][2], r = a[2][1];
        d[1] = c * c * p + 2.0 * c * s * r + s * s * q;
        d[2] = s * s * p - 2.0 * c * s * r + c * c * q;
        e[0] = b;
        e[1] = c * s * (p - q) + (s * s - c * c) * r;
        Z[1][1] = c; Z[1][2] = s;
        Z[2][1] = s; Z[2][2] = -c;
    }
    const double eps = std::numeric_limits<double>::epsilon();
    const double tiny = std::numeric_limits<double>::min();
    // dense symmetric tridiagonal T = Z^T A Z; rotations applied as T <- G^T T G
    double T[3][3] = {{d[0], e[0], 0.0}, {e[0], d[1], e[1]}, {0.0, e[1], d[2]}};
    for (int iter = 0; iter < 90; ++iter) {
        for (int i = 0; i < 2; ++i)
            if (std::fabs(T[i + 1][i]) <= 2.0 * eps * (std::fabs(T[i][i]) + std::fabs(T[i + 1][i + 1])) ||
                std::fabs(T[i + 1][i]) <= tiny)
                T[i + 1][i] = T[i][i + 1] = 0.0;
        int end = 2;
        while (end > 0 && T[end][end - 1] == 0.0) --end;
        if (end == 0) break;
        int start = end - 1;
        while (start > 0 && T[start][start - 1] != 0.0) --start;
        // Wilkinson shift from the trailing 2x2 of the active block
        const double ta = T[end - 1][end - 1], tc = T[end][end], tb = T[end][end - 1];
        const double td = (ta - tc) * 0.5;
        double mu;
        if (td == 0.0) {
            mu = tc - std::fabs(tb);
        } else {
            const double h = std::hypot(td, tb);
            mu = tc - tb * tb / (td + (td > 0 ? h : -h));
        }
        double x = T[start][start] - mu, z = T[start + 1][start];
        for (int k = start; k < end; ++k) {
            // G = [[c, s], [-s, c]] on (k, k+1) with G^T [x; z] = [r; 0]
            double c = 1.0, s = 0.0;
            if (z != 0.0) {
                const double r = std::hypot(x, z);
                c = x / r;
                s = -z / r;
            }
            for (int j = 0; j < 3; ++j) {
                const double tk = T[k][j], tk1 = T[k + 1][j];
                T[k][j] = c * tk - s * tk1;
                T[k + 1][j] = s * tk + c * tk1;
            }
            for (int i = 0; i < 3; ++i) {
                const double tk = T[i][k], tk1 = T[i][k + 1];
                T[i][k] = c * tk - s * tk1;
                T[i][k + 1] = s * tk + c * tk1;
            }
            for (int r = 0; r < 3; ++r) {
                const double zk = Z[r][k], zk1 = Z[r][k + 1];
                Z[r][k] = c * zk - s * zk1;
                Z[r][k + 1] = s * zk + c * zk1;
            }
            if (k > start) T[k + 1][k - 1] = T[k - 1][k + 1] = 0.0;  // bulge annihilated
            if (k + 1 < end) {
                x = T[k + 1][k];
                z = T[k + 2][k];
            }
        }
        T[2][0] = T[0][2] = 0.0;
    }
    for (int i = 0; i < 3; ++i) d[i] = T[i][i];
    int idx[3] = {0, 1, 2};
    // selection sort, ascending (stable for ties)
    for (int i = 0; i < 2; ++i) {
        int k = i;
        for (int j = i + 1; j < 3; ++j)
            if (d[idx[j]] < d[idx[k]]) k = j;
        std::swap(idx[i], idx[k]);
    }
    for (int c = 0; c < 3; ++c) {
        ev[c] = d[idx[c]] * scale;
        for (int r = 0; r < 3; ++r) Q[3 * r + c] = Z[r][idx[c]];
    }
}

// utils/jacobi.H:7-158 restated: cyclic Jacobi with the threshold strategy
// (tresh = 0.2 sm^2 for the first three sweeps, skip of negligible
// off-diagonals after four), at most 50 sweeps.  Returns the eigenvalues of
// theta (not their exponential) in Jacobi order.
void eigJacobiH(const double* th, double D[3], double Vout[9]) {
    double A[3][3] = {{th[0], th[1], th[2]}, {th[1], th[3], th[4]}, {th[2], th[4], th[5]}};
    double V[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    double B[3], Zs[3];
    for (int p = 0; p < 3; ++p) {
        B[p] = D[p] = A[p][p];
        Zs[p] = 0;
    }
    for (int sweep = 1; sweep <= 50; ++sweep) {
        double sm = 0;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) sm += std::fabs(A[p][q]);
        const double tresh = sweep < 4 ? 0.2 * sm * sm : 0.0;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                const double g = 100.0 * std::fabs(A[p][q]);
                if (sweep > 4 && std::fabs(D[p]) + g == std::fabs(D[p]) &&
                    std::fabs(D[q]) + g == std::fabs(D[q])) {
                    A[p][q] = 0;
                } else if (std::fabs(A[p][q]) > tresh) {
                    double h = D[q] - D[p], t;
                    if (std::fabs(h) + g == std::fabs(h)) {
                        t = A[p][q] / h;
                    } else {
                        const double thet = 0.5 * h / A[p][q];
                        t = 1.0 / (std::fabs(thet) + std::sqrt(1.0 + thet * thet));
                        if (thet < 0) t = -t;
                    }
                    const double c = 1.0 / std::sqrt(1 + t * t), s = t * c, tau = s / (1.0 + c);
                    h = t * A[p][q];
                    Zs[p] -= h; Zs[q] += h;
                    D[p] -= h; D[q] += h;
                    A[p][q] = 0;
                    auto rot = [&](double& x, double& y) {
                        const double gg = x, hh = y;
                        x = gg - s * (hh + gg * tau);
                        y = hh + s * (gg - hh * tau);
                    };
                    for (int j = 0; j < p; ++j) rot(A[j][p], A[j][q]);
                    for (int j = p + 1; j < q; ++j) rot(A[p][j], A[j][q]);
                    for (int j = q + 1; j < 3; ++j) rot(A[p][j], A[q][j]);
                    for (int j = 0; j < 3; ++j) rot(V[j][p], V[j][q]);
                }
            }
        for (int p = 0; p < 3; ++p) {
            B[p] += Zs[p];
            D[p] = B[p];
            Zs[p] = 0;
        }
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Vout[3 * i + j] = V[i][j];
}

// ---------------- [OF9] lduMatrix primitives -----------------------------
struct Ldu {
    int N, F;
    const int32_t *l, *u;
    const double *diag, *lower, *upper;
};

void Amul(const Ldu& A, const double* x, double* y) {
    for (int c = 0; c < A.N; ++c) y[c] = A.diag[c] * x[c];
    for (int f = 0; f < A.F; ++f) {
        y[A.u[f]] += A.lower[f] * x[A.l[f]];
        y[A.l[f]] += A.upper[f] * x[A.u[f]];
    }
}
void Tmul(const Ldu& A, const double* x, double* y) {
    for (int c = 0; c < A.N; ++c) y[c] = A.diag[c] * x[c];
    for (int f = 0; f < A.F; ++f) {
        y[A.u[f]] += A.upper[f] * x[A.l[f]];
        y[A.l[f]] += A.lower[f] * x[A.u[f]];
    }
}
void sumA(const Ldu& A, double* s) {
    for (int c = 0; c < A.N; ++c) s[c] = A.diag[c];
    for (int f = 0; f < A.F; ++f) {
        s[A.u[f]] += A.lower[f];
        s[A.l[f]] += A.upper[f];
    }
}
double sumMag(const std::vector<double>& v) {
    double s = 0;
    for (double x : v) s += std::fabs(x);
    return s;
}
double sumProd(const double* a, const double* b, int n) {
    double s = 0;
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}
// [OF9] lduMatrix::solver::normFactor
double normFactor(const Ldu& A, const double* psi, const double* b, const double* Apsi) {
    std::vector<double> t(A.N);
    sumA(A, t.data());
    double avg = 0;
    for (int c = 0; c < A.N; ++c) avg += psi[c];
    avg /= A.N;
    for (int c = 0; c < A.N; ++c) t[c] *= avg;
    double s = 0;
    for (int c = 0; c < A.N; ++c) s += std::fabs(Apsi[c] - t[c]) + std::fabs(b[c] - t[c]);
    return s + 1e-20;
}
bool converged(double finalRes, double initRes, const orc_controls* c) {
    return finalRes < c->tolerance || (c->relTol > 0 && finalRes < c->relTol * initRes);
}

// [OF9] preconditioners: none, diagonal, DILU (with losort forward sweep)
struct Precond {
    const Ldu& A;
    int kind;
    std::vector<double> rD;
    std::vector<int32_t> losort;
    Precond(const Ldu& a, int k) : A(a), kind(k) {
        if (kind == ORC_PC_DIAGONAL) {
            rD.resize(A.N);
            for (int c = 0; c < A.N; ++c) rD[c] = 1.0 / A.diag[c];
        } else if (kind == ORC_PC_DILU) {
            rD.assign(A.diag, A.diag + A.N);
            for (int f = 0; f < A.F; ++f) rD[A.u[f]] -= A.upper[f] * A.lower[f] / rD[A.l[f]];
            for (int c = 0; c < A.N; ++c) rD[c] = 1.0 / rD[c];
            // faces sorted by upper address, stable in face index
            std::vector<int32_t> cnt(A.N + 1, 0);
            for (int f = 0; f < A.F; ++f) cnt[A.u[f] + 1]++;
            for (int c = 0; c < A.N; ++c) cnt[c + 1] += cnt[c];
            losort.resize(A.F);
            for (int f = 0; f < A.F; ++f) losort[cnt[A.u[f]]++] = f;
        }
    }
    void apply(double* w, const double* r) const {
        if (kind == ORC_PC_NONE) {
            std::memcpy(w, r, sizeof(double) * A.N);
        } else if (kind == ORC_PC_DIAGONAL) {
            for (int c = 0; c < A.N; ++c) w[c] = rD[c] * r[c];
        } else {
            for (int c = 0; c < A.N; ++c) w[c] = rD[c] * r[c];
            for (int k = 0; k < A.F; ++k) {
                const int f = losort[k];
                w[A.u[f]] -= rD[A.u[f]] * A.lower[f] * w[A.l[f]];
            }
            for (int f = A.F - 1; f >= 0; --f) w[A.l[f]] -= rD[A.l[f]] * A.upper[f] * w[A.u[f]];
        }
    }
    void applyT(double* w, const double* r) const {
        if (kind != ORC_PC_DILU) {
            apply(w, r);
            return;
        }
        for (int c = 0; c < A.N; ++c) w[c] = rD[c] * r[c];
        for (int k = 0; k < A.F; ++k) {
            const int f = losort[k];
            w[A.u[f]] -= rD[A.u[f]] * A.upper[f] * w[A.l[f]];
        }
        for (int f = A.F - 1; f >= 0; --f) w[A.l[f]] -= rD[A.l[f]] * A.lower[f] * w[A.u[f]];
    }
};

// [OF9] PBiCGStab::solve (van der Vorst BiCGStab, OpenFOAM residual control)
orc_perf pbicgstab(const Ldu& A, const double* b, const orc_controls* ctl, double* psi) {
    const int N = A.N;
    orc_perf perf{0, 0, 0, 0, 1};
    std::vector<double> yA(N), pA(N), rA(N);
    Amul(A, psi, yA.data());
    for (int c = 0; c < N; ++c) rA[c] = b[c] - yA[c];
    const double nf = normFactor(A, psi, b, yA.data());
    perf.initialResidual = sumMag(rA) / nf;
    perf.finalResidual = perf.initialResidual;
    if (ctl->minIter > 0 || !converged(perf.finalResidual, perf.initialResidual, ctl)) {
        std::vector<double> AyA(N), sA(N), zA(N), tA(N), rA0(rA);
        double rA0rA = 0, alpha = 0, omega = 0;
        Precond P(A, ctl->precond);
        do {
            const double rA0rAold = rA0rA;
            rA0rA = sumProd(rA0.data(), rA.data(), N);
            if (std::fabs(rA0rA) < 1e-300) break;
            if (perf.nIterations == 0) {
                pA = rA;
            } else {
                if (std::fabs(omega) < 1e-300) break;
                const double beta = (rA0rA / rA0rAold) * (alpha / omega);
                for (int c = 0; c < N; ++c) pA[c] = rA[c] + beta * (pA[c] - omega * AyA[c]);
            }
            P.apply(yA.data(), pA.data());
            Amul(A, yA.data(), AyA.data());
            const double rA0AyA = sumProd(rA0.data(), AyA.data(), N);
            alpha = rA0rA / rA0AyA;
            for (int c = 0; c < N; ++c) sA[c] = rA[c] - alpha * AyA[c];
            perf.finalResidual = sumMag(sA) / nf;
            if (converged(perf.finalResidual, perf.initialResidual, ctl)) {
                for (int c = 0; c < N; ++c) psi[c] += alpha * yA[c];
                perf.nIterations++;
                perf.converged = 1;
                return perf;
            }
            P.apply(zA.data(), sA.data());
            Amul(A, zA.data(), tA.data());
            const double tAtA = sumProd(tA.data(), tA.data(), N);
            omega = sumProd(tA.data(), sA.data(), N) / tAtA;
            for (int c = 0; c < N; ++c) {
                psi[c] += alpha * yA[c] + omega * zA[c];
                rA[c] = sA[c] - omega * tA[c];
            }
            perf.finalResidual = sumMag(rA) / nf;
        } while ((++perf.nIterations < ctl->maxIter &&
                  !converged(perf.finalResidual, perf.initialResidual, ctl)) ||
                 perf.nIterations < ctl->minIter);
    }
    perf.converged = converged(perf.finalResidual, perf.initialResidual, ctl) ? 1 : 0;
    return perf;
}

// [OF9] PBiCG::solve (bi-conjugate gradients, the tutorials' solver)
orc_perf pbicg(const Ldu& A, const double* b, const orc_controls* ctl, double* psi) {
    const int N = A.N;
    orc_perf perf{0, 0, 0, 0, 1};
    std::vector<double> pA(N), wA(N), qA(N), rA(N), rT(N);
    Amul(A, psi, wA.data());
    Tmul(A, psi, qA.data());
    for (int c = 0; c < N; ++c) {
        rA[c] = b[c] - wA[c];
        rT[c] = b[c] - qA[c];
    }
    const double nf = normFactor(A, psi, b, wA.data());
    perf.initialResidual = sumMag(rA) / nf;
    perf.finalResidual = perf.initialResidual;
    if (ctl->minIter > 0 || !converged(perf.finalResidual, perf.initialResidual, ctl)) {
        std::vector<double> pT(N, 0.0), wT(N);
        double wArT = 1e20, wArTold;
        Precond P(A, ctl->precond);
        do {
            wArTold = wArT;
            P.apply(wA.data(), rA.data());
            P.applyT(wT.data(), rT.data());
            wArT = sumProd(wA.data(), rT.data(), N);
            if (perf.nIterations == 0) {
                pA = wA;
                pT = wT;
            } else {
                const double beta = wArT / wArTold;
                for (int c = 0; c < N; ++c) {
                    pA[c] = wA[c] + beta * pA[c];
                    pT[c] = wT[c] + beta * pT[c];
                }
            }
            Amul(A, pA.data(), wA.data());
            Tmul(A, pT.data(), wT.data());
            const double wApT = sumProd(wA.data(), pT.data(), N);
            if (std::fabs(wApT) / nf < 1e-300) break;
            const double alpha = wArT / wApT;
            for (int c = 0; c < N; ++c) {
                psi[c] += alpha * pA[c];
                rA[c] -= alpha * wA[c];
                rT[c] -= alpha * wT[c];
            }
            perf.finalResidual = sumMag(rA) / nf;
        } while ((++perf.nIterations < ctl->maxIter &&
                  !converged(perf.finalResidual, perf.initialResidual, ctl)) ||
                 perf.nIterations < ctl->minIter);
    }
    perf.converged = converged(perf.finalResidual, perf.initialResidual, ctl) ? 1 : 0;
    return perf;
}

}  // namespace

extern "C" {

void orc_grad_gauss(const orc_mesh* m, const double* U, const double* Ub, double* g) {
    const int N = m->nCells, F = m->nInternal;
    std::fill(g, g + 9 * (size_t)N, 0.0);
    for (int f = 0; f < F; ++f) {
        const int o = m->owner[f], n = m->neighbour[f];
        double uf[3];
        for (int d = 0; d < 3; ++d) uf[d] = m->w[f] * (U[3 * o + d] - U[3 * n + d]) + U[3 * n + d];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                const double v = m->Sf[3 * f + i] * uf[j];
                g[9 * (size_t)o + 3 * i + j] += v;
                g[9 * (size_t)n + 3 * i + j] -= v;
            }
    }
    for (int p = 0; p < m->nPatches; ++p) {
        if (m->patchEmpty[p]) continue;
        for (int k = 0; k < m->patchSize[p]; ++k) {
            const int b = m->patchStart[p] + k;
            const int64_t f = (int64_t)F + b;
            const int c = m->owner[f];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) g[9 * (size_t)c + 3 * i + j] += m->Sf[3 * f + i] * Ub[3 * b + j];
        }
    }
    for (int c = 0; c < N; ++c)
        for (int q = 0; q < 9; ++q) g[9 * (size_t)c + q] /= m->V[c];
}

void orc_rhs(const orc_params* p, int32_t n, const double* gradU, const double* theta,
             const double* eigVals, const double* eigVecs, double* S, double* fOut) {
    orc_rhs_tau(p, n, gradU, theta, eigVals, eigVecs, nullptr, nullptr, S, fOut);
}

// PTTLog.C:137-152: gamma(beta), then gamma(alpha k + beta) for k = 0, 1, ...
// while k < maxIterMittagLeffler and the last value stays below 1e100; the
// series then runs over as many terms as the table holds.
static std::vector<double> mlGammaTable(const orc_params* p, int* maxIter) {
    std::vector<double> g{std::tgamma(p->mlBeta)};
    int k = 0;
    while (k < p->mlMaxIter && g.back() < 1e+100) {
        g.push_back(std::tgamma(p->mlAlpha * k + p->mlBeta));
        k++;
    }
    *maxIter = k;
    return g;
}

// PTTLog.C:207-227: E_{alpha,beta}(z) summed term by term until the relative
// change of the sum falls to rTolMittagLeffler (or the table ends)
static double mittagLeffler(const std::vector<double>& g, int maxIter, double rtol, double z) {
    double sum = 0.0, sumOld = 0.0, error = 1.0;
    int k = 0;
    while (k < maxIter && error > rtol) {
        const double Eabk = std::pow(z, k) / g[k + 1];
        sumOld = sum;
        sum += Eabk;
        error = std::fabs((sumOld - sum) / (sumOld + 1e-12));
        k++;
    }
    return sum;
}

// SaramitoLog.C:181-212: the plasticity factor from the second invariant of
// the deviatoric part of tau (over the valid dimensions)
static double saramitoFac(const orc_params* p, const double* tau) {
    double t[6] = {0, 0, 0, 0, 0, 0};
    if (tau)
        for (int q = 0; q < 6; ++q) t[q] = tau[q];
    const double nDims = p->dimX + p->dimY + p->dimZ;
    const double tr = t[0] + t[3] + t[5];
    double d[6] = {t[0] - (p->dimX * tr) / nDims, t[1] - (0.0 * tr) / nDims, t[2] - (0.0 * tr) / nDims,
                   t[3] - (p->dimY * tr) / nDims, t[4] - (0.0 * tr) / nDims, t[5] - (p->dimZ * tr) / nDims};
    const double magSqr = d[0] * d[0] + 2 * (d[1] * d[1]) + 2 * (d[2] * d[2]) + d[3] * d[3] + 2 * (d[4] * d[4]) +
                          d[5] * d[5];
    const double tauDMag = std::sqrt(magSqr) / std::sqrt(2.);
    const double small = 1e-16;
    if (p->sarN == 1.) return std::max(0., (tauDMag - p->tau0) / (p->sarK * tauDMag + small));
    return std::pow(std::max(0., (tauDMag - p->tau0) / (p->sarK * std::pow(tauDMag, p->sarN) * 1. + small)),
                    1. / p->sarN);
}

void orc_rhs_tau(const orc_params* p, int32_t n, const double* gradU, const double* theta,
                 const double* eigVals, const double* eigVecs, const double* tauOld, const double* aux,
                 double* S, double* fOut) {
    bool mask[3] = {false, false, false};
    if (p->ofInvQuirk && n > 0) invShiftMask(eigVals, mask);
    const bool ptt = p->model == ORC_PTT_LINEAR || p->model == ORC_PTT_EXP || p->model == ORC_PTT_GEN;
    int mlMax = 0;
    const std::vector<double> mlG = p->model == ORC_PTT_GEN ? mlGammaTable(p, &mlMax) : std::vector<double>();
    for (int c = 0; c < n; ++c) {
        const double* L = gradU + 9 * (size_t)c;
        if (p->model == ORC_FLUIDITY) {
            // BMPLog.C:152-160: Phi0/lambda + k (PhiInf - Phi) (tau && symm(L)); the
            // implicit -Sp(1/lambda, Phi) goes to the diagonal
            double t[6] = {0, 0, 0, 0, 0, 0}, d[6];
            if (tauOld)
                for (int q = 0; q < 6; ++q) t[q] = tauOld[6 * (size_t)c + q];
            symmT(L, d);
            const double tdd = t[0] * d[0] + 2 * (t[1] * d[1]) + 2 * (t[2] * d[2]) + t[3] * d[3] +
                               2 * (t[4] * d[4]) + t[5] * d[5];
            const double Phi = theta[6 * (size_t)c];
            double* s = S + 6 * (size_t)c;
            s[0] = p->Phi0 / p->lambda + (p->kBMP * (p->PhiInf - Phi)) * tdd;
            for (int q = 1; q < 6; ++q) s[q] = 0.0;
            continue;
        }
        const double* V = eigVecs + 9 * (size_t)c;
        const double* Lam = eigVals + 9 * (size_t)c;
        // boilerLog.H:25-31: M = V^T (L^T [- zeta symm(L)]) V
        T9 LT, M;
        transp(L, LT);
        if (ptt) {
            double sL[6];
            symmT(L, sL);
            T9 zs;
            expandSymm(sL, zs);
            for (int q = 0; q < 9; ++q) LT[q] = LT[q] - p->zeta * zs[q];
        }
        innerP(V, LT, true, M);
        // decomposeGradU (constitutiveEq.C:323-358)
        T9 Bt = {0, 0, 0, 0, 0, 0, 0, 0, 0}, Ot = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        Bt[0] = M[0];
        Bt[4] = M[4];
        Bt[8] = M[8];
        const double lx = Lam[0], ly = Lam[4], lz = Lam[8];
        Ot[1] = (ly * M[1] + lx * M[3]) / (ly - lx + 1e-16);
        Ot[2] = (lz * M[2] + lx * M[6]) / (lz - lx + 1e-16);
        Ot[5] = (lz * M[5] + ly * M[7]) / (lz - ly + 1e-16);
        Ot[3] = -Ot[1];
        Ot[6] = -Ot[2];
        Ot[7] = -Ot[5];
        T9 Om, B;
        innerP(V, Ot, false, Om);
        innerP(V, Bt, false, B);
        // (omega & theta) - (theta & omega) + 2 B
        T9 th, Ot1, Ot2, R;
        expandSymm(theta + 6 * (size_t)c, th);
        dotTT(Om, th, Ot1);
        dotTT(th, Om, Ot2);
        if (p->model == ORC_ROOTK) {
            // Oldroyd_BRootk.C:184-187: (omega & theta) - (theta & omega) + (2/k) (B & theta)
            T9 Bth;
            dotTT(B, th, Bth);
            const double c2k = 2.0 / p->rootK;
            for (int q = 0; q < 9; ++q) R[q] = Ot1[q] - Ot2[q] + c2k * Bth[q];
        } else {
            for (int q = 0; q < 9; ++q) R[q] = Ot1[q] - Ot2[q] + 2.0 * B[q];
        }
        // model term
        T9 invL, X;
        invField(Lam, mask, invL);
        const double rl = 1.0 / p->lambda;
        if (p->model == ORC_BMP) {
            // BMPLog.C:166-187: (Phi G0) V (inv(Lam) - I) V^T with the step's fluidity
            const double Phi = aux ? aux[6 * (size_t)c] : 0.0;
            T9 D, Y;
            for (int q = 0; q < 9; ++q) D[q] = invL[q] - I9[q];
            innerP(V, D, false, Y);
            const double coef = Phi * p->G0;
            for (int q = 0; q < 9; ++q) X[q] = coef * Y[q];
        } else if (p->model == ORC_SARAMITO) {
            // SaramitoLog.C:231-275: (fac etaP/lambda) [PTT g(tr C)] V (inv(Lam) - I) V^T
            const double fac = saramitoFac(p, tauOld ? tauOld + 6 * (size_t)c : nullptr);
            if (fOut) fOut[c] = fac;
            T9 D, Y;
            for (int q = 0; q < 9; ++q) D[q] = invL[q] - I9[q];
            innerP(V, D, false, Y);
            double coef = (fac * p->etaP) / p->lambda;
            if (p->sarPTT != 0) {
                T9 VT, t1, C;
                transp(V, VT);
                dotTT(V, Lam, t1);
                dotTT(t1, VT, C);
                const double trC = C[0] + C[4] + C[8];
                const double z = (p->epsilon / (1 - p->zeta)) * (trC - 3.);
                coef = coef * (p->sarPTT == 1 ? 1. + z : std::exp(z));
            }
            for (int q = 0; q < 9; ++q) X[q] = coef * Y[q];
        } else if (p->model == ORC_OLDROYDB) {
            T9 D, Y;
            for (int q = 0; q < 9; ++q) D[q] = invL[q] - I9[q];
            innerP(V, D, false, Y);
            for (int q = 0; q < 9; ++q) X[q] = rl * Y[q];
        } else if (p->model == ORC_GIESEKUS) {
            T9 D, trhs, t1, C, tt, CT;
            for (int q = 0; q < 9; ++q) D[q] = invL[q] - I9[q];
            T9 VT;
            transp(V, VT);
            dotTT(V, D, t1);
            dotTT(t1, VT, trhs);
            dotTT(V, Lam, t1);
            dotTT(t1, VT, C);
            dotTT(trhs, trhs, tt);
            dotTT(C, tt, CT);
            for (int q = 0; q < 9; ++q) X[q] = rl * (trhs[q] - p->alpha * CT[q]);
        } else if (p->model == ORC_FENEP) {
            T9 VT, t1, C, D, Y;
            transp(V, VT);
            dotTT(V, Lam, t1);
            dotTT(t1, VT, C);
            const double f = p->L2 / (p->L2 - (C[0] + C[4] + C[8]));
            const double a = p->L2 / (p->L2 - 3.);
            if (fOut) fOut[c] = f;
            for (int q = 0; q < 9; ++q) D[q] = a * invL[q] - f * I9[q];
            dotTT(V, D, t1);
            dotTT(t1, VT, Y);
            for (int q = 0; q < 9; ++q) X[q] = rl * Y[q];
        } else if (p->model == ORC_PTT_LINEAR || p->model == ORC_PTT_EXP || p->model == ORC_PTT_GEN) {
            T9 VT, t1, C, D, Y;
            transp(V, VT);
            for (int q = 0; q < 9; ++q) D[q] = invL[q] - I9[q];
            dotTT(V, D, t1);
            dotTT(t1, VT, Y);
            dotTT(V, Lam, t1);
            dotTT(t1, VT, C);
            const double k = p->epsilon / (1 - p->zeta);
            const double z = k * (C[0] + C[4] + C[8] - 3.);
            double fac;
            if (p->model == ORC_PTT_LINEAR) fac = 1. + z;
            else if (p->model == ORC_PTT_EXP) fac = std::exp(z);
            else fac = mlG[0] * mittagLeffler(mlG, mlMax, p->mlRtol, z);  // PTTLog.C:229, gammaBeta*Eab
            for (int q = 0; q < 9; ++q) X[q] = (rl * Y[q]) * fac;
        } else if (p->model == ORC_ROOTK) {
            // Oldroyd_BRootk.C:165-173,188: opert = V diag(Lam^(1/k - 1)) V^T - theta,
            // scaled by 1/(lambda k)
            T9 VT, t1, Od, Y;
            transp(V, VT);
            const double e = 1. / p->rootK - 1.;
            for (int q = 0; q < 9; ++q) Od[q] = 0.0;
            Od[0] = std::pow(Lam[0], e);
            Od[4] = std::pow(Lam[4], e);
            Od[8] = std::pow(Lam[8], e);
            dotTT(V, Od, t1);
            dotTT(t1, VT, Y);
            const double cf = 1.0 / (p->lambda * p->rootK);
            for (int q = 0; q < 9; ++q) X[q] = cf * (Y[q] - th[q]);
        } else if (p->model == ORC_FENECR) {
            // FENE_CRLog.C:142-164: f = L2/(L2 - tr(V Lam V^T)), X = (1/lambda) f V (inv(Lam) - I) V^T
            T9 VT, t1, C, D, Y;
            transp(V, VT);
            dotTT(V, Lam, t1);
            dotTT(t1, VT, C);
            const double f = p->L2 / (p->L2 - (C[0] + C[4] + C[8]));
            if (fOut) fOut[c] = f;
            for (int q = 0; q < 9; ++q) D[q] = invL[q] - I9[q];
            dotTT(V, D, t1);
            dotTT(t1, VT, Y);
            for (int q = 0; q < 9; ++q) X[q] = (rl * f) * Y[q];
        } else if (p->model == ORC_WMCY) {
            // WhiteMetznerCYLog.C:154-187: shear-rate dependent etaP and lambda
            double sL[6];
            symmT(L, sL);
            const double magS = std::sqrt(sL[0] * sL[0] + 2 * (sL[1] * sL[1]) + 2 * (sL[2] * sL[2]) + sL[3] * sL[3] +
                                          2 * (sL[4] * sL[4]) + sL[5] * sL[5]);
            const double etaP =
                p->etaP * std::pow(1 + std::pow(p->wmK * std::sqrt(2.0) * magS, p->wmA), (p->wmN - 1) / p->wmA);
            const double lam =
                p->lambda * std::pow(1 + std::pow(p->wmL * std::sqrt(2.0) * magS, p->wmB), (p->wmM - 1) / p->wmB);
            if (fOut) fOut[c] = etaP / lam;
            T9 VT, t1, D, Y;
            transp(V, VT);
            for (int q = 0; q < 9; ++q) D[q] = invL[q] - I9[q];
            dotTT(V, D, t1);
            dotTT(t1, VT, Y);
            const double rlc = 1.0 / lam;
            for (int q = 0; q < 9; ++q) X[q] = rlc * Y[q];
        } else if (p->model == ORC_ROLIEPOLY || p->model == ORC_XPP) {
            // A = symm(V Lam V^T); X = -(1/lambda_) (V inv(Lam) V^T) & Z(A)
            T9 VT, t1, C, A, Ci, Z;
            transp(V, VT);
            dotTT(V, Lam, t1);
            dotTT(t1, VT, C);
            double A6[6];
            symmT(C, A6);
            expandSymm(A6, A);
            const double trA = A[0] + A[4] + A[8];
            dotTT(V, invL, t1);
            dotTT(t1, VT, Ci);
            double rlm;
            if (p->model == ORC_ROLIEPOLY) {  // RoliePolyLog.C:140-185
                double M1 = 2. * (1. - std::sqrt(3. / trA)) / p->lambdaR;
                if (p->chiMax > 1.) {
                    const double c2 = p->chiMax * p->chiMax;
                    M1 *= (3. - (trA / 3.) / c2) * (1. - 1. / c2) / ((1. - (trA / 3.) / c2) * (3. - 1. / c2));
                }
                const double bp = p->beta * std::pow(trA / 3., p->delta);
                for (int q = 0; q < 9; ++q) {
                    const double AmI = A[q] - I9[q];
                    Z[q] = AmI + (M1 * p->lambdaD) * (A[q] + bp * AmI);
                }
                rlm = 1.0 / p->lambdaD;
            } else {  // XPomPomLog.C:140-185
                const double lb = std::sqrt(trA / 3.);
                T9 AA;
                dotTT(A, A, AA);
                const double trAA = AA[0] + AA[4] + AA[8];
                const double g = p->nXPP == 0 ? (1. - 1. / lb) : (1. - 1. / std::pow(lb, p->nXPP + 1));
                const double f = 2. * (p->lambdaB / p->lambdaS) * std::exp((2. / p->q) * (lb - 1.)) * g +
                                 (1. / (lb * lb)) * (1. - p->alpha - (p->alpha / 3.) * (trAA - 2. * trA));
                for (int q = 0; q < 9; ++q)
                    Z[q] = A[q] * (f - 2. * p->alpha) + p->alpha * AA[q] + I9[q] * (p->alpha - 1.);
                rlm = 1.0 / p->lambdaB;
            }
            T9 P;
            dotTT(Ci, Z, P);
            for (int q = 0; q < 9; ++q) X[q] = -(rlm * P[q]);
        }
        for (int q = 0; q < 9; ++q) R[q] = R[q] + X[q];
        symmT(R, S + 6 * (size_t)c);
    }
}

// fvm::ddt(theta) + fvm::div(phi, theta) == S (Oldroyd_BLog.C:141-164).
// scheme: the GaussDefCmpw limiter; souT (NULL for upwind/none) the explicit
// correction, entering the div matrix's source as -souT (:166).
// [OF9] basicSymmetryFvPatchField::evaluate: (theta_P + T theta_P T^T)/2 with
// T = I - 2 n n, n the unit normal of boundary face b (F + b in Sf)
static void symmetryValue(const orc_mesh* m, int b, const double* th, double* out) {
    const int64_t f = (int64_t)m->nInternal + b;
    const double sx = m->Sf[3 * f], sy = m->Sf[3 * f + 1], sz = m->Sf[3 * f + 2];
    const double mag = std::sqrt(sx * sx + sy * sy + sz * sz);
    const double n[3] = {sx / mag, sy / mag, sz / mag};
    double T[3][3], A[3][3], M[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) T[i][j] = (i == j ? 1.0 : 0.0) - 2.0 * n[i] * n[j];
    A[0][0] = th[0]; A[0][1] = A[1][0] = th[1]; A[0][2] = A[2][0] = th[2];
    A[1][1] = th[3]; A[1][2] = A[2][1] = th[4]; A[2][2] = th[5];
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) M[i][k] = T[i][0] * A[0][k] + T[i][1] * A[1][k] + T[i][2] * A[2][k];
    static const int I6[6] = {0, 0, 0, 1, 1, 2}, J6[6] = {0, 1, 2, 1, 2, 2};
    for (int q = 0; q < 6; ++q) {
        const int i = I6[q], l = J6[q];
        const double r = M[i][0] * T[l][0] + M[i][1] * T[l][1] + M[i][2] * T[l][2];
        out[q] = (th[q] + r) / 2.0;
    }
}
// [OF9] wedgePolyPatch::calcGeometry; faceT = rotationTensor(cn, n) of
// [OF9] transform.H: s I + (1 - s) a a / |a|^2 + (n cn - cn n), s = cn . n,
// a = cn ^ n
int orc_wedge_face_t(const orc_mesh* m, int32_t patch, double faceT[9]) {
    const int n0 = m->patchStart[patch], nf = m->patchSize[patch];
    if (nf <= 0) return 3;
    double n[3] = {0, 0, 0};
    for (int k = 0; k < nf; ++k) {
        const int64_t f = (int64_t)m->nInternal + n0 + k;
        const double* sf = m->Sf + 3 * f;
        const double mag = std::sqrt(sf[0] * sf[0] + sf[1] * sf[1] + sf[2] * sf[2]);
        for (int d = 0; d < 3; ++d) n[d] += sf[d] / mag;
    }
    for (int d = 0; d < 3; ++d) n[d] /= nf;  // gAverage
    double cn[3];
    for (int d = 0; d < 3; ++d) {
        const double sg = n[d] > 0 ? 1.0 : (n[d] < 0 ? -1.0 : 0.0);
        cn[d] = sg * (std::max(std::fabs(n[d]), 0.5) - 0.5);
    }
    const double mc = std::sqrt(cn[0] * cn[0] + cn[1] * cn[1] + cn[2] * cn[2]);
    if (mc == 0) return 1;
    for (int d = 0; d < 3; ++d) cn[d] /= mc;
    if (std::fabs(cn[0] + cn[1] + cn[2]) < 1 - 1e-15) return 1;
    const double s = cn[0] * n[0] + cn[1] * n[1] + cn[2] * n[2];
    const double a[3] = {cn[1] * n[2] - cn[2] * n[1], cn[2] * n[0] - cn[0] * n[2], cn[0] * n[1] - cn[1] * n[0]};
    const double ma = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
    if (std::sqrt(ma) < 1e-15) return 2;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            faceT[3 * i + j] = (i == j ? s : 0.0) + (1 - s) * a[i] * a[j] / ma + (n[i] * cn[j] - cn[i] * n[j]);
    return 0;
}
// [OF9] wedgeFvPatchField::evaluate: transform(faceT, theta_P) = T theta_P T^T
static void wedgeValue(const double* T, const double* th, double* out) {
    double A[3][3], M[3][3];
    A[0][0] = th[0]; A[0][1] = A[1][0] = th[1]; A[0][2] = A[2][0] = th[2];
    A[1][1] = th[3]; A[1][2] = A[2][1] = th[4]; A[2][2] = th[5];
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) M[i][k] = T[3 * i] * A[0][k] + T[3 * i + 1] * A[1][k] + T[3 * i + 2] * A[2][k];
    static const int I6[6] = {0, 0, 0, 1, 1, 2}, J6[6] = {0, 1, 2, 1, 2, 2};
    for (int q = 0; q < 6; ++q) {
        const int i = I6[q], l = J6[q];
        out[q] = M[i][0] * T[3 * l] + M[i][1] * T[3 * l + 1] + M[i][2] * T[3 * l + 2];
    }
}
// [OF9] wedgeFvPatchField::snGradTransformDiag for a symmTensor: d_i d_j,
// d = diag(I - cellT)/2 with cellT = faceT & faceT; valueInternalCoeffs = 1
// minus it
static double wedgeInternalCoeff(const double* T, int q) {
    double d[3];
    for (int i = 0; i < 3; ++i) {
        const double cii = T[3 * i] * T[i] + T[3 * i + 1] * T[3 + i] + T[3 * i + 2] * T[6 + i];
        d[i] = 0.5 * (1.0 - cii);
    }
    static const int I6[6] = {0, 0, 0, 1, 1, 2}, J6[6] = {0, 1, 2, 1, 2, 2};
    return 1.0 - d[I6[q]] * d[J6[q]];
}
// [OF9] basicSymmetryFvPatchField::snGradTransformDiag for a symmTensor:
// |n_i||n_j| of component (i, j); valueInternalCoeffs = 1 minus it
static double symmetryInternalCoeff(const orc_mesh* m, int b, int q) {
    const int64_t f = (int64_t)m->nInternal + b;
    const double sx = m->Sf[3 * f], sy = m->Sf[3 * f + 1], sz = m->Sf[3 * f + 2];
    const double mag = std::sqrt(sx * sx + sy * sy + sz * sz);
    const double d[3] = {std::fabs(sx / mag), std::fabs(sy / mag), std::fabs(sz / mag)};
    static const int I6[6] = {0, 0, 0, 1, 1, 2}, J6[6] = {0, 1, 2, 1, 2, 2};
    return 1.0 - d[I6[q]] * d[J6[q]];
}

static void assembleImpl(const orc_mesh* m, const double* phi, const double* thetaOld, const double* S,
                         double deltaT, const int32_t* bcType, const double* bcValue, int scheme,
                         const double* souT, double* diag, double* lower, double* upper,
                         double* source, double* iC, double* bC, const double* thetaCur = nullptr) {
    const int N = m->nCells, F = m->nInternal;
    const double rDeltaT = 1.0 / deltaT;
    const bool bounded = (scheme & ORC_DC_BOUNDED) != 0;
    scheme &= ~ORC_DC_BOUNDED;
    const bool conv = scheme != ORC_DC_NONE;  // "none": fvm.diag() = 0, nothing else (:90)
    // fvm::ddt Euler
    std::vector<double> ddtDiag(N), divDiag(N, 0.0);
    for (int c = 0; c < N; ++c) ddtDiag[c] = rDeltaT * m->V[c];
    // fvm::div upwind (gaussDefCmpwConvectionScheme.C:93-100); pos(0) = 1
    for (int f = 0; f < F; ++f) {
        const double upw = phi[f] >= 0 ? 1.0 : 0.0;
        lower[f] = conv ? -upw * phi[f] : 0.0;
        upper[f] = conv ? (1.0 - upw) * phi[f] : 0.0;
    }
    for (int f = 0; f < F; ++f) {  // negSumDiag
        divDiag[m->owner[f]] -= lower[f];
        divDiag[m->neighbour[f]] -= upper[f];
    }
    if (bounded) {
        // [OF9] boundedConvectionScheme::fvmDiv: scheme - fvm::Sp(surfaceIntegrate(phi)):
        // surfaceIntegrate sums phi out of each cell (internal faces in face
        // order, then the non-empty patches) over V; Sp's diagonal is V*sp
        std::vector<double> net(N, 0.0);
        for (int f = 0; f < F; ++f) {
            net[m->owner[f]] += phi[f];
            net[m->neighbour[f]] -= phi[f];
        }
        for (int p = 0; p < m->nPatches; ++p) {
            if (m->patchEmpty[p]) continue;
            for (int k = 0; k < m->patchSize[p]; ++k) {
                const int b = m->patchStart[p] + k;
                net[m->owner[(int64_t)F + b]] += phi[(int64_t)F + b];
            }
        }
        for (int c = 0; c < N; ++c) divDiag[c] -= m->V[c] * (net[c] / m->V[c]);
    }
    for (int c = 0; c < N; ++c) diag[c] = ddtDiag[c] + divDiag[c];
    // patch coefficients with plim = 1 (:102-116)
    const double* thP = thetaCur ? thetaCur : thetaOld;  // the patch fields' internal values
    for (int p = 0; p < m->nPatches; ++p) {
        double wT[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        if (bcType[p] == ORC_BC_WEDGE && m->patchSize[p] > 0 && orc_wedge_face_t(m, p, wT) != 0) {
            std::fprintf(stderr, "oracle: patch %d is no valid wedge (orc_wedge_face_t)\n", p);
            std::abort();
        }
        for (int k = 0; k < m->patchSize[p]; ++k) {
            const int b = m->patchStart[p] + k;
            const double pphi = phi[(int64_t)F + b];
            double thb[6] = {0, 0, 0, 0, 0, 0};
            const double* thc = thP + 6 * (size_t)m->owner[(int64_t)F + b];
            if (conv && bcType[p] == ORC_BC_SYMMETRY) symmetryValue(m, b, thc, thb);
            if (conv && bcType[p] == ORC_BC_WEDGE) wedgeValue(wT, thc, thb);
            for (int q = 0; q < 6; ++q) {
                double ic = 0, bc = 0;
                if (!conv) {
                } else if (bcType[p] == ORC_BC_WEDGE) {
                    // [OF9] transformFvPatchField::valueInternalCoeffs /
                    // valueBoundaryCoeffs
                    const double vic = wedgeInternalCoeff(wT, q);
                    ic = pphi * vic;
                    bc = -pphi * (thb[q] - vic * thc[q]);
                } else if (bcType[p] == ORC_BC_SYMMETRY) {
                    // [OF9] transformFvPatchField::valueInternalCoeffs /
                    // valueBoundaryCoeffs
                    const double vic = symmetryInternalCoeff(m, b, q);
                    ic = pphi * vic;
                    bc = -pphi * (thb[q] - vic * thc[q]);
                } else if (bcType[p] == ORC_BC_ZEROGRADIENT) {
                    ic = pphi * 1.0;
                    bc = -pphi * 0.0;
                } else if (bcType[p] == ORC_BC_FIXEDVALUE) {
                    ic = pphi * 0.0;
                    bc = -pphi * bcValue[6 * (size_t)b + q];
                } else if (bcType[p] == ORC_BC_COUPLED) {
                    // [OF9] coupled upwind: weights pos0(phi); the boundary
                    // coefficient becomes the interface coefficient, not a source
                    ic = pphi * (pphi >= 0 ? 1.0 : 0.0);
                }
                iC[6 * (size_t)b + q] = ic;
                bC[6 * (size_t)b + q] = bc;
            }
        }
    }
    // sources: ddt (rDeltaT*theta_old*V) [+ div: -souT] + (== S): V*S
    for (int c = 0; c < N; ++c)
        for (int q = 0; q < 6; ++q) {
            const double sd = rDeltaT * thetaOld[6 * (size_t)c + q] * m->V[c];  // theta.oldTime()
            const double vs = m->V[c] * S[6 * (size_t)c + q];
            source[6 * (size_t)c + q] = souT ? (sd + (-souT[6 * (size_t)c + q])) + vs : sd + vs;
        }
}

void orc_assemble(const orc_mesh* m, const double* phi, const double* theta, const double* S,
                  double deltaT, const int32_t* bcType, const double* bcValue, double* diag,
                  double* lower, double* upper, double* source, double* iC, double* bC) {
    assembleImpl(m, phi, theta, S, deltaT, bcType, bcValue, ORC_DC_UPWIND, nullptr, diag, lower, upper,
                 source, iC, bC);
}

void orc_grad_gauss_symm(const orc_mesh* m, const double* th, const double* thB, double* g) {
    // [OF9] gaussGrad::gradf on interpolate(vfc), linear weights, per component
    const int N = m->nCells, F = m->nInternal;
    std::fill(g, g + 18 * (size_t)N, 0.0);
    for (int f = 0; f < F; ++f) {
        const int o = m->owner[f], n = m->neighbour[f];
        for (int q = 0; q < 6; ++q) {
            const double tf = m->w[f] * (th[6 * (size_t)o + q] - th[6 * (size_t)n + q]) + th[6 * (size_t)n + q];
            for (int i = 0; i < 3; ++i) {
                const double v = m->Sf[3 * (size_t)f + i] * tf;
                g[18 * (size_t)o + 3 * q + i] += v;
                g[18 * (size_t)n + 3 * q + i] -= v;
            }
        }
    }
    for (int p = 0; p < m->nPatches; ++p) {
        if (m->patchEmpty[p]) continue;
        for (int k = 0; k < m->patchSize[p]; ++k) {
            const int b = m->patchStart[p] + k;
            const int64_t f = (int64_t)F + b;
            const int c = m->owner[f];
            for (int q = 0; q < 6; ++q)
                for (int i = 0; i < 3; ++i)
                    g[18 * (size_t)c + 3 * q + i] += m->Sf[3 * f + i] * thB[6 * (size_t)b + q];
        }
    }
    for (int c = 0; c < N; ++c)
        for (int q = 0; q < 18; ++q) g[18 * (size_t)c + q] /= m->V[c];
}

// limiters.H:38-100: alpha, beta per NVD segment and the segment bounds
static int limiterTable(int scheme, double al[3], double be[3], double bl[2]) {
    switch (scheme) {
    case ORC_DC_CUBISTA:
        al[0] = 7. / 4.; al[1] = 3. / 4.; al[2] = 1. / 4.;
        be[0] = 0.; be[1] = 3. / 8.; be[2] = 3. / 4.;
        bl[0] = 3. / 8.; bl[1] = 3. / 4.;
        return 1;
    case ORC_DC_MINMOD:
        al[0] = 1.5; al[1] = .5; al[2] = .5;
        be[0] = 0.; be[1] = .5; be[2] = .5;
        bl[0] = .5; bl[1] = 1.;
        return 1;
    case ORC_DC_SMART:
        al[0] = 3.; al[1] = 3. / 4.; al[2] = 0.;
        be[0] = 0.; be[1] = 3. / 8.; be[2] = 1.;
        bl[0] = 1. / 6.; bl[1] = 5. / 6.;
        return 1;
    case ORC_DC_WACEB:
        al[0] = 2.; al[1] = 3. / 4.; al[2] = 0.;
        be[0] = 0.; be[1] = 3. / 8.; be[2] = 1.;
        bl[0] = 3. / 10.; bl[1] = 5. / 6.;
        return 1;
    case ORC_DC_SUPERBEE:
        al[0] = 0.5; al[1] = 1.5; al[2] = 0.;
        be[0] = 0.5; be[1] = 0.; be[2] = 1.;
        bl[0] = 1. / 2.; bl[1] = 2. / 3.;
        return 1;
    default:
        return 0;  // upwind / none: no correction
    }
}

void orc_defc_faces(const orc_mesh* m, int32_t scheme, const double* phi, const double* th,
                    const double* g, const double* C, double* dc) {
    const int F = m->nInternal;
    double al[3], be[3], bl[2];
    if (!limiterTable(scheme, al, be, bl)) {
        std::fill(dc, dc + 6 * (size_t)F, 0.0);
        return;
    }
    const double swit = 1.0;  // onlyDCphi
    for (int f = 0; f < F; ++f) {
        const int o = m->owner[f], n = m->neighbour[f];
        const double upw = phi[f] >= 0 ? 1.0 : 0.0;
        const double d[3] = {C[3 * (size_t)n] - C[3 * (size_t)o], C[3 * (size_t)n + 1] - C[3 * (size_t)o + 1],
                             C[3 * (size_t)n + 2] - C[3 * (size_t)o + 2]};
        for (int q = 0; q < 6; ++q) {
            const double* go = g + 18 * (size_t)o + 3 * q;
            const double* gn = g + 18 * (size_t)n + 3 * q;
            const double vo = th[6 * (size_t)o + q], vn = th[6 * (size_t)n + q];
            double gu[3];
            for (int i = 0; i < 3; ++i) gu[i] = go[i] * upw + (1.0 - upw) * gn[i];
            const double gud = gu[0] * d[0] + gu[1] * d[1] + gu[2] * d[2];
            const double god = go[0] * d[0] + go[1] * d[1] + go[2] * d[2];
            const double gnd = gn[0] * d[0] + gn[1] * d[1] + gn[2] * d[2];
            // :251-255 the normalised variable of the upwind cell, :257-260 the segment
            const double phitc = 1.0 - ((vn - vo) / (2.0 * gud + 1e-18));
            double alpha, beta;
            if (phitc <= 0. || phitc >= 1.) {
                alpha = 1.;
                beta = 0.;
            } else if (phitc < bl[0]) {
                alpha = al[0];
                beta = be[0];
            } else if (phitc < bl[1]) {
                alpha = al[1];
                beta = be[1];
            } else {
                alpha = al[2];
                beta = be[2];
            }
            // :262-266 limited face value minus the upwind value
            dc[6 * (size_t)f + q] = (1.0 - alpha - beta) * (vn - 2.0 * god) * upw +
                                    (1.0 - alpha - beta) * (vo + 2.0 * gnd) * (1.0 - upw) +
                                    ((alpha - 1.0 * swit) * upw + beta * (1.0 - upw)) * vo +
                                    (beta * upw + (alpha - 1.0 * swit) * (1.0 - upw)) * vn;
        }
    }
}

void orc_defc_sout(const orc_mesh* m, const double* phi, const double* dc, double* souT) {
    const int N = m->nCells, F = m->nInternal;
    std::fill(souT, souT + 6 * (size_t)N, 0.0);
    for (int f = 0; f < F; ++f)  // :145-149
        for (int q = 0; q < 6; ++q) {
            souT[6 * (size_t)m->owner[f] + q] += dc[6 * (size_t)f + q] * phi[f];
            souT[6 * (size_t)m->neighbour[f] + q] -= dc[6 * (size_t)f + q] * phi[f];
        }
}

void orc_div_tau(const orc_mesh* m, const double* tau, const double* tauB, double rho, double* out) {
    const int N = m->nCells, F = m->nInternal;
    std::vector<double> t(6 * (size_t)N);
    for (size_t k = 0; k < 6 * (size_t)N; ++k) t[k] = tau[k] / rho;  // tau()/rho()
    std::fill(out, out + 3 * (size_t)N, 0.0);
    // Sf & T for a symmTensor T (xx xy xz yy yz zz)
    auto dot = [](const double* sf, const double* T, double* v) {
        v[0] = sf[0] * T[0] + sf[1] * T[1] + sf[2] * T[2];
        v[1] = sf[0] * T[1] + sf[1] * T[3] + sf[2] * T[4];
        v[2] = sf[0] * T[2] + sf[1] * T[4] + sf[2] * T[5];
    };
    for (int f = 0; f < F; ++f) {  // linear interpolate, surfaceIntegrate
        const int o = m->owner[f], n = m->neighbour[f];
        double tf[6], v[3];
        for (int q = 0; q < 6; ++q) tf[q] = m->w[f] * (t[6 * (size_t)o + q] - t[6 * (size_t)n + q]) + t[6 * (size_t)n + q];
        dot(m->Sf + 3 * (size_t)f, tf, v);
        for (int j = 0; j < 3; ++j) {
            out[3 * (size_t)o + j] += v[j];
            out[3 * (size_t)n + j] -= v[j];
        }
    }
    for (int p = 0; p < m->nPatches; ++p) {
        if (m->patchEmpty[p]) continue;
        for (int k = 0; k < m->patchSize[p]; ++k) {
            const int b = m->patchStart[p] + k;
            const int64_t f = (int64_t)F + b;
            const int c = m->owner[f];
            double tb[6], v[3];
            for (int q = 0; q < 6; ++q) tb[q] = tauB ? tauB[6 * (size_t)b + q] / rho : t[6 * (size_t)c + q];
            dot(m->Sf + 3 * f, tb, v);
            for (int j = 0; j < 3; ++j) out[3 * (size_t)c + j] += v[j];
        }
    }
    for (int c = 0; c < N; ++c)
        for (int j = 0; j < 3; ++j) out[3 * (size_t)c + j] /= m->V[c];
}

void orc_div_tau_stab(const orc_mesh* m, const double* C, const double* Cb, const double* U, const double* Ub,
                      const double* tau, const double* tauB, double rho, double etaP, int32_t stab,
                      int32_t lapScheme, double* out) {
    orc_div_tau(m, tau, tauB, rho, out);  // fvc::div(tau()/rho(), "div(tau)")
    if (stab == ORC_STAB_NONE) return;
    const int N = m->nCells, F = m->nInternal;
    std::vector<double> G(9 * (size_t)N);
    orc_grad_gauss(m, U, Ub, G.data());  // fvc::grad(U), cell values
    const double gam = etaP / rho;       // etaP()/rho(), uniform (Constant thermo)
    std::vector<double> S(3 * (size_t)N, 0.0);
    auto vdotT = [](const double* v, const double* T, double* o) {  // v & T
        for (int j = 0; j < 3; ++j) o[j] = v[0] * T[j] + v[1] * T[3 + j] + v[2] * T[6 + j];
    };
    auto mag3 = [](const double* v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); };
    // patch delta coefficient of a boundary face b of cell c: 1/|n (n & (Cf - C))|
    auto patchDelta = [&](int64_t f, int b, int c, const double* n) {
        const double d[3] = {Cb[3 * (size_t)b] - C[3 * (size_t)c], Cb[3 * (size_t)b + 1] - C[3 * (size_t)c + 1],
                             Cb[3 * (size_t)b + 2] - C[3 * (size_t)c + 2]};
        const double nd = n[0] * d[0] + n[1] * d[1] + n[2] * d[2];
        const double dl[3] = {n[0] * nd, n[1] * nd, n[2] * nd};
        (void)f;
        return 1.0 / mag3(dl);
    };
    if (stab == ORC_STAB_COUPLING) {
        // fvc::div((etaP/rho)*fvc::grad(U), "div(grad(U))"), Gauss linear
        std::vector<double> Gp(9 * (size_t)N);
        for (size_t k = 0; k < 9 * (size_t)N; ++k) Gp[k] = gam * G[k];
        for (int f = 0; f < F; ++f) {
            const int o = m->owner[f], n = m->neighbour[f];
            double gf[9], v[3];
            for (int q = 0; q < 9; ++q)
                gf[q] = m->w[f] * (Gp[9 * (size_t)o + q] - Gp[9 * (size_t)n + q]) + Gp[9 * (size_t)n + q];
            vdotT(m->Sf + 3 * (size_t)f, gf, v);
            for (int j = 0; j < 3; ++j) {
                S[3 * (size_t)o + j] += v[j];
                S[3 * (size_t)n + j] -= v[j];
            }
        }
        for (int p = 0; p < m->nPatches; ++p) {
            if (m->patchEmpty[p]) continue;
            for (int k = 0; k < m->patchSize[p]; ++k) {
                const int b = m->patchStart[p] + k;
                const int64_t f = (int64_t)F + b;
                const int c = m->owner[f];
                const double* sf = m->Sf + 3 * f;
                const double ms = mag3(sf);
                const double n[3] = {sf[0] / ms, sf[1] / ms, sf[2] / ms};
                const double dcb = patchDelta(f, b, c, n);
                const double* gP = G.data() + 9 * (size_t)c;
                double sn[3], nG[3], gb[9], gpb[9], v[3];
                for (int j = 0; j < 3; ++j) sn[j] = dcb * (Ub[3 * (size_t)b + j] - U[3 * (size_t)c + j]);
                vdotT(n, gP, nG);
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) gb[3 * i + j] = gP[3 * i + j] + n[i] * (sn[j] - nG[j]);
                for (int q = 0; q < 9; ++q) gpb[q] = gam * gb[q];
                vdotT(sf, gpb, v);
                for (int j = 0; j < 3; ++j) S[3 * (size_t)c + j] += v[j];
            }
        }
    } else {
        // fvc::laplacian(etaP/rho, U, "laplacian(eta,U)"): Gauss linear <snGrad>
        for (int f = 0; f < F; ++f) {
            const int o = m->owner[f], n = m->neighbour[f];
            const double* sf = m->Sf + 3 * (size_t)f;
            const double ms = mag3(sf);
            const double nf[3] = {sf[0] / ms, sf[1] / ms, sf[2] / ms};
            const double d[3] = {C[3 * (size_t)n] - C[3 * (size_t)o], C[3 * (size_t)n + 1] - C[3 * (size_t)o + 1],
                                 C[3 * (size_t)n + 2] - C[3 * (size_t)o + 2]};
            double dc;
            if (lapScheme == ORC_SNGRAD_ORTHOGONAL) dc = 1.0 / mag3(d);
            else dc = 1.0 / std::max(nf[0] * d[0] + nf[1] * d[1] + nf[2] * d[2], 0.05 * mag3(d));
            double sn[3];
            for (int j = 0; j < 3; ++j) sn[j] = dc * (U[3 * (size_t)n + j] - U[3 * (size_t)o + j]);
            if (lapScheme == ORC_SNGRAD_CORRECTED) {
                const double kv[3] = {nf[0] - d[0] * dc, nf[1] - d[1] * dc, nf[2] - d[2] * dc};
                double gf[9], cr[3];
                for (int q = 0; q < 9; ++q)
                    gf[q] = m->w[f] * (G[9 * (size_t)o + q] - G[9 * (size_t)n + q]) + G[9 * (size_t)n + q];
                vdotT(kv, gf, cr);
                for (int j = 0; j < 3; ++j) sn[j] = sn[j] + cr[j];
            }
            const double gms = gam * ms;  // gammaMagSf
            for (int j = 0; j < 3; ++j) {
                const double v = gms * sn[j];
                S[3 * (size_t)o + j] += v;
                S[3 * (size_t)n + j] -= v;
            }
        }
        for (int p = 0; p < m->nPatches; ++p) {
            if (m->patchEmpty[p]) continue;
            for (int k = 0; k < m->patchSize[p]; ++k) {
                const int b = m->patchStart[p] + k;
                const int64_t f = (int64_t)F + b;
                const int c = m->owner[f];
                const double* sf = m->Sf + 3 * f;
                const double ms = mag3(sf);
                const double n[3] = {sf[0] / ms, sf[1] / ms, sf[2] / ms};
                const double dcb = patchDelta(f, b, c, n);
                const double gms = gam * ms;
                for (int j = 0; j < 3; ++j)
                    S[3 * (size_t)c + j] += gms * (dcb * (Ub[3 * (size_t)b + j] - U[3 * (size_t)c + j]));
            }
        }
    }
    for (int c = 0; c < N; ++c)
        for (int j = 0; j < 3; ++j) out[3 * (size_t)c + j] = out[3 * (size_t)c + j] - S[3 * (size_t)c + j] / m->V[c];
}

void orc_tau_total(const orc_mesh* m, const double* U, const double* Ub, const double* tau, double etaS,
                   double* out) {
    const int N = m->nCells;
    std::vector<double> L(9 * (size_t)N);
    orc_grad_gauss(m, U, Ub, L.data());
    for (int c = 0; c < N; ++c) {
        const double* l = L.data() + 9 * (size_t)c;
        double T[9];  // L + L.T()
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) T[3 * i + j] = l[3 * i + j] + l[3 * j + i];
        // symm(T)
        const double s[6] = {T[0], 0.5 * (T[1] + T[3]), 0.5 * (T[2] + T[6]), T[4], 0.5 * (T[5] + T[7]), T[8]};
        const double tr = l[0] + l[4] + l[8];
        const double d = (2. / 3.) * tr;
        const double I[6] = {d, 0, 0, d, 0, d};  // (2/3) tr(L) * I
        for (int q = 0; q < 6; ++q) out[6 * (size_t)c + q] = tau[6 * (size_t)c + q] + etaS * (s[q] - I[q]);
    }
}

void orc_relax(const orc_mesh* m, const int32_t* bcType, double alpha, const double* psi,
               double* D, const double* lower, const double* upper, double* S,
               const double* iC) {
    if (alpha <= 0) return;
    const int N = m->nCells, F = m->nInternal;
    std::vector<double> D0(D, D + N), sumOff(N, 0.0);
    for (int f = 0; f < F; ++f) {
        sumOff[m->neighbour[f]] += std::fabs(lower[f]);
        sumOff[m->owner[f]] += std::fabs(upper[f]);
    }
    for (int p = 0; p < m->nPatches; ++p) {
        if (m->patchEmpty[p] || bcType[p] == ORC_BC_EMPTY) continue;
        for (int k = 0; k < m->patchSize[p]; ++k) {
            const int b = m->patchStart[p] + k;
            double mx = 0;
            for (int q = 0; q < 6; ++q) mx = std::max(mx, std::fabs(iC[6 * (size_t)b + q]));
            D[m->owner[(int64_t)F + b]] += mx;
        }
    }
    for (int c = 0; c < N; ++c) D[c] = std::max(std::fabs(D[c]), sumOff[c]);
    for (int c = 0; c < N; ++c) D[c] /= alpha;
    for (int p = 0; p < m->nPatches; ++p) {
        if (m->patchEmpty[p] || bcType[p] == ORC_BC_EMPTY) continue;
        for (int k = 0; k < m->patchSize[p]; ++k) {
            const int b = m->patchStart[p] + k;
            double mn = iC[6 * (size_t)b];
            for (int q = 1; q < 6; ++q) mn = std::min(mn, iC[6 * (size_t)b + q]);
            D[m->owner[(int64_t)F + b]] -= mn;
        }
    }
    for (int c = 0; c < N; ++c)
        for (int q = 0; q < 6; ++q) S[6 * (size_t)c + q] += (D[c] - D0[c]) * psi[6 * (size_t)c + q];
}

orc_perf orc_ldu_solve(int32_t nCells, int32_t nFaces, const int32_t* l, const int32_t* u,
                       const double* diag, const double* lower, const double* upper,
                       const double* b, const orc_controls* c, double* x) {
    Ldu A{nCells, nFaces, l, u, diag, lower, upper};
    return c->solver == ORC_PBICG ? pbicg(A, b, c, x) : pbicgstab(A, b, c, x);
}

double orc_foam_residual(int32_t nCells, int32_t nFaces, const int32_t* l, const int32_t* u,
                         const double* diag, const double* lower, const double* upper,
                         const double* b, const double* x) {
    Ldu A{nCells, nFaces, l, u, diag, lower, upper};
    std::vector<double> Ax(nCells), r(nCells);
    Amul(A, x, Ax.data());
    for (int c = 0; c < nCells; ++c) r[c] = b[c] - Ax[c];
    return sumMag(r) / normFactor(A, x, b, Ax.data());
}

void orc_solve_segregated(const orc_mesh* m, const double* diag, const double* lower,
                          const double* upper, const double* source, const double* iC,
                          const double* bC, const int32_t* bcType, const int32_t valid[6],
                          const orc_controls* c, double* psi, orc_perf perf[6]) {
    const int N = m->nCells, F = m->nInternal;
    std::vector<double> src(source, source + 6 * (size_t)N);
    // addBoundarySource (non-coupled patches; empty patches have no faces)
    for (int p = 0; p < m->nPatches; ++p) {
        if (m->patchEmpty[p] || bcType[p] == ORC_BC_EMPTY) continue;
        for (int k = 0; k < m->patchSize[p]; ++k) {
            const int b = m->patchStart[p] + k;
            const int cell = m->owner[(int64_t)F + b];
            for (int q = 0; q < 6; ++q) src[6 * (size_t)cell + q] += bC[6 * (size_t)b + q];
        }
    }
    std::vector<double> d(N), x(N), bb(N);
    for (int q = 0; q < 6; ++q) {
        perf[q] = orc_perf{0, 0, 0, 0, 0};
        if (valid[q] == -1) continue;
        std::copy(diag, diag + N, d.begin());
        // addBoundaryDiag
        for (int p = 0; p < m->nPatches; ++p) {
            if (m->patchEmpty[p] || bcType[p] == ORC_BC_EMPTY) continue;
            for (int k = 0; k < m->patchSize[p]; ++k) {
                const int b = m->patchStart[p] + k;
                d[m->owner[(int64_t)F + b]] += iC[6 * (size_t)b + q];
            }
        }
        for (int cl = 0; cl < N; ++cl) {
            x[cl] = psi[6 * (size_t)cl + q];
            bb[cl] = src[6 * (size_t)cl + q];
        }
        perf[q] = orc_ldu_solve(N, F, m->owner, m->neighbour, d.data(), lower, upper, bb.data(), c,
                                x.data());
        for (int cl = 0; cl < N; ++cl) psi[6 * (size_t)cl + q] = x[cl];
    }
}

void orc_calc_eig(int32_t n, const double* theta, double* eigVals, double* eigVecs, int32_t method) {
    for (int c = 0; c < n; ++c) {
        double ev[3], Q[9];
        if (method == 1) {
            double D[3], Vv[9];
            eigJacobiH(theta + 6 * (size_t)c, D, Vv);
            int idx[3] = {0, 1, 2};
            std::sort(idx, idx + 3, [&](int a, int b) { return D[a] < D[b]; });
            for (int k = 0; k < 3; ++k) {
                ev[k] = D[idx[k]];
                for (int r = 0; r < 3; ++r) Q[3 * r + k] = Vv[3 * r + idx[k]];
            }
        } else {
            eigQR(theta + 6 * (size_t)c, ev, Q);
        }
        double* e = eigVals + 9 * (size_t)c;
        for (int q = 0; q < 9; ++q) e[q] *= 0.;
        e[0] = std::exp(ev[0]);
        e[4] = std::exp(ev[1]);
        e[8] = std::exp(ev[2]);
        for (int q = 0; q < 9; ++q) eigVecs[9 * (size_t)c + q] = Q[q];
    }
}

// Oldroyd_BRootk.C:194-200: after calcEig (eigVals = exp(eig theta)) the
// conformation eigenvalues are log(eigVals)^k
static void rootkEigVals(const orc_params* p, int32_t n, double* eigVals) {
    if (p->model != ORC_ROOTK) return;
    for (int c = 0; c < n; ++c) {
        double* e = eigVals + 9 * (size_t)c;
        e[0] = std::pow(std::log(e[0]), p->rootK);
        e[4] = std::pow(std::log(e[4]), p->rootK);
        e[8] = std::pow(std::log(e[8]), p->rootK);
    }
}

void orc_tau(const orc_params* p, int32_t n, const double* eigVals, const double* eigVecs,
             const double* fOld, double* tau) {
    for (int c = 0; c < n; ++c) {
        const double* V = eigVecs + 9 * (size_t)c;
        const double* Lam = eigVals + 9 * (size_t)c;
        T9 VT, t1, C, D;
        transp(V, VT);
        dotTT(V, Lam, t1);
        dotTT(t1, VT, C);
        double coef;
        double post = 1.0;  // Rolie-Poly finite-extensibility factor
        if (p->model == ORC_FENEP) {
            const double a = p->L2 / (p->L2 - 3.);
            const double f = fOld[c];
            for (int q = 0; q < 9; ++q) D[q] = f * C[q] - a * I9[q];
            coef = p->etaP / p->lambda;
        } else {
            for (int q = 0; q < 9; ++q) D[q] = C[q] - I9[q];
            if (p->model == ORC_PTT_LINEAR || p->model == ORC_PTT_EXP || p->model == ORC_PTT_GEN ||
                p->model == ORC_SARAMITO)
                coef = p->etaP / (p->lambda * (1 - p->zeta));  // PTTLog.C, SaramitoLog.C:292
            else if (p->model == ORC_BMP)  // BMPLog.C:199
                coef = p->G0;
            else if (p->model == ORC_FENECR)  // FENE_CRLog.C:176
                coef = (p->etaP / p->lambda) * fOld[c];
            else if (p->model == ORC_WMCY)  // WhiteMetznerCYLog.C:187, etaP/lambda of the old L
                coef = fOld[c];
            else if (p->model == ORC_ROLIEPOLY) {  // RoliePolyLog.C:171-183
                coef = p->etaP / p->lambdaD;
                if (p->chiMax > 1.) {
                    double s6[6];
                    symmT(C, s6);
                    const double trA = s6[0] + s6[3] + s6[5];
                    const double c2 = p->chiMax * p->chiMax;
                    post = (3. - (trA / 3.) / c2) * (1. - 1. / c2) / ((1. - (trA / 3.) / c2) * (3. - 1. / c2));
                }
            } else if (p->model == ORC_XPP)  // XPomPomLog.C:180
                coef = p->etaP / p->lambdaB;
            else
                coef = p->etaP / p->lambda;
        }
        double s[6];
        symmT(D, s);
        for (int q = 0; q < 6; ++q) tau[6 * (size_t)c + q] = p->model == ORC_ROLIEPOLY && p->chiMax > 1.
                                                                 ? (coef * s[q]) * post
                                                                 : coef * s[q];
    }
}

void orc_correct(const orc_mesh* m, const orc_params* p, const double* U, const double* Ub,
                 const double* phi, const double* gradU, double deltaT, const int32_t* bcType,
                 const double* bcValue, const int32_t valid[6], const orc_controls* c,
                 double relaxFactor, double* theta, double* eigVals, double* eigVecs, double* tau,
                 orc_perf perf[6]) {
    orc_correct_scheme(m, p, U, Ub, phi, gradU, deltaT, bcType, bcValue, valid, c, relaxFactor,
                       ORC_DC_UPWIND, nullptr, theta, eigVals, eigVecs, tau, perf);
}

void orc_correct_scheme(const orc_mesh* m, const orc_params* p, const double* U, const double* Ub,
                        const double* phi, const double* gradU, double deltaT, const int32_t* bcType,
                        const double* bcValue, const int32_t valid[6], const orc_controls* c,
                        double relaxFactor, int32_t scheme, const double* C, double* theta,
                        double* eigVals, double* eigVecs, double* tau, orc_perf perf[6]) {
    orc_correct_iter(m, p, U, Ub, phi, gradU, deltaT, bcType, bcValue, valid, c, relaxFactor, scheme, C, nullptr,
                     theta, eigVals, eigVecs, tau, perf);
}

void orc_correct_iter(const orc_mesh* m, const orc_params* p, const double* U, const double* Ub,
                      const double* phi, const double* gradU, double deltaT, const int32_t* bcType,
                      const double* bcValue, const int32_t valid[6], const orc_controls* c,
                      double relaxFactor, int32_t scheme, const double* C, const double* thetaOld,
                      double* theta, double* eigVals, double* eigVecs, double* tau, orc_perf perf[6]) {
    orc_correct_aux(m, p, U, Ub, phi, gradU, deltaT, bcType, bcValue, valid, c, relaxFactor, scheme, C, thetaOld,
                    nullptr, theta, eigVals, eigVecs, tau, perf);
}

void orc_correct_aux(const orc_mesh* m, const orc_params* p, const double* U, const double* Ub,
                     const double* phi, const double* gradU, double deltaT, const int32_t* bcType,
                     const double* bcValue, const int32_t valid[6], const orc_controls* c,
                     double relaxFactor, int32_t scheme, const double* C, const double* thetaOld,
                     const double* aux, double* theta, double* eigVals, double* eigVecs, double* tau,
                     orc_perf perf[6]) {
    const int N = m->nCells, F = m->nInternal, NB = m->nBoundary;
    std::vector<double> L;
    if (!gradU) {
        L.resize(9 * (size_t)N);
        orc_grad_gauss(m, U, Ub, L.data());
        gradU = L.data();
    }
    std::vector<double> S(6 * (size_t)N), f(N, 0.0);
    orc_rhs_tau(p, N, gradU, theta, eigVals, eigVecs, tau, aux, S.data(), f.data());
    std::vector<double> diag(N), lower(F), upper(F), source(6 * (size_t)N), iC(6 * (size_t)NB),
        bC(6 * (size_t)NB);
    std::vector<double> souT;
    double al[3], be[3], bl[2];
    if (limiterTable(scheme & ~ORC_DC_BOUNDED, al, be, bl)) {
        // explicit correction from the old theta: patch values (zeroGradient:
        // the cell value; fixedValue: the given value), Gauss gradients, face
        // corrections, souT
        std::vector<double> thB(6 * (size_t)NB, 0.0), g(18 * (size_t)N), dc(6 * (size_t)F);
        for (int pp = 0; pp < m->nPatches; ++pp)
            for (int k = 0; k < m->patchSize[pp]; ++k) {
                const int b = m->patchStart[pp] + k;
                const int cell = m->owner[(int64_t)F + b];
                for (int q = 0; q < 6; ++q)
                    thB[6 * (size_t)b + q] = bcType[pp] == ORC_BC_FIXEDVALUE ? bcValue[6 * (size_t)b + q]
                                                                              : theta[6 * (size_t)cell + q];
                if (bcType[pp] == ORC_BC_SYMMETRY)
                    symmetryValue(m, b, theta + 6 * (size_t)cell, thB.data() + 6 * (size_t)b);
                if (bcType[pp] == ORC_BC_WEDGE) {
                    double wT[9];
                    if (orc_wedge_face_t(m, pp, wT) != 0) {
                        std::fprintf(stderr, "oracle: patch %d is no valid wedge (orc_wedge_face_t)\n", pp);
                        std::abort();
                    }
                    wedgeValue(wT, theta + 6 * (size_t)cell, thB.data() + 6 * (size_t)b);
                }
            }
        orc_grad_gauss_symm(m, theta, thB.data(), g.data());
        orc_defc_faces(m, scheme & ~ORC_DC_BOUNDED, phi, theta, g.data(), C, dc.data());
        souT.resize(6 * (size_t)N);
        orc_defc_sout(m, phi, dc.data(), souT.data());
    }
    assembleImpl(m, phi, thetaOld ? thetaOld : theta, S.data(), deltaT, bcType, bcValue, scheme,
                 souT.empty() ? nullptr : souT.data(),
                 diag.data(), lower.data(), upper.data(), source.data(), iC.data(), bC.data(), theta);
    if (p->model == ORC_FLUIDITY)  // == -fvm::Sp(1/lambda, Phi) (BMPLog.C:157)
        for (int cc = 0; cc < N; ++cc) diag[cc] = diag[cc] - (-(m->V[cc] * (1.0 / p->lambda)));
    if (relaxFactor > 0)
        orc_relax(m, bcType, relaxFactor, theta, diag.data(), lower.data(), upper.data(),
                  source.data(), iC.data());
    orc_solve_segregated(m, diag.data(), lower.data(), upper.data(), source.data(), iC.data(),
                         bC.data(), bcType, valid, c, theta, perf);
    if (p->model == ORC_FLUIDITY) return;
    orc_calc_eig(N, theta, eigVals, eigVecs, 0);
    rootkEigVals(p, N, eigVals);
    orc_tau(p, N, eigVals, eigVecs, f.data(), tau);
}

}  // extern "C"

// ---------------- the step of a decomposed run ---------------------------
// [OF9] decomposePar + processor patches: one thread per domain stands for
// one MPI rank.  Threads share memory, so the halo "exchange" of a processor
// patch is a read of the other domain's values after a barrier, and gSum is
// a sum of per-domain partials in domain order (every thread forms the same
// total, as every rank does after the all-reduce).
namespace {

struct Domain {
    int n = 0, off = 0;          // cells [off, off + n) of the domain-major order
    int Fi = 0, Nb = 0, Np = 0;  // internal, boundary and processor faces
    std::vector<int32_t> owner, neighbour;  // local cells; owner [Fi + Nb + Np]
    std::vector<double> Sf, w, V;           // Sf [3 (Fi + Nb + Np)], outward for processor faces
    std::vector<int32_t> patchStart, patchSize, patchEmpty;  // the mesh's patches, then the processor patch
    std::vector<int64_t> intFace, bndFace;  // original internal face / boundary-face index
    // processor faces: original owner and neighbour (domain-major), linear
    // weight, the other domain's cell (domain-major), the original face and
    // whether the local cell is its owner
    std::vector<int32_t> pO, pN, pRemote;
    std::vector<int64_t> pFace;
    std::vector<int8_t> pOwnerSide;
    std::vector<double> pW;
    std::vector<double> Ub, phi, bcValue;  // [Nb + Np] face values (processor part of bcValue unused)
    std::vector<int32_t> bcType;           // [nPatches + 1]
    orc_mesh mesh{};
};

// gSum over domains of up to four values; two slot banks alternate so a
// bank is rewritten only after every thread has passed the barrier that
// follows its last read.
struct Reducer {
    double* slots;
    int P, s, k;
    void sum(double* v, int m) {
        double* bank = slots + (size_t)(k++ & 1) * P * 4;
        for (int i = 0; i < m; ++i) bank[4 * s + i] = v[i];
#pragma omp barrier
        for (int i = 0; i < m; ++i) {
            double a = 0;
            for (int r = 0; r < P; ++r) a += bank[4 * r + i];
            v[i] = a;
        }
    }
    double sum1(double x) {
        sum(&x, 1);
        return x;
    }
};

// [OF9] lduMatrix::Amul: the domain's own product, then
// updateMatrixInterfaces: result[faceCell] -= interfaceBouCoeff * neighbour value.
void amulDD(const Ldu& A, const Domain& D, const double* cc, const double* xAll, double* y) {
    Amul(A, xAll + D.off, y);
    const int32_t* fc = D.owner.data() + D.Fi + D.Nb;
    for (int k = 0; k < D.Np; ++k) y[fc[k]] -= cc[k] * xAll[D.pRemote[k]];
}

// [OF9] lduMatrix::solver::normFactor with sumA's interface terms
// (sumA[faceCell] -= interfaceBouCoeff), gAverage and gSum.
double normFactorDD(const Ldu& A, const Domain& D, const double* cc, const double* psi, const double* b,
                    const double* Apsi, int nTotal, Reducer& R) {
    std::vector<double> t(A.N);
    sumA(A, t.data());
    const int32_t* fc = D.owner.data() + D.Fi + D.Nb;
    for (int k = 0; k < D.Np; ++k) t[fc[k]] -= cc[k];
    double avg = 0;
    for (int c = 0; c < A.N; ++c) avg += psi[c];
    avg = R.sum1(avg);
    avg /= nTotal;
    for (int c = 0; c < A.N; ++c) t[c] *= avg;
    double s = 0;
    for (int c = 0; c < A.N; ++c) s += std::fabs(Apsi[c] - t[c]) + std::fabs(b[c] - t[c]);
    return R.sum1(s) + 1e-20;
}

// pbicgstab() above on one domain: the same recurrence, with Amul across
// the interfaces, DILU on the domain's own matrix and global reductions.
// psi, yA and zA live in the shared domain-major vectors xAll/yAll/zAll
// because Amul reads them across the interfaces.
orc_perf pbicgstabDD(const Ldu& A, const Domain& D, const double* cc, double* xAll, double* yAll,
                     double* zAll, const double* b, const orc_controls* ctl, int nTotal, Reducer& R) {
    const int N = A.N;
    double* psi = xAll + D.off;
    double* yA = yAll + D.off;
    double* zA = zAll + D.off;
    orc_perf perf{0, 0, 0, 0, 1};
    std::vector<double> pA(N), rA(N);
#pragma omp barrier
    amulDD(A, D, cc, xAll, yA);
    for (int c = 0; c < N; ++c) rA[c] = b[c] - yA[c];
    const double nf = normFactorDD(A, D, cc, psi, b, yA, nTotal, R);
    perf.initialResidual = R.sum1(sumMag(rA)) / nf;
    perf.finalResidual = perf.initialResidual;
    if (ctl->minIter > 0 || !converged(perf.finalResidual, perf.initialResidual, ctl)) {
        std::vector<double> AyA(N), sA(N), tA(N), rA0(rA);
        double rA0rA = 0, alpha = 0, omega = 0;
        Precond P(A, ctl->precond);
        do {
            const double rA0rAold = rA0rA;
            rA0rA = R.sum1(sumProd(rA0.data(), rA.data(), N));
            if (std::fabs(rA0rA) < 1e-300) break;
            if (perf.nIterations == 0) {
                pA = rA;
            } else {
                if (std::fabs(omega) < 1e-300) break;
                const double beta = (rA0rA / rA0rAold) * (alpha / omega);
                for (int c = 0; c < N; ++c) pA[c] = rA[c] + beta * (pA[c] - omega * AyA[c]);
            }
            P.apply(yA, pA.data());
#pragma omp barrier
            amulDD(A, D, cc, yAll, AyA.data());
            const double rA0AyA = R.sum1(sumProd(rA0.data(), AyA.data(), N));
            alpha = rA0rA / rA0AyA;
            for (int c = 0; c < N; ++c) sA[c] = rA[c] - alpha * AyA[c];
            perf.finalResidual = R.sum1(sumMag(sA)) / nf;
            if (converged(perf.finalResidual, perf.initialResidual, ctl)) {
                for (int c = 0; c < N; ++c) psi[c] += alpha * yA[c];
                perf.nIterations++;
                perf.converged = 1;
                return perf;
            }
            P.apply(zA, sA.data());
#pragma omp barrier
            amulDD(A, D, cc, zAll, tA.data());
            double tt[2] = {sumProd(tA.data(), tA.data(), N), sumProd(tA.data(), sA.data(), N)};
            R.sum(tt, 2);
            omega = tt[1] / tt[0];
            for (int c = 0; c < N; ++c) {
                psi[c] += alpha * yA[c] + omega * zA[c];
                rA[c] = sA[c] - omega * tA[c];
            }
            perf.finalResidual = R.sum1(sumMag(rA)) / nf;
        } while ((++perf.nIterations < ctl->maxIter &&
                  !converged(perf.finalResidual, perf.initialResidual, ctl)) ||
                 perf.nIterations < ctl->minIter);
    }
    perf.converged = converged(perf.finalResidual, perf.initialResidual, ctl) ? 1 : 0;
    return perf;
}

}  // namespace

struct orc_decomp {
    int P = 0, N = 0, nPatches = 0, F = 0;
    std::vector<int32_t> order, pos;  // domain-major position <-> original cell
    std::vector<Domain> dom;
    // domain-major fields and work vectors, each page first written by the
    // thread of the domain that owns it
    std::vector<double*> bufs;
    double *U = nullptr, *theta = nullptr, *eigVals = nullptr, *eigVecs = nullptr, *tau = nullptr;
    double *g = nullptr, *S = nullptr, *f = nullptr, *xh = nullptr, *yh = nullptr, *zh = nullptr;
    std::vector<double> red;
    ~orc_decomp() {
        for (double* b : bufs) delete[] b;
    }
};

namespace {

// one domain's step: orc_correct_aux's phases on the domain's cells
void domainStep(orc_decomp* d, int s, const orc_params* p, double deltaT, const int32_t valid[6],
                const orc_controls* ctl, int scheme, orc_perf perf[6]) {
    Domain& D = d->dom[s];
    const int n = D.n, off = D.off, Fi = D.Fi, NB = D.Nb + D.Np;
    Reducer R{d->red.data(), d->P, s, 0};
    // coupled patch: the linear interpolate of U from both sides, formed
    // as the undecomposed face loop forms it
    for (int k = 0; k < D.Np; ++k) {
        const int o = D.pO[k], nb = D.pN[k];
        for (int q = 0; q < 3; ++q)
            D.Ub[3 * (size_t)(D.Nb + k) + q] = D.pW[k] * (d->U[3 * (size_t)o + q] - d->U[3 * (size_t)nb + q]) +
                                               d->U[3 * (size_t)nb + q];
    }
    double* g = d->g + 9 * (size_t)off;
    double* th = d->theta + 6 * (size_t)off;
    double* ev = d->eigVals + 9 * (size_t)off;
    double* vec = d->eigVecs + 9 * (size_t)off;
    double* tau = d->tau + 6 * (size_t)off;
    double* S = d->S + 6 * (size_t)off;
    double* f = d->f + off;
    orc_grad_gauss(&D.mesh, d->U + 3 * (size_t)off, D.Ub.data(), g);
    std::fill(f, f + n, 0.0);
    orc_rhs_tau(p, n, g, th, ev, vec, tau, nullptr, S, f);
    std::vector<double> diag(n), lower(Fi), upper(Fi), source(6 * (size_t)n), iC(6 * (size_t)NB),
        bC(6 * (size_t)NB);
    assembleImpl(&D.mesh, D.phi.data(), th, S, deltaT, D.bcType.data(), D.bcValue.data(), scheme, nullptr,
                 diag.data(), lower.data(), upper.data(), source.data(), iC.data(), bC.data());
    // interfaceBouCoeffs: the coupled patch's boundaryCoeffs, -phi_b (1 - pos0(phi_b))
    std::vector<double> cc(D.Np);
    for (int k = 0; k < D.Np; ++k) {
        const double pphi = D.phi[(size_t)Fi + D.Nb + k];
        cc[k] = -pphi * (1.0 - (pphi >= 0 ? 1.0 : 0.0));
    }
    // orc_solve_segregated on the domain.  [OF9] solveSegregated adds the
    // coupled patches' boundaryCoeffs x neighbour value to the source and
    // the interface update takes the same product off again; for these
    // component-wise coefficients the two cancel and both are left out.
    for (int pp = 0; pp < D.mesh.nPatches; ++pp) {
        if (D.patchEmpty[pp] || D.bcType[pp] == ORC_BC_EMPTY || D.bcType[pp] == ORC_BC_COUPLED) continue;
        for (int k = 0; k < D.patchSize[pp]; ++k) {
            const int b = D.patchStart[pp] + k;
            const int cell = D.owner[(size_t)Fi + b];
            for (int q = 0; q < 6; ++q) source[6 * (size_t)cell + q] += bC[6 * (size_t)b + q];
        }
    }
    std::vector<double> dg(n), bb(n);
    double* x = d->xh + off;
    for (int q = 0; q < 6; ++q) {
        orc_perf pf{0, 0, 0, 0, 0};
        if (valid[q] != -1) {
            std::copy(diag.begin(), diag.end(), dg.begin());
            for (int pp = 0; pp < D.mesh.nPatches; ++pp) {
                if (D.patchEmpty[pp] || D.bcType[pp] == ORC_BC_EMPTY) continue;
                for (int k = 0; k < D.patchSize[pp]; ++k) {
                    const int b = D.patchStart[pp] + k;
                    dg[D.owner[(size_t)Fi + b]] += iC[6 * (size_t)b + q];
                }
            }
            for (int c = 0; c < n; ++c) {
                x[c] = th[6 * (size_t)c + q];
                bb[c] = source[6 * (size_t)c + q];
            }
            Ldu A{n, Fi, D.owner.data(), D.neighbour.data(), dg.data(),
                  lower.data(), upper.data()};
            pf = pbicgstabDD(A, D, cc.data(), d->xh, d->yh, d->zh, bb.data(), ctl, d->N, R);
            for (int c = 0; c < n; ++c) th[6 * (size_t)c + q] = x[c];
        }
        if (s == 0) perf[q] = pf;
    }
    orc_calc_eig(n, th, ev, vec, 0);
    rootkEigVals(p, n, ev);
    orc_tau(p, n, ev, vec, f, tau);
}

}  // namespace

extern "C" {

orc_decomp* orc_decomp_create(const orc_mesh* m, int32_t P, const int32_t* domainOf) {
    const int N = m->nCells, F = m->nInternal;
    if (P < 1) return nullptr;
    for (int c = 0; c < N; ++c)
        if (domainOf[c] < 0 || domainOf[c] >= P) return nullptr;
    auto* d = new orc_decomp;
    d->P = P;
    d->N = N;
    d->F = F;
    d->nPatches = m->nPatches;
    d->dom.resize(P);
    std::vector<int32_t> start(P + 1, 0);
    for (int c = 0; c < N; ++c) start[domainOf[c] + 1]++;
    for (int r = 0; r < P; ++r) start[r + 1] += start[r];
    d->order.resize(N);
    d->pos.resize(N);
    {
        std::vector<int32_t> next(start.begin(), start.end() - 1);
        for (int c = 0; c < N; ++c) {
            const int k = next[domainOf[c]]++;
            d->order[k] = c;
            d->pos[c] = k;
        }
    }
    // faces of each domain, in the undecomposed face order
    std::vector<std::vector<int64_t>> iF(P), pF(P);
    std::vector<std::vector<std::vector<int64_t>>> bF(P, std::vector<std::vector<int64_t>>(m->nPatches));
    for (int f = 0; f < F; ++f) {
        const int a = domainOf[m->owner[f]], b = domainOf[m->neighbour[f]];
        if (a == b) {
            iF[a].push_back(f);
        } else {
            pF[a].push_back(2 * (int64_t)f + 1);
            pF[b].push_back(2 * (int64_t)f);
        }
    }
    for (int pp = 0; pp < m->nPatches; ++pp)
        for (int k = 0; k < m->patchSize[pp]; ++k) {
            const int b = m->patchStart[pp] + k;
            bF[domainOf[m->owner[(int64_t)F + b]]][pp].push_back(b);
        }
    d->red.assign(2 * (size_t)P * 4, 0.0);
    d->bufs.assign(11, nullptr);
#pragma omp parallel for schedule(static, 1) num_threads(P)
    for (int r = 0; r < P; ++r) {
        Domain& D = d->dom[r];
        D.off = start[r];
        D.n = start[r + 1] - start[r];
        D.Fi = (int)iF[r].size();
        D.Np = (int)pF[r].size();
        D.Nb = 0;
        for (int pp = 0; pp < m->nPatches; ++pp) D.Nb += (int)bF[r][pp].size();
        const size_t nf = (size_t)D.Fi + D.Nb + D.Np;
        D.owner.resize(nf);
        D.neighbour.resize(D.Fi);
        D.Sf.resize(3 * nf);
        D.w.resize(D.Fi);
        D.V.resize(D.n);
        for (int c = 0; c < D.n; ++c) D.V[c] = m->V[d->order[D.off + c]];
        D.intFace = iF[r];
        for (int k = 0; k < D.Fi; ++k) {
            const int64_t f = iF[r][k];
            D.owner[k] = d->pos[m->owner[f]] - D.off;
            D.neighbour[k] = d->pos[m->neighbour[f]] - D.off;
            D.w[k] = m->w[f];
            for (int q = 0; q < 3; ++q) D.Sf[3 * (size_t)k + q] = m->Sf[3 * f + q];
        }
        D.patchStart.resize(m->nPatches + 1);
        D.patchSize.resize(m->nPatches + 1);
        D.patchEmpty.resize(m->nPatches + 1);
        int run = 0;
        for (int pp = 0; pp < m->nPatches; ++pp) {
            D.patchStart[pp] = run;
            D.patchSize[pp] = (int)bF[r][pp].size();
            D.patchEmpty[pp] = m->patchEmpty[pp];
            for (int64_t b : bF[r][pp]) {
                const size_t lf = (size_t)D.Fi + run;
                D.owner[lf] = d->pos[m->owner[(int64_t)F + b]] - D.off;
                for (int q = 0; q < 3; ++q) D.Sf[3 * lf + q] = m->Sf[3 * ((int64_t)F + b) + q];
                D.bndFace.push_back(b);
                ++run;
            }
        }
        D.patchStart[m->nPatches] = run;
        D.patchSize[m->nPatches] = D.Np;
        D.patchEmpty[m->nPatches] = 0;
        D.pO.resize(D.Np);
        D.pN.resize(D.Np);
        D.pRemote.resize(D.Np);
        D.pFace.resize(D.Np);
        D.pOwnerSide.resize(D.Np);
        D.pW.resize(D.Np);
        for (int k = 0; k < D.Np; ++k) {
            const int64_t f = pF[r][k] / 2;
            const bool own = (pF[r][k] & 1) != 0;
            const int o = d->pos[m->owner[f]], nb = d->pos[m->neighbour[f]];
            const size_t lf = (size_t)D.Fi + D.Nb + k;
            D.owner[lf] = (own ? o : nb) - D.off;
            const double sg = own ? 1.0 : -1.0;
            for (int q = 0; q < 3; ++q) D.Sf[3 * lf + q] = sg * m->Sf[3 * f + q];
            D.pO[k] = o;
            D.pN[k] = nb;
            D.pRemote[k] = own ? nb : o;
            D.pFace[k] = f;
            D.pOwnerSide[k] = own ? 1 : 0;
            D.pW[k] = m->w[f];
        }
        D.Ub.assign(3 * (size_t)(D.Nb + D.Np), 0.0);
        D.phi.assign(nf, 0.0);
        D.bcValue.assign(6 * (size_t)(D.Nb + D.Np), 0.0);
        D.bcType.assign(m->nPatches + 1, ORC_BC_COUPLED);
        D.mesh = orc_mesh{D.n,
                          D.Fi,
                          D.Nb + D.Np,
                          m->nPatches + 1,
                          D.owner.data(),
                          D.neighbour.data(),
                          D.Sf.data(),
                          D.w.data(),
                          D.V.data(),
                          D.patchStart.data(),
                          D.patchSize.data(),
                          D.patchEmpty.data()};
    }
    const size_t widths[11] = {3, 6, 9, 9, 6, 9, 6, 1, 1, 1, 1};
    for (int i = 0; i < 11; ++i) d->bufs[i] = new double[widths[i] * (size_t)N];
#pragma omp parallel for schedule(static, 1) num_threads(P)
    for (int r = 0; r < P; ++r)
        for (int i = 0; i < 11; ++i)
            std::fill(d->bufs[i] + widths[i] * (size_t)start[r], d->bufs[i] + widths[i] * (size_t)start[r + 1],
                      0.0);
    d->U = d->bufs[0];
    d->theta = d->bufs[1];
    d->eigVals = d->bufs[2];
    d->eigVecs = d->bufs[3];
    d->tau = d->bufs[4];
    d->g = d->bufs[5];
    d->S = d->bufs[6];
    d->f = d->bufs[7];
    d->xh = d->bufs[8];
    d->yh = d->bufs[9];
    d->zh = d->bufs[10];
    return d;
}

void orc_decomp_destroy(orc_decomp* d) { delete d; }

void orc_decomp_set_fields(orc_decomp* d, const double* U, const double* Ub, const double* phi,
                           const double* theta, const double* eigVals, const double* eigVecs,
                           const double* tau, const double* bcValue) {
    const int N = d->N, F = d->F;
    for (int k = 0; k < N; ++k) {
        const int c = d->order[k];
        std::copy(U + 3 * (size_t)c, U + 3 * (size_t)c + 3, d->U + 3 * (size_t)k);
        std::copy(theta + 6 * (size_t)c, theta + 6 * (size_t)c + 6, d->theta + 6 * (size_t)k);
        std::copy(eigVals + 9 * (size_t)c, eigVals + 9 * (size_t)c + 9, d->eigVals + 9 * (size_t)k);
        std::copy(eigVecs + 9 * (size_t)c, eigVecs + 9 * (size_t)c + 9, d->eigVecs + 9 * (size_t)k);
        for (int q = 0; q < 6; ++q) d->tau[6 * (size_t)k + q] = tau ? tau[6 * (size_t)c + q] : 0.0;
    }
    for (Domain& D : d->dom) {
        for (int k = 0; k < D.Fi; ++k) D.phi[k] = phi[D.intFace[k]];
        for (int k = 0; k < D.Nb; ++k) {
            const int64_t b = D.bndFace[k];
            D.phi[(size_t)D.Fi + k] = phi[F + b];
            for (int q = 0; q < 3; ++q) D.Ub[3 * (size_t)k + q] = Ub ? Ub[3 * b + q] : 0.0;
            for (int q = 0; q < 6; ++q) D.bcValue[6 * (size_t)k + q] = bcValue ? bcValue[6 * b + q] : 0.0;
        }
        for (int k = 0; k < D.Np; ++k)
            D.phi[(size_t)D.Fi + D.Nb + k] = D.pOwnerSide[k] ? phi[D.pFace[k]] : -phi[D.pFace[k]];
    }
}

int32_t orc_decomp_correct(orc_decomp* d, const orc_params* p, double deltaT, const int32_t* bcType,
                           const int32_t valid[6], const orc_controls* c, int32_t scheme, orc_perf perf[6]) {
    if ((scheme & ~ORC_DC_BOUNDED) != ORC_DC_UPWIND || c->solver != ORC_PBICGSTAB ||
        p->model == ORC_BMP || p->model == ORC_FLUIDITY)
        return -1;
    for (Domain& D : d->dom)
        for (int pp = 0; pp < d->nPatches; ++pp) D.bcType[pp] = bcType[pp];
    int bad = 0;
    const int P = d->P;
#pragma omp parallel num_threads(P)
    {
        if (omp_get_num_threads() == P)
            domainStep(d, omp_get_thread_num(), p, deltaT, valid, c, scheme, perf);
        else if (omp_get_thread_num() == 0)
            bad = 1;
    }
    return bad ? -1 : 0;
}

void orc_decomp_get(const orc_decomp* d, double* theta, double* eigVals, double* eigVecs, double* tau) {
    for (int k = 0; k < d->N; ++k) {
        const int c = d->order[k];
        std::copy(d->theta + 6 * (size_t)k, d->theta + 6 * (size_t)k + 6, theta + 6 * (size_t)c);
        std::copy(d->eigVals + 9 * (size_t)k, d->eigVals + 9 * (size_t)k + 9, eigVals + 9 * (size_t)c);
        std::copy(d->eigVecs + 9 * (size_t)k, d->eigVecs + 9 * (size_t)k + 9, eigVecs + 9 * (size_t)c);
        std::copy(d->tau + 6 * (size_t)k, d->tau + 6 * (size_t)k + 6, tau + 6 * (size_t)c);
    }
}

}  // extern "C"
