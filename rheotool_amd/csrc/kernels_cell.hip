// kernels_cell.hip -- per-cell kernels of the log-conformation stress step (gfx950).
//
// k_assemble: one row (cell) per lane.  Fuses what the reference does in ~25
// field passes (boilerLog.H, constitutiveEq.C:323-358 decomposeGradU, the
// model source of Oldroyd_BLog.C:141-164 / GiesekusLog.C:142-160 /
// FENE_PLog.C:142-166 / PTTLog.C:184-196,241-253, the upwind coefficients of
// gaussDefCmpwConvectionScheme.C:93-116, Euler ddt and `== S`) into one
// pass: read theta, eigen state, U (+ neighbours via the ELL rows) and the
// face data, write the ELL values, the solve diagonal and the 6 sources.
// All arithmetic is FP64 in the reference's operation order and the file is
// built with -ffp-contract=off, so each term rounds as the reference's does.
//
// k_eig_tau: calcEig (constitutiveEq.C:360-428) with a cyclic Jacobi
// eigensolver per cell, exp of the eigenvalues, and the tau rebuild
// (Oldroyd_BLog.C:175, GiesekusLog.C:172, FENE_PLog.C:178, PTTLog.C:264).
#include "rt_device.h"

namespace rt {

// ------------------------------------------------------ tensor algebra
// OpenFOAM inner product a & b (row-major 3x3), sum order k = 0,1,2.
__device__ __forceinline__ void dotTT(const double* a, const double* b, double* c) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            c[3 * i + j] = a[3 * i] * b[j] + a[3 * i + 1] * b[3 + j] + a[3 * i + 2] * b[6 + j];
}
// V & D & V^T for diagonal D (exact: the off-diagonal products are zeros)
__device__ __forceinline__ void vdvt(const double* V, const double* d, double* r) {
    double VD[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) VD[3 * i + j] = V[3 * i + j] * d[j];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            r[3 * i + j] = VD[3 * i] * V[3 * j] + VD[3 * i + 1] * V[3 * j + 1] + VD[3 * i + 2] * V[3 * j + 2];
}
__device__ __forceinline__ void symmT(const double* t, double* s) {
    s[0] = t[0];
    s[1] = 0.5 * (t[1] + t[3]);
    s[2] = 0.5 * (t[2] + t[6]);
    s[3] = t[4];
    s[4] = 0.5 * (t[5] + t[7]);
    s[5] = t[8];
}

// [OF9] inv() of the diagonal eigVals tensor via cofactors / determinant,
// with the optional first-cell component shift (mask bits).
__device__ __forceinline__ void invDiag(const double* l, int mask, double* r) {
    double a = l[0], b = l[1], c = l[2];
    if (mask & 1) a += 1.0;
    if (mask & 2) b += 1.0;
    if (mask & 4) c += 1.0;
    const double d = a * b * c;
    r[0] = (b * c) / d;
    r[1] = (a * c) / d;
    r[2] = (a * b) / d;
    if (mask & 1) r[0] -= 1.0;
    if (mask & 2) r[1] -= 1.0;
    if (mask & 4) r[2] -= 1.0;
}

// Log-conformation right-hand side S (6 comps) of one cell; KIND is the
// model (ModelKind), fixed at compile time so only its branch holds registers.
template <int KIND>
__device__ __forceinline__ void cell_rhs(const ModelParams& p, int invMask, const double* L, const double* th6,
                                         const double* lam, const double* V, const double* tau6, double aux0,
                                         double* S, double& fOut) {
    constexpr bool ptt = KIND == M_PTT_LINEAR || KIND == M_PTT_EXP || KIND == M_PTT_GEN;
    if constexpr (KIND == M_FLUIDITY) {
        // BMPLog.C:152-160: Phi0/lambda + k (PhiInf - Phi)(tau && symm(L)); the
        // implicit -Sp(1/lambda) is on the diagonal
        double d6[6];
        symmT(L, d6);
        const double tdd = tau6[0] * d6[0] + 2 * (tau6[1] * d6[1]) + 2 * (tau6[2] * d6[2]) + tau6[3] * d6[3] +
                           2 * (tau6[4] * d6[4]) + tau6[5] * d6[5];
        S[0] = p.Phi0 / p.lambda + (p.kBMP * (p.PhiInf - th6[0])) * tdd;
#pragma unroll
        for (int q = 1; q < 6; ++q) S[q] = 0.0;
        fOut = 0.0;
        return;
    }
    double LT[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) LT[3 * i + j] = L[3 * j + i];
    if constexpr (ptt) {
        const double zs[9] = {L[0], 0.5 * (L[1] + L[3]), 0.5 * (L[2] + L[6]),
                              0.5 * (L[1] + L[3]), L[4], 0.5 * (L[5] + L[7]),
                              0.5 * (L[2] + L[6]), 0.5 * (L[5] + L[7]), L[8]};
#pragma unroll
        for (int q = 0; q < 9; ++q) LT[q] = LT[q] - p.zeta * zs[q];
    }
    // M = V^T & L^T & V
    double VT[9], tmp[9], M[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) VT[3 * i + j] = V[3 * j + i];
    dotTT(VT, LT, tmp);
    dotTT(tmp, V, M);
    // decomposeGradU
    const double lx = lam[0], ly = lam[1], lz = lam[2];
    double Ot[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    Ot[1] = (ly * M[1] + lx * M[3]) / (ly - lx + 1e-16);
    Ot[2] = (lz * M[2] + lx * M[6]) / (lz - lx + 1e-16);
    Ot[5] = (lz * M[5] + ly * M[7]) / (lz - ly + 1e-16);
    Ot[3] = -Ot[1];
    Ot[6] = -Ot[2];
    Ot[7] = -Ot[5];
    double Om[9], B[9];
    dotTT(V, Ot, tmp);
    dotTT(tmp, VT, Om);
    {
        const double bd[3] = {M[0], M[4], M[8]};
        vdvt(V, bd, B);
    }
    const double th[9] = {th6[0], th6[1], th6[2], th6[1], th6[3], th6[4], th6[2], th6[4], th6[5]};
    double R[9], t2[9];
    dotTT(Om, th, tmp);
    dotTT(th, Om, t2);
    if constexpr (KIND == M_ROOTK) {
        // Oldroyd_BRootk.C:184-187: (omega & theta) - (theta & omega) + (2/k) (B & theta)
        double Bth[9];
        dotTT(B, th, Bth);
        const double c2k = 2.0 / p.rootK;
#pragma unroll
        for (int q = 0; q < 9; ++q) R[q] = tmp[q] - t2[q] + c2k * Bth[q];
    } else {
#pragma unroll
        for (int q = 0; q < 9; ++q) R[q] = tmp[q] - t2[q] + 2.0 * B[q];
    }
    double il[3];
    invDiag(lam, invMask, il);
    const double rl = 1.0 / p.lambda;
    double X[9];
    fOut = 0.0;
    if constexpr (KIND == M_OLDROYDB) {
        const double dd[3] = {il[0] - 1.0, il[1] - 1.0, il[2] - 1.0};
        double Y[9];
        vdvt(V, dd, Y);
#pragma unroll
        for (int q = 0; q < 9; ++q) X[q] = rl * Y[q];
    } else if constexpr (KIND == M_GIESEKUS) {
        const double dd[3] = {il[0] - 1.0, il[1] - 1.0, il[2] - 1.0};
        double trhs[9], C[9], tt[9], CT[9];
        vdvt(V, dd, trhs);
        vdvt(V, lam, C);
        dotTT(trhs, trhs, tt);
        dotTT(C, tt, CT);
#pragma unroll
        for (int q = 0; q < 9; ++q) X[q] = rl * (trhs[q] - p.alpha * CT[q]);
    } else if constexpr (KIND == M_FENEP) {
        double C[9], Y[9];
        vdvt(V, lam, C);
        const double f = p.L2 / (p.L2 - (C[0] + C[4] + C[8]));
        const double a = p.L2 / (p.L2 - 3.);
        fOut = f;
        const double dd[3] = {a * il[0] - f, a * il[1] - f, a * il[2] - f};
        vdvt(V, dd, Y);
#pragma unroll
        for (int q = 0; q < 9; ++q) X[q] = rl * Y[q];
    } else if constexpr (KIND == M_PTT_LINEAR || KIND == M_PTT_EXP || KIND == M_PTT_GEN) {
        const double dd[3] = {il[0] - 1.0, il[1] - 1.0, il[2] - 1.0};
        double Y[9], C[9];
        vdvt(V, dd, Y);
        vdvt(V, lam, C);
        const double k = p.epsilon / (1 - p.zeta);
        const double z = k * (C[0] + C[4] + C[8] - 3.);
        double fac;
        if constexpr (KIND == M_PTT_LINEAR) {
            fac = 1. + z;
        } else if constexpr (KIND == M_PTT_EXP) {
            fac = exp(z);
        } else {
            // PTTLog.C:207-229: gamma(beta) E_{alpha,beta}(z), summed term by
            // term until the relative change falls to rTolMittagLeffler
            double sum = 0.0, sumOld = 0.0, error = 1.0;
            for (int j = 0; j < p.mlMaxIter && error > p.mlRtol; ++j) {
                const double Eabk = pow(z, (double)j) / p.mlGamma[j];
                sumOld = sum;
                sum += Eabk;
                error = fabs((sumOld - sum) / (sumOld + 1e-12));
            }
            fac = p.mlGammaBeta * sum;
        }
#pragma unroll
        for (int q = 0; q < 9; ++q) X[q] = (rl * Y[q]) * fac;
    } else if constexpr (KIND == M_ROOTK) {
        // Oldroyd_BRootk.C:165-173,188: (1/(lambda k)) (V diag(Lam^(1/k - 1)) V^T - theta)
        const double e = 1. / p.rootK - 1.;
        const double od[3] = {pow(lam[0], e), pow(lam[1], e), pow(lam[2], e)};
        double Y[9];
        vdvt(V, od, Y);
        const double cf = 1.0 / (p.lambda * p.rootK);
#pragma unroll
        for (int q = 0; q < 9; ++q) X[q] = cf * (Y[q] - th[q]);
    } else if constexpr (KIND == M_BMP) {
        // BMPLog.C:166-187: (Phi G0) V (inv(Lam) - I) V^T with the step's fluidity
        const double dd[3] = {il[0] - 1.0, il[1] - 1.0, il[2] - 1.0};
        double Y[9];
        vdvt(V, dd, Y);
        const double coef = aux0 * p.G0;
#pragma unroll
        for (int q = 0; q < 9; ++q) X[q] = coef * Y[q];
    } else if constexpr (KIND == M_SARAMITO) {
        // SaramitoLog.C:181-275: plasticity factor from the second invariant of
        // the deviator (over the valid dimensions) of the tau before the step,
        // then (fac etaP/lambda) [PTT g(tr C)] V (inv(Lam) - I) V^T
        const double nD = p.dimX + p.dimY + p.dimZ;
        const double tr = tau6[0] + tau6[3] + tau6[5];
        const double d0 = tau6[0] - (p.dimX * tr) / nD, d3 = tau6[3] - (p.dimY * tr) / nD,
                     d5 = tau6[5] - (p.dimZ * tr) / nD;
        const double magSqr = d0 * d0 + 2 * (tau6[1] * tau6[1]) + 2 * (tau6[2] * tau6[2]) + d3 * d3 +
                              2 * (tau6[4] * tau6[4]) + d5 * d5;
        const double tauDMag = sqrt(magSqr) / sqrt(2.);
        const double fac = p.sarN == 1. ? fmax(0., (tauDMag - p.tau0) / (p.sarK * tauDMag + 1e-16))
                                        : pow(fmax(0., (tauDMag - p.tau0) / (p.sarK * pow(tauDMag, p.sarN) * 1. + 1e-16)),
                                              1. / p.sarN);
        fOut = fac;
        double coef = (fac * p.etaP) / p.lambda;
        if (p.sarPTT != 0) {
            double C[9];
            vdvt(V, lam, C);
            const double z = (p.epsilon / (1 - p.zeta)) * (C[0] + C[4] + C[8] - 3.);
            coef = coef * (p.sarPTT == 1 ? 1. + z : exp(z));
        }
        const double dd[3] = {il[0] - 1.0, il[1] - 1.0, il[2] - 1.0};
        double Y[9];
        vdvt(V, dd, Y);
#pragma unroll
        for (int q = 0; q < 9; ++q) X[q] = coef * Y[q];
    } else if constexpr (KIND == M_FENECR) {
        // FENE_CRLog.C:142-164: X = (1/lambda) f V (inv(Lam) - I) V^T, f = L2/(L2 - tr C)
        double C[9], Y[9];
        vdvt(V, lam, C);
        const double f = p.L2 / (p.L2 - (C[0] + C[4] + C[8]));
        fOut = f;
        const double dd[3] = {il[0] - 1.0, il[1] - 1.0, il[2] - 1.0};
        vdvt(V, dd, Y);
#pragma unroll
        for (int q = 0; q < 9; ++q) X[q] = (rl * f) * Y[q];
    } else if constexpr (KIND == M_WMCY) {
        // WhiteMetznerCYLog.C:154-187: etaP(gd), lambda(gd), gd = sqrt(2)|symm(L)|
        double sL[6];
        symmT(L, sL);
        const double magS = sqrt(sL[0] * sL[0] + 2 * (sL[1] * sL[1]) + 2 * (sL[2] * sL[2]) + sL[3] * sL[3] +
                                 2 * (sL[4] * sL[4]) + sL[5] * sL[5]);
        const double etaP = p.etaP * pow(1 + pow(p.wmK * sqrt(2.0) * magS, p.wmA), (p.wmN - 1) / p.wmA);
        const double lamc = p.lambda * pow(1 + pow(p.wmL * sqrt(2.0) * magS, p.wmB), (p.wmM - 1) / p.wmB);
        fOut = etaP / lamc;
        const double dd[3] = {il[0] - 1.0, il[1] - 1.0, il[2] - 1.0};
        double Y[9];
        vdvt(V, dd, Y);
        const double rlc = 1.0 / lamc;
#pragma unroll
        for (int q = 0; q < 9; ++q) X[q] = rlc * Y[q];
    } else {
        // Rolie-Poly (RoliePolyLog.C:140-165) and XPP (XPomPomLog.C:140-172):
        // X = -(1/lambda_) (V inv(Lam) V^T) & Z(A), A = symm(V Lam V^T)
        double C[9], Ci[9], A[9], Z[9];
        vdvt(V, lam, C);
        vdvt(V, il, Ci);
        {
            double a6[6];
            symmT(C, a6);
            const double a9[9] = {a6[0], a6[1], a6[2], a6[1], a6[3], a6[4], a6[2], a6[4], a6[5]};
#pragma unroll
            for (int q = 0; q < 9; ++q) A[q] = a9[q];
        }
        const double trA = A[0] + A[4] + A[8];
        const double I9[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        double rlm;
        if constexpr (KIND == M_ROLIEPOLY) {
            double M1 = 2. * (1. - sqrt(3. / trA)) / p.lambdaR;
            if (p.chiMax > 1.) {
                const double c2 = p.chiMax * p.chiMax;
                M1 *= (3. - (trA / 3.) / c2) * (1. - 1. / c2) / ((1. - (trA / 3.) / c2) * (3. - 1. / c2));
            }
            const double bp = p.beta * pow(trA / 3., p.delta);
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                const double AmI = A[q] - I9[q];
                Z[q] = AmI + (M1 * p.lambdaD) * (A[q] + bp * AmI);
            }
            rlm = 1.0 / p.lambdaD;
        } else {
            const double lb = sqrt(trA / 3.);
            double AA[9];
            dotTT(A, A, AA);
            const double trAA = AA[0] + AA[4] + AA[8];
            const double g = p.nXPP == 0 ? (1. - 1. / lb) : (1. - 1. / pow(lb, p.nXPP + 1));
            const double f = 2. * (p.lambdaB / p.lambdaS) * exp((2. / p.q) * (lb - 1.)) * g +
                             (1. / (lb * lb)) * (1. - p.alpha - (p.alpha / 3.) * (trAA - 2. * trA));
#pragma unroll
            for (int q = 0; q < 9; ++q) Z[q] = A[q] * (f - 2. * p.alpha) + p.alpha * AA[q] + I9[q] * (p.alpha - 1.);
            rlm = 1.0 / p.lambdaB;
        }
        double P[9];
        dotTT(Ci, Z, P);
#pragma unroll
        for (int q = 0; q < 9; ++q) X[q] = -(rlm * P[q]);
    }
#pragma unroll
    for (int q = 0; q < 9; ++q) R[q] = R[q] + X[q];
    symmT(R, S);
}

// Widest row whose matrix values are staged in registers and written in
// column order; wider rows write each value to its slot directly.
constexpr int STAGE_W = 8;

template <class T, int KIND>
__global__ void __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(3, 8))) k_assemble(AssembleArgs a) {
    // per-thread running sums of the initial residual live in LDS, not in
    // registers: the row body needs every VGPR it can get for occupancy
    __shared__ double qs[19][BLK];
#pragma unroll
    for (int i = 0; i < 19; ++i) qs[i][threadIdx.x] = 0;
    double xbar[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) xbar[c] = a.cs[c].xbar;
    const int64_t nE = a.nEll;
    T* rOut = static_cast<T*>(a.r);
    T* r0Out = static_cast<T*>(a.r0);
    for (int64_t it = 0;; ++it) {
        const int64_t tl = tile_of(it, a.N, a.nRed);
        if (tl < 0) {
            if (tl == -1) break;
            continue;
        }
        const int64_t t0 = tl * BLK;
        const int32_t r = (int32_t)(t0 + threadIdx.x);
        if (r >= a.N) continue;
        const int32_t N = a.N;
        double th[6];
        ld6(a.theta + (int64_t)NCMP * r, th);
        const double vol = a.V[r];
        const double2 u01 = reinterpret_cast<const double2*>(a.U)[2 * (int64_t)r];
        const double Ui[3] = {u01.x, u01.y, a.U[4 * (int64_t)r + 2]};
        const int s = r >> 6, lane = r & 63;
        const int64_t base = a.ellOff[s] + lane;
        const int width = a.ellWidth[s];
        const bool staged = width <= STAGE_W;
        double g[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        double av[STAGE_W];
#pragma unroll
        for (int k = 0; k < STAGE_W; ++k) av[k] = 0.0;
        double offY[6] = {0, 0, 0, 0, 0, 0};
        double divDiag = 0.0, sumOff = 0.0, sumAv = 0.0, netPhi = 0.0;
        const bool haveGrad = a.gradU != nullptr;
        for (int k = 0; k < width; ++k) {
            const int64_t e = base + (int64_t)SLICE * k;  // face-order entry
            const int32_t fid = a.ellFid[e];
            const int slot = a.ellSlot[e];
            if (fid == RT_FID_PAD) {
                if (!staged) a.aval[base + (int64_t)SLICE * slot] = 0.0;
                if (a.atval) a.atval[base + (int64_t)SLICE * slot] = 0.0;
                continue;
            }
            const bool own = fid >= 0;
            const int64_t ec = base + (int64_t)SLICE * slot;  // column-order slot
            const double phRaw = a.ePhi[e];
            netPhi += own ? phRaw : -phRaw;
            const double ph = a.noConv ? 0.0 : phRaw;
            const double upw = ph >= 0 ? 1.0 : 0.0;
            const double lower = -upw * ph;
            const double upper = (1.0 - upw) * ph;
            const double val = own ? upper : lower;
            if (a.atval) a.atval[ec] = own ? lower : upper;  // A^T: the face's other coefficient
            const int32_t c = a.ellCol[ec];
            if (staged) {
#pragma unroll
                for (int j = 0; j < STAGE_W; ++j)
                    if (j == slot) av[j] = val;
            } else {
                a.aval[ec] = val;
            }
            divDiag -= own ? lower : upper;
            sumOff += fabs(val);
            sumAv += val;
            if (val != 0.0) {  // upwind: only the upstream side of a face couples
                double thc[6];
                ld6(a.theta + (int64_t)NCMP * c, thc);
#pragma unroll
                for (int q6 = 0; q6 < 6; ++q6) offY[q6] += val * thc[q6];
            }
            if (!haveGrad) {
                const double wf = a.eW[e];
                const double2 c01 = reinterpret_cast<const double2*>(a.U)[2 * (int64_t)c];
                const double Uc[3] = {c01.x, c01.y, a.U[4 * (int64_t)c + 2]};
                double uf[3];
#pragma unroll
                for (int q3 = 0; q3 < 3; ++q3) {
                    const double uo = own ? Ui[q3] : Uc[q3], un = own ? Uc[q3] : Ui[q3];
                    uf[q3] = wf * (uo - un) + un;
                }
                // Sf already carries the row's side: g += (+-Sf) (x) U_f
                const double sf[3] = {a.eSf[e], a.eSf[nE + e], a.eSf[2 * nE + e]};
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j) g[3 * i + j] += sf[i] * uf[j];
            }
        }
        // Upwinding leaves one nonzero per face, on the downstream row, so many
        // column-order slots are zero in every row of a slice (the 64 rows are
        // the lanes of this wavefront): bit k of the slice's live-slot mask is
        // set when slot k holds a nonzero in any row, and the solve's kernels
        // skip the others without loading them.  Slots past 31 count as live.
        if (staged) {
            uint32_t mask = 0;
            for (int k = 0; k < width; ++k) {
                double v = 0.0;
#pragma unroll
                for (int j = 0; j < STAGE_W; ++j)
                    if (j == k) v = av[j];
                a.aval[base + (int64_t)SLICE * k] = v;
                if (__ballot(v != 0.0) != 0) mask |= 1u << k;
            }
            if (lane == 0) {
                a.slotMask[s] = mask;
                qs[18][threadIdx.x] += (double)(SLICE * __popc(mask));
            }
        } else if (lane == 0) {
            a.slotMask[s] = 0xffffffffu;
            qs[18][threadIdx.x] += (double)(SLICE * width);
        }
        const int32_t b0 = a.bPtr[r], b1 = a.bPtr[r + 1];
        const int64_t NB = a.NB > 0 ? a.NB : 1;
        if (!haveGrad) {
            for (int32_t qb = b0; qb < b1; ++qb) {
                const int32_t bf = a.bFace[qb];
                const double sf[3] = {a.bSf[bf], a.bSf[NB + bf], a.bSf[2 * NB + bf]};
                const double ub[3] = {a.Ub[bf], a.Ub[NB + bf], a.Ub[2 * NB + bf]};
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j) g[3 * i + j] += sf[i] * ub[j];
            }
#pragma unroll
            for (int qq = 0; qq < 9; ++qq) g[qq] /= vol;
        } else {
#pragma unroll
            for (int qq = 0; qq < 9; ++qq) g[qq] = a.gradU[(int64_t)qq * N + r];
        }
        double S[6], f;
        {
            double lam[3], V[9];
#pragma unroll
            for (int qq = 0; qq < 3; ++qq) lam[qq] = a.eigVal[(int64_t)qq * N + r];
#pragma unroll
            for (int qq = 0; qq < 9; ++qq) V[qq] = a.eigVec[(int64_t)qq * N + r];
            double to6[6] = {0, 0, 0, 0, 0, 0};
            if constexpr (KIND == M_SARAMITO || KIND == M_FLUIDITY) {
                if (a.tauOld) ld6(a.tauOld + (int64_t)NCMP * r, to6);
            }
            double aux0 = 0.0;
            if constexpr (KIND == M_BMP) aux0 = a.aux[(int64_t)NCMP * r];
            cell_rhs<KIND>(a.prm, a.invMask, g, th, lam, V, to6, aux0, S, f);
        }
        if (a.fOld) a.fOld[r] = f;
        // matrix diagonal: ddt + div, optional relax, then boundary internalCoeffs
        if (a.bounded) {
            // bounded wrapper: div - fvm::Sp(surfaceIntegrate(phi)); the net
            // outflux adds the non-empty patch faces after the internal ones
            for (int32_t qb = b0; qb < b1; ++qb) {
                const int32_t bf = a.bFace[qb];
                if (a.bcType[a.bPatch[bf]] != RT_BC_EMPTY) netPhi += a.phiB[bf];
            }
            divDiag -= vol * (netPhi / vol);
        }
        double D = a.rDeltaT * vol + divDiag;
        if constexpr (KIND == M_FLUIDITY) D = D - (-(vol * (1.0 / a.prm.lambda)));  // == -fvm::Sp(1/lambda, Phi)
        double src[6];
        double tho[6];  // theta.oldTime() of the step: theta unless this is an inner iteration
        if (a.thetaOld) ld6(a.thetaOld + (int64_t)NCMP * r, tho);
        else {
#pragma unroll
            for (int qq = 0; qq < 6; ++qq) tho[qq] = th[qq];
        }
        if (a.souT) {
            // fvm::ddt + fvm::div (source -souT) == S, summed in that order
            double so[6];
            ld6(a.souT + (int64_t)NCMP * r, so);
#pragma unroll
            for (int qq = 0; qq < 6; ++qq) src[qq] = (a.rDeltaT * tho[qq] * vol + (-so[qq])) + vol * S[qq];
        } else {
#pragma unroll
            for (int qq = 0; qq < 6; ++qq) src[qq] = a.rDeltaT * tho[qq] * vol + vol * S[qq];
        }
        if (a.relax > 0) {
            const double D0 = D;
            for (int32_t qb = b0; qb < b1; ++qb) {
                const int32_t bf = a.bFace[qb];
                const int t = a.bcType[a.bPatch[bf]];
                if (t == RT_BC_EMPTY) continue;
                const double ph = a.noConv ? 0.0 : a.phiB[bf];
                D += t != RT_BC_FIXEDVALUE ? fabs(ph * 1.0) : fabs(ph * 0.0);
            }
            D = fmax(fabs(D), sumOff);
            D /= a.relax;
            for (int32_t qb = b0; qb < b1; ++qb) {
                const int32_t bf = a.bFace[qb];
                const int t = a.bcType[a.bPatch[bf]];
                if (t == RT_BC_EMPTY) continue;
                const double ph = a.noConv ? 0.0 : a.phiB[bf];
                D -= t != RT_BC_FIXEDVALUE ? ph * 1.0 : ph * 0.0;
            }
#pragma unroll
            for (int qq = 0; qq < 6; ++qq) src[qq] += (D - D0) * th[qq];
        }
        for (int32_t qb = b0; qb < b1; ++qb) {
            const int32_t bf = a.bFace[qb];
            const int t = a.bcType[a.bPatch[bf]];
            if (t == RT_BC_EMPTY) continue;
            const double ph = a.noConv ? 0.0 : a.phiB[bf];
            if (t == RT_BC_ZEROGRADIENT) {
                D += ph * 1.0;
            } else if (t == RT_BC_SYMMETRY) {
                // the reference's coefficients are phi_b (1 - |n_i||n_j|) on the
                // diagonal of component (i, j) and -phi_b (theta_b - (1 -
                // |n_i||n_j|) theta_P) in its source; the lockstep solve shares
                // one diagonal, so it takes phi_b and -phi_b (theta_b - theta_P),
                // which differ from them only through phi_b (zero on a symmetry
                // plane, up to rounding) times (theta_new - theta_P)
                D += ph * 1.0;
                const double sf[3] = {a.bSf[bf], a.bSf[NB + bf], a.bSf[2 * NB + bf]};
                double sym[6];
                symmetry_value(sf, th, sym);
#pragma unroll
                for (int c = 0; c < 6; ++c) src[c] += -ph * (sym[c] - th[c]);
            } else if (t == RT_BC_WEDGE) {
                // the same shared-diagonal form as symmetry: the reference's
                // per-component coefficient is 1 - d_i d_j, d = diag(I - cellT)/2
                D += ph * 1.0;
                double wv[6];
                wedge_value(a.patchT + 9 * (int64_t)a.bPatch[bf], th, wv);
#pragma unroll
                for (int c = 0; c < 6; ++c) src[c] += -ph * (wv[c] - th[c]);
            } else {
                D += ph * 0.0;
#pragma unroll
                for (int c = 0; c < 6; ++c) src[c] += -ph * a.bcVal[(int64_t)c * NB + bf];
            }
        }
        a.diag[r] = D;
        a.rD[r] = 1.0 / D;
#pragma unroll
        for (int qq = 0; qq < 6; ++qq)
            if (a.valid[qq] <= 0) src[qq] = 0.0;
        st6(a.b + (int64_t)NCMP * r, src);
        // initial residual of the solve: r = b - A theta, and the sums of the
        // normalisation factor (sparseSolver.C:129-180)
        const double sumA = D + sumAv;
        T rt[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            const double y = D * th[c] + offY[c];
            const double res = src[c] - y;
            rt[c] = (T)res;
            if (a.valid[c] > 0) {
                qs[c][threadIdx.x] += fabs(res);
                qs[6 + c][threadIdx.x] += (double)rt[c] * (double)rt[c];
                const double pA = sumA * xbar[c];
                qs[12 + c][threadIdx.x] += fabs(y - pA) + fabs(src[c] - pA);
            }
        }
        if (rOut) {  // NULL: the solve's first sweep writes r and r0 itself (red-black)
            st6(rOut + (int64_t)NCMP * r, rt);
            st6(r0Out + (int64_t)NCMP * r, rt);
        }
    }
    double q[19];
#pragma unroll
    for (int i = 0; i < 19; ++i) q[i] = qs[i][threadIdx.x];
    write_partials<19>(q, a.partials);
}

__global__ void __launch_bounds__(BLK) k_fin_assemble(AssembleArgs a, int G) {
    double sums[19];
    sum_partials<19>(a.partials, G, sums);
    if (threadIdx.x == 0) {
        for (int c = 0; c < 6; ++c)
            if (a.valid[c] > 0) finish_initial(a.cs[c], a.ctl, sums[c], sums[6 + c], sums[12 + c]);
        *a.nnzOut = sums[18];
    }
}

// ----------------------------------------------------------- calcEig + tau
// Symmetric 3x3 eigen-decomposition by cyclic Jacobi rotations (Rutishauser's
// stable update), ascending order; column eigenvectors in V (row-major).
// W (may be NULL): an orthonormal basis to start from (columns), normally the
// previous step's eigenvectors.  The rotations then act on W^T t W, which is
// nearly diagonal when theta changed little, and converge in one or two
// sweeps instead of four to six; the result is the same decomposition (to
// rounding, and up to the order and signs of the vectors, which are sorted
// and do not enter the model terms).
__device__ void jacobi3(const double* t6, const double* W, double* ev, double* V) {
    double a00 = t6[0], a01 = t6[1], a02 = t6[2], a11 = t6[3], a12 = t6[4], a22 = t6[5];
    double v[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    if (W) {
        // A = W^T T W
        const double T[9] = {a00, a01, a02, a01, a11, a12, a02, a12, a22};
        double TW[9];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) TW[3 * i + j] = T[3 * i] * W[j] + T[3 * i + 1] * W[3 + j] + T[3 * i + 2] * W[6 + j];
        double A[9];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) A[3 * i + j] = W[i] * TW[j] + W[3 + i] * TW[3 + j] + W[6 + i] * TW[6 + j];
        a00 = A[0], a11 = A[4], a22 = A[8];
        a01 = 0.5 * (A[1] + A[3]), a02 = 0.5 * (A[2] + A[6]), a12 = 0.5 * (A[5] + A[7]);
#pragma unroll
        for (int q = 0; q < 9; ++q) v[q] = W[q];
    }
    for (int sweep = 0; sweep < 12; ++sweep) {
        const double off = fabs(a01) + fabs(a02) + fabs(a12);
        const double dia = fabs(a00) + fabs(a11) + fabs(a22);
        if (off == 0.0 || off <= 1e-19 * dia) break;
        // rotate pairs (0,1), (0,2), (1,2)
#pragma unroll
        for (int pq = 0; pq < 3; ++pq) {
            double *app, *aqq, *apq, *ar1, *ar2;  // ar1 = a[r][p], ar2 = a[r][q] for the third index r
            int p, q;
            if (pq == 0) { app = &a00; aqq = &a11; apq = &a01; ar1 = &a02; ar2 = &a12; p = 0; q = 1; }
            else if (pq == 1) { app = &a00; aqq = &a22; apq = &a02; ar1 = &a01; ar2 = &a12; p = 0; q = 2; }
            else { app = &a11; aqq = &a22; apq = &a12; ar1 = &a01; ar2 = &a02; p = 1; q = 2; }
            const double x = *apq;
            if (x == 0.0) continue;
            const double h = *aqq - *app;
            double t;
            if (fabs(h) + 100.0 * fabs(x) == fabs(h)) {
                t = x / h;
            } else {
                const double theta = 0.5 * h / x;
                t = 1.0 / (fabs(theta) + sqrt(1.0 + theta * theta));
                if (theta < 0.0) t = -t;
            }
            const double c = 1.0 / sqrt(1.0 + t * t), s = t * c, tau = s / (1.0 + c);
            const double hh = t * x;
            *app -= hh;
            *aqq += hh;
            *apq = 0.0;
            const double g1 = *ar1, g2 = *ar2;
            *ar1 = g1 - s * (g2 + g1 * tau);
            *ar2 = g2 + s * (g1 - g2 * tau);
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                const double vp = v[3 * r + p], vq = v[3 * r + q];
                v[3 * r + p] = vp - s * (vq + vp * tau);
                v[3 * r + q] = vq + s * (vp - vq * tau);
            }
        }
    }
    double d[3] = {a00, a11, a22};
    int i0 = 0, i1 = 1, i2 = 2;
    // sort ascending (3-element network)
    if (d[i1] < d[i0]) { int t = i0; i0 = i1; i1 = t; }
    if (d[i2] < d[i1]) { int t = i1; i1 = i2; i2 = t; }
    if (d[i1] < d[i0]) { int t = i0; i0 = i1; i1 = t; }
    const int idx[3] = {i0, i1, i2};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        ev[c] = d[idx[c]];
#pragma unroll
        for (int r = 0; r < 3; ++r) V[3 * r + c] = v[3 * r + idx[c]];
    }
}

// Gram-Schmidt on the columns of W (row-major) in place: false when W holds a
// non-finite or (nearly) dependent set, which is then not used as a start.
__device__ __forceinline__ bool orthonormalize3(double* W) {
    double n0 = W[0] * W[0] + W[3] * W[3] + W[6] * W[6];
    if (!(n0 > 1e-12) || !isfinite(n0)) return false;
    n0 = 1.0 / sqrt(n0);
    W[0] *= n0, W[3] *= n0, W[6] *= n0;
    const double d = W[0] * W[1] + W[3] * W[4] + W[6] * W[7];
    W[1] -= d * W[0], W[4] -= d * W[3], W[7] -= d * W[6];
    double n1 = W[1] * W[1] + W[4] * W[4] + W[7] * W[7];
    if (!(n1 > 1e-12) || !isfinite(n1)) return false;
    n1 = 1.0 / sqrt(n1);
    W[1] *= n1, W[4] *= n1, W[7] *= n1;
    // third column: the cross product of the first two
    W[2] = W[3] * W[7] - W[6] * W[4];
    W[5] = W[6] * W[1] - W[0] * W[7];
    W[8] = W[0] * W[4] - W[3] * W[1];
    return true;
}

__global__ void __launch_bounds__(BLK) k_eig_tau(int32_t N, ModelParams p, const double* __restrict__ theta,
                                                 double* __restrict__ eigVal, double* __restrict__ eigVec,
                                                 double* __restrict__ tau, const double* __restrict__ fOld) {
    const int64_t r = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (r >= N) return;
    double th[6], ev[3], V[9], W[9];
    ld6(theta + (int64_t)NCMP * r, th);
#pragma unroll
    for (int q = 0; q < 9; ++q) W[q] = eigVec[q * (int64_t)N + r];
    jacobi3(th, orthonormalize3(W) ? W : nullptr, ev, V);
    double lam[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        lam[q] = exp(ev[q]);
        // Oldroyd_BRootk.C:194-200: the conformation eigenvalues log(exp(.))^k
        if (p.kind == M_ROOTK) lam[q] = pow(log(lam[q]), p.rootK);
        eigVal[q * (int64_t)N + r] = lam[q];
    }
#pragma unroll
    for (int q = 0; q < 9; ++q) eigVec[q * (int64_t)N + r] = V[q];
    if (!tau) return;  // calcEig only (the eigen state of an uploaded theta)
    double C[9], D[9];
    vdvt(V, lam, C);
    double coef;
    double post = 1.0;
    const double I9[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    if (p.kind == M_FENEP) {
        const double a = p.L2 / (p.L2 - 3.);
        const double f = fOld[r];
#pragma unroll
        for (int q = 0; q < 9; ++q) D[q] = f * C[q] - a * I9[q];
        coef = p.etaP / p.lambda;
    } else {
#pragma unroll
        for (int q = 0; q < 9; ++q) D[q] = C[q] - I9[q];
        if (p.kind == M_PTT_LINEAR || p.kind == M_PTT_EXP || p.kind == M_PTT_GEN || p.kind == M_SARAMITO) {
            coef = p.etaP / (p.lambda * (1 - p.zeta));  // PTTLog.C, SaramitoLog.C:292
        } else if (p.kind == M_FENECR) {  // FENE_CRLog.C:176, f of the old state
            coef = (p.etaP / p.lambda) * fOld[r];
        } else if (p.kind == M_WMCY) {  // WhiteMetznerCYLog.C:187, etaP/lambda of the old L
            coef = fOld[r];
        } else if (p.kind == M_ROLIEPOLY) {  // RoliePolyLog.C:171-183
            coef = p.etaP / p.lambdaD;
            if (p.chiMax > 1.) {
                double s6[6];
                symmT(C, s6);
                const double trA = s6[0] + s6[3] + s6[5];
                const double c2 = p.chiMax * p.chiMax;
                post = (3. - (trA / 3.) / c2) * (1. - 1. / c2) / ((1. - (trA / 3.) / c2) * (3. - 1. / c2));
            }
        } else if (p.kind == M_XPP) {  // XPomPomLog.C:180
            coef = p.etaP / p.lambdaB;
        } else if (p.kind == M_BMP) {  // BMPLog.C:199
            coef = p.G0;
        } else {
            coef = p.etaP / p.lambda;
        }
    }
    double s[6];
    symmT(D, s);
    const bool chi = p.kind == M_ROLIEPOLY && p.chiMax > 1.;
#pragma unroll
    for (int q = 0; q < 6; ++q) s[q] = chi ? (coef * s[q]) * post : coef * s[q];
    st6(tau + (int64_t)NCMP * r, s);
}

// ------------------------------------------------------------- layout helpers
__global__ void k_aos_to_soa(int32_t N, int nc, const double* __restrict__ aos,
                             const int32_t* __restrict__ perm, double* __restrict__ soa) {
    const int64_t r = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (r >= N) return;
    const int64_t c = perm ? perm[r] : r;
    for (int q = 0; q < nc; ++q) soa[q * (int64_t)N + r] = aos[c * nc + q];
}
__global__ void k_soa_to_aos(int32_t N, int nc, const double* __restrict__ soa,
                             const int32_t* __restrict__ perm, double* __restrict__ aos) {
    const int64_t r = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (r >= N) return;
    const int64_t c = perm ? perm[r] : r;
    for (int q = 0; q < nc; ++q) aos[c * nc + q] = soa[q * (int64_t)N + r];
}
__global__ void k_eigvals_from_tensor(int32_t N, const double* __restrict__ aos9,
                                      const int32_t* __restrict__ perm, double* __restrict__ ev) {
    const int64_t r = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (r >= N) return;
    const int64_t c = perm ? perm[r] : r;
    ev[r] = aos9[9 * c];
    ev[N + r] = aos9[9 * c + 4];
    ev[2 * (int64_t)N + r] = aos9[9 * c + 8];
}
__global__ void k_eigvals_to_tensor(int32_t N, const double* __restrict__ ev,
                                    const int32_t* __restrict__ perm, double* __restrict__ aos9) {
    const int64_t r = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (r >= N) return;
    const int64_t c = perm ? perm[r] : r;
    double* t = aos9 + 9 * c;
    t[0] = ev[r];
    t[1] = t[2] = t[3] = 0.0;
    t[4] = ev[N + r];
    t[5] = t[6] = t[7] = 0.0;
    t[8] = ev[2 * (int64_t)N + r];
}
__global__ void k_add(int64_t n, const double* __restrict__ in, double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i < n) out[i] = out[i] + in[i];
}

__global__ void k_fill(double* p, int64_t n, double v) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i < n) p[i] = v;
}
// face-order entry e of row r sits at ellOff[s] + 64k + lane; its value goes to
// the column-order slot ellSlot[e] of the same row
__device__ __forceinline__ int64_t col_slot(int64_t e, const int64_t* __restrict__ off,
                                            const uint8_t* __restrict__ slot, int32_t nSlices) {
    // locate the slice of e by binary search over the slice offsets
    int32_t lo = 0, hi = nSlices - 1;
    while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (off[mid] <= e) lo = mid;
        else hi = mid - 1;
    }
    const int64_t within = e - off[lo];
    const int64_t lane = within & 63;
    return off[lo] + 64 * (int64_t)slot[e] + lane;
}
__global__ void k_ldu_to_ell(int64_t nEll, int32_t nSlices, const int64_t* __restrict__ off,
                             const int32_t* __restrict__ fid, const uint8_t* __restrict__ slot,
                             const double* __restrict__ lower, const double* __restrict__ upper,
                             double* __restrict__ aval) {
    const int64_t e = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (e >= nEll) return;
    const int32_t f = fid[e];
    aval[col_slot(e, off, slot, nSlices)] = f == RT_FID_PAD ? 0.0 : (f >= 0 ? upper[f] : lower[-f - 1]);
}
__global__ void k_face_to_entry(int64_t nEll, const int32_t* __restrict__ fid,
                                const double* __restrict__ phi, double* __restrict__ ePhi) {
    const int64_t e = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (e >= nEll) return;
    const int32_t f = fid[e];
    ePhi[e] = f == RT_FID_PAD ? 0.0 : phi[f >= 0 ? f : -f - 1];
}
__global__ void k_permute_rows(int32_t N, int nc, int ncOut, const double* __restrict__ in,
                               const int32_t* __restrict__ perm, double* __restrict__ out) {
    const int64_t r = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (r >= N) return;
    const int64_t c = perm ? perm[r] : r;
    for (int q = 0; q < ncOut; ++q) out[r * ncOut + q] = q < nc ? in[c * nc + q] : 0.0;
}
__global__ void k_unpermute_rows(int32_t N, int ncIn, int nc, const double* __restrict__ in,
                                 const int32_t* __restrict__ perm, double* __restrict__ out) {
    const int64_t r = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (r >= N) return;
    const int64_t c = perm ? perm[r] : r;
    for (int q = 0; q < nc; ++q) out[c * nc + q] = in[r * ncIn + q];
}

// plain streaming copy, the HBM calibration kernel: each lane moves COPY_U
// 16-byte vectors, all loads issued before the stores
constexpr int COPY_U = 4;
__global__ void __launch_bounds__(BLK) k_copy(int64_t n16, const double2* __restrict__ src,
                                              double2* __restrict__ dst) {
    const int64_t base = (int64_t)blockIdx.x * BLK * COPY_U + threadIdx.x;
    double2 v[COPY_U];
#pragma unroll
    for (int u = 0; u < COPY_U; ++u) {
        const int64_t i = base + (int64_t)u * BLK;
        if (i < n16) v[u] = src[i];
    }
#pragma unroll
    for (int u = 0; u < COPY_U; ++u) {
        const int64_t i = base + (int64_t)u * BLK;
        if (i < n16) dst[i] = v[u];
    }
}

// ---------------------------------------------------------------- launchers
static inline unsigned nblk(int64_t n) { return (unsigned)((n + BLK - 1) / BLK); }

// ------------------------------------------------ GaussDefCmpw correction
// Gauss-linear gradient of each theta component ([OF9] gaussGrad::gradf on
// the linear interpolate, as phifDefC takes it, gaussDefCmpwConvectionScheme.C:
// 243-250): per row, faces in increasing face id (the reference's scatter
// order), then the row's boundary faces (zeroGradient: the cell value,
// fixedValue: the patch value), then / V.  gradTh[r][3q + i] = d theta_q / dx_i.
__global__ void __launch_bounds__(BLK) k_grad_theta(AssembleArgs a, double* __restrict__ gradTh) {
    const int64_t NB = a.NB > 0 ? a.NB : 1;
    for (int64_t it = 0;; ++it) {
        const int64_t tl = tile_of(it, a.N, a.nRed);
        if (tl < 0) {
            if (tl == -1) break;
            continue;
        }
        const int32_t r = (int32_t)(tl * BLK + threadIdx.x);
        if (r >= a.N) continue;
        double th[6];
        ld6(a.theta + (int64_t)NCMP * r, th);
        double g[18];
#pragma unroll
        for (int q = 0; q < 18; ++q) g[q] = 0.0;
        const int s = r >> 6, lane = r & 63;
        const int64_t base = a.ellOff[s] + lane;
        const int width = a.ellWidth[s];
        for (int k = 0; k < width; ++k) {
            const int64_t e = base + (int64_t)SLICE * k;
            const int32_t fid = a.ellFid[e];
            if (fid == RT_FID_PAD) continue;
            const bool own = fid >= 0;
            const int32_t c = a.ellCol[base + (int64_t)SLICE * a.ellSlot[e]];
            double thc[6];
            ld6(a.theta + (int64_t)NCMP * c, thc);
            const double wf = a.eW[e];
            const double sf[3] = {a.eSf[e], a.eSf[a.nEll + e], a.eSf[2 * a.nEll + e]};
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                const double vo = own ? th[q] : thc[q], vn = own ? thc[q] : th[q];
                const double tf = wf * (vo - vn) + vn;
#pragma unroll
                for (int i = 0; i < 3; ++i) g[3 * q + i] += sf[i] * tf;
            }
        }
        const int32_t b0 = a.bPtr[r], b1 = a.bPtr[r + 1];
        for (int32_t qb = b0; qb < b1; ++qb) {
            const int32_t bf = a.bFace[qb];
            const int t = a.bcType[a.bPatch[bf]];
            if (t == RT_BC_EMPTY) continue;
            const double sf[3] = {a.bSf[bf], a.bSf[NB + bf], a.bSf[2 * NB + bf]};
            double sym[6];
            if (t == RT_BC_SYMMETRY) symmetry_value(sf, th, sym);
            else if (t == RT_BC_WEDGE) wedge_value(a.patchT + 9 * (int64_t)a.bPatch[bf], th, sym);
            const bool transformed = t == RT_BC_SYMMETRY || t == RT_BC_WEDGE;
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                const double tb = t == RT_BC_FIXEDVALUE ? a.bcVal[(int64_t)q * NB + bf]
                                                        : (transformed ? sym[q] : th[q]);
#pragma unroll
                for (int i = 0; i < 3; ++i) g[3 * q + i] += sf[i] * tb;
            }
        }
        const double vol = a.V[r];
        double2* out = reinterpret_cast<double2*>(gradTh + (int64_t)18 * r);
#pragma unroll
        for (int q = 0; q < 9; ++q) out[q] = make_double2(g[2 * q] / vol, g[2 * q + 1] / vol);
    }
}

// limiters.H:38-100: NVD segments (alpha, beta) and their bounds
template <int S> struct Limiter;
template <> struct Limiter<DS_CUBISTA> {
    static constexpr double al[3] = {7. / 4., 3. / 4., 1. / 4.}, be[3] = {0., 3. / 8., 3. / 4.},
                            bl[2] = {3. / 8., 3. / 4.};
};
template <> struct Limiter<DS_MINMOD> {
    static constexpr double al[3] = {1.5, .5, .5}, be[3] = {0., .5, .5}, bl[2] = {.5, 1.};
};
template <> struct Limiter<DS_SMART> {
    static constexpr double al[3] = {3., 3. / 4., 0.}, be[3] = {0., 3. / 8., 1.}, bl[2] = {1. / 6., 5. / 6.};
};
template <> struct Limiter<DS_WACEB> {
    static constexpr double al[3] = {2., 3. / 4., 0.}, be[3] = {0., 3. / 8., 1.}, bl[2] = {3. / 10., 5. / 6.};
};
template <> struct Limiter<DS_SUPERBEE> {
    static constexpr double al[3] = {0.5, 1.5, 0.}, be[3] = {0.5, 0., 1.}, bl[2] = {1. / 2., 2. / 3.};
};

// Explicit correction of fvmDiv (gaussDefCmpwConvectionScheme.C:120-169) with
// the face values of phifDefC, onlyDCphi (:238-266): per row, in face order,
// souT += dc*phi on the owner side and -= on the neighbour side.  Only the
// upwind cell's gradient enters a face (the other one is multiplied by 0), so
// a row gathers a neighbour's gradient only where the neighbour is upwind.
template <int S>
__global__ void __launch_bounds__(BLK) k_defc(AssembleArgs a, const double* __restrict__ C4,
                                              const double* __restrict__ gradTh, double* __restrict__ souT) {
    using L = Limiter<S>;
    for (int64_t it = 0;; ++it) {
        const int64_t tl = tile_of(it, a.N, a.nRed);
        if (tl < 0) {
            if (tl == -1) break;
            continue;
        }
        const int32_t r = (int32_t)(tl * BLK + threadIdx.x);
        if (r >= a.N) continue;
        double th[6];
        ld6(a.theta + (int64_t)NCMP * r, th);
        const double2 cr01 = reinterpret_cast<const double2*>(C4)[2 * (int64_t)r];
        const double Cr[3] = {cr01.x, cr01.y, C4[4 * (int64_t)r + 2]};
        double so[6] = {0, 0, 0, 0, 0, 0};
        const int s = r >> 6, lane = r & 63;
        const int64_t base = a.ellOff[s] + lane;
        const int width = a.ellWidth[s];
        for (int k = 0; k < width; ++k) {
            const int64_t e = base + (int64_t)SLICE * k;
            const int32_t fid = a.ellFid[e];
            if (fid == RT_FID_PAD) continue;
            const bool own = fid >= 0;
            const int32_t c = a.ellCol[base + (int64_t)SLICE * a.ellSlot[e]];
            const double ph = a.ePhi[e];
            const double upw = ph >= 0 ? 1.0 : 0.0;
            const double2 cc01 = reinterpret_cast<const double2*>(C4)[2 * (int64_t)c];
            const double Cc[3] = {cc01.x, cc01.y, C4[4 * (int64_t)c + 2]};
            double d[3];  // C[neighbour] - C[owner]
#pragma unroll
            for (int i = 0; i < 3; ++i) d[i] = own ? Cc[i] - Cr[i] : Cr[i] - Cc[i];
            double thc[6];
            ld6(a.theta + (int64_t)NCMP * c, thc);
            const bool rowUp = own == (upw == 1.0);
            const double* gU = gradTh + (int64_t)18 * (rowUp ? r : c);
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                const double vo = own ? th[q] : thc[q], vn = own ? thc[q] : th[q];
                const double gud = gU[3 * q] * d[0] + gU[3 * q + 1] * d[1] + gU[3 * q + 2] * d[2];
                const double phitc = 1.0 - ((vn - vo) / (2.0 * gud + 1e-18));
                double alpha, beta;
                if (phitc <= 0. || phitc >= 1.) {
                    alpha = 1.;
                    beta = 0.;
                } else if (phitc < L::bl[0]) {
                    alpha = L::al[0];
                    beta = L::be[0];
                } else if (phitc < L::bl[1]) {
                    alpha = L::al[1];
                    beta = L::be[1];
                } else {
                    alpha = L::al[2];
                    beta = L::be[2];
                }
                const double god = upw == 1.0 ? gud : 0.0, gnd = upw == 1.0 ? 0.0 : gud;
                const double dc = (1.0 - alpha - beta) * (vn - 2.0 * god) * upw +
                                  (1.0 - alpha - beta) * (vo + 2.0 * gnd) * (1.0 - upw) +
                                  ((alpha - 1.0) * upw + beta * (1.0 - upw)) * vo +
                                  (beta * upw + (alpha - 1.0) * (1.0 - upw)) * vn;
                const double fl = dc * ph;
                so[q] = own ? so[q] + fl : so[q] - fl;
            }
        }
        st6(souT + (int64_t)NCMP * r, so);
    }
}

void launch_defc(Ctx* ctx, const AssembleArgs& a, int scheme, const double* C4, double* gradTh, double* souT,
                 int grid) {
    const double N = a.N, ne = (double)a.nEll;
    {
        // theta own + gathers (counted once), face data, boundary, gradient out
        ProfScope ps(ctx, "k_grad_theta", N * (48 + 8 + 144) + ne * (4 + 1 + 4 + 32));
        hipLaunchKernelGGL(k_grad_theta, dim3(grid), dim3(BLK), 0, ctx->stream, a, gradTh);
    }
    ProfScope ps(ctx, "k_defc", N * (48 + 32 + 144 + 48) + ne * (4 + 1 + 4 + 8));
    switch (scheme) {
    case DS_CUBISTA:
        hipLaunchKernelGGL(k_defc<DS_CUBISTA>, dim3(grid), dim3(BLK), 0, ctx->stream, a, C4, gradTh, souT);
        break;
    case DS_MINMOD:
        hipLaunchKernelGGL(k_defc<DS_MINMOD>, dim3(grid), dim3(BLK), 0, ctx->stream, a, C4, gradTh, souT);
        break;
    case DS_SMART:
        hipLaunchKernelGGL(k_defc<DS_SMART>, dim3(grid), dim3(BLK), 0, ctx->stream, a, C4, gradTh, souT);
        break;
    case DS_WACEB:
        hipLaunchKernelGGL(k_defc<DS_WACEB>, dim3(grid), dim3(BLK), 0, ctx->stream, a, C4, gradTh, souT);
        break;
    case DS_SUPERBEE:
        hipLaunchKernelGGL(k_defc<DS_SUPERBEE>, dim3(grid), dim3(BLK), 0, ctx->stream, a, C4, gradTh, souT);
        break;
    default:
        throw Error{RT_E_BAD_ARG, "launch_defc: not a limited scheme"};
    }
    RT_HIP(hipGetLastError());
}

// ------------------------------------------- stress terms of the momentum eq.
// The explicit part of constitutiveEq::divTau(U) (constitutiveEq.C:72-132):
// fvc::div(tau/rho) with Gauss linear, per row the faces in increasing face id
// then the row's boundary faces (patch value tauB [6][NB], or the cell value
// when tauB is NULL: zeroGradient), then / V; minus, for STAB,
//  * BSD (1): fvc::laplacian(etaP/rho, U), Gauss linear with the snGrad scheme
//    LAP (0 corrected, 1 orthogonal, 2 uncorrected): per face
//    gamma |Sf| (dc (U_n - U_o) [+ k & interpolate(grad U)]), the boundary
//    faces' patch snGrad dcB (Ub - U_P);
//  * coupling (2): fvc::div((etaP/rho) grad(U)), Gauss linear, the patch value
//    of grad(U) being G_P + n (snGrad - n & G_P).
// Each face term is evaluated in the row's frame (d = C_other - C_row, Sf
// signed for the row's side), which negates the owner's value exactly, so
// both rows of a face see the reference's face value with its sign.
template <int STAB, int LAP>
__global__ void __launch_bounds__(BLK) k_div_tau(AssembleArgs a, const double* __restrict__ tau,
                                                 const double* __restrict__ tauB, double rho, double gam,
                                                 const double* __restrict__ G, const double* __restrict__ C4,
                                                 const double* __restrict__ bDc, double* __restrict__ out) {
    const int64_t NB = a.NB > 0 ? a.NB : 1;
    for (int64_t it = 0;; ++it) {
        const int64_t tl = tile_of(it, a.N, a.nRed);
        if (tl < 0) {
            if (tl == -1) break;
            continue;
        }
        const int32_t r = (int32_t)(tl * BLK + threadIdx.x);
        if (r >= a.N) continue;
        double t[6];
        ld6(tau + (int64_t)NCMP * r, t);
#pragma unroll
        for (int q = 0; q < 6; ++q) t[q] = t[q] / rho;
        double acc[3] = {0, 0, 0}, st[3] = {0, 0, 0};
        double Ur[3] = {0, 0, 0}, Cr[3] = {0, 0, 0}, Gr[9];
        if constexpr (STAB != 0) {
            Ur[0] = a.U[4 * (int64_t)r], Ur[1] = a.U[4 * (int64_t)r + 1], Ur[2] = a.U[4 * (int64_t)r + 2];
#pragma unroll
            for (int q = 0; q < 9; ++q) Gr[q] = G[9 * (int64_t)r + q];
            if constexpr (STAB == 1) Cr[0] = C4[4 * (int64_t)r], Cr[1] = C4[4 * (int64_t)r + 1], Cr[2] = C4[4 * (int64_t)r + 2];
        }
        const int s = r >> 6, lane = r & 63;
        const int64_t base = a.ellOff[s] + lane;
        const int width = a.ellWidth[s];
        for (int k = 0; k < width; ++k) {
            const int64_t e = base + (int64_t)SLICE * k;
            const int32_t fid = a.ellFid[e];
            if (fid == RT_FID_PAD) continue;
            const bool own = fid >= 0;
            const int32_t c = a.ellCol[base + (int64_t)SLICE * a.ellSlot[e]];
            double tc[6];
            ld6(tau + (int64_t)NCMP * c, tc);
            const double wf = a.eW[e];
            double tf[6];
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                const double tcq = tc[q] / rho;
                const double vo = own ? t[q] : tcq, vn = own ? tcq : t[q];
                tf[q] = wf * (vo - vn) + vn;
            }
            // (+-Sf) & tf, the sign carried by the entry
            const double sf[3] = {a.eSf[e], a.eSf[a.nEll + e], a.eSf[2 * a.nEll + e]};
            acc[0] += sf[0] * tf[0] + sf[1] * tf[1] + sf[2] * tf[2];
            acc[1] += sf[0] * tf[1] + sf[1] * tf[3] + sf[2] * tf[4];
            acc[2] += sf[0] * tf[2] + sf[1] * tf[4] + sf[2] * tf[5];
            if constexpr (STAB != 0) {
                double Gc[9];
#pragma unroll
                for (int q = 0; q < 9; ++q) Gc[q] = G[9 * (int64_t)c + q];
                if constexpr (STAB == 2) {
                    double gf[9];
#pragma unroll
                    for (int q = 0; q < 9; ++q) {
                        const double pr = gam * Gr[q], pc = gam * Gc[q];
                        const double vo = own ? pr : pc, vn = own ? pc : pr;
                        gf[q] = wf * (vo - vn) + vn;
                    }
#pragma unroll
                    for (int j = 0; j < 3; ++j) st[j] += sf[0] * gf[j] + sf[1] * gf[3 + j] + sf[2] * gf[6 + j];
                } else {
                    const double Cc[3] = {C4[4 * (int64_t)c], C4[4 * (int64_t)c + 1], C4[4 * (int64_t)c + 2]};
                    const double Uc[3] = {a.U[4 * (int64_t)c], a.U[4 * (int64_t)c + 1], a.U[4 * (int64_t)c + 2]};
                    const double d[3] = {Cc[0] - Cr[0], Cc[1] - Cr[1], Cc[2] - Cr[2]};
                    const double ms = sqrt(sf[0] * sf[0] + sf[1] * sf[1] + sf[2] * sf[2]);
                    const double nf[3] = {sf[0] / ms, sf[1] / ms, sf[2] / ms};
                    const double md = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                    const double dc = LAP == 1 ? 1.0 / md : 1.0 / fmax(nf[0] * d[0] + nf[1] * d[1] + nf[2] * d[2], 0.05 * md);
                    double sn[3];
#pragma unroll
                    for (int j = 0; j < 3; ++j) sn[j] = dc * (Uc[j] - Ur[j]);
                    if constexpr (LAP == 0) {
                        const double kv[3] = {nf[0] - d[0] * dc, nf[1] - d[1] * dc, nf[2] - d[2] * dc};
                        double gf[9];
#pragma unroll
                        for (int q = 0; q < 9; ++q) {
                            const double vo = own ? Gr[q] : Gc[q], vn = own ? Gc[q] : Gr[q];
                            gf[q] = wf * (vo - vn) + vn;
                        }
#pragma unroll
                        for (int j = 0; j < 3; ++j) sn[j] = sn[j] + (kv[0] * gf[j] + kv[1] * gf[3 + j] + kv[2] * gf[6 + j]);
                    }
                    const double gms = gam * ms;
#pragma unroll
                    for (int j = 0; j < 3; ++j) st[j] += gms * sn[j];
                }
            }
        }
        const int32_t b0 = a.bPtr[r], b1 = a.bPtr[r + 1];
        for (int32_t qb = b0; qb < b1; ++qb) {
            const int32_t bf = a.bFace[qb];
            const double sf[3] = {a.bSf[bf], a.bSf[NB + bf], a.bSf[2 * NB + bf]};
            double tb[6];
#pragma unroll
            for (int q = 0; q < 6; ++q) tb[q] = tauB ? tauB[(int64_t)q * NB + bf] / rho : t[q];
            if (!tauB) {
                // a constraint patch gives every field its type: tau takes the
                // transform theta takes there (symmetry: mirrored, wedge: rotated)
                const int bt = a.bcType[a.bPatch[bf]];
                if (bt == RT_BC_SYMMETRY) symmetry_value(sf, t, tb);
                else if (bt == RT_BC_WEDGE) wedge_value(a.patchT + 9 * (int64_t)a.bPatch[bf], t, tb);
            }
            acc[0] += sf[0] * tb[0] + sf[1] * tb[1] + sf[2] * tb[2];
            acc[1] += sf[0] * tb[1] + sf[1] * tb[3] + sf[2] * tb[4];
            acc[2] += sf[0] * tb[2] + sf[1] * tb[4] + sf[2] * tb[5];
            if constexpr (STAB != 0) {
                const double ub[3] = {a.Ub[bf], a.Ub[NB + bf], a.Ub[2 * NB + bf]};
                const double ms = sqrt(sf[0] * sf[0] + sf[1] * sf[1] + sf[2] * sf[2]);
                const double dcb = bDc[bf];
                double sn[3];
#pragma unroll
                for (int j = 0; j < 3; ++j) sn[j] = dcb * (ub[j] - Ur[j]);
                if constexpr (STAB == 1) {
                    const double gms = gam * ms;
#pragma unroll
                    for (int j = 0; j < 3; ++j) st[j] += gms * sn[j];
                } else {
                    const double n[3] = {sf[0] / ms, sf[1] / ms, sf[2] / ms};
                    double nG[3], gpb[9];
#pragma unroll
                    for (int j = 0; j < 3; ++j) nG[j] = n[0] * Gr[j] + n[1] * Gr[3 + j] + n[2] * Gr[6 + j];
#pragma unroll
                    for (int i = 0; i < 3; ++i)
#pragma unroll
                        for (int j = 0; j < 3; ++j) gpb[3 * i + j] = gam * (Gr[3 * i + j] + n[i] * (sn[j] - nG[j]));
#pragma unroll
                    for (int j = 0; j < 3; ++j) st[j] += sf[0] * gpb[j] + sf[1] * gpb[3 + j] + sf[2] * gpb[6 + j];
                }
            }
        }
        const double vol = a.V[r];
#pragma unroll
        for (int j = 0; j < 3; ++j) out[3 * (int64_t)r + j] = STAB == 0 ? acc[j] / vol : acc[j] / vol - st[j] / vol;
    }
}

// fvc::grad(U) Gauss linear of row r (boilerLog.H:1): the row's faces in
// increasing face id, then its boundary faces with the patch values Ub, then
// / V -- the same sums, in the same order, as the assembly's gather.
__device__ __forceinline__ void gauss_grad_u(const AssembleArgs& a, int32_t r, double* g) {
    const int64_t NB = a.NB > 0 ? a.NB : 1;
    const int64_t nE = a.nEll;
    const double2 u01 = reinterpret_cast<const double2*>(a.U)[2 * (int64_t)r];
    const double Ui[3] = {u01.x, u01.y, a.U[4 * (int64_t)r + 2]};
#pragma unroll
    for (int q = 0; q < 9; ++q) g[q] = 0.0;
    const int s = r >> 6, lane = r & 63;
    const int64_t base = a.ellOff[s] + lane;
    const int width = a.ellWidth[s];
    for (int k = 0; k < width; ++k) {
        const int64_t e = base + (int64_t)SLICE * k;
        const int32_t fid = a.ellFid[e];
        if (fid == RT_FID_PAD) continue;
        const bool own = fid >= 0;
        const int32_t c = a.ellCol[base + (int64_t)SLICE * a.ellSlot[e]];
        const double wf = a.eW[e];
        const double2 c01 = reinterpret_cast<const double2*>(a.U)[2 * (int64_t)c];
        const double Uc[3] = {c01.x, c01.y, a.U[4 * (int64_t)c + 2]};
        double uf[3];
#pragma unroll
        for (int q3 = 0; q3 < 3; ++q3) {
            const double uo = own ? Ui[q3] : Uc[q3], un = own ? Uc[q3] : Ui[q3];
            uf[q3] = wf * (uo - un) + un;
        }
        const double sf[3] = {a.eSf[e], a.eSf[nE + e], a.eSf[2 * nE + e]};
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) g[3 * i + j] += sf[i] * uf[j];
    }
    const int32_t b0 = a.bPtr[r], b1 = a.bPtr[r + 1];
    for (int32_t qb = b0; qb < b1; ++qb) {
        const int32_t bf = a.bFace[qb];
        const double sf[3] = {a.bSf[bf], a.bSf[NB + bf], a.bSf[2 * NB + bf]};
        const double ub[3] = {a.Ub[bf], a.Ub[NB + bf], a.Ub[2 * NB + bf]};
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) g[3 * i + j] += sf[i] * ub[j];
    }
    const double vol = a.V[r];
#pragma unroll
    for (int q = 0; q < 9; ++q) g[q] /= vol;
}

// tauTotal() = tau + etaS (symm(L + L^T) - 2/3 tr(L) I), L = fvc::grad(U) Gauss
// linear, the same gather as the assembly's (constitutiveEq.C:534-547).
__global__ void __launch_bounds__(BLK) k_tau_total(AssembleArgs a, const double* __restrict__ tau, double etaS,
                                                   double* __restrict__ out) {
    for (int64_t it = 0;; ++it) {
        const int64_t tl = tile_of(it, a.N, a.nRed);
        if (tl < 0) {
            if (tl == -1) break;
            continue;
        }
        const int32_t r = (int32_t)(tl * BLK + threadIdx.x);
        if (r >= a.N) continue;
        double g[9];
        gauss_grad_u(a, r, g);
        double T[9];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) T[3 * i + j] = g[3 * i + j] + g[3 * j + i];
        const double sy[6] = {T[0], 0.5 * (T[1] + T[3]), 0.5 * (T[2] + T[6]), T[4], 0.5 * (T[5] + T[7]), T[8]};
        const double d = (2. / 3.) * (g[0] + g[4] + g[8]);
        const double I[6] = {d, 0.0, 0.0, d, 0.0, d};
        double t[6];
        ld6(tau + (int64_t)NCMP * r, t);
#pragma unroll
        for (int q = 0; q < 6; ++q) t[q] = t[q] + etaS * (sy[q] - I[q]);
        st6(out + (int64_t)NCMP * r, t);
    }
}

// fvc::grad(U) of every row -> G [N][9] (the cell values the stabilization
// terms of divTau interpolate to the faces)
__global__ void __launch_bounds__(BLK) k_grad_u(AssembleArgs a, double* __restrict__ G) {
    for (int64_t it = 0;; ++it) {
        const int64_t tl = tile_of(it, a.N, a.nRed);
        if (tl < 0) {
            if (tl == -1) break;
            continue;
        }
        const int32_t r = (int32_t)(tl * BLK + threadIdx.x);
        if (r >= a.N) continue;
        double g[9];
        gauss_grad_u(a, r, g);
#pragma unroll
        for (int q = 0; q < 9; ++q) G[9 * (int64_t)r + q] = g[q];
    }
}

void launch_div_tau(Ctx* ctx, const AssembleArgs& a, const double* tau, const double* tauB, double rho,
                    int stab, int lap, double gam, double* G, const double* C4, const double* bDc, double* out,
                    int grid) {
    if (stab != 0) {
        ProfScope ps(ctx, "k_grad_u", (double)a.N * (32 + 8 + 72) + (double)a.nEll * (4 + 1 + 4 + 32));
        hipLaunchKernelGGL(k_grad_u, dim3(grid), dim3(BLK), 0, ctx->stream, a, G);
    }
    ProfScope ps(ctx, "k_div_tau", (double)a.N * (48 + 8 + 24) + (double)a.nEll * (4 + 1 + 4 + 32));
#define RT_DIVTAU(S, L) \
    hipLaunchKernelGGL((k_div_tau<S, L>), dim3(grid), dim3(BLK), 0, ctx->stream, a, tau, tauB, rho, gam, G, C4, bDc, out)
    if (stab == 0) RT_DIVTAU(0, 0);
    else if (stab == 2) RT_DIVTAU(2, 0);
    else if (lap == 0) RT_DIVTAU(1, 0);
    else if (lap == 1) RT_DIVTAU(1, 1);
    else RT_DIVTAU(1, 2);
#undef RT_DIVTAU
    RT_HIP(hipGetLastError());
}

void launch_tau_total(Ctx* ctx, const AssembleArgs& a, const double* tau, double etaS, double* out, int grid) {
    ProfScope ps(ctx, "k_tau_total", (double)a.N * (32 + 8 + 48 + 48) + (double)a.nEll * (4 + 1 + 4 + 32));
    hipLaunchKernelGGL(k_tau_total, dim3(grid), dim3(BLK), 0, ctx->stream, a, tau, etaS, out);
    RT_HIP(hipGetLastError());
}

// Workgroups of k_assemble<T, KIND> that fit on the chip at once (VGPRs allow
// three per CU).  The kernel walks its tiles in a loop, so a grid larger than
// that runs in rounds: the later workgroups start when the first have walked
// their whole share, and neither walk keeps the neighbours it gathers in L2.
template <class T, int KIND>
static int resident_grid(const Ctx* ctx, int cap) {
    static int perCU = 0;  // the same on every gfx950 device
    if (perCU == 0) {
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_assemble<T, KIND>, BLK, 0) != hipSuccess || nb < 1)
            nb = 1;
        perCU = nb;
    }
    int64_t g = std::min<int64_t>(cap, (int64_t)ctx->numCUs * perCU);
    if (g >= 64) g = g / 8 * 8;
    return (int)std::max<int64_t>(1, g);
}

template <class T, int KIND>
static int launch_assemble_one(Ctx* ctx, const AssembleArgs& a, int cap) {
    const int grid = resident_grid<T, KIND>(ctx, cap);
    hipLaunchKernelGGL((k_assemble<T, KIND>), dim3(grid), dim3(BLK), 0, ctx->stream, a);
    return grid;
}

template <class T>
static int launch_assemble_kind(Ctx* ctx, const AssembleArgs& a, int cap) {
    switch (a.prm.kind) {
    case M_OLDROYDB: return launch_assemble_one<T, M_OLDROYDB>(ctx, a, cap);
    case M_GIESEKUS: return launch_assemble_one<T, M_GIESEKUS>(ctx, a, cap);
    case M_FENEP: return launch_assemble_one<T, M_FENEP>(ctx, a, cap);
    case M_PTT_LINEAR: return launch_assemble_one<T, M_PTT_LINEAR>(ctx, a, cap);
    case M_PTT_EXP: return launch_assemble_one<T, M_PTT_EXP>(ctx, a, cap);
    case M_PTT_GEN: return launch_assemble_one<T, M_PTT_GEN>(ctx, a, cap);
    case M_ROOTK: return launch_assemble_one<T, M_ROOTK>(ctx, a, cap);
    case M_FENECR: return launch_assemble_one<T, M_FENECR>(ctx, a, cap);
    case M_WMCY: return launch_assemble_one<T, M_WMCY>(ctx, a, cap);
    case M_ROLIEPOLY: return launch_assemble_one<T, M_ROLIEPOLY>(ctx, a, cap);
    case M_SARAMITO: return launch_assemble_one<T, M_SARAMITO>(ctx, a, cap);
    case M_BMP: return launch_assemble_one<T, M_BMP>(ctx, a, cap);
    case M_FLUIDITY: return launch_assemble_one<T, M_FLUIDITY>(ctx, a, cap);
    default: return launch_assemble_one<T, M_XPP>(ctx, a, cap);
    }
}

void launch_assemble(Ctx* ctx, const AssembleArgs& a, int grid) {
    // algorithmic bytes per cell: theta 48, eigen state 96, U 32, V 8, b 48, diag 8, 1/diag 8,
    // r and r0 (2 x 6 x w, unless the solve writes them), theta.oldTime() 48 more in inner
    // iterations, the old tau 48 for Saramito; per ELL entry: face id 4, slot 1, column 4,
    // phi 8, value 8 (+ Sf 24 and w 8 for the gradient)
    const double w = a.mixed ? 4 : 8;
    const double by = (double)a.N * (48 + 96 + 32 + 8 + 48 + 8 + 8 + (a.r ? 12 * w : 0) + (a.thetaOld ? 48 : 0) +
                                     (a.tauOld ? 48 : 0) + (a.aux ? 8 : 0)) +
                      (double)a.nEll * (4 + 1 + 4 + 8 + 8 + (a.gradU ? 0 : 24 + 8));
    int g;
    {
        ProfScope ps(ctx, a.mixed ? "k_assemble<f32>" : "k_assemble<f64>", by);
        g = a.mixed ? launch_assemble_kind<float>(ctx, a, grid) : launch_assemble_kind<double>(ctx, a, grid);
    }
    hipLaunchKernelGGL(k_fin_assemble, dim3(1), dim3(BLK), 0, ctx->stream, a, g);
    RT_HIP(hipGetLastError());
}
void launch_eig_tau(Ctx* ctx, int32_t N, const ModelParams& prm, const double* theta,
                    double* eigVal, double* eigVec, double* tau, const double* fOld) {
    // theta 48 + previous eigenvectors 72 in; eigenvalues 24, eigenvectors 72, tau 48 out
    // (tau NULL: the eigen state only)
    const double tauBytes = tau ? 48 + (modelKeepsFactor(prm.kind) ? 8 : 0) : 0;
    ProfScope ps(ctx, "k_eig_tau", (double)N * (48 + 72 + 24 + 72 + tauBytes));
    hipLaunchKernelGGL(k_eig_tau, dim3(nblk(N)), dim3(BLK), 0, ctx->stream, N, prm, theta, eigVal,
                       eigVec, tau, fOld);
    RT_HIP(hipGetLastError());
}
void launch_aos_to_soa(Ctx* ctx, int32_t N, int nc, const double* aos, const int32_t* perm, double* soa) {
    hipLaunchKernelGGL(k_aos_to_soa, dim3(nblk(N)), dim3(BLK), 0, ctx->stream, N, nc, aos, perm, soa);
    RT_HIP(hipGetLastError());
}
void launch_soa_to_aos(Ctx* ctx, int32_t N, int nc, const double* soa, const int32_t* perm, double* aos) {
    hipLaunchKernelGGL(k_soa_to_aos, dim3(nblk(N)), dim3(BLK), 0, ctx->stream, N, nc, soa, perm, aos);
    RT_HIP(hipGetLastError());
}
void launch_eigvals_from_tensor(Ctx* ctx, int32_t N, const double* aos9, const int32_t* perm, double* ev) {
    hipLaunchKernelGGL(k_eigvals_from_tensor, dim3(nblk(N)), dim3(BLK), 0, ctx->stream, N, aos9, perm, ev);
    RT_HIP(hipGetLastError());
}
void launch_eigvals_to_tensor(Ctx* ctx, int32_t N, const double* ev, const int32_t* perm, double* aos9) {
    hipLaunchKernelGGL(k_eigvals_to_tensor, dim3(nblk(N)), dim3(BLK), 0, ctx->stream, N, ev, perm, aos9);
    RT_HIP(hipGetLastError());
}
void launch_copy(Ctx* ctx, const void* src, void* dst, size_t bytes) {
    const int64_t n16 = (int64_t)(bytes / 16);
    const unsigned grid = (unsigned)((n16 + (int64_t)BLK * COPY_U - 1) / ((int64_t)BLK * COPY_U));
    ProfScope ps(ctx, "k_copy", 2.0 * (double)n16 * 16);
    hipLaunchKernelGGL(k_copy, dim3(grid), dim3(BLK), 0, ctx->stream, n16, (const double2*)src, (double2*)dst);
    RT_HIP(hipGetLastError());
}
void launch_add(Ctx* ctx, int64_t n, const double* in, double* out) {
    if (n <= 0) return;
    ProfScope ps(ctx, "k_add", (double)n * 24);
    hipLaunchKernelGGL(k_add, dim3((unsigned)((n + BLK - 1) / BLK)), dim3(BLK), 0, ctx->stream, n, in, out);
    RT_HIP(hipGetLastError());
}

void launch_fill(Ctx* ctx, double* p, int64_t n, double v) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill, dim3(nblk(n)), dim3(BLK), 0, ctx->stream, p, n, v);
    RT_HIP(hipGetLastError());
}
void launch_ldu_to_ell(Ctx* ctx, const Mesh& m, const double* lower, const double* upper, double* aval) {
    if (m.nEll <= 0) return;
    hipLaunchKernelGGL(k_ldu_to_ell, dim3(nblk(m.nEll)), dim3(BLK), 0, ctx->stream, m.nEll, m.nSlices,
                       m.ellOff.as<int64_t>(), m.ellFid.as<int32_t>(), m.ellSlot.as<uint8_t>(), lower,
                       upper, aval);
    RT_HIP(hipGetLastError());
}
void launch_face_to_entry(Ctx* ctx, const Mesh& m, const double* phi, double* ePhi) {
    if (m.nEll <= 0) return;
    hipLaunchKernelGGL(k_face_to_entry, dim3(nblk(m.nEll)), dim3(BLK), 0, ctx->stream, m.nEll,
                       m.ellFid.as<int32_t>(), phi, ePhi);
    RT_HIP(hipGetLastError());
}
void launch_permute_rows(Ctx* ctx, int32_t N, int nc, int ncOut, const double* in, const int32_t* perm,
                         double* out) {
    hipLaunchKernelGGL(k_permute_rows, dim3(nblk(N)), dim3(BLK), 0, ctx->stream, N, nc, ncOut, in, perm, out);
    RT_HIP(hipGetLastError());
}
void launch_unpermute_rows(Ctx* ctx, int32_t N, int ncIn, int nc, const double* in, const int32_t* perm,
                           double* out) {
    hipLaunchKernelGGL(k_unpermute_rows, dim3(nblk(N)), dim3(BLK), 0, ctx->stream, N, ncIn, nc, in, perm,
                       out);
    RT_HIP(hipGetLastError());
}

}  // namespace rt
