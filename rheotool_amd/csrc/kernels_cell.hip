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
#include "rt_internal.h"

namespace rt {

constexpr int BLK = 256;

__device__ __forceinline__ int xcd_block(int G) {
    // Consecutive logical blocks on one XCD so neighbouring rows share an L2.
    const int b = blockIdx.x;
    if (G % 8 != 0) return b;
    return (b % 8) * (G / 8) + b / 8;
}

__device__ __forceinline__ void block_rows(int32_t N, int32_t& r0, int32_t& r1) {
    const int G = gridDim.x;
    const int b = xcd_block(G);
    int64_t per = ((int64_t)N + G - 1) / G;
    per = (per + BLK - 1) / BLK * BLK;
    int64_t a = (int64_t)b * per;
    r0 = (int32_t)(a < N ? a : N);
    r1 = (int32_t)(a + per < N ? a + per : N);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Block reduction of NQ values per thread into partials[block][NQ]; the last
// block to finish reduces all blocks in a fixed order (deterministic) and
// returns true with sums[] filled (thread 0 valid).
template <int NQ>
__device__ bool block_reduce_last(double (&v)[NQ], double* partials, int32_t* counter,
                                  double (&sums)[NQ]) {
    __shared__ double red[NQ][BLK / 64];
    __shared__ int amLast;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        double s = wave_sum(v[q]);
        if (lane == 0) red[q][wid] = s;
    }
    __syncthreads();
    if (threadIdx.x < NQ) {
        double s = 0;
        for (int w = 0; w < BLK / 64; ++w) s += red[threadIdx.x][w];
        partials[(int64_t)blockIdx.x * NQ + threadIdx.x] = s;
    }
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) amLast = atomicAdd(counter, 1) == (int)gridDim.x - 1;
    __syncthreads();
    if (!amLast) return false;
    __threadfence();
    const volatile double* vp = partials;
    double acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = 0;
    for (int i = threadIdx.x; i < (int)gridDim.x; i += BLK)
#pragma unroll
        for (int q = 0; q < NQ; ++q) acc[q] += vp[(int64_t)i * NQ + q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        double s = wave_sum(acc[q]);
        if (lane == 0) red[q][wid] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            double s = 0;
            for (int w = 0; w < BLK / 64; ++w) s += red[q][w];
            sums[q] = s;
        }
        *counter = 0;
    }
    return true;
}

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

// Log-conformation right-hand side S (6 comps) of one cell.
__device__ void cell_rhs(const ModelParams& p, int invMask, const double* L, const double* th6,
                         const double* lam, const double* V, double* S, double& fOut) {
    const bool ptt = p.kind == M_PTT_LINEAR || p.kind == M_PTT_EXP;
    double LT[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) LT[3 * i + j] = L[3 * j + i];
    if (ptt) {
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
#pragma unroll
    for (int q = 0; q < 9; ++q) R[q] = tmp[q] - t2[q] + 2.0 * B[q];
    double il[3];
    invDiag(lam, invMask, il);
    const double rl = 1.0 / p.lambda;
    double X[9];
    fOut = 0.0;
    if (p.kind == M_OLDROYDB) {
        const double dd[3] = {il[0] - 1.0, il[1] - 1.0, il[2] - 1.0};
        double Y[9];
        vdvt(V, dd, Y);
#pragma unroll
        for (int q = 0; q < 9; ++q) X[q] = rl * Y[q];
    } else if (p.kind == M_GIESEKUS) {
        const double dd[3] = {il[0] - 1.0, il[1] - 1.0, il[2] - 1.0};
        double trhs[9], C[9], tt[9], CT[9];
        vdvt(V, dd, trhs);
        vdvt(V, lam, C);
        dotTT(trhs, trhs, tt);
        dotTT(C, tt, CT);
#pragma unroll
        for (int q = 0; q < 9; ++q) X[q] = rl * (trhs[q] - p.alpha * CT[q]);
    } else if (p.kind == M_FENEP) {
        double C[9], Y[9];
        vdvt(V, lam, C);
        const double f = p.L2 / (p.L2 - (C[0] + C[4] + C[8]));
        const double a = p.L2 / (p.L2 - 3.);
        fOut = f;
        const double dd[3] = {a * il[0] - f, a * il[1] - f, a * il[2] - f};
        vdvt(V, dd, Y);
#pragma unroll
        for (int q = 0; q < 9; ++q) X[q] = rl * Y[q];
    } else {
        const double dd[3] = {il[0] - 1.0, il[1] - 1.0, il[2] - 1.0};
        double Y[9], C[9];
        vdvt(V, dd, Y);
        vdvt(V, lam, C);
        const double k = p.epsilon / (1 - p.zeta);
        const double z = k * (C[0] + C[4] + C[8] - 3.);
        const double fac = p.kind == M_PTT_LINEAR ? 1. + z : exp(z);
#pragma unroll
        for (int q = 0; q < 9; ++q) X[q] = (rl * Y[q]) * fac;
    }
#pragma unroll
    for (int q = 0; q < 9; ++q) R[q] = R[q] + X[q];
    symmT(R, S);
}

__device__ __forceinline__ void load6(const double* p, double* v) {
    const double2* q = reinterpret_cast<const double2*>(p);
    const double2 a = q[0], b = q[1], c = q[2];
    v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y; v[4] = c.x; v[5] = c.y;
}
__device__ __forceinline__ void store6(double* p, const double* v) {
    double2* q = reinterpret_cast<double2*>(p);
    q[0] = make_double2(v[0], v[1]);
    q[1] = make_double2(v[2], v[3]);
    q[2] = make_double2(v[4], v[5]);
}

__global__ void __launch_bounds__(BLK) k_assemble(AssembleArgs a) {
    int32_t r0, r1;
    block_rows(a.N, r0, r1);
    double thsum[6] = {0, 0, 0, 0, 0, 0};
    const int64_t nE = a.nEll;
    for (int32_t r = r0 + threadIdx.x; r < r1; r += BLK) {
        const int32_t N = a.N;
        double th[6], lam[3], V[9];
        load6(a.theta + (int64_t)NCMP * r, th);
#pragma unroll
        for (int q = 0; q < 3; ++q) lam[q] = a.eigVal[(int64_t)q * N + r];
#pragma unroll
        for (int q = 0; q < 9; ++q) V[q] = a.eigVec[(int64_t)q * N + r];
        const double vol = a.V[r];
        const double2 u01 = reinterpret_cast<const double2*>(a.U)[2 * (int64_t)r];
        const double Ui[3] = {u01.x, u01.y, a.U[4 * (int64_t)r + 2]};
        const int s = r >> 6, lane = r & 63;
        const int64_t base = a.ellOff[s] + lane;
        const int width = a.ellWidth[s];
        double g[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        double divDiag = 0.0, sumOff = 0.0;
        const bool haveGrad = a.gradU != nullptr;
        for (int k = 0; k < width; ++k) {
            const int64_t e = base + (int64_t)SLICE * k;  // face-order entry
            const int32_t fid = a.ellFid[e];
            if (fid == RT_FID_PAD) {
                a.aval[base + (int64_t)SLICE * a.ellSlot[e]] = 0.0;
                continue;
            }
            const bool own = fid >= 0;
            const int64_t ec = base + (int64_t)SLICE * a.ellSlot[e];  // column-order slot
            const double ph = a.ePhi[e];
            const double upw = ph >= 0 ? 1.0 : 0.0;
            const double lower = -upw * ph;
            const double upper = (1.0 - upw) * ph;
            const double val = own ? upper : lower;
            a.aval[ec] = val;
            divDiag -= own ? lower : upper;
            sumOff += fabs(val);
            if (!haveGrad) {
                const int32_t c = a.ellCol[ec];
                const double wf = a.eW[e];
                const double2 c01 = reinterpret_cast<const double2*>(a.U)[2 * (int64_t)c];
                const double Uc[3] = {c01.x, c01.y, a.U[4 * (int64_t)c + 2]};
                double uf[3];
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    const double uo = own ? Ui[q] : Uc[q], un = own ? Uc[q] : Ui[q];
                    uf[q] = wf * (uo - un) + un;
                }
                // Sf already carries the row's side: g += (+-Sf) (x) U_f
                const double sf[3] = {a.eSf[e], a.eSf[nE + e], a.eSf[2 * nE + e]};
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j) g[3 * i + j] += sf[i] * uf[j];
            }
        }
        const int32_t b0 = a.bPtr[r], b1 = a.bPtr[r + 1];
        const int64_t NB = a.NB > 0 ? a.NB : 1;
        if (!haveGrad) {
            for (int32_t q = b0; q < b1; ++q) {
                const int32_t bf = a.bFace[q];
                const double sf[3] = {a.bSf[bf], a.bSf[NB + bf], a.bSf[2 * NB + bf]};
                const double ub[3] = {a.Ub[bf], a.Ub[NB + bf], a.Ub[2 * NB + bf]};
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j) g[3 * i + j] += sf[i] * ub[j];
            }
#pragma unroll
            for (int q = 0; q < 9; ++q) g[q] /= vol;
        } else {
#pragma unroll
            for (int q = 0; q < 9; ++q) g[q] = a.gradU[(int64_t)q * N + r];
        }
        double S[6], f;
        cell_rhs(a.prm, a.invMask, g, th, lam, V, S, f);
        if (a.fOld) a.fOld[r] = f;
        // matrix diagonal: ddt + div, optional relax, then boundary internalCoeffs
        double D = a.rDeltaT * vol + divDiag;
        double src[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) src[q] = a.rDeltaT * th[q] * vol + vol * S[q];
        if (a.relax > 0) {
            const double D0 = D;
            for (int32_t q = b0; q < b1; ++q) {
                const int32_t bf = a.bFace[q];
                const int t = a.bcType[a.bPatch[bf]];
                if (t == RT_BC_EMPTY) continue;
                const double ph = a.phiB[bf];
                D += t == RT_BC_ZEROGRADIENT ? fabs(ph * 1.0) : fabs(ph * 0.0);
            }
            D = fmax(fabs(D), sumOff);
            D /= a.relax;
            for (int32_t q = b0; q < b1; ++q) {
                const int32_t bf = a.bFace[q];
                const int t = a.bcType[a.bPatch[bf]];
                if (t == RT_BC_EMPTY) continue;
                const double ph = a.phiB[bf];
                D -= t == RT_BC_ZEROGRADIENT ? ph * 1.0 : ph * 0.0;
            }
#pragma unroll
            for (int q = 0; q < 6; ++q) src[q] += (D - D0) * th[q];
        }
        for (int32_t q = b0; q < b1; ++q) {
            const int32_t bf = a.bFace[q];
            const int t = a.bcType[a.bPatch[bf]];
            if (t == RT_BC_EMPTY) continue;
            const double ph = a.phiB[bf];
            if (t == RT_BC_ZEROGRADIENT) {
                D += ph * 1.0;
            } else {
                D += ph * 0.0;
#pragma unroll
                for (int c = 0; c < 6; ++c) src[c] += -ph * a.bcVal[(int64_t)c * NB + bf];
            }
        }
        a.diag[r] = D;
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            if (a.valid[q] <= 0) src[q] = 0.0;
            thsum[q] += th[q];
        }
        store6(a.b + (int64_t)NCMP * r, src);
    }
    double sums[6];
    if (block_reduce_last<6>(thsum, a.partials, a.counter, sums) && threadIdx.x == 0) {
        for (int q = 0; q < 6; ++q) a.cs[q].xbar = sums[q] / a.N;
    }
}

// ----------------------------------------------------------- calcEig + tau
// Symmetric 3x3 eigen-decomposition by cyclic Jacobi rotations (Rutishauser's
// stable update), ascending order; column eigenvectors in V (row-major).
__device__ void jacobi3(const double* t6, double* ev, double* V) {
    double a00 = t6[0], a01 = t6[1], a02 = t6[2], a11 = t6[3], a12 = t6[4], a22 = t6[5];
    double v[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
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

__global__ void __launch_bounds__(BLK) k_eig_tau(int32_t N, ModelParams p, const double* __restrict__ theta,
                                                 double* __restrict__ eigVal, double* __restrict__ eigVec,
                                                 double* __restrict__ tau, const double* __restrict__ fOld) {
    const int64_t r = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (r >= N) return;
    double th[6], ev[3], V[9];
    load6(theta + (int64_t)NCMP * r, th);
    jacobi3(th, ev, V);
    double lam[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        lam[q] = exp(ev[q]);
        eigVal[q * (int64_t)N + r] = lam[q];
    }
#pragma unroll
    for (int q = 0; q < 9; ++q) eigVec[q * (int64_t)N + r] = V[q];
    double C[9], D[9];
    vdvt(V, lam, C);
    double coef;
    if (p.kind == M_FENEP) {
        const double a = p.L2 / (p.L2 - 3.);
        const double f = fOld[r];
        const double I9[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
#pragma unroll
        for (int q = 0; q < 9; ++q) D[q] = f * C[q] - a * I9[q];
        coef = p.etaP / p.lambda;
    } else {
        const double I9[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
#pragma unroll
        for (int q = 0; q < 9; ++q) D[q] = C[q] - I9[q];
        coef = (p.kind == M_PTT_LINEAR || p.kind == M_PTT_EXP) ? p.etaP / (p.lambda * (1 - p.zeta))
                                                               : p.etaP / p.lambda;
    }
    double s[6];
    symmT(D, s);
#pragma unroll
    for (int q = 0; q < 6; ++q) s[q] = coef * s[q];
    store6(tau + (int64_t)NCMP * r, s);
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

// ---------------------------------------------------------------- launchers
static inline unsigned nblk(int64_t n) { return (unsigned)((n + BLK - 1) / BLK); }

void launch_assemble(Ctx* ctx, const AssembleArgs& a, int grid) {
    // algorithmic bytes per cell: theta 48, eigen state 96, U 32, V 8, b 48, diag 8;
    // per ELL entry: face id 4, slot 1, Sf 24, w 8, phi 8, value 8 (+ column 4 for the gradient)
    const double by = (double)a.N * (48 + 96 + 32 + 8 + 48 + 8) +
                      (double)a.nEll * (4 + 1 + 8 + 8 + (a.gradU ? 0 : 24 + 8 + 4));
    ProfScope ps(ctx, "k_assemble", by);
    hipLaunchKernelGGL(k_assemble, dim3(grid), dim3(BLK), 0, ctx->stream, a);
    RT_HIP(hipGetLastError());
}
void launch_eig_tau(Ctx* ctx, int32_t N, const ModelParams& prm, const double* theta,
                    double* eigVal, double* eigVec, double* tau, const double* fOld) {
    ProfScope ps(ctx, "k_eig_tau", (double)N * (48 + 24 + 72 + 48 + (prm.kind == M_FENEP ? 8 : 0)));
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
