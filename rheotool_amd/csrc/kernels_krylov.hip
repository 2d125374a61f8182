// kernels_krylov.hip -- lockstep multi-component PBiCGStab for the theta solve (gfx950).
//
// The reference solves each valid symmTensor component separately with
// OpenFOAM's PBiCGStab (the segregated loop restated in
// sparseMatrixSolvers/segregated/eigenSolver/eigenSolver.C:422-562, residual
// normalisation sparseSolver.C:129-180).  The components share the matrix, so
// here one pass over the matrix serves every component still iterating:
// each kernel walks the sliced-ELL rows once and updates all active
// components (stored interleaved, [N][6]) with per-component convergence
// masks.  Every scalar of the recurrence (rho, alpha, omega, beta, residuals,
// flags) stays on the device: each row kernel writes per-workgroup partial
// sums and a one-workgroup finalize kernel on the same stream reduces them in
// a fixed order and updates the component records, so no host round trip is
// needed inside an iteration.
//
// Preconditioning is right-Jacobi with the column-scaled matrix
// Ahat = A D^-1 stored explicitly: A M^-1 p = p + offdiag(Ahat) p.
//
// Precision: the vector type T is double (FP64 drop-in) or float (inner
// solve of the FP64-refined "mixed" solve).  x (theta), b, the matrix used
// for the true residual and the grid-level reductions are FP64; per-thread
// partial sums use T.
#include "rt_device.h"

namespace rt {

template <class T>
struct KryArgs {
    int32_t N;
    const int64_t* ellOff;
    const int32_t* ellWidth;
    const int32_t* ellCol;
    const double* aval;  // FP64 matrix (true residual)
    const uint32_t* slotMask;  // [nSlices] live column-order slots, NULL: all live
    const double* diag;
    const T* ahat;       // A D^-1 off-diagonal values
    const double* rD;    // 1 / diag
    const double* b;     // [N][6]
    double* x;           // [N][6]
    T *r, *r0, *p, *v, *t;  // [N][6]
    T* e;                // [N][6] mixed: the inner solve's correction to x (NULL in FP64)
    // preconditioner M (right preconditioning of PBiCGStab, M^-1 of PBiCG):
    // rM = M^-1 as a diagonal (1/diag for Jacobi, NULL for none: 1), and
    // dg = diag(A M^-1) (NULL: 1, Jacobi) -- D itself for none
    const double* rM;
    const double* dg;
    // PBiCG: wT = s, the off-diagonals of A and A^T in T, FP64 A^T for rT
    T* s;
    const T* aA;
    const T* aT;
    const double* atval;
    double* partials;
    CompState* cs;
    int32_t valid[6];
    KryCtl ctl;
    int32_t G;           // workgroups of the row kernels (partials rows)
};

// ---- true residual r = b - A x (FP64): MODE 0 initial, MODE 1 refinement ----
// A refinement sweep (MODE 1) waits until no inner solve is running; the
// gate is read on the device so a sweep may be queued ahead of knowing.
// Matrix streams of the FP32 solve's passes are read once per pass: load
// them non-temporal, which leaves more of L2 to the gathered neighbour rows
// (measured 1.5% faster on those passes at 256^3; the FP64 passes keep
// ordinary loads).
template <class T, class U>
__device__ __forceinline__ U ld_stream(const U* p) {
    if constexpr (std::is_same<T, float>::value) return __builtin_nontemporal_load(p);
    else return *p;
}

template <class T>
__device__ __forceinline__ bool inner_running(const KryArgs<T>& a) {
    bool run = false;
#pragma unroll
    for (int c = 0; c < 6; ++c) run |= a.valid[c] > 0 && a.cs[c].active;
    return run;
}

template <class T, int MODE>
__global__ void __launch_bounds__(BLK) k_residual(KryArgs<T> a) {
    constexpr int NQ = 18;
    bool on[6];
    bool any = false;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        on[c] = a.valid[c] > 0 && (MODE == 0 || a.cs[c].needRefine);
        any |= on[c];
    }
    if (!any || (MODE == 1 && inner_running(a))) return;
    double q[NQ];
#pragma unroll
    for (int i = 0; i < NQ; ++i) q[i] = 0;
    double xbar[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) xbar[c] = MODE == 0 ? a.cs[c].xbar : 0.0;
    for (int64_t it = 0;; ++it) {
        const int64_t tl = tile_of(it, a.N);
        if (tl < 0) break;
        const int64_t t0 = tl * BLK;
        const int32_t r = (int32_t)(t0 + threadIdx.x);
        if (r >= a.N) continue;
        const double d = a.diag[r];
        double y[6], xv[6];
        ld6(a.x + (int64_t)NCMP * r, xv);
#pragma unroll
        for (int c = 0; c < 6; ++c) y[c] = d * xv[c];
        double sumA = d;
        const int s = r >> 6, lane = r & 63;
        const int64_t base = a.ellOff[s] + lane;
        const int width = a.ellWidth[s];
        const uint32_t live = a.slotMask ? a.slotMask[s] : 0xffffffffu;
        for (int k = 0; k < width; ++k) {
            if (k < 32 && !((live >> k) & 1u)) continue;  // zero in every row of the slice
            const int64_t e = base + (int64_t)SLICE * k;
            const int32_t col = ld_stream<T>(a.ellCol + e);
            const double av = ld_stream<T>(a.aval + e);
            if (MODE == 0) sumA += av;
            if (av == 0.0) continue;  // upwind: the downwind side of a face has no coupling
            double xc[6];
            ld6(a.x + (int64_t)NCMP * col, xc);
#pragma unroll
            for (int c = 0; c < 6; ++c) y[c] += av * xc[c];
        }
        double bb[6];
        ld6(a.b + (int64_t)NCMP * r, bb);
        // r and r0 are written whole (three 2-wide stores each): the entries of
        // components that are off belong to finished or frozen solves, which
        // no later kernel reads
        T rt[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            const double res = bb[c] - y[c];
            rt[c] = (T)res;
            if (!on[c]) continue;
            q[c] += fabs(res);
            q[6 + c] += (double)rt[c] * (double)rt[c];
            if (MODE == 0) {
                const double pA = sumA * xbar[c];
                q[12 + c] += fabs(y[c] - pA) + fabs(bb[c] - pA);
            } else {
                // (M^-1 r, r): the rho of a PBiCG restart with rT = r
                const double rm = a.rM ? a.rM[r] : 1.0;
                q[12 + c] += (rm * (double)rt[c]) * (double)rt[c];
            }
        }
        st6(a.r + (int64_t)NCMP * r, rt);
        st6(a.r0 + (int64_t)NCMP * r, rt);
    }
    write_partials<NQ>(q, a.partials);
}

template <class T, int MODE>
__global__ void __launch_bounds__(BLK) k_fin_residual(KryArgs<T> a) {
    constexpr int NQ = 18;
    bool on[6];
    bool any = false;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        on[c] = a.valid[c] > 0 && (MODE == 0 || a.cs[c].needRefine);
        any |= on[c];
    }
    if (!any || (MODE == 1 && inner_running(a))) return;
    double sums[NQ];
    sum_partials<NQ>(a.partials, a.G, sums);
    if (threadIdx.x == 0) {
        for (int c = 0; c < 6; ++c) {
            if (!on[c]) continue;
            CompState& cs = a.cs[c];
            if (MODE == 0) {
                finish_initial(cs, a.ctl, sums[c], sums[6 + c], sums[12 + c]);
            } else {
                cs.needRefine = 0;
                cs.nRefine += 1;
                cs.finalRes = sums[c] / cs.normFactor;
                if (conv_outer(cs, a.ctl, cs.finalRes) && cs.iter >= a.ctl.minIter) {
                    cs.converged = 1;
                    cs.active = 0;
                } else if (cs.nRefine >= a.ctl.maxRefine || cs.iter >= a.ctl.maxIter) {
                    cs.active = 0;
                } else {
                    start_inner(cs, a.ctl, a.ctl.bicg ? sums[12 + c] : sums[6 + c]);
                }
            }
        }
    }
}

// ---- p = r  |  p = r + beta (p - omega v) ---------------------------------
template <class T>
__global__ void __launch_bounds__(BLK) k_kry_p(KryArgs<T> a) {
    bool on[6];
    bool first[6];
    T beta[6], omega[6];
    bool any = false, update = false;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        on[c] = a.cs[c].active;
        first[c] = a.cs[c].innerIter == 0;
        beta[c] = (T)a.cs[c].beta;
        omega[c] = (T)a.cs[c].omega;
        any |= on[c];
        update |= on[c] && !first[c];
    }
    if (!any) return;
    const int64_t N = a.N;
    if (!update) {
        // every active component starts an inner solve: p = r.  The other
        // components' p is dead until their own restart sets it.
        for (int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x; i < N; i += (int64_t)gridDim.x * BLK) {
            T rr[6];
            ld6(a.r + NCMP * i, rr);
            st6(a.p + NCMP * i, rr);
        }
        return;
    }
    for (int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x; i < N; i += (int64_t)gridDim.x * BLK) {
        T rr[6], pp[6], vv[6];
        ld6(a.r + NCMP * i, rr);
        ld6(a.p + NCMP * i, pp);
        ld6(a.v + NCMP * i, vv);
#pragma unroll
        for (int c = 0; c < 6; ++c)
            if (on[c]) pp[c] = first[c] ? rr[c] : rr[c] + beta[c] * (pp[c] - omega[c] * vv[c]);
        st6(a.p + NCMP * i, pp);
    }
}

// ---- v = A M^-1 p = p + Ahat p ;  sigma = (r0, v) -------------------------
template <class T>
__global__ void __launch_bounds__(BLK) k_kry_v(KryArgs<T> a) {
    bool on[6];
    bool any = false;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        on[c] = a.cs[c].active;
        any |= on[c];
    }
    if (!any) return;
    T q[6] = {0, 0, 0, 0, 0, 0};
    for (int64_t it = 0;; ++it) {
        const int64_t tl = tile_of(it, a.N);
        if (tl < 0) break;
        const int64_t t0 = tl * BLK;
        const int32_t r = (int32_t)(t0 + threadIdx.x);
        if (r >= a.N) continue;
        T acc[6];
        ld6(a.p + (int64_t)NCMP * r, acc);
        if (a.dg) {
            const T dgr = (T)a.dg[r];
#pragma unroll
            for (int c = 0; c < 6; ++c) acc[c] = dgr * acc[c];
        }
        const int s = r >> 6, lane = r & 63;
        const int64_t base = a.ellOff[s] + lane;
        const int width = a.ellWidth[s];
        const uint32_t live = a.slotMask ? a.slotMask[s] : 0xffffffffu;
        for (int k = 0; k < width; ++k) {
            if (k < 32 && !((live >> k) & 1u)) continue;  // zero in every row of the slice
            const int64_t e = base + (int64_t)SLICE * k;
            const int32_t col = ld_stream<T>(a.ellCol + e);
            const T av = ld_stream<T>(a.ahat + e);
            if (av == (T)0) continue;
            T pc[6];
            ld6(a.p + (int64_t)NCMP * col, pc);
#pragma unroll
            for (int c = 0; c < 6; ++c) acc[c] += av * pc[c];
        }
        st6(a.v + (int64_t)NCMP * r, acc);
        T rr0[6];
        ld6(a.r0 + (int64_t)NCMP * r, rr0);
#pragma unroll
        for (int c = 0; c < 6; ++c)
            if (on[c]) q[c] += rr0[c] * acc[c];
    }
    double qd[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) qd[c] = q[c];
    write_partials<6>(qd, a.partials);
}

template <class T>
__global__ void __launch_bounds__(BLK) k_fin_v(KryArgs<T> a) {
    bool on[6];
    bool any = false;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        on[c] = a.cs[c].active;
        any |= on[c];
    }
    if (!any) return;
    double sums[6];
    sum_partials<6>(a.partials, a.G, sums);
    if (threadIdx.x == 0)
        for (int c = 0; c < 6; ++c)
            if (on[c]) a.cs[c].alpha = a.cs[c].rho / sums[c];
}

// ---- s = r - alpha v ; t = A M^-1 s ; |s|, (t,s), (t,t) --------------------
template <class T>
__global__ void __launch_bounds__(BLK) k_kry_t(KryArgs<T> a) {
    bool on[6];
    T alpha[6];
    bool any = false;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        on[c] = a.cs[c].active;
        alpha[c] = (T)a.cs[c].alpha;
        any |= on[c];
    }
    if (!any) return;
    T q[18];
#pragma unroll
    for (int i = 0; i < 18; ++i) q[i] = 0;
    for (int64_t it = 0;; ++it) {
        const int64_t tl = tile_of(it, a.N);
        if (tl < 0) break;
        const int64_t t0 = tl * BLK;
        const int32_t r = (int32_t)(t0 + threadIdx.x);
        if (r >= a.N) continue;
        T sv[6], acc[6], rr[6], vv[6];
        ld6(a.r + (int64_t)NCMP * r, rr);
        ld6(a.v + (int64_t)NCMP * r, vv);
        const T dgr = a.dg ? (T)a.dg[r] : (T)1;
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            sv[c] = rr[c] - alpha[c] * vv[c];
            acc[c] = a.dg ? dgr * sv[c] : sv[c];
        }
        const int s = r >> 6, lane = r & 63;
        const int64_t base = a.ellOff[s] + lane;
        const int width = a.ellWidth[s];
        const uint32_t live = a.slotMask ? a.slotMask[s] : 0xffffffffu;
        for (int k = 0; k < width; ++k) {
            if (k < 32 && !((live >> k) & 1u)) continue;  // zero in every row of the slice
            const int64_t e = base + (int64_t)SLICE * k;
            const int32_t col = ld_stream<T>(a.ellCol + e);
            const T av = ld_stream<T>(a.ahat + e);
            if (av == (T)0) continue;
            T rc[6], vc[6];
            ld6(a.r + (int64_t)NCMP * col, rc);
            ld6(a.v + (int64_t)NCMP * col, vc);
#pragma unroll
            for (int c = 0; c < 6; ++c) acc[c] += av * (rc[c] - alpha[c] * vc[c]);
        }
        st6(a.t + (int64_t)NCMP * r, acc);
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            if (!on[c]) continue;
            q[c] += fabs(sv[c]);
            q[6 + c] += acc[c] * sv[c];
            q[12 + c] += acc[c] * acc[c];
        }
    }
    double qd[18];
#pragma unroll
    for (int i = 0; i < 18; ++i) qd[i] = q[i];
    write_partials<18>(qd, a.partials);
}

template <class T>
__global__ void __launch_bounds__(BLK) k_fin_t(KryArgs<T> a) {
    bool on[6];
    bool any = false;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        on[c] = a.cs[c].active;
        any |= on[c];
    }
    if (!any) return;
    double sums[18];
    sum_partials<18>(a.partials, a.G, sums);
    if (threadIdx.x == 0) {
        for (int c = 0; c < 6; ++c) {
            if (!on[c]) continue;
            CompState& cs = a.cs[c];
            const double res = sums[c] / cs.normFactor;
            if (!a.ctl.mixed) cs.finalRes = res;
            if (conv_inner(cs, a.ctl, res)) {
                cs.exitAtS = 1;
            } else {
                cs.omega = sums[6 + c] / sums[12 + c];
            }
        }
    }
}

// ---- x += alpha M^-1 p + omega M^-1 s ; r = s - omega t ; |r|, (r0, r) ------
template <class T>
__global__ void __launch_bounds__(BLK) k_kry_xr(KryArgs<T> a) {
    // FP64: x += alpha M^-1 p + omega M^-1 s in place.  Mixed: the inner
    // solve accumulates the same update into its FP32 correction e (from 0 at
    // the solve's first iteration) and k_fold adds e to x in FP64 before the
    // refinement sweep, so the pass moves 4-byte instead of 8-byte x words.
    constexpr bool MIX = std::is_same<T, float>::value;
    bool on[6], ex[6], first[6];
    double alpha[6], omega[6];
    bool any = false;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        on[c] = a.cs[c].active;
        ex[c] = a.cs[c].exitAtS;
        first[c] = a.cs[c].innerIter == 0;
        alpha[c] = a.cs[c].alpha;
        omega[c] = a.cs[c].omega;
        any |= on[c];
    }
    if (!any) return;
    T q[12];
#pragma unroll
    for (int i = 0; i < 12; ++i) q[i] = 0;
    for (int64_t it = 0;; ++it) {
        const int64_t tl = tile_of(it, a.N);
        if (tl < 0) break;
        const int64_t t0 = tl * BLK;
        const int32_t r = (int32_t)(t0 + threadIdx.x);
        if (r >= a.N) continue;
        const double rd = a.rM ? a.rM[r] : 1.0;
        const int64_t j = (int64_t)NCMP * r;
        double xv[6];
        T ev[6];
        T pp[6], rr[6], vv[6], tt[6], rr0[6];
        if constexpr (MIX) {
            ld6(a.e + j, ev);
#pragma unroll
            for (int c = 0; c < 6; ++c) xv[c] = first[c] ? 0.0 : (double)ev[c];
        } else {
            ld6(a.x + j, xv);
        }
        ld6(a.p + j, pp);
        ld6(a.r + j, rr);
        ld6(a.v + j, vv);
        ld6(a.t + j, tt);
        ld6(a.r0 + j, rr0);
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            if (!on[c]) continue;
            const double y = rd * (double)pp[c];
            if (ex[c]) {
                xv[c] += alpha[c] * y;
                continue;
            }
            const T sv = rr[c] - (T)alpha[c] * vv[c];
            const double z = rd * (double)sv;
            xv[c] += alpha[c] * y + omega[c] * z;
            const T rn = sv - (T)omega[c] * tt[c];
            rr[c] = rn;
            q[c] += fabs(rn);
            q[6 + c] += rr0[c] * rn;
        }
        if constexpr (MIX) {
#pragma unroll
            for (int c = 0; c < 6; ++c) ev[c] = (T)xv[c];
            st6(a.e + j, ev);
        } else {
            st6(a.x + j, xv);
        }
        st6(a.r + j, rr);
    }
    double qd[12];
#pragma unroll
    for (int i = 0; i < 12; ++i) qd[i] = q[i];
    write_partials<12>(qd, a.partials);
}

// ---- mixed: x += e (FP64) for the components about to be refined ----------
// Gated like the refinement sweep it precedes (no inner solve running, the
// component waiting for refinement); the sweep's finalize then clears the
// wait, so a correction is added once.
template <class T>
__global__ void __launch_bounds__(BLK) k_fold(KryArgs<T> a) {
    bool on[6];
    bool any = false;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        on[c] = a.valid[c] > 0 && a.cs[c].needRefine;
        any |= on[c];
    }
    if (!any || inner_running(a)) return;
    for (int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x; i < a.N; i += (int64_t)gridDim.x * BLK) {
        double xv[6];
        T ev[6];
        ld6(a.x + NCMP * i, xv);
        ld6(a.e + NCMP * i, ev);
#pragma unroll
        for (int c = 0; c < 6; ++c)
            if (on[c]) xv[c] += (double)ev[c];
        st6(a.x + NCMP * i, xv);
    }
}

template <class T>
__global__ void __launch_bounds__(BLK) k_fin_xr(KryArgs<T> a) {
    bool on[6];
    bool any = false;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        on[c] = a.cs[c].active;
        any |= on[c];
    }
    if (!any) return;
    double sums[12];
    sum_partials<12>(a.partials, a.G, sums);
    if (threadIdx.x == 0) {
        for (int c = 0; c < 6; ++c) {
            if (!on[c]) continue;
            CompState& cs = a.cs[c];
            cs.iter += 1;
            cs.innerIter += 1;
            bool stop, breakdown = false;
            if (cs.exitAtS) {
                cs.exitAtS = 0;
                stop = true;
            } else {
                const double res = sums[c] / cs.normFactor;
                if (!a.ctl.mixed) cs.finalRes = res;
                const bool conv = conv_inner(cs, a.ctl, res);
                bool cont = (!conv && cs.iter < a.ctl.maxIter) ||
                            (!a.ctl.mixed && cs.iter < a.ctl.minIter);
                if (a.ctl.mixed && cs.innerIter >= a.ctl.maxInner) cont = false;
                stop = !cont;
                if (cont) {
                    cs.rhoOld = cs.rho;
                    cs.rho = sums[6 + c];
                    if (fabs(cs.rho) < 1e-300 || fabs(cs.omega) < 1e-300) {
                        stop = true;
                        breakdown = true;
                    } else {
                        cs.beta = (cs.rho / cs.rhoOld) * (cs.alpha / cs.omega);
                    }
                }
            }
            if (stop) {
                cs.active = 0;
                // A breakdown ((r0, r) or omega vanishing before convergence)
                // restarts from the true residual with r0 = r, through the same
                // pass as a refinement sweep (at most maxRefine times); [OF9]
                // PBiCGStab stops unconverged there instead (checkSingularity), but its
                // DILU preconditioner meets breakdowns in other places than
                // right Jacobi does, so stopping would leave results that differ
                // from it by more than the tolerance.
                if (a.ctl.mixed || breakdown) {
                    cs.needRefine = 1;
                } else {
                    cs.converged = conv_outer(cs, a.ctl, cs.finalRes);
                }
            }
        }
    }
}

// ================================================= PBiCG ([OF9] PBiCG::solve)
// The reference's tutorials solve theta with PBiCG (bi-conjugate gradients,
// e.g. tutorials/rheoFoam/Channel/Oldroyd-BLog/system/fvSolution:32): per
// iteration wA = M^-1 rA, wT = M^-T rT, wArT = (wA, rT), pA = wA + beta pA,
// pT = wT + beta pT, wA = A pA, wT = A^T pT, alpha = wArT / (wA, pT),
// psi += alpha pA, rA -= alpha wA, rT -= alpha wT.  Lockstep over the
// components like the PBiCGStab above, with M diagonal (rM).  Vectors:
// rA = r, rT = r0, pA = p, pT = v, wA = t, wT = s.

// rT = b - A^T x (FP64) and the first wArT = (M^-1 rA, rT), after the
// initial residual has written rA
template <class T>
__global__ void __launch_bounds__(BLK) k_bicg_init(KryArgs<T> a) {
    bool on[6];
    bool any = false;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        on[c] = a.cs[c].active;
        any |= on[c];
    }
    if (!any) return;
    double q[6] = {0, 0, 0, 0, 0, 0};
    for (int64_t it = 0;; ++it) {
        const int64_t tl = tile_of(it, a.N);
        if (tl < 0) break;
        const int32_t r = (int32_t)(tl * BLK + threadIdx.x);
        if (r >= a.N) continue;
        const double d = a.diag[r];
        double xv[6], y[6];
        ld6(a.x + (int64_t)NCMP * r, xv);
#pragma unroll
        for (int c = 0; c < 6; ++c) y[c] = d * xv[c];
        const int s = r >> 6, lane = r & 63;
        const int64_t base = a.ellOff[s] + lane;
        const int width = a.ellWidth[s];
        for (int k = 0; k < width; ++k) {
            const int64_t e = base + (int64_t)SLICE * k;
            const double av = a.atval[e];
            if (av == 0.0) continue;
            const int32_t col = a.ellCol[e];
            double xc[6];
            ld6(a.x + (int64_t)NCMP * col, xc);
#pragma unroll
            for (int c = 0; c < 6; ++c) y[c] += av * xc[c];
        }
        double bb[6];
        ld6(a.b + (int64_t)NCMP * r, bb);
        T rA[6], rT[6];
        ld6(a.r + (int64_t)NCMP * r, rA);
        const double rm = a.rM ? a.rM[r] : 1.0;
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            rT[c] = (T)(bb[c] - y[c]);
            if (on[c]) q[c] += (rm * (double)rA[c]) * (double)rT[c];
        }
        st6(a.r0 + (int64_t)NCMP * r, rT);
    }
    write_partials<6>(q, a.partials);
}

template <class T>
__global__ void __launch_bounds__(BLK) k_fin_bicg_init(KryArgs<T> a) {
    bool on[6];
    bool any = false;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        on[c] = a.cs[c].active;
        any |= on[c];
    }
    if (!any) return;
    double sums[6];
    sum_partials<6>(a.partials, a.G, sums);
    if (threadIdx.x == 0)
        for (int c = 0; c < 6; ++c)
            if (on[c]) a.cs[c].rho = sums[c];
}

// pA = M^-1 rA (+ beta pA), pT = M^-1 rT (+ beta pT)
template <class T>
__global__ void __launch_bounds__(BLK) k_bicg_p(KryArgs<T> a) {
    bool on[6], first[6];
    T beta[6];
    bool any = false;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        on[c] = a.cs[c].active;
        first[c] = a.cs[c].innerIter == 0;
        beta[c] = (T)a.cs[c].beta;
        any |= on[c];
    }
    if (!any) return;
    for (int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x; i < a.N; i += (int64_t)gridDim.x * BLK) {
        const T rm = a.rM ? (T)a.rM[i] : (T)1;
        T rA[6], rT[6], pA[6], pT[6];
        ld6(a.r + NCMP * i, rA);
        ld6(a.r0 + NCMP * i, rT);
        ld6(a.p + NCMP * i, pA);
        ld6(a.v + NCMP * i, pT);
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            if (!on[c]) continue;
            const T wA = rm * rA[c], wT = rm * rT[c];
            pA[c] = first[c] ? wA : wA + beta[c] * pA[c];
            pT[c] = first[c] ? wT : wT + beta[c] * pT[c];
        }
        st6(a.p + NCMP * i, pA);
        st6(a.v + NCMP * i, pT);
    }
}

// wA = A pA, wT = A^T pT; (wA, pT)
template <class T>
__global__ void __launch_bounds__(BLK) k_bicg_w(KryArgs<T> a) {
    bool on[6];
    bool any = false;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        on[c] = a.cs[c].active;
        any |= on[c];
    }
    if (!any) return;
    T q[6] = {0, 0, 0, 0, 0, 0};
    for (int64_t it = 0;; ++it) {
        const int64_t tl = tile_of(it, a.N);
        if (tl < 0) break;
        const int32_t r = (int32_t)(tl * BLK + threadIdx.x);
        if (r >= a.N) continue;
        const T d = (T)a.diag[r];
        T wA[6], wT[6], pT[6];
        ld6(a.p + (int64_t)NCMP * r, wA);
        ld6(a.v + (int64_t)NCMP * r, pT);
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            wA[c] = d * wA[c];
            wT[c] = d * pT[c];
        }
        const int s = r >> 6, lane = r & 63;
        const int64_t base = a.ellOff[s] + lane;
        const int width = a.ellWidth[s];
        const uint32_t live = a.slotMask ? a.slotMask[s] : 0xffffffffu;
        for (int k = 0; k < width; ++k) {
            const int64_t e = base + (int64_t)SLICE * k;
            const T av = (k < 32 && !((live >> k) & 1u)) ? (T)0 : ld_stream<T>(a.aA + e);
            const T at = ld_stream<T>(a.aT + e);
            if (av == (T)0 && at == (T)0) continue;
            const int32_t col = ld_stream<T>(a.ellCol + e);
            if (av != (T)0) {
                T pc[6];
                ld6(a.p + (int64_t)NCMP * col, pc);
#pragma unroll
                for (int c = 0; c < 6; ++c) wA[c] += av * pc[c];
            }
            if (at != (T)0) {
                T pc[6];
                ld6(a.v + (int64_t)NCMP * col, pc);
#pragma unroll
                for (int c = 0; c < 6; ++c) wT[c] += at * pc[c];
            }
        }
        st6(a.t + (int64_t)NCMP * r, wA);
        st6(a.s + (int64_t)NCMP * r, wT);
#pragma unroll
        for (int c = 0; c < 6; ++c)
            if (on[c]) q[c] += wA[c] * pT[c];
    }
    double qd[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) qd[c] = q[c];
    write_partials<6>(qd, a.partials);
}

template <class T>
__global__ void __launch_bounds__(BLK) k_fin_bicg_w(KryArgs<T> a) {
    bool on[6];
    bool any = false;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        on[c] = a.cs[c].active;
        any |= on[c];
    }
    if (!any) return;
    double sums[6];
    sum_partials<6>(a.partials, a.G, sums);
    if (threadIdx.x == 0)
        for (int c = 0; c < 6; ++c) {
            if (!on[c]) continue;
            CompState& cs = a.cs[c];
            // checkSingularity(mag(wApT)/normFactor): stop; restarted from the
            // true residual as after a PBiCGStab breakdown
            if (fabs(sums[c]) / cs.normFactor < 1e-300) {
                cs.active = 0;
                cs.needRefine = 1;
            } else {
                cs.alpha = cs.rho / sums[c];
            }
        }
}

// psi += alpha pA, rA -= alpha wA, rT -= alpha wT; |rA|, (M^-1 rA, rT)
template <class T>
__global__ void __launch_bounds__(BLK) k_bicg_xr(KryArgs<T> a) {
    constexpr bool MIX = std::is_same<T, float>::value;
    bool on[6], first[6];
    double alpha[6];
    bool any = false;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        on[c] = a.cs[c].active;
        first[c] = a.cs[c].innerIter == 0;
        alpha[c] = a.cs[c].alpha;
        any |= on[c];
    }
    if (!any) return;
    T q[12];
#pragma unroll
    for (int i = 0; i < 12; ++i) q[i] = 0;
    for (int64_t it = 0;; ++it) {
        const int64_t tl = tile_of(it, a.N);
        if (tl < 0) break;
        const int32_t r = (int32_t)(tl * BLK + threadIdx.x);
        if (r >= a.N) continue;
        const int64_t j = (int64_t)NCMP * r;
        const T rm = a.rM ? (T)a.rM[r] : (T)1;
        double xv[6];
        T ev[6], pA[6], wA[6], wT[6], rA[6], rT[6];
        if constexpr (MIX) {
            ld6(a.e + j, ev);
#pragma unroll
            for (int c = 0; c < 6; ++c) xv[c] = first[c] ? 0.0 : (double)ev[c];
        } else {
            ld6(a.x + j, xv);
        }
        ld6(a.p + j, pA);
        ld6(a.t + j, wA);
        ld6(a.s + j, wT);
        ld6(a.r + j, rA);
        ld6(a.r0 + j, rT);
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            if (!on[c]) continue;
            xv[c] += alpha[c] * (double)pA[c];
            rA[c] = rA[c] - (T)alpha[c] * wA[c];
            rT[c] = rT[c] - (T)alpha[c] * wT[c];
            q[c] += fabs(rA[c]);
            q[6 + c] += (rm * rA[c]) * rT[c];
        }
        if constexpr (MIX) {
#pragma unroll
            for (int c = 0; c < 6; ++c) ev[c] = (T)xv[c];
            st6(a.e + j, ev);
        } else {
            st6(a.x + j, xv);
        }
        st6(a.r + j, rA);
        st6(a.r0 + j, rT);
    }
    double qd[12];
#pragma unroll
    for (int i = 0; i < 12; ++i) qd[i] = q[i];
    write_partials<12>(qd, a.partials);
}

template <class T>
__global__ void __launch_bounds__(BLK) k_fin_bicg_xr(KryArgs<T> a) {
    bool on[6];
    bool any = false;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        on[c] = a.cs[c].active;
        any |= on[c];
    }
    if (!any) return;
    double sums[12];
    sum_partials<12>(a.partials, a.G, sums);
    if (threadIdx.x == 0) {
        for (int c = 0; c < 6; ++c) {
            if (!on[c]) continue;
            CompState& cs = a.cs[c];
            cs.iter += 1;
            cs.innerIter += 1;
            const double res = sums[c] / cs.normFactor;
            if (!a.ctl.mixed) cs.finalRes = res;
            const bool conv = conv_inner(cs, a.ctl, res);
            bool cont = (!conv && cs.iter < a.ctl.maxIter) || (!a.ctl.mixed && cs.iter < a.ctl.minIter);
            if (a.ctl.mixed && cs.innerIter >= a.ctl.maxInner) cont = false;
            bool breakdown = false;
            if (cont) {
                cs.rhoOld = cs.rho;
                cs.rho = sums[6 + c];
                if (fabs(cs.rho) < 1e-300) {
                    cont = false;
                    breakdown = true;
                } else {
                    cs.beta = cs.rho / cs.rhoOld;
                }
            }
            if (!cont) {
                cs.active = 0;
                if (a.ctl.mixed || breakdown) cs.needRefine = 1;
                else cs.converged = conv_outer(cs, a.ctl, cs.finalRes);
            }
        }
    }
}

// float copies of the off-diagonals of A and A^T (mixed PBiCG)
__global__ void k_to_f32(int64_t n, const double* __restrict__ a, const double* __restrict__ b,
                         float* __restrict__ fa, float* __restrict__ fb) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    fa[i] = (float)a[i];
    fb[i] = (float)b[i];
}

// ---- Ahat = offdiag(A) D^-1 (column scaled), rD = 1/diag ---------------------
__global__ void k_rdiag(int32_t N, const double* __restrict__ diag, double* __restrict__ rD) {
    const int64_t r = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (r < N) rD[r] = 1.0 / diag[r];
}
template <class T>
__global__ void k_scale(int32_t N, const int64_t* __restrict__ off, const int32_t* __restrict__ width,
                        const uint32_t* __restrict__ slotMask, const int32_t* __restrict__ col,
                        const double* __restrict__ aval, const double* __restrict__ rD, T* __restrict__ ahat) {
    const int64_t r = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (r >= N) return;
    const int s = (int)(r >> 6), lane = (int)(r & 63);
    const int64_t base = off[s] + lane;
    const int w = width[s];
    const uint32_t live = slotMask ? slotMask[s] : 0xffffffffu;
    for (int k = 0; k < w; ++k) {
        if (k < 32 && !((live >> k) & 1u)) continue;
        const int64_t e = base + (int64_t)SLICE * k;
        ahat[e] = rD ? (T)(aval[e] * rD[col[e]]) : (T)aval[e];
    }
}

// ---- mean of x per component (normalisation factor's gAverage(psi)) ---------
__global__ void __launch_bounds__(BLK) k_mean6(int32_t N, const double* __restrict__ x, double* partials) {
    double q[6] = {0, 0, 0, 0, 0, 0};
    for (int64_t it = 0;; ++it) {
        const int64_t tl = tile_of(it, N);
        if (tl < 0) break;
        const int64_t t0 = tl * BLK;
        const int64_t r = t0 + threadIdx.x;
        if (r >= N) continue;
        double v[6];
        ld6(x + NCMP * r, v);
#pragma unroll
        for (int c = 0; c < 6; ++c) q[c] += v[c];
    }
    write_partials<6>(q, partials);
}

__global__ void __launch_bounds__(BLK) k_fin_mean6(int32_t N, int G, const double* partials, CompState* cs) {
    double sums[6];
    sum_partials<6>(partials, G, sums);
    if (threadIdx.x == 0)
        for (int c = 0; c < 6; ++c) cs[c].xbar = sums[c] / N;
}

// ----------------------------------------------------------------- driver
static inline unsigned kblk(int64_t n) { return (unsigned)((n + BLK - 1) / BLK); }

void launch_mean6(Ctx* ctx, int32_t N, const double* x, double* partials, CompState* cs,
                  int grid) {
    {
        ProfScope ps(ctx, "k_mean6", (double)N * 48);
        hipLaunchKernelGGL(k_mean6, dim3(grid), dim3(BLK), 0, ctx->stream, N, x, partials);
    }
    hipLaunchKernelGGL(k_fin_mean6, dim3(1), dim3(BLK), 0, ctx->stream, N, grid, partials, cs);
    RT_HIP(hipGetLastError());
}

KryCtl kryCtl(const Controls& C, bool mixed) {
    KryCtl k{};
    k.tol = C.tolerance;
    k.relTol = C.relTol;
    k.innerRelTol = C.innerRelTol;
    k.maxIter = C.maxIter;
    k.minIter = C.minIter;
    k.maxInner = 200;
    k.maxRefine = C.maxRefine;
    k.mixed = mixed ? 1 : 0;
    k.bicg = C.solver == SV_PBICG ? 1 : 0;
    k.precond = C.precond;
    return k;
}

void Krylov::ensure(Mesh& m, bool mixed, bool bicg) {
    if (!ctx) ctx = m.ctx;
    const size_t vec = (size_t)6 * m.N;
    if (noAlloc) {
        const bool ok = N == m.N && partials.p && state.p && rD.p && (mixed ? rs.p && r0s.p && ps.p && vs.p &&
                                                                          ts.p && es.p && ahat32.p
                                                                    : r.p && r0.p && p.p && v.p && t.p && ahat64.p) &&
                        (!bicg || (atval.p && (mixed ? s32.p && a32.p && at32.p : s64.p != nullptr)));
        if (!ok) throw Error{RT_E_BAD_ARG, "Krylov workspace bound to scratch memory does not fit the solve"};
        return;
    }
    if (N != m.N) {
        N = m.N;
        grid = grid_for(ctx, N);
        r = DBuf(); r0 = DBuf(); p = DBuf(); v = DBuf(); t = DBuf();
        rs = DBuf(); r0s = DBuf(); ps = DBuf(); vs = DBuf(); ts = DBuf(); es = DBuf();
        ahat32 = DBuf(); ahat64 = DBuf();
        partials = ctx->alloc((size_t)grid * 24 * sizeof(double));
        state = ctx->alloc(sizeof(SolveState));
        RT_HIP(hipMemsetAsync(state.p, 0, sizeof(SolveState), ctx->stream));
        rD = ctx->alloc(std::max<size_t>(1, (size_t)m.N) * sizeof(double));
    }
    if (!mixed && !r.p) {
        r = ctx->alloc(vec * 8); r0 = ctx->alloc(vec * 8); p = ctx->alloc(vec * 8);
        v = ctx->alloc(vec * 8); t = ctx->alloc(vec * 8);
        ahat64 = ctx->alloc(std::max<int64_t>(1, m.nEll) * 8);
    }
    if (mixed && !rs.p) {
        rs = ctx->alloc(vec * 4); r0s = ctx->alloc(vec * 4); ps = ctx->alloc(vec * 4);
        vs = ctx->alloc(vec * 4); ts = ctx->alloc(vec * 4); es = ctx->alloc(vec * 4);
        ahat32 = ctx->alloc(std::max<int64_t>(1, m.nEll) * 4);
    }
    if (bicg) {
        if (!atval.p) atval = ctx->alloc(std::max<int64_t>(1, m.nEll) * 8);
        if (!mixed && !s64.p) s64 = ctx->alloc(vec * 8);
        if (mixed && !s32.p) {
            s32 = ctx->alloc(vec * 4);
            a32 = ctx->alloc(std::max<int64_t>(1, m.nEll) * 4);
            at32 = ctx->alloc(std::max<int64_t>(1, m.nEll) * 4);
        }
    }
}

// Resumable solve: begin() launches the initial residual (unless the assembly
// already produced it), every advance() reads the device state of the
// previous round (non-blocking when asked) and launches the next round, so
// several solves can progress on their own streams at once (the serving loop)
// or one can run to completion (rt_model_correct / rt_symm_solve).
struct SolveRun {
    bool mixed = false;
    KryArgs<float> af{};
    KryArgs<double> ad{};
    int G = 1, pgrid = 1;
    double N = 0, ne = 0, nValid = 0;
    int32_t valid[6];
    Ctx* ctx = nullptr;
    SolveState* hostState = nullptr;  // pinned block of ctx
    CompState* host = nullptr;        // = hostState->c
    bool neFromDevice = false;        // matrix entry count comes with the first status pull
    bool needScale = false;           // Ahat (or the float copies of A, A^T) not built yet
    bool bicg = false;                // PBiCG instead of PBiCGStab
    int64_t nEll = 0;                 // all ELL entries (the float copies of A, A^T)
    hipEvent_t ev = nullptr;
    bool ownsHost = false;            // hostState is a pinned block of ctx
    bool pending = false, done = false;
    int rounds = 0;
    ~SolveRun() {
        if (hostState && ownsHost) ctx->givePinnedBlock(hostState);
        if (ev) (void)hipEventDestroy(ev);
    }
};

void SolveRunDeleter::operator()(SolveRun* r) const { delete r; }

namespace {

template <class T>
KryArgs<T>& argsOf(SolveRun& R);
template <>
KryArgs<float>& argsOf<float>(SolveRun& R) { return R.af; }
template <>
KryArgs<double>& argsOf<double>(SolveRun& R) { return R.ad; }

template <class T>
void fill_args(Ctx* ctx, const SolveArgs& sa, SolveRun& R) {
    Mesh& m = *sa.mesh;
    Krylov& K = *sa.kry;
    const bool mixed = std::is_same<T, float>::value;
    const int32_t N = m.N;
    SolveState* st = K.state.as<SolveState>();
    double* rD = K.rD.as<double>();
    if (!sa.initDone) {
        ProfScope ps(ctx, "k_rdiag", (double)N * 16);
        hipLaunchKernelGGL(k_rdiag, dim3(kblk(N)), dim3(BLK), 0, ctx->stream, N, sa.diag, rD);
    }
    T* ahat = mixed ? (T*)K.ahat32.p : (T*)K.ahat64.p;
    R.needScale = m.nEll > 0;  // launched with the first round, once the entry count is known
    RT_HIP(hipGetLastError());
    KryArgs<T>& a = argsOf<T>(R);
    a = KryArgs<T>{};
    a.N = N;
    a.ellOff = m.ellOff.as<int64_t>();
    a.ellWidth = m.ellWidth.as<int32_t>();
    a.ellCol = m.ellCol.as<int32_t>();
    a.aval = sa.aval;
    a.slotMask = sa.slotMask;
    a.diag = sa.diag;
    a.ahat = ahat;
    a.rD = rD;
    a.b = sa.b;
    a.x = sa.x;
    if (mixed) {
        a.r = (T*)K.rs.p; a.r0 = (T*)K.r0s.p; a.p = (T*)K.ps.p; a.v = (T*)K.vs.p; a.t = (T*)K.ts.p;
        a.e = (T*)K.es.p;
    } else {
        a.r = (T*)K.r.p; a.r0 = (T*)K.r0.p; a.p = (T*)K.p.p; a.v = (T*)K.v.p; a.t = (T*)K.t.p;
        a.e = nullptr;
    }
    a.partials = K.partials.as<double>();
    a.cs = st->c;
    for (int c = 0; c < 6; ++c) a.valid[c] = sa.valid[c];
    a.ctl = kryCtl(*sa.ctl, mixed);
    a.G = K.grid;
    // the preconditioner: Jacobi M = D (M^-1 = rD, A M^-1 has a unit diagonal)
    // or none (M^-1 = 1, diag(A M^-1) = D)
    const bool jac = sa.ctl->precond == PC_JACOBI;
    a.rM = jac ? rD : nullptr;
    a.dg = jac ? nullptr : sa.diag;
    R.bicg = sa.ctl->solver == SV_PBICG;
    if (R.bicg) {
        a.atval = sa.atval;
        a.s = mixed ? (T*)K.s32.p : (T*)K.s64.p;
        a.aA = mixed ? (const T*)K.a32.p : (const T*)sa.aval;
        a.aT = mixed ? (const T*)K.at32.p : (const T*)sa.atval;
    }
}

// algorithmic bytes per launch for nc participating components: ELL entries
// (col + value) once per SpMV, each vector once per pass (interleaved rows
// move all 6 components; gathered vectors counted once)
struct Bytes {
    double N, ne, w;
    double res(double nc) const { return N * 8.0 + ne * 12 + nc * N * (8 + 8 + 2 * w); }
    double p(double nc, bool start) const { return nc * N * (start ? 2 : 4) * w; }
    double v(double nc) const { return ne * (4 + w) + nc * N * 3 * w; }
    double t(double nc) const { return ne * (4 + w) + nc * N * 3 * w; }
    // FP64: x read and written in place; mixed: the FP32 correction instead
    double xr(double nc) const { return N * 8.0 + nc * N * (w == 4 ? 8 * w : 16 + 6 * w); }
    double fold(double nc) const { return nc * N * (8 + 8 + w); }
    // PBiCG: rT = b - A^T x; the p pass; A pA and A^T pT; the x/r pass
    double binit(double nc) const { return N * 16.0 + ne * 12 + nc * N * (8 + 8 + 2 * w); }
    double bp(double nc) const { return N * 8.0 + nc * N * 6 * w; }
    double bw(double nc) const { return N * 8.0 + 2 * ne * (4 + w) + nc * N * 4 * w; }
    double bxr(double nc) const { return N * 8.0 + nc * N * (w == 4 ? 9 * w : 16 + 7 * w); }
};

template <class T>
void launch_init(Ctx* ctx, SolveRun& R) {
    const bool mixed = std::is_same<T, float>::value;
    Bytes B{R.N, R.ne, (double)sizeof(T)};
    {
        ProfScope ps(ctx, mixed ? "k_residual<f32>" : "k_residual<f64>", B.res(R.nValid));
        hipLaunchKernelGGL((k_residual<T, 0>), dim3(R.G), dim3(BLK), 0, ctx->stream, argsOf<T>(R));
    }
    hipLaunchKernelGGL((k_fin_residual<T, 0>), dim3(1), dim3(BLK), 0, ctx->stream, argsOf<T>(R));
}

// PBiCG: rT = b - A^T x and the first wArT, once the initial residual is in
template <class T>
void launch_bicg_init(Ctx* ctx, SolveRun& R) {
    const bool mixed = std::is_same<T, float>::value;
    Bytes B{R.N, R.ne, (double)sizeof(T)};
    {
        ProfScope ps(ctx, mixed ? "k_bicg_init<f32>" : "k_bicg_init<f64>", B.binit(R.nValid));
        hipLaunchKernelGGL(k_bicg_init<T>, dim3(R.G), dim3(BLK), 0, ctx->stream, argsOf<T>(R));
    }
    hipLaunchKernelGGL(k_fin_bicg_init<T>, dim3(1), dim3(BLK), 0, ctx->stream, argsOf<T>(R));
}

// The refinement sweep: mixed-precision fold, then the FP64 true residual and
// its finalize (which converges, restarts or stops each waiting component).
template <class T>
void launch_refine(Ctx* ctx, SolveRun& R, double nRef) {
    const bool mixed = std::is_same<T, float>::value;
    Bytes B{R.N, R.ne, (double)sizeof(T)};
    KryArgs<T>& a = argsOf<T>(R);
    if (mixed) {
        ProfScope ps(ctx, "k_fold<f32>", B.fold(nRef));
        hipLaunchKernelGGL(k_fold<T>, dim3(R.pgrid), dim3(BLK), 0, ctx->stream, a);
    }
    {
        ProfScope ps(ctx, mixed ? "k_residual<f32>" : "k_residual<f64>", B.res(nRef));
        hipLaunchKernelGGL((k_residual<T, 1>), dim3(R.G), dim3(BLK), 0, ctx->stream, a);
    }
    hipLaunchKernelGGL((k_fin_residual<T, 1>), dim3(1), dim3(BLK), 0, ctx->stream, a);
}

// One lockstep PBiCGStab iteration of the nAct active components.
template <class T>
void launch_iter(Ctx* ctx, SolveRun& R, double nAct, bool start) {
    const bool mixed = std::is_same<T, float>::value;
    Bytes B{R.N, R.ne, (double)sizeof(T)};
    KryArgs<T>& a = argsOf<T>(R);
    if (R.bicg) {
        if (R.needScale && mixed) {
            const int64_t n = std::max<int64_t>(1, R.nEll);
            ProfScope ps(ctx, "k_to_f32", (double)n * 24);
            hipLaunchKernelGGL(k_to_f32, dim3(kblk(n)), dim3(BLK), 0, ctx->stream, n, a.aval, a.atval,
                               (float*)const_cast<T*>(a.aA), (float*)const_cast<T*>(a.aT));
        }
        R.needScale = false;
        {
            ProfScope ps(ctx, mixed ? "k_bicg_p<f32>" : "k_bicg_p<f64>", B.bp(nAct));
            hipLaunchKernelGGL(k_bicg_p<T>, dim3(R.pgrid), dim3(BLK), 0, ctx->stream, a);
        }
        {
            ProfScope ps(ctx, mixed ? "k_bicg_w<f32>" : "k_bicg_w<f64>", B.bw(nAct));
            hipLaunchKernelGGL(k_bicg_w<T>, dim3(R.G), dim3(BLK), 0, ctx->stream, a);
        }
        hipLaunchKernelGGL(k_fin_bicg_w<T>, dim3(1), dim3(BLK), 0, ctx->stream, a);
        {
            ProfScope ps(ctx, mixed ? "k_bicg_xr<f32>" : "k_bicg_xr<f64>", B.bxr(nAct));
            hipLaunchKernelGGL(k_bicg_xr<T>, dim3(R.G), dim3(BLK), 0, ctx->stream, a);
        }
        hipLaunchKernelGGL(k_fin_bicg_xr<T>, dim3(1), dim3(BLK), 0, ctx->stream, a);
        return;
    }
    if (R.needScale) {
        ProfScope ps(ctx, mixed ? "k_scale<f32>" : "k_scale<f64>", R.ne * (4 + 8 + sizeof(T)) + R.N * 8);
        hipLaunchKernelGGL(k_scale<T>, dim3(kblk((int64_t)R.N)), dim3(BLK), 0, ctx->stream, a.N, a.ellOff,
                           a.ellWidth, a.slotMask, a.ellCol, a.aval, a.rM, const_cast<T*>(a.ahat));
        R.needScale = false;
    }
    {
        ProfScope ps(ctx, mixed ? "k_kry_p<f32>" : "k_kry_p<f64>", B.p(nAct, start));
        hipLaunchKernelGGL(k_kry_p<T>, dim3(R.pgrid), dim3(BLK), 0, ctx->stream, a);
    }
    {
        ProfScope ps(ctx, mixed ? "k_kry_v<f32>" : "k_kry_v<f64>", B.v(nAct));
        hipLaunchKernelGGL(k_kry_v<T>, dim3(R.G), dim3(BLK), 0, ctx->stream, a);
    }
    hipLaunchKernelGGL(k_fin_v<T>, dim3(1), dim3(BLK), 0, ctx->stream, a);
    {
        ProfScope ps(ctx, mixed ? "k_kry_t<f32>" : "k_kry_t<f64>", B.t(nAct));
        hipLaunchKernelGGL(k_kry_t<T>, dim3(R.G), dim3(BLK), 0, ctx->stream, a);
    }
    hipLaunchKernelGGL(k_fin_t<T>, dim3(1), dim3(BLK), 0, ctx->stream, a);
    {
        ProfScope ps(ctx, mixed ? "k_kry_xr<f32>" : "k_kry_xr<f64>", B.xr(nAct));
        hipLaunchKernelGGL(k_kry_xr<T>, dim3(R.G), dim3(BLK), 0, ctx->stream, a);
    }
    hipLaunchKernelGGL(k_fin_xr<T>, dim3(1), dim3(BLK), 0, ctx->stream, a);
}

// rounds a non-blocking advance queues per host round trip
constexpr int kRoundsAhead = 3;

void queue_pull(Ctx* ctx, Krylov& K, SolveRun& R) {
    SolveState* st = K.state.as<SolveState>();
    RT_HIP(hipMemcpyAsync(R.hostState, st, sizeof(SolveState), hipMemcpyDeviceToHost, ctx->stream));
    if (R.ev) RT_HIP(hipEventRecord(R.ev, ctx->stream));
    R.pending = true;
}

}  // namespace

void solve_begin(Ctx* ctx, const SolveArgs& sa) {
    Mesh& m = *sa.mesh;
    Krylov& K = *sa.kry;
    const bool mixed = sa.ctl->mixed;
    const bool bicg = sa.ctl->solver == SV_PBICG;
    K.ensure(m, mixed, bicg);
    if (bicg && !sa.atval) throw Error{RT_E_BAD_ARG, "PBiCG needs the transposed matrix"};
    if (K.run && sa.hostStatus && K.run->hostState != sa.hostStatus) K.run.reset();
    if (!K.run) {
        static_assert(sizeof(SolveState) <= Ctx::kPinnedBlock, "status block too small");
        K.run.reset(new SolveRun);
        K.run->ctx = ctx;
        if (sa.hostStatus) {
            // the caller's pageable block, no event: a blocking solve waits on
            // the stream (the no-allocation path)
            K.run->hostState = sa.hostStatus;
        } else {
            K.run->hostState = static_cast<SolveState*>(ctx->takePinnedBlock());
            K.run->ownsHost = true;
            RT_HIP(hipEventCreateWithFlags(&K.run->ev, hipEventDisableTiming));
        }
        K.run->host = K.run->hostState->c;
    }
    SolveRun& R = *K.run;
    R.mixed = mixed;
    R.G = K.grid;
    R.pgrid = std::min<int>(K.grid, (int)kblk(m.N));
    R.N = m.N;
    R.ne = (double)m.nEll;
    R.nEll = m.nEll;
    R.neFromDevice = sa.slotMask != nullptr;
    R.nValid = 0;
    for (int c = 0; c < 6; ++c) {
        R.valid[c] = sa.valid[c];
        R.nValid += sa.valid[c] > 0;
    }
    R.done = false;
    R.rounds = 0;
    if (mixed) {
        fill_args<float>(ctx, sa, R);
        if (!sa.initDone) launch_init<float>(ctx, R);
        if (bicg) launch_bicg_init<float>(ctx, R);
    } else {
        fill_args<double>(ctx, sa, R);
        if (!sa.initDone) launch_init<double>(ctx, R);
        if (bicg) launch_bicg_init<double>(ctx, R);
    }
    RT_HIP(hipGetLastError());
    queue_pull(ctx, K, R);
}

bool solve_advance(Ctx* ctx, Krylov& K, bool block) {
    SolveRun& R = *K.run;
    if (R.done) return true;
    if (R.pending) {
        if (block) {
            if (R.ev) RT_HIP(hipEventSynchronize(R.ev));
            else RT_HIP(hipStreamSynchronize(ctx->stream));
        } else if (!R.ev) {
            throw Error{RT_E_BAD_ARG, "a solve without a status event can only be advanced blocking"};
        } else {
            hipError_t e = hipEventQuery(R.ev);
            if (e == hipErrorNotReady) return false;
            RT_HIP(e);
        }
        R.pending = false;
        if (R.neFromDevice) {
            R.ne = R.hostState->misc[0];
            R.neFromDevice = false;
        }
    }
    double nAct = 0, nRef = 0;
    bool start = true;  // every component of the round begins an inner solve (p = r)
    for (int c = 0; c < 6; ++c) {
        if (R.valid[c] <= 0) continue;
        nAct += R.host[c].active != 0;
        nRef += R.host[c].needRefine != 0;
        if (R.host[c].active && R.host[c].innerIter != 0) start = false;
    }
    if ((nAct == 0 && nRef == 0) || R.rounds >= 100000) {
        R.done = true;
        return true;
    }
    // A refinement sweep costs a full FP64 pass whatever the number of
    // components in it: components whose inner solve stopped wait (idle, no
    // state change) until every inner solve has stopped, then refine together.
    if (nAct > 0) nRef = 0;
    if (block) {
        // Blocking callers (rt_model_correct, rt_symm_solve, every profiled
        // run) launch exactly what the host's view of the state asks for: a
        // round is either one refinement sweep or one Krylov iteration of the
        // components active now.  A sweep may converge every component, so the
        // iteration after it is only launched once its state is back; no
        // kernel is launched (or charged bytes) for a finished solve.
        if (nAct > 0) {
            if (R.mixed) launch_iter<float>(ctx, R, nAct, start);
            else launch_iter<double>(ctx, R, nAct, start);
        } else {
            if (R.mixed) launch_refine<float>(ctx, R, nRef);
            else launch_refine<double>(ctx, R, nRef);
        }
        ++R.rounds;
    } else {
        // Non-blocking callers (the serving loop) queue kRoundsAhead rounds of
        // sweep + iteration before the state comes back: every kernel reads
        // the component records on the device and does nothing for finished
        // components, and the sweep is gated there on no inner solve running,
        // so a round launched ahead does what the host would have launched.
        // This hides the host round trip, which dominates small solves.
        for (int k = 0; k < kRoundsAhead; ++k) {
            const double nR = k == 0 ? nRef : 1, nA = k == 0 ? nAct + nRef : nAct;
            if (R.mixed) {
                if (nR > 0) launch_refine<float>(ctx, R, nR);
                launch_iter<float>(ctx, R, nA, k == 0 && start && nRef == 0);
            } else {
                if (nR > 0) launch_refine<double>(ctx, R, nR);
                launch_iter<double>(ctx, R, nA, k == 0 && start && nRef == 0);
            }
            ++R.rounds;
        }
    }
    RT_HIP(hipGetLastError());
    queue_pull(ctx, K, R);
    return false;
}

void solve_finish(Krylov& K, rt_solver_perf perf[6]) {
    SolveRun& R = *K.run;
    for (int c = 0; c < 6; ++c) {
        rt_solver_perf& pf = perf[c];
        pf = rt_solver_perf{0, 0, 0, 0, 0, 0};
        if (R.valid[c] <= 0) continue;
        pf.initialResidual = R.host[c].initRes;
        pf.finalResidual = R.host[c].finalRes;
        pf.nIterations = R.host[c].iter;
        pf.nRefinements = R.host[c].nRefine;
        pf.converged = R.host[c].converged;
        pf.solved = 1;
    }
}

void solve_lockstep(Ctx* ctx, const SolveArgs& a, rt_solver_perf perf[6]) {
    solve_begin(ctx, a);
    while (!solve_advance(ctx, *a.kry, true)) {
    }
    solve_finish(*a.kry, perf);
}

int grid_for(const Ctx* ctx, int32_t N) {
    const int64_t need = ((int64_t)N + BLK - 1) / BLK;
    const int64_t cap = (int64_t)ctx->numCUs * 8;
    int64_t g = need < cap ? need : cap;
    if (g >= 64) g = g / 8 * 8;
    return (int)std::max<int64_t>(1, g);
}

}  // namespace rt
