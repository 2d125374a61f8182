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
    int32_t nRed;        // colour-blocked rows: the first colour's count (0: not coloured)
    int32_t rowBegin;    // k_residual: rows [rowBegin, N) (the red-black sweep's second colour)
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
__device__ __forceinline__ void k_residual_body(const KryArgs<T>& a, int gG, int gB) {
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
    // rows [rowBegin, N): all rows, or the second colour of a red-black sweep
    // (the first colour's residual is zero once it has been eliminated)
    const int64_t nRows = (int64_t)a.N - a.rowBegin;
    for (int64_t it = 0;; ++it) {
        const int64_t tl = a.rowBegin > 0 ? tile_of_g(it, nRows, 0, gG, gB) : tile_of_g(it, a.N, a.nRed, gG, gB);
        if (tl < 0) {
            if (tl == -1) break;
            continue;
        }
        const int64_t t0 = tl * BLK;
        const int32_t r = (int32_t)(a.rowBegin + t0 + threadIdx.x);
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
    write_partials<NQ>(q, a.partials, gB);
}

template <class T, int MODE>
__global__ void __launch_bounds__(BLK) k_residual(KryArgs<T> a) {
    k_residual_body<T, MODE>(a, gridDim.x, blockIdx.x);
}

template <class T, int MODE>
__device__ __forceinline__ void k_fin_residual_body(const KryArgs<T>& a, int gG, int gB) {
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
                cs.foldE = 0;
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

template <class T, int MODE>
__global__ void __launch_bounds__(BLK) k_fin_residual(KryArgs<T> a) {
    k_fin_residual_body<T, MODE>(a, gridDim.x, blockIdx.x);
}

// ---- p = r  |  p = r + beta (p - omega v) ---------------------------------
template <class T>
__device__ __forceinline__ void k_kry_p_body(const KryArgs<T>& a, int gG, int gB) {
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
        for (int64_t i = (int64_t)gB * BLK + threadIdx.x; i < N; i += (int64_t)gG * BLK) {
            T rr[6];
            ld6(a.r + NCMP * i, rr);
            st6(a.p + NCMP * i, rr);
        }
        return;
    }
    for (int64_t i = (int64_t)gB * BLK + threadIdx.x; i < N; i += (int64_t)gG * BLK) {
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

template <class T>
__global__ void __launch_bounds__(BLK) k_kry_p(KryArgs<T> a) {
    k_kry_p_body<T>(a, gridDim.x, blockIdx.x);
}

// ---- v = A M^-1 p = p + Ahat p ;  sigma = (r0, v) -------------------------
template <class T>
__device__ __forceinline__ void k_kry_v_body(const KryArgs<T>& a, int gG, int gB) {
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
        const int64_t tl = tile_of_g(it, a.N, a.nRed, gG, gB);
        if (tl < 0) {
            if (tl == -1) break;
            continue;
        }
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
    write_partials<6>(qd, a.partials, gB);
}

template <class T>
__global__ void __launch_bounds__(BLK) k_kry_v(KryArgs<T> a) {
    k_kry_v_body<T>(a, gridDim.x, blockIdx.x);
}

template <class T>
__device__ __forceinline__ void k_fin_v_body(const KryArgs<T>& a, int gG, int gB) {
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

template <class T>
__global__ void __launch_bounds__(BLK) k_fin_v(KryArgs<T> a) {
    k_fin_v_body<T>(a, gridDim.x, blockIdx.x);
}

// ---- s = r - alpha v ; t = A M^-1 s ; |s|, (t,s), (t,t) --------------------
template <class T>
__device__ __forceinline__ void k_kry_t_body(const KryArgs<T>& a, int gG, int gB) {
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
        const int64_t tl = tile_of_g(it, a.N, a.nRed, gG, gB);
        if (tl < 0) {
            if (tl == -1) break;
            continue;
        }
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
    write_partials<18>(qd, a.partials, gB);
}

template <class T>
__global__ void __launch_bounds__(BLK) k_kry_t(KryArgs<T> a) {
    k_kry_t_body<T>(a, gridDim.x, blockIdx.x);
}

template <class T>
__device__ __forceinline__ void k_fin_t_body(const KryArgs<T>& a, int gG, int gB) {
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

template <class T>
__global__ void __launch_bounds__(BLK) k_fin_t(KryArgs<T> a) {
    k_fin_t_body<T>(a, gridDim.x, blockIdx.x);
}

// ---- x += alpha M^-1 p + omega M^-1 s ; r = s - omega t ; |r|, (r0, r) ------
template <class T>
__device__ __forceinline__ void k_kry_xr_body(const KryArgs<T>& a, int gG, int gB) {
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
        const int64_t tl = tile_of_g(it, a.N, a.nRed, gG, gB);
        if (tl < 0) {
            if (tl == -1) break;
            continue;
        }
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
    write_partials<12>(qd, a.partials, gB);
}

template <class T>
__global__ void __launch_bounds__(BLK) k_kry_xr(KryArgs<T> a) {
    k_kry_xr_body<T>(a, gridDim.x, blockIdx.x);
}

// ---- mixed: x += e (FP64) for the components about to be refined ----------
// Gated like the refinement sweep it precedes (no inner solve running, the
// component waiting for refinement); the sweep's finalize then clears the
// wait, so a correction is added once.
template <class T>
__device__ __forceinline__ void k_fold_body(const KryArgs<T>& a, int gG, int gB) {
    bool on[6];
    bool any = false;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        on[c] = a.valid[c] > 0 && a.cs[c].needRefine && a.cs[c].foldE;
        any |= on[c];
    }
    if (!any || inner_running(a)) return;
    for (int64_t i = (int64_t)gB * BLK + threadIdx.x; i < a.N; i += (int64_t)gG * BLK) {
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
__global__ void __launch_bounds__(BLK) k_fold(KryArgs<T> a) {
    k_fold_body<T>(a, gridDim.x, blockIdx.x);
}

template <class T>
__device__ __forceinline__ void k_fin_xr_body(const KryArgs<T>& a, int gG, int gB) {
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
                if (a.ctl.mixed || breakdown || a.ctl.rb) {
                    // mixed: fold the inner solve's correction, then refine;
                    // red-black: the closing sweep eliminates the first colour
                    cs.needRefine = 1;
                    cs.foldE = a.ctl.mixed && cs.innerIter > 0 ? 1 : 0;
                } else {
                    cs.converged = conv_outer(cs, a.ctl, cs.finalRes);
                }
            }
        }
    }
}

template <class T>
__global__ void __launch_bounds__(BLK) k_fin_xr(KryArgs<T> a) {
    k_fin_xr_body<T>(a, gridDim.x, blockIdx.x);
}

// ================================== red-black ILU0 (multicolour ILU0, two colours)
// On a two-colourable cell graph stored colour by colour (rows [0, nRed) =: R,
// [nRed, N) =: B; A_RR = D_R and A_BB = D_B are diagonal), ILU(0) in this
// ordering factors A exactly except for the pivots of B, which become
// Ds = diag(S) of the Schur complement S = D_B - A_BR D_R^-1 A_RB (D_B itself
// when every product A_ij A_ji vanishes, as for the upwind theta matrix).
// ILU(0)-preconditioned Krylov is the same method as Krylov on S with the
// Jacobi preconditioner Ds and R eliminated exactly,
// x_R = D_R^-1 (b_R - A_RB x_B): the residual of the whole system is then zero
// on R and S's residual on B.  So the lockstep PBiCGStab runs on B's rows with
// right preconditioning M = Ds, each product S M^-1 y in two passes:
//   R rows gather  w   = D_R^-1 A_RB Ds^-1 y   (ahat = A_jk / (D_j Ds_k) there)
//   B rows gather  S M^-1 y = (D_B Ds^-1) y - A_BR w   (ahat = A_ij there).
// r, r0, p, v, t live on B's rows; w uses R's rows of v, w' R's rows of t.
// Every solve opens with a sweep (x_R from x_B, then the true residual on B)
// and closes with one, so x_R is always the exact elimination of x_B.

// R rows: w = D_R^-1 A_RB Ds^-1 y, y = p (SRC 0) or s = r - alpha v (SRC 1)
template <class T, int SRC>
__device__ __forceinline__ void k_rb_w_body(const KryArgs<T>& a, int gG, int gB) {
    T* __restrict__ w = SRC == 0 ? a.v : a.t;
    T alpha[6];
    bool any = false;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        alpha[c] = (T)a.cs[c].alpha;
        any |= a.cs[c].active != 0;
    }
    if (!any) return;
    const int64_t n = a.nRed;
    for (int64_t it = 0;; ++it) {
        const int64_t tl = tile_of_g(it, n, 0, gG, gB);
        if (tl < 0) {
            if (tl == -1) break;
            continue;
        }
        const int32_t r = (int32_t)(tl * BLK + threadIdx.x);
        if (r >= n) continue;
        T acc[6] = {0, 0, 0, 0, 0, 0};
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
            if constexpr (SRC == 0) {
                T pc[6];
                ld6(a.p + (int64_t)NCMP * col, pc);
#pragma unroll
                for (int c = 0; c < 6; ++c) acc[c] += av * pc[c];
            } else {
                T rc[6], vc[6];
                ld6(a.r + (int64_t)NCMP * col, rc);
                ld6(a.v + (int64_t)NCMP * col, vc);
#pragma unroll
                for (int c = 0; c < 6; ++c) acc[c] += av * (rc[c] - alpha[c] * vc[c]);
            }
        }
        st6(w + (int64_t)NCMP * r, acc);
    }
}

template <class T, int SRC>
__global__ void __launch_bounds__(BLK) k_rb_w(KryArgs<T> a) {
    k_rb_w_body<T, SRC>(a, gridDim.x, blockIdx.x);
}

// B rows: v = S M^-1 p = (D_B Ds^-1) p - A_BR w ;  sigma = (r0, v)
template <class T>
__device__ __forceinline__ void k_rb_v_body(const KryArgs<T>& a, int gG, int gB) {
    const T* __restrict__ w = a.v;
    bool on[6];
    bool any = false;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        on[c] = a.cs[c].active;
        any |= on[c];
    }
    if (!any) return;
    T q[6] = {0, 0, 0, 0, 0, 0};
    const int64_t n = (int64_t)a.N - a.nRed;
    for (int64_t it = 0;; ++it) {
        const int64_t tl = tile_of_g(it, n, 0, gG, gB);
        if (tl < 0) {
            if (tl == -1) break;
            continue;
        }
        const int32_t r = (int32_t)(a.nRed + tl * BLK + threadIdx.x);
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
            if (k < 32 && !((live >> k) & 1u)) continue;
            const int64_t e = base + (int64_t)SLICE * k;
            const int32_t col = ld_stream<T>(a.ellCol + e);
            const T av = ld_stream<T>(a.ahat + e);
            if (av == (T)0) continue;
            T wc[6];
            ld6(w + (int64_t)NCMP * col, wc);
#pragma unroll
            for (int c = 0; c < 6; ++c) acc[c] -= av * wc[c];
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
    write_partials<6>(qd, a.partials, gB);
}

template <class T>
__global__ void __launch_bounds__(BLK) k_rb_v(KryArgs<T> a) {
    k_rb_v_body<T>(a, gridDim.x, blockIdx.x);
}

// B rows: s = r - alpha v ; t = S M^-1 s = (D_B Ds^-1) s - A_BR w' ;
// |s|, (t, s), (t, t)
template <class T>
__device__ __forceinline__ void k_rb_t_body(const KryArgs<T>& a, int gG, int gB) {
    const T* __restrict__ w = a.t;
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
    const int64_t n = (int64_t)a.N - a.nRed;
    for (int64_t it = 0;; ++it) {
        const int64_t tl = tile_of_g(it, n, 0, gG, gB);
        if (tl < 0) {
            if (tl == -1) break;
            continue;
        }
        const int32_t r = (int32_t)(a.nRed + tl * BLK + threadIdx.x);
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
            if (k < 32 && !((live >> k) & 1u)) continue;
            const int64_t e = base + (int64_t)SLICE * k;
            const int32_t col = ld_stream<T>(a.ellCol + e);
            const T av = ld_stream<T>(a.ahat + e);
            if (av == (T)0) continue;
            T wc[6];
            ld6(w + (int64_t)NCMP * col, wc);
#pragma unroll
            for (int c = 0; c < 6; ++c) acc[c] -= av * wc[c];
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
    write_partials<18>(qd, a.partials, gB);
}

template <class T>
__global__ void __launch_bounds__(BLK) k_rb_t(KryArgs<T> a) {
    k_rb_t_body<T>(a, gridDim.x, blockIdx.x);
}

// The sweep's elimination of R (FP64): x_R = D_R^-1 (b_R - A_RB x_B) for the
// components about to be refined; gated like the refinement sweep it opens.
template <class T>
__device__ __forceinline__ void k_rb_backsub_body(const KryArgs<T>& a, int gG, int gB) {
    bool on[6];
    bool any = false;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        on[c] = a.valid[c] > 0 && a.cs[c].needRefine;
        any |= on[c];
    }
    if (!any || inner_running(a)) return;
    const int64_t n = a.nRed;
    for (int64_t it = 0;; ++it) {
        const int64_t tl = tile_of_g(it, n, 0, gG, gB);
        if (tl < 0) {
            if (tl == -1) break;
            continue;
        }
        const int32_t r = (int32_t)(tl * BLK + threadIdx.x);
        if (r >= n) continue;
        double acc[6];
        ld6(a.b + (int64_t)NCMP * r, acc);
        const int s = r >> 6, lane = r & 63;
        const int64_t base = a.ellOff[s] + lane;
        const int width = a.ellWidth[s];
        const uint32_t live = a.slotMask ? a.slotMask[s] : 0xffffffffu;
        for (int k = 0; k < width; ++k) {
            if (k < 32 && !((live >> k) & 1u)) continue;
            const int64_t e = base + (int64_t)SLICE * k;
            const int32_t col = a.ellCol[e];
            const double av = a.aval[e];
            if (av == 0.0) continue;
            double xc[6];
            ld6(a.x + (int64_t)NCMP * col, xc);
#pragma unroll
            for (int c = 0; c < 6; ++c) acc[c] -= av * xc[c];
        }
        double xv[6];
        ld6(a.x + (int64_t)NCMP * r, xv);
        const double rd = a.rD[r];
#pragma unroll
        for (int c = 0; c < 6; ++c)
            if (on[c]) xv[c] = rd * acc[c];
        st6(a.x + (int64_t)NCMP * r, xv);
    }
}

template <class T>
__global__ void __launch_bounds__(BLK) k_rb_backsub(KryArgs<T> a) {
    k_rb_backsub_body<T>(a, gridDim.x, blockIdx.x);
}

// ahat of the red-black solve: A_jk / (D_j Ds_k) on R's rows, A_ij on B's
template <class T>
__global__ void k_rb_scale(int32_t N, int32_t nRed, const int64_t* __restrict__ off,
                           const int32_t* __restrict__ width, const uint32_t* __restrict__ slotMask,
                           const int32_t* __restrict__ col, const double* __restrict__ aval,
                           const double* __restrict__ rD, const double* __restrict__ rDs, T* __restrict__ ahat) {
    const int64_t r = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (r >= N) return;
    const int s = (int)(r >> 6), lane = (int)(r & 63);
    const int64_t base = off[s] + lane;
    const int w = width[s];
    const uint32_t live = slotMask ? slotMask[s] : 0xffffffffu;
    const bool red = r < nRed;
    const double rdr = red ? rD[r] : 1.0;
    for (int k = 0; k < w; ++k) {
        if (k < 32 && !((live >> k) & 1u)) continue;
        const int64_t e = base + (int64_t)SLICE * k;
        ahat[e] = red ? (T)((aval[e] * rdr) * rDs[col[e]]) : (T)aval[e];
    }
}

// Pivots of B for a general LDU matrix (the sparseSolver surface):
// Ds_i = D_i - sum over i's faces of lower_f upper_f / D_other, in face order;
// out[0, N) = 1/Ds (1/D on R), out[N, 2N) = D Ds^-1 (1 on R)
__global__ void k_rb_schur_diag(int32_t N, int32_t nRed, const int64_t* __restrict__ off,
                                const int32_t* __restrict__ width, const int32_t* __restrict__ fid,
                                const uint8_t* __restrict__ slot, const int32_t* __restrict__ col,
                                const double* __restrict__ lower, const double* __restrict__ upper,
                                const double* __restrict__ diag, double* __restrict__ out) {
    const int64_t r = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (r >= N) return;
    const double D = diag[r];
    double ds = D;
    if (r >= nRed) {
        const int s = (int)(r >> 6), lane = (int)(r & 63);
        const int64_t base = off[s] + lane;
        const int w = width[s];
        for (int k = 0; k < w; ++k) {
            const int64_t e = base + (int64_t)SLICE * k;
            const int32_t f = fid[e];
            if (f == RT_FID_PAD) continue;
            const int32_t face = f >= 0 ? f : -f - 1;
            const int32_t other = col[base + (int64_t)SLICE * slot[e]];
            ds = ds - (lower[face] * upper[face]) / diag[other];
        }
    }
    out[r] = 1.0 / ds;
    out[N + r] = D / ds;
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
        const int64_t tl = tile_of(it, a.N, a.nRed);
        if (tl < 0) {
            if (tl == -1) break;
            continue;
        }
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
        const int64_t tl = tile_of(it, a.N, a.nRed);
        if (tl < 0) {
            if (tl == -1) break;
            continue;
        }
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
                // a correction to fold exists once this inner solve has updated e
                cs.foldE = a.ctl.mixed && cs.innerIter > 0 ? 1 : 0;
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
        const int64_t tl = tile_of(it, a.N, a.nRed);
        if (tl < 0) {
            if (tl == -1) break;
            continue;
        }
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
                if (a.ctl.mixed || breakdown) {
                    cs.needRefine = 1;
                    cs.foldE = a.ctl.mixed && cs.innerIter > 0 ? 1 : 0;
                } else {
                    cs.converged = conv_outer(cs, a.ctl, cs.finalRes);
                }
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
        if (tl < 0) {
            if (tl == -1) break;
            continue;
        }
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

void launch_rb_schur_diag(Ctx* ctx, const Mesh& m, const double* lower, const double* upper, const double* diag,
                          double* out) {
    if (m.N <= 0) return;
    ProfScope ps(ctx, "k_rb_schur_diag", (double)m.N * 32 + (double)m.nEll * 9 + (double)m.F * 16);
    hipLaunchKernelGGL(k_rb_schur_diag, dim3(kblk(m.N)), dim3(BLK), 0, ctx->stream, m.N, m.nRed,
                       m.ellOff.as<int64_t>(), m.ellWidth.as<int32_t>(), m.ellFid.as<int32_t>(),
                       m.ellSlot.as<uint8_t>(), m.ellCol.as<int32_t>(), lower, upper, diag, out);
    RT_HIP(hipGetLastError());
}

void launch_mean6(Ctx* ctx, int32_t N, const double* x, double* partials, CompState* cs,
                  int grid) {
    {
        ProfScope ps(ctx, "k_mean6", (double)N * 48);
        hipLaunchKernelGGL(k_mean6, dim3(grid), dim3(BLK), 0, ctx->stream, N, x, partials);
    }
    hipLaunchKernelGGL(k_fin_mean6, dim3(1), dim3(BLK), 0, ctx->stream, N, grid, partials, cs);
    RT_HIP(hipGetLastError());
}

KryCtl kryCtl(const Controls& C, bool mixed, const Mesh& m) {
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
    // multicolour ILU0 runs as the red-black solve on a colour-blocked mesh
    // with PBiCGStab; the DILU keyword takes it where it is available and the
    // diagonal preconditioner elsewhere, multicolourILU0 demands it
    k.rb = C.precond == PC_MCILU0 && m.nRed > 0 && C.solver == SV_PBICGSTAB ? 1 : 0;
    if (C.precond == PC_MCILU0 && !k.rb) {
        if (C.precondStrict)
            throw Error{RT_E_UNSUPPORTED,
                        std::string("preconditioner multicolourILU0 needs solver PBiCGStab and a mesh whose cell "
                                    "graph is two-colourable (this one is ") +
                            (m.nRed > 0 ? "two-coloured; the solver is " + C.solverName : std::string("not")) + ")"};
        k.precond = PC_JACOBI;
    }
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
    int Gw = 1;                       // red-black: the first colour's gather passes (no partial sums)
    double N = 0, ne = 0, nValid = 0;
    int32_t valid[6];
    Ctx* ctx = nullptr;
    SolveState* hostState = nullptr;  // pinned block of ctx
    CompState* host = nullptr;        // = hostState->c
    bool neFromDevice = false;        // matrix entry count comes with the first status pull
    bool needScale = false;           // Ahat (or the float copies of A, A^T) not built yet
    bool bicg = false;                // PBiCG instead of PBiCGStab
    bool rb = false;                  // red-black ILU0: Krylov on the second colour's Schur complement
    double nRed = 0;                  // first colour's rows
    KryArgs<float> bf{};              // red-black: af/ad shifted to the second colour's rows
    KryArgs<double> bd{};             // (the pointwise passes p, x/r and the fold)
    KryArgs<float> sf{};              // red-black: the residual restricted to the second colour
    KryArgs<double> sd{};
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
KryArgs<T>& blackOf(SolveRun& R);
template <>
KryArgs<float>& blackOf<float>(SolveRun& R) { return R.bf; }
template <>
KryArgs<double>& blackOf<double>(SolveRun& R) { return R.bd; }
template <class T>
KryArgs<T>& sweepOf(SolveRun& R);
template <>
KryArgs<float>& sweepOf<float>(SolveRun& R) { return R.sf; }
template <>
KryArgs<double>& sweepOf<double>(SolveRun& R) { return R.sd; }

// Workgroups of the solve's row kernels (the passes that launch with the
// solve grid and write partial sums) that fit on the chip at once: the least
// over those kernels of the workgroups per CU their VGPRs allow, times the CUs.
// Each workgroup walks its share of the tiles in a loop, so a grid larger than
// that runs in rounds, and the workgroups of a later round find the
// neighbours they gather gone from L2.
template <class T>
int solve_grid(const Ctx* ctx, int cap, bool rb, bool bicg) {
    static int perCU[2][2] = {{0, 0}, {0, 0}};  // [rb][bicg], the same on every gfx950 device
    int& pc = perCU[rb ? 1 : 0][bicg ? 1 : 0];
    if (pc == 0) {
        int least = 8;
        auto take = [&](auto fn) {
            int nb = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, BLK, 0) == hipSuccess && nb > 0)
                least = std::min(least, nb);
        };
        take(k_residual<T, 0>);
        take(k_residual<T, 1>);
        if (bicg) {
            take(k_bicg_init<T>);
            take(k_bicg_w<T>);
            take(k_bicg_xr<T>);
        } else if (rb) {
            take(k_rb_backsub<T>);
            take(k_rb_w<T, 0>);
            take(k_rb_w<T, 1>);
            take(k_rb_v<T>);
            take(k_rb_t<T>);
            take(k_kry_xr<T>);
        } else {
            take(k_kry_v<T>);
            take(k_kry_t<T>);
            take(k_kry_xr<T>);
        }
        pc = least;
    }
    int64_t g = std::min<int64_t>(cap, (int64_t)ctx->numCUs * pc);
    if (g >= 64) g = g / 8 * 8;
    return (int)std::max<int64_t>(1, g);
}

// The red-black passes over the first colour's rows (w, w', the elimination)
// reduce nothing, so they take a grid of their own: every workgroup their
// VGPRs allow at once.
template <class T>
int rb_wide_grid(const Ctx* ctx, int cap) {
    static int perCU = 0;
    if (perCU == 0) {
        int least = 8;
        auto take = [&](auto fn) {
            int nb = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, BLK, 0) == hipSuccess && nb > 0)
                least = std::min(least, nb);
        };
        take(k_rb_w<T, 0>);
        take(k_rb_w<T, 1>);
        take(k_rb_backsub<T>);
        perCU = least;
    }
    int64_t g = std::min<int64_t>(cap, (int64_t)ctx->numCUs * perCU);
    if (g >= 64) g = g / 8 * 8;
    return (int)std::max<int64_t>(1, g);
}

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
    a.ctl = kryCtl(*sa.ctl, mixed, m);
    R.G = solve_grid<T>(ctx, K.grid, a.ctl.rb != 0, sa.ctl->solver == SV_PBICG);
    R.Gw = a.ctl.rb ? rb_wide_grid<T>(ctx, K.grid) : R.G;
    a.G = R.G;
    a.nRed = m.nRed;
    a.rowBegin = 0;
    // the preconditioner: Jacobi M = D (M^-1 = rD, A M^-1 has a unit diagonal)
    // or none (M^-1 = 1, diag(A M^-1) = D); red-black: M = Ds on the second
    // colour (M^-1 = 1/Ds, diag(S M^-1) = D Ds^-1, 1 when Ds = D)
    const bool jac = a.ctl.precond == PC_JACOBI;
    a.rM = jac ? rD : nullptr;
    a.dg = jac ? nullptr : sa.diag;
    R.rb = a.ctl.rb != 0;
    R.nRed = m.nRed;
    if (R.rb) {
        a.rM = sa.rDs ? sa.rDs : rD;
        a.dg = sa.rDs ? sa.rDs + N : nullptr;
        // pointwise passes on the second colour's rows: the same kernels on
        // pointers shifted by nRed rows
        const int64_t o = (int64_t)NCMP * m.nRed;
        KryArgs<T>& b = blackOf<T>(R);
        b = a;
        b.N = N - m.nRed;
        b.nRed = 0;
        b.x = a.x + o;
        b.b = a.b + o;
        b.r += o; b.r0 += o; b.p += o; b.v += o; b.t += o;
        if (b.e) b.e += o;
        b.rM = a.rM + m.nRed;
        b.dg = a.dg ? a.dg + m.nRed : nullptr;
        b.rD = a.rD + m.nRed;
        b.diag = a.diag + m.nRed;
        KryArgs<T>& sw = sweepOf<T>(R);
        sw = a;
        sw.rowBegin = m.nRed;
    }
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
    // red-black (N = rows of the colour the pass runs on, ne its ELL entries):
    // w gathers one or two vectors of the other colour and writes one; the
    // second colour's v and t passes gather w and stream their own vectors;
    // the elimination of the first colour and the sweep's residual in FP64
    double rbw(double nc, int src) const { return ne * (4 + w) + nc * N * (src ? 3 : 2) * w; }
    double rbv(double nc) const { return ne * (4 + w) + nc * N * 4 * w; }
    double rbt(double nc) const { return ne * (4 + w) + nc * N * 5 * w; }
    double rbx(double nc) const { return N * 8.0 + ne * 12 + nc * N * (8 + 8 + 8); }
};

// the scaled off-diagonals of the solve (Jacobi: A D^-1; red-black: its two
// colours' forms), built before the first iteration
template <class T>
void launch_scale_of(Ctx* ctx, SolveRun& R) {
    const bool mixed = std::is_same<T, float>::value;
    KryArgs<T>& a = argsOf<T>(R);
    if (R.rb) {
        ProfScope ps(ctx, mixed ? "k_rb_scale<f32>" : "k_rb_scale<f64>", R.ne * (4 + 8 + sizeof(T)) + R.N * 8);
        hipLaunchKernelGGL(k_rb_scale<T>, dim3(kblk((int64_t)R.N)), dim3(BLK), 0, ctx->stream, a.N, a.nRed, a.ellOff,
                           a.ellWidth, a.slotMask, a.ellCol, a.aval, a.rD, a.rM, const_cast<T*>(a.ahat));
    } else {
        ProfScope ps(ctx, mixed ? "k_scale<f32>" : "k_scale<f64>", R.ne * (4 + 8 + sizeof(T)) + R.N * 8);
        hipLaunchKernelGGL(k_scale<T>, dim3(kblk((int64_t)R.N)), dim3(BLK), 0, ctx->stream, a.N, a.ellOff,
                           a.ellWidth, a.slotMask, a.ellCol, a.aval, a.rM, const_cast<T*>(a.ahat));
    }
    RT_HIP(hipGetLastError());
    R.needScale = false;
}

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
// fold: a component about to be refined holds an inner correction (the
// blocking path knows; queued rounds launch it gated on the device).
// Red-black: the fold and the residual run on the second colour's rows and
// the elimination of the first colour runs between them.
template <class T>
void launch_refine(Ctx* ctx, SolveRun& R, double nRef, bool fold = true) {
    const bool mixed = std::is_same<T, float>::value;
    Bytes B{R.N, R.ne, (double)sizeof(T)};
    KryArgs<T>& a = argsOf<T>(R);
    if (R.rb) {
        const double w = (double)sizeof(T);
        Bytes Bb{R.N - R.nRed, 0.5 * R.ne, w}, Br{R.nRed, 0.5 * R.ne, w};
        const int pg = std::max(1, std::min<int>(R.pgrid, (int)((R.N - R.nRed + BLK - 1) / BLK)));
        if (mixed && fold) {
            ProfScope ps(ctx, "k_fold<f32>", Bb.fold(nRef));
            hipLaunchKernelGGL(k_fold<T>, dim3(pg), dim3(BLK), 0, ctx->stream, blackOf<T>(R));
        }
        {
            ProfScope ps(ctx, mixed ? "k_rb_backsub<f32>" : "k_rb_backsub<f64>", Br.rbx(nRef));
            hipLaunchKernelGGL(k_rb_backsub<T>, dim3(R.Gw), dim3(BLK), 0, ctx->stream, a);
        }
        {
            ProfScope ps(ctx, mixed ? "k_residual<f32>" : "k_residual<f64>", Bb.res(nRef));
            hipLaunchKernelGGL((k_residual<T, 1>), dim3(R.G), dim3(BLK), 0, ctx->stream, sweepOf<T>(R));
        }
        hipLaunchKernelGGL((k_fin_residual<T, 1>), dim3(1), dim3(BLK), 0, ctx->stream, a);
        return;
    }
    if (mixed && fold) {
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
    if (R.rb) {
        if (R.needScale) launch_scale_of<T>(ctx, R);
        const double w = (double)sizeof(T);
        Bytes Bb{R.N - R.nRed, 0.5 * R.ne, w}, Br{R.nRed, 0.5 * R.ne, w};
        KryArgs<T>& b = blackOf<T>(R);
        const int pg = std::max(1, std::min<int>(R.pgrid, (int)((R.N - R.nRed + BLK - 1) / BLK)));
        // w lives on the first colour's rows of v, w' on those of t
        {
            ProfScope ps(ctx, mixed ? "k_kry_p<f32>" : "k_kry_p<f64>", Bb.p(nAct, start));
            hipLaunchKernelGGL(k_kry_p<T>, dim3(pg), dim3(BLK), 0, ctx->stream, b);
        }
        {
            ProfScope ps(ctx, mixed ? "k_rb_w<f32>" : "k_rb_w<f64>", Br.rbw(nAct, 0));
            hipLaunchKernelGGL((k_rb_w<T, 0>), dim3(R.Gw), dim3(BLK), 0, ctx->stream, a);
        }
        {
            ProfScope ps(ctx, mixed ? "k_rb_v<f32>" : "k_rb_v<f64>", Bb.rbv(nAct));
            hipLaunchKernelGGL(k_rb_v<T>, dim3(R.G), dim3(BLK), 0, ctx->stream, a);
        }
        hipLaunchKernelGGL(k_fin_v<T>, dim3(1), dim3(BLK), 0, ctx->stream, a);
        {
            ProfScope ps(ctx, mixed ? "k_rb_ws<f32>" : "k_rb_ws<f64>", Br.rbw(nAct, 1));
            hipLaunchKernelGGL((k_rb_w<T, 1>), dim3(R.Gw), dim3(BLK), 0, ctx->stream, a);
        }
        {
            ProfScope ps(ctx, mixed ? "k_rb_t<f32>" : "k_rb_t<f64>", Bb.rbt(nAct));
            hipLaunchKernelGGL(k_rb_t<T>, dim3(R.G), dim3(BLK), 0, ctx->stream, a);
        }
        hipLaunchKernelGGL(k_fin_t<T>, dim3(1), dim3(BLK), 0, ctx->stream, a);
        {
            ProfScope ps(ctx, mixed ? "k_kry_xr<f32>" : "k_kry_xr<f64>", Bb.xr(nAct));
            hipLaunchKernelGGL(k_kry_xr<T>, dim3(R.G), dim3(BLK), 0, ctx->stream, b);
        }
        hipLaunchKernelGGL(k_fin_xr<T>, dim3(1), dim3(BLK), 0, ctx->stream, a);
        return;
    }
    if (R.needScale) launch_scale_of<T>(ctx, R);
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
            bool fold = false;
            for (int c = 0; c < 6; ++c) fold |= R.valid[c] > 0 && R.host[c].needRefine && R.host[c].foldE;
            if (R.mixed) launch_refine<float>(ctx, R, nRef, fold);
            else launch_refine<double>(ctx, R, nRef, fold);
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

// ---------------------------------------------------------- grouped rounds
namespace {

// the member of workgroup blk: starts[i] <= blk < starts[i + 1]
__device__ __forceinline__ int group_of(const int32_t* __restrict__ starts, int n, int blk) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (starts[mid] <= blk) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <class T, void (*Body)(const KryArgs<T>&, int, int)>
__global__ void __launch_bounds__(BLK) k_group(const KryArgs<T>* __restrict__ tab, const int32_t* __restrict__ starts,
                                               int n) {
    const int i = group_of(starts, n, blockIdx.x);
    const KryArgs<T> a = tab[i];
    Body(a, starts[i + 1] - starts[i], blockIdx.x - starts[i]);
}

// device tables of a group: the members' arguments (whole rows; the red-black
// second colour's rows; its sweep residual) and the workgroup offsets of each
// launch shape (row kernels, pointwise passes, finalizers)
template <class T>
struct GroupDev {
    const KryArgs<T>* A;
    const KryArgs<T>* B;
    const KryArgs<T>* S;
    const int32_t *G, *P, *One, *W;
    int n, totG, totP, totW;
};

template <class T, void (*Body)(const KryArgs<T>&, int, int)>
void glaunch(hipStream_t st, const KryArgs<T>* tab, const int32_t* starts, int total, int n) {
    if (total > 0) hipLaunchKernelGGL((k_group<T, Body>), dim3(total), dim3(BLK), 0, st, tab, starts, n);
}

// One round as solve_advance queues it ahead for one solve: the refinement
// sweep (every kernel gated on the device) and one iteration.
template <class T>
void group_round(bool rb, hipStream_t st, const GroupDev<T>& d) {
    const int n = d.n;
    if (rb) {
        glaunch<T, &k_fold_body<T>>(st, d.B, d.P, d.totP, n);
        glaunch<T, &k_rb_backsub_body<T>>(st, d.A, d.W, d.totW, n);
        glaunch<T, &k_residual_body<T, 1>>(st, d.S, d.G, d.totG, n);
        glaunch<T, &k_fin_residual_body<T, 1>>(st, d.A, d.One, n, n);
        glaunch<T, &k_kry_p_body<T>>(st, d.B, d.P, d.totP, n);
        glaunch<T, &k_rb_w_body<T, 0>>(st, d.A, d.W, d.totW, n);
        glaunch<T, &k_rb_v_body<T>>(st, d.A, d.G, d.totG, n);
        glaunch<T, &k_fin_v_body<T>>(st, d.A, d.One, n, n);
        glaunch<T, &k_rb_w_body<T, 1>>(st, d.A, d.W, d.totW, n);
        glaunch<T, &k_rb_t_body<T>>(st, d.A, d.G, d.totG, n);
        glaunch<T, &k_fin_t_body<T>>(st, d.A, d.One, n, n);
        glaunch<T, &k_kry_xr_body<T>>(st, d.B, d.G, d.totG, n);
        glaunch<T, &k_fin_xr_body<T>>(st, d.A, d.One, n, n);
    } else {
        glaunch<T, &k_fold_body<T>>(st, d.A, d.P, d.totP, n);
        glaunch<T, &k_residual_body<T, 1>>(st, d.A, d.G, d.totG, n);
        glaunch<T, &k_fin_residual_body<T, 1>>(st, d.A, d.One, n, n);
        glaunch<T, &k_kry_p_body<T>>(st, d.A, d.P, d.totP, n);
        glaunch<T, &k_kry_v_body<T>>(st, d.A, d.G, d.totG, n);
        glaunch<T, &k_fin_v_body<T>>(st, d.A, d.One, n, n);
        glaunch<T, &k_kry_t_body<T>>(st, d.A, d.G, d.totG, n);
        glaunch<T, &k_fin_t_body<T>>(st, d.A, d.One, n, n);
        glaunch<T, &k_kry_xr_body<T>>(st, d.A, d.G, d.totG, n);
        glaunch<T, &k_fin_xr_body<T>>(st, d.A, d.One, n, n);
    }
    RT_HIP(hipGetLastError());
}

// the workgroups of a member's pointwise passes (p, fold)
int point_grid(const SolveRun& R) {
    if (!R.rb) return R.pgrid;
    return std::max(1, std::min<int>(R.pgrid, (int)((R.N - R.nRed + BLK - 1) / BLK)));
}

template <class T>
void group_launch(SolveGroup& g) {
    const int n = (int)g.members.size();
    const size_t szA = sizeof(KryArgs<T>) * n, szI = sizeof(int32_t) * (n + 1);
    const size_t offB = szA, offS = 2 * szA, offG = 3 * szA, offP = offG + szI, offO = offP + szI,
                 offW = offO + szI;
    const size_t bytes = offW + szI;
    int totG = 0, totP = 0, totW = 0;
    if (g.dirty) {
        if (g.hostBytes < bytes) {
            if (g.hostTables) (void)hipHostFree(g.hostTables);
            g.hostTables = nullptr;
            RT_HIP(hipHostMalloc(&g.hostTables, bytes, hipHostMallocDefault));
            g.hostBytes = bytes;
        }
        if (g.tables.bytes < bytes) g.tables = g.ctx->alloc(bytes);
        unsigned char* h = static_cast<unsigned char*>(g.hostTables);
        auto* tA = reinterpret_cast<KryArgs<T>*>(h);
        auto* tB = reinterpret_cast<KryArgs<T>*>(h + offB);
        auto* tS = reinterpret_cast<KryArgs<T>*>(h + offS);
        auto* sG = reinterpret_cast<int32_t*>(h + offG);
        auto* sP = reinterpret_cast<int32_t*>(h + offP);
        auto* sO = reinterpret_cast<int32_t*>(h + offO);
        auto* sW = reinterpret_cast<int32_t*>(h + offW);
        sG[0] = sP[0] = sO[0] = sW[0] = 0;
        for (int i = 0; i < n; ++i) {
            SolveRun& R = *g.members[i].K->run;
            tA[i] = argsOf<T>(R);
            tB[i] = g.rb ? blackOf<T>(R) : argsOf<T>(R);
            tS[i] = g.rb ? sweepOf<T>(R) : argsOf<T>(R);
            sG[i + 1] = sG[i] + R.G;
            sP[i + 1] = sP[i] + point_grid(R);
            sO[i + 1] = i + 1;
            sW[i + 1] = sW[i] + R.Gw;
        }
        // the previous launch's kernels have all finished (their pulls are back)
        RT_HIP(hipMemcpyAsync(g.tables.p, h, bytes, hipMemcpyHostToDevice, g.ctx->stream));
        g.dirty = false;
    }
    for (auto& mb : g.members) {
        SolveRun& R = *mb.K->run;
        totG += R.G;
        totP += point_grid(R);
        totW += R.Gw;
    }
    const unsigned char* d = static_cast<const unsigned char*>(g.tables.p);
    GroupDev<T> dv{reinterpret_cast<const KryArgs<T>*>(d), reinterpret_cast<const KryArgs<T>*>(d + offB),
                   reinterpret_cast<const KryArgs<T>*>(d + offS), reinterpret_cast<const int32_t*>(d + offG),
                   reinterpret_cast<const int32_t*>(d + offP), reinterpret_cast<const int32_t*>(d + offO),
                   reinterpret_cast<const int32_t*>(d + offW), n, totG, totP, totW};
    for (int k = 0; k < kRoundsAhead; ++k) group_round<T>(g.rb, g.ctx->stream, dv);
    for (auto& mb : g.members) {
        SolveRun& R = *mb.K->run;
        R.rounds += kRoundsAhead;
        queue_pull(g.ctx, *mb.K, R);
    }
    g.launched = true;
    ++g.launches;
}

}  // namespace

SolveGroup::~SolveGroup() {
    if (hostTables) (void)hipHostFree(hostTables);
}

bool solve_groupable(const Krylov& K, bool* mixed, bool* rb) {
    if (!K.run || K.run->bicg || !K.run->ev) return false;
    if (mixed) *mixed = K.run->mixed;
    if (rb) *rb = K.run->rb;
    return true;
}

void group_join(SolveGroup& g, Krylov& K, hipStream_t home) {
    SolveRun& R = *K.run;
    if (R.bicg || R.mixed != g.mixed || R.rb != g.rb || !R.ev)
        throw Error{RT_E_BAD_ARG, "group_join: the solve does not belong to this group"};
    // the group's stream waits for the solve's own queued work (the assembly,
    // the initial residual and the first state pull)
    RT_HIP(hipStreamWaitEvent(g.ctx->stream, R.ev, 0));
    if (R.needScale) {  // the scaled matrix its first iteration would build
        if (R.mixed) launch_scale_of<float>(g.ctx, R);
        else launch_scale_of<double>(g.ctx, R);
    }
    g.members.push_back({&K, home});
    g.dirty = true;
}

void group_remove(SolveGroup& g, Krylov& K) {
    for (size_t i = 0; i < g.members.size(); ++i)
        if (g.members[i].K == &K) {
            RT_HIP(hipStreamSynchronize(g.ctx->stream));
            if (K.run) K.run->pending = false;
            g.members.erase(g.members.begin() + i);
            g.dirty = true;
            return;
        }
}

void group_advance(SolveGroup& g, std::vector<Krylov*>& finished) {
    bool allBack = true;
    for (auto& mb : g.members) {
        SolveRun& R = *mb.K->run;
        if (!R.pending) continue;
        hipError_t e = hipEventQuery(R.ev);
        if (e == hipErrorNotReady) {
            allBack = false;
            continue;
        }
        RT_HIP(e);
        R.pending = false;
        if (R.neFromDevice) {
            R.ne = R.hostState->misc[0];
            R.neFromDevice = false;
        }
    }
    if (!allBack) return;
    g.launched = false;
    // members whose solve has finished leave; their streams go on after the
    // group's queued work (which reads them) has drained
    hipEvent_t drained = nullptr;
    for (size_t i = 0; i < g.members.size();) {
        SolveRun& R = *g.members[i].K->run;
        double nAct = 0, nRef = 0;
        for (int c = 0; c < 6; ++c) {
            if (R.valid[c] <= 0) continue;
            nAct += R.host[c].active != 0;
            nRef += R.host[c].needRefine != 0;
        }
        if ((nAct == 0 && nRef == 0) || R.rounds >= 100000) {
            if (!drained) {
                RT_HIP(hipEventCreateWithFlags(&drained, hipEventDisableTiming));
                RT_HIP(hipEventRecord(drained, g.ctx->stream));
            }
            RT_HIP(hipStreamWaitEvent(g.members[i].home, drained, 0));
            R.done = true;
            finished.push_back(g.members[i].K);
            g.members.erase(g.members.begin() + i);
            g.dirty = true;
        } else {
            ++i;
        }
    }
    if (drained) RT_HIP(hipEventDestroy(drained));  // released once the waits are satisfied
    if (g.members.empty()) return;
    if (g.mixed) group_launch<float>(g);
    else group_launch<double>(g);
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
