// rt_device.h -- device helpers shared by the cell and Krylov kernels (gfx950).
#pragma once

#include "rt_internal.h"

namespace rt {

constexpr int BLK = 256;

// Row tiles of BLK rows.  Workgroups are dealt round-robin to the 8 XCDs
// (XCD = blockIdx.x % 8), each with its own L2.  XCD x owns the contiguous
// slab of tiles [x*S, (x+1)*S), S = ceil(T/8), and its G/8 workgroups walk
// the slab together: tile i*(G/8) + blockIdx.x/8 of the slab at step i.  The
// rows a tile gathers (bandwidth-reduced numbering keeps them within a level
// of the row) were touched moments before by the same XCD, so they are still
// in its L2.
// nRed > 0: the rows are colour-blocked (rows [0, nRed) have neighbours only
// in [nRed, N) and the other way round).  The tiles of each block are cut
// into 8 slabs and XCD x walks slab x of both blocks, alternating, so the
// other colour's rows a tile gathers are the ones the same XCD walks at the
// same time.
// Returns the tile, -2 for a step with no tile (a slab shorter than the
// others), -1 once the workgroup's share is done.
// tile_of_g: the same for workgroup b of G (a request's share of a grouped launch)
__device__ __forceinline__ int64_t tile_of_g(int64_t i, int64_t N, int64_t nRed, int G, int b) {
    const int64_t T = (N + BLK - 1) / BLK;
    if (G % 8 != 0) {
        const int64_t t = i * G + b;
        return t < T ? t : -1;
    }
    const int64_t local = i * (G / 8) + b / 8;
    const int64_t x = b % 8;
    const int64_t Ts = nRed / BLK;  // tiles [0, Ts) hold first-colour rows only
    if (nRed <= 0 || Ts == 0 || Ts >= T) {
        const int64_t S = (T + 7) / 8;
        if (local >= S) return -1;
        const int64_t t = x * S + local;
        return t < T ? t : -1;
    }
    const int64_t S1 = (Ts + 7) / 8, S2 = (T - Ts + 7) / 8;
    if (local >= S1 + S2) return -1;
    const int64_t m = S1 < S2 ? S1 : S2;
    bool first;
    int64_t k;
    if (local < 2 * m) {
        first = (local & 1) == 0;
        k = local >> 1;
    } else {
        first = S1 > S2;
        k = local - m;
    }
    if (first) {
        const int64_t t = x * S1 + k;
        return t < Ts ? t : -2;
    }
    const int64_t t = Ts + x * S2 + k;
    return t < T ? t : -2;
}
__device__ __forceinline__ int64_t tile_of(int64_t i, int64_t N, int64_t nRed = 0) {
    return tile_of_g(i, N, nRed, gridDim.x, blockIdx.x);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Grid reductions in two launches: every workgroup of the row kernel writes
// its NQ partial sums to partials[block][NQ] (plain stores, no fence), then a
// one-workgroup finalize kernel on the same stream reduces all partials in a
// fixed order (deterministic) and updates the solver state.  The kernel
// boundary makes the partials visible across the XCDs' L2s without per-block
// agent-scope release fences.
template <int NQ>
__device__ __forceinline__ void write_partials(const double (&v)[NQ], double* partials, int blk) {
    __shared__ double red[NQ][BLK / 64];
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
        partials[(int64_t)blk * NQ + threadIdx.x] = s;
    }
}
template <int NQ>
__device__ __forceinline__ void write_partials(const double (&v)[NQ], double* partials) {
    write_partials<NQ>(v, partials, blockIdx.x);
}

// In a one-workgroup finalize kernel: sums over the G blocks' partials
// (thread 0 holds the result).
template <int NQ>
__device__ __forceinline__ void sum_partials(const double* partials, int G, double (&sums)[NQ]) {
    __shared__ double red[NQ][BLK / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    double acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = 0;
    for (int i = threadIdx.x; i < G; i += BLK) {
        const double* row = partials + (int64_t)i * NQ;
        double tmp[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) tmp[q] = row[q];
#pragma unroll
        for (int q = 0; q < NQ; ++q) acc[q] += tmp[q];
    }
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
    }
}

// interleaved [N][6] rows: three 2-wide vector loads/stores per row
template <class T> struct Vec2;
template <> struct Vec2<float> { typedef float2 type; };
template <> struct Vec2<double> { typedef double2 type; };

template <class T>
__device__ __forceinline__ void ld6(const T* p, T* v) {
    typedef typename Vec2<T>::type V2;
    const V2* q = reinterpret_cast<const V2*>(p);
    const V2 a = q[0], b = q[1], c = q[2];
    v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y; v[4] = c.x; v[5] = c.y;
}
template <class T>
__device__ __forceinline__ void st6(T* p, const T* v) {
    typedef typename Vec2<T>::type V2;
    V2* q = reinterpret_cast<V2*>(p);
    V2 a, b, c;
    a.x = v[0]; a.y = v[1]; b.x = v[2]; b.y = v[3]; c.x = v[4]; c.y = v[5];
    q[0] = a; q[1] = b; q[2] = c;
}

// wedge patch value of a symmTensor ([OF9] wedgeFvPatchField::evaluate,
// transform(faceT, theta_P)): T theta_P T^T, T (row-major) the patch's faceT;
// the same operations as the oracle's
__device__ __forceinline__ void wedge_value(const double* __restrict__ T, const double th[6], double out[6]) {
    double A[3][3], M[3][3];
    A[0][0] = th[0]; A[0][1] = A[1][0] = th[1]; A[0][2] = A[2][0] = th[2];
    A[1][1] = th[3]; A[1][2] = A[2][1] = th[4]; A[2][2] = th[5];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) M[i][k] = T[3 * i] * A[0][k] + T[3 * i + 1] * A[1][k] + T[3 * i + 2] * A[2][k];
    constexpr int I6[6] = {0, 0, 0, 1, 1, 2}, J6[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const int i = I6[q], l = J6[q];
        out[q] = M[i][0] * T[3 * l] + M[i][1] * T[3 * l + 1] + M[i][2] * T[3 * l + 2];
    }
}

// symmetryPlane / symmetry patch value of a symmTensor ([OF9]
// basicSymmetryFvPatchField::evaluate): (theta_P + T theta_P T^T)/2 with
// T = I - 2 n n, n = Sf/|Sf|; the same operations as the oracle's
__device__ __forceinline__ void symmetry_value(const double sf[3], const double th[6], double out[6]) {
    const double mag = sqrt(sf[0] * sf[0] + sf[1] * sf[1] + sf[2] * sf[2]);
    const double n[3] = {sf[0] / mag, sf[1] / mag, sf[2] / mag};
    double T[3][3], A[3][3], M[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) T[i][j] = (i == j ? 1.0 : 0.0) - 2.0 * n[i] * n[j];
    A[0][0] = th[0]; A[0][1] = A[1][0] = th[1]; A[0][2] = A[2][0] = th[2];
    A[1][1] = th[3]; A[1][2] = A[2][1] = th[4]; A[2][2] = th[5];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) M[i][k] = T[i][0] * A[0][k] + T[i][1] * A[1][k] + T[i][2] * A[2][k];
    constexpr int I6[6] = {0, 0, 0, 1, 1, 2}, J6[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const int i = I6[q], l = J6[q];
        const double r = M[i][0] * T[l][0] + M[i][1] * T[l][1] + M[i][2] * T[l][2];
        out[q] = (th[q] + r) / 2.0;
    }
}

// ------------------------------------------------- Krylov control (OpenFOAM)
// SolverPerformance::checkConvergence: final < tolerance, or relTol > 0 and
// final < relTol * initial
__device__ __forceinline__ bool conv_outer(const CompState& c, const KryCtl& k, double res) {
    return res < k.tol || (k.relTol > 0 && res < k.relTol * c.initRes);
}
__device__ __forceinline__ bool conv_inner(const CompState& c, const KryCtl& k, double res) {
    return k.mixed ? res < c.innerTarget : conv_outer(c, k, res);
}

// begin an inner Krylov solve on a component whose r, r0 were just set
__device__ __forceinline__ void start_inner(CompState& c, const KryCtl& k, double rr) {
    c.innerIter = 0;
    c.rho = rr;
    c.exitAtS = 0;
    if (k.mixed) c.innerTarget = fmax(k.innerRelTol * c.finalRes, 0.3 * k.tol);
    if (fabs(rr) < 1e-300) {  // singular: nothing to iterate on
        c.active = 0;
        c.converged = conv_outer(c, k, c.finalRes);
        return;
    }
    c.active = 1;
}

// after the initial residual: normalisation factor, initial residual, start
__device__ __forceinline__ void finish_initial(CompState& cs, const KryCtl& k, double absR, double rr,
                                               double nrm) {
    cs.normFactor = nrm + 1e-20;
    cs.initRes = absR / cs.normFactor;
    cs.finalRes = cs.initRes;
    cs.iter = 0;
    cs.nRefine = 0;
    cs.solved = 1;
    cs.converged = 0;
    cs.needRefine = 0;
    cs.foldE = 0;
    if (k.minIter > 0 || !conv_outer(cs, k, cs.initRes)) {
        if (k.rb) {
            // red-black: a first sweep eliminates the first colour (x there
            // from the second colour's x) and leaves the Schur residual
            cs.active = 0;
            cs.needRefine = 1;
        } else {
            start_inner(cs, k, rr);
        }
    } else {
        cs.active = 0;
        cs.converged = 1;
    }
}

}  // namespace rt
