// rt_mesh.cpp -- mesh upload: locality renumbering, sliced-ELL rows, boundary lists.
//
// The reference solves on OpenFOAM's lduAddressing (owner/neighbour per
// internal face, sparseMatrixSolvers/segregated/eigenSolver/eigenSolver.C:
// 600-641).  On the device every cell is a row of a sliced ELL matrix
// (64-row slices, one row per lane of a wavefront) whose entries keep the face
// id, so the face loops of the reference (Gauss gradient, upwind
// coefficients, negSumDiag, Amul) become coalesced row gathers that add the
// faces of a cell in the same (increasing face) order as the reference's
// scatter loops.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <queue>

#include "rt_internal.h"

namespace rt {

namespace {

// Reverse Cuthill-McKee on the cell graph (faces as edges).
// spreads the low 21 bits of v to every third bit
uint64_t spread3(uint64_t v) {
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffULL;
    v = (v | v << 16) & 0x1f0000ff0000ffULL;
    v = (v | v << 8) & 0x100f00f00f00f00fULL;
    v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
    v = (v | v << 2) & 0x1249249249249249ULL;
    return v;
}

// cells sorted by the Morton key of their centre, quantised to 21 bits per
// axis over the bounding box; ties keep the input order
std::vector<int32_t> mortonOrder(int32_t N, const double* C) {
    double lo[3] = {C[0], C[1], C[2]}, hi[3] = {C[0], C[1], C[2]};
    for (int32_t c = 0; c < N; ++c)
        for (int q = 0; q < 3; ++q) {
            lo[q] = std::min(lo[q], C[3 * (size_t)c + q]);
            hi[q] = std::max(hi[q], C[3 * (size_t)c + q]);
        }
    const double span = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1e-300});
    const double scale = ((1 << 21) - 1) / span;  // one scale: cubic cells stay cubic
    std::vector<std::pair<uint64_t, int32_t>> key(N);
    for (int32_t c = 0; c < N; ++c) {
        uint64_t k = 0;
        for (int q = 0; q < 3; ++q) {
            const uint64_t x = (uint64_t)((C[3 * (size_t)c + q] - lo[q]) * scale);
            k |= spread3(x) << q;
        }
        key[c] = {k, c};
    }
    std::sort(key.begin(), key.end());
    std::vector<int32_t> order(N);
    for (int32_t r = 0; r < N; ++r) order[r] = key[r].second;
    return order;
}

std::vector<int32_t> rcmOrder(int32_t N, const std::vector<int64_t>& ptr,
                              const std::vector<int32_t>& adj) {
    std::vector<int32_t> deg(N);
    for (int32_t c = 0; c < N; ++c) deg[c] = (int32_t)(ptr[c + 1] - ptr[c]);
    std::vector<int32_t> order;
    order.reserve(N);
    std::vector<char> seen(N, 0);
    std::vector<int32_t> level(N, -1);
    auto bfsLast = [&](int32_t s) {  // pseudo-peripheral search helper
        std::vector<int32_t> q{s};
        std::vector<int32_t> touched{s};
        level[s] = 0;
        size_t h = 0;
        int32_t last = s;
        while (h < q.size()) {
            int32_t c = q[h++];
            last = c;
            for (int64_t e = ptr[c]; e < ptr[c + 1]; ++e) {
                int32_t o = adj[e];
                if (level[o] < 0 && !seen[o]) {
                    level[o] = level[c] + 1;
                    q.push_back(o);
                    touched.push_back(o);
                }
            }
        }
        // among the deepest level pick the minimum degree
        int32_t best = last;
        for (int32_t c : touched)
            if (level[c] == level[last] && deg[c] < deg[best]) best = c;
        for (int32_t c : touched) level[c] = -1;
        return best;
    };
    std::vector<int32_t> nb;
    for (int32_t s0 = 0; s0 < N; ++s0) {
        if (seen[s0]) continue;
        int32_t s = bfsLast(bfsLast(s0));
        size_t head = order.size();
        order.push_back(s);
        seen[s] = 1;
        while (head < order.size()) {
            int32_t c = order[head++];
            nb.clear();
            for (int64_t e = ptr[c]; e < ptr[c + 1]; ++e)
                if (!seen[adj[e]]) nb.push_back(adj[e]);
            std::sort(nb.begin(), nb.end(),
                      [&](int32_t a, int32_t b) { return deg[a] != deg[b] ? deg[a] < deg[b] : a < b; });
            for (int32_t o : nb) {
                if (!seen[o]) {
                    seen[o] = 1;
                    order.push_back(o);
                }
            }
        }
    }
    std::reverse(order.begin(), order.end());
    return order;  // order[k] = external cell at internal position k
}

template <class T>
void upload(Ctx* ctx, DBuf& d, const std::vector<T>& h) {
    d = ctx->alloc(std::max<size_t>(1, h.size()) * sizeof(T));
    if (!h.empty())
        RT_HIP(hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
}

}  // namespace

void buildMesh(Ctx* ctx, const rt_mesh_desc* d, int renumber, Mesh& m) {
    if (!d || d->nCells <= 0 || d->nInternalFaces < 0 || d->nBoundaryFaces < 0 ||
        (d->nInternalFaces > 0 && (!d->owner || !d->neighbour)) || !d->V ||
        (d->nInternalFaces + d->nBoundaryFaces > 0 && !d->Sf) ||
        (d->nInternalFaces > 0 && !d->weights) || (d->nPatches > 0 && !d->patches))
        throw Error{RT_E_BAD_ARG, "rt_mesh_create: incomplete mesh description"};
    m.ctx = ctx;
    const int32_t N = d->nCells, F = d->nInternalFaces, NB = d->nBoundaryFaces;
    const int64_t NF = (int64_t)F + NB;
    m.N = N;
    m.F = F;
    m.NB = NB;
    m.nPatches = d->nPatches;
    m.patches.assign(d->patches, d->patches + d->nPatches);
    {
        int64_t tot = 0;
        for (auto& p : m.patches) {
            if (p.start < 0 || p.size < 0 || (int64_t)p.start + p.size > NB)
                throw Error{RT_E_BAD_ARG, "rt_mesh_create: patch outside the boundary faces"};
            tot += p.size;
        }
        if (tot != NB) throw Error{RT_E_BAD_ARG, "rt_mesh_create: patch sizes do not cover the boundary faces"};
    }
    // validate addressing
    for (int32_t f = 0; f < F; ++f) {
        int32_t o = d->owner[f], n = d->neighbour[f];
        if (o < 0 || n < 0 || o >= N || n >= N || o >= n)
            throw Error{RT_E_BAD_ARG, "rt_mesh_create: internal face " + std::to_string(f) +
                                          " violates 0 <= owner < neighbour < nCells"};
    }
    for (int64_t b = 0; b < NB; ++b) {
        int32_t o = d->owner[F + b];
        if (o < 0 || o >= N) throw Error{RT_E_BAD_ARG, "rt_mesh_create: boundary face owner out of range"};
    }
    m.hOwner.assign(d->owner, d->owner + NF);
    m.hNeighbour.assign(d->neighbour, d->neighbour + F);

    // empty direction -> validComponents
    double tot[3] = {0, 0, 0};
    for (auto& p : m.patches)
        if (p.kind == RT_PATCH_EMPTY)
            for (int32_t k = 0; k < p.size; ++k)
                for (int q = 0; q < 3; ++q) tot[q] += std::fabs(d->Sf[3 * ((int64_t)F + p.start + k) + q]);
    m.emptyDir = -1;
    if (tot[0] + tot[1] + tot[2] > 0) m.emptyDir = (int)(std::max_element(tot, tot + 3) - tot);
    {
        int dir[3] = {1, 1, 1};
        if (m.emptyDir >= 0) dir[m.emptyDir] = -1;
        const int ij[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
        for (int q = 0; q < 6; ++q) m.valid[q] = dir[ij[q][0]] * dir[ij[q][1]];
    }

    // row faces (external numbering), ascending face id per row
    std::vector<int64_t> rptr(N + 1, 0);
    for (int32_t f = 0; f < F; ++f) {
        rptr[d->owner[f] + 1]++;
        rptr[d->neighbour[f] + 1]++;
    }
    for (int32_t c = 0; c < N; ++c) rptr[c + 1] += rptr[c];
    std::vector<int32_t> rface(2 * (size_t)F), rcol(2 * (size_t)F);
    {
        std::vector<int64_t> fill(rptr.begin(), rptr.end() - 1);
        for (int32_t f = 0; f < F; ++f) {
            int32_t o = d->owner[f], n = d->neighbour[f];
            rface[fill[o]] = f;
            rcol[fill[o]++] = n;
            rface[fill[n]] = -f - 1;
            rcol[fill[n]++] = o;
        }
    }
    // renumbering decision
    double band = 0;
    for (int32_t f = 0; f < F; ++f) band += (double)(d->neighbour[f] - d->owner[f]);
    m.meanAbsBandwidth = F ? band / F : 0.0;
    const bool colour = !(renumber & RT_RENUMBER_NO_COLOUR);
    renumber &= ~RT_RENUMBER_NO_COLOUR;
    bool doRcm = renumber == RT_RENUMBER_RCM;
    bool doSfc = renumber == RT_RENUMBER_SFC;
    if (doSfc && !d->C) throw Error{RT_E_BAD_ARG, "rt_mesh_create: SFC renumbering needs the cell centres"};
    if (renumber == RT_RENUMBER_AUTO) {
        const double ref = std::pow((double)N, 2.0 / 3.0);
        doRcm = F > 0 && m.meanAbsBandwidth > 8.0 * ref;
    }
    m.iperm.resize(N);
    m.perm.resize(N);
    if (doSfc) {
        m.iperm = mortonOrder(N, d->C);
        m.renumbered = true;
    } else if (doRcm) {
        m.iperm = rcmOrder(N, rptr, rcol);
        m.renumbered = true;
    } else {
        std::iota(m.iperm.begin(), m.iperm.end(), 0);
    }
    // Two-colourable cell graphs (every hexahedral or quadrilateral mesh) are
    // stored colour by colour, each colour in the order chosen above: rows
    // [0, nRed) then [nRed, N), no face inside a colour.  The red-black ILU0
    // solve eliminates the first colour exactly; every other kernel is
    // indifferent to the order.
    if (colour && F > 0) {
        std::vector<int8_t> col(N, -1);  // by external cell
        std::vector<int32_t> q;
        q.reserve(N);
        bool ok = true;
        for (int32_t r0 = 0; r0 < N && ok; ++r0) {
            const int32_t seed = m.iperm[r0];
            if (col[seed] >= 0) continue;
            col[seed] = 0;
            q.clear();
            q.push_back(seed);
            for (size_t h = 0; h < q.size() && ok; ++h) {
                const int32_t c = q[h];
                for (int64_t e = rptr[c]; e < rptr[c + 1]; ++e) {
                    const int32_t o = rcol[e];
                    if (col[o] < 0) {
                        col[o] = (int8_t)(1 - col[c]);
                        q.push_back(o);
                    } else if (col[o] == col[c]) {
                        ok = false;
                        break;
                    }
                }
            }
        }
        if (ok) {
            std::vector<int32_t> ip;
            ip.reserve(N);
            for (int k = 0; k < 2; ++k)
                for (int32_t r = 0; r < N; ++r)
                    if (col[m.iperm[r]] == k) ip.push_back(m.iperm[r]);
            int32_t nRed = 0;
            for (int32_t c = 0; c < N; ++c) nRed += col[c] == 0;
            if (nRed > 0 && nRed < N) {
                m.iperm.swap(ip);
                m.nRed = nRed;
                m.renumbered = true;
            }
        }
    }
    for (int32_t r = 0; r < N; ++r) m.perm[m.iperm[r]] = r;

    // sliced ELL in internal order
    m.nSlices = (N + SLICE - 1) / SLICE;
    std::vector<int64_t> off(m.nSlices + 1, 0);
    std::vector<int32_t> width(m.nSlices, 0);
    for (int32_t s = 0; s < m.nSlices; ++s) {
        int wmax = 0;
        for (int l = 0; l < SLICE; ++l) {
            int32_t r = s * SLICE + l;
            if (r >= N) break;
            int32_t c = m.iperm[r];
            wmax = std::max<int>(wmax, (int)(rptr[c + 1] - rptr[c]));
        }
        if (wmax > 255) throw Error{RT_E_UNSUPPORTED, "rt_mesh_create: a cell has more than 255 faces"};
        width[s] = wmax;
        off[s + 1] = off[s] + (int64_t)SLICE * wmax;
        m.maxWidth = std::max(m.maxWidth, wmax);
    }
    m.nEll = off[m.nSlices];
    const int64_t nE = std::max<int64_t>(1, m.nEll);
    std::vector<int32_t> ecol(nE), efid(nE);
    std::vector<uint8_t> eslot(nE, 0);
    std::vector<double> esf(3 * nE, 0.0), ew(nE, 0.0);
    std::vector<std::pair<int32_t, int>> byCol;  // (column, face-order index)
    for (int32_t s = 0; s < m.nSlices; ++s) {
        for (int l = 0; l < SLICE; ++l) {
            const int32_t r = s * SLICE + l;
            const int W = width[s];
            auto at = [&](int k) { return off[s] + (int64_t)SLICE * k + l; };
            if (r >= N) {
                for (int k = 0; k < W; ++k) {
                    ecol[at(k)] = 0;
                    efid[at(k)] = RT_FID_PAD;
                    eslot[at(k)] = (uint8_t)k;
                }
                continue;
            }
            const int32_t c = m.iperm[r];
            const int deg = (int)(rptr[c + 1] - rptr[c]);
            byCol.clear();
            for (int k = 0; k < deg; ++k) byCol.push_back({m.perm[rcol[rptr[c] + k]], k});
            std::sort(byCol.begin(), byCol.end());
            std::vector<int> slotOf(deg);
            for (int k = 0; k < deg; ++k) {
                ecol[at(k)] = byCol[k].first;
                slotOf[byCol[k].second] = k;
            }
            for (int k = 0; k < deg; ++k) {  // face order
                const int32_t fid = rface[rptr[c] + k];
                const bool own = fid >= 0;
                const int64_t f = own ? fid : -(int64_t)fid - 1;
                efid[at(k)] = fid;
                eslot[at(k)] = (uint8_t)slotOf[k];
                for (int q = 0; q < 3; ++q) esf[q * nE + at(k)] = own ? d->Sf[3 * f + q] : -d->Sf[3 * f + q];
                ew[at(k)] = d->weights[f];
            }
            for (int k = deg; k < W; ++k) {
                ecol[at(k)] = r;
                efid[at(k)] = RT_FID_PAD;
                eslot[at(k)] = (uint8_t)k;
            }
        }
    }
    // boundary faces per row (non-empty patches), ascending boundary index
    std::vector<int32_t> bpatch(std::max(1, NB), 0);
    std::vector<char> bEmpty(std::max(1, NB), 0);
    for (int32_t p = 0; p < m.nPatches; ++p)
        for (int32_t k = 0; k < m.patches[p].size; ++k) {
            bpatch[m.patches[p].start + k] = p;
            bEmpty[m.patches[p].start + k] = m.patches[p].kind == RT_PATCH_EMPTY;
        }
    std::vector<int32_t> bptr(N + 1, 0);
    for (int32_t b = 0; b < NB; ++b)
        if (!bEmpty[b]) bptr[m.perm[d->owner[F + b]] + 1]++;
    for (int32_t r = 0; r < N; ++r) bptr[r + 1] += bptr[r];
    std::vector<int32_t> bface(std::max(1, bptr[N]));
    {
        std::vector<int32_t> fill(bptr.begin(), bptr.end() - 1);
        for (int32_t b = 0; b < NB; ++b)
            if (!bEmpty[b]) bface[fill[m.perm[d->owner[F + b]]]++] = b;
    }
    std::vector<double> bsf(3 * (size_t)std::max(1, NB), 0.0);
    for (int32_t b = 0; b < NB; ++b)
        for (int q = 0; q < 3; ++q) bsf[(size_t)q * std::max(1, NB) + b] = d->Sf[3 * ((int64_t)F + b) + q];
    std::vector<double> V(N);
    for (int32_t r = 0; r < N; ++r) V[r] = d->V[m.iperm[r]];

    upload(ctx, m.ellOff, off);
    upload(ctx, m.ellWidth, width);
    upload(ctx, m.ellCol, ecol);
    upload(ctx, m.ellFid, efid);
    upload(ctx, m.ellSlot, eslot);
    upload(ctx, m.eSf, esf);
    upload(ctx, m.eW, ew);
    upload(ctx, m.bPtr, bptr);
    upload(ctx, m.bFace, bface);
    upload(ctx, m.bPatch, bpatch);
    upload(ctx, m.bSf, bsf);
    upload(ctx, m.V, V);
    if (d->C) {
        std::vector<double> c4(4 * (size_t)N, 0.0);
        for (int32_t r = 0; r < N; ++r)
            for (int q = 0; q < 3; ++q) c4[4 * (size_t)r + q] = d->C[3 * (size_t)m.iperm[r] + q];
        upload(ctx, m.C4, c4);
    }
    if (d->C && d->Cb && NB > 0) {
        // [OF9] fvPatch::delta() = n (n & (Cf - C_P)) on non-coupled patches;
        // the patch snGrad is deltaCoeffs (value - internal), deltaCoeffs =
        // 1/|delta| (surfaceInterpolation::makeDeltaCoeffs)
        std::vector<double> dcb(NB);
        for (int32_t b = 0; b < NB; ++b) {
            const double* sf = d->Sf + 3 * ((int64_t)F + b);
            const int32_t c = d->owner[F + b];
            const double ms = std::sqrt(sf[0] * sf[0] + sf[1] * sf[1] + sf[2] * sf[2]);
            const double n[3] = {sf[0] / ms, sf[1] / ms, sf[2] / ms};
            const double dd[3] = {d->Cb[3 * (size_t)b] - d->C[3 * (size_t)c], d->Cb[3 * (size_t)b + 1] - d->C[3 * (size_t)c + 1],
                                  d->Cb[3 * (size_t)b + 2] - d->C[3 * (size_t)c + 2]};
            const double nd = n[0] * dd[0] + n[1] * dd[1] + n[2] * dd[2];
            const double dl[3] = {n[0] * nd, n[1] * nd, n[2] * nd};
            dcb[b] = 1.0 / std::sqrt(dl[0] * dl[0] + dl[1] * dl[1] + dl[2] * dl[2]);
        }
        upload(ctx, m.bDc, dcb);
    }
    upload(ctx, m.permD, m.iperm);
    RT_HIP(hipStreamSynchronize(ctx->stream));
}

}  // namespace rt
