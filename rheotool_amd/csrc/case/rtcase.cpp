// rtcase.cpp -- synthetic stress-step cases: meshes and seeded fields.
// Host-only C++17.  See rtcase.h for the conventions.
#include "rtcase.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

namespace {

struct V3 {
    double x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double mag(V3 a) { return std::sqrt(dot(a, a)); }

// ---- counter-based random numbers -------------------------------------
inline uint64_t splitmix(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
inline uint64_t rkey(uint64_t seed, uint64_t stream, uint64_t id) {
    return splitmix(splitmix(splitmix(seed) ^ (stream * 0xD1B54A32D192ED03ull)) ^ id);
}
// k-th uniform in (0,1) of a key
inline double runif(uint64_t key, uint64_t k) {
    uint64_t v = splitmix(key + 0x632BE59BD9B4E019ull * (k + 1));
    return ((double)(v >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}
// k-th standard normal of a key (Box-Muller on uniforms 2k, 2k+1)
inline double rnorm(uint64_t key, uint64_t k) {
    double u1 = runif(key, 2 * k), u2 = runif(key, 2 * k + 1);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
}

// ---- OpenFOAM-style face geometry (triangle decomposition) -------------
void faceGeom(const V3* p, int n, V3& Cf, V3& Sf) {
    V3 fc{0, 0, 0};
    for (int i = 0; i < n; ++i) fc = fc + p[i];
    fc = (1.0 / n) * fc;
    V3 sumN{0, 0, 0}, sumAc{0, 0, 0};
    double sumA = 0;
    for (int i = 0; i < n; ++i) {
        const V3& a = p[i];
        const V3& b = p[(i + 1) % n];
        V3 c = a + b + fc;
        V3 nn = cross(b - a, fc - a);
        double ar = mag(nn);
        sumN = sumN + nn;
        sumA += ar;
        sumAc = sumAc + ar * c;
    }
    if (sumA < 1e-150) {
        Cf = fc;
    } else {
        Cf = (1.0 / 3.0) * ((1.0 / sumA) * sumAc);
    }
    Sf = 0.5 * sumN;
}

// ---- symmetric 3x3 eigen-decomposition (cyclic Jacobi) ------------------
void symEig(const double t[6], double ev[3], double R[9]) {
    double a[3][3] = {{t[0], t[1], t[2]}, {t[1], t[3], t[4]}, {t[2], t[4], t[5]}};
    double v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int sweep = 0; sweep < 50; ++sweep) {
        double off = std::fabs(a[0][1]) + std::fabs(a[0][2]) + std::fabs(a[1][2]);
        double dia = std::fabs(a[0][0]) + std::fabs(a[1][1]) + std::fabs(a[2][2]);
        if (off <= 1e-300 || off <= 1e-18 * dia) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                double apq = a[p][q];
                if (apq == 0.0) continue;
                double theta = (a[q][q] - a[p][p]) / (2.0 * apq);
                double tt = (theta >= 0 ? 1.0 : -1.0) /
                            (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                double c = 1.0 / std::sqrt(tt * tt + 1.0), s = tt * c;
                for (int k = 0; k < 3; ++k) {  // A <- A G
                    double akp = a[k][p], akq = a[k][q];
                    a[k][p] = c * akp - s * akq;
                    a[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < 3; ++k) {  // A <- G^T A
                    double apk = a[p][k], aqk = a[q][k];
                    a[p][k] = c * apk - s * aqk;
                    a[q][k] = s * apk + c * aqk;
                }
                a[p][q] = a[q][p] = 0.0;
                for (int k = 0; k < 3; ++k) {
                    double vkp = v[k][p], vkq = v[k][q];
                    v[k][p] = c * vkp - s * vkq;
                    v[k][q] = s * vkp + c * vkq;
                }
            }
    }
    int idx[3] = {0, 1, 2};
    std::sort(idx, idx + 3, [&](int i, int j) { return a[i][i] < a[j][j]; });
    for (int c = 0; c < 3; ++c) {
        ev[c] = a[idx[c]][idx[c]];
        for (int r = 0; r < 3; ++r) R[3 * r + c] = v[r][idx[c]];
    }
}

}  // namespace

struct rtc_mesh {
    int32_t nCells = 0, nInternal = 0, nBoundary = 0;
    std::vector<int32_t> owner, neighbour;
    std::vector<double> Sf, Cf, w, V, C;
    std::vector<int32_t> pStart, pSize, pKind;
    std::vector<std::string> pName;
    std::vector<int32_t> origId;
    // Cartesian metadata (for analytic boundary values)
    bool twoD = false;
};

namespace {

// finish a mesh whose owner/neighbour/Sf/Cf are set: cell centres, volumes, weights
void finishGeometry(rtc_mesh& m) {
    const int N = m.nCells, F = m.nInternal, NF = m.nInternal + m.nBoundary;
    std::vector<V3> cEst(N, V3{0, 0, 0});
    std::vector<int> nf(N, 0);
    auto Cf = [&](int f) { return V3{m.Cf[3 * f], m.Cf[3 * f + 1], m.Cf[3 * f + 2]}; };
    auto Sf = [&](int f) { return V3{m.Sf[3 * f], m.Sf[3 * f + 1], m.Sf[3 * f + 2]}; };
    for (int f = 0; f < NF; ++f) {
        cEst[m.owner[f]] = cEst[m.owner[f]] + Cf(f);
        nf[m.owner[f]]++;
    }
    for (int f = 0; f < F; ++f) {
        cEst[m.neighbour[f]] = cEst[m.neighbour[f]] + Cf(f);
        nf[m.neighbour[f]]++;
    }
    for (int c = 0; c < N; ++c) cEst[c] = (1.0 / nf[c]) * cEst[c];
    std::vector<V3> cc(N, V3{0, 0, 0});
    std::vector<double> vol(N, 0.0);
    for (int f = 0; f < NF; ++f) {
        int o = m.owner[f];
        double pyr3 = dot(Sf(f), Cf(f) - cEst[o]);
        V3 pc = (3.0 / 4.0) * Cf(f) + (1.0 / 4.0) * cEst[o];
        cc[o] = cc[o] + pyr3 * pc;
        vol[o] += pyr3;
    }
    for (int f = 0; f < F; ++f) {
        int n = m.neighbour[f];
        double pyr3 = dot(Sf(f), cEst[n] - Cf(f));
        V3 pc = (3.0 / 4.0) * Cf(f) + (1.0 / 4.0) * cEst[n];
        cc[n] = cc[n] + pyr3 * pc;
        vol[n] += pyr3;
    }
    m.C.resize(3 * N);
    m.V.resize(N);
    for (int c = 0; c < N; ++c) {
        V3 x = std::fabs(vol[c]) > 1e-300 ? (1.0 / vol[c]) * cc[c] : cEst[c];
        m.C[3 * c] = x.x;
        m.C[3 * c + 1] = x.y;
        m.C[3 * c + 2] = x.z;
        m.V[c] = vol[c] * (1.0 / 3.0);
    }
    m.w.resize(F);
    for (int f = 0; f < F; ++f) {
        V3 Co{m.C[3 * m.owner[f]], m.C[3 * m.owner[f] + 1], m.C[3 * m.owner[f] + 2]};
        V3 Cn{m.C[3 * m.neighbour[f]], m.C[3 * m.neighbour[f] + 1],
              m.C[3 * m.neighbour[f] + 2]};
        double sfo = std::fabs(dot(Sf(f), Cf(f) - Co));
        double sfn = std::fabs(dot(Sf(f), Cn - Cf(f)));
        m.w[f] = sfn / (sfo + sfn);
    }
}

}  // namespace

extern "C" {

rtc_mesh* rtc_mesh_cartesian(int nx, int ny, int nz, double Lx, double Ly, double Lz,
                             int twoD, int channel, double jitter, uint64_t jitterSeed,
                             uint64_t permSeed) {
    if (nx < 1 || ny < 1 || nz < 1 || (twoD && nz != 1)) return nullptr;
    auto* m = new rtc_mesh;
    m->twoD = twoD != 0;
    const int64_t N = (int64_t)nx * ny * nz;
    if (N > INT32_MAX / 2) {
        delete m;
        return nullptr;
    }
    m->nCells = (int32_t)N;
    const double dx = Lx / nx, dy = Ly / ny, dz = Lz / nz;
    const int px = nx + 1, py = ny + 1;
    auto pid = [&](int i, int j, int k) -> int64_t { return i + (int64_t)px * (j + (int64_t)py * k); };
    auto point = [&](int i, int j, int k) -> V3 {
        V3 p{i * dx, j * dy, k * dz};
        bool inX = i > 0 && i < nx, inY = j > 0 && j < ny, inZ = k > 0 && k < nz;
        if (jitter > 0 && inX && inY && (twoD || inZ)) {
            // 2D: both z layers share the (i,j) jitter so faces stay planar
            uint64_t key = rkey(jitterSeed, 11, (uint64_t)pid(i, j, twoD ? 0 : k));
            p.x += (2.0 * runif(key, 0) - 1.0) * jitter * dx;
            p.y += (2.0 * runif(key, 1) - 1.0) * jitter * dy;
            if (!twoD) p.z += (2.0 * runif(key, 2) - 1.0) * jitter * dz;
        }
        return p;
    };
    auto cid = [&](int i, int j, int k) -> int32_t { return (int32_t)(i + (int64_t)nx * (j + (int64_t)ny * k)); };

    // permutation orig -> new
    std::vector<int32_t> perm(N);
    for (int64_t c = 0; c < N; ++c) perm[c] = (int32_t)c;
    if (permSeed) {
        for (int64_t c = N - 1; c > 0; --c) {
            uint64_t r = splitmix(rkey(permSeed, 13, (uint64_t)c));
            int64_t j = (int64_t)(r % (uint64_t)(c + 1));
            std::swap(perm[c], perm[j]);
        }
    }
    m->origId.resize(N);
    for (int64_t c = 0; c < N; ++c) m->origId[perm[c]] = (int32_t)c;

    // ---- internal faces (in original numbering), then sorted ----------
    const int64_t Fx = (int64_t)(nx - 1) * ny * nz, Fy = (int64_t)nx * (ny - 1) * nz,
                  Fz = twoD ? 0 : (int64_t)nx * ny * (nz - 1);
    const int64_t F = Fx + Fy + Fz;
    m->nInternal = (int32_t)F;
    struct FaceRec {
        int32_t o, n;
        double Cf[3], Sf[3];
    };
    std::vector<FaceRec> faces;
    faces.reserve(F);
    auto addFace = [&](int32_t o, int32_t n, V3 p0, V3 p1, V3 p2, V3 p3) {
        V3 pts[4] = {p0, p1, p2, p3};
        V3 cf, sf;
        int32_t a = perm[o], b = perm[n];
        if (a > b) {  // flip so the normal leaves the new owner
            std::swap(a, b);
            std::swap(pts[1], pts[3]);
        }
        faceGeom(pts, 4, cf, sf);
        faces.push_back(FaceRec{a, b, {cf.x, cf.y, cf.z}, {sf.x, sf.y, sf.z}});
    };
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                int32_t c = cid(i, j, k);
                if (i < nx - 1)
                    addFace(c, cid(i + 1, j, k), point(i + 1, j, k), point(i + 1, j + 1, k),
                            point(i + 1, j + 1, k + 1), point(i + 1, j, k + 1));
                if (j < ny - 1)
                    addFace(c, cid(i, j + 1, k), point(i, j + 1, k), point(i, j + 1, k + 1),
                            point(i + 1, j + 1, k + 1), point(i + 1, j + 1, k));
                if (!twoD && k < nz - 1)
                    addFace(c, cid(i, j, k + 1), point(i, j, k + 1), point(i + 1, j, k + 1),
                            point(i + 1, j + 1, k + 1), point(i, j + 1, k + 1));
            }
    // counting sort by owner, then by neighbour within owner
    std::vector<int64_t> start(N + 1, 0);
    for (auto& f : faces) start[f.o + 1]++;
    for (int64_t c = 0; c < N; ++c) start[c + 1] += start[c];
    std::vector<int64_t> order(F);
    {
        std::vector<int64_t> fill(start.begin(), start.end() - 1);
        for (int64_t f = 0; f < F; ++f) order[fill[faces[f].o]++] = f;
    }
    for (int64_t c = 0; c < N; ++c)
        std::sort(order.begin() + start[c], order.begin() + start[c + 1],
                  [&](int64_t a, int64_t b) { return faces[a].n < faces[b].n; });

    // ---- boundary faces by patch --------------------------------------
    struct BRec {
        int32_t o;
        double Cf[3], Sf[3];
    };
    std::vector<std::vector<BRec>> patches;
    std::vector<std::string> names;
    std::vector<int32_t> kinds;
    auto bface = [&](std::vector<BRec>& P, int32_t o, V3 p0, V3 p1, V3 p2, V3 p3) {
        V3 pts[4] = {p0, p1, p2, p3}, cf, sf;
        faceGeom(pts, 4, cf, sf);
        P.push_back(BRec{perm[o], {cf.x, cf.y, cf.z}, {sf.x, sf.y, sf.z}});
    };
    std::vector<BRec> xmin, xmax, ymin, ymax, zmin, zmax;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j) {
            bface(xmin, cid(0, j, k), point(0, j, k), point(0, j, k + 1), point(0, j + 1, k + 1),
                  point(0, j + 1, k));
            bface(xmax, cid(nx - 1, j, k), point(nx, j, k), point(nx, j + 1, k),
                  point(nx, j + 1, k + 1), point(nx, j, k + 1));
        }
    for (int k = 0; k < nz; ++k)
        for (int i = 0; i < nx; ++i) {
            bface(ymin, cid(i, 0, k), point(i, 0, k), point(i + 1, 0, k), point(i + 1, 0, k + 1),
                  point(i, 0, k + 1));
            bface(ymax, cid(i, ny - 1, k), point(i, ny, k), point(i, ny, k + 1),
                  point(i + 1, ny, k + 1), point(i + 1, ny, k));
        }
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            bface(zmin, cid(i, j, 0), point(i, j, 0), point(i, j + 1, 0), point(i + 1, j + 1, 0),
                  point(i + 1, j, 0));
            bface(zmax, cid(i, j, nz - 1), point(i, j, nz), point(i + 1, j, nz),
                  point(i + 1, j + 1, nz), point(i, j + 1, nz));
        }
    auto cat = [](std::vector<BRec> a, const std::vector<BRec>& b) {
        a.insert(a.end(), b.begin(), b.end());
        return a;
    };
    if (channel) {
        patches.push_back(xmin); names.push_back("inlet"); kinds.push_back(RTC_PATCH);
        patches.push_back(xmax); names.push_back("outlet"); kinds.push_back(RTC_PATCH);
        patches.push_back(cat(ymin, ymax)); names.push_back("walls"); kinds.push_back(RTC_WALL);
        if (twoD) {
            patches.push_back(cat(zmin, zmax)); names.push_back("frontAndBack"); kinds.push_back(RTC_EMPTY);
        } else {
            patches.push_back(cat(zmin, zmax)); names.push_back("sides"); kinds.push_back(RTC_WALL);
        }
    } else {
        patches.push_back(xmin); names.push_back("xmin"); kinds.push_back(RTC_PATCH);
        patches.push_back(xmax); names.push_back("xmax"); kinds.push_back(RTC_PATCH);
        patches.push_back(ymin); names.push_back("ymin"); kinds.push_back(RTC_PATCH);
        patches.push_back(ymax); names.push_back("ymax"); kinds.push_back(RTC_PATCH);
        if (twoD) {
            patches.push_back(cat(zmin, zmax)); names.push_back("frontAndBack"); kinds.push_back(RTC_EMPTY);
        } else {
            patches.push_back(zmin); names.push_back("zmin"); kinds.push_back(RTC_PATCH);
            patches.push_back(zmax); names.push_back("zmax"); kinds.push_back(RTC_PATCH);
        }
    }
    int64_t NB = 0;
    for (auto& p : patches) NB += (int64_t)p.size();
    m->nBoundary = (int32_t)NB;
    const int64_t NF = F + NB;
    m->owner.resize(NF);
    m->neighbour.resize(F);
    m->Sf.resize(3 * NF);
    m->Cf.resize(3 * NF);
    for (int64_t f = 0; f < F; ++f) {
        const FaceRec& r = faces[order[f]];
        m->owner[f] = r.o;
        m->neighbour[f] = r.n;
        for (int d = 0; d < 3; ++d) {
            m->Sf[3 * f + d] = r.Sf[d];
            m->Cf[3 * f + d] = r.Cf[d];
        }
    }
    faces.clear();
    faces.shrink_to_fit();
    int64_t f = F;
    int32_t bstart = 0;
    for (size_t p = 0; p < patches.size(); ++p) {
        m->pStart.push_back(bstart);
        m->pSize.push_back((int32_t)patches[p].size());
        m->pKind.push_back(kinds[p]);
        m->pName.push_back(names[p]);
        bstart += (int32_t)patches[p].size();
        for (auto& r : patches[p]) {
            m->owner[f] = r.o;
            for (int d = 0; d < 3; ++d) {
                m->Sf[3 * f + d] = r.Sf[d];
                m->Cf[3 * f + d] = r.Cf[d];
            }
            ++f;
        }
    }
    finishGeometry(*m);
    return m;
}

rtc_mesh* rtc_mesh_random_graph(int nCells, int kNear, double sigma, uint64_t seed) {
    if (nCells < 2 || kNear < 1) return nullptr;
    auto* m = new rtc_mesh;
    const int N = nCells;
    m->nCells = N;
    std::vector<V3> x(N);
    for (int c = 0; c < N; ++c) {
        uint64_t key = rkey(seed, 21, (uint64_t)c);
        x[c] = V3{runif(key, 0), runif(key, 1), runif(key, 2)};
    }
    // bucket grid with ~2 cells per bucket
    int G = std::max(1, (int)std::cbrt(N / 2.0));
    auto bucketOf = [&](const V3& p) {
        int i = std::min(G - 1, (int)(p.x * G)), j = std::min(G - 1, (int)(p.y * G)),
            k = std::min(G - 1, (int)(p.z * G));
        return i + G * (j + G * k);
    };
    std::vector<int> bstart((size_t)G * G * G + 1, 0), bcell(N);
    for (int c = 0; c < N; ++c) bstart[bucketOf(x[c]) + 1]++;
    for (size_t b = 0; b + 1 < bstart.size(); ++b) bstart[b + 1] += bstart[b];
    {
        std::vector<int> fill(bstart.begin(), bstart.end() - 1);
        for (int c = 0; c < N; ++c) bcell[fill[bucketOf(x[c])]++] = c;
    }
    std::vector<uint64_t> edges;
    edges.reserve((size_t)N * kNear);
    std::vector<std::pair<double, int>> cand;
    for (int c = 0; c < N; ++c) {
        cand.clear();
        int bi = std::min(G - 1, (int)(x[c].x * G)), bj = std::min(G - 1, (int)(x[c].y * G)),
            bk = std::min(G - 1, (int)(x[c].z * G));
        for (int r = 1; r <= G && (int)cand.size() < kNear; ++r) {
            cand.clear();
            for (int k = std::max(0, bk - r); k <= std::min(G - 1, bk + r); ++k)
                for (int j = std::max(0, bj - r); j <= std::min(G - 1, bj + r); ++j)
                    for (int i = std::max(0, bi - r); i <= std::min(G - 1, bi + r); ++i) {
                        int b = i + G * (j + G * k);
                        for (int q = bstart[b]; q < bstart[b + 1]; ++q) {
                            int o = bcell[q];
                            if (o == c) continue;
                            V3 d = x[o] - x[c];
                            cand.push_back({dot(d, d), o});
                        }
                    }
        }
        int kk = std::min(kNear, (int)cand.size());
        std::partial_sort(cand.begin(), cand.begin() + kk, cand.end());
        for (int q = 0; q < kk; ++q) {
            uint32_t a = (uint32_t)std::min(c, cand[q].second),
                     b = (uint32_t)std::max(c, cand[q].second);
            edges.push_back(((uint64_t)a << 32) | b);
        }
    }
    std::sort(edges.begin(), edges.end());
    edges.erase(std::unique(edges.begin(), edges.end()), edges.end());
    const int64_t F = (int64_t)edges.size();
    m->nInternal = (int32_t)F;
    m->nBoundary = 0;
    m->owner.resize(F);
    m->neighbour.resize(F);
    m->Sf.resize(3 * F);
    m->Cf.resize(3 * F);
    m->w.assign(F, 0.5);
    const double A0 = std::pow((double)N, -2.0 / 3.0);
    for (int64_t f = 0; f < F; ++f) {
        int a = (int)(edges[f] >> 32), b = (int)(edges[f] & 0xffffffffu);
        m->owner[f] = a;
        m->neighbour[f] = b;
        V3 d = x[b] - x[a];
        double area = A0 * std::exp(sigma * rnorm(rkey(seed, 22, edges[f]), 0));
        V3 s = (area / mag(d)) * d;
        V3 mid = 0.5 * (x[a] + x[b]);
        m->Sf[3 * f] = s.x; m->Sf[3 * f + 1] = s.y; m->Sf[3 * f + 2] = s.z;
        m->Cf[3 * f] = mid.x; m->Cf[3 * f + 1] = mid.y; m->Cf[3 * f + 2] = mid.z;
    }
    m->V.assign(N, 1.0 / N);
    m->C.resize(3 * N);
    for (int c = 0; c < N; ++c) {
        m->C[3 * c] = x[c].x; m->C[3 * c + 1] = x[c].y; m->C[3 * c + 2] = x[c].z;
    }
    m->origId.resize(N);
    for (int c = 0; c < N; ++c) m->origId[c] = c;
    return m;
}

rtc_mesh* rtc_mesh_from_poly(int32_t nPoints, const double* points, int32_t nFaces,
                             const int32_t* faceStart, const int32_t* faceLabels,
                             const int32_t* owner, int32_t nInternal, const int32_t* neighbour,
                             int32_t nCells, int32_t nPatches, const int32_t* patchStart,
                             const int32_t* patchSize, const int32_t* patchKind,
                             const char* patchNames) {
    if (nPoints < 1 || nCells < 1 || nInternal < 0 || nFaces < nInternal || nPatches < 0 || !points ||
        !faceStart || !faceLabels || !owner || (nInternal && !neighbour))
        return nullptr;
    // patches tile [nInternal, nFaces)
    int32_t next = 0;
    for (int32_t p = 0; p < nPatches; ++p) {
        if (patchStart[p] != next || patchSize[p] < 0) return nullptr;
        next += patchSize[p];
    }
    if (next != nFaces - nInternal) return nullptr;
    for (int32_t f = 0; f < nFaces; ++f) {
        if (faceStart[f + 1] - faceStart[f] < 3 || owner[f] < 0 || owner[f] >= nCells) return nullptr;
        if (f < nInternal && (neighbour[f] <= owner[f] || neighbour[f] >= nCells)) return nullptr;
        for (int32_t k = faceStart[f]; k < faceStart[f + 1]; ++k)
            if (faceLabels[k] < 0 || faceLabels[k] >= nPoints) return nullptr;
    }
    auto* m = new rtc_mesh;
    m->nCells = nCells;
    m->nInternal = nInternal;
    m->nBoundary = nFaces - nInternal;
    m->owner.assign(owner, owner + nFaces);
    m->neighbour.assign(neighbour, neighbour + nInternal);
    m->Sf.resize(3 * (size_t)nFaces);
    m->Cf.resize(3 * (size_t)nFaces);
    std::vector<V3> pts;
    for (int32_t f = 0; f < nFaces; ++f) {
        pts.clear();
        for (int32_t k = faceStart[f]; k < faceStart[f + 1]; ++k) {
            const double* x = points + 3 * (size_t)faceLabels[k];
            pts.push_back(V3{x[0], x[1], x[2]});
        }
        V3 cf, sf;
        faceGeom(pts.data(), (int)pts.size(), cf, sf);
        m->Cf[3 * f] = cf.x, m->Cf[3 * f + 1] = cf.y, m->Cf[3 * f + 2] = cf.z;
        m->Sf[3 * f] = sf.x, m->Sf[3 * f + 1] = sf.y, m->Sf[3 * f + 2] = sf.z;
    }
    const char* nm = patchNames;
    for (int32_t p = 0; p < nPatches; ++p) {
        m->pStart.push_back(patchStart[p]);
        m->pSize.push_back(patchSize[p]);
        m->pKind.push_back(patchKind[p]);
        m->pName.push_back(nm ? std::string(nm) : "patch" + std::to_string(p));
        if (nm) nm += std::strlen(nm) + 1;
    }
    m->origId.resize(nCells);
    for (int32_t c = 0; c < nCells; ++c) m->origId[c] = c;
    finishGeometry(*m);
    return m;
}

void rtc_mesh_free(rtc_mesh* m) { delete m; }

void rtc_mesh_sizes(const rtc_mesh* m, int32_t out[4]) {
    out[0] = m->nCells;
    out[1] = m->nInternal;
    out[2] = m->nBoundary;
    out[3] = (int32_t)m->pStart.size();
}

void rtc_mesh_export(const rtc_mesh* m, int32_t* owner, int32_t* neighbour, double* Sf,
                     double* Cf, double* w, double* V, double* C, int32_t* patchStart,
                     int32_t* patchSize, int32_t* patchKind, int32_t* origId) {
    auto cp = [](auto* dst, const auto& v) {
        if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
    };
    cp(owner, m->owner);
    cp(neighbour, m->neighbour);
    cp(Sf, m->Sf);
    cp(Cf, m->Cf);
    cp(w, m->w);
    cp(V, m->V);
    cp(C, m->C);
    cp(patchStart, m->pStart);
    cp(patchSize, m->pSize);
    cp(patchKind, m->pKind);
    cp(origId, m->origId);
}

const char* rtc_mesh_patch_name(const rtc_mesh* m, int patch) {
    if (patch < 0 || patch >= (int)m->pName.size()) return "";
    return m->pName[patch].c_str();
}

void rtc_sym_eig(const double theta[6], double ev[3], double R[9]) { symEig(theta, ev, R); }

int rtc_fields(const rtc_mesh* m, int kind, uint64_t seed, double noiseAmp, double lnLo,
               double lnHi, double* U, double* Ub, double* phi, double* gradU,
               double* theta, double* eigVals, double* eigVecs) {
    const int N = m->nCells, F = m->nInternal, NB = m->nBoundary;
    const double k2p = 2.0 * M_PI;
    // global coefficients
    double a[4][4], pkl[4][4];
    for (int k = 0; k < 4; ++k)
        for (int l = 0; l < 4; ++l) {
            uint64_t key = rkey(seed, 4, (uint64_t)(4 * k + l));
            a[k][l] = rnorm(key, 0);
            pkl[k][l] = 2.0 * M_PI * runif(key, 2);
        }
    double ph[3];
    {
        uint64_t key = rkey(seed, 5, 0);
        for (int d = 0; d < 3; ++d) ph[d] = 2.0 * M_PI * runif(key, d);
    }
    // streamfunction derivatives at a point: psi_x, psi_y, psi_xx, psi_yy, psi_xy
    auto psiD = [&](double x, double y, double out[5]) {
        for (int q = 0; q < 5; ++q) out[q] = 0;
        for (int k = 0; k < 4; ++k)
            for (int l = 0; l < 4; ++l) {
                double K = M_PI * (k + 1), Lw = 2.0 * M_PI * (l + 1);
                double sy = std::sin(K * y), cy = std::cos(K * y);
                double sx = std::sin(Lw * x + pkl[k][l]), cx = std::cos(Lw * x + pkl[k][l]);
                out[0] += a[k][l] * sy * Lw * cx;
                out[1] += a[k][l] * K * cy * sx;
                out[2] += -a[k][l] * sy * Lw * Lw * sx;
                out[3] += -a[k][l] * K * K * sy * sx;
                out[4] += a[k][l] * K * cy * Lw * cx;
            }
    };
    double scale = 0.0;
    if (kind == RTC_FIELDS_CHANNEL) {
        double mx = 0;
        for (int c = 0; c < N; ++c) {
            double d[5];
            psiD(m->C[3 * c], m->C[3 * c + 1], d);
            mx = std::max(mx, std::sqrt(d[1] * d[1] + d[0] * d[0]));
        }
        scale = mx > 0 ? noiseAmp / mx : 0.0;
    }
    auto velocity = [&](double x, double y, double z, double u[3]) {
        if (kind == RTC_FIELDS_CHANNEL) {
            double d[5];
            psiD(x, y, d);
            u[0] = 1.5 * (1.0 - (2.0 * y - 1.0) * (2.0 * y - 1.0)) + scale * d[1];
            u[1] = -scale * d[0];
            u[2] = 0.0;
        } else {
            double X = k2p * x + ph[0], Y = k2p * y + ph[1], Z = k2p * z + ph[2];
            u[0] = std::sin(X) * std::cos(Y) * std::cos(Z) + (y - 0.5);
            u[1] = -std::cos(X) * std::sin(Y) * std::cos(Z);
            u[2] = 0.0;
        }
    };
    auto velGrad = [&](double x, double y, double z, double g[9]) {
        for (int q = 0; q < 9; ++q) g[q] = 0;
        if (kind == RTC_FIELDS_CHANNEL) {
            double d[5];
            psiD(x, y, d);
            g[0] = scale * d[4];                              // dx Ux
            g[1] = -scale * d[2];                             // dx Uy
            g[3] = -6.0 * (2.0 * y - 1.0) + scale * d[3];     // dy Ux
            g[4] = -scale * d[4];                             // dy Uy
        } else {
            double X = k2p * x + ph[0], Y = k2p * y + ph[1], Z = k2p * z + ph[2];
            double sX = std::sin(X), cX = std::cos(X), sY = std::sin(Y), cY = std::cos(Y),
                   sZ = std::sin(Z), cZ = std::cos(Z);
            g[0] = k2p * cX * cY * cZ;
            g[1] = k2p * sX * sY * cZ;
            g[3] = -k2p * sX * sY * cZ + 1.0;
            g[4] = -k2p * cX * cY * cZ;
            g[6] = -k2p * sX * cY * sZ;
            g[7] = k2p * cX * sY * sZ;
        }
    };
    std::vector<double> Uc;
    double* Uw = U;
    if (!Uw && phi) {
        Uc.resize(3 * (size_t)N);
        Uw = Uc.data();
    }
    if (Uw)
        for (int c = 0; c < N; ++c) velocity(m->C[3 * c], m->C[3 * c + 1], m->C[3 * c + 2], Uw + 3 * c);
    if (gradU)
        for (int c = 0; c < N; ++c) velGrad(m->C[3 * c], m->C[3 * c + 1], m->C[3 * c + 2], gradU + 9 * c);
    // boundary values (0 on empty patches)
    std::vector<double> ubLocal;
    double* Ubw = Ub;
    if (!Ubw && phi && NB) {
        ubLocal.resize(3 * (size_t)NB);
        Ubw = ubLocal.data();
    }
    if (Ubw) {
        for (size_t p = 0; p < m->pStart.size(); ++p)
            for (int i = 0; i < m->pSize[p]; ++i) {
                int b = m->pStart[p] + i;
                int64_t f = (int64_t)F + b;
                if (m->pKind[p] == RTC_EMPTY) {
                    Ubw[3 * b] = Ubw[3 * b + 1] = Ubw[3 * b + 2] = 0.0;
                } else {
                    velocity(m->Cf[3 * f], m->Cf[3 * f + 1], m->Cf[3 * f + 2], Ubw + 3 * b);
                }
            }
    }
    if (phi) {
        for (int f = 0; f < F; ++f) {
            const double* uo = Uw + 3 * m->owner[f];
            const double* un = Uw + 3 * m->neighbour[f];
            double wf = m->w[f], s = 0;
            for (int d = 0; d < 3; ++d) {
                double uf = wf * (uo[d] - un[d]) + un[d];
                s += uf * m->Sf[3 * f + d];
            }
            phi[f] = s;
        }
        for (int b = 0; b < NB; ++b) {
            int64_t f = (int64_t)F + b;
            phi[f] = Ubw[3 * b] * m->Sf[3 * f] + Ubw[3 * b + 1] * m->Sf[3 * f + 1] +
                     Ubw[3 * b + 2] * m->Sf[3 * f + 2];
        }
        for (size_t p = 0; p < m->pStart.size(); ++p)
            if (m->pKind[p] == RTC_EMPTY)
                for (int i = 0; i < m->pSize[p]; ++i) phi[(int64_t)F + m->pStart[p] + i] = 0.0;
    }
    if (theta || eigVals || eigVecs) {
        for (int c = 0; c < N; ++c) {
            const uint64_t id = (uint64_t)m->origId[c];
            double th[6], ev[3], R[9];
            if (kind == RTC_FIELDS_CHANNEL) {
                uint64_t key = rkey(seed, 1, id);
                double A[9];
                for (int q = 0; q < 9; ++q) A[q] = rnorm(key, q);
                double B[6];  // I + 0.1 A A^T
                const int ij[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
                for (int q = 0; q < 6; ++q) {
                    int i = ij[q][0], j = ij[q][1];
                    double s = 0;
                    for (int k = 0; k < 3; ++k) s += A[3 * i + k] * A[3 * j + k];
                    B[q] = (i == j ? 1.0 : 0.0) + 0.1 * s;
                }
                double mu[3], Q[9];
                symEig(B, mu, Q);
                for (int q = 0; q < 6; ++q) {
                    int i = ij[q][0], j = ij[q][1];
                    double s = 0;
                    for (int k = 0; k < 3; ++k) s += Q[3 * i + k] * std::log(mu[k]) * Q[3 * j + k];
                    th[q] = s;
                }
                symEig(th, ev, R);  // eigen state = decomposition of theta itself
            } else {
                uint64_t key = rkey(seed, 2, id);
                double l[3];
                for (int d = 0; d < 3; ++d) l[d] = lnLo + (lnHi - lnLo) * runif(key, d);
                uint64_t kq = rkey(seed, 3, id);
                double q0 = rnorm(kq, 0), q1 = rnorm(kq, 1), q2 = rnorm(kq, 2), q3 = rnorm(kq, 3);
                double qn = std::sqrt(q0 * q0 + q1 * q1 + q2 * q2 + q3 * q3);
                q0 /= qn; q1 /= qn; q2 /= qn; q3 /= qn;
                double Rq[9] = {1 - 2 * (q2 * q2 + q3 * q3), 2 * (q1 * q2 - q0 * q3), 2 * (q1 * q3 + q0 * q2),
                                2 * (q1 * q2 + q0 * q3), 1 - 2 * (q1 * q1 + q3 * q3), 2 * (q2 * q3 - q0 * q1),
                                2 * (q1 * q3 - q0 * q2), 2 * (q2 * q3 + q0 * q1), 1 - 2 * (q1 * q1 + q2 * q2)};
                const int ij[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
                for (int q = 0; q < 6; ++q) {
                    int i = ij[q][0], j = ij[q][1];
                    double s = 0;
                    for (int k = 0; k < 3; ++k) s += Rq[3 * i + k] * l[k] * Rq[3 * j + k];
                    th[q] = s;
                }
                int idx[3] = {0, 1, 2};
                std::sort(idx, idx + 3, [&](int i, int j) { return l[i] < l[j]; });
                for (int cc = 0; cc < 3; ++cc) {
                    ev[cc] = l[idx[cc]];
                    for (int r = 0; r < 3; ++r) R[3 * r + cc] = Rq[3 * r + idx[cc]];
                }
            }
            if (theta)
                for (int q = 0; q < 6; ++q) theta[6 * (size_t)c + q] = th[q];
            if (eigVals) {
                double* e = eigVals + 9 * (size_t)c;
                for (int q = 0; q < 9; ++q) e[q] = 0.0;
                e[0] = std::exp(ev[0]);
                e[4] = std::exp(ev[1]);
                e[8] = std::exp(ev[2]);
            }
            if (eigVecs)
                for (int q = 0; q < 9; ++q) eigVecs[9 * (size_t)c + q] = R[q];
        }
    }
    return 0;
}

}  // extern "C"
