// Synthetic Newton-step workload: Kuhn-split cube mesh, seeded Gaussian node perturbation and the
// compressible neo-Hookean element gradient/Hessian (linear tets, neohookean.h).  The reference
// ships no energy code (its caller supplies eval_He, SystemAssembler.hh:559-565).
#include "mfs_internal.h"
#include "neohookean.h"

#include <algorithm>
#include <cmath>
#include <thread>

namespace mfs {

namespace {
// Axis permutations of the Kuhn split; tet = {0, e_p0, e_p0 + e_p1, (1,1,1)}, sign = parity.
const int kPerm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
const int kSign[6] = {+1, -1, -1, +1, +1, -1};

inline uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
inline double uniform_open0(uint64_t& s) {   // (0, 1]
    return (double)((splitmix64(s) >> 11) + 1) * (1.0 / 9007199254740992.0);
}

void element(const double* Xr, const double* X, const int32_t* t, double E, double nu,
             double* He, double* g12, double* energy) {
    double xr[4][3], x[4][3];
    for (int a = 0; a < 4; ++a)
        for (int r = 0; r < 3; ++r) {
            xr[a][r] = Xr[3 * t[a] + r];
            x[a][r] = X[3 * t[a] + r];
        }
    NHState st;
    nh_state(xr, x, E, nu, st);
    if (energy) *energy = nh_energy(st);
    if (g12)
        for (int a = 0; a < 4; ++a)
            for (int i = 0; i < 3; ++i) g12[3 * a + i] = nh_grad(st, a, i);
    if (He)
        for (int b = 0; b < 4; ++b)
            for (int j = 0; j < 3; ++j)
                for (int a = 0; a < 4; ++a)
                    for (int i = 0; i < 3; ++i) He[(3 * b + j) * 12 + 3 * a + i] = nh_hess(st, a, i, b, j);
}
} // namespace

void cube_mesh(int N, int32_t* tets, double* X) {
    const int64_t Np = N + 1;
    auto nid = [&](int64_t x, int64_t y, int64_t z) { return (int32_t)(x + Np * (y + Np * z)); };
    if (X) {
        const double h = 1.0 / N;
        for (int64_t z = 0; z <= N; ++z)
            for (int64_t y = 0; y <= N; ++y)
                for (int64_t x = 0; x <= N; ++x) {
                    double* p = X + 3 * nid(x, y, z);
                    p[0] = x * h; p[1] = y * h; p[2] = z * h;
                }
    }
    if (!tets) return;
    for (int64_t z = 0; z < N; ++z)
        for (int64_t y = 0; y < N; ++y)
            for (int64_t x = 0; x < N; ++x) {
                const int64_t hex = x + N * (y + N * z);
                for (int q = 0; q < 6; ++q) {
                    int c[3] = {0, 0, 0};
                    int v[4];
                    v[0] = nid(x, y, z);
                    c[kPerm[q][0]] = 1;
                    v[1] = nid(x + c[0], y + c[1], z + c[2]);
                    c[kPerm[q][1]] = 1;
                    v[2] = nid(x + c[0], y + c[1], z + c[2]);
                    v[3] = nid(x + 1, y + 1, z + 1);
                    if (kSign[q] < 0) std::swap(v[2], v[3]);
                    int32_t* t = tets + 4 * (6 * hex + q);
                    for (int k = 0; k < 4; ++k) t[k] = v[k];
                }
            }
}

void perturb_nodes(int N, uint64_t seed, double sigma_rel, double* X) {
    const int64_t nn = (int64_t)(N + 1) * (N + 1) * (N + 1);
    const double s = sigma_rel / N;
    uint64_t st = seed;
    const int64_t m = 3 * nn;
    for (int64_t q = 0; q < m; q += 2) {
        const double u1 = uniform_open0(st), u2 = uniform_open0(st);
        const double r = std::sqrt(-2.0 * std::log(u1));
        const double th = 6.283185307179586 * u2;
        X[q] += s * r * std::cos(th);
        if (q + 1 < m) X[q + 1] += s * r * std::sin(th);
    }
}

void neohookean_host(const double* Xr, const double* X, const int32_t* tets, int64_t ne,
                     const double* Ee, const double* nue, double E, double nu,
                     double* He, double* grad, double* energy, int nthreads) {
    if (nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
    nthreads = (int)std::min<int64_t>(nthreads, std::max<int64_t>(1, ne / 1024));
    // gradient is assembled serially (element order) from per-element vectors to stay
    // deterministic; Hessians are independent per element.
    std::vector<double> g12;
    if (grad) g12.resize((size_t)ne * 12);
    std::vector<double> en;
    if (energy) en.resize(ne);
    auto work = [&](int64_t e0, int64_t e1) {
        for (int64_t e = e0; e < e1; ++e)
            element(Xr, X, tets + 4 * e, Ee ? Ee[e] : E, nue ? nue[e] : nu,
                    He ? He + 144 * e : nullptr, grad ? g12.data() + 12 * e : nullptr,
                    energy ? &en[e] : nullptr);
    };
    std::vector<std::thread> th;
    const int64_t chunk = (ne + nthreads - 1) / nthreads;
    for (int t = 0; t < nthreads; ++t) {
        const int64_t a = t * chunk, b = std::min(ne, a + chunk);
        if (a < b) th.emplace_back(work, a, b);
    }
    for (auto& x : th) x.join();
    if (grad) {
        for (int64_t e = 0; e < ne; ++e)
            for (int a = 0; a < 4; ++a)
                for (int i = 0; i < 3; ++i) grad[3 * tets[4 * e + a] + i] += g12[12 * e + 3 * a + i];
    }
    if (energy) {
        double s = 0;
        for (int64_t e = 0; e < ne; ++e) s += en[e];
        *energy = s;
    }
}

} // namespace mfs
