// Compressible neo-Hookean energy of a linear tetrahedron, shared by the host workload generator
// and the device element kernel:
//   psi(F) = mu/2 (tr(F^T F) - 3) - mu ln J + lambda/2 (ln J)^2,   E_e = V_e psi(F_e),
// F = Ds Dm^{-1}, gradients dE/dx_a = V (mu F g_a + (lambda ln J - mu) F^{-T} g_a) and Hessian
// K_{ai,bj} = V [mu d_ij (g_a . g_b) + (mu - lambda ln J) p_b[i] p_a[j] + lambda p_a[i] p_b[j]],
// p_a = F^{-T} g_a, g_a the rest-frame shape-function gradients.  The reference supplies its
// element Hessians through a caller callback (SystemAssembler.hh:559-565) and ships no energy.
#pragma once
#include <cmath>

#if defined(__HIPCC__)
#define MFS_HD __host__ __device__
#else
#define MFS_HD
#endif

namespace mfs {

MFS_HD inline double det3(const double* A) {         // column-major 3x3
    return A[0] * (A[4] * A[8] - A[7] * A[5]) - A[3] * (A[1] * A[8] - A[7] * A[2]) +
           A[6] * (A[1] * A[5] - A[4] * A[2]);
}
MFS_HD inline void inv3(const double* A, double* R, double det) {
    const double id = 1.0 / det;
    R[0] = (A[4] * A[8] - A[5] * A[7]) * id;
    R[1] = (A[2] * A[7] - A[1] * A[8]) * id;
    R[2] = (A[1] * A[5] - A[2] * A[4]) * id;
    R[3] = (A[5] * A[6] - A[3] * A[8]) * id;
    R[4] = (A[0] * A[8] - A[2] * A[6]) * id;
    R[5] = (A[2] * A[3] - A[0] * A[5]) * id;
    R[6] = (A[3] * A[7] - A[4] * A[6]) * id;
    R[7] = (A[1] * A[6] - A[0] * A[7]) * id;
    R[8] = (A[0] * A[4] - A[1] * A[3]) * id;
}

struct NHState {
    double V, mu, lam, lnJ;
    double F[9];
    double g[4][3], p[4][3];
};

// xr[a], x[a]: rest / current positions of the element's 4 nodes (3 doubles each)
MFS_HD inline void nh_state(const double (&xr)[4][3], const double (&x)[4][3], double E, double nu, NHState& s) {
    double Dm[9], Ds[9], Dmi[9];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) {
            Dm[3 * c + r] = xr[c + 1][r] - xr[0][r];
            Ds[3 * c + r] = x[c + 1][r] - x[0][r];
        }
    const double detDm = det3(Dm);
    inv3(Dm, Dmi, detDm);
    s.V = detDm / 6.0;
    for (int a = 1; a < 4; ++a)
        for (int k = 0; k < 3; ++k) s.g[a][k] = Dmi[3 * k + (a - 1)];
    for (int k = 0; k < 3; ++k) s.g[0][k] = -(s.g[1][k] + s.g[2][k] + s.g[3][k]);
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) {
            double v = 0;
            for (int k = 0; k < 3; ++k) v += Ds[3 * k + r] * Dmi[3 * c + k];
            s.F[3 * c + r] = v;
        }
    const double J = det3(s.F);
    double Fi[9];
    inv3(s.F, Fi, J);
    s.mu = E / (2.0 * (1.0 + nu));
    s.lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    s.lnJ = std::log(J);
    for (int a = 0; a < 4; ++a)
        for (int i = 0; i < 3; ++i) {
            double v = 0;
            for (int k = 0; k < 3; ++k) v += Fi[3 * i + k] * s.g[a][k];
            s.p[a][i] = v;
        }
}

MFS_HD inline double nh_energy(const NHState& s) {
    double I1 = 0;
    for (int q = 0; q < 9; ++q) I1 += s.F[q] * s.F[q];
    return s.V * (0.5 * s.mu * (I1 - 3.0) - s.mu * s.lnJ + 0.5 * s.lam * s.lnJ * s.lnJ);
}

// dE/dx_{a,i}
MFS_HD inline double nh_grad(const NHState& s, int a, int i) {
    double Fg = 0;
    for (int k = 0; k < 3; ++k) Fg += s.F[3 * k + i] * s.g[a][k];
    return s.V * (s.mu * Fg + (s.lam * s.lnJ - s.mu) * s.p[a][i]);
}

// d^2E / dx_{a,i} dx_{b,j}
MFS_HD inline double nh_hess(const NHState& s, int a, int i, int b, int j) {
    const double c2 = s.mu - s.lam * s.lnJ;
    double v = c2 * s.p[b][i] * s.p[a][j] + s.lam * s.p[a][i] * s.p[b][j];
    if (i == j) v += s.mu * (s.g[a][0] * s.g[b][0] + s.g[a][1] * s.g[b][1] + s.g[a][2] * s.g[b][2]);
    return s.V * v;
}

} // namespace mfs
