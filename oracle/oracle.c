/*
 * oracle.c -- CPU restatement of the reference's per-Newton-step linear-solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg load this library, and only as the checker / the timed CPU baseline; the product path
 * (meshfemsparse_amd) never calls it.
 *
 * Functions and what pins each one (tests/test_oracle*.py):
 *   orc_cube_mesh / orc_perturb   Kuhn 6-tet cube and splitmix64+Box-Muller perturbation
 *                                 (workload definition; pinned by volume sum = 1, positive
 *                                 orientation, node-id formula, Gaussian moments).
 *   orc_neohookean_energy         psi = mu/2 (I1-3) - mu lnJ + lam/2 lnJ^2 per tet (pinned by
 *                                 closed forms: zero at rest, rigid-rotation invariance,
 *                                 uniform-stretch energy; gradient/Hessian of the product are
 *                                 then pinned by finite differences of this energy).
 *   orc_pattern                   SystemAssembler<3>::blockSparsityPattern
 *                                 (src/lib/MeshFEMSparse/SystemAssembler.hh:251-465); pinned by
 *                                 brute-force pair sets on tiny meshes.
 *   orc_assemble                  ElementHessianContribAssembler<true>::run, serial element
 *                                 order (SystemAssembler.hh:60-97; ColumnScanner
 *                                 BlockCSCHessian.hh:323-381); pinned by dense sum_e P_e^T He P_e.
 *   orc_row_col_removal           CSCMatrix::rowColRemoval (SparseMatrices.hh:2020-2076) on the
 *                                 block pattern; pinned by index bookkeeping identities.
 *   orc_spmv                      BlockCSCHessian::applyRawColumnRangeAccum
 *                                 (BlockCSCHessian.hh:1085-1120); pinned by the dense product.
 *   orc_permuted_lower            P (A_red + sigma I or sigma B) P^T as scalar lower CSC (value
 *                                 injection, CatamariConversionPlan.hh:402-541 semantics).
 *   orc_chol                      up-looking sparse Cholesky (elimination-tree row patterns); the
 *                                 factor of a fixed ordering is unique, pinned against LAPACK
 *                                 potrf of the dense matrix on small cases.
 *   orc_lsolve / orc_ltsolve      triangular solves with that factor (pinned by dense solves).
 *   orc_multiply_lower_block      trs_kernels::MultiplyLowerBlock (BlockCatamari, external;
 *                                 semantics from tests/test_catamari_solve_kernels.cc:33-113);
 *                                 pinned by the reference's known-answer test.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------------------------------- */
/* workload                                                                                   */
/* ---------------------------------------------------------------------------------------- */
void orc_cube_mesh(int N, int32_t* tets, double* X) {
    const int64_t M = N + 1;
    const double h = 1.0 / N;
    if (X)
        for (int64_t z = 0; z <= N; ++z)
            for (int64_t y = 0; y <= N; ++y)
                for (int64_t x = 0; x <= N; ++x) {
                    const int64_t id = x + M * (y + M * z);
                    X[3 * id] = x * h;
                    X[3 * id + 1] = y * h;
                    X[3 * id + 2] = z * h;
                }
    if (!tets) return;
    /* permutations of the axes, in lexicographic order, and their parity */
    static const int P[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    static const int odd[6] = {0, 1, 1, 0, 0, 1};
    int64_t t = 0;
    for (int64_t z = 0; z < N; ++z)
        for (int64_t y = 0; y < N; ++y)
            for (int64_t x = 0; x < N; ++x)
                for (int q = 0; q < 6; ++q) {
                    int64_t c[3] = {x, y, z};
                    int32_t v[4];
                    v[0] = (int32_t)(c[0] + M * (c[1] + M * c[2]));
                    for (int s = 0; s < 2; ++s) {
                        c[P[q][s]] += 1;
                        v[1 + s] = (int32_t)(c[0] + M * (c[1] + M * c[2]));
                    }
                    v[3] = (int32_t)((x + 1) + M * ((y + 1) + M * (z + 1)));
                    if (odd[q]) { int32_t tmp = v[2]; v[2] = v[3]; v[3] = tmp; }
                    memcpy(tets + 4 * t, v, sizeof v);
                    ++t;
                }
}

static uint64_t sm64(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void orc_perturb(int N, uint64_t seed, double sigma_rel, double* X) {
    const int64_t m = 3 * (int64_t)(N + 1) * (N + 1) * (N + 1);
    const double s = sigma_rel / N;
    uint64_t st = seed;
    for (int64_t q = 0; q < m; q += 2) {
        const double u1 = (double)((sm64(&st) >> 11) + 1) / 9007199254740992.0;
        const double u2 = (double)((sm64(&st) >> 11) + 1) / 9007199254740992.0;
        const double r = sqrt(-2.0 * log(u1)), th = 6.283185307179586 * u2;
        X[q] += s * r * cos(th);
        if (q + 1 < m) X[q + 1] += s * r * sin(th);
    }
}

static double det3(const double a[3][3]) {   /* a[row][col] */
    return a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
           a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
           a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
}

double orc_neohookean_energy(const double* Xr, const double* X, const int32_t* tets, int64_t ne,
                             const double* Ee, const double* nue, double E0, double nu0) {
    double total = 0.0;
    for (int64_t e = 0; e < ne; ++e) {
        const int32_t* t = tets + 4 * e;
        double Dm[3][3], Ds[3][3], Dmi[3][3], F[3][3];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                Dm[r][c] = Xr[3 * t[c + 1] + r] - Xr[3 * t[0] + r];
                Ds[r][c] = X[3 * t[c + 1] + r] - X[3 * t[0] + r];
            }
        const double dm = det3(Dm);
        /* adjugate / det */
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                const int r1 = (c + 1) % 3, r2 = (c + 2) % 3, c1 = (r + 1) % 3, c2 = (r + 2) % 3;
                Dmi[r][c] = (Dm[r1][c1] * Dm[r2][c2] - Dm[r1][c2] * Dm[r2][c1]) / dm;
            }
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                double s = 0;
                for (int k = 0; k < 3; ++k) s += Ds[r][k] * Dmi[k][c];
                F[r][c] = s;
            }
        const double E = Ee ? Ee[e] : E0, nu = nue ? nue[e] : nu0;
        const double mu = E / (2 * (1 + nu)), lam = E * nu / ((1 + nu) * (1 - 2 * nu));
        const double J = det3(F), lnJ = log(J);
        double I1 = 0;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) I1 += F[r][c] * F[r][c];
        total += dm / 6.0 * (0.5 * mu * (I1 - 3) - mu * lnJ + 0.5 * lam * lnJ * lnJ);
    }
    return total;
}

/* ---------------------------------------------------------------------------------------- */
/* block sparsity pattern (SystemAssembler.hh:251-465)                                       */
/* ---------------------------------------------------------------------------------------- */
static int cmp_u32(const void* a, const void* b) {
    const uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
    return (x > y) - (x < y);
}
static void sort_small(int32_t* v, int n) {
    for (int i = 1; i < n; ++i)
        for (int j = i; j > 0 && v[j - 1] > v[j]; --j) { int32_t t = v[j]; v[j] = v[j - 1]; v[j - 1] = t; }
}

/* Returns nz; Ap[nb+1], Ai must hold `cap` entries (returns -nz if cap is too small). */
int64_t orc_pattern(const int32_t* en, int64_t ne, int npe, int64_t nb, int64_t* Ap, int64_t* Ai,
                    int64_t cap) {
    int64_t* bucketStart = calloc(nb + 1, sizeof(int64_t));
    int32_t bv[16];
    for (int64_t e = 0; e < ne; ++e) {           /* sizes[bvars[i]] += i + 1 on sorted stencils */
        memcpy(bv, en + npe * e, npe * sizeof(int32_t));
        sort_small(bv, npe);
        for (int i = 0; i < npe; ++i) bucketStart[bv[i] + 1] += i + 1;
    }
    int64_t orig = 0;
    for (int64_t j = 1; j <= nb; ++j) { int64_t s = bucketStart[j]; bucketStart[j] = orig; orig += s; }
    uint32_t* buckets = malloc((orig ? orig : 1) * sizeof(uint32_t));
    int64_t* back = bucketStart + 1;
    for (int64_t e = 0; e < ne; ++e) {
        memcpy(bv, en + npe * e, npe * sizeof(int32_t));
        sort_small(bv, npe);
        for (int i = 0; i < npe; ++i)
            for (int k = 0; k <= i; ++k) buckets[back[bv[i]]++] = (uint32_t)bv[k];
    }
    /* bucketStart[j] .. bucketStart[j+1] now spans column j's bucket */
    int64_t* flags = malloc(nb * sizeof(int64_t));
    for (int64_t i = 0; i < nb; ++i) flags[i] = -1;
    int64_t* cnt = malloc((nb + 1) * sizeof(int64_t));
    for (int64_t j = 0; j < nb; ++j) {
        uint32_t *start = buckets + (j ? bucketStart[j] : 0), *end = buckets + bucketStart[j + 1];
        uint32_t* w = start;
        for (uint32_t* c = start; c < end; ++c)
            if (flags[*c] != j) { flags[*c] = j; *w++ = *c; }
        qsort(start, (size_t)(w - start), sizeof(uint32_t), cmp_u32);
        cnt[j] = w - start;
    }
    int64_t nz = 0;
    for (int64_t j = 0; j < nb; ++j) nz += cnt[j];
    if (nz > cap) { free(bucketStart); free(buckets); free(flags); free(cnt); return -nz; }
    Ap[0] = 0;
    for (int64_t j = 0; j < nb; ++j) {
        Ap[j + 1] = Ap[j] + cnt[j];
        const uint32_t* start = buckets + (j ? bucketStart[j] : 0);
        for (int64_t k = 0; k < cnt[j]; ++k) Ai[Ap[j] + k] = start[k];
    }
    free(bucketStart); free(buckets); free(flags); free(cnt);
    return nz;
}

/* ---------------------------------------------------------------------------------------- */
/* Hessian assembly, reference semantics (SystemAssembler.hh:60-97), serial element order     */
/* He: 144 doubles per element (12x12 column-major).  Ax accumulates (9 doubles per block).   */
/* ---------------------------------------------------------------------------------------- */
void orc_assemble(const int64_t* Ap, const int64_t* Ai, int64_t nb, const int32_t* en, int64_t ne,
                  const double* He, double* Ax) {
    (void)nb;
    for (int64_t e = 0; e < ne; ++e) {
        const int32_t* bv = en + 4 * e;
        const double* H = He + 144 * e;
        int order[4] = {0, 1, 2, 3};               /* argsort(blockVars) */
        for (int i = 1; i < 4; ++i)
            for (int j = i; j > 0 && bv[order[j - 1]] > bv[order[j]]; --j) {
                int t = order[j]; order[j] = order[j - 1]; order[j - 1] = t;
            }
        for (int lbj_i = 0; lbj_i < 4; ++lbj_i) {
            const int lbj = order[lbj_i];
            const int64_t bj = bv[lbj];
            int64_t bloc = Ap[bj];                   /* ColumnScanner over column bj */
            int64_t sloc = 9 * Ap[bj];
            for (int lbi_i = 0; lbi_i < lbj_i; ++lbi_i) {
                const int lbi = order[lbi_i];
                const int64_t bi = bv[lbi];
                const int64_t old = bloc;             /* advanceToBlock: linear scan */
                while (Ai[bloc] < bi) ++bloc;
                sloc += 9 * (bloc - old);
                for (int c = 0; c < 3; ++c)
                    for (int r = 0; r < 3; ++r) {
                        double v;
                        if (lbi < lbj) v = H[(3 * lbj + c) * 12 + 3 * lbi + r];   /* He(lbi,lbj) */
                        else           v = H[(3 * lbi + r) * 12 + 3 * lbj + c];   /* He(lbj,lbi)^T */
                        Ax[sloc + 3 * c + r] += v;
                    }
            }
            /* upper triangle of the diagonal block: last block of column bj */
            const int64_t dloc = 9 * Ap[bj + 1] - 9;
            for (int c = 0; c < 3; ++c)
                for (int r = 0; r <= c; ++r) Ax[dloc + 3 * c + r] += H[(3 * lbj + c) * 12 + 3 * lbj + r];
        }
    }
}

/* ---------------------------------------------------------------------------------------- */
/* rowColRemoval on the block pattern (SparseMatrices.hh:2020-2076)                          */
/* removed[b] != 0 marks removed blocks.  Outputs reduced Ap/Ai (in place copies), the        */
/* reducedRowForRow map and entryForReducedEntry.  Returns the reduced nz.                   */
/* ---------------------------------------------------------------------------------------- */
int64_t orc_row_col_removal(const int64_t* Ap, const int64_t* Ai, int64_t n, const uint8_t* removed,
                            int64_t* Rp, int64_t* Ri, int64_t* reducedRowForRow,
                            int64_t* entryForReducedEntry) {
    int64_t ridx = 0;
    for (int64_t i = 0; i < n; ++i) reducedRowForRow[i] = removed[i] ? -1 : ridx++;
    int64_t back = 0, cb = 0;
    Rp[0] = 0;
    for (int64_t j = 0; j < n; ++j) {
        if (removed[j]) continue;
        for (int64_t k = Ap[j]; k < Ap[j + 1]; ++k) {
            const int64_t i = Ai[k];
            if (removed[i]) continue;
            Ri[back] = reducedRowForRow[i];
            entryForReducedEntry[back] = k;
            ++back;
        }
        Rp[++cb] = back;
    }
    return back;
}

/* ---------------------------------------------------------------------------------------- */
/* symmetric block SpMV from upper blocks (BlockCSCHessian.hh:1085-1120)                     */
/* ---------------------------------------------------------------------------------------- */
void orc_spmv(const int64_t* Ap, const int64_t* Ai, int64_t nb, const double* Ax, const double* x,
              double* y) {
    memset(y, 0, sizeof(double) * 3 * nb);
    for (int64_t bj = 0; bj < nb; ++bj) {
        const double* xj = x + 3 * bj;
        for (int64_t k = Ap[bj]; k < Ap[bj + 1]; ++k) {
            const int64_t bi = Ai[k];
            const double* B = Ax + 9 * k;
            if (bi < bj) {
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) {
                        y[3 * bi + r] += B[3 * c + r] * xj[c];
                        y[3 * bj + c] += B[3 * c + r] * x[3 * bi + r];
                    }
            } else {
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) {
                        const double v = r <= c ? B[3 * c + r] : B[3 * r + c];
                        y[3 * bi + r] += v * xj[c];
                    }
            }
        }
    }
}

/* ---------------------------------------------------------------------------------------- */
/* P (A_red + shift) P^T as scalar lower CSC.                                                 */
/*   red_of_full[b]: reduced block or -1; pos[r]: factor position of reduced block r.         */
/*   Lower CSC: column j holds rows i >= j, sorted.  Returns nnz (or -nnz if cap too small).  */
/* ---------------------------------------------------------------------------------------- */
int64_t orc_permuted_lower(const int64_t* Ap, const int64_t* Ai, int64_t nb, const double* Ax,
                           const int64_t* red_of_full, const int32_t* pos, int64_t nred,
                           double sigma, const double* Bx, int64_t* Cp, int64_t* Ci, double* Cx,
                           int64_t cap) {
    const int64_t n = 3 * nred;
    int64_t* cnt = calloc(n + 1, sizeof(int64_t));
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1) {
            Cp[0] = 0;
            for (int64_t j = 0; j < n; ++j) Cp[j + 1] = Cp[j] + cnt[j];
            if (Cp[n] > cap) { int64_t r = -Cp[n]; free(cnt); return r; }
            memcpy(cnt, Cp, n * sizeof(int64_t));
        }
        for (int64_t bj = 0; bj < nb; ++bj) {
            if (red_of_full[bj] < 0) continue;
            for (int64_t k = Ap[bj]; k < Ap[bj + 1]; ++k) {
                const int64_t bi = Ai[k];
                if (red_of_full[bi] < 0) continue;
                const int64_t a = pos[red_of_full[bi]], b = pos[red_of_full[bj]];
                for (int c = 0; c < 3; ++c)
                    for (int r = 0; r < 3; ++r) {
                        /* stored entry A(3bi + r, 3bj + c) */
                        if (bi == bj && r > c) continue;          /* diagonal: upper only */
                        const int64_t gi = 3 * a + r, gj = 3 * b + c;
                        const int64_t row = gi > gj ? gi : gj, col = gi > gj ? gj : gi;
                        double v = Ax[9 * k + 3 * c + r];
                        if (Bx) v += sigma * Bx[9 * k + 3 * c + r];
                        else if (bi == bj && r == c) v += sigma;
                        if (pass == 0) cnt[col]++;
                        else { Ci[cnt[col]] = row; Cx[cnt[col]] = v; cnt[col]++; }
                    }
            }
        }
    }
    /* sort each column by row (insertion sort; columns are short) */
    for (int64_t j = 0; j < n; ++j)
        for (int64_t p = Cp[j] + 1; p < Cp[j + 1]; ++p)
            for (int64_t q = p; q > Cp[j] && Ci[q - 1] > Ci[q]; --q) {
                int64_t ti = Ci[q]; Ci[q] = Ci[q - 1]; Ci[q - 1] = ti;
                double tx = Cx[q]; Cx[q] = Cx[q - 1]; Cx[q - 1] = tx;
            }
    int64_t nz = Cp[n];
    free(cnt);
    return nz;
}

/* ---------------------------------------------------------------------------------------- */
/* Up-looking sparse Cholesky of a lower CSC matrix C (rows >= col, sorted).                 */
/* Row k of L is found by a sparse triangular solve whose pattern is the elimination-tree     */
/* reach of row k's off-diagonal entries.  L: lower CSC, diagonal first, rows sorted.         */
/* Returns nnz(L), -nnz if cap too small, or INT64_MIN + k + 1 on a non-positive pivot at k.  */
/* ---------------------------------------------------------------------------------------- */
int64_t orc_chol(int64_t n, const int64_t* Cp, const int64_t* Ci, const double* Cx, int64_t* Lp,
                 int64_t* Li, double* Lx, int64_t cap) {
    /* transpose to get rows of the lower triangle (= columns of the upper) */
    const int64_t nz = Cp[n];
    int64_t* Rp = calloc(n + 1, sizeof(int64_t));
    int64_t* Ri = malloc((nz ? nz : 1) * sizeof(int64_t));
    double* Rx = malloc((nz ? nz : 1) * sizeof(double));
    for (int64_t p = 0; p < nz; ++p) Rp[Ci[p] + 1]++;
    for (int64_t i = 0; i < n; ++i) Rp[i + 1] += Rp[i];
    int64_t* w = malloc((n + 1) * sizeof(int64_t));
    memcpy(w, Rp, (n + 1) * sizeof(int64_t));
    for (int64_t j = 0; j < n; ++j)
        for (int64_t p = Cp[j]; p < Cp[j + 1]; ++p) { Ri[w[Ci[p]]] = j; Rx[w[Ci[p]]++] = Cx[p]; }
    /* elimination tree */
    int64_t* parent = malloc(n * sizeof(int64_t));
    int64_t* anc = malloc(n * sizeof(int64_t));
    for (int64_t k = 0; k < n; ++k) {
        parent[k] = -1; anc[k] = -1;
        for (int64_t p = Rp[k]; p < Rp[k + 1]; ++p) {
            int64_t i = Ri[p];
            while (i != -1 && i < k) {
                const int64_t nx = anc[i];
                anc[i] = k;
                if (nx == -1) { parent[i] = k; break; }
                i = nx;
            }
        }
    }
    int64_t* mark = malloc(n * sizeof(int64_t));
    int64_t* stack = malloc(n * sizeof(int64_t));
    int64_t* cc = calloc(n, sizeof(int64_t));
    /* counts: column j gets one entry per row k whose reach contains j (+ diagonal) */
    for (int64_t k = 0; k < n; ++k) mark[k] = -1;
    for (int64_t k = 0; k < n; ++k) {
        mark[k] = k;
        cc[k]++;
        for (int64_t p = Rp[k]; p < Rp[k + 1]; ++p) {
            int64_t i = Ri[p];
            if (i >= k) continue;
            for (; mark[i] != k; i = parent[i]) { mark[i] = k; cc[i]++; }
        }
    }
    Lp[0] = 0;
    for (int64_t j = 0; j < n; ++j) Lp[j + 1] = Lp[j] + cc[j];
    const int64_t lnz = Lp[n];
    int64_t result = lnz;
    if (lnz > cap) { result = -lnz; goto done; }
    {
        double* x = calloc(n, sizeof(double));
        int64_t* c = malloc(n * sizeof(int64_t));
        for (int64_t j = 0; j < n; ++j) { c[j] = Lp[j]; mark[j] = -1; }
        for (int64_t k = 0; k < n; ++k) {
            /* reach of row k in topological order */
            int64_t top = n;
            mark[k] = k;
            for (int64_t p = Rp[k]; p < Rp[k + 1]; ++p) {
                int64_t i = Ri[p];
                if (i > k) continue;
                x[i] += Rx[p];
                if (i == k) continue;
                int64_t len = 0;
                for (; mark[i] != k; i = parent[i]) { w[len++] = i; mark[i] = k; }
                while (len > 0) stack[--top] = w[--len];
            }
            double d = x[k];
            x[k] = 0;
            for (; top < n; ++top) {
                const int64_t j = stack[top];
                const double lkj = x[j] / Lx[Lp[j]];
                x[j] = 0;
                for (int64_t p = Lp[j] + 1; p < c[j]; ++p) x[Li[p]] -= Lx[p] * lkj;
                d -= lkj * lkj;
                Li[c[j]] = k;
                Lx[c[j]++] = lkj;
            }
            if (!(d > 0)) { result = INT64_MIN + k + 1; free(x); free(c); goto done; }
            Li[c[k]] = k;
            Lx[c[k]++] = sqrt(d);
        }
        free(x);
        free(c);
    }
done:
    free(Rp); free(Ri); free(Rx); free(w); free(parent); free(anc); free(mark); free(stack); free(cc);
    return result;
}

/* L x = b in place, L y = ... (diagonal first in each column) */
void orc_lsolve(int64_t n, const int64_t* Lp, const int64_t* Li, const double* Lx, double* x) {
    for (int64_t j = 0; j < n; ++j) {
        x[j] /= Lx[Lp[j]];
        for (int64_t p = Lp[j] + 1; p < Lp[j + 1]; ++p) x[Li[p]] -= Lx[p] * x[j];
    }
}
void orc_ltsolve(int64_t n, const int64_t* Lp, const int64_t* Li, const double* Lx, double* x) {
    for (int64_t j = n - 1; j >= 0; --j) {
        for (int64_t p = Lp[j] + 1; p < Lp[j + 1]; ++p) x[j] -= Lx[p] * x[Li[p]];
        x[j] /= Lx[Lp[j]];
    }
}

/* trs_kernels::MultiplyLowerBlock<Field, BS>::run semantics:
 *   for i < degree: row = indices[i / BS] + i % BS;
 *   B(row, :) -= sum_c A(i, c) * B(supernode_start + c, :)                                  */
void orc_multiply_lower_block(int BS, const int64_t* indices, int64_t start, int64_t size,
                              int64_t degree, const double* A, int64_t lda, int nrhs, double* B,
                              int64_t ldb) {
    for (int j = 0; j < nrhs; ++j) {
        double* Bj = B + (int64_t)j * ldb;
        for (int64_t i = 0; i < degree; ++i) {
            double s = 0;
            for (int64_t c = 0; c < size; ++c) s += A[c * lda + i] * Bj[start + c];
            Bj[indices[i / BS] + i % BS] -= s;
        }
    }
}
