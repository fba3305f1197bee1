/*
 * supernodal.c -- multifrontal supernodal Cholesky on the host: the CPU baseline of the
 * per-Newton-step factorization and solve.
 *
 * TEST INFRASTRUCTURE ONLY.  bench.py's cpu_baseline leg times it and tests/ check it; the
 * product path (meshfemsparse_amd) never calls it.
 *
 * It restates the reference's CPU algorithm class: a relaxed-supernodal LL^T whose dense work per
 * supernode is POTRF of the diagonal block, TRSM of the rows below and a SYRK Schur update
 * scattered into the parent (CatamariFactorizer.cc:97-109 selects BlockCatamari's supernodal
 * Cholesky; CatamariFactorizer.cc:634-714 runs its RefactorWithFixedSparsityPattern), with the
 * dense kernels from a sequential BLAS on each supernode and the elimination tree processed in
 * parallel (benchmark_linear_systems.cc:68-69 pins MKL to sequential mode under TBB).  Here the
 * tree runs level by level under OpenMP; levels with fewer supernodes than threads hand the
 * threads to the BLAS instead (the top separators).  The BLAS is the OpenBLAS that ships with
 * scipy, loaded at run time (orc_blas_load); nothing else of it is used.
 *
 * Input: the scalar lower CSC (rows >= column, sorted, diagonal first) of P A_red P^T, three
 * scalars per block, as orc_permuted_lower writes it.  The symbolic phase works on the block
 * pattern: elimination tree, block column counts by row subtrees, supernodes = chains of
 * consecutive columns j -> j + 1 = parent(j) that either are fundamental (count(j) =
 * count(j + 1) + 1) or keep the supernode at most `relax` blocks wide; each supernode's row
 * structure is the union of its columns' rows and its children's structures.  Pinned by
 * tests/test_oracle_supernodal.py against orc_chol (the unique factor of the same ordering)
 * and dense solves.
 */
#include <dlfcn.h>
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef void (*dpotrf_f)(const char*, const int*, double*, const int*, int*, size_t);
typedef void (*dtrsm_f)(const char*, const char*, const char*, const char*, const int*, const int*,
                        const double*, const double*, const int*, double*, const int*, size_t, size_t,
                        size_t, size_t);
typedef void (*dsyrk_f)(const char*, const char*, const int*, const int*, const double*, const double*,
                        const int*, const double*, double*, const int*, size_t, size_t);
typedef void (*dgemv_f)(const char*, const int*, const int*, const double*, const double*, const int*,
                        const double*, const int*, const double*, double*, const int*, size_t);
typedef void (*dtrsv_f)(const char*, const char*, const char*, const int*, const double*, const int*,
                        double*, const int*, size_t, size_t, size_t);
typedef void (*setnt_f)(int);

static dpotrf_f p_dpotrf;
static dtrsm_f p_dtrsm;
static dsyrk_f p_dsyrk;
static dgemv_f p_dgemv;
static dtrsv_f p_dtrsv;
static setnt_f p_setnt;

/* 0 on success; -1 if the library or a symbol is missing */
int orc_blas_load(const char* path) {
    void* h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!h) return -1;
    p_dpotrf = (dpotrf_f)dlsym(h, "scipy_dpotrf_");
    p_dtrsm = (dtrsm_f)dlsym(h, "scipy_dtrsm_");
    p_dsyrk = (dsyrk_f)dlsym(h, "scipy_dsyrk_");
    p_dgemv = (dgemv_f)dlsym(h, "scipy_dgemv_");
    p_dtrsv = (dtrsv_f)dlsym(h, "scipy_dtrsv_");
    p_setnt = (setnt_f)dlsym(h, "scipy_openblas_set_num_threads");
    return (p_dpotrf && p_dtrsm && p_dsyrk && p_dgemv && p_dtrsv && p_setnt) ? 0 : -1;
}

typedef struct {
    int64_t nb, ns;
    int64_t* sfirst;    /* [ns + 1] first block column of each supernode */
    int64_t* sparent;   /* [ns] parent supernode or -1 */
    int64_t* soff;      /* [ns + 1] offsets of the structure rows (blocks) */
    int64_t* srows;     /* structure rows: blocks below the supernode, ascending */
    int64_t* poff;      /* [ns + 1] offset of the supernode's panel ((W + D) x W, scalars) */
    int64_t* lev_ptr;   /* [nlev + 1] supernodes by level (leaves first) */
    int64_t* lev_sn;
    int64_t nlev;
    double flops;       /* sum W^3/3 + W^2 D + W D^2 (scalars) */
    int64_t* uoff;      /* [ns] offset of the update matrix (D x D) in the arena */
    int64_t usize;      /* arena size (scalars) */
    int64_t* chp;       /* [ns + 1] children (CSR), ascending */
    int64_t* chi;
    double* L;          /* panels, allocated on the first factorization and kept */
    double* U;          /* update-matrix arena, likewise */
} orc_sn;

/* First-fit placement of the update matrices: a supernode's U is placed when its level starts
 * and released after its parent's level has assembled it, so U's that are live together never
 * overlap. */
typedef struct { int64_t off, size; } span;

static int64_t arena_alloc(span** fl, int64_t* nfl, int64_t* end, int64_t size) {
    for (int64_t i = 0; i < *nfl; ++i)
        if ((*fl)[i].size >= size) {
            const int64_t off = (*fl)[i].off;
            (*fl)[i].off += size;
            (*fl)[i].size -= size;
            if ((*fl)[i].size == 0) {
                memmove(*fl + i, *fl + i + 1, (*nfl - i - 1) * sizeof(span));
                --*nfl;
            }
            return off;
        }
    const int64_t off = *end;
    *end += size;
    return off;
}

static void arena_release(span** fl, int64_t* nfl, int64_t* cap, int64_t off, int64_t size) {
    int64_t i = 0;
    while (i < *nfl && (*fl)[i].off < off) ++i;
    if (*nfl + 1 > *cap) {
        *cap = 2 * *cap + 16;
        *fl = realloc(*fl, *cap * sizeof(span));
    }
    memmove(*fl + i + 1, *fl + i, (*nfl - i) * sizeof(span));
    (*fl)[i].off = off;
    (*fl)[i].size = size;
    ++*nfl;
    /* merge with the neighbours */
    if (i + 1 < *nfl && (*fl)[i].off + (*fl)[i].size == (*fl)[i + 1].off) {
        (*fl)[i].size += (*fl)[i + 1].size;
        memmove(*fl + i + 1, *fl + i + 2, (*nfl - i - 2) * sizeof(span));
        --*nfl;
    }
    if (i > 0 && (*fl)[i - 1].off + (*fl)[i - 1].size == (*fl)[i].off) {
        (*fl)[i - 1].size += (*fl)[i].size;
        memmove(*fl + i, *fl + i + 1, (*nfl - i - 1) * sizeof(span));
        --*nfl;
    }
}

/* seconds spent per phase by all threads of the last factorization: front assembly (with the
 * extend-add), POTRF, TRSM, SYRK, panel/update copies */
static double g_phase[7];   /* + wall seconds in tree-parallel levels and in BLAS-parallel levels */

static int cmp_i64(const void* a, const void* b) {
    const int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
    return (x > y) - (x < y);
}

/* Symbolic analysis of the scalar lower CSC (n = 3 nb); relax = widest merged chain (blocks). */
orc_sn* orc_sn_analyze(int64_t n, const int64_t* Cp, const int64_t* Ci, int relax) {
    if (n % 3) return NULL;
    const int64_t nb = n / 3;
    /* block pattern: rows of scalar column 3J divided by 3, unique (rows >= J) */
    int64_t* Bp = malloc((nb + 1) * sizeof(int64_t));
    int64_t* Bi = malloc((Cp[n] / 3 + nb + 1) * sizeof(int64_t));
    Bp[0] = 0;
    for (int64_t J = 0; J < nb; ++J) {
        int64_t q = Bp[J], last = -1;
        for (int64_t p = Cp[3 * J]; p < Cp[3 * J + 1]; ++p) {
            const int64_t I = Ci[p] / 3;
            if (I != last) { Bi[q++] = I; last = I; }
        }
        Bp[J + 1] = q;
    }
    /* transpose (row lists of the strict lower part) for the elimination tree */
    int64_t* Tp = calloc(nb + 1, sizeof(int64_t));
    for (int64_t p = 0; p < Bp[nb]; ++p) Tp[Bi[p] + 1]++;
    for (int64_t i = 0; i < nb; ++i) Tp[i + 1] += Tp[i];
    int64_t* Ti = malloc((Bp[nb] + 1) * sizeof(int64_t));
    int64_t* w = malloc((nb + 1) * sizeof(int64_t));
    memcpy(w, Tp, (nb + 1) * sizeof(int64_t));
    for (int64_t J = 0; J < nb; ++J)
        for (int64_t p = Bp[J]; p < Bp[J + 1]; ++p) Ti[w[Bi[p]]++] = J;
    int64_t* parent = malloc(nb * sizeof(int64_t));
    int64_t* anc = malloc(nb * sizeof(int64_t));
    for (int64_t k = 0; k < nb; ++k) {
        parent[k] = -1; anc[k] = -1;
        for (int64_t p = Tp[k]; p < Tp[k + 1]; ++p) {
            int64_t i = Ti[p];
            while (i != -1 && i < k) {
                const int64_t nx = anc[i];
                anc[i] = k;
                if (nx == -1) { parent[i] = k; break; }
                i = nx;
            }
        }
    }
    /* block column counts (with the diagonal) by row subtrees */
    int64_t* cnt = calloc(nb, sizeof(int64_t));
    int64_t* mark = malloc(nb * sizeof(int64_t));
    for (int64_t k = 0; k < nb; ++k) mark[k] = -1;
    for (int64_t k = 0; k < nb; ++k) {
        mark[k] = k;
        cnt[k]++;
        for (int64_t p = Tp[k]; p < Tp[k + 1]; ++p)
            for (int64_t i = Ti[p]; i < k && mark[i] != k; i = parent[i]) { mark[i] = k; cnt[i]++; }
    }
    /* supernodes: chains j -> j + 1 */
    int64_t* sfirst = malloc((nb + 1) * sizeof(int64_t));
    int64_t* sn_of = malloc(nb * sizeof(int64_t));
    int64_t ns = 0;
    for (int64_t j = 0; j < nb; ++j) {
        int merge = 0;
        if (j > 0 && parent[j - 1] == j) {
            const int64_t width = j - sfirst[ns - 1];
            merge = (cnt[j - 1] == cnt[j] + 1) || (width < relax);
        }
        if (!merge) sfirst[ns++] = j;
        sn_of[j] = ns - 1;
    }
    sfirst[ns] = nb;
    orc_sn* S = calloc(1, sizeof(orc_sn));
    S->nb = nb;
    S->ns = ns;
    S->sfirst = realloc(sfirst, (ns + 1) * sizeof(int64_t));
    S->sparent = malloc(ns * sizeof(int64_t));
    for (int64_t s = 0; s < ns; ++s) {
        const int64_t pj = parent[S->sfirst[s + 1] - 1];
        S->sparent[s] = pj < 0 ? -1 : sn_of[pj];
    }
    /* children lists */
    int64_t* chp = calloc(ns + 1, sizeof(int64_t));
    for (int64_t s = 0; s < ns; ++s) if (S->sparent[s] >= 0) chp[S->sparent[s] + 1]++;
    for (int64_t s = 0; s < ns; ++s) chp[s + 1] += chp[s];
    int64_t* chi = malloc((ns + 1) * sizeof(int64_t));
    int64_t* cw = malloc((ns + 1) * sizeof(int64_t));
    memcpy(cw, chp, (ns + 1) * sizeof(int64_t));
    for (int64_t s = 0; s < ns; ++s) if (S->sparent[s] >= 0) chi[cw[S->sparent[s]]++] = s;
    /* structures: union of the columns' rows and the children's structures, rows > last */
    int64_t cap = 1024, used = 0;
    int64_t* rows = malloc(cap * sizeof(int64_t));
    S->soff = malloc((ns + 1) * sizeof(int64_t));
    for (int64_t k = 0; k < nb; ++k) mark[k] = -1;
    int64_t* tmp = malloc((nb + 1) * sizeof(int64_t));
    for (int64_t s = 0; s < ns; ++s) {
        const int64_t f = S->sfirst[s], l = S->sfirst[s + 1] - 1;
        int64_t m = 0;
        for (int64_t j = f; j <= l; ++j)
            for (int64_t p = Bp[j]; p < Bp[j + 1]; ++p) {
                const int64_t I = Bi[p];
                if (I > l && mark[I] != s) { mark[I] = s; tmp[m++] = I; }
            }
        for (int64_t q = chp[s]; q < chp[s + 1]; ++q) {
            const int64_t c = chi[q];
            for (int64_t p = S->soff[c]; p < S->soff[c + 1]; ++p) {
                const int64_t I = rows[p];
                if (I > l && mark[I] != s) { mark[I] = s; tmp[m++] = I; }
            }
        }
        qsort(tmp, (size_t)m, sizeof(int64_t), cmp_i64);
        if (used + m > cap) {
            while (used + m > cap) cap *= 2;
            rows = realloc(rows, cap * sizeof(int64_t));
        }
        S->soff[s] = used;
        memcpy(rows + used, tmp, m * sizeof(int64_t));
        used += m;
        S->soff[s + 1] = used;
    }
    S->srows = rows;
    /* panels, flops, levels */
    S->poff = malloc((ns + 1) * sizeof(int64_t));
    int64_t* lev = malloc(ns * sizeof(int64_t));
    int64_t nlev = 0;
    S->poff[0] = 0;
    S->flops = 0;
    for (int64_t s = 0; s < ns; ++s) {
        const double W = 3.0 * (S->sfirst[s + 1] - S->sfirst[s]), D = 3.0 * (S->soff[s + 1] - S->soff[s]);
        S->poff[s + 1] = S->poff[s] + (int64_t)((W + D) * W);
        S->flops += W * W * W / 3.0 + W * W * D + W * D * D;
        int64_t lv = 0;
        for (int64_t q = chp[s]; q < chp[s + 1]; ++q) if (lev[chi[q]] + 1 > lv) lv = lev[chi[q]] + 1;
        lev[s] = lv;
        if (lv + 1 > nlev) nlev = lv + 1;
    }
    S->nlev = nlev;
    S->lev_ptr = calloc(nlev + 1, sizeof(int64_t));
    for (int64_t s = 0; s < ns; ++s) S->lev_ptr[lev[s] + 1]++;
    for (int64_t l = 0; l < nlev; ++l) S->lev_ptr[l + 1] += S->lev_ptr[l];
    S->lev_sn = malloc((ns + 1) * sizeof(int64_t));
    memcpy(cw, S->lev_ptr, (nlev + 1) * sizeof(int64_t));
    for (int64_t s = 0; s < ns; ++s) S->lev_sn[cw[lev[s]]++] = s;
    /* update-matrix arena */
    S->uoff = malloc((ns + 1) * sizeof(int64_t));
    {
        span* fl = NULL;
        int64_t nfl = 0, cap = 0, end = 0;
        for (int64_t lv = 0; lv < nlev; ++lv) {
            for (int64_t q = S->lev_ptr[lv]; q < S->lev_ptr[lv + 1]; ++q) {
                const int64_t t = S->lev_sn[q], D = 3 * (S->soff[t + 1] - S->soff[t]);
                S->uoff[t] = D > 0 ? arena_alloc(&fl, &nfl, &end, D * D) : 0;
            }
            for (int64_t q = S->lev_ptr[lv]; q < S->lev_ptr[lv + 1]; ++q) {
                const int64_t t = S->lev_sn[q];
                for (int64_t r = chp[t]; r < chp[t + 1]; ++r) {
                    const int64_t c = chi[r], D = 3 * (S->soff[c + 1] - S->soff[c]);
                    if (D > 0) arena_release(&fl, &nfl, &cap, S->uoff[c], D * D);
                }
            }
        }
        S->usize = end;
        free(fl);
    }
    S->chp = chp;
    S->chi = chi;
    free(Bp); free(Bi); free(Tp); free(Ti); free(w); free(parent); free(anc); free(cnt); free(mark);
    free(sn_of); free(cw); free(tmp); free(lev);
    return S;
}

/* out[0..5] = supernodes, levels, panel scalars (factor storage), flops, widest supernode
 * (scalars), update-matrix arena (scalars) */
void orc_sn_info(const orc_sn* S, double* out) {
    out[5] = (double)S->usize;
    int64_t mw = 0;
    for (int64_t s = 0; s < S->ns; ++s)
        if (S->sfirst[s + 1] - S->sfirst[s] > mw) mw = S->sfirst[s + 1] - S->sfirst[s];
    out[0] = (double)S->ns;
    out[1] = (double)S->nlev;
    out[2] = (double)S->poff[S->ns];
    out[3] = S->flops;
    out[4] = 3.0 * (double)mw;
}

/* One supernode, in place: the panel P ((W + D) x W) receives the matrix columns and the
 * children's update-matrix entries in its columns, the update matrix U (D x D) their entries in
 * the structure rows (children in ascending order, fixed entry order); then POTRF of P's top,
 * TRSM of its rows below, and U += -L21 L21^T (SYRK).  fpos: per-thread block -> front position
 * map.  Returns -1 or the first failing scalar column. */
static int64_t factor_one(orc_sn* S, int64_t s, const int64_t* Cp, const int64_t* Ci, const double* Cx,
                          int64_t* fpos, int par) {
    const int64_t f = S->sfirst[s], l = S->sfirst[s + 1] - 1;
    const int W = (int)(3 * (l - f + 1)), D = (int)(3 * (S->soff[s + 1] - S->soff[s]));
    const int ld = W + D;
    double tp[6];
    tp[0] = omp_get_wtime();
    double* P = S->L + S->poff[s];
    double* U = S->U + S->uoff[s];
    /* par > 1: the supernode has the threads to itself (a level with few supernodes); columns
     * are independent, so they are split among the threads */
#pragma omp parallel for num_threads(par) schedule(static) if (par > 1)
    for (int64_t j = 0; j < W; ++j) memset(P + j * ld, 0, (size_t)ld * sizeof(double));
#pragma omp parallel for num_threads(par) schedule(static) if (par > 1)
    for (int64_t j = 0; j < D; ++j) memset(U + j * D, 0, (size_t)D * sizeof(double));
    for (int64_t j = f; j <= l; ++j) fpos[j] = 3 * (j - f);
    for (int64_t p = S->soff[s]; p < S->soff[s + 1]; ++p) fpos[S->srows[p]] = W + 3 * (p - S->soff[s]);
#pragma omp parallel for num_threads(par) schedule(static) if (par > 1)
    for (int64_t c = 3 * f; c <= 3 * l + 2; ++c) {
        double* col = P + (c - 3 * f) * ld;
        for (int64_t p = Cp[c]; p < Cp[c + 1]; ++p) {
            const int64_t r = Ci[p];
            col[fpos[r / 3] + r % 3] += Cx[p];
        }
    }
    for (int64_t q = S->chp[s]; q < S->chp[s + 1]; ++q) {
        const int64_t c = S->chi[q];
        const int64_t nc = S->soff[c + 1] - S->soff[c];
        const int Dc = (int)(3 * nc);
        const double* Uc = S->U + S->uoff[c];
        const int64_t* rc = S->srows + S->soff[c];
#pragma omp parallel for num_threads(par) schedule(dynamic, 16) if (par > 1)
        for (int64_t bj = 0; bj < nc; ++bj) {
            const int64_t pj = fpos[rc[bj]];
            for (int cj = 0; cj < 3; ++cj) {
                const int64_t fj = pj + cj;
                double* dst = fj < W ? P + fj * ld : U + (fj - W) * D - W;   /* indexed by front row */
                const double* src = Uc + (3 * bj + cj) * Dc;
                for (int64_t bi = bj; bi < nc; ++bi) {
                    const int64_t fi = fpos[rc[bi]];
                    for (int ci = (bi == bj ? cj : 0); ci < 3; ++ci) dst[fi + ci] += src[3 * bi + ci];
                }
            }
        }
    }
    int info = 0;
    tp[1] = omp_get_wtime();
    /* diagonal block: right-looking in steps of NBK columns (POTRF of the step's block, TRSM of
     * the rows below inside the diagonal block, SYRK of the rest of it) */
    const int NBK = 256;
    const double one = 1.0, mone = -1.0;
    for (int k = 0; k < W; k += NBK) {
        const int kb = W - k < NBK ? W - k : NBK, rest = W - k - kb;
        double* Akk = P + (int64_t)k * ld + k;
        p_dpotrf("L", &kb, Akk, &ld, &info, 1);
        if (info != 0) return 3 * f + k + info - 1;
        if (rest > 0) {
            p_dtrsm("R", "L", "T", "N", &rest, &kb, &one, Akk, &ld, Akk + kb, &ld, 1, 1, 1, 1);
            p_dsyrk("L", "N", &rest, &kb, &mone, Akk + kb, &ld, &one, Akk + (int64_t)kb * ld + kb, &ld, 1, 1);
        }
    }
    tp[2] = omp_get_wtime();
    tp[3] = tp[4] = tp[2];
    if (D > 0) {
        p_dtrsm("R", "L", "T", "N", &D, &W, &one, P, &ld, P + W, &ld, 1, 1, 1, 1);
        tp[3] = omp_get_wtime();
        p_dsyrk("L", "N", &D, &W, &mone, P + W, &ld, &one, U, &D, 1, 1);
        tp[4] = omp_get_wtime();
    }
    tp[5] = omp_get_wtime();
#pragma omp atomic
    g_phase[0] += tp[1] - tp[0];
#pragma omp atomic
    g_phase[1] += tp[2] - tp[1];
#pragma omp atomic
    g_phase[2] += tp[3] - tp[2];
#pragma omp atomic
    g_phase[3] += tp[4] - tp[3];
#pragma omp atomic
    g_phase[4] += tp[5] - tp[4];
    return -1;
}

void orc_sn_phase_times(double* out7) {
    for (int i = 0; i < 7; ++i) out7[i] = g_phase[i];
}

/* Numeric factorization with `nthreads` threads; returns -1 or the first failing scalar column
 * found (a pivot failure stops the factorization at the end of its level).  The factor and the
 * update-matrix arena are allocated on the first call and reused. */
int64_t orc_sn_factor(orc_sn* S, const int64_t* Cp, const int64_t* Ci, const double* Cx, int nthreads) {
    if (!p_dpotrf) return -2;
    if (!S->L) S->L = malloc((size_t)(S->poff[S->ns] ? S->poff[S->ns] : 1) * sizeof(double));
    if (!S->U) S->U = malloc((size_t)(S->usize ? S->usize : 1) * sizeof(double));
    if (nthreads < 1) nthreads = 1;
    int64_t** fpos = malloc(nthreads * sizeof(int64_t*));
    for (int t = 0; t < nthreads; ++t) fpos[t] = malloc(S->nb * sizeof(int64_t));
    int64_t fail = -1;
    for (int i = 0; i < 7; ++i) g_phase[i] = 0.0;
    for (int64_t lv = 0; lv < S->nlev && fail < 0; ++lv) {
        const int64_t a = S->lev_ptr[lv], b = S->lev_ptr[lv + 1];
        const double t0 = omp_get_wtime();
        const int tree = b - a >= nthreads;
        if (tree) {
            p_setnt(1);
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
            for (int64_t q = a; q < b; ++q) {
                const int64_t r = factor_one(S, S->lev_sn[q], Cp, Ci, Cx, fpos[omp_get_thread_num()], 1);
                if (r >= 0) {
#pragma omp critical
                    if (fail < 0 || r < fail) fail = r;
                }
            }
        } else {
            p_setnt(nthreads);
            for (int64_t q = a; q < b; ++q) {
                const int64_t r = factor_one(S, S->lev_sn[q], Cp, Ci, Cx, fpos[0], nthreads);
                if (r >= 0 && (fail < 0 || r < fail)) fail = r;
            }
        }
        g_phase[tree ? 5 : 6] += omp_get_wtime() - t0;
    }
    p_setnt(nthreads);
    for (int t = 0; t < nthreads; ++t) free(fpos[t]);
    free(fpos);
    return fail;
}

/* x := (L L^T)^{-1} x in place (scalar factor order). */
void orc_sn_solve(const orc_sn* S, double* x) {
    const int one = 1;
    const double d1 = 1.0, dm1 = -1.0;
    double* t = malloc((size_t)(3 * S->nb + 1) * sizeof(double));
    for (int64_t s = 0; s < S->ns; ++s) {
        const int64_t f = S->sfirst[s], l = S->sfirst[s + 1] - 1;
        const int W = (int)(3 * (l - f + 1)), D = (int)(3 * (S->soff[s + 1] - S->soff[s])), ld = W + D;
        const double* P = S->L + S->poff[s];
        double* xs = x + 3 * f;
        p_dtrsv("L", "N", "N", &W, P, &ld, xs, &one, 1, 1, 1);
        if (D > 0) {
            const double zero = 0.0;
            p_dgemv("N", &D, &W, &d1, P + W, &ld, xs, &one, &zero, t, &one, 1);
            const int64_t* r = S->srows + S->soff[s];
            for (int64_t b = 0; b < D / 3; ++b)
                for (int c = 0; c < 3; ++c) x[3 * r[b] + c] -= t[3 * b + c];
        }
    }
    for (int64_t s = S->ns - 1; s >= 0; --s) {
        const int64_t f = S->sfirst[s], l = S->sfirst[s + 1] - 1;
        const int W = (int)(3 * (l - f + 1)), D = (int)(3 * (S->soff[s + 1] - S->soff[s])), ld = W + D;
        const double* P = S->L + S->poff[s];
        double* xs = x + 3 * f;
        if (D > 0) {
            const int64_t* r = S->srows + S->soff[s];
            for (int64_t b = 0; b < D / 3; ++b)
                for (int c = 0; c < 3; ++c) t[3 * b + c] = x[3 * r[b] + c];
            p_dgemv("T", &D, &W, &dm1, P + W, &ld, t, &one, &d1, xs, &one, 1);
        }
        p_dtrsv("L", "T", "N", &W, P, &ld, xs, &one, 1, 1, 1);
    }
    free(t);
}

/* Dense copy of L (n x n column-major, n = 3 nb) for small cases. */
void orc_sn_dense_L(const orc_sn* S, double* Ld) {
    const int64_t n = 3 * S->nb;
    memset(Ld, 0, (size_t)(n * n) * sizeof(double));
    for (int64_t s = 0; s < S->ns; ++s) {
        const int64_t f = S->sfirst[s], l = S->sfirst[s + 1] - 1;
        const int64_t W = 3 * (l - f + 1), D = 3 * (S->soff[s + 1] - S->soff[s]), ld = W + D;
        const double* P = S->L + S->poff[s];
        const int64_t* r = S->srows + S->soff[s];
        for (int64_t j = 0; j < W; ++j) {
            const int64_t cj = 3 * f + j;
            for (int64_t i = j; i < W; ++i) Ld[cj * n + 3 * f + i] = P[j * ld + i];
            for (int64_t i = 0; i < D; ++i) Ld[cj * n + 3 * r[i / 3] + i % 3] = P[j * ld + W + i];
        }
    }
}

void orc_sn_free(orc_sn* S) {
    if (!S) return;
    free(S->U); free(S->sfirst); free(S->sparent); free(S->soff); free(S->srows); free(S->poff);
    free(S->lev_ptr); free(S->lev_sn); free(S->L); free(S->uoff); free(S->chp); free(S->chi);
    free(S);
}
