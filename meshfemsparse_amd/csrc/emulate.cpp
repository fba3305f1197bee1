// Host replay of the device factorization schedule: the same task lists (injection plan,
// extend-add tasks, diagonal-tile POTRF + inverse, tile GEMMs) executed by plain loops.  Used by
// CPU tests to check the symbolic schedule independently of the GPU kernels.
#include <algorithm>
#include <cmath>
#include <climits>
#include <cstdlib>

#include "mfs_internal.h"

namespace mfs {

int64_t emulate_numeric(const Symbolic& S, const double* Ax, double sigma, std::vector<double>& L,
                        std::vector<double>& inv) {
    const int TS = 64;
    // the Schur-update tiles' own lists of children's update-matrix blocks, added after the tile
    // (child order), in place of the extend-add passes into U, and the panel tiles' lists in place
    // of the passes into the panels; MFS_PULL=0 selects the passes, MFS_PULL=2 makes an analysis
    // without complete lists an error (replay checks)
    const char* pe = getenv("MFS_PULL");
    const int pmode = pe ? atoi(pe) : 1;
    const bool pull = S.pulls_complete && pmode != 0, ppull = S.panel_pulls_complete && pmode != 0;
    if (pmode == 2 && !(pull && ppull)) throw Error(-1, "replay: the tiles carry no complete children lists");
    if (ppull) {
        // as on the device only the tiles the lists do not rebuild are cleared; the rest starts
        // as NaN, so an entry read before its tile was rebuilt shows up in the result
        L.assign(S.L_size, std::nan(""));
        for (const FillTile& f : S.fill_tiles)
            for (int c = 0; c < f.N; ++c)
                for (int r = 0; r < f.M; ++r) L[f.off + (int64_t)c * f.ld + r] = 0.0;
    } else {
        L.assign(S.L_size, 0.0);
    }
    inv.assign((size_t)S.inv_slots * TS * TS, 0.0);
    // the update arena is never cleared on the device: start from NaNs so that any read of an
    // entry that was not written first shows up in the result
    std::vector<double> U(S.U_size, std::nan(""));
    int64_t status = INT64_MAX;
    // injection: put(r, c, v) for the stored entries of plan entry q's 3x3 block
    auto inject_entry = [&](size_t q, auto&& put) {
        const int mode = (int)(S.plan_dst[q] & 3);
        const double* a = Ax + 9 * (int64_t)S.plan_src[q];
        const int mr = S.masked() ? S.comp_mask[S.plan_row[q]] : 0;
        const int mc = S.masked() ? S.comp_mask[S.plan_col[q]] : 0;
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) {
                if (mode == 2 && r < c) continue;
                double v = mode == 0 ? a[3 * c + r] : a[3 * r + c];
                if (mode == 2 && r == c) v += sigma;
                if (((mr >> r) | (mc >> c)) & 1) v = (mode == 2 && r == c) ? 1.0 : 0.0;   // masked: identity
                put(r, c, v);
            }
    };
    for (size_t q = 0; q < S.plan_src.size(); ++q) {
        const int64_t dst = S.plan_dst[q] >> 2;
        const int32_t s = (int32_t)(std::upper_bound(S.panel_off.begin(), S.panel_off.end(), dst) -
                                    S.panel_off.begin()) - 1;
        const int64_t ld = 3 * ((int64_t)S.swidth[s] + S.sdeg[s]);
        inject_entry(q, [&](int r, int c, double v) { L[dst + c * ld + r] = v; });
    }
    auto gemm = [&](const GemmTask& t) {
        const double* A = L.data() + t.a_off;
        const double* B = (t.flags & GT_B_IN_INV) ? inv.data() + t.b_off : L.data() + t.b_off;
        double* C = (t.flags & GT_C_IN_U) ? U.data() + t.c_off : L.data() + t.c_off;
        const double alpha = (t.flags & GT_ALPHA_POS) ? 1.0 : -1.0;
        std::vector<double> tmp((size_t)t.M * t.N, 0.0);
        for (int n = 0; n < t.N; ++n)
            for (int k = 0; k < t.K; ++k) {
                const double b = B[(int64_t)k * t.ldb + n];
                for (int m = 0; m < t.M; ++m) tmp[(size_t)n * t.M + m] += A[(int64_t)k * t.lda + m] * b;
            }
        for (int n = 0; n < t.N; ++n)
            for (int m = 0; m < t.M; ++m) {
                double& c = C[(int64_t)n * t.ldc + m];
                const double v = alpha * tmp[(size_t)n * t.M + m];
                c = (t.flags & GT_BETA0) ? v : c + v;
            }
    };
    // rows [row0, row1) of the child's column block kj (row1 = 3 dc for a whole column)
    auto extend_rows = [&](const ExtendTask& t, int64_t row0, int64_t row1) {
        const int32_t* relc = S.rel.data() + t.rel_off;
        const int kj = t.col_block, tj = relc[kj];
        const int64_t Dc = 3 * (int64_t)t.dc;
        for (int64_t i = row0; i < row1; ++i) {
            const int ki = (int)(i / 3), r = (int)(i % 3), ti = relc[ki];
            double* dst;
            int64_t ldd;
            if (tj < t.wp) { dst = L.data() + t.p_panel + (int64_t)(3 * tj) * t.ldp + 3 * ti + r; ldd = t.ldp; }
            else {
                const int64_t Dp = 3 * (int64_t)t.dp;
                dst = U.data() + t.u_parent + (int64_t)(3 * (tj - t.wp)) * Dp + 3 * (ti - t.wp) + r;
                ldd = Dp;
            }
            const double* src = U.data() + t.u_child + 3 * (int64_t)kj * Dc + i;
            for (int c = 0; c < 3; ++c)
                if (ki != kj || r >= c) dst[c * ldd] += src[c * Dc];
        }
    };
    auto extend_one = [&](const ExtendTask& t) {
        extend_rows(t, t.row0, std::min<int64_t>(3 * (int64_t)t.dc, (int64_t)t.row0 + kExtendRows));
    };
    auto potrf = [&](const PotrfTask& t) {
        double* D = L.data() + t.diag_off;
        const int kb = t.kb;
        const int64_t ld = t.ld;
        for (int j = 0; j < kb; ++j) {
            double d = D[j * ld + j];
            for (int k = 0; k < j; ++k) d -= D[k * ld + j] * D[k * ld + j];
            if (!(d > 0.0)) { status = std::min<int64_t>(status, t.col0 + j); d = 1.0; }
            d = std::sqrt(d);
            D[j * ld + j] = d;
            for (int r = j + 1; r < kb; ++r) {
                double s = D[j * ld + r];
                for (int k = 0; k < j; ++k) s -= D[k * ld + r] * D[k * ld + j];
                D[j * ld + r] = s / d;
            }
        }
        double* out = inv.data() + (int64_t)t.slot * TS * TS;
        std::fill(out, out + TS * TS, 0.0);
        for (int c = 0; c < kb; ++c) {
            out[c * TS + c] = 1.0 / D[c * ld + c];
            for (int i = c + 1; i < kb; ++i) {
                double s = 0;
                for (int k = c; k < i; ++k) s += D[k * ld + i] * out[c * TS + k];
                out[c * TS + i] = -s / D[i * ld + i];
            }
        }
    };
    // a fused task: the children's extend-add split by destination (panel columns first, U
    // columns after the Schur update), tile POTRF, TRSM with the tile inverse, U = -L21 L21^T
    auto fused = [&](const FusedTask& f) {
        auto children = [&](bool to_u) {
            for (int32_t q = f.child_begin; q < f.child_end; ++q) {
                const int32_t c = S.children[q];
                for (int32_t kj = 0; kj < S.sdeg[c]; ++kj) {
                    if ((S.rel[S.rel_off[c] + kj] >= f.wblk) != to_u) continue;
                    ExtendTask t;
                    t.u_child = S.u_off[c];
                    t.p_panel = f.panel_off;
                    t.u_parent = f.u_off;
                    t.dc = S.sdeg[c];
                    t.ldp = f.W + f.D;
                    t.wp = f.wblk;
                    t.dp = f.D / 3;
                    t.rel_off = S.rel_off[c];
                    t.col_block = kj;
                    t.row0 = 3 * kj;
                    extend_rows(t, 3 * kj, 3 * (int64_t)t.dc);
                }
            }
        };
        children(false);
        potrf(S.potrf_tasks[f.slot]);
        const int64_t ld = f.W + f.D;
        double* P = L.data() + f.panel_off;
        const double* X = inv.data() + (int64_t)f.slot * TS * TS;
        std::vector<double> row(f.W);
        for (int32_t r = f.W; r < f.W + f.D; ++r) {
            for (int k = 0; k < f.W; ++k) row[k] = P[(int64_t)k * ld + r];
            for (int c = 0; c < f.W; ++c) {
                double v = 0;
                for (int k = 0; k <= c; ++k) v += row[k] * X[k * TS + c];
                P[(int64_t)c * ld + r] = v;
            }
        }
        if (f.D > 0) {
            double* Us = U.data() + f.u_off;
            for (int32_t j = 0; j < f.D; ++j)
                for (int32_t i = j; i < f.D; ++i) {
                    double v = 0;
                    for (int k = 0; k < f.W; ++k) v += P[(int64_t)k * ld + f.W + i] * P[(int64_t)k * ld + f.W + j];
                    Us[(int64_t)j * f.D + i] = -v;
                }
        }
        children(true);
    };
    auto add_lists = [&](double* C, int64_t ldc, int M, int N, int32_t e0, int32_t n) {
        for (int32_t e = e0; e < e0 + n; ++e) {
            const PullEnt& p = S.pull_ents[e];
            const int32_t* rl = S.rel.data() + p.rel_off;
            for (int32_t jb = p.jb0; jb < p.jb0 + p.njb; ++jb)
                for (int32_t ib = std::max(jb, p.ib0); ib < p.ib0 + p.nib; ++ib)
                    for (int cs = 0; cs < 3; ++cs)
                        for (int r = (ib == jb ? cs : 0); r < 3; ++r) {
                            const int64_t row = 3 * (int64_t)rl[ib] + r - p.base_r;
                            const int64_t col = 3 * (int64_t)rl[jb] + cs - p.base_c;
                            if (row < 0 || row >= M || col < 0 || col >= N) continue;
                            C[col * ldc + row] += U[p.u_child + (int64_t)(3 * jb + cs) * p.dc3 + 3 * ib + r];
                        }
        }
    };
    auto pull_tile = [&](const GemmTask& t) {
        add_lists(U.data() + t.c_off, t.ldc, t.M, t.N, t.pad, t.flags >> GT_PULL_SHIFT);
    };
    for (const LevelPlan& P : S.levels) {
        for (int64_t q = P.fused.first; q < P.fused.second; ++q) fused(S.fused_tasks[q]);
        if (ppull) {
            for (int64_t q = P.panel_pulls.first; q < P.panel_pulls.second; ++q) {
                const PanelPullTask& t = S.panel_pull_tasks[q];
                // the tile rebuilt from its injection entries (zero elsewhere), then the lists
                double* C = L.data() + t.c_off;
                for (int c = 0; c < t.N; ++c)
                    for (int r = 0; r < t.M; ++r) C[(int64_t)c * t.ldc + r] = 0.0;
                for (int32_t i = t.inj_begin; i < t.inj_begin + t.ninj; ++i) {
                    const TileInj& ti = S.tile_inj[i];
                    inject_entry((size_t)ti.plan, [&](int r, int c, double v) {
                        const int lr = ti.lr + r, lc = ti.lc + c;
                        if (lr >= 0 && lr < t.M && lc >= 0 && lc < t.N) C[(int64_t)lc * t.ldc + lr] = v;
                    });
                }
                add_lists(C, t.ldc, t.M, t.N, t.pull_begin, t.npull);
            }
        } else {
            for (const auto& pass : P.extend_passes)
                for (int64_t q = pass.first; q < pass.second; ++q) extend_one(S.extend_tasks[q]);
        }
        for (size_t st = 0; st < P.potrf_steps.size(); ++st) {
            for (int64_t q = P.potrf_steps[st].first; q < P.potrf_steps[st].second; ++q)
                potrf(S.potrf_tasks[q]);
            for (int64_t q = P.trsm_steps[st].b; q < P.trsm_steps[st].e; ++q) gemm(S.gemm_tasks[q]);
            for (int64_t q = P.update_steps[st].b; q < P.update_steps[st].e; ++q) gemm(S.gemm_tasks[q]);
            for (int64_t q = P.rest_steps[st].b; q < P.rest_steps[st].e; ++q) gemm(S.gemm_tasks[q]);
        }
        for (int64_t q = P.syrk.b; q < P.syrk.e; ++q) {
            gemm(S.gemm_tasks[q]);
            if (pull) pull_tile(S.gemm_tasks[q]);
        }
        if (!pull)
            for (const auto& pass : P.extend_u_passes)
                for (int64_t q = pass.first; q < pass.second; ++q) extend_one(S.extend_tasks[q]);
    }
    return status == INT64_MAX ? -1 : status;
}

} // namespace mfs

namespace mfs {

void panels_to_csc(const Symbolic& S, const std::vector<double>& h, int64_t* Lp, int64_t* Li, double* Lx) {
    // columns and rows in factor order over the free scalars (masked ones dropped: they hold the
    // identity), n_free columns
    const int64_t ni = 3 * S.n_red, n = S.n_free;
    std::vector<int64_t> uidx(ni, -1);
    for (int64_t u = 0; u < n; ++u) uidx[S.uperm_to_int[u]] = u;
    std::vector<int64_t> cnt(n + 1, 0);
    for (int32_t sn = 0; sn < S.ns; ++sn) {
        const int64_t W = 3 * (int64_t)S.swidth[sn], D = 3 * (int64_t)S.sdeg[sn];
        const int64_t f = 3 * (int64_t)S.sfirst[sn];
        int64_t struct_free = 0;
        for (int64_t r = 0; r < D; ++r) struct_free += uidx[3 * (int64_t)S.rows[S.srows_off[sn] + r / 3] + r % 3] >= 0;
        int64_t below = 0;
        for (int64_t c = W - 1; c >= 0; --c)
            if (uidx[f + c] >= 0) cnt[uidx[f + c] + 1] = ++below + struct_free;
    }
    for (int64_t j = 0; j < n; ++j) cnt[j + 1] += cnt[j];
    for (int64_t j = 0; j <= n; ++j) Lp[j] = cnt[j];
    for (int32_t sn = 0; sn < S.ns; ++sn) {
        const int64_t W = 3 * (int64_t)S.swidth[sn], D = 3 * (int64_t)S.sdeg[sn], ld = W + D;
        const int64_t f = 3 * (int64_t)S.sfirst[sn];
        for (int64_t c = 0; c < W; ++c) {
            if (uidx[f + c] < 0) continue;
            int64_t q = Lp[uidx[f + c]];
            for (int64_t r = c; r < W + D; ++r) {
                const int64_t row = r < W ? f + r : 3 * (int64_t)S.rows[S.srows_off[sn] + (r - W) / 3] + (r - W) % 3;
                if (uidx[row] < 0) continue;
                Li[q] = uidx[row];
                Lx[q] = h[S.panel_off[sn] + c * ld + r];
                ++q;
            }
        }
    }
}

} // namespace mfs

namespace mfs {

// Host replay of the device solve schedule on a factor-order vector x (one RHS), using the
// panels L and the stored tile inverses `inv` produced by emulate_numeric_full.
// Block inverses X_b = L_bb^{-1} by the tile recurrence the device uses: X_jj = (tile inverse),
// X_ij = -X_ii sum_{m=j}^{i-1} L_im X_mj.
void emulate_block_inverses(const Symbolic& S, const std::vector<double>& L, const std::vector<double>& inv,
                            std::vector<double>& Xb) {
    const int TS = 64;
    Xb.assign(S.xb_size, 0.0);
    for (const SolveBlock& b : S.sblocks) {
        const int nb = b.nb, nt = (nb + TS - 1) / TS;
        double* X = Xb.data() + b.xb_off;
        const double* Lb = L.data() + b.l_off;
        for (int j = 0; j < nt; ++j) {
            const int cj = std::min(TS, nb - TS * j);
            const double* Vj = inv.data() + (int64_t)S.tile_slot[b.ts_off + j] * TS * TS;
            for (int c = 0; c < cj; ++c)
                for (int r = c; r < cj; ++r) X[(int64_t)(TS * j + c) * nb + TS * j + r] = Vj[c * TS + r];
            for (int i = j + 1; i < nt; ++i) {
                const int ri = std::min(TS, nb - TS * i);
                std::vector<double> acc((size_t)TS * TS, 0.0);   // acc[c * TS + r]
                for (int m = j; m < i; ++m) {
                    const int cm = std::min(TS, nb - TS * m);
                    for (int c = 0; c < cj; ++c)
                        for (int k = 0; k < cm; ++k) {
                            const double xv = X[(int64_t)(TS * j + c) * nb + TS * m + k];
                            for (int r = 0; r < ri; ++r) acc[c * TS + r] += Lb[(int64_t)(TS * m + k) * b.ld + TS * i + r] * xv;
                        }
                }
                const double* Vi = inv.data() + (int64_t)S.tile_slot[b.ts_off + i] * TS * TS;
                for (int c = 0; c < cj; ++c)
                    for (int r = 0; r < ri; ++r) {
                        double v = 0;
                        for (int k = 0; k <= r; ++k) v += Vi[k * TS + r] * acc[c * TS + k];
                        X[(int64_t)(TS * j + c) * nb + TS * i + r] = -v;
                    }
            }
        }
    }
}

// Host replay of the device solve schedule (one right-hand side, factor order, in place).
void emulate_solve(const Symbolic& S, const std::vector<double>& L, const std::vector<double>& Xb,
                   std::vector<double>& x) {
    const int TS = 64;
    const int SB = S.solve_block;
    std::vector<double> u(std::max<int64_t>(S.u_solve_size, 1), 0.0), Y(x.size(), 0.0);
    std::vector<double> part((size_t)S.bpart_size * SB, 0.0);
    for (const LevelPlan& P : S.levels) {
        for (int64_t q = P.gather.first; q < P.gather.second; ++q) {
            const SolveGatherTask& t = S.gather_tasks[q];
            std::fill(u.begin() + t.u_off, u.begin() + t.u_off + t.d, 0.0);
            for (int32_t k = t.child_begin; k < t.child_end; ++k) {
                const int32_t c = S.children[k];
                const int Dc = 3 * S.sdeg[c];
                for (int i = 0; i < Dc; ++i) {
                    const int tgt = 3 * S.rel[S.rel_off[c] + i / 3] + i % 3;
                    const double v = u[S.u_solve_off[c] + i];
                    if (tgt < t.w) x[t.x_off + tgt] -= v;
                    else u[t.u_off + tgt - t.w] += v;
                }
            }
        }
        for (size_t st = 0; st < P.fdiag_steps.size(); ++st) {
            for (int64_t q = P.fdiag_steps[st].first; q < P.fdiag_steps[st].second; ++q) {
                const SolveTile& tl = S.fdiag_tiles[q];
                const SolveBlock& b = S.sblocks[tl.blk];
                const int r0 = TS * tl.t, nr = std::min(TS, b.nb - r0), nc = std::min(b.nb, r0 + TS);
                for (int r = 0; r < nr; ++r) {
                    double v = 0;
                    for (int c = 0; c < nc; ++c) v += Xb[b.xb_off + (int64_t)c * b.nb + r0 + r] * x[b.x_off + c];
                    Y[b.x_off + r0 + r] = v;
                }
            }
            for (int64_t q = P.fupd_steps[st].first; q < P.fupd_steps[st].second; ++q) {
                const SolveRowTask& t = S.fupd_tasks[q];
                for (int r = 0; r < t.nr; ++r) {
                    double v = 0;
                    for (int c = 0; c < t.nc; ++c) v += L[t.l_off + (int64_t)c * t.ld + r] * Y[t.v_off + c];
                    if (t.mode == 0) x[t.d_off + r] -= v;
                    else u[t.d_off + r] += v;
                }
            }
        }
    }
    for (int lv = (int)S.levels.size() - 1; lv >= 0; --lv) {
        const LevelPlan& P = S.levels[lv];
        for (int64_t q = P.gather.first; q < P.gather.second; ++q) {
            const SolveGatherTask& t = S.gather_tasks[q];
            for (int i = 0; i < t.d; ++i) u[t.u_off + i] = x[3 * (int64_t)S.rows[t.rows_off + i / 3] + i % 3];
        }
        for (size_t st = 0; st < P.bdiag_steps.size(); ++st) {
            for (int64_t q = P.bpart_steps[st].first; q < P.bpart_steps[st].second; ++q) {
                const SolveRowTask& t = S.bpart_tasks[q];
                const double* src = (t.mode == 0 ? x.data() : u.data()) + t.v_off;
                for (int c = 0; c < t.nc; ++c) {
                    double acc = 0;
                    for (int r = 0; r < t.nr; ++r) acc += L[t.l_off + (int64_t)c * t.ld + r] * src[r];
                    part[t.d_off + c] = acc;
                }
            }
            for (int64_t q = P.bdiag_steps[st].first; q < P.bdiag_steps[st].second; ++q) {
                const SolveTile& tl = S.bdiag_tiles[q];
                const SolveBlock& b = S.sblocks[tl.blk];
                const int c0 = TS * tl.t, nc = std::min(TS, b.nb - c0);
                for (int c = c0; c < c0 + nc; ++c) {
                    double v = 0;
                    for (int r = c0; r < b.nb; ++r) {
                        double z = Y[b.x_off + r];
                        for (int p = b.part_begin; p < b.part_end; ++p) z -= part[(int64_t)p * SB + r];
                        v += Xb[b.xb_off + (int64_t)c * b.nb + r] * z;
                    }
                    x[b.x_off + c] = v;
                }
            }
        }
    }
}

} // namespace mfs
