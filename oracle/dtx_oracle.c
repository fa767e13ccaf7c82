/*
 * dtx_oracle.c -- FP64 CPU restatement of DiplodocusTransport.jl's explicit-Euler hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path (see dtx_oracle.h).  It is never
 * linked into, loaded by or called from the product library.
 *
 * Julia semantics reproduced here (they come from Julia Base, not from the reference):
 *   - `x * (cond)` with a Bool is a "strong zero": true*x == x (NaN stays NaN),
 *     false*x == copysign(0.0, x).  So `f*(f>=thr)` maps negatives to -0.0, NaN to +-0.0 and
 *     small positives to +0.0.  Implemented with a select + copysign, never a multiply.
 *   - `1f-10` / `1f-20` are Float32 literals: the comparison promotes them exactly to
 *     Float64, i.e. 1.00000001335143196e-10 and 9.99999968e-21.
 *   - `@. a += b + c*s` is `a = a + (b + (c*s))`; `@. df *= invAp*s` is `df = df*(invAp*s)`;
 *     `-x*y*z` is `((-x)*y)*z`.  No fused multiply-add anywhere (built with
 *     -ffp-contract=off).
 *   - SparseMatrixCSC * Vector (SparseArrays' generic kernel) accumulates column by column:
 *     y[rowval[k]] += nzval[k] * x[j]; every row therefore sums its terms in ascending
 *     column order.
 *   - Dense Matrix * Vector is taken in the same column order (y += A[:,j] * x[j]), which
 *     fixes the per-row summation order as ascending j.
 */
#include "dtx_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Float32 literals of the reference, promoted exactly to Float64. */
#define THR_1F_10 ((double)1e-10f)
#define THR_1F_20 ((double)1e-20f)

struct dtxo_ws {
    int64_t n, N;
    double* f;
    double* df;
    double* df_flux;   /* df_Flux (ForwardEuler) or df_PFlux (symplectic) */
    double* df_emi;
    double* df_bin;    /* persistent across steps, zero-initialised (SteppingStructs.jl:94) */
    double* df_tmp;
    double* T;         /* M_Bin_Mul_Step: N^2 (SteppingStructs.jl:90-91) */
    /* symplectic extras (SteppingStructs.jl:214-224) */
    double* df_mom;
    double* df_space;
    double* df_xflux;
    double* f_mom;
    double* f_space;
    double* pre;       /* the state before the final clamps of the last step */
    /* channel form evaluated for all cells at once: operator-row blocks x cells */
    double* Tb;        /* rows x CELL_BATCH of the current channel          */
    double* Fb;        /* f of the batch's cells, state-major: [N][CELL_BATCH] */
    int64_t Tb_rows;
};

/* Cells evaluated together in the batched channel form: each operator row is read once per
 * batch of this many cells. */
#define CELL_BATCH 64

int dtxo_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

static double* zalloc(int64_t count) {
    double* p = (double*)calloc((size_t)(count > 0 ? count : 1), sizeof(double));
    if (!p) {
        fprintf(stderr, "dtx_oracle: out of memory (%lld doubles)\n", (long long)count);
        abort();
    }
    return p;
}

dtxo_ws* dtxo_ws_create(const dtxo_problem* p) {
    dtxo_ws* w = (dtxo_ws*)calloc(1, sizeof(dtxo_ws));
    int64_t n = p->N * p->C;
    w->n = n;
    w->N = p->N;
    w->f = zalloc(n);
    w->df = zalloc(n);
    w->df_flux = zalloc(n);
    w->df_emi = zalloc(n);
    w->df_bin = zalloc(n);
    w->df_tmp = zalloc(n);
    w->T = zalloc(p->N * p->N);
    w->df_mom = zalloc(n);
    w->df_space = zalloc(n);
    w->df_xflux = zalloc(n);
    w->f_mom = zalloc(n);
    w->f_space = zalloc(n);
    w->pre = zalloc(n);
    if (p->batched && p->M_Bin == NULL) {
        int64_t rows = 0;
        for (int32_t g = 0; g < p->n_gain; ++g) {
            int64_t r = p->species_size[p->gain[g].s_out] * p->species_size[p->gain[g].s1];
            if (r > rows) rows = r;
        }
        w->Tb_rows = rows;
        w->Tb = zalloc(rows * CELL_BATCH);
        w->Fb = zalloc(p->N * CELL_BATCH);
    }
    return w;
}

void dtxo_ws_destroy(dtxo_ws* w) {
    if (!w) return;
    free(w->f); free(w->df); free(w->df_flux); free(w->df_emi); free(w->df_bin);
    free(w->df_tmp); free(w->T); free(w->df_mom); free(w->df_space); free(w->df_xflux);
    free(w->f_mom); free(w->f_space); free(w->pre); free(w->Tb); free(w->Fb);
    free(w);
}

double* dtxo_ws_f(dtxo_ws* w) { return w->f; }
const double* dtxo_ws_df_bin(dtxo_ws* w) { return w->df_bin; }
const double* dtxo_ws_pre_clamp(dtxo_ws* w) { return w->pre; }

/* Julia's Bool strong-zero product  x * (cond). */
static inline double bool_mul(double x, int cond) { return cond ? x : copysign(0.0, x); }

/* Julia's sign(x) for Float64. */
static inline double jl_sign(double x) {
    if (x > 0) return 1.0;
    if (x < 0) return -1.0;
    return x; /* +-0.0 and NaN */
}

/* mul!(y, A::SparseMatrixCSC, x)  (SparseArrays generic kernel, column order). */
static void csc_mul(const dtxo_csc* A, int64_t n, const double* x, double* y) {
    memset(y, 0, (size_t)n * sizeof(double));
    if (A->ncols == 0) return;
    for (int64_t col = 0; col < A->ncols; ++col) {
        double xj = x[col];
        for (int64_t k = A->colptr[col] - 1; k < A->colptr[col + 1] - 1; ++k) {
            int64_t r = A->rowval[k] - 1;
            y[r] = y[r] + A->nzval[k] * xj;
        }
    }
}

/* mul!(y, A::Matrix, x) with A rows x cols column-major; per-row order ascending column.
 * Parallel over row blocks: every row's sum keeps the same sequential order. */
static void dense_mul(const double* A, int64_t rows, int64_t cols, const double* x, double* y) {
    const int64_t RB = 4096;
    int64_t nblk = (rows + RB - 1) / RB;
#pragma omp parallel for schedule(static)
    for (int64_t b = 0; b < nblk; ++b) {
        int64_t r0 = b * RB, r1 = r0 + RB < rows ? r0 + RB : rows;
        for (int64_t r = r0; r < r1; ++r) y[r] = 0.0;
        for (int64_t j = 0; j < cols; ++j) {
            double xj = x[j];
            const double* Aj = A + j * rows;
            for (int64_t r = r0; r < r1; ++r) y[r] = y[r] + Aj[r] * xj;
        }
    }
}

static int in_domain(const dtxo_problem* p, int64_t off) {
    for (int64_t d = 0; d < p->n_domain; ++d)
        if (p->domain[d] == off) return 1;
    return 0;
}

static int has_emission(const dtxo_problem* p) { return p->M_Emi != NULL || p->Emi.ncols > 0; }

/* mul!(df_Emi, M_Emi, f) with the dense or the sparse M_Emi */
static void emission_mul(const dtxo_problem* p, int64_t n, const double* f, double* y) {
    if (p->Emi.ncols > 0) csc_mul(&p->Emi, n, f, y);
    else dense_mul(p->M_Emi, n, n, f, y);
}

static int has_binary(const dtxo_problem* p) {
    return p->M_Bin != NULL || p->n_gain > 0 || p->n_loss > 0;
}

/* Channel form of one cell's binary term (same mathematics as the monolithic tensor:
 * GainMatrix_to_M_Bin! writes G/2 at (k,i,j) and (k,j,i); LossMatrix_to_M_Bin! writes
 * -L/2 at (a,a,c) and (a,c,a); src/Collisions/BuildBinaryMatrices.jl:181-196,261-276). */
static void channel_cell(const dtxo_problem* p, const double* fc, double* dB, double* T) {
    int64_t N = p->N;
    for (int64_t k = 0; k < N; ++k) dB[k] = 0.0;
    for (int32_t g = 0; g < p->n_gain; ++g) {
        const dtxo_gain* ch = &p->gain[g];
        int64_t No = p->species_size[ch->s_out], N1 = p->species_size[ch->s1],
                N2 = p->species_size[ch->s2];
        const double* f1 = fc + p->species_offset[ch->s1];
        const double* f2 = fc + p->species_offset[ch->s2];
        double* out = dB + p->species_offset[ch->s_out];
        /* T[k,i] = sum_j G[k,i,j] f2_j  (first GEMV), then out[k] += sum_i T[k,i] f1_i */
        int64_t rows = No * N1;
        dense_mul(ch->G, rows, N2, f2, T);
        for (int64_t k = 0; k < No; ++k) {
            double acc = 0.0;
            for (int64_t i = 0; i < N1; ++i) acc = acc + T[k + No * i] * f1[i];
            out[k] = out[k] + acc;
        }
    }
    for (int32_t l = 0; l < p->n_loss; ++l) {
        const dtxo_loss* ch = &p->loss[l];
        int64_t N1 = p->species_size[ch->s1], N2 = p->species_size[ch->s2];
        const double* f1 = fc + p->species_offset[ch->s1];
        const double* f2 = fc + p->species_offset[ch->s2];
        double* out = dB + p->species_offset[ch->s1];
        for (int64_t a = 0; a < N1; ++a) {
            double acc = 0.0;
            for (int64_t c = 0; c < N2; ++c) acc = acc + ch->L[a + N1 * c] * f2[c];
            out[a] = out[a] - f1[a] * acc;
        }
    }
}

/* channel_cell for a batch of cells (cells[0..nc), nc <= CELL_BATCH) with the same arithmetic
 * per (row, cell): T[r] = sum_j G[r, j] f2_j from 0.0 in ascending j, then
 * out[k] = out[k] + (sum_i T[k + No i] f1_i from 0.0 in ascending i), channels in list order,
 * then the losses.  Only the loop nest differs: each operator row block is read once for the
 * batch instead of once per cell. */
static void channel_batch(const dtxo_problem* p, dtxo_ws* w, const int64_t* cells, int64_t nc) {
    const int64_t N = p->N;
    double* Fb = w->Fb;  /* Fb[k * CELL_BATCH + c] = f of cell c, state k */
    for (int64_t c = 0; c < nc; ++c)
        for (int64_t k = 0; k < N; ++k) {
            Fb[k * CELL_BATCH + c] = w->f[cells[c] * N + k];
            w->df_bin[cells[c] * N + k] = 0.0;
        }
    for (int32_t g = 0; g < p->n_gain; ++g) {
        const dtxo_gain* ch = &p->gain[g];
        const int64_t No = p->species_size[ch->s_out], N1 = p->species_size[ch->s1],
                      N2 = p->species_size[ch->s2];
        const int64_t o1 = p->species_offset[ch->s1], o2 = p->species_offset[ch->s2],
                      oo = p->species_offset[ch->s_out];
        const int64_t rows = No * N1, RB = 256;
        double* Tb = w->Tb;
#pragma omp parallel for schedule(static)
        for (int64_t b = 0; b < (rows + RB - 1) / RB; ++b) {
            const int64_t r0 = b * RB, r1 = r0 + RB < rows ? r0 + RB : rows;
            for (int64_t r = r0; r < r1; ++r)
                for (int64_t c = 0; c < nc; ++c) Tb[r * CELL_BATCH + c] = 0.0;
            for (int64_t j = 0; j < N2; ++j) {
                const double* Gj = ch->G + rows * j;
                const double* fj = Fb + (o2 + j) * CELL_BATCH;
                for (int64_t r = r0; r < r1; ++r) {
                    const double a = Gj[r];
                    double* t = Tb + r * CELL_BATCH;
                    for (int64_t c = 0; c < nc; ++c) t[c] = t[c] + a * fj[c];
                }
            }
        }
#pragma omp parallel for schedule(static)
        for (int64_t k = 0; k < No; ++k) {
            double acc[CELL_BATCH];
            for (int64_t c = 0; c < nc; ++c) acc[c] = 0.0;
            for (int64_t i = 0; i < N1; ++i) {
                const double* t = Tb + (k + No * i) * CELL_BATCH;
                const double* fi = Fb + (o1 + i) * CELL_BATCH;
                for (int64_t c = 0; c < nc; ++c) acc[c] = acc[c] + t[c] * fi[c];
            }
            for (int64_t c = 0; c < nc; ++c) {
                double* out = w->df_bin + cells[c] * N + oo + k;
                *out = *out + acc[c];
            }
        }
    }
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < nc; ++c)
        for (int32_t l = 0; l < p->n_loss; ++l) {
            const dtxo_loss* ch = &p->loss[l];
            const int64_t N1 = p->species_size[ch->s1], N2 = p->species_size[ch->s2];
            const int64_t o1 = p->species_offset[ch->s1], o2 = p->species_offset[ch->s2];
            double* out = w->df_bin + cells[c] * N + o1;
            for (int64_t a = 0; a < N1; ++a) {
                double acc = 0.0;
                for (int64_t q = 0; q < N2; ++q)
                    acc = acc + ch->L[a + N1 * q] * Fb[(o2 + q) * CELL_BATCH + c];
                out[a] = out[a] - Fb[(o1 + a) * CELL_BATCH + c] * acc;
            }
        }
}

/* update_Big_Bin!  src/SteppingMethods.jl:332-397 */
static void update_big_bin(const dtxo_problem* p, dtxo_ws* w) {
    int64_t N = p->N;
    if (p->batched && p->M_Bin == NULL) {
        int64_t cells[CELL_BATCH], nc = 0;
        for (int64_t off = 0; off < p->C; ++off) {
            const double* fv = w->f + off * N;
            int all_zero = 1;
            for (int64_t k = 0; k < N; ++k)
                if (fv[k] != 0.0) { all_zero = 0; break; }
            if (!(p->n_domain < 0 || in_domain(p, off) || all_zero)) continue;  /* :369 */
            cells[nc++] = off;
            if (nc == CELL_BATCH || off == p->C - 1) {
                channel_batch(p, w, cells, nc);
                nc = 0;
            }
        }
        if (nc) channel_batch(p, w, cells, nc);
        for (int64_t off = 0; off < p->C; ++off) {
            const double* fv = w->f + off * N;
            int all_zero = 1;
            for (int64_t k = 0; k < N; ++k)
                if (fv[k] != 0.0) { all_zero = 0; break; }
            if (!(p->n_domain < 0 || in_domain(p, off) || all_zero)) continue;
            double* dB = w->df_bin + off * N;
            const double vol = p->Vol[off];
            for (int64_t k = 0; k < N; ++k) dB[k] = dB[k] * vol;  /* :380 */
        }
        return;
    }
    for (int64_t off = 0; off < p->C; ++off) {          /* x, y, z loop, z fastest (:360-362) */
        const double* fv = w->f + off * N;              /* fView (:364-367) */
        int all_zero = 1;
        for (int64_t k = 0; k < N; ++k)
            if (fv[k] != 0.0) { all_zero = 0; break; }  /* isnothing(findfirst(!iszero, fView)) */
        if (!(p->n_domain < 0 || in_domain(p, off) || all_zero)) continue;  /* :369 */
        double* dB = w->df_bin + off * N;
        if (p->M_Bin) {
            dense_mul(p->M_Bin, N * N, N, fv, w->T);    /* :372  T(N^2) = M_Bin * fView   */
            dense_mul(w->T, N, N, fv, dB);              /* :373  dB = reshape(T,N,N) * fView */
        } else {
            channel_cell(p, fv, dB, w->T);
        }
        double vol = p->Vol[off];
        for (int64_t k = 0; k < N; ++k) dB[k] = dB[k] * vol;  /* :380 */
    }
}

void dtxo_binary_term(const dtxo_problem* p, const double* f, double* dB) {
    dtxo_ws* w = dtxo_ws_create(p);
    memcpy(w->f, f, (size_t)(p->N * p->C) * sizeof(double));
    update_big_bin(p, w);
    memcpy(dB, w->df_bin, (size_t)(p->N * p->C) * sizeof(double));
    dtxo_ws_destroy(w);
}

static double seq_sum(const double* x, int64_t n) {
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s = s + x[i];
    return s;
}

/* -mapreduce(x -> isnan(x) ? Inf : x, min, num ./ den) */
static double neg_min_ratio(const double* num, const double* den, int64_t n) {
    double m = INFINITY;
    for (int64_t i = 0; i < n; ++i) {
        double r = num[i] / den[i];
        if (isnan(r)) r = INFINITY;
        if (r < m || (r == m && signbit(r))) m = r;
    }
    return -m;
}

static void clip_dt(double t_start, double t_stop, double* dt, int* save) {
    /* src/SteppingMethods.jl:18-23 */
    if (t_start + *dt >= t_stop) {
        *dt = t_stop - t_start;
        *save = 1;
    } else {
        *save = 0;
    }
}

static void report_cr(int verbose, const dtxo_stats* st) {
    if (verbose == 1 && st->cr > 1.0)
        printf("Cr = %g, t=%.17g, dt=%.17g, system may be unstable\n", st->cr, st->t_start, st->dt);
    else if (verbose == 2)
        printf("\rCr = %g, t=%.17g, dt=%.17g\n", st->cr, st->t_start, st->dt);
    else if (verbose == 3)
        printf("Cr = %g,Cr_Bin = %g, Cr_Emi = %g, Cr_Flux = %g, t=%.17g, dt=%.17g\n", st->cr,
               st->cr_bin, st->cr_emi, st->cr_flux, st->t_start, st->dt);
}

/* ForwardEuler functor  src/SteppingMethods.jl:12-107 */
int dtxo_forward_euler(const dtxo_problem* p, dtxo_ws* w, double t_start, double t_stop,
                       double dt, int verbose, double* dt_out, int* save_out, dtxo_stats* st) {
    int64_t n = p->N * p->C;
    int save;
    double dt0 = p->dt0;
    clip_dt(t_start, t_stop, &dt, &save);
    double s = dt / dt0;                                         /* :27 */

    csc_mul(&p->F, n, w->f, w->df_flux);                         /* :31 */
    for (int64_t i = 0; i < n; ++i) w->df[i] = -w->df_flux[i];   /* :32 */
    int emission = has_emission(p);
    int binary = has_binary(p);
    if (emission) {                                              /* :35-38 */
        emission_mul(p, n, w->f, w->df_emi);
        for (int64_t i = 0; i < n; ++i) w->df[i] = w->df[i] + w->df_emi[i];
    }
    if (binary) {                                                /* :41-44 */
        update_big_bin(p, w);
        for (int64_t i = 0; i < n; ++i) w->df[i] = w->df[i] + w->df_bin[i];
    }
    for (int64_t i = 0; i < n; ++i) w->df[i] = w->df[i] * (p->invAp[i] * s);  /* :46 */

    dtxo_stats local;
    memset(&local, 0, sizeof(local));
    local.finite = isfinite(seq_sum(w->df, n)) ? 1 : 0;         /* :48 */
    if (!local.finite) printf("non-finite value in df calculation\n");
    local.dt = dt;
    local.t_start = t_start;
    local.save = save;
    if (verbose == 1 || verbose == 2 || verbose == 3) {          /* :52-94 */
        if (seq_sum(w->f, n) != 0.0) {
            local.cr_evaluated = 1;
            if (verbose == 3) {
                if (binary) local.cr_bin = neg_min_ratio(w->df_bin, w->f, n);
                if (emission) local.cr_emi = neg_min_ratio(w->df_emi, w->f, n);
                for (int64_t i = 0; i < n; ++i) w->df_tmp[i] = -w->df_flux[i];
                local.cr_flux = neg_min_ratio(w->df_tmp, w->f, n);
            }
            local.cr = neg_min_ratio(w->df, w->f, n);
        }
        report_cr(verbose, &local);
    }
    for (int64_t i = 0; i < n; ++i) {                            /* :97-101 */
        double x = w->f[i] + (w->df[i] + p->df_inj[i] * s);
        w->pre[i] = x;
        x = bool_mul(x, x >= THR_1F_10);
        x = bool_mul(x, x != INFINITY);
        w->f[i] = x;
    }
    if (st) *st = local;
    *dt_out = dt;
    *save_out = save;
    return 0;
}

/* ForwardSymplecticEuler functor  src/SteppingMethods.jl:120-243 */
int dtxo_forward_symplectic_euler(const dtxo_problem* p, dtxo_ws* w, double t_start,
                                  double t_stop, double dt, int verbose, double* dt_out,
                                  int* save_out, dtxo_stats* st) {
    int64_t n = p->N * p->C;
    int save;
    clip_dt(t_start, t_stop, &dt, &save);                        /* :126-131 */
    double s = dt / p->dt0;                                      /* :135 */
    int emission = has_emission(p);
    int binary = has_binary(p);

    csc_mul(&p->P, n, w->f, w->df_flux);                         /* :140 df_PFlux */
    for (int64_t i = 0; i < n; ++i) w->df_mom[i] = -w->df_flux[i];
    if (emission) {
        emission_mul(p, n, w->f, w->df_emi);
        for (int64_t i = 0; i < n; ++i) w->df_mom[i] = w->df_mom[i] + w->df_emi[i];
    }
    if (binary) {
        update_big_bin(p, w);
        for (int64_t i = 0; i < n; ++i) w->df_mom[i] = w->df_mom[i] + w->df_bin[i];
    }
    for (int64_t i = 0; i < n; ++i) w->df_mom[i] = w->df_mom[i] * (p->invAp[i] * s);  /* :155 */
    for (int64_t i = 0; i < n; ++i) {                            /* :158-160 */
        double x = w->f[i] + w->df_mom[i];
        w->f_mom[i] = bool_mul(x, x >= THR_1F_10);
    }
    csc_mul(&p->X, n, w->f_mom, w->df_xflux);                    /* :165 */
    for (int64_t i = 0; i < n; ++i)                              /* :166 ((-dX)*invAp)*s */
        w->df_space[i] = ((-w->df_xflux[i]) * p->invAp[i]) * s;
    for (int64_t i = 0; i < n; ++i) {                            /* :169-171 */
        double x = w->f_mom[i] + w->df_space[i];
        w->f_space[i] = bool_mul(x, x >= THR_1F_20);
    }
    dtxo_stats local;
    memset(&local, 0, sizeof(local));
    local.finite = 1;
    local.dt = dt;
    local.t_start = t_start;
    local.save = save;
    if (verbose == 1 || verbose == 2 || verbose == 3) {          /* :175-229 */
        if (seq_sum(w->f, n) != 0.0) {
            local.cr_evaluated = 1;
            if (verbose == 3) {
                if (binary) {
                    for (int64_t i = 0; i < n; ++i)
                        w->df_tmp[i] = (w->df_bin[i] * p->invAp[i]) * s;
                    local.cr_bin = neg_min_ratio(w->df_tmp, w->f, n);
                }
                if (emission) {
                    for (int64_t i = 0; i < n; ++i)
                        w->df_tmp[i] = (w->df_emi[i] * p->invAp[i]) * s;
                    local.cr_emi = neg_min_ratio(w->df_tmp, w->f, n);
                }
                for (int64_t i = 0; i < n; ++i) {
                    double r = (((-w->df_flux[i]) * p->invAp[i]) * s) / w->f[i];
                    w->df_tmp[i] = isnan(r) ? 0.0 : r;     /* replace!(df_tmp, NaN=>0.0) */
                }
                /* after replace!(NaN => 0.0) the isnan -> Inf map of the reduction finds no
                 * NaN left, so it is a plain minimum */
                {
                    double m = INFINITY;
                    for (int64_t i = 0; i < n; ++i) {
                        double r = w->df_tmp[i];
                        if (r < m || (r == m && signbit(r))) m = r;
                    }
                    local.cr_flux = -m;
                }
                /* Cr_Momentum = -min(df_Momentum ./ f)                          (:218-219) */
                local.cr_momentum = neg_min_ratio(w->df_mom, w->f, n);
                /* Cr_Space = -min(df_Space ./ f_Momentum)                        (:222-223) */
                local.cr_space = neg_min_ratio(w->df_space, w->f_mom, n);
            }
            for (int64_t i = 0; i < n; ++i) w->df_tmp[i] = w->df_space[i] + w->df_mom[i];
            local.cr = neg_min_ratio(w->df_tmp, w->f, n);
        }
        report_cr(verbose, &local);
    }
    for (int64_t i = 0; i < n; ++i) {                            /* :233-237 */
        double x = w->f_space[i] + p->df_inj[i] * s;
        w->pre[i] = x;
        x = bool_mul(x, x >= THR_1F_20) * jl_sign(x);
        x = bool_mul(x, x != INFINITY);
        w->f[i] = x;
    }
    if (st) *st = local;
    *dt_out = dt;
    *save_out = save;
    return 0;
}

/* Solve  src/Solver.jl:1-64 (without the JLD2 write) */
int dtxo_solve(const dtxo_problem* p, int method, const double* f_init, double dt_initial,
               const double* t_save, int64_t n_save, int verbose, double* f_out,
               double* t_out, dtxo_stats* stats, int64_t max_stats, int64_t* n_steps) {
    int64_t n = p->N * p->C;
    dtxo_ws* w = dtxo_ws_create(p);
    memcpy(f_out, f_init, (size_t)n * sizeof(double));           /* :9-15 */
    t_out[0] = t_save[0];
    memcpy(w->f, f_init, (size_t)n * sizeof(double));            /* :22 */
    double dt = dt_initial;                                       /* :25 */
    int64_t steps = 0;
    int rc = 0;
    for (int64_t i = 1; i < n_save; ++i) {                        /* :27 */
        double t_start = t_save[i - 1], t_stop = t_save[i];
        int save = 0;
        /* a carried dt that collapsed by rounding would loop (practically) forever */
        if (!(dt > 0.0) || (t_stop - t_start) / dt > 1e8) {
            rc = -1;
            break;
        }
        while (!save) {                                           /* :34-37 */
            dtxo_stats st;
            if (method == 0)
                dtxo_forward_euler(p, w, t_start, t_stop, dt, verbose, &dt, &save, &st);
            else
                dtxo_forward_symplectic_euler(p, w, t_start, t_stop, dt, verbose, &dt, &save, &st);
            if (stats && steps < max_stats) stats[steps] = st;
            ++steps;
            t_start += dt;
        }
        memcpy(f_out + i * n, w->f, (size_t)n * sizeof(double));  /* :40-43 */
        t_out[i] = t_stop;
    }
    if (n_steps) *n_steps = steps;
    dtxo_ws_destroy(w);
    return rc;
}

/* N^0 (src/DistributionMoments.jl:31-33) and T^00 (:141-162) of one (cell, species) block.
 * Loop order p, u, h with h innermost, as `for p in 1:p_num, u in 1:u_num, h in 1:h_num`. */
void dtxo_moments(const double* f_block, int64_t p_num, int64_t u_num, int64_t h_num,
                  const double* pr, double m, double* number, double* energy) {
    double N0 = 0.0, T00 = 0.0;
    for (int64_t p = 0; p < p_num; ++p)
        for (int64_t u = 0; u < u_num; ++u)
            for (int64_t h = 0; h < h_num; ++h) {
                double fv = f_block[p + p_num * (u + u_num * h)];
                N0 += fv;
                double p0 = pr[p], p1 = pr[p + 1], w;
                if (m != 0.0)
                    w = (p0 * sqrt(m * m + p0 * p0) - p1 * sqrt(m * m + p1 * p1) +
                         m * m * (asinh(p0 / m) - asinh(p1 / m))) / (2 * (p0 - p1));
                else
                    w = (p0 + p1) / 2;
                T00 += fv * w;
            }
    *number = N0;
    *energy = T00;
}
