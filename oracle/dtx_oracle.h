/*
 * dtx_oracle.h -- FP64 CPU restatement of DiplodocusTransport.jl's explicit-Euler hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  Nothing in the product
 * (the dtx C-ABI library and its Python/Julia mirrors) links, loads or calls it.
 *
 * It restates, in plain C with the reference's own loop structure:
 *   - Solve                      src/Solver.jl:1-64
 *   - ForwardEuler functor       src/SteppingMethods.jl:12-107
 *   - update_Big_Bin!            src/SteppingMethods.jl:332-397 (two GEMVs per cell on the
 *                                monolithic (N^2 x N) column-major tensor)
 *   - ForwardSymplecticEuler     src/SteppingMethods.jl:120-243
 *   - moments N^0 / T^00         src/DistributionMoments.jl:8-42, 94-212
 * with the Julia semantics listed at each function.
 */
#ifndef DTX_ORACLE_H
#define DTX_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* A sparse matrix in Julia's SparseMatrixCSC form: 1-based Int64 colptr (ncols+1) and rowval. */
typedef struct {
    int64_t ncols;
    const int64_t* colptr;
    const int64_t* rowval;
    const double*  nzval;
} dtxo_csc;

/* A gain channel: dB[k] += sum_{i in s1, j in s2} G[k + No*(i + N1*j)] f_i f_j  (k in s_out).
 * A loss channel: dB[a] -= f_a * sum_{c in s2} L[a + N1*c] f_c                 (a in s1).   */
typedef struct {
    int32_t s_out, s1, s2;
    const double* G;
} dtxo_gain;
typedef struct {
    int32_t s1, s2;
    const double* L;
} dtxo_loss;

typedef struct {
    int64_t N;            /* momentum states per cell (all species) */
    int64_t C;            /* spatial cells                          */
    double  dt0;          /* PhaseSpace.Time.dt0                    */
    int32_t num_species;
    int64_t species_offset[16];
    int64_t species_size[16];
    /* F_Flux = X_Flux + P_Flux (ForwardEuler); the symplectic method uses X and P separately. */
    dtxo_csc F, X, P;     /* ncols == 0 means "absent" (treated as an all-zero matrix)   */
    const double* M_Emi;  /* dense n x n column-major, or NULL (Emission_Interactions=false) */
    const double* M_Bin;  /* dense N^2 x N column-major, or NULL                             */
    /* channel form of M_Bin (used when M_Bin == NULL and n_gain+n_loss > 0) */
    int32_t n_gain, n_loss;
    const dtxo_gain* gain;
    const dtxo_loss* loss;
    const int64_t* domain; int64_t n_domain;   /* n_domain < 0: Domain === nothing */
    const double* invAp;  /* n   (1 ./ Ap_Flux)   */
    const double* Vol;    /* C                    */
    const double* df_inj; /* n                    */
    /* Channel form only: 1 = evaluate the channels for all cells at once (each operator row
     * read once per step, every (row, cell) sum in the same order as the per-cell loop, so the
     * result is bitwise the same), 0 = cell by cell as update_Big_Bin! loops. */
    int32_t batched;
    /* M_Emi as SparseMatrixCSC (Emi_sparse, src/Collisions/BuildEmissionMatrices.jl:28-39):
     * used instead of the dense M_Emi when Emi.ncols > 0.  For finite states both give the same
     * df_Emi; a non-finite f reaches every row of the dense product through its stored zeros
     * (0 * Inf = NaN) but only the stored entries' rows of the sparse one. */
    dtxo_csc Emi;
} dtxo_problem;

typedef struct {
    int    finite;        /* isfinite(sum(df))  (ForwardEuler; 1 for the symplectic method)  */
    int    cr_evaluated;  /* Verbose in 1..3 and sum(f) != 0                    */
    double cr, cr_bin, cr_emi;
    double cr_flux;       /* Cr_Flux (ForwardEuler) or Cr_PFlux (symplectic)    */
    double dt;            /* dt actually used (after the clip)                  */
    double t_start;
    int    save;
    double cr_momentum;   /* Cr_Momentum (symplectic, Verbose 3)                */
    double cr_space;      /* Cr_Space    (symplectic, Verbose 3)                */
} dtxo_stats;

typedef struct dtxo_ws dtxo_ws;

dtxo_ws* dtxo_ws_create(const dtxo_problem* p);
void     dtxo_ws_destroy(dtxo_ws* w);
/* f (length n) is the state; it is updated in place. */
double*  dtxo_ws_f(dtxo_ws* w);
/* binary term of the last step (length n), as kept in method.df_Bin */
const double* dtxo_ws_df_bin(dtxo_ws* w);
/* The last step's state before its clamps (length n): f + (df + df_Inj s) for ForwardEuler,
 * f_Space + df_Inj s for the symplectic method.  A NaN there is clamped to a zero whose sign
 * is the NaN's, which depends on the hardware that produced the NaN. */
const double* dtxo_ws_pre_clamp(dtxo_ws* w);

/* One ForwardEuler functor call: returns (dt, save) through dt_out/save_out. */
int dtxo_forward_euler(const dtxo_problem* p, dtxo_ws* w, double t_start, double t_stop,
                       double dt, int verbose, double* dt_out, int* save_out, dtxo_stats* st);
/* One ForwardSymplecticEuler functor call (src/SteppingMethods.jl:120-243). */
int dtxo_forward_symplectic_euler(const dtxo_problem* p, dtxo_ws* w, double t_start,
                                  double t_stop, double dt, int verbose, double* dt_out,
                                  int* save_out, dtxo_stats* st);

/* Solve: f_out is n_save x n (row per save), t_out n_save.  method 0 = ForwardEuler,
 * 1 = ForwardSymplecticEuler.  stats (optional) receives one record per step, up to
 * max_stats; the number of steps taken is returned through n_steps. */
int dtxo_solve(const dtxo_problem* p, int method, const double* f_init, double dt_initial,
               const double* t_save, int64_t n_save, int verbose, double* f_out,
               double* t_out, dtxo_stats* stats, int64_t max_stats, int64_t* n_steps);

/* update_Big_Bin! alone, for the binary-contraction parity tests: dB (length n). */
void dtxo_binary_term(const dtxo_problem* p, const double* f, double* dB);

/* Moments of one (cell, species) block: N^0 = sum f, T^00 = sum f * E_bar(p).
 * pr: p bounds (p_num+1), mass m; layout p fastest, then u, then phi. */
void dtxo_moments(const double* f_block, int64_t p_num, int64_t u_num, int64_t h_num,
                  const double* pr, double m, double* number, double* energy);

/* number of OpenMP threads the oracle uses */
int dtxo_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
