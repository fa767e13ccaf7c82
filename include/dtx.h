/*
 * dtx.h -- C ABI of the MI355X-native explicit-Euler stepping library (libdtx.so).
 *
 * This is the drop-in boundary for DiplodocusTransport.jl's hot path.  The Julia side keeps its
 * API (ForwardEulerStruct / ForwardSymplecticEulerStruct, the functor call, Solve, the moments)
 * and replaces the body of the step with `ccall`s into these entry points.  All arguments are
 * plain C types; host arrays are owned by the caller and copied during the call; no host pointer
 * is retained.  Device memory belongs to the handle and is released by dtx_destroy.
 *
 * Arrays follow Julia's layouts: column-major dense matrices, SparseMatrixCSC with 1-based Int64
 * colptr/rowval, state vectors indexed as in GlobalIndices_To_StateIndex
 * (src/GlobalToStateIndices.jl:7-25): cell-major, species blocks in name_list order, px fastest.
 *
 * Errors: every int-returning call returns DTX_OK (0) or a negative DTX_ERR_* code; the message
 * is available from dtx_last_error(h) (or dtx_last_error(NULL) when creation failed).  As in the
 * reference, a non-finite df or a CFL number above one is not an error: it is reported in the
 * step statistics and printed (src/SteppingMethods.jl:48-50, 87-93).
 *
 * Threading: a handle is not thread-safe; its calls are ordered on one HIP stream.  Different
 * handles may be used from different threads.  The dtx_server_* API is thread-safe.
 */
#ifndef DTX_H
#define DTX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DTX_MAX_SPECIES 16

enum dtx_status {
    DTX_OK = 0,
    DTX_ERR_INVALID = -1,     /* bad argument or shape                                   */
    DTX_ERR_OOM = -2,         /* device or host allocation failed                        */
    DTX_ERR_HIP = -3,         /* a HIP runtime call failed                               */
    DTX_ERR_STATE = -4,       /* call out of order (e.g. stepping before set_vectors)    */
    DTX_ERR_UNSUPPORTED = -5, /* valid input the library does not support                */
    DTX_ERR_BUSY = -6         /* server: request table or memory budget full             */
};

/* Arithmetic of the binary contraction.  f, df, the Euler update, losses and the moments are
 * FP64 in every mode.  With one spatial cell the contraction is a GEMV (no MFMA): the operator
 * is stored in the mode's format (FP64, FP32, or BF16 hi + lo) and its products are summed in
 * FP64 against the FP64 f. */
enum dtx_precision {
    DTX_PREC_FP64 = 0,   /* FP64 MFMA                                                      */
    DTX_PREC_FP32 = 1,   /* FP32 MFMA, FP32 accumulation within a pair block, FP64 across   */
    DTX_PREC_BF16X3 = 2  /* split-BF16 hi/lo, 3 MFMA products, same accumulation as FP32    */
};

enum dtx_method {
    DTX_METHOD_FORWARD_EULER = 0,            /* src/SteppingMethods.jl:12-107   */
    DTX_METHOD_FORWARD_SYMPLECTIC_EULER = 1  /* src/SteppingMethods.jl:120-243  */
};

enum dtx_flux_which {
    DTX_FLUX_F = 0,  /* F_Flux = X_Flux + P_Flux (ForwardEuler, src/SteppingStructs.jl:60)  */
    DTX_FLUX_X = 1,  /* X_Flux (space; symplectic, src/SteppingStructs.jl:170)             */
    DTX_FLUX_P = 2   /* P_Flux (momentum; symplectic, src/SteppingStructs.jl:171)          */
};

typedef struct dtx_handle dtx_handle;

/* Shape and mode of one simulation.  Replaces the globals the reference reads from Main
 * (Backend, Precision, px/py/pz_num_list, x/y/z_num; src/StructsAndDictionaries.jl:70-121,
 * src/SteppingStructs.jl:35-36) and PhaseSpace.Time.dt0 (src/StructsAndDictionaries.jl:55). */
typedef struct {
    int32_t num_species;
    int32_t px_num[DTX_MAX_SPECIES];
    int32_t py_num[DTX_MAX_SPECIES];
    int32_t pz_num[DTX_MAX_SPECIES];
    int32_t x_num, y_num, z_num;
    int32_t precision;   /* enum dtx_precision */
    int32_t method;      /* enum dtx_method    */
    int32_t device;      /* HIP device ordinal */
    double dt0;
} dtx_config;

/* Per-step diagnostics of the functor (src/SteppingMethods.jl:48-94, 175-229). */
typedef struct {
    int32_t finite;        /* isfinite(sum(df))                     (ForwardEuler only; 1 else) */
    int32_t cr_evaluated;  /* Verbose in 1..3 and sum(f) != 0                                   */
    int32_t save;          /* the functor's `save` result                                       */
    int32_t step_index;    /* 0-based step counter since dtx_reset                              */
    double cr;             /* Cr                                                                */
    double cr_bin;         /* Cr_Bin   (Verbose == 3)                                           */
    double cr_emi;         /* Cr_Emi   (Verbose == 3)                                           */
    double cr_flux;        /* Cr_Flux (ForwardEuler) / Cr_PFlux (symplectic) (Verbose == 3)      */
    double cr_momentum;    /* Cr_Momentum (symplectic, Verbose == 3)                            */
    double cr_space;       /* Cr_Space    (symplectic, Verbose == 3)                            */
    double dt;             /* dt used (after the clip, src/SteppingMethods.jl:18-23)            */
    double t_start;
    double sum_f;          /* sum(f) before the update                                          */
} dtx_step_stats;

/* ---- lifetime -------------------------------------------------------------------------- */

/* Allocate a handle (replaces the device placement in ForwardEulerStruct /
 * ForwardSymplecticEulerStruct, src/SteppingStructs.jl:33-100, 142-227). */
int dtx_create(const dtx_config* cfg, dtx_handle** out);
/* Release all device memory of the handle. */
void dtx_destroy(dtx_handle* h);
/* Message of the last failed call on h; h == NULL gives the last failure of dtx_create. */
const char* dtx_last_error(const dtx_handle* h);

/* ---- operators (inputs of the struct constructors) ------------------------------------- */

/* A flux matrix as Julia's SparseMatrixCSC{Float64,Int64}: n x n, 1-based colptr (n+1) and
 * rowval (nnz), nzval (nnz).  Stored on the device as column-sorted ELL rows.
 * Replaces CuSparseMatrixCSC(FluxM.X_Flux + FluxM.P_Flux) (src/SteppingStructs.jl:68) and the
 * symplectic X_Flux / P_Flux uploads (:188-189).  A method that needs an absent matrix treats it
 * as zero. */
int dtx_set_flux_csc(dtx_handle* h, int32_t which, int64_t n, const int64_t* colptr,
                     const int64_t* rowval, const double* nzval);

/* The monolithic binary tensor BinM.M_Bin: (N^2 x N) column-major Float64 with
 * M_Bin[(i-1)N+k, j] = M[k,i,j] (src/Collisions/BuildBinaryMatrices.jl:54, 181-196, 261-276).
 * Replaces CuArray(BinM.M_Bin) (src/SteppingStructs.jl:66). */
int dtx_set_binary_dense(dtx_handle* h, int64_t N, const double* M_Bin);

/* The same tensor as Julia's SparseMatrixCSC{Float64,Int64} (Bin_sparse, N^2 x N, 1-based
 * colptr (N+1), rowval and nzval (nnz); src/Collisions/BuildBinaryMatrices.jl:44-52), without
 * densifying it.  Replaces CuSparseMatrixCSC(BinM.M_Bin) (src/SteppingStructs.jl:178-179).
 * The stored entries are folded into the same packed operator as dtx_set_binary_dense builds
 * from the dense array, bit for bit. */
int dtx_set_binary_csc(dtx_handle* h, int64_t N, const int64_t* colptr, const int64_t* rowval,
                       const double* nzval);

/* Channel form of the same tensor, for sizes where the monolithic array does not fit.
 * Gain: dB[k] += sum_{i in s1, j in s2} G[k + No*(i + N1*j)] f_i f_j for k in s_out, which is
 * what GainMatrix_to_M_Bin! contributes (G/2 at (k,i,j) and (k,j,i),
 * src/Collisions/BuildBinaryMatrices.jl:114-204).  Species are 0-based. */
int dtx_add_binary_gain(dtx_handle* h, int32_t s_out, int32_t s1, int32_t s2, const double* G);
/* Loss: dB[a] -= f_a * sum_{c in s2} L[a + N1*c] f_c for a in s1, which is what
 * LossMatrix_to_M_Bin! contributes (-L/2 at (a,a,c) and (a,c,a), :206-284). */
int dtx_add_binary_loss(dtx_handle* h, int32_t s1, int32_t s2, const double* L);

/* Bin_Mode: how the collision arrays are averaged over the angles (src/Collisions/
 * BuildBinaryMatrices.jl:7): ANI keeps them, AXI averages over phi (pz), ISO over u and phi. */
enum dtx_bin_mode { DTX_BIN_ANI = 0, DTX_BIN_AXI = 1, DTX_BIN_ISO = 2 };

/* One binary interaction name1 + name2 -> name3 + name4 as DiplodocusCollisions' arrays hold it
 * (the Output of BinaryFileLoad_Matrix, src/Collisions/LoadMatrices.jl:122-133).  Arrays are
 * Float64 in Julia's (column-major) memory order; px/py/pz counts are those of dtx_config. */
typedef struct {
    int32_t species[4];   /* name1_loc .. name4_loc, 0-based                                  */
    int32_t mode;         /* enum dtx_bin_mode                                                */
    int32_t sparse;       /* 1: Bin_sparse assembly (ANI gains below 1e-16 of their column's
                             maximum over px3 are dropped, BuildBinaryMatrices.jl:163-173)   */
    double gain_scale;    /* GainScale * Bin_Norm: every gain entry is multiplied by it first */
    double loss_scale;    /* LossScale * Bin_Norm                                             */
    const double* gain3;  /* GainMatrix3 (px3+2, py3, pz3, px1, py1, pz1, px2, py2, pz2): the
                             first and last px3 bins (under/overflow) are not used           */
    const double* gain4;  /* GainMatrix4 (px4+2, py4, pz4, px1, py1, pz1, px2, py2, pz2)      */
    const double* loss1;  /* LossMatrix1 (px1, py1, pz1, px2, py2, pz2)                       */
    const double* loss2;  /* LossMatrix2 (px2, py2, pz2, px1, py1, pz1)                       */
    const double* dpy;    /* u bin widths of every species in order, py_num[s] each
                             (Grids.dpy_list, deltaVector of the u bounds)                   */
    const double* dpz;    /* phi bin widths, pz_num[s] each (Grids.dpz_list)                  */
} dtx_binary_interaction;

/* Adds one interaction's gain and loss terms to the handle's operator, as Fill_M_Bin! adds them
 * to M_Bin (src/Collisions/BuildBinaryMatrices.jl:81-112, GainMatrix_to_M_Bin! :114-204,
 * LossMatrix_to_M_Bin! :206-284): val * w / 2 at M[a, b, c] and M[a, c, b] for every gain entry
 * and -val * w / 2 at M[a, a, c] and M[a, c, a] for every loss entry, with the ANI / AXI / ISO
 * averages and weights.  The terms go straight into the packed operator's FP64 channels and loss
 * matrix (no N^2 x N array); calls accumulate, in the order of Binary_list.  Replaces the
 * Fill_M_Bin! call of LoadMatrices_Binary (src/Collisions/LoadMatrices.jl:136) and the upload
 * of the M_Bin it fills. */
int dtx_add_binary_interaction(dtx_handle* h, const dtx_binary_interaction* x);
/* BinM.Domain: 0-based cell offsets that carry binary interactions; count < 0 means `nothing`
 * (all cells).  (src/Collisions/CollisionDomains.jl:6-38, src/SteppingMethods.jl:369) */
int dtx_set_binary_domain(dtx_handle* h, int64_t count, const int64_t* domain);

/* EmiM.M_Emi: dense n x n column-major Float64 (src/Collisions/BuildEmissionMatrices.jl:43).
 * It must be block diagonal per cell (the structure BuildEmissionMatrices produces); the
 * library keeps only the C diagonal N x N blocks and rejects off-block non-zeros.
 * Replaces CuArray(EmiM.M_Emi) (src/SteppingStructs.jl:67). */
int dtx_set_emission_dense(dtx_handle* h, int64_t n, const double* M_Emi);
/* EmiM.M_Emi as SparseMatrixCSC{Float64,Int64} (Emi_sparse, n x n, 1-based;
 * src/Collisions/BuildEmissionMatrices.jl:28-39), without densifying it.  Replaces
 * CuSparseMatrixCSC(EmiM.M_Emi) (src/SteppingStructs.jl:183-184).  Stored entries outside the
 * per-cell diagonal blocks must be zero. */
int dtx_set_emission_csc(dtx_handle* h, int64_t n, const int64_t* colptr, const int64_t* rowval,
                         const double* nzval);
/* The same operator given directly as C blocks of N x N column-major (cell-major order). */
int dtx_set_emission_blocks(dtx_handle* h, const double* blocks);

/* FluxM.Ap_Flux (n; the library forms 1 ./ Ap as src/SteppingStructs.jl:55 does), FluxM.Vol
 * (C), Initial (n, f_init) and Injection (n, df_Inj).  f is set to f_init.  Ap and Vol may both
 * be NULL after dtx_synth_flux, which has written them on the device. */
int dtx_set_vectors(dtx_handle* h, const double* Ap, const double* Vol, const double* f_init,
                    const double* df_inj);

/* Per-state T^00 weights E_bar(p) of one cell (N), i.e. Grids.dE_list expanded over (u, phi)
 * (src/StructsAndDictionaries.jl:259; src/DistributionMoments.jl:145,162).  Optional: without
 * them the energy moment is reported as 0. */
int dtx_set_moment_weights(dtx_handle* h, const double* energy_weight);

/* FP64 moment-rate correction of the FP32 / BF16x3 gain contraction (default: on).  Per gain
 * channel the library keeps W_N[i,j] = sum_k G[k,i,j] and W_E[i,j] = sum_k E_k G[k,i,j] in FP64
 * (E = the moment weights, when set before the first step) and evaluates the exact number and
 * energy rates f^T W f of the gain term every step; the reduced-precision gain sums of each
 * (cell, species) are rescaled per state by (1 + a + b E_k) so that their number and energy
 * sums equal those rates.  No counterpart in the reference, which has no reduced-precision
 * path; FP64 handles ignore it.  Must be called before the first step.
 * enable: 0 off; 1 on, Forward Euler steps summing the slots, solving the coefficients and
 * updating in one launch where the species layout allows; 2 on, always as three launches
 * (bitwise the same results; for checking). */
int dtx_set_moment_correction(dtx_handle* h, int32_t enable);

/* Synthetic workload support: generate lognormal(0,1)*scale gain channels for every species
 * pair (a <= b) -> (a, b) directly in device memory, with number-conserving losses
 * (sum_k G[k,i,j] = L[i,j]).  No counterpart in the reference, which reads DiplodocusCollisions
 * files; used for the benchmark workloads whose tensors are too large to build on the host. */
int dtx_synth_binary(dtx_handle* h, uint64_t seed, double scale);

/* Synthetic Cartesian flux operators built on the device, in place of the CSC that
 * BuildFluxMatrices assembles on the host (src/Fluxes/BuildFluxMatrices.jl:10-184) and
 * dtx_set_flux_csc uploads: upwind advection along x and / or y with the bin-averaged velocity
 * component (the B / C stencils, :708-974; Cartesian face fluxes,
 * src/Fluxes/FluxFunctionsCoordinate.jl:264-267) and synchrotron radiation reaction along p
 * and u (the I and J stencils, :275-557, with src/Fluxes/FluxFunctionsEmissionForces.jl:10-150),
 * closed p and u boundaries.  Builds the matrices the method reads (ForwardEuler: F = X + P; symplectic: X and
 * P) with the entries, explicit zeros, duplicate sums and column order that dtx_set_flux_csc
 * gives for the host-assembled CSC, and writes 1 ./ Ap (Ap = dx dy dz per cell, :145-154) and
 * Vol = dt0 dx dy dz (src/Fluxes/FluxFunctionsCoordinate.jl:313-315), so dtx_set_vectors may
 * then take NULL Ap and Vol.  The bounds arrays are read during the call only. */
typedef struct {
    const double* xr;  /* x_num + 1 cell bounds; likewise yr (y_num + 1) and zr (z_num + 1)  */
    const double* yr;
    const double* zr;
    const double* pr;  /* momentum bounds of every species in order: px_num[s] + 1 each      */
    const double* ur;  /* u bounds: py_num[s] + 1 each                                       */
    const double* hr;  /* phi bounds: pz_num[s] + 1 each                                     */
    double mass[DTX_MAX_SPECIES];
    double charge[DTX_MAX_SPECIES];
    double sync_K;     /* synchrotron coefficient; 0: no momentum fluxes (P empty)           */
    int32_t axes;      /* bit 0: advection along x, bit 1: along y (axes of one cell skipped) */
    int32_t periodic;  /* 1: periodic space boundaries, 0: closed                            */
} dtx_flux_synth;
int dtx_synth_flux(dtx_handle* h, const dtx_flux_synth* spec);

/* ---- stepping ---------------------------------------------------------------------------- */

/* One functor call `(method)(t_start, t_stop, dt, Verbose) -> (dt, save)`
 * (src/SteppingMethods.jl:12-107 / 120-243).  Returns the clipped dt and save.  With
 * stats != NULL the call waits for the step and fills *stats; otherwise it only enqueues the
 * step (its diagnostics are printed when the handle next synchronises). */
int dtx_step(dtx_handle* h, double t_start, double t_stop, double dt, int32_t verbose,
             double* dt_out, int32_t* save_out, dtx_step_stats* stats);

/* dtx_step without allocation.  No counterpart in the reference: for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same arguments, semantics and results as dtx_step, and it calls no
 * hipMalloc, hipFree or other allocating routine: the contraction's partial sums live in a
 * 128 MiB array that the library holds on each device it runs on until the process ends.
 * Limits: the handle's operators must already be packed (any dtx_step, dtx_solve or dtx_work
 * call packs them; DTX_ERR_STATE otherwise); a step whose partial sums exceed 128 MiB is
 * refused with DTX_ERR_UNSUPPORTED (C x N x 8 bytes per partial slot).  Every call on a device
 * shares that device's array: no-alloc steps of different handles that overlap in time race,
 * so order them on one stream or synchronise between them (steps of one handle are ordered on
 * its stream).  Kernel timing (dtx_kernel_timing) does not cover these steps. */
int dtx_step_noalloc(dtx_handle* h, double t_start, double t_stop, double dt, int32_t verbose,
                     double* dt_out, int32_t* save_out, dtx_step_stats* stats);

/* Solve(method, dt_initial, t_save; Verbose) (src/Solver.jl:1-64) without the JLD2 write:
 * resets f to f_init, steps through every save interval without host synchronisation and writes
 * the n_save states to f_out (n_save x n, row i = output.f[i+1]) and t_out (n_save).  The
 * clipped dt carries over between intervals as in the reference (:35).  *n_steps receives the
 * number of steps taken (may be NULL). */
int dtx_solve(dtx_handle* h, double dt_initial, int64_t n_save, const double* t_save,
              int32_t verbose, double* f_out, double* t_out, int64_t* n_steps);

/* `@. method.f = method.f_init` (src/Solver.jl:22) and clears the step counter. */
int dtx_reset(dtx_handle* h);
/* copyto!(f, method.f) (src/Solver.jl:40) and the reverse, for resume. */
int dtx_get_f(dtx_handle* h, double* f);
int dtx_set_f(dtx_handle* h, const double* f);
/* Number N^0 and energy T^00 of the current f per (cell, species): arrays of C*num_species,
 * cell-major (src/DistributionMoments.jl:33, 145/162), computed in FP64 by the step kernel. */
int dtx_moments(dtx_handle* h, double* number, double* energy);
/* Wait for all enqueued work of h and print pending step diagnostics. */
int dtx_synchronize(dtx_handle* h);
/* Statistics of the most recent max_steps steps (oldest first) as kept on the device. */
int dtx_recent_stats(dtx_handle* h, dtx_step_stats* out, int32_t max_steps, int32_t* count);

/* ---- introspection ----------------------------------------------------------------------- */

/* Bytes of device memory the handle holds (operators, state and scratch). */
int64_t dtx_device_bytes(const dtx_handle* h);
/* Algorithmic work of one step: binary FLOPs (math, counted once) and HBM bytes per part. */
typedef struct {
    double bin_flops;      /* contraction kernels: 2 * No * pairs * cells + FP64 loss GEMM */
    double bin_bytes;      /* packed gain bytes + loss matrix + f reads + partial traffic */
    double update_bytes;   /* fused update kernel: f, partials, invAp, df_inj, ELL, out   */
    double gain_bytes;     /* packed gain operator bytes (streamed once per step)         */
    int64_t pairs;         /* total packed pairs over channels                            */
    int32_t bin_launch_items;
    int32_t fused_correction;  /* 1: corrected Forward Euler steps sum the slots, solve the
                                  coefficients and update in one launch                    */
    double emission_bytes; /* emission GEMV: the per-cell blocks + f + df_Emi (0: none)    */
} dtx_work_info;
int dtx_work(const dtx_handle* h, dtx_work_info* out);
/* Device time of the last synchronised window, per kernel class, in milliseconds (HIP events
 * on the handle's stream); reset by passing reset != 0.  enable = k > 0 times the kernels of
 * every k-th step from then on (1: every step; each timed step costs the stream a few
 * microseconds of event records), 0 stops timing. */
typedef struct {
    double bin_ms, update_ms;
    int64_t bin_launches, update_launches;
    double emission_ms;          /* the emission GEMV (handles with emission blocks) */
    int64_t emission_launches;
} dtx_kernel_times;
int dtx_kernel_timing(dtx_handle* h, int32_t enable, int32_t reset, dtx_kernel_times* out);

/* ---- serving ----------------------------------------------------------------------------- */

typedef struct dtx_server dtx_server;

typedef struct {
    int32_t device;
    int32_t precision;       /* enum dtx_precision used for every request            */
    int64_t hbm_budget;      /* bytes of device memory the admitted requests may hold  */
    int32_t max_active;      /* requests admitted at once (building, stepping, retiring) */
    int32_t no_moment_correction;  /* 1: FP32 / BF16x3 requests without the FP64 moment-rate
                                      correction (dtx_set_moment_correction); 0: with it    */
} dtx_server_config;

typedef struct {
    dtx_config cfg;          /* shape of the simulation                               */
    uint64_t seed;           /* synthetic workload seed (operators, f0)               */
    int64_t steps;           /* Euler steps to run                                    */
    double theta;            /* Maxwell-Juttner temperature of f0                     */
    int32_t with_flux;       /* x advection + synchrotron cooling when x_num > 1      */
} dtx_request;

typedef struct {
    int64_t id;
    int32_t status;          /* 0 queued, 1 running, 2 done, negative = failed        */
    double submit_ms, start_ms, finish_ms;  /* server clock                           */
    double number_total, energy_total;       /* moments of the final f               */
    int64_t steps_done;
} dtx_request_result;

int dtx_server_create(const dtx_server_config* cfg, dtx_server** out);
/* Enqueue a request; returns its id through *id.  Thread-safe. */
int dtx_server_submit(dtx_server* s, const dtx_request* req, int64_t* id);
/* Enqueue a simulation the caller built (dtx_create, operators and dtx_set_vectors, in the
 * server's precision, ForwardEuler) to be stepped `steps` times with step dt (no clipping) in
 * the grouped launches; the server takes ownership of the handle and destroys it when the
 * request finishes (its final f and moments stay available through poll / fetch).  The caller
 * must not use the handle after this call.  Thread-safe.  No counterpart in the reference:
 * this is how arriving user simulations, with their own operators and f0, join the loop. */
int dtx_server_submit_handle(dtx_server* s, dtx_handle* h, int64_t steps, double dt,
                             int64_t* id);
/* Result of a request (status 2 when done).  Thread-safe. */
int dtx_server_poll(dtx_server* s, int64_t id, dtx_request_result* out);
/* Final f of a finished request (n doubles). */
int dtx_server_fetch(dtx_server* s, int64_t id, double* f, int64_t n);
/* Block until every submitted request has finished. */
int dtx_server_drain(dtx_server* s);
void dtx_server_destroy(dtx_server* s);
const char* dtx_server_last_error(const dtx_server* s);

/* The simulation a server builds for a synthetic request, as a stand-alone handle with the
 * server config's device, precision and moment-rate correction, and its step *dt: stepped by
 * dtx_step with that dt (and a t_stop it never reaches) it gives the served result bit for
 * bit.  No counterpart in the reference. */
int dtx_request_create(const dtx_request* req, const dtx_server_config* cfg, dtx_handle** out,
                       double* dt);

#ifdef __cplusplus
}
#endif
#endif /* DTX_H */
