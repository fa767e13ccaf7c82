// dtx_internal.h -- device data structures shared by the runtime and the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/dtx.h"

namespace dtx {

// ---- binary contraction tiling --------------------------------------------------------------
// The contraction follows the two-GEMV association of update_Big_Bin!: for every pair row i,
// T_i[k, c] = sum_j S[k, (i, j)] f_j^c is a GEMM whose B operand is the state itself, and
// dB[k, c] += f_i^c * T_i[k, c] is folded into an FP64 accumulator when the row ends.  Rows are
// cut into chunks of KC consecutive j; a work item multiplies a TK-row tile of one channel with
// a TC-cell tile over a range of chunks and writes its partial sum to its own slot.
constexpr int TK = 64;          // output rows (k) per work item
constexpr int KC = 32;          // j per chunk (one LDS stage)
constexpr int SPECIES_ALIGN = 32;  // species blocks of the B operand start on this multiple
constexpr int MAX_RANGE = 1024;    // most chunks one work item covers (its table slice lives in LDS)

// Channel kinds: TRI = one species with itself, pairs i <= j; RECT = s1 < s2, all (i, j);
// LOSS = rows of one species against all states, multiplied by the row's own f at the end.
// CH_MOMENT items are LOSS-layout FP64 items of a moment-rate matrix W: they write, per cell,
// the sum over their rows of f_row (W f)_row instead of the rows themselves.
enum ChannelKind : int32_t { CH_TRI = 0, CH_RECT = 1, CH_LOSS = 2, CH_MOMENT = 3 };

// A work item of the grouped contraction kernel (one workgroup).  Items of several problems
// (serving) can be mixed in one launch; everything an item needs is reachable from it.
struct BinWork {
    const void* A;            // packed A of the channel, k-tile 0: [ktile][chunk][...]
    const uint32_t* chunks;   // chunk table: i_pad | (j0_pad << 16)  (padded state indices)
    const double* fT;         // FP64 f, state-major over padded states: fT[kp * ldc + cell]
    const void* B;            // B operand: FP64 -> fT; FP32 -> fT32 (FP32 f, state-major),
                              // BF16 -> hi bits in fragment order: B[fb_index(cell, kp, ldb)]
    int64_t b_lo;             // BF16: elements from the hi to the lo array
    double* part;             // partial sums of this slot: part[cell * N + row]
                              // (CH_MOMENT: the item's per-cell sums, part[cell]; wide
                              // split-BF16 gain items: FP32 sums, ((float*)part)[cell * N + row])
    int64_t a_ktile_stride;   // elements between consecutive k-tiles of A
    int32_t ktile;            // k-tile inside the channel
    int32_t ctile;            // first cell of the tile = ctile * TC
    int32_t c_begin, c_end;   // chunk range
    int32_t row_base;         // state index (unpadded) of the channel's first output row
    int32_t row_pad_base;     // padded state index of the first output row (LOSS multiplier)
    int32_t rows;             // valid output rows of the channel (No)
    int32_t cells;            // valid cells of the problem (C)
    int32_t ldc;              // leading dimension of fT (cells, padded)
    int32_t ldb;              // 16-cell blocks of the BF16 B arrays (ldc / 16)
    int32_t N;                // states per cell (partial buffer row pitch)
    int32_t kind;             // ChannelKind
    int32_t tc;               // cells per tile of the item (1: GEMV, 16, 32, 64, or WIDE)
};

// Packed operator of one channel on the device.
struct PackedChannel {
    int32_t s_out, s1, s2, kind, precision;
    int32_t No, N1, N2;
    int64_t npairs;         // real (i, j) pairs
    int64_t nchunks;        // chunks of KC
    int32_t nktiles;        // ceil(No / TK)
    int32_t moment = -1;    // moment-rate channel: 0 number, 1 energy (-1: operator channel)
    bool synth = false;     // generated on the device from (synth_key, synth_scale)
    uint64_t synth_key = 0;
    double synth_scale = 0.0;
    void* A = nullptr;      // device: packed A
    uint32_t* chunks = nullptr;  // device chunk table
    int64_t a_bytes = 0;
};

// ELL form of a CSC flux matrix: slot-major, column-sorted rows, col = -1 pads.
struct EllMatrix {
    int64_t n = 0;
    int32_t width = 0;      // slots per row
    int32_t* col = nullptr; // device [width][n]
    double* val = nullptr;  // device [width][n]
    int64_t nnz = 0;
    bool present() const { return width > 0; }
};

// Arguments of the fused update kernels.
struct UpdateArgs {
    const double* f;        // current state (cell-major)
    double* f_out;          // next state
    double* fT;             // FP64 state-major copy of f_out (padded states) for the contraction
    float* fT32;            // FP32 copy in the same layout (FP32 contraction), or null
    uint16_t* fhi;          // BF16 hi/lo split in fragment order (fb_index), or null
    uint16_t* flo;
    const int32_t* pad_index;  // padded index of every state (N)
    int32_t ldb;            // 16-cell blocks of fhi / flo (ldc / 16; fb_index)
    const double* inv_ap;
    const double* df_inj;
    const double* vol;
    const double* part;     // binary partial sums [slot][cell][N] (moment-rate corrected:
                            // the loss sums [cell][N])
    // moment-rate correction (null gsum: off): gain sums [cell][N], the moment partials of the
    // gain sums [cell][S][zb][3], the moment-rate item sums [index][ldc] and their lists
    const double* gsum;
    const double* cpart;
    const double* mpart;
    const int32_t* mom_off;
    const int32_t* mom_list;
    int32_t zb, zspan;      // sums workgroups per species and the states each covers
    double* coef;           // (alpha, beta) [cell][S][2] from correct_coef_kernel, or null: each
                            // update workgroup derives its own (grids of at most one round)
    const int32_t* nslot;   // slots per state (N)
    const int32_t* nslot32; // leading slots of each state holding FP32 sums (N), or null
    int64_t part_stride;    // doubles between slots
    const double* emi;      // emission blocks [cell][N x N] or null
    double* emi_part;       // emission GEMV part sums [emi_q][cell][N] when emi_q > 1
    int32_t emi_q;          // workgroups per (cell, row tile) of the emission GEMV
    const uint8_t* active;  // binary domain per cell (1 = evaluated) or null = all
    const int32_t* ell_col; const double* ell_val; int32_t ell_width;
    const double* eweight;  // per-state energy weight (N) or null
    const int32_t* state_species;  // species of each state (N)
    const int32_t* species_off;    // (S+1)
    double* moments;        // [cell][S][2] (number, energy)
    double* mom_part;       // per-block moment partials [cell][segment][S][2]
    double* block_part;     // per-block partials [nblocks][BP_WORDS], then the finalize
                            // fold partials [ceil(max(nblocks, C S) / 256)][BP_WORDS]
    double* status;         // status ring [ring][16]
    // symplectic intermediates
    double* f_mom;          // f after the momentum update
    double* df_mom;         // df_Momentum (scaled)
    double* dpflux;         // df_PFlux (unscaled) kept for Cr_PFlux
    double* dbin;           // df_Bin (unscaled) kept for Cr_Bin
    double* demi;           // df_Emi (unscaled) of the step's emission GEMV (n), read by the
                            // update kernels and by the space half's Cr_Emi
    int64_t n;
    int32_t N, C, S, ldc;
    int32_t binary, emission, verbose;
    int32_t step_index, ring;
    double s;               // dt_scale
    double t_start, dt;
    int32_t save;
    // batched finalisation (Forward Euler, small grids): the update writes its step's scalars
    // to fin_rec[fin_slot] (FIN_WORDS doubles) for finalize_batch; null: not batched
    double* fin_rec;
    int32_t fin_slot;
};

// Steps whose finalisation is batched into one launch (their block partials and scalars are
// kept in rings of this many slots), and the doubles of one step's record.
constexpr int FIN_BATCH = 64;
constexpr int FIN_WORDS = 8;

constexpr int EMI_MAX_N = 16384;   // states per cell with emission blocks (f of a cell in LDS)
// Workgroups of the emission GEMV per (cell, row tile): enough for two per CU over the whole
// grid, each keeping at least 64 columns
int emission_parts(int N, int C, int num_cu);
constexpr int UPDATE_STATES = 64;  // states per update workgroup (four threads each)
constexpr int UPDATE_THREADS = 256;  // threads per update and finalize workgroup
constexpr int STATUS_RING = 4096;
constexpr int STATUS_WORDS = 16;  // doubles per status record
// status record layout
enum StatusSlot {
    ST_STEP = 0, ST_FINITE, ST_CR_EVAL, ST_CR, ST_CR_BIN, ST_CR_EMI, ST_CR_FLUX, ST_CR_MOM,
    ST_CR_SPACE, ST_SUMF, ST_DT, ST_TSTART, ST_SAVE, ST_SUMDF
};
// per-block partial layout
enum BlockSlot { BP_SUMDF = 0, BP_SUMF, BP_CR, BP_CR_BIN, BP_CR_EMI, BP_CR_FLUX, BP_CR_MOM,
                 BP_CR_SPACE, BP_WORDS };

// FP64 moment-rate correction of a reduced-precision gain term, per (cell, output species s).
// dBg = sum of a state's gain slots (gsum), dBl = sum of its loss slots (lsum); R_N and R_E are
// the exact FP64 number and energy rates of the gain term (sums of the moment-rate item sums
// listed for (s, moment) in mom_off / mom_list).  The update kernels take the corrected binary
// sum dBg (1 + a + b E) + dBl, with (a, b) chosen so that sum dBg' = R_N and sum E dBg' = R_E
// over the species (correction_coef).  CorrectArgs drives the launch that sums the slots.
struct CorrectArgs {
    const double* part;        // partial slots [slot][cell][N]
    int64_t part_stride;
    double* gsum;              // gain sums [cell][N]
    double* lsum;              // loss sums [cell][N]
    double* cpart;             // moment partials of the gain sums [cell][S][zb][3]
    int32_t zb;                // workgroups per species in the sums launch: ceil(max ns / zspan)
    int32_t tps;               // threads per state in the sums launch (4 or 16); zspan = 256 / tps
    const int32_t* species_off;  // (S+1)
    const int32_t* gslots;     // gain slots per species (S)
    const int32_t* nslots;     // gain + loss slots per species (S)
    int32_t gain_f32;          // the gain slots hold FP32 sums (wide split-BF16 items)
    const double* eweight;     // per-state energy weight (N) or null (number only)
    int32_t N, C, S;
};

// ---- kernel launchers (kernels_*.hip) ----------------------------------------------------
hipError_t launch_gain_correct(const CorrectArgs& a, hipStream_t stream);
// Corrected Forward Euler in one launch (slot sums, coefficients, update; bitwise the same as
// launch_gain_correct + launch_euler_update), one workgroup per (cell, species), for layouts
// fused_update_fits accepts: four threads per state in the sums, every species starting on an
// update segment and at most 16 segments long.  finalize: also launch the step's finalisation.
bool fused_update_fits(int N, int S, const int32_t* species_off, int tps);
hipError_t launch_euler_update_fused(const UpdateArgs& a, const CorrectArgs& ca, bool finalize,
                                     hipStream_t stream);
// Grouped forms (serving): member g of the table owns blocks [start[g], start[g+1]) of the
// launch, which runs the member's own grid with its own arithmetic.  The coefficient launch
// covers the members whose update grid is too large to derive them per workgroup (coef set).
hipError_t launch_gain_correct_group(const CorrectArgs* table, const int32_t* start, int count,
                                     int total_blocks, hipStream_t stream);
hipError_t launch_coef_group(const UpdateArgs* table, const int32_t* start, int count,
                             int total_blocks, hipStream_t stream);
hipError_t launch_binary(int precision, int tc, const BinWork* items, int nitems,
                         hipStream_t stream);
// Items of the narrow 64-row shapes (tc 1, 16 and 64, per item) in one launch: grouped steps
// mix single-cell requests (operator streaming, HBM-bound) with many-cell ones (MFMA-bound).
hipError_t launch_binary_mixed(int precision, const BinWork* items, int nitems,
                               hipStream_t stream);
// Contraction with (64 ns) x WIDE (rows x cells) gain items for C > 64: split-BF16 with ns = 2
// or 4 k tiles per item, FP64 with ns = 2.  LOSS-layout items keep the 64 x 64 FP64 body.
constexpr int WIDE = 128;   // cells per wide item
// Groups of four waves in a wide FP64 workgroup: each group takes its own share of the item's
// chunk range, and the groups' sums are added in LDS before the single partial is written, so
// an item needs one slot per G chunk shares.  With G = 2 LOSS items take the wide shape too.
#ifndef DTX_F64_WIDE_GROUPS
#define DTX_F64_WIDE_GROUPS 2
#endif
constexpr int F64_WIDE_GROUPS = DTX_F64_WIDE_GROUPS;
hipError_t launch_binary_wide(int precision, int ns, const BinWork* items, int nitems,
                              hipStream_t stream);
hipError_t launch_pack_loss(const double* D, int N, const int32_t* pad_index, int Np, int row0,
                            int rows, int nktiles, double* A, hipStream_t stream);
hipError_t launch_euler_update(const UpdateArgs& a, hipStream_t stream);  // update + finalisation
// Forward Euler update without its finalisation (the step's record goes to a.fin_rec); only for
// grids where finalize_is_single holds.  finalize_batch later finalises `count` such steps
// whose records and block partials sit in ring slots first, first + 1, ... (mod FIN_BATCH;
// block partials of slot q at base.block_part + q * bp_words), one workgroup per step, and
// writes the moments of the last of them (base.mom_part holds that step's partials).
hipError_t launch_euler_update_only(const UpdateArgs& a, hipStream_t stream);
hipError_t launch_finalize_batch(const UpdateArgs& base, int64_t bp_words, int first, int count,
                                 hipStream_t stream);
bool finalize_is_single(const UpdateArgs& a);
hipError_t launch_finalize(const UpdateArgs& a, hipStream_t stream);
hipError_t launch_symplectic_momentum(const UpdateArgs& a, hipStream_t stream);
// df_Emi = M_Emi f over the per-cell blocks into a.demi (nothing without emission)
hipError_t launch_emission(const UpdateArgs& a, hipStream_t stream);
hipError_t launch_symplectic_space(const UpdateArgs& a, hipStream_t stream);
hipError_t launch_euler_update_group(const UpdateArgs* table, const int32_t* block_start,
                                    int count, int total_blocks, hipStream_t stream);
hipError_t launch_fill_zero(double* p, int64_t count, hipStream_t stream);
hipError_t launch_publish_state(const double* f, int N, int C, int ldc, const int32_t* pad_index,
                                double* fT, float* fT32, uint16_t* fhi, uint16_t* flo,
                                hipStream_t stream);

// synthetic generator (synth.hip)
hipError_t launch_synth_gain_packed(int precision, void* A, const uint32_t* chunks,
                                    int64_t nchunks, int No, int nktiles, int kind, int N1,
                                    int N2, int pad1, int pad2, uint64_t key, double scale,
                                    hipStream_t stream);
hipError_t launch_synth_loss(double* loss, int N, int off1, int N1, int off2, int N2,
                             int No, uint64_t key, double scale, int kind_same,
                             hipStream_t stream);
// Moment-rate matrices W_N[i,j] = sum_k G'[k,i,j], W_E[i,j] = sum_k E_k G'[k,i,j] of a
// synthetic gain channel (G' its symmetric pair form), packed as LOSS rows i over the j chunks
// of species 2.  ew (No weights of the output species) or AE may be null.
hipError_t launch_synth_moment_packed(double* AN, double* AE, int64_t nchunks, int nktiles,
                                      int kind, int No, int N1, int N2, const double* ew,
                                      uint64_t key, double scale, hipStream_t stream);

// Synthetic Cartesian flux stencils built on the device (synth_flux_kernel): upwind advection
// along x and / or y, synchrotron radiation reaction along p.
constexpr int FLUX_MAX_ROW = 10;  // entries per row: diagonal, 2 per space axis, p +- 1, u +- 1
struct FluxSynthArgs {
    const double *xr, *yr, *zr;        // cell bounds (x_num + 1, ...)
    const double *pr, *ur, *hr;        // momentum bounds of every species, concatenated
    int32_t pr_off[DTX_MAX_SPECIES], ur_off[DTX_MAX_SPECIES], hr_off[DTX_MAX_SPECIES];
    int32_t px[DTX_MAX_SPECIES], py[DTX_MAX_SPECIES];
    int32_t sp_off[DTX_MAX_SPECIES + 1];
    double mass[DTX_MAX_SPECIES], charge[DTX_MAX_SPECIES];
    double dt0, sync_K;
    int64_t n;
    int32_t N, S, x_num, y_num, z_num, axes, periodic;
    int32_t which;                     // DTX_FLUX_F (X + P), DTX_FLUX_X or DTX_FLUX_P
    int32_t width;                     // 0: count pass, else the ELL width of the fill pass
    int32_t* col;                      // fill pass: [width][n]
    double* val;
    unsigned long long* stats;         // count pass: longest row, entries
    double* inv_ap;                    // count pass, when set: 1/Ap [n] and Vol [C]
    double* vol;
};
hipError_t launch_synth_flux(const FluxSynthArgs& a, hipStream_t stream);

uint64_t splitmix_key(uint64_t seed, uint64_t stream);

// grouped stepping of several handles (runtime.hip)
struct Group;
Group* group_create(int device);
void group_destroy(Group* g);
const char* group_error(const Group* g);
hipStream_t group_stream(const Group* g);
// Packs a handle's operators and builds its work tables (from any thread; the handle must not be
// in use elsewhere); required before a handle joins a group.
int group_prepare(dtx_handle* h);
// Members must be prepared; the tables are enqueued on the group stream (no device wait).
int group_set(Group* g, dtx_handle* const* hs, int count);
int group_step(Group* g, const double* dt, int verbose);
int group_sync(Group* g);
// read-only views of a handle for the server
const dtx_config* handle_config(const dtx_handle* h);
bool handle_has_vectors(const dtx_handle* h);
// records msg as the handle's last error and returns code
int handle_fail(dtx_handle* h, int code, const char* msg);
// first state of every species in a cell (S + 1 entries)
const int32_t* handle_species_offsets(const dtx_handle* h);

// Host-side assembly of a handle's FP64 binary operator (runtime.hip).  A channel (s_out; s1 <=
// s2) holds S[pair * No + k], pair = (a, b) of states a in s1, b in s2 (a <= b when s1 == s2:
// upper-triangle row-major, else a * N2 + b); the loss matrix is N x N, [k + N * other].
struct BinaryHostView {
    double* S;
    int32_t No, N1, N2;
    bool tri;
};
int binary_host_begin(dtx_handle* h, const char* who);  // loss matrix allocated when absent
BinaryHostView binary_host_channel(dtx_handle* h, int s_out, int s1, int s2);
double* binary_host_loss(dtx_handle* h);
void binary_host_end(dtx_handle* h, bool loss_nonzero);  // marks the operator for repacking

}  // namespace dtx
