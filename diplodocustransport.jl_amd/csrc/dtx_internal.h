// dtx_internal.h -- device data structures shared by the runtime and the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/dtx.h"

namespace dtx {

// ---- binary contraction tiling --------------------------------------------------------------
// One work item multiplies a TK-row tile of a packed gain channel with the pair products of a
// TC-cell tile over a range of pairs, and writes a partial sum for its (slot, cell, row).
constexpr int TK = 64;          // output rows (k) per work item
constexpr int KC64 = 16;        // pairs per LDS stage, FP64 and FP32 kernels
constexpr int KC16 = 32;        // pairs per LDS stage, split-BF16 kernel
constexpr int PAIR_ALIGN = 32;  // channel pair lists are padded to this multiple
constexpr int FLUSH_PAIRS = 512;  // FP32 accumulators are folded into FP64 every this many pairs
constexpr uint32_t PAIR_PAD = 0xFFFFFFFFu;  // padded pair: product forced to zero

// Channel kinds: TRI = one species with itself, pairs i <= j; RECT = s1 < s2, all (i, j).
enum ChannelKind : int32_t { CH_TRI = 0, CH_RECT = 1 };

// A work item of the grouped contraction kernel (one workgroup).  Items of several problems
// (serving) can be mixed in one launch; everything an item needs is reachable from it.
struct BinWork {
    const void* A;          // packed gain of the channel: [ktile][...] (see pack layout)
    const uint32_t* pairs;  // pair table of the channel: (i | j << 16), cell-local state index
    const double* fT;       // state-major f of the problem: fT[state * ldc + cell]
    double* part;           // partial sums of this slot: part[cell * N + row]
    int64_t a_ktile_stride; // elements between consecutive k-tiles of A
    int32_t ktile;          // k-tile index inside the channel
    int32_t ctile;          // first cell of the tile = ctile * TC
    int32_t p_begin, p_end; // pair range, multiples of the stage size
    int32_t row_base;       // state index of the channel's first output row inside a cell
    int32_t rows;           // valid output rows of the channel (No)
    int32_t cells;          // valid cells of the problem (C)
    int32_t ldc;            // leading dimension of fT
    int32_t N;              // states per cell (partial buffer row pitch)
    int32_t pad_;
};

// Packed operator of one channel on the device.
struct PackedChannel {
    int32_t s_out, s1, s2, kind;
    int32_t No, N1, N2;
    int64_t npairs;         // real pairs
    int64_t npairs_pad;     // padded to PAIR_ALIGN
    int32_t nktiles;        // ceil(No / TK)
    void* A = nullptr;      // device: packed gain (precision dependent)
    uint32_t* pairs = nullptr;  // device pair table
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
    double* fT;             // state-major copy of f_out for the next contraction
    const double* inv_ap;
    const double* df_inj;
    const double* vol;
    const double* part;     // binary partial sums [slot][cell][N]
    const int32_t* nslot;   // slots per state (N)
    int64_t part_stride;    // doubles between slots
    const double* loss;     // FP64 N x N diagonal/loss matrix (column-major) or null
    const double* emi;      // emission blocks [cell][N x N] or null
    const uint8_t* active;  // binary domain per cell (1 = evaluated) or null = all
    const int32_t* ell_col; const double* ell_val; int32_t ell_width;
    const double* eweight;  // per-state energy weight (N) or null
    const int32_t* state_species;  // species of each state (N)
    const int32_t* species_off;    // (S+1)
    double* moments;        // [cell][S][2] (number, energy)
    double* block_part;     // per-block partials [nblocks][8]
    double* status;         // status ring [ring][16]
    uint32_t* counter;      // last-block counter
    // symplectic intermediates
    double* f_mom;          // f after the momentum update
    double* df_mom;         // df_Momentum (scaled)
    double* dpflux;         // df_PFlux (unscaled) kept for Cr_PFlux
    double* dbin;           // df_Bin (unscaled) kept for Cr_Bin
    double* demi;           // df_Emi (unscaled) kept for Cr_Emi
    int64_t n;
    int32_t N, C, S, ldc;
    int32_t binary, emission, verbose;
    int32_t step_index, ring;
    double s;               // dt_scale
    double t_start, dt;
    int32_t save;
};

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

// ---- kernel launchers (kernels_*.hip) ----------------------------------------------------
hipError_t launch_binary(int precision, int tc, const BinWork* items, int nitems,
                         hipStream_t stream);
hipError_t launch_euler_update(const UpdateArgs& a, hipStream_t stream);
hipError_t launch_symplectic_momentum(const UpdateArgs& a, hipStream_t stream);
hipError_t launch_symplectic_space(const UpdateArgs& a, hipStream_t stream);
hipError_t launch_fill_zero(double* p, int64_t count, hipStream_t stream);

// synthetic generator (synth.hip)
hipError_t launch_synth_gain_packed(int precision, void* A, const uint32_t* pairs,
                                    int64_t npairs_pad, int No, int nktiles, int kind,
                                    int N1, int N2, uint64_t key, double scale,
                                    hipStream_t stream);
hipError_t launch_synth_loss(double* loss, int N, int off1, int N1, int off2, int N2,
                             int No, uint64_t key, double scale, int kind_same,
                             hipStream_t stream);

uint64_t splitmix_key(uint64_t seed, uint64_t stream);

}  // namespace dtx
