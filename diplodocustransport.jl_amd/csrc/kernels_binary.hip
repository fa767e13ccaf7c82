// kernels_binary.hip -- grouped binary-collision contraction for gfx950.
//
// Replaces the per-cell pair of GEMVs of update_Big_Bin! (src/SteppingMethods.jl:360-386),
//     T[k,i] = sum_j M[k,i,j] f_j ;  dB[k] = sum_i T[k,i] f_i ,
// for all cells at once.  The tensor is packed per channel in symmetric pair rows: row i holds
// S[k,(i,j)] for its j range, cut into chunks of KC = 32 consecutive j.  For each chunk the
// kernel runs an MFMA GEMM  acc[k,c] += sum_j S[k,(i,j)] f_j^c  whose B operand is the state
// itself, scaled by f_i^c: FP64 and FP32 fold f_i into the staged B tile, split-BF16 keeps B
// as the bf16 split of f and folds tot[k,c] += f_i^c * acc[k,c] in FP64 when the row ends.  LOSS
// channels (the diagonal part, FP64 in every mode) accumulate sum_j D[k,j] f_j^c over all chunks
// and are multiplied by the row's own f_k^c.
//
// Every workgroup takes one work item (channel, row tile, cell tile, chunk range) from a work
// table and writes its partial sum to its own slot (FP64; FP32 for wide split-BF16 gain items,
// whose sums are FP32), so the reduction order is fixed by the item list alone: a simulation
// gives bit-identical results alone or grouped with others.  LOSS-layout items share the launch
// with the gain items of any precision (they take the FP64 body).
//
// Item shapes: 64 rows x 16, 32 or 64 cells (four waves = WK (rows) x WC (cells), each owning
// MR x MC 16x16 MFMA tiles), and for more than 64 cells the wide shapes, 128 or 256 rows x 128
// cells with 64 x 64 per wave (wide FP64 and split-BF16 sections below).
// MFMA operand layout (16x16xK): lane l supplies A[l%16][kgroup(l/16)], B[kgroup(l/16)][l%16].
// The 16x16 result tile is held as D[4*(l/16)+e][l%16] (e = 0..3) by the FP32 and BF16
// instructions, and as D[(l/16)+4*e][l%16] by v_mfma_f64_16x16x4_f64 (measured on gfx950).
#include "bf16_util.h"
#include "dtx_internal.h"
#include "wave_ops.h"

namespace dtx {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef double d2 __attribute__((ext_vector_type(2)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));  // one 16-byte load / LDS store

template <int TC>
struct Tiling {
    static constexpr int WC = TC >= 32 ? 2 : 1;   // waves along cells
    static constexpr int WK = 4 / WC;             // waves along rows
    static constexpr int WROWS = TK / WK;
    static constexpr int WCOLS = TC / WC;
    static constexpr int MR = WROWS / 16;
    static constexpr int MC = WCOLS / 16;
};

// Row pitch of the epilogue's cell-major transpose tiles in LDS: 64 rows + 1 against bank
// conflicts of the result-layout stores.
constexpr int EPI_PITCH = 65;

// The work table's pointers are generic to the compiler, which would emit flat memory
// instructions for them.  Flat loads count against the LDS counter as well, so every wait for an
// LDS read would also wait for the operator loads prefetched for the next stage, and their
// latency would never overlap the MFMAs.  All global accesses of the contraction kernels
// therefore go through global-address-space pointers.
#define GLB __attribute__((address_space(1)))

struct ItemView {
    const GLB char* A;
    const GLB uint32_t* chunks;
    const GLB double* fT;
    const GLB void* B;
    int64_t b_lo;
    GLB double* part;
    int c_begin, c_end, krow0, row_base, row_pad_base, rows, cells, ldc, ldb, N, c0, kind;
};

__device__ inline ItemView load_item(const BinWork& w, int tc, int elem_bytes) {
    ItemView v;
    v.A = (const GLB char*)w.A + (int64_t)w.ktile * w.a_ktile_stride * elem_bytes;
    v.chunks = (const GLB uint32_t*)w.chunks;
    v.fT = (const GLB double*)w.fT;
    v.B = (const GLB void*)w.B;
    v.b_lo = w.b_lo;
    v.part = (GLB double*)w.part;
    v.c_begin = w.c_begin;
    v.c_end = w.c_end;
    v.krow0 = w.ktile * TK;
    v.row_base = w.row_base;
    v.row_pad_base = w.row_pad_base;
    v.rows = w.rows;
    v.cells = w.cells;
    v.ldc = w.ldc;
    v.ldb = w.ldb;
    v.N = w.N;
    v.c0 = w.ctile * tc;
    v.kind = w.kind;
    return v;
}

// Row (within the 64-row tile) of accumulator element e of a 16x16 tile at row16.
template <int ROWSTEP>
__device__ inline int tile_row(int row16, int lane, int e) {
    return row16 + (ROWSTEP == 1 ? 4 * (lane >> 4) + e : (lane >> 4) + 4 * e);
}

// Write an item's 64 x TC tile of FP64 partial sums through LDS.  Per column c of 16-cell
// tiles, every wave puts its result tiles (val(r, c, e) at row tile_row<ROWSTEP>, cell l % 16;
// LOSS rows multiplied by their own f) into T cell-major; then lane l writes row l of one cell
// per store, 512 contiguous bytes instead of 16 pieces of 32.  T (WC x 16 x EPI_PITCH doubles)
// may alias the staging tiles once the main loop has ended with a barrier.
template <int TC, int ROWSTEP, typename Get>
__device__ inline void store_item(const ItemView& v, double (*T)[16 * EPI_PITCH], Get val) {
    using TL = Tiling<TC>;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wk = wave % TL::WK, wc = wave / TL::WK;
    const bool loss = v.kind == CH_LOSS;
    const int krow_out = v.krow0 + lane;
#pragma unroll
    for (int c = 0; c < TL::MC; ++c) {
        const int cl = wc * TL::WCOLS + c * 16 + (lane & 15);   // cell within the tile
        double* t = T[wc] + (lane & 15) * EPI_PITCH;
#pragma unroll
        for (int r = 0; r < TL::MR; ++r)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int row = tile_row<ROWSTEP>(wk * TL::WROWS + r * 16, lane, e);
                double x = val(r, c, e);
                if (loss && v.krow0 + row < v.rows && v.c0 + cl < v.cells)
                    x = x * v.fT[(int64_t)(v.row_pad_base + v.krow0 + row) * v.ldc + v.c0 + cl];
                t[row] = x;
            }
        __syncthreads();
        // the 16 WC cells of this column of tiles, every fourth to each wave
        for (int q = wave; q < 16 * TL::WC; q += 4) {
            const int g = q >> 4, ci = q & 15;
            const int cell = v.c0 + g * TL::WCOLS + c * 16 + ci;
            if (cell < v.cells && krow_out < v.rows)
                v.part[(int64_t)cell * v.N + v.row_base + krow_out] = T[g][ci * EPI_PITCH + lane];
        }
        __syncthreads();
    }
}

// f_i of the lane's column in the cell tile (row-end multiplier).
__device__ inline double f_at(const ItemView& v, int ip, int col) {
    return v.fT[(int64_t)ip * v.ldc + v.c0 + col];
}

// Copy the item's slice of the chunk table into LDS, so the main loop never waits on a global
// load for its control flow.
__device__ inline void load_ctab(const ItemView& v, uint32_t* ctab) {
    for (int q = threadIdx.x; q < v.c_end - v.c_begin; q += blockDim.x)
        ctab[q] = v.chunks[v.c_begin + q];
    __syncthreads();
}

// Direct-to-LDS load (global_load_lds): lane l of the wave lands at lds + BYTES l.
#define DTX_LDS(p) ((__attribute__((address_space(3))) void*)(p))

template <int BYTES>
__device__ __forceinline__ void dma_lds(const GLB void* src, void* lds) {
    static_assert(BYTES == 16 || BYTES == 4, "16- or 4-byte lanes");
    if constexpr (BYTES == 16) __builtin_amdgcn_global_load_lds((GLB void*)src, DTX_LDS(lds), 16, 0, 0);
    else __builtin_amdgcn_global_load_lds((GLB void*)src, DTX_LDS(lds), 4, 0, 0);
}

// ------------------------------------------------------------------------------- FP64 -----
// The row multiplier f_i is folded into the B tile while it is staged: B'[t][c] = f_i^c f_j^c
// (LOSS channels stage f_j^c and multiply by their row's f at the end), so the accumulators
// hold the whole item's contribution and no epilogue runs inside the loop.  Each 32-j chunk is
// staged through LDS in two halves of 16 j, double-buffered (32 KB of LDS).
//
// With 64-cell tiles both tiles sit in LDS in MFMA operand order, as in the wide FP64 item: one
// 64-double block per (k step of 4, 16 rows or cells), element l of a block being what lane l
// supplies, so every LDS store and operand read is a linear 512-byte block (the row-major
// staging stores cost about 7 bank-conflict cycles per instruction; the 16-cell shape, which
// streams its operator from HBM, keeps them, as operand order measured 3% slower there).
template <int TC>
struct alignas(16) SmemF64 {
    static constexpr int HK = KC / 2;   // j per LDS stage
    union {
        struct {
            double As[2][HK * TK];
            double Bs[2][HK * TC];
        } s;
        double T[Tiling<TC>::WC][16 * EPI_PITCH];   // epilogue transpose
    };
    uint32_t ctab[MAX_RANGE];
};

template <int TC>
__device__ __forceinline__ void f64_item(const BinWork& w, SmemF64<TC>& sm) {
    using T = Tiling<TC>;
    constexpr int HK = SmemF64<TC>::HK;
    constexpr int AP = HK * TK / 256;   // 4 A elements per thread
    constexpr int BP = HK * TC / 256;   // B elements per thread (consecutive cells)
    constexpr int BT = TC / BP;         // threads per B row
    auto& As = sm.s.As;
    auto& Bs = sm.s.Bs;
    uint32_t* ctab = sm.ctab;

    const ItemView v = load_item(w, TC, 8);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wk = wave % T::WK, wc = wave / T::WK;
    const GLB double* A = (const GLB double*)v.A + (int64_t)v.c_begin * KC * TK;
    const GLB double* B = (const GLB double*)v.B;
    const int bq = tid / BT, bcc = (tid % BT) * BP;
    const bool loss = v.kind >= CH_LOSS;   // B = f_j; LOSS and MOMENT rows take their own f
    // operand order (TC = 64), in 16-byte pieces: elements q 512 + 2 tid and the next (q = 0, 1)
    // of a stage tile are in block 8 q + tid / 32 = 4 ks + b with k step ks = 2 q + tid / 128 and
    // row / cell block b = (tid / 32) % 4, at lanes l0 = 2 (tid % 32) and l0 + 1: the same k,
    // 4 ks + l0 / 16, and consecutive rows (A: 16 b + l0 % 16) or cells (B, likewise)
    constexpr bool frag = TC == 64;
    const int l0 = 2 * (tid & 31);
    const int prow = 16 * ((tid >> 5) & 3) + (l0 & 15), pk = 4 * (tid >> 7) + (l0 >> 4);
    load_ctab(v, ctab);

    double ra[AP], rb[BP], rf[BP];
    d2 pb[2], pf = d2{0.0, 0.0};
    // stage st = half (st & 1) of local chunk (st >> 1), into buffer buf (64-cell tiles: the
    // operator straight into LDS, lane l of wave w at piece q's 128 w + 2 l, where the register
    // path stored it; the B' stores wait for the state loads issued after it)
    auto load_stage = [&](int st, int buf) {
        const int ch = st >> 1, h = st & 1;
        const uint32_t ce = ctab[ch];
        const int j0 = (int)(ce >> 16) + h * HK;
        if constexpr (frag) {
            const GLB double* src = A + (int64_t)ch * KC * TK + (int64_t)(h * HK + pk) * TK + prow;
#pragma unroll
            for (int q = 0; q < 2; ++q)
                dma_lds<16>(src + 8 * q * TK, &As[buf][q * 512 + 128 * wave]);
            const GLB double* b = B + (int64_t)(j0 + pk) * v.ldc + v.c0 + prow;
#pragma unroll
            for (int q = 0; q < 2; ++q) pb[q] = *(const GLB d2*)(b + (int64_t)(8 * q) * v.ldc);
            if (!loss) pf = *(const GLB d2*)(v.fT + (int64_t)(ce & 0xFFFFu) * v.ldc + v.c0 + prow);
        } else {
            const GLB double* src = A + (int64_t)ch * KC * TK + h * HK * TK + tid * AP;
#pragma unroll
            for (int q = 0; q < AP; ++q) ra[q] = src[q];
            const GLB double* b = B + (int64_t)(j0 + bq) * v.ldc + v.c0 + bcc;
#pragma unroll
            for (int q = 0; q < BP; ++q) rb[q] = b[q];
            if (!loss) {
                const GLB double* fi = v.fT + (int64_t)(ce & 0xFFFFu) * v.ldc + v.c0 + bcc;
#pragma unroll
                for (int q = 0; q < BP; ++q) rf[q] = fi[q];
            }
        }
    };
    auto store_stage = [&](int buf) {
        if constexpr (frag) {
#pragma unroll
            for (int q = 0; q < 2; ++q) *(d2*)&Bs[buf][q * 512 + 2 * tid] = loss ? pb[q] : pf * pb[q];
        } else {
#pragma unroll
            for (int q = 0; q < AP; ++q) As[buf][tid * AP + q] = ra[q];
#pragma unroll
            for (int q = 0; q < BP; ++q)
                Bs[buf][bq * TC + bcc + q] = loss ? rb[q] : rf[q] * rb[q];
        }
    };

    d4 acc[T::MR][T::MC];
#pragma unroll
    for (int r = 0; r < T::MR; ++r)
#pragma unroll
        for (int c = 0; c < T::MC; ++c) acc[r][c] = d4{0.0, 0.0, 0.0, 0.0};
    const int nst = 2 * (v.c_end - v.c_begin);
    if (nst > 0) {
        load_stage(0, 0);
        store_stage(0);
    }
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    int buf = 0;
    for (int st = 0; st < nst; ++st) {
        const bool more = st + 1 < nst;
        if (more) load_stage(st + 1, buf ^ 1);
#pragma unroll
        for (int kk = 0; kk < HK; kk += 4) {
            const int kr = kk + (lane >> 4);
            double a[T::MR], b[T::MC];
#pragma unroll
            for (int r = 0; r < T::MR; ++r)
                a[r] = frag ? As[buf][((kk / 4) * (TK / 16) + wk * T::MR + r) * 64 + lane]
                            : As[buf][kr * TK + wk * T::WROWS + r * 16 + (lane & 15)];
#pragma unroll
            for (int c = 0; c < T::MC; ++c)
                b[c] = frag ? Bs[buf][((kk / 4) * (TC / 16) + wc * T::MC + c) * 64 + lane]
                            : Bs[buf][kr * TC + wc * T::WCOLS + c * 16 + (lane & 15)];
#pragma unroll
            for (int r = 0; r < T::MR; ++r)
#pragma unroll
                for (int c = 0; c < T::MC; ++c)
                    acc[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[r], b[c], acc[r][c], 0, 0, 0);
        }
        if (more) store_stage(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
    if (v.kind == CH_MOMENT) {
        // per cell, the sum over the item's rows of f_row (W f)_row: each lane adds its rows of
        // the column (tiles r, elements e in order), the four lanes of a column combine by
        // butterfly, and the WK waves of a column in wave order through LDS (As is free now)
        double cs[T::MC];
#pragma unroll
        for (int c = 0; c < T::MC; ++c) {
            const int cell = v.c0 + wc * T::WCOLS + c * 16 + (lane & 15);
            cs[c] = 0.0;
#pragma unroll
            for (int r = 0; r < T::MR; ++r)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int krow = v.krow0 + tile_row<4>(wk * T::WROWS + r * 16, lane, e);
                    if (krow < v.rows && cell < v.cells)
                        cs[c] = cs[c] +
                                acc[r][c][e] * v.fT[(int64_t)(v.row_pad_base + krow) * v.ldc + cell];
                }
            cs[c] = cs[c] + xor_lane<16>(cs[c]);
            cs[c] = cs[c] + xor_lane<32>(cs[c]);
        }
        double* msum = &As[0][0];
        if (lane < 16)
#pragma unroll
            for (int c = 0; c < T::MC; ++c) msum[wk * TC + wc * T::WCOLS + c * 16 + lane] = cs[c];
        __syncthreads();
        if (tid < TC && v.c0 + tid < v.cells) {
            double t = msum[tid];
            for (int k = 1; k < T::WK; ++k) t = t + msum[k * TC + tid];
            v.part[v.c0 + tid] = t;
        }
        return;
    }
    store_item<TC, 4>(v, sm.T, [&](int r, int c, int e) { return acc[r][c][e]; });
}

template <int TC>
__global__ __launch_bounds__(256) void bin_f64_kernel(const BinWork* __restrict__ items) {
    __shared__ SmemF64<TC> sm;
    f64_item<TC>(items[blockIdx.x], sm);
}

// ------------------------------------------------------------------------------- FP32 -----
// As FP64, with B'[t][c] = (float)(f_i^c * f_j^c) formed in FP64 and rounded once, FP32 MFMA
// accumulation, and the FP32 accumulators folded into FP64 every FLUSH_CHUNKS chunks.
constexpr int FLUSH_CHUNKS = 16;

template <int TC>
struct alignas(16) SmemF32 {
    union {
        struct {
            float As[2][KC * TK];
            float Bs[2][KC * TC];
        } s;
        double T[Tiling<TC>::WC][16 * EPI_PITCH];   // epilogue transpose
    };
    uint32_t ctab[MAX_RANGE];
};

template <int TC>
__device__ __forceinline__ void f32_item(const BinWork& w, SmemF32<TC>& sm) {
    using T = Tiling<TC>;
    auto& As = sm.s.As;
    auto& Bs = sm.s.Bs;
    uint32_t* ctab = sm.ctab;

    const ItemView v = load_item(w, TC, 4);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wk = wave % T::WK, wc = wave / T::WK;
    const GLB float* A = (const GLB float*)v.A + (int64_t)v.c_begin * KC * TK;
    const GLB float* B = (const GLB float*)v.B;   // FP32 f_j; f_i from the FP64 fT
    // Operand order in 16-byte pieces: piece p of a chunk tile holds elements 4 p .. 4 p + 3 of
    // block p / 16 = ks RB + blk (A) or ks CB + blk (B), k step ks and row or cell block blk, at
    // lanes l0 = 4 (p % 16) .. l0 + 3: the same k, 4 ks + l0 / 16, and four consecutive rows or
    // cells from 16 blk + l0 % 16.  Every LDS store and operand read is a linear block, free of
    // bank conflicts (row-major tiles cost about 7.5 conflict cycles per LDS instruction).
    constexpr int RB = TK / 16, CB = TC / 16;          // row / cell blocks
    constexpr int APC = KC * TK / 4 / 256;             // A pieces per thread (2)
    constexpr int BPC = (KC * TC / 4 + 255) / 256;     // B pieces per thread (2, or 1 for TC 16)
    constexpr int BPIECES = KC * TC / 4;
    int a_off[APC], b_k[BPC], b_cell[BPC];
#pragma unroll
    for (int q = 0; q < APC; ++q) {
        const int p = q * 256 + tid, blk = p >> 4, l0 = 4 * (p & 15);
        a_off[q] = (4 * (blk / RB) + (l0 >> 4)) * TK + 16 * (blk % RB) + (l0 & 15);
    }
#pragma unroll
    for (int q = 0; q < BPC; ++q) {
        const int p = q * 256 + tid, blk = p >> 4, l0 = 4 * (p & 15);
        b_k[q] = 4 * (blk / CB) + (l0 >> 4);
        b_cell[q] = 16 * (blk % CB) + (l0 & 15);
    }
    load_ctab(v, ctab);

    f4 ra[APC], rb[BPC];
    d4 rf[BPC];
    auto load_chunk = [&](int ch) {
        const uint32_t ce = ctab[ch];
        const GLB float* src = A + (int64_t)ch * KC * TK;
#pragma unroll
        for (int q = 0; q < APC; ++q) ra[q] = *(const GLB f4*)(src + a_off[q]);
        const GLB double* fi = v.fT + (int64_t)(ce & 0xFFFFu) * v.ldc + v.c0;
#pragma unroll
        for (int q = 0; q < BPC; ++q) {
            if (BPIECES % 256 == 0 || q * 256 + tid < BPIECES) {
                rb[q] = *(const GLB f4*)(B + (int64_t)((int)(ce >> 16) + b_k[q]) * v.ldc + v.c0 +
                                         b_cell[q]);
                rf[q] = *(const GLB d4*)(fi + b_cell[q]);
            }
        }
    };
    auto store_chunk = [&](int buf) {
#pragma unroll
        for (int q = 0; q < APC; ++q) *(f4*)&As[buf][4 * (q * 256 + tid)] = ra[q];
#pragma unroll
        for (int q = 0; q < BPC; ++q)
            if (BPIECES % 256 == 0 || q * 256 + tid < BPIECES) {
                const d4 x = rf[q] * d4{(double)rb[q][0], (double)rb[q][1], (double)rb[q][2],
                                        (double)rb[q][3]};
                *(f4*)&Bs[buf][4 * (q * 256 + tid)] =
                    f4{(float)x[0], (float)x[1], (float)x[2], (float)x[3]};
            }
    };

    f4 acc[T::MR][T::MC];
    double tot[T::MR][T::MC][4];
#pragma unroll
    for (int r = 0; r < T::MR; ++r)
#pragma unroll
        for (int c = 0; c < T::MC; ++c) {
            acc[r][c] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int e = 0; e < 4; ++e) tot[r][c][e] = 0.0;
        }
    const int nch = v.c_end - v.c_begin;
    if (nch > 0) {
        load_chunk(0);
        store_chunk(0);
    }
    __syncthreads();
    int buf = 0;
    for (int ch = 0; ch < nch; ++ch) {
        const bool more = ch + 1 < nch;
        if (more) load_chunk(ch + 1);
#pragma unroll
        for (int kk = 0; kk < KC; kk += 4) {
            float a[T::MR], b[T::MC];
#pragma unroll
            for (int r = 0; r < T::MR; ++r)
                a[r] = As[buf][((kk / 4) * RB + wk * T::MR + r) * 64 + lane];
#pragma unroll
            for (int c = 0; c < T::MC; ++c)
                b[c] = Bs[buf][((kk / 4) * CB + wc * T::MC + c) * 64 + lane];
#pragma unroll
            for (int r = 0; r < T::MR; ++r)
#pragma unroll
                for (int c = 0; c < T::MC; ++c)
                    acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[r], b[c], acc[r][c], 0, 0, 0);
        }
        if ((ch + 1) % FLUSH_CHUNKS == 0 || !more) {
#pragma unroll
            for (int r = 0; r < T::MR; ++r)
#pragma unroll
                for (int c = 0; c < T::MC; ++c) {
#pragma unroll
                    for (int e = 0; e < 4; ++e) tot[r][c][e] += (double)acc[r][c][e];
                    acc[r][c] = f4{0.f, 0.f, 0.f, 0.f};
                }
        }
        if (more) store_chunk(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
    store_item<TC, 1>(v, sm.T, [&](int r, int c, int e) { return tot[r][c][e]; });
}

// LOSS items (FP64 in every mode) run the FP64 body inside the same launch.
template <int TC>
__global__ __launch_bounds__(256) void bin_f32_kernel(const BinWork* __restrict__ items) {
    __shared__ union {
        SmemF64<TC> d;
        SmemF32<TC> s;
    } sm;
    const BinWork& w = items[blockIdx.x];
    if (w.kind >= CH_LOSS) f64_item<TC>(w, sm.d);
    else f32_item<TC>(w, sm.s);
}

// ------------------------------------------------------------------------ split BF16 x3 -----
// B is the cell-major hi/lo split of f (no per-element work while staging); the row multiplier
// f_i is staged in LDS beside each chunk and folded into FP64 when the row ends.  Packed A per
// k-tile and chunk: hi then lo, each 64 rows x 32 j in MFMA fragment order (bf16_frag_index).
// The B tile is staged in the same order (16 cells x 8 j per lane quad), so LDS is written and
// read as whole linear 1 KB blocks: no bank conflicts, and no padding.
template <int TC>
struct SmemBF16 {
    union {
        struct {
            u32x4 Ah[2][TK * KC / 8];   // 16-byte lane quads: [row block][lane]
            u32x4 Al[2][TK * KC / 8];
            u32x4 Bh[2][TC * KC / 8];   // [cell block][lane]
            u32x4 Bl[2][TC * KC / 8];
            double Fs[2][TC];
        } s;
        double T[Tiling<TC>::WC][16 * EPI_PITCH];   // epilogue transpose
    };
    uint32_t ctab[MAX_RANGE];
};

template <int TC>
__device__ __forceinline__ void bf16x3_item(const BinWork& w, SmemBF16<TC>& sm) {
    using T = Tiling<TC>;
    auto& Ah = sm.s.Ah;
    auto& Al = sm.s.Al;
    auto& Bh = sm.s.Bh;
    auto& Bl = sm.s.Bl;
    auto& Fs = sm.s.Fs;
    uint32_t* ctab = sm.ctab;

    const ItemView v = load_item(w, TC, 2);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wk = wave % T::WK, wc = wave / T::WK;
    const GLB uint16_t* A = (const GLB uint16_t*)v.A + (int64_t)v.c_begin * TK * KC * 2;
    const GLB uint16_t* Bhi = (const GLB uint16_t*)v.B;
    const GLB uint16_t* Blo = Bhi + v.b_lo;
    // B quad of this thread: lane tid % 64 of cell block tid / 64 in the fragment-ordered split
    // of f (fb_index): one linear 1 KB block per wave.
    const bool bthread = tid < TC * 4;
    const int64_t bq0 = (int64_t)(v.c0 >> 4) * 64 + tid;
    load_ctab(v, ctab);

    u32x4 rah, ral, rbh, rbl;
    double rfi = 0.0;
    auto load_chunk = [&](int ch) {
        const uint32_t ce = ctab[ch];
        const GLB u32x4* src = (const GLB u32x4*)(A + (int64_t)ch * TK * KC * 2);
        rah = src[tid];
        ral = src[256 + tid];
        if (bthread) {
            const int64_t off = (int64_t)((ce >> 16) >> 5) * v.ldb * 64 + bq0;
            rbh = ((const GLB u32x4*)Bhi)[off];
            rbl = ((const GLB u32x4*)Blo)[off];
        }
        if (tid < TC) rfi = v.fT[(int64_t)(ce & 0xFFFFu) * v.ldc + v.c0 + tid];
    };
    auto store_chunk = [&](int buf) {
        Ah[buf][tid] = rah;
        Al[buf][tid] = ral;
        if (bthread) {
            Bh[buf][tid] = rbh;
            Bl[buf][tid] = rbl;
        }
        if (tid < TC) Fs[buf][tid] = rfi;
    };

    f4 acc[T::MR][T::MC];
    double tot[T::MR][T::MC][4];
#pragma unroll
    for (int r = 0; r < T::MR; ++r)
#pragma unroll
        for (int c = 0; c < T::MC; ++c) {
            acc[r][c] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int e = 0; e < 4; ++e) tot[r][c][e] = 0.0;
        }
    const int nch = v.c_end - v.c_begin;
    if (nch > 0) {
        load_chunk(0);
        store_chunk(0);
    }
    __syncthreads();
    int buf = 0;
    for (int ch = 0; ch < nch; ++ch) {
        const bool more = ch + 1 < nch;
        if (more) load_chunk(ch + 1);
        bf16x8 ah[T::MR], al[T::MR], bh[T::MC], bl[T::MC];
#pragma unroll
        for (int r = 0; r < T::MR; ++r) {
            const int q = (wk * T::MR + r) * 64 + lane;  // row block wk*MR + r
            ah[r] = __builtin_bit_cast(bf16x8, Ah[buf][q]);
            al[r] = __builtin_bit_cast(bf16x8, Al[buf][q]);
        }
#pragma unroll
        for (int c = 0; c < T::MC; ++c) {
            const int q = (wc * T::MC + c) * 64 + lane;  // cell block wc*MC + c
            bh[c] = __builtin_bit_cast(bf16x8, Bh[buf][q]);
            bl[c] = __builtin_bit_cast(bf16x8, Bl[buf][q]);
        }
#pragma unroll
        for (int r = 0; r < T::MR; ++r)
#pragma unroll
            for (int c = 0; c < T::MC; ++c) {
                acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[r], bh[c], acc[r][c], 0, 0, 0);
                acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[r], bl[c], acc[r][c], 0, 0, 0);
                acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[r], bh[c], acc[r][c], 0, 0, 0);
            }
        const uint32_t ip = ctab[ch] & 0xFFFFu;
        if (!more || (ctab[ch + 1] & 0xFFFFu) != ip) {
            // the row ends: tot += f_i * acc in FP64
#pragma unroll
            for (int c = 0; c < T::MC; ++c) {
                const double fi = Fs[buf][wc * T::WCOLS + c * 16 + (lane & 15)];
#pragma unroll
                for (int r = 0; r < T::MR; ++r) {
#pragma unroll
                    for (int e = 0; e < 4; ++e) tot[r][c][e] += fi * (double)acc[r][c][e];
                    acc[r][c] = f4{0.f, 0.f, 0.f, 0.f};
                }
            }
        }
        if (more) store_chunk(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
    store_item<TC, 1>(v, sm.T, [&](int r, int c, int e) { return tot[r][c][e]; });
}

template <int TC>
__global__ __launch_bounds__(256) void bin_bf16x3_kernel(const BinWork* __restrict__ items) {
    __shared__ union {
        SmemF64<TC> d;
        SmemBF16<TC> b;
    } sm;
    const BinWork& w = items[blockIdx.x];
    if (w.kind >= CH_LOSS) f64_item<TC>(w, sm.d);
    else bf16x3_item<TC>(w, sm.b);
}

// --------------------------------------------------------------- wide split BF16 x3 -----
// For many cells (C > 64) the 64 x 64 tile moves too many bytes per MFMA through global loads,
// LDS and barriers.  A wide item covers NS packed 64-row slices (k tiles NS w .. NS w + NS - 1)
// x 128 cells with 2 NS waves, each owning a 64 x 64 block (16 MFMA tiles): every staged A
// byte feeds two waves and every B byte NS waves.  NS = 4 (eight waves, one workgroup per CU)
// halves the B bytes per MFMA again but needs enough items to fill the chip; NS = 2 (four
// waves, two workgroups per CU) serves fewer cells.  Both operands are staged as plain 16-byte
// copies in MFMA fragment order: A from the packed operator, B from the cell-major hi/lo split
// of f.  The row multiplier f_i is staged beside each chunk and folded in FP32 when the row
// ends, tot += f_i * acc; tot is written to the item's FP64 slot at the end.
template <int NS>
struct SmemWide {
    union {
        struct {
            u32x4 Ah[2][4 * NS * 64];   // [row block][lane]: blocks 4s..4s+3 from slice s
            u32x4 Al[2][4 * NS * 64];
            u32x4 Bh[2][8 * 64];        // [cell block 0..7][lane]
            u32x4 Bl[2][8 * 64];
            double Fs[2][WIDE];         // f_i of the chunk's row, per cell of the tile
        } s;
        float T[2 * NS][16 * EPI_PITCH];   // epilogue: each wave's 16 cells x 64 rows
    };
    uint32_t ctab[MAX_RANGE];
};

template <int NS>
__device__ __forceinline__ void wide_item(const BinWork& w, SmemWide<NS>& sm) {
    constexpr int NT = 128 * NS;      // threads
    constexpr int BQ = 4 / NS;        // B cell blocks staged per thread
    auto& Ah = sm.s.Ah;
    auto& Al = sm.s.Al;
    auto& Bh = sm.s.Bh;
    auto& Bl = sm.s.Bl;
    auto& Fs = sm.s.Fs;
    uint32_t* ctab = sm.ctab;
    const ItemView v = load_item(w, WIDE, 2);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wk = wave % NS, wc = wave / NS;   // this wave: rows 64 wk.., cells 64 wc..
    // A quads of this thread: slices sa and sa + NS / 2 (sa = tid / 256), quad tid % 256 of
    // each slice's hi and lo halves.  Slice s of the item is k tile NS w.ktile + s; a slice
    // beyond the channel's rows reads slice 0 again, and the rows it feeds are never written.
    const int64_t slice = w.a_ktile_stride;     // bf16 elements per packed 64-row slice
    const int kt0 = NS * w.ktile;
    const int sa = tid >> 8, qa = tid & 255;
    auto slice_ptr = [&](int s) {
        const bool present = (kt0 + s) * TK < v.rows;
        return (const GLB u32x4*)((const GLB uint16_t*)((const GLB char*)w.A) +
                                  (int64_t)(present ? kt0 + s : kt0) * slice +
                                  (int64_t)v.c_begin * TK * KC * 2) + qa;
    };
    const GLB u32x4* As0 = slice_ptr(sa);
    const GLB u32x4* As1 = slice_ptr(sa + NS / 2);
    const GLB uint16_t* Bhi = (const GLB uint16_t*)v.B;
    const GLB uint16_t* Blo = Bhi + v.b_lo;
    // B quads of this thread: cell blocks wave + 2 NS q (q < BQ), fragment lane `lane` of each
    // (cell lane % 16 of the block, j octet lane / 16); the four lanes of one cell read one
    // contiguous 64-byte segment, and the LDS stores are linear 1 KB blocks
    const int64_t boff = ((int64_t)(v.c0 >> 4) + wave) * 64 + lane;   // u32x4 of quad 0
    constexpr int bstep = 2 * NS * 64;                                 // u32x4 between quads
    load_ctab(v, ctab);

    u32x4 rh0, rl0, rh1, rl1;   // A quads (slices sa, sa + NS / 2)
    u32x4 bh[BQ], bl[BQ];       // B quads
    double rfi = 0.0;
    auto load_chunk = [&](int ch) {
        const uint32_t ce = ctab[ch];
        const int64_t co = (int64_t)ch * (TK * KC * 2 / 8);   // u32x4 per slice and chunk
        rh0 = As0[co];
        rl0 = As0[co + 256];
        rh1 = As1[co];
        rl1 = As1[co + 256];
        const int j0 = (int)(ce >> 16);
#pragma unroll
        for (int q = 0; q < BQ; ++q) {
            const int64_t off = (int64_t)(j0 >> 5) * v.ldb * 64 + boff + q * bstep;
            bh[q] = ((const GLB u32x4*)Bhi)[off];
            bl[q] = ((const GLB u32x4*)Blo)[off];
        }
        if (tid < WIDE) rfi = v.fT[(int64_t)(ce & 0xFFFFu) * v.ldc + v.c0 + tid];
    };
    auto store_chunk = [&](int buf) {
        Ah[buf][sa * 256 + qa] = rh0;
        Al[buf][sa * 256 + qa] = rl0;
        Ah[buf][(sa + NS / 2) * 256 + qa] = rh1;
        Al[buf][(sa + NS / 2) * 256 + qa] = rl1;
#pragma unroll
        for (int q = 0; q < BQ; ++q) {
            Bh[buf][q * NT + tid] = bh[q];
            Bl[buf][q * NT + tid] = bl[q];
        }
        if (tid < WIDE) Fs[buf][tid] = rfi;
    };

    f4 acc[4][4], tot[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] = tot[r][c] = f4{0.f, 0.f, 0.f, 0.f};
    const int nch = v.c_end - v.c_begin;
    if (nch > 0) {
        load_chunk(0);
        store_chunk(0);
    }
    __syncthreads();
    int buf = 0;
    for (int ch = 0; ch < nch; ++ch) {
        const bool more = ch + 1 < nch;
        if (more) load_chunk(ch + 1);
        bf16x8 ah[4], al[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            ah[r] = __builtin_bit_cast(bf16x8, Ah[buf][(4 * wk + r) * 64 + lane]);
            al[r] = __builtin_bit_cast(bf16x8, Al[buf][(4 * wk + r) * 64 + lane]);
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const bf16x8 bhc = __builtin_bit_cast(bf16x8, Bh[buf][(4 * wc + c) * 64 + lane]);
            const bf16x8 blc = __builtin_bit_cast(bf16x8, Bl[buf][(4 * wc + c) * 64 + lane]);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[r], bhc, acc[r][c], 0, 0, 0);
                acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[r], blc, acc[r][c], 0, 0, 0);
                acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[r], bhc, acc[r][c], 0, 0, 0);
            }
        }
        const uint32_t ip = ctab[ch] & 0xFFFFu;
        if (!more || (ctab[ch + 1] & 0xFFFFu) != ip) {
            // the row ends: tot += f_i * acc (FP32)
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const float fi = (float)Fs[buf][64 * wc + 16 * c + (lane & 15)];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        tot[r][c][e] = __builtin_fmaf(fi, acc[r][c][e], tot[r][c][e]);
                    acc[r][c] = f4{0.f, 0.f, 0.f, 0.f};
                }
            }
        }
        if (more) store_chunk(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
    // write the item's partial sums, FP32 as they are (the slot is read as floats): per group
    // of 16 cells each wave puts its FP32 result tiles (layout D[4 (l / 16) + e][l % 16]) into
    // LDS cell-major, then lane l writes row l of one cell per store, 256 contiguous bytes
    float* tw = sm.T[wave];
    const int krow = kt0 * TK + 64 * wk + lane;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int e = 0; e < 4; ++e)
                tw[(lane & 15) * EPI_PITCH + 16 * r + 4 * (lane >> 4) + e] = tot[r][c][e];
        __syncthreads();
        for (int q = 0; q < 16; ++q) {
            const int cell = v.c0 + 64 * wc + 16 * c + q;
            if (cell < v.cells && krow < v.rows)
                ((GLB float*)v.part)[(int64_t)cell * v.N + v.row_base + krow] =
                    tw[q * EPI_PITCH + lane];
        }
        __syncthreads();
    }
}

// The wide item with the operator loaded straight into LDS (global_load_lds_dwordx4: lane l of
// a wave's load lands at LDS base + 16 l, so each wave moves one 1 KB fragment-order block of
// the packed operator per instruction, with no registers in between).  Three operator stages
// let the loads of chunk ch + 2 be in flight while chunk ch computes and chunk ch + 1 arrives;
// with two workgroups per CU (NS = 2) a register stage of that depth does not fit in the 256
// VGPRs of two waves per SIMD.  The state split (B, from L2) and the row multipliers f_i come
// the same way, B one chunk ahead: into a single buffer, free once every wave has read its
// operands, at NS = 2 (two stages would not fit two workgroups in the LDS), and into two
// alternating buffers at NS = 4.  Per chunk each wave issues 4 operator loads (16 bytes a lane),
// 1 f_i load (4 bytes a lane: four waves move the 1 KB of a chunk's f_i, and at NS = 4 the
// other four load a second copy, so every wave issues the same loads; with the 2-byte form,
// 2 bytes a lane over eight waves, the multipliers came out wrong) and 8 / NS state loads, so
// its wait counts
// are uniform: at the top of an iteration only the operator group of the next chunk (5 loads)
// may be in flight.
template <int NS>
struct SmemWideDma {
    static constexpr int NB = NS == 4 ? 2 : 1;   // state buffers
    union {
        struct {
            u32x4 A0[2][4 * NS * 64];   // stage 0: hi, lo; 4 NS blocks (slices x row blocks)
            u32x4 A1[2][4 * NS * 64];
            u32x4 A2[2][4 * NS * 64];
            double F0[WIDE * NS / 2];   // f_i of the chunk's row per cell of the tile (NS / 2
            double F1[WIDE * NS / 2];   // copies, the first one read)
            double F2[WIDE * NS / 2];
            u32x4 B[NB][2][8 * 64];     // hi, lo; cell blocks 0..7
        } s;
        float T[2 * NS][16 * EPI_PITCH];   // epilogue: each wave's 16 cells x 64 rows
    };
    uint32_t ctab[MAX_RANGE];
};


template <int NS>
__device__ __forceinline__ void wide_item_dma(const BinWork& w, SmemWideDma<NS>& sm) {
    using Sm = SmemWideDma<NS>;
    constexpr int NB = Sm::NB;
    constexpr int BQ = 8 / NS;        // state blocks a wave
    uint32_t* ctab = sm.ctab;
    const ItemView v = load_item(w, WIDE, 2);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wk = wave % NS, wc = wave / NS;   // this wave: rows 64 wk.., cells 64 wc..
    const int64_t slice = w.a_ktile_stride;     // bf16 elements per packed 64-row slice
    const int kt0 = NS * w.ktile;
    // this wave's operator loads: half (hi / lo) wave / NS of slice wave % NS, its 4 row blocks
    const int dh = wave / NS, ds = wave % NS;
    const bool present = (kt0 + ds) * TK < v.rows;   // else slice 0 again (rows never written)
    const GLB u32x4* Ag = (const GLB u32x4*)((const GLB uint16_t*)((const GLB char*)w.A) +
                                             (int64_t)(present ? kt0 + ds : kt0) * slice +
                                             (int64_t)v.c_begin * TK * KC * 2) +
                          dh * 256 + lane;
    // this wave's state loads: half wave / NS, cell blocks BQ (wave % NS) .. + BQ - 1
    const GLB u32x4* Bg = (const GLB u32x4*)((const GLB uint16_t*)v.B + (dh ? v.b_lo : 0)) +
                          ((int64_t)(v.c0 >> 4) + BQ * ds) * 64 + lane;
    load_ctab(v, ctab);

    auto issue_a = [&](int ch, u32x4 (&A)[2][4 * NS * 64], double* F) {
        const GLB u32x4* src = Ag + (int64_t)ch * (TK * KC * 2 / 8);
#pragma unroll
        for (int r = 0; r < 4; ++r) dma_lds<16>(src + r * 64, &A[dh][(4 * ds + r) * 64]);
        // wave w: bytes 256 (w % 4) .. + 255 of f_i into copy w / 4
        const uint32_t ce = ctab[ch];
        const GLB char* fsrc = (const GLB char*)(v.fT + (int64_t)(ce & 0xFFFFu) * v.ldc + v.c0) +
                               256 * (wave & 3) + 4 * lane;
        dma_lds<4>(fsrc, (char*)F + 1024 * (wave >> 2) + 256 * (wave & 3));
    };
    auto issue_b = [&](int ch) {
        const int j0 = (int)(ctab[ch] >> 16);
        const GLB u32x4* src = Bg + (int64_t)(j0 >> 5) * v.ldb * 64;
#pragma unroll
        for (int q = 0; q < BQ; ++q)
            dma_lds<16>(src + q * 64, &sm.s.B[NB == 1 ? 0 : (ch & 1)][dh][(BQ * ds + q) * 64]);
    };

    f4 acc[4][4], tot[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] = tot[r][c] = f4{0.f, 0.f, 0.f, 0.f};
    const int nch = v.c_end - v.c_begin;
    // one chunk: wait for its loads (only the next chunk's operator group may stay in flight),
    // read the operands, issue the loads of the state one chunk and of the operator two chunks
    // ahead (with one state buffer, after a second barrier that frees it), then the MFMAs and
    // the row-end fold tot += f_i * acc (FP32)
    auto step = [&](int ch, u32x4 (&A)[2][4 * NS * 64], const double* F,
                    u32x4 (&Anext)[2][4 * NS * 64], double* Fnext) {
        if (ch + 1 < nch) asm volatile("s_waitcnt vmcnt(5)\n\ts_barrier" ::: "memory");
        else asm volatile("s_waitcnt vmcnt(0)\n\ts_barrier" ::: "memory");
        const int bb = NB == 1 ? 0 : (ch & 1);
        bf16x8 ah[4], al[4], bh[4], bl[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            ah[r] = __builtin_bit_cast(bf16x8, A[0][(4 * wk + r) * 64 + lane]);
            al[r] = __builtin_bit_cast(bf16x8, A[1][(4 * wk + r) * 64 + lane]);
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            bh[c] = __builtin_bit_cast(bf16x8, sm.s.B[bb][0][(4 * wc + c) * 64 + lane]);
            bl[c] = __builtin_bit_cast(bf16x8, sm.s.B[bb][1][(4 * wc + c) * 64 + lane]);
        }
        if constexpr (NB == 1) asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
        if (ch + 1 < nch) issue_b(ch + 1);
        if (ch + 2 < nch) issue_a(ch + 2, Anext, Fnext);
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[r], bh[c], acc[r][c], 0, 0, 0);
                acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[r], bl[c], acc[r][c], 0, 0, 0);
                acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[r], bh[c], acc[r][c], 0, 0, 0);
            }
        const uint32_t ip = ctab[ch] & 0xFFFFu;
        if (ch + 1 >= nch || (ctab[ch + 1] & 0xFFFFu) != ip) {
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const float fi = (float)F[64 * wc + 16 * c + (lane & 15)];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        tot[r][c][e] = __builtin_fmaf(fi, acc[r][c][e], tot[r][c][e]);
                    acc[r][c] = f4{0.f, 0.f, 0.f, 0.f};
                }
            }
        }
    };
    if (nch > 0) {
        issue_a(0, sm.s.A0, sm.s.F0);
        issue_b(0);
        if (nch > 1) issue_a(1, sm.s.A1, sm.s.F1);
    }
    // chunk ch uses stage ch % 3 and loads chunk ch + 2 into stage (ch + 2) % 3 (unrolled by
    // three, so every stage is a fixed LDS array and the compiler's own waits can tell them
    // apart)
    for (int ch = 0; ch < nch; ch += 3) {
        step(ch, sm.s.A0, sm.s.F0, sm.s.A2, sm.s.F2);
        if (ch + 1 >= nch) break;
        step(ch + 1, sm.s.A1, sm.s.F1, sm.s.A0, sm.s.F0);
        if (ch + 2 >= nch) break;
        step(ch + 2, sm.s.A2, sm.s.F2, sm.s.A1, sm.s.F1);
    }
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();   // the epilogue's T aliases the stages
    // write the item's FP32 partial sums through LDS, as wide_item does
    float* tw = sm.T[wave];
    const int krow = kt0 * TK + 64 * wk + lane;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int e = 0; e < 4; ++e)
                tw[(lane & 15) * EPI_PITCH + 16 * r + 4 * (lane >> 4) + e] = tot[r][c][e];
        __syncthreads();
        for (int q = 0; q < 16; ++q) {
            const int cell = v.c0 + 64 * wc + 16 * c + q;
            if (cell < v.cells && krow < v.rows)
                ((GLB float*)v.part)[(int64_t)cell * v.N + v.row_base + krow] =
                    tw[q * EPI_PITCH + lane];
        }
        __syncthreads();
    }
}

// LOSS items take the 256-thread FP64 body.  With eight waves the upper four end at once, and
// the barriers of the body then wait for the four that remain (a workgroup barrier counts only
// waves that have not ended).
template <int NS>
__global__ __launch_bounds__(128 * NS) void bin_bf16x3_wide_kernel(
    const BinWork* __restrict__ items) {
    __shared__ union {
        SmemF64<64> d;
        SmemWide<NS> b;
        SmemWideDma<NS> m;
    } sm;
    const BinWork& w = items[blockIdx.x];
    if (w.kind >= CH_LOSS) {
        if (threadIdx.x >= 256) return;
        f64_item<64>(w, sm.d);
    } else {
        wide_item_dma<NS>(w, sm.m);
    }
}


// ------------------------------------------------------------------------- wide FP64 -----
// FP64 items for C > 64: two packed 64-row slices (k tiles 2w and 2w+1) x 128 cells, per group
// of four waves 64 x 64 each (16 tiles of v_mfma_f64_16x16x4_f64), so every staged byte feeds
// twice the MFMAs of a 64 x 64 item.  B' = f_i f_j is formed while staging, as in f64_item
// (LOSS items stage f_j and multiply each row by its own f when written).  Both tiles sit in
// LDS in MFMA operand order: one 64-double block per (k step of 4, 16 rows or 16 cells),
// element l of a block being what lane l supplies (A[row l % 16][k l / 16],
// B[k l / 16][cell l % 16]), so every LDS store and operand read of a wave is one linear
// 512-byte block.
//
// With G groups the item's chunk range is cut into G consecutive shares, one per group, each
// with its own double-buffered stages; the groups meet at the same barriers.  At the end group
// g hands the other group the cell blocks it will not write, and adds the ones it receives:
// the item's sum is share 0 + share 1, and one partial slot covers G shares.
template <int G>
struct alignas(16) SmemWide64 {
    static constexpr int HK = KC / 2;   // j per LDS stage
    union {
        struct {
            double As[G][2][HK * 2 * TK];  // block ks * 8 + row block (0..7)
            double Bs[G][2][HK * WIDE];    // block ks * 8 + cell block (0..7)
        } s;
        double X[4 * G][32 * 64];          // the exchanged half of each wave's tiles (G = 2)
        double T[4 * G][16 * EPI_PITCH];   // epilogue: each wave's 16 cells x 64 rows
    };
    uint32_t ctab[MAX_RANGE];
    uint32_t arrived[G];   // per group: waves arrived at its stage barriers so far
};

// Stage barrier of one group of four waves (G > 1): the workgroup barrier would hold every
// group at each stage, so that no group's MFMAs cover another's barrier.  Each wave waits for
// its own LDS reads and writes, adds one to the group's counter and waits until all four have
// arrived at this barrier (target = 4 x barriers so far).  The four waves of a group run the
// same stages and every wave of the workgroup is resident, so every wait ends.
__device__ __forceinline__ void group_barrier(uint32_t* arrived, uint32_t target) {
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    if ((threadIdx.x & 63) == 0)
        __hip_atomic_fetch_add(arrived, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    while (__hip_atomic_load(arrived, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) < target)
        __builtin_amdgcn_s_sleep(1);
    asm volatile("" ::: "memory");
}

template <int G>
__device__ __forceinline__ void wide64_item(const BinWork& w, SmemWide64<G>& sm) {
    constexpr int HK = SmemWide64<G>::HK;
    uint32_t* ctab = sm.ctab;
    const ItemView v = load_item(w, WIDE, 8);
    const int tid = threadIdx.x & 255, grp = threadIdx.x >> 8;
    const int lane = tid & 63, wave = tid >> 6;
    auto& As = sm.s.As[grp];
    auto& Bs = sm.s.Bs[grp];
    const int wk = wave & 1, wc = wave >> 1;    // this wave: rows 64 wk.., cells 64 wc..
    const bool loss = v.kind == CH_LOSS;
    // 16-byte pieces: elements q 512 + 2 tid and the next (q = 0..3) of a stage tile are in
    // block 8 q + tid / 32 = 8 ks + b (ks = q) at lanes l0 = 2 (tid % 32) and l0 + 1: the same k,
    // 4 ks + l0 / 16, and consecutive rows or cells.  A: row block b is slice b / 4, rows
    // 16 (b % 4) + l0 % 16 of the slice.  B: cells 16 b + l0 % 16.  A slice beyond the channel's
    // rows (odd k-tile count) reads slice 0 again, and the rows it feeds are never written.
    const int64_t slice = w.a_ktile_stride;   // doubles per packed 64-row slice
    const int kt0 = 2 * w.ktile;
    const bool has1 = (kt0 + 1) * TK < v.rows;
    const int b8 = tid >> 5, l0 = 2 * (tid & 31), kl = l0 >> 4;
    const GLB double* Ab = (const GLB double*)w.A +
                           (int64_t)(kt0 + ((b8 >> 2) && has1 ? 1 : 0)) * slice +
                           (int64_t)v.c_begin * KC * TK + 16 * (b8 & 3) + (l0 & 15);
    const int bcell = 16 * b8 + (l0 & 15);
    if (threadIdx.x < G) sm.arrived[threadIdx.x] = 0;
    load_ctab(v, ctab);   // ends with a barrier
    // this group's share of the range: local chunks [cb, ce)
    const int nch = v.c_end - v.c_begin, share = (nch + G - 1) / G;
    const int cb = grp * share, ce_ = min(nch, cb + share);
    const int nst = 2 * max(0, ce_ - cb), nst_all = 2 * share;   // stages: own, every group

    d2 rb[4], rf;
    // stage st = half (st & 1) of share chunk (st >> 1), into buffer buf: the operator straight
    // into LDS (global_load_lds_dwordx4: lane l of wave w lands at block q's 128 w + 2 l, the
    // pair the register path stored), the state into registers for the f_i fold
    auto load_stage = [&](int st, int buf) {
        const int ch = cb + (st >> 1), h = st & 1;
        const uint32_t ce = ctab[ch];
        const int64_t ao = (int64_t)ch * KC * TK + (int64_t)(h * HK + kl) * TK;
#pragma unroll
        for (int q = 0; q < 4; ++q)
            dma_lds<16>(Ab + ao + 4 * q * TK, &As[buf][q * 512 + 128 * wave]);
        const GLB double* bp = (const GLB double*)v.B +
                               (int64_t)((int)(ce >> 16) + h * HK + kl) * v.ldc + v.c0 + bcell;
#pragma unroll
        for (int q = 0; q < 4; ++q) rb[q] = *(const GLB d2*)(bp + (int64_t)(4 * q) * v.ldc);
        if (!loss) rf = *(const GLB d2*)(v.fT + (int64_t)(ce & 0xFFFFu) * v.ldc + v.c0 + bcell);
    };
    // the B' stores wait for the state loads, issued after the operator's, so the operator
    // block has landed too
    auto store_stage = [&](int buf) {
#pragma unroll
        for (int q = 0; q < 4; ++q) *(d2*)&Bs[buf][q * 512 + 2 * tid] = loss ? rb[q] : rf * rb[q];
    };

    d4 acc[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] = d4{0.0, 0.0, 0.0, 0.0};
    if (nst > 0) {
        load_stage(0, 0);
        store_stage(0);
    }
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    int buf = 0;
    // G = 1: the workgroup barrier after each stage; G > 1: each group its own stages and
    // barriers
    const int nloop = G == 1 ? nst_all : nst;
    for (int st = 0; st < nloop; ++st) {
        const bool mine = st < nst;
        const bool more = st + 1 < nst;
        if (more) load_stage(st + 1, buf ^ 1);
        if (mine) {
#pragma unroll
            for (int ks = 0; ks < HK / 4; ++ks) {
                double a[4], b[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) a[r] = As[buf][(8 * ks + 4 * wk + r) * 64 + lane];
#pragma unroll
                for (int c = 0; c < 4; ++c) b[c] = Bs[buf][(8 * ks + 4 * wc + c) * 64 + lane];
#pragma unroll
                for (int r = 0; r < 4; ++r)
#pragma unroll
                    for (int c = 0; c < 4; ++c)
                        acc[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[r], b[c], acc[r][c], 0, 0, 0);
            }
        }
        if (more) store_stage(buf ^ 1);
        if constexpr (G == 1) __syncthreads();
        else group_barrier(&sm.arrived[grp], 4u * (uint32_t)(st + 1));
        buf ^= 1;
    }
    if constexpr (G > 1) __syncthreads();   // both groups' sums are complete
    // the cell blocks c this group writes: all (G = 1), or 2 grp and 2 grp + 1 (G = 2); the
    // block indices are compile-time constants (a run-time index would put acc in scratch)
    constexpr int NC = 4 / G;
    const int krow = kt0 * TK + 64 * wk + lane;
    double* tw = sm.T[threadIdx.x >> 6];
    auto finish = [&](auto first) {
        constexpr int c_first = decltype(first)::value;
        if constexpr (G == 2) {
            // hand over the other group's two cell blocks (lane-minor, 64 doubles a store), then
            // add the two received: share 0 + share 1 for every tile (addition is commutative,
            // so both groups form the same sums)
            constexpr int give = 2 - c_first;
            double* mine_x = sm.X[threadIdx.x >> 6];
            const double* other_x = sm.X[(threadIdx.x >> 6) ^ 4];
#pragma unroll
            for (int cc = 0; cc < 2; ++cc)
#pragma unroll
                for (int r = 0; r < 4; ++r)
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        mine_x[((cc * 4 + r) * 4 + e) * 64 + lane] = acc[r][give + cc][e];
            __syncthreads();
#pragma unroll
            for (int cc = 0; cc < 2; ++cc)
#pragma unroll
                for (int r = 0; r < 4; ++r)
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        acc[r][c_first + cc][e] =
                            acc[r][c_first + cc][e] + other_x[((cc * 4 + r) * 4 + e) * 64 + lane];
            __syncthreads();   // T aliases X
        }
        // write the item's partial sums: per group of 16 cells each wave puts its result tiles
        // (FP64 layout D[(l / 16) + 4 e][l % 16]) into LDS cell-major, then lane l writes row l
        // of one cell per store, 512 contiguous bytes instead of 16 pieces of 32 (the epilogue
        // of all items runs at about the same time: measured -3% contraction time at
        // transport_1d); LOSS rows are multiplied by their own f
#pragma unroll
        for (int cc = 0; cc < NC; ++cc) {
            const int c = c_first + cc;
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    tw[(lane & 15) * EPI_PITCH + 16 * r + (lane >> 4) + 4 * e] = acc[r][c][e];
            __syncthreads();
            for (int q = 0; q < 16; ++q) {
                const int cell = v.c0 + 64 * wc + 16 * c + q;
                if (cell < v.cells && krow < v.rows) {
                    double x = tw[q * EPI_PITCH + lane];
                    if (loss) x = x * v.fT[(int64_t)(v.row_pad_base + krow) * v.ldc + cell];
                    v.part[(int64_t)cell * v.N + v.row_base + krow] = x;
                }
            }
            __syncthreads();
        }
    };
    if (grp == 0) finish(std::integral_constant<int, 0>{});
    else if constexpr (G == 2) finish(std::integral_constant<int, 2>{});
}

// G = 1: LOSS items take the 64-cell FP64 body.  G = 2: every item is wide.
__global__ __launch_bounds__(256 * F64_WIDE_GROUPS) void bin_f64_wide_kernel(
    const BinWork* __restrict__ items) {
    __shared__ union {
        SmemF64<64> d;
        SmemWide64<F64_WIDE_GROUPS> w;
    } sm;
    const BinWork& w = items[blockIdx.x];
    if (F64_WIDE_GROUPS == 1 && w.kind >= CH_LOSS) f64_item<64>(w, sm.d);
    else wide64_item<F64_WIDE_GROUPS>(w, sm.w);
}

hipError_t launch_binary_wide(int precision, int ns, const BinWork* items, int nitems,
                              hipStream_t stream) {
    if (nitems <= 0) return hipSuccess;
    if (precision == DTX_PREC_FP64 && ns == 2)
        bin_f64_wide_kernel<<<nitems, 256 * F64_WIDE_GROUPS, 0, stream>>>(items);
    else if (precision != DTX_PREC_BF16X3) return hipErrorInvalidValue;
    else if (ns == 4) bin_bf16x3_wide_kernel<4><<<nitems, 512, 0, stream>>>(items);
    else if (ns == 2) bin_bf16x3_wide_kernel<2><<<nitems, 256, 0, stream>>>(items);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

// Pack rows [row0, row0+rows) of the FP64 loss matrix D (N x N, column-major) as a LOSS
// channel: A[kt][chunk][t][r] = D[row0 + kt*TK + r, state at padded column chunk*KC + t].
__global__ void pack_loss_kernel(const double* D, int N, const int32_t* unpad, int Np, int row0,
                                 int rows, int nktiles, double* A) {
    const int64_t nch = Np / KC;
    const int64_t total = (int64_t)nktiles * nch * KC * TK;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(e % TK);
        const int t = (int)((e / TK) % KC);
        const int64_t ch = (e / ((int64_t)TK * KC)) % nch;
        const int kt = (int)(e / ((int64_t)TK * KC * nch));
        const int row = kt * TK + r;
        const int jp = (int)(ch * KC + t);
        const int j = unpad[jp];
        A[e] = (row < rows && j >= 0) ? D[(int64_t)(row0 + row) + (int64_t)N * j] : 0.0;
    }
}

hipError_t launch_pack_loss(const double* D, int N, const int32_t* unpad, int Np, int row0,
                            int rows, int nktiles, double* A, hipStream_t stream) {
    pack_loss_kernel<<<4096, 256, 0, stream>>>(D, N, unpad, Np, row0, rows, nktiles, A);
    return hipGetLastError();
}

// ------------------------------------------------------------------- single-cell GEMV -----
// With one cell the contraction is a GEMV that streams the operator from HBM once per step:
// MFMA tiles would compute on 15 padding cells out of 16 and pay LDS staging and barriers per
// chunk for nothing.  A GEMV item (channel, 64-row tile, chunk range) has its four waves take
// every fourth chunk of the range; lane r of a wave owns row r, loads the chunk's 32 operator
// values of its row straight from HBM (each load instruction of a wave reads consecutive rows)
// and accumulates them against f_j (read by lanes 0..31 and broadcast with readlane) in FP64,
// whatever the stored width (FP64, FP32, or BF16 hi + lo, which is exact in FP32).  The
// operator is read once per step, so its loads are nontemporal (measured 7-13% faster than
// plain loads on the single-cell configurations, above a plain streaming read).  Gain rows
// fold f_i per chunk; LOSS rows take their own f at the end; MOMENT items sum f_row (W f)_row.
// The four wave partials of a row are added in wave order.
__device__ inline double lane_value(double x, int l) {
    const uint64_t u = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), l);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

struct SmemGemv {
    double red[4][TK];
};

template <int P>
__device__ __forceinline__ void gemv_item(const BinWork& w, SmemGemv& sm) {
    const bool fp64 = P == DTX_PREC_FP64 || w.kind >= CH_LOSS;
    const int eb = fp64 ? 8 : P == DTX_PREC_FP32 ? 4 : 2;
    const ItemView v = load_item(w, 16, eb);   // one cell: c0 = 0
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const bool gain = v.kind < CH_LOSS;
    const int nch = v.c_end - v.c_begin;
    double tot = 0.0;
    for (int ch = wave; ch < nch; ch += 4) {
        const uint32_t ce = v.chunks[v.c_begin + ch];
        const int j0 = (int)(ce >> 16);
        const double fv = v.fT[(int64_t)(j0 + (lane & 31)) * v.ldc];
        const int64_t co = (int64_t)(v.c_begin + ch) * KC * TK;   // chunk offset (BF16: x 2)
        double acc = 0.0;
        if (fp64) {
            const GLB double* a = (const GLB double*)v.A + co + lane;
            double x[KC];
#pragma unroll
            for (int j = 0; j < KC; ++j) x[j] = __builtin_nontemporal_load(a + j * TK);
#pragma unroll
            for (int j = 0; j < KC; ++j) acc = __builtin_fma(x[j], lane_value(fv, j), acc);
        } else if (P == DTX_PREC_FP32) {
            const GLB float* a = (const GLB float*)v.A + co + lane;
            float x[KC];
#pragma unroll
            for (int j = 0; j < KC; ++j) x[j] = __builtin_nontemporal_load(a + j * TK);
#pragma unroll
            for (int j = 0; j < KC; ++j) acc = __builtin_fma((double)x[j], lane_value(fv, j), acc);
        } else {
            // split-BF16 chunk: hi then lo, each in MFMA fragment order (bf16_frag_index): the
            // 32 j of a row are four 16-byte groups
            const GLB uint16_t* a = (const GLB uint16_t*)v.A + 2 * co;
            u32x4 hq[4], lq[4];
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const int o = bf16_frag_index(lane, 8 * g);
                hq[g] = __builtin_nontemporal_load((const GLB u32x4*)(a + o));
                lq[g] = __builtin_nontemporal_load((const GLB u32x4*)(a + TK * KC + o));
            }
#pragma unroll
            for (int g = 0; g < 4; ++g)
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const uint32_t hw = hq[g][e >> 1], lw = lq[g][e >> 1];
                    const uint16_t hb = (uint16_t)((e & 1) ? hw >> 16 : hw);
                    const uint16_t lb = (uint16_t)((e & 1) ? lw >> 16 : lw);
                    const float x = bf16_to_f32(hb) + bf16_to_f32(lb);   // exact
                    acc = __builtin_fma((double)x, lane_value(fv, 8 * g + e), acc);
                }
        }
        if (gain) tot = __builtin_fma(v.fT[(int64_t)(ce & 0xFFFFu) * v.ldc], acc, tot);
        else tot = tot + acc;
    }
    sm.red[wave][lane] = tot;
    __syncthreads();
    if (wave != 0) return;
    double t = sm.red[0][lane];
#pragma unroll
    for (int q = 1; q < 4; ++q) t = t + sm.red[q][lane];
    const int krow = v.krow0 + lane;
    const bool valid = krow < v.rows;
    if (v.kind == CH_MOMENT) {
        double m = valid ? t * v.fT[(int64_t)(v.row_pad_base + krow) * v.ldc] : 0.0;
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) m = m + xor_lane_rt(m, d);
        if (lane == 0) v.part[0] = m;
        return;
    }
    if (!valid) return;
    if (v.kind == CH_LOSS) t = t * v.fT[(int64_t)(v.row_pad_base + krow) * v.ldc];
    v.part[v.row_base + krow] = t;
}

template <int P>
__global__ __launch_bounds__(256) void bin_gemv_kernel(const BinWork* __restrict__ items) {
    __shared__ SmemGemv sm;
    gemv_item<P>(items[blockIdx.x], sm);
}


// One launch over items of the narrow shapes (single-cell GEMV, 16, 32 and 64 cells): each
// workgroup takes the body of its item's shape (the same device code as the single-shape kernels, so a result does not depend
// on the launch it ran in).
template <int P>
__global__ __launch_bounds__(256) void bin_mixed_kernel(const BinWork* __restrict__ items) {
    __shared__ union {
        SmemF64<16> d16;
        SmemF64<32> d32;
        SmemF64<64> d64;
        SmemF32<16> s16;
        SmemF32<32> s32;
        SmemF32<64> s64;
        SmemBF16<16> b16;
        SmemBF16<32> b32;
        SmemBF16<64> b64;
        SmemGemv g;
    } sm;
    const BinWork& w = items[blockIdx.x];
    const bool fp64_body = P == DTX_PREC_FP64 || w.kind >= CH_LOSS;
    if (w.tc == 1) {
        gemv_item<P>(w, sm.g);
    } else if (w.tc == 16) {
        if (fp64_body) f64_item<16>(w, sm.d16);
        else if (P == DTX_PREC_FP32) f32_item<16>(w, sm.s16);
        else bf16x3_item<16>(w, sm.b16);
    } else if (w.tc == 32) {
        if (fp64_body) f64_item<32>(w, sm.d32);
        else if (P == DTX_PREC_FP32) f32_item<32>(w, sm.s32);
        else bf16x3_item<32>(w, sm.b32);
    } else {
        if (fp64_body) f64_item<64>(w, sm.d64);
        else if (P == DTX_PREC_FP32) f32_item<64>(w, sm.s64);
        else bf16x3_item<64>(w, sm.b64);
    }
}

hipError_t launch_binary_mixed(int precision, const BinWork* items, int nitems,
                               hipStream_t stream) {
    if (nitems <= 0) return hipSuccess;
    if (precision == DTX_PREC_FP64) bin_mixed_kernel<DTX_PREC_FP64><<<nitems, 256, 0, stream>>>(items);
    else if (precision == DTX_PREC_FP32) bin_mixed_kernel<DTX_PREC_FP32><<<nitems, 256, 0, stream>>>(items);
    else bin_mixed_kernel<DTX_PREC_BF16X3><<<nitems, 256, 0, stream>>>(items);
    return hipGetLastError();
}

hipError_t launch_binary(int precision, int tc, const BinWork* items, int nitems,
                         hipStream_t stream) {
    if (nitems <= 0) return hipSuccess;
    dim3 grid(nitems), block(256);
    if (tc == 1) {
        if (precision == DTX_PREC_FP64) bin_gemv_kernel<DTX_PREC_FP64><<<grid, block, 0, stream>>>(items);
        else if (precision == DTX_PREC_FP32) bin_gemv_kernel<DTX_PREC_FP32><<<grid, block, 0, stream>>>(items);
        else bin_gemv_kernel<DTX_PREC_BF16X3><<<grid, block, 0, stream>>>(items);
    } else if (precision == DTX_PREC_FP64) {
        if (tc == 16) bin_f64_kernel<16><<<grid, block, 0, stream>>>(items);
        else if (tc == 32) bin_f64_kernel<32><<<grid, block, 0, stream>>>(items);
        else bin_f64_kernel<64><<<grid, block, 0, stream>>>(items);
    } else if (precision == DTX_PREC_FP32) {
        if (tc == 16) bin_f32_kernel<16><<<grid, block, 0, stream>>>(items);
        else if (tc == 32) bin_f32_kernel<32><<<grid, block, 0, stream>>>(items);
        else bin_f32_kernel<64><<<grid, block, 0, stream>>>(items);
    } else {
        if (tc == 16) bin_bf16x3_kernel<16><<<grid, block, 0, stream>>>(items);
        else if (tc == 32) bin_bf16x3_kernel<32><<<grid, block, 0, stream>>>(items);
        else bin_bf16x3_kernel<64><<<grid, block, 0, stream>>>(items);
    }
    return hipGetLastError();
}

}  // namespace dtx
