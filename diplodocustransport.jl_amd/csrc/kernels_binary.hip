// kernels_binary.hip -- grouped binary-collision contraction for gfx950.
//
// Replaces the per-cell pair of GEMVs of update_Big_Bin! (src/SteppingMethods.jl:360-386):
//     dB[k, c] = sum_{(i,j)} S[k, (i,j)] * f_i^c * f_j^c
// as one GEMM per channel with K = the channel's (i,j) pairs.  The A operand is the packed gain
// S (streamed from HBM once per step); the B operand, the per-cell pair products f_i f_j, is
// generated on the fly into LDS from the state-major copy fT and never materialised in HBM.
// Every workgroup takes one work item (channel, 64-row tile, cell tile, pair range) from a work
// table and writes its partial sum to its own slot, so the reduction order is fixed by the
// item list alone: a simulation gives bit-identical results alone or grouped with others.
//
// Wave tiling: 4 waves = WK (rows) x WC (cells); each wave owns MR x MC 16x16 MFMA tiles.
// MFMA operand layout (16x16xK): lane l supplies A[l%16][kgroup(l/16)], B[kgroup(l/16)][l%16].
// The 16x16 result tile is held as D[4*(l/16)+e][l%16] (e = 0..3) by the FP32 and BF16
// instructions, and as D[(l/16)+4*e][l%16] by v_mfma_f64_16x16x4_f64 (measured on gfx950).
#include "dtx_internal.h"
#include "bf16_util.h"

namespace dtx {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

template <int TC>
struct Tiling {
    static constexpr int WC = TC >= 32 ? 2 : 1;   // waves along cells
    static constexpr int WK = 4 / WC;             // waves along rows
    static constexpr int WROWS = TK / WK;
    static constexpr int WCOLS = TC / WC;
    static constexpr int MR = WROWS / 16;
    static constexpr int MC = WCOLS / 16;
};

struct ItemView {
    const void* A;
    const uint32_t* pairs;
    const double* fT;
    double* part;
    int p_begin, p_end, krow0, row_base, rows, cells, ldc, N, c0;
};

__device__ inline ItemView load_item(const BinWork* items, int tc) {
    const BinWork& w = items[blockIdx.x];
    ItemView v;
    v.A = w.A;
    v.pairs = w.pairs;
    v.fT = w.fT;
    v.part = w.part;
    v.p_begin = w.p_begin;
    v.p_end = w.p_end;
    v.krow0 = w.ktile * TK;
    v.row_base = w.row_base;
    v.rows = w.rows;
    v.cells = w.cells;
    v.ldc = w.ldc;
    v.N = w.N;
    v.c0 = w.ctile * tc;
    return v;
}

// Store one 16x16 tile of partial sums: rows krow0 + row, cells c0 + col.  ROWSTEP is the row
// distance between a lane's consecutive accumulator elements (1 for FP32/BF16, 4 for FP64).
template <int ROWSTEP>
__device__ inline void store_tile(const ItemView& v, int row16, int col16, int lane,
                                  const double* acc4) {
    const int col = col16 + (lane & 15);
    const int cell = v.c0 + col;
    if (cell >= v.cells) return;
    double* dst = v.part + (int64_t)cell * v.N + v.row_base;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int krow = v.krow0 + row16 + (ROWSTEP == 1 ? 4 * (lane >> 4) + e
                                                         : (lane >> 4) + 4 * e);
        if (krow < v.rows) dst[krow] = acc4[e];
    }
}

// ------------------------------------------------------------------------------- FP64 -----
template <int TC>
__global__ __launch_bounds__(256) void bin_f64_kernel(const BinWork* __restrict__ items) {
    using T = Tiling<TC>;
    constexpr int KC = KC64;
    constexpr int AP = KC * TK / 256;     // A elements per thread
    constexpr int BP = KC * TC / 256;     // pair products per thread (consecutive cells)
    constexpr int BT = TC / BP;           // threads per pair row
    __shared__ double As[2][KC * TK];
    __shared__ double Bs[2][KC * TC];

    const ItemView v = load_item(items, TC);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wk = wave % T::WK, wc = wave / T::WK;
    const double* A = (const double*)v.A + (int64_t)items[blockIdx.x].a_ktile_stride *
                                               items[blockIdx.x].ktile;
    const int bpp = tid / BT, bcc = (tid % BT) * BP;

    double ra[AP], rbi[BP], rbj[BP];
    auto load_chunk = [&](int p) {
        const double* src = A + (int64_t)p * TK + tid * AP;
#pragma unroll
        for (int q = 0; q < AP; ++q) ra[q] = src[q];
        const uint32_t pr = v.pairs[p + bpp];
        if (pr != PAIR_PAD) {
            const double* fi = v.fT + (int64_t)(pr & 0xFFFFu) * v.ldc + v.c0 + bcc;
            const double* fj = v.fT + (int64_t)(pr >> 16) * v.ldc + v.c0 + bcc;
#pragma unroll
            for (int q = 0; q < BP; ++q) { rbi[q] = fi[q]; rbj[q] = fj[q]; }
        } else {
#pragma unroll
            for (int q = 0; q < BP; ++q) { rbi[q] = 0.0; rbj[q] = 0.0; }
        }
    };
    auto store_chunk = [&](int buf) {
#pragma unroll
        for (int q = 0; q < AP; ++q) As[buf][tid * AP + q] = ra[q];
#pragma unroll
        for (int q = 0; q < BP; ++q) Bs[buf][bpp * TC + bcc + q] = rbi[q] * rbj[q];
    };

    d4 acc[T::MR][T::MC];
#pragma unroll
    for (int r = 0; r < T::MR; ++r)
#pragma unroll
        for (int c = 0; c < T::MC; ++c) acc[r][c] = d4{0.0, 0.0, 0.0, 0.0};

    if (v.p_begin < v.p_end) {
        load_chunk(v.p_begin);
        store_chunk(0);
    }
    __syncthreads();
    int buf = 0;
    for (int p = v.p_begin; p < v.p_end; p += KC) {
        const bool more = p + KC < v.p_end;
        if (more) load_chunk(p + KC);
#pragma unroll
        for (int kk = 0; kk < KC; kk += 4) {
            const int kr = kk + (lane >> 4);
            double a[T::MR], b[T::MC];
#pragma unroll
            for (int r = 0; r < T::MR; ++r)
                a[r] = As[buf][kr * TK + wk * T::WROWS + r * 16 + (lane & 15)];
#pragma unroll
            for (int c = 0; c < T::MC; ++c)
                b[c] = Bs[buf][kr * TC + wc * T::WCOLS + c * 16 + (lane & 15)];
#pragma unroll
            for (int r = 0; r < T::MR; ++r)
#pragma unroll
                for (int c = 0; c < T::MC; ++c)
                    acc[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[r], b[c], acc[r][c], 0, 0, 0);
        }
        if (more) store_chunk(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
#pragma unroll
    for (int r = 0; r < T::MR; ++r)
#pragma unroll
        for (int c = 0; c < T::MC; ++c) {
            double t[4] = {acc[r][c][0], acc[r][c][1], acc[r][c][2], acc[r][c][3]};
            store_tile<4>(v, wk * T::WROWS + r * 16, wc * T::WCOLS + c * 16, lane, t);
        }
}

// ------------------------------------------------------------------------------- FP32 -----
template <int TC>
__global__ __launch_bounds__(256) void bin_f32_kernel(const BinWork* __restrict__ items) {
    using T = Tiling<TC>;
    constexpr int KC = KC64;
    constexpr int AP = KC * TK / 256;
    constexpr int BP = KC * TC / 256;
    constexpr int BT = TC / BP;
    constexpr int FLUSH = FLUSH_PAIRS / KC;
    __shared__ float As[2][KC * TK];
    __shared__ float Bs[2][KC * TC];

    const ItemView v = load_item(items, TC);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wk = wave % T::WK, wc = wave / T::WK;
    const float* A = (const float*)v.A + (int64_t)items[blockIdx.x].a_ktile_stride *
                                             items[blockIdx.x].ktile;
    const int bpp = tid / BT, bcc = (tid % BT) * BP;

    float ra[AP];
    double rbi[BP], rbj[BP];
    auto load_chunk = [&](int p) {
        const float* src = A + (int64_t)p * TK + tid * AP;
#pragma unroll
        for (int q = 0; q < AP; ++q) ra[q] = src[q];
        const uint32_t pr = v.pairs[p + bpp];
        if (pr != PAIR_PAD) {
            const double* fi = v.fT + (int64_t)(pr & 0xFFFFu) * v.ldc + v.c0 + bcc;
            const double* fj = v.fT + (int64_t)(pr >> 16) * v.ldc + v.c0 + bcc;
#pragma unroll
            for (int q = 0; q < BP; ++q) { rbi[q] = fi[q]; rbj[q] = fj[q]; }
        } else {
#pragma unroll
            for (int q = 0; q < BP; ++q) { rbi[q] = 0.0; rbj[q] = 0.0; }
        }
    };
    auto store_chunk = [&](int buf) {
#pragma unroll
        for (int q = 0; q < AP; ++q) As[buf][tid * AP + q] = ra[q];
#pragma unroll
        for (int q = 0; q < BP; ++q) Bs[buf][bpp * TC + bcc + q] = (float)(rbi[q] * rbj[q]);
    };

    f4 acc[T::MR][T::MC];
    double acc64[T::MR][T::MC][4];
#pragma unroll
    for (int r = 0; r < T::MR; ++r)
#pragma unroll
        for (int c = 0; c < T::MC; ++c) {
            acc[r][c] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int e = 0; e < 4; ++e) acc64[r][c][e] = 0.0;
        }
    if (v.p_begin < v.p_end) {
        load_chunk(v.p_begin);
        store_chunk(0);
    }
    __syncthreads();
    int buf = 0, it = 0;
    for (int p = v.p_begin; p < v.p_end; p += KC) {
        const bool more = p + KC < v.p_end;
        if (more) load_chunk(p + KC);
#pragma unroll
        for (int kk = 0; kk < KC; kk += 4) {
            const int kr = kk + (lane >> 4);
            float a[T::MR], b[T::MC];
#pragma unroll
            for (int r = 0; r < T::MR; ++r)
                a[r] = As[buf][kr * TK + wk * T::WROWS + r * 16 + (lane & 15)];
#pragma unroll
            for (int c = 0; c < T::MC; ++c)
                b[c] = Bs[buf][kr * TC + wc * T::WCOLS + c * 16 + (lane & 15)];
#pragma unroll
            for (int r = 0; r < T::MR; ++r)
#pragma unroll
                for (int c = 0; c < T::MC; ++c)
                    acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[r], b[c], acc[r][c], 0, 0, 0);
        }
        if (++it == FLUSH || !more) {
            it = 0;
#pragma unroll
            for (int r = 0; r < T::MR; ++r)
#pragma unroll
                for (int c = 0; c < T::MC; ++c) {
#pragma unroll
                    for (int e = 0; e < 4; ++e) acc64[r][c][e] += (double)acc[r][c][e];
                    acc[r][c] = f4{0.f, 0.f, 0.f, 0.f};
                }
        }
        if (more) store_chunk(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
#pragma unroll
    for (int r = 0; r < T::MR; ++r)
#pragma unroll
        for (int c = 0; c < T::MC; ++c)
            store_tile<1>(v, wk * T::WROWS + r * 16, wc * T::WCOLS + c * 16, lane, acc64[r][c]);
}

// ------------------------------------------------------------------------ split BF16 x3 -----
// Packed A per k-tile and 32-pair chunk: hi[TK][32] then lo[TK][32] (bf16 bits).
// LDS rows are padded to 40 elements so the 16-byte fragment reads of 16 rows hit distinct
// banks.
constexpr int BPITCH = 40;

template <int TC>
__global__ __launch_bounds__(256) void bin_bf16x3_kernel(const BinWork* __restrict__ items) {
    using T = Tiling<TC>;
    constexpr int KC = KC16;                 // 32 pairs = one MFMA K step
    constexpr int FLUSH = FLUSH_PAIRS / KC;
    constexpr int BG = KC * TC / 8;          // groups of 8 products
    __shared__ __attribute__((aligned(16))) uint16_t Ah[2][TK * BPITCH];
    __shared__ __attribute__((aligned(16))) uint16_t Al[2][TK * BPITCH];
    __shared__ __attribute__((aligned(16))) uint16_t Bh[2][TC * BPITCH];
    __shared__ __attribute__((aligned(16))) uint16_t Bl[2][TC * BPITCH];

    const ItemView v = load_item(items, TC);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wk = wave % T::WK, wc = wave / T::WK;
    const uint16_t* A = (const uint16_t*)v.A + (int64_t)items[blockIdx.x].a_ktile_stride *
                                                   items[blockIdx.x].ktile;
    const bool bthread = tid < BG;
    const int bcc = tid % TC, bk8 = (tid / TC) * 8;
    const int arow = tid >> 2, ak8 = (tid & 3) * 8;

    uint4 rah, ral;
    double rbi[8], rbj[8];
    auto load_chunk = [&](int p) {
        const uint4* src = (const uint4*)(A + (int64_t)p * TK * 2);  // chunk = 2*TK*32 bf16
        rah = src[tid];
        ral = src[256 + tid];
        if (bthread) {
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const uint32_t pr = v.pairs[p + bk8 + q];
                if (pr != PAIR_PAD) {
                    rbi[q] = v.fT[(int64_t)(pr & 0xFFFFu) * v.ldc + v.c0 + bcc];
                    rbj[q] = v.fT[(int64_t)(pr >> 16) * v.ldc + v.c0 + bcc];
                } else {
                    rbi[q] = 0.0;
                    rbj[q] = 0.0;
                }
            }
        }
    };
    auto store_chunk = [&](int buf) {
        *(uint4*)&Ah[buf][arow * BPITCH + ak8] = rah;
        *(uint4*)&Al[buf][arow * BPITCH + ak8] = ral;
        if (bthread) {
            uint16_t h[8], l[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) split_bf16(rbi[q] * rbj[q], &h[q], &l[q]);
            uint4 vh, vl;
            vh.x = h[0] | ((uint32_t)h[1] << 16); vh.y = h[2] | ((uint32_t)h[3] << 16);
            vh.z = h[4] | ((uint32_t)h[5] << 16); vh.w = h[6] | ((uint32_t)h[7] << 16);
            vl.x = l[0] | ((uint32_t)l[1] << 16); vl.y = l[2] | ((uint32_t)l[3] << 16);
            vl.z = l[4] | ((uint32_t)l[5] << 16); vl.w = l[6] | ((uint32_t)l[7] << 16);
            *(uint4*)&Bh[buf][bcc * BPITCH + bk8] = vh;
            *(uint4*)&Bl[buf][bcc * BPITCH + bk8] = vl;
        }
    };

    f4 acc[T::MR][T::MC];
    double acc64[T::MR][T::MC][4];
#pragma unroll
    for (int r = 0; r < T::MR; ++r)
#pragma unroll
        for (int c = 0; c < T::MC; ++c) {
            acc[r][c] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int e = 0; e < 4; ++e) acc64[r][c][e] = 0.0;
        }
    if (v.p_begin < v.p_end) {
        load_chunk(v.p_begin);
        store_chunk(0);
    }
    __syncthreads();
    int buf = 0, it = 0;
    const int kq = 8 * (lane >> 4);
    for (int p = v.p_begin; p < v.p_end; p += KC) {
        const bool more = p + KC < v.p_end;
        if (more) load_chunk(p + KC);
        bf16x8 ah[T::MR], al[T::MR], bh[T::MC], bl[T::MC];
#pragma unroll
        for (int r = 0; r < T::MR; ++r) {
            const int row = wk * T::WROWS + r * 16 + (lane & 15);
            ah[r] = *(const bf16x8*)&Ah[buf][row * BPITCH + kq];
            al[r] = *(const bf16x8*)&Al[buf][row * BPITCH + kq];
        }
#pragma unroll
        for (int c = 0; c < T::MC; ++c) {
            const int col = wc * T::WCOLS + c * 16 + (lane & 15);
            bh[c] = *(const bf16x8*)&Bh[buf][col * BPITCH + kq];
            bl[c] = *(const bf16x8*)&Bl[buf][col * BPITCH + kq];
        }
#pragma unroll
        for (int r = 0; r < T::MR; ++r)
#pragma unroll
            for (int c = 0; c < T::MC; ++c) {
                acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[r], bh[c], acc[r][c], 0, 0, 0);
                acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[r], bl[c], acc[r][c], 0, 0, 0);
                acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[r], bh[c], acc[r][c], 0, 0, 0);
            }
        if (++it == FLUSH || !more) {
            it = 0;
#pragma unroll
            for (int r = 0; r < T::MR; ++r)
#pragma unroll
                for (int c = 0; c < T::MC; ++c) {
#pragma unroll
                    for (int e = 0; e < 4; ++e) acc64[r][c][e] += (double)acc[r][c][e];
                    acc[r][c] = f4{0.f, 0.f, 0.f, 0.f};
                }
        }
        if (more) store_chunk(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
#pragma unroll
    for (int r = 0; r < T::MR; ++r)
#pragma unroll
        for (int c = 0; c < T::MC; ++c)
            store_tile<1>(v, wk * T::WROWS + r * 16, wc * T::WCOLS + c * 16, lane, acc64[r][c]);
}

hipError_t launch_binary(int precision, int tc, const BinWork* items, int nitems,
                         hipStream_t stream) {
    if (nitems <= 0) return hipSuccess;
    dim3 grid(nitems), block(256);
    if (precision == DTX_PREC_FP64) {
        if (tc == 16) bin_f64_kernel<16><<<grid, block, 0, stream>>>(items);
        else bin_f64_kernel<64><<<grid, block, 0, stream>>>(items);
    } else if (precision == DTX_PREC_FP32) {
        if (tc == 16) bin_f32_kernel<16><<<grid, block, 0, stream>>>(items);
        else bin_f32_kernel<64><<<grid, block, 0, stream>>>(items);
    } else {
        if (tc == 16) bin_bf16x3_kernel<16><<<grid, block, 0, stream>>>(items);
        else bin_bf16x3_kernel<64><<<grid, block, 0, stream>>>(items);
    }
    return hipGetLastError();
}

}  // namespace dtx
