// kernels_update.hip -- fused per-cell update kernels for gfx950 (HBM-bound).
//
// One workgroup per (cell, segment of UPDATE_STATES = 64 states), four threads per state: the
// four split the state's partial-slot loads, and the first of them evaluates, in FP64 and with
// the reference's operation order (this file is built with -ffp-contract=off, so no FMA):
//   ForwardEuler, src/SteppingMethods.jl:31-101
//     dF  = F f                           (ELL rows, columns ascending = CSC column order)
//     df  = -dF ; df = df + dE ; df = df + dB
//     df  = df * (invAp * s)
//     f   = f + (df + df_Inj * s) ; f = f*(f >= 1f-10) ; f = f*(f != Inf)   (Bool strong zero)
//   plus Cr / Cr_Bin / Cr_Emi / Cr_Flux, isfinite(sum(df)), sum(f) and the per-(cell, species)
//   number and energy moments of the new f.
// dE is the cell's emission block times f (ascending column order) and dB the binary term:
// the sum of the contraction's partial slots (gain and FP64 loss channels), times Vol.
// Cross-cell reductions go to per-block partials that a one-workgroup kernel folds in block
// order, so every reduction is deterministic.
#include "bf16_util.h"
#include "dtx_internal.h"
#include "wave_ops.h"

namespace dtx {

__device__ constexpr double THR_1F_10 = (double)1e-10f;  // Julia Float32 literal 1f-10
__device__ constexpr double THR_1F_20 = (double)1e-20f;  // Julia Float32 literal 1f-20

constexpr int UB = UPDATE_THREADS;

__device__ inline double strong_zero(double x, bool cond) { return cond ? x : copysign(0.0, x); }

__device__ inline double jl_sign(double x) {
    if (x > 0.0) return 1.0;
    if (x < 0.0) return -1.0;
    return x;
}

// isnan(r) ? Inf : r, folded into a running minimum with Julia's min (-0.0 < 0.0).
__device__ inline double cr_min(double m, double r) {
    if (isnan(r)) r = INFINITY;
    if (r < m || (r == m && signbit(r))) return r;
    return m;
}

// Wave reductions: a butterfly over the 64 lanes of each wave (xor_lane: DPP and permlane
// exchanges, no LDS).
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) v = v + xor_lane_rt(v, m);
    return v;
}

__device__ inline double wave_min(double v) {
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) v = cr_min(v, xor_lane_rt(v, m));
    return v;
}

template <int NT = UB>
__device__ int block_or(int v, int* red) {
    const int tid = threadIdx.x;
    v = __any(v) ? 1 : 0;
    if ((tid & 63) == 0) red[tid >> 6] = v;
    __syncthreads();
    int r = 0;
    for (int w = 0; w < NT / 64; ++w) r |= red[w];
    __syncthreads();
    return r;
}

// SpMV row: sum over ELL slots in ascending column order, separate multiply and add.  A row's
// column and value loads are all issued before the first gather, and the gathers before the
// first add, so a row costs two memory round trips whatever its width.  Padding slots (column
// -1) sit after a row's entries and are skipped.
__device__ inline double ell_row(const int32_t* col, const double* val, int width, int64_t n,
                                 int64_t r, const double* x) {
    constexpr int EB = 8;
    double acc = 0.0;
    for (int s0 = 0; s0 < width; s0 += EB) {
        int32_t c[EB];
        double v[EB], xv[EB];
#pragma unroll
        for (int u = 0; u < EB; ++u) {
            const bool in = s0 + u < width;
            c[u] = in ? col[(int64_t)(s0 + u) * n + r] : -1;
            v[u] = in ? val[(int64_t)(s0 + u) * n + r] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < EB; ++u) xv[u] = c[u] >= 0 ? x[c[u]] : 0.0;
#pragma unroll
        for (int u = 0; u < EB; ++u)
            if (c[u] >= 0) acc = acc + v[u] * xv[u];
    }
    return acc;
}

// Binary term of state k of cell c (unscaled by invAp): the contraction's partial slots (gain
// channels and the FP64 loss channels) times Vol.  The four threads q = 0..3 of the state sum
// slots q, q+4, ...; the four sums are combined as (s0 + s1) + (s2 + s3) by a butterfly, so every
// lane holds the same, order-fixed result.
__device__ inline double binary_term(const UpdateArgs& a, int c, int k, int q, bool valid,
                                     const double* coef) {
#ifndef DTX_UPDATE_SB
#define DTX_UPDATE_SB 8
#endif
    constexpr int SB = DTX_UPDATE_SB;  // slot loads in flight per thread
    const int64_t idx = (int64_t)c * a.N + k;
    double g = 0.0;
    if (a.gsum) {
        // moment-rate corrected: the state's gain sum rescaled by (1 + alpha + beta E_k) of its
        // species (coef: the workgroup's (alpha, beta) per species), plus its loss sum (thread
        // q = 0; the others add zeros)
        if (valid && q == 0) {
            const double* cf = coef + 2 * a.state_species[k];
            const double e = a.eweight ? a.eweight[k] : 0.0;
            g = a.gsum[idx] * (1.0 + (cf[0] + cf[1] * e)) + a.part[idx];
        }
    } else if (valid) {
        // slots q, q + 4, ... in order: the leading FP32 ones, then the FP64 ones
        const int ns = a.nslot[k], n32 = a.nslot32 ? a.nslot32[k] : 0;
        int s0 = q;
        for (; s0 < n32; s0 += 4 * SB) {
            // past the last FP32 slot the load repeats slot s0 (valid, unused), so the loads
            // carry no branch and stay in flight together
            float v[SB];
#pragma unroll
            for (int u = 0; u < SB; ++u) {
                const int s = s0 + 4 * u < n32 ? s0 + 4 * u : s0;
                v[u] = ((const float*)(a.part + (int64_t)s * a.part_stride))[idx];
            }
#pragma unroll
            for (int u = 0; u < SB; ++u)
                if (s0 + 4 * u < n32) g = g + (double)v[u];
        }
        // first FP64 slot of this thread: the smallest q + 4 i >= n32
        s0 = n32 + ((q - n32) % 4 + 4) % 4;
        for (; s0 < ns; s0 += 4 * SB) {
            double v[SB];
#pragma unroll
            for (int u = 0; u < SB; ++u) {
                const int s = s0 + 4 * u;
                v[u] = s < ns ? a.part[(int64_t)s * a.part_stride + idx] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < SB; ++u)
                if (s0 + 4 * u < ns) g = g + v[u];
        }
    }
    g = g + xor_lane<1>(g);
    g = g + xor_lane<2>(g);
    return g * (valid ? a.vol[c] : 0.0);
}

// Copies of a new state value for the next contraction: FP64 state-major (padded states), and
// the BF16 hi/lo split in the contraction's fragment order (fb_index) when that precision is in
// use.
__device__ inline void publish(const UpdateArgs& a, int c, int k, double x) {
    const int kp = a.pad_index[k];
    a.fT[(int64_t)kp * a.ldc + c] = x;
    if (a.fT32) a.fT32[(int64_t)kp * a.ldc + c] = (float)x;
    if (a.fhi) {
        uint16_t hi, lo;
        split_bf16(x, &hi, &lo);
        const int64_t q = fb_index(c, kp, a.ldb);
        a.fhi[q] = hi;
        a.flo[q] = lo;
    }
}

// ------------------------------------------------------------------ emission GEMV ---------
// df_Emi = M_Emi f (src/SteppingMethods.jl:35-38) over the per-cell N x N blocks, into demi
// before the update reads it.  The blocks are column-major: a wave reads 64 R consecutive rows
// of one column per load (R = 2, 16-byte loads of two rows, when N is even) and EMI_ROWS = 2
// such row groups per column, so each column is read in runs of 2 KB.  Workgroup (cell, column
// range q of Q, tile of 128 R rows): EMI_WAVES waves, wave p summing its share of the range in
// ascending column order.  The loads run EMI_BATCH columns at a time in two alternating register
// sets, so the next batch is in flight while the current one is added; the cell's f sits in LDS
// (every lane reads the same f_j).  The wave sums are added pairwise (((s0 + s1) + (s2 + s3)) +
// ...).  Q keeps the launch at about one workgroup per CU when there are few cells; with Q > 1
// the range sums go to emi_part and emission_sum_kernel adds them in range order.  Measured at
// 128 cells x 512 states (268 MB): 1 KB runs (one row group) 60 us, 2 KB 50 us, 4 KB 55 us; two
// workgroups per CU (Q = 2) 67 us; 8 to 16 waves and batches of 2 to 8 columns alike.
constexpr int EMI_WAVES = 8;
constexpr int EMI_BATCH = 4;
constexpr int EMI_ROWS = 2;   // loads of 64 R rows per lane and column
constexpr int EMI_WG_PER_CU = 1;

int emission_parts(int N, int C, int num_cu) {
    const int64_t tiles = (int64_t)C * ((N + 128 * EMI_ROWS - 1) / (128 * EMI_ROWS));
    const int64_t want = (EMI_WG_PER_CU * (int64_t)num_cu + tiles - 1) / tiles;
    return (int)std::max<int64_t>(1, std::min<int64_t>(want, N / 64));
}
static_assert((EMI_WAVES & (EMI_WAVES - 1)) == 0, "part sums are added pairwise");

template <int R, int T>
__global__ __launch_bounds__(64 * EMI_WAVES) void emission_gemv_kernel(UpdateArgs a) {
    extern __shared__ double fl[];   // f of the cell
    __shared__ double part[EMI_WAVES][64 * R * T];
    constexpr int B = EMI_BATCH;
    constexpr int RW = 64 * R;         // rows of one load of the wave
    const int N = a.N, Q = a.emi_q, c = blockIdx.x / Q, q = blockIdx.x % Q;
    const int lane = threadIdx.x & 63;
    const int p = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int kt = blockIdx.y * RW * T + lane * R;   // rows kt + RW t (+ 1 when R = 2) of the lane
    const int q0 = (int)((int64_t)N * q / Q), q1 = (int)((int64_t)N * (q + 1) / Q);
    const int j0 = q0 + (int)((int64_t)(q1 - q0) * p / EMI_WAVES);
    const int j1 = q0 + (int)((int64_t)(q1 - q0) * (p + 1) / EMI_WAVES);
    const double* E = a.emi + (int64_t)c * N * N;
    typedef double d2v __attribute__((ext_vector_type(2)));
    double ea[B][T][R], eb[B][T][R];
    // columns jc .. jc + B - 1 into e; past j1 the last column is read again and not added, and
    // rows past N read row 0 and are not written, so the loads carry no branch and the compiler
    // counts them exactly
    auto issue = [&](int jc, double (&e)[B][T][R]) {
#pragma unroll
        for (int u = 0; u < B; ++u)
#pragma unroll
            for (int t = 0; t < T; ++t) {
                const int k = kt + RW * t;
                const double* col = E + (int64_t)min(jc + u, j1 - 1) * N + (k < N ? k : 0);
                if constexpr (R == 2) {
                    const d2v x = *(const d2v*)col;
                    e[u][t][0] = x[0];
                    e[u][t][1] = x[1];
                } else {
                    e[u][t][0] = col[0];
                }
            }
    };
    double acc[T][R];
#pragma unroll
    for (int t = 0; t < T; ++t)
#pragma unroll
        for (int r = 0; r < R; ++r) acc[t][r] = 0.0;
    auto consume = [&](int jc, const double (&e)[B][T][R]) {
        double fj[B];
#pragma unroll
        for (int u = 0; u < B; ++u) fj[u] = fl[min(jc + u, j1 - 1)];
#pragma unroll
        for (int u = 0; u < B; ++u) {
            const bool in = jc + u < j1;
#pragma unroll
            for (int t = 0; t < T; ++t)
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const double x = acc[t][r] + e[u][t][r] * fj[u];
                    acc[t][r] = in ? x : acc[t][r];
                }
        }
    };
    if (j0 < j1) issue(j0, ea);
    const double* fc = a.f + (int64_t)c * N;
    for (int j = threadIdx.x; j < N; j += 64 * EMI_WAVES) fl[j] = fc[j];
    __syncthreads();
    // the next batch is issued before the current one is added, also past j1 (a few repeated
    // loads of the last column at the end), so every wait is for one batch exactly; the
    // scheduling barriers keep the compiler from sinking the loads below the adds
    for (int jc = j0; jc < j1;) {
        issue(jc + B, eb);
        __builtin_amdgcn_sched_barrier(0);
        consume(jc, ea);
        jc += B;
        if (jc >= j1) break;
        issue(jc + B, ea);
        __builtin_amdgcn_sched_barrier(0);
        consume(jc, eb);
        jc += B;
    }
#pragma unroll
    for (int t = 0; t < T; ++t)
#pragma unroll
        for (int r = 0; r < R; ++r) part[p][RW * t + lane * R + r] = acc[t][r];
    __syncthreads();
    if (p == 0)
#pragma unroll
        for (int t = 0; t < T; ++t)
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = RW * t + lane * R + r, k = kt + RW * t + r;
                double v[EMI_WAVES];
#pragma unroll
                for (int w = 0; w < EMI_WAVES; ++w) v[w] = part[w][i];
#pragma unroll
                for (int h = EMI_WAVES / 2; h >= 1; h /= 2)
#pragma unroll
                    for (int w = 0; w < h; ++w) v[w] = v[2 * w] + v[2 * w + 1];
                if (kt + RW * t < N) {
                    if (Q == 1) a.demi[(int64_t)c * N + k] = v[0];
                    else a.emi_part[((int64_t)q * a.C + c) * N + k] = v[0];
                }
            }
}

// demi = the Q range sums of each state added in range order
__global__ void emission_sum_kernel(UpdateArgs a) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n;
         i += (int64_t)gridDim.x * blockDim.x) {
        double s = a.emi_part[i];
        for (int q = 1; q < a.emi_q; ++q) s = s + a.emi_part[(int64_t)q * a.n + i];
        a.demi[i] = s;
    }
}

hipError_t launch_emission(const UpdateArgs& a, hipStream_t stream) {
    if (!a.emission) return hipSuccess;
    if (a.N > EMI_MAX_N || a.emi_q < 1 || (a.emi_q > 1 && !a.emi_part))
        return hipErrorInvalidValue;
    constexpr int T = EMI_ROWS;
    const size_t lds = (size_t)a.N * sizeof(double);
    if (a.N % 2 == 0) {
        const dim3 grid(a.C * a.emi_q, (a.N + 128 * T - 1) / (128 * T));
        emission_gemv_kernel<2, T><<<grid, 64 * EMI_WAVES, lds, stream>>>(a);
    } else {
        const dim3 grid(a.C * a.emi_q, (a.N + 64 * T - 1) / (64 * T));
        emission_gemv_kernel<1, T><<<grid, 64 * EMI_WAVES, lds, stream>>>(a);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || a.emi_q == 1) return e;
    const int64_t blocks = std::min<int64_t>((a.n + 255) / 256, 4096);
    emission_sum_kernel<<<(unsigned)blocks, 256, 0, stream>>>(a);
    return hipGetLastError();
}

// Values reduced together by one workgroup: per-species number and energy sums, then the block
// partial words (sums first, minima after).
constexpr int RED_VALUES = 2 * DTX_MAX_SPECIES + BP_WORDS;

// Fused workgroup reductions.  Each value is first reduced over its wave by a butterfly, and
// after one barrier thread t combines the four wave results of value t in wave order
// (wave_combine).  The pattern is fixed, so the reductions are deterministic.
//
// wave_reduce_batch runs the butterflies of WB values at once, with the arithmetic of wave_sum /
// wave_min (bit i of summask selects the sum for value i) and the shuffles of one level issued
// together, so a batch costs about the latency of one butterfly.
constexpr int WB = 8;
__device__ inline void wave_reduce_batch(double* v, uint32_t summask) {
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) {
        double x[WB];
#pragma unroll
        for (int i = 0; i < WB; ++i) x[i] = xor_lane_rt(v[i], m);
#pragma unroll
        for (int i = 0; i < WB; ++i)
            v[i] = ((summask >> i) & 1u) ? v[i] + x[i] : cr_min(v[i], x[i]);
    }
}

// t: the thread's index in its group of UB threads (red: that group's array)
__device__ inline void wave_put(double x, int i, bool is_sum, double* red, int t = threadIdx.x) {
    const double w = is_sum ? wave_sum(x) : wave_min(x);
    if ((t & 63) == 0) red[(t >> 6) * RED_VALUES + i] = w;
}

__device__ inline double wave_combine(int t, bool is_sum, const double* red) {
    double r = red[t];
    for (int w = 1; w < UB / 64; ++w)
        r = is_sum ? r + red[w * RED_VALUES + t] : cr_min(r, red[w * RED_VALUES + t]);
    return r;
}

// The moment-rate correction's (alpha, beta) of (cell c, species s) into out[0..1], by one wave:
// the moment partials of the gain sums (workgroups of the sums launch, lane-strided in order)
// and the moment-rate item sums (lane-strided in list order), one butterfly, then the 2x2
// system.  Every update workgroup of the cell runs the same arithmetic, so they agree exactly.
// v[3 + m], m = 0 (number) and 1 (energy): the lane's moment-rate item sums of (cell c,
// species s), lane-strided in list order.
__device__ inline void mom_rate_sums(const UpdateArgs& a, int c, int s, int lane, double* v) {
#pragma unroll
    for (int m = 0; m < 2; ++m) {
        const int t0 = a.mom_off[2 * s + m], t1 = a.mom_off[2 * s + m + 1];
        for (int t = t0 + lane; t < t1; t += 64)
            v[3 + m] = v[3 + m] + a.mpart[(int64_t)a.mom_list[t] * a.ldc + c];
    }
}

// v[0..2]: the lane's sums of the gain-sum moment partials cp[z][3], z = lane, lane + 64, ...
__device__ inline void cpart_sums(const double* cp, int nz, int lane, double* v) {
    for (int z = lane; z < nz; z += 64)
#pragma unroll
        for (int i = 0; i < 3; ++i) v[i] = v[i] + cp[z * 3 + i];
}

__device__ inline void coef_solve(const UpdateArgs& a, double* v, int lane, double* out);

__device__ inline void correction_coef(const UpdateArgs& a, int c, int s, double* out) {
    const int lane = threadIdx.x & 63;
    const int64_t p = (int64_t)c * a.S + s;
    const int ns = a.species_off[s + 1] - a.species_off[s];
    const int nz = (ns + a.zspan - 1) / a.zspan;
    double v[WB] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    cpart_sums(a.cpart + p * a.zb * 3, nz, lane, v);
    mom_rate_sums(a, c, s, lane, v);
    coef_solve(a, v, lane, out);
}

// One butterfly over the wave, then (lane 0) the 2x2 system into out[0..1].
__device__ inline void coef_solve(const UpdateArgs& a, double* v, int lane, double* out) {
    wave_reduce_batch(v, 0xFFu);
    if (lane != 0) return;
    const double S0 = v[0], S1 = v[1], S2 = v[2], RN = v[3], RE = v[4];
    double alpha = 0.0, beta = 0.0;
    if (S0 != 0.0) {
        const double dn = RN - S0, de = RE - S1;
        const double det = S0 * S2 - S1 * S1;
        if (a.eweight && fabs(det) > 1e-12 * fabs(S0 * S2)) {
            alpha = (dn * S2 - de * S1) / det;
            beta = (de * S0 - dn * S1) / det;
        } else {
            alpha = dn / S0;
        }
        if (!isfinite(alpha) || !isfinite(beta)) alpha = beta = 0.0;
    }
    out[0] = alpha;
    out[1] = beta;
}

// The coefficients of cell c for binary_term (indexed 2 species): those of correct_coef_kernel
// when it ran, else those of every species with states in [k0, k1), derived into the
// workgroup's LDS array one wave per species (ends with a barrier).
__device__ inline const double* load_correction(const UpdateArgs& a, int c, int k0, int k1,
                                                double* lds) {
    if (a.coef) return a.coef + (int64_t)c * a.S * 2;
    if (k0 < k1) {
        const int sa = a.state_species[k0], sb = a.state_species[k1 - 1];
        for (int s = sa + (int)(threadIdx.x >> 6); s <= sb; s += UB / 64)
            correction_coef(a, c, s, lds + 2 * s);
    }
    __syncthreads();
    return lds;
}

// One wave per (cell, species): the coefficients for large update grids, where deriving them in
// every update workgroup would repeat the work once per round of workgroups.
__global__ __launch_bounds__(UB) void correct_coef_kernel(UpdateArgs a) {
    const int64_t p = (int64_t)blockIdx.x * (UB / 64) + (threadIdx.x >> 6);
    if (p >= (int64_t)a.C * a.S) return;  // the whole wave
    correction_coef(a, (int)(p / a.S), (int)(p % a.S), a.coef + p * 2);
}

// The species and energy weight of state k (eweight absent: 0), for finish_block and the gain
// correction.
struct StateMeta {
    int sp;
    double ew;
};
__device__ inline StateMeta state_meta(const UpdateArgs& a, int k, bool mine) {
    StateMeta m{-1, 0.0};
    if (mine) {
        m.sp = a.state_species[k];
        m.ew = a.eweight ? a.eweight[k] : 0.0;
    }
    return m;
}

// Per-species moment partials of this block's new states and the per-block partials (sums and
// Cr minima), in one fused reduction.  Thread (tid & 3) == 0 of each state holds its values;
// the finalize kernels fold the partials in a fixed order.
// t: the thread's index in its group of UB threads (red: the group's array); write: the group
// owns a segment (every group of the workgroup takes part in the barrier).  sp and ew: the
// state's species and energy weight (state_meta), loaded by the caller with the state's other
// values so that no load waits at the end of the kernel.
__device__ void finish_block(const UpdateArgs& a, int c, int seg, int nseg, bool mine, int sp_in,
                             double ew_in, double fnew, double* red, double sumdf, double sumf,
                             double cr, double crb, double cre, double crf, double crm,
                             double crs, int t = threadIdx.x, bool write = true) {
    const int tid = t;
    const int64_t b = (int64_t)c * nseg + seg;
    const int S = a.S;
    const int sp = mine ? sp_in : -1;
    const double ew = mine ? ew_in : 0.0;
    for (int s = 0; s < S; ++s) {
        wave_put(s == sp ? fnew : 0.0, 2 * s, true, red, t);
        wave_put((s == sp && a.eweight) ? fnew * ew : 0.0, 2 * s + 1, true, red, t);
    }
    const bool want_cr = a.verbose >= 1 && a.verbose <= 3;
    const int w0 = 2 * S;
    wave_put(sumdf, w0 + BP_SUMDF, true, red, t);
    wave_put(sumf, w0 + BP_SUMF, true, red, t);
    if (want_cr) {
        wave_put(cr, w0 + BP_CR, false, red, t);
        if (a.verbose == 3) {
            wave_put(crb, w0 + BP_CR_BIN, false, red, t);
            wave_put(cre, w0 + BP_CR_EMI, false, red, t);
            wave_put(crf, w0 + BP_CR_FLUX, false, red, t);
            wave_put(crm, w0 + BP_CR_MOM, false, red, t);
            wave_put(crs, w0 + BP_CR_SPACE, false, red, t);
        }
    }
    __syncthreads();
    if (!write) return;
    if (tid < w0) {
        a.mom_part[(b * S + (tid >> 1)) * 2 + (tid & 1)] = wave_combine(tid, true, red);
    } else if (tid < w0 + BP_WORDS) {
        const int w = tid - w0;
        const bool reduced = w <= BP_SUMF || (want_cr && (w == BP_CR || a.verbose == 3));
        a.block_part[b * BP_WORDS + w] =
            reduced ? wave_combine(tid, w <= BP_SUMF, red) : INFINITY;
    }
}

// Finalisation of a step: the per-block partials of the update grid are folded into the
// moments of every (cell, species) and into the status record.  The fold has two fixed levels,
// so it runs on many workgroups and is still deterministic: fold group g reduces blocks
// [UB g, UB g + UB) (one per thread, fixed workgroup pattern) into fold partial g, and the
// fold partials are then reduced by one workgroup (thread t takes partials t, t + UB, ... in
// order, then the same workgroup pattern).  The moments of (cell, species) q add the segment
// partials of the cell in segment order.

// Reduce BP_WORDS values per thread over the workgroup (the sums, then the Cr minima); thread
// q < BP_WORDS returns word q.
__device__ inline double wg_reduce_words(const double* v, double* red) {
    double x[WB];
#pragma unroll
    for (int q = 0; q < BP_WORDS; ++q) x[q] = v[q];
    wave_reduce_batch(x, 0x3u);
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int q = 0; q < BP_WORDS; ++q) red[(threadIdx.x >> 6) * RED_VALUES + q] = x[q];
    __syncthreads();
    const int tid = threadIdx.x;
    const double r = tid < BP_WORDS ? wave_combine(tid, tid <= BP_SUMF, red) : 0.0;
    __syncthreads();
    return r;
}

__device__ inline void words_identity(double* v) {
#pragma unroll
    for (int q = 0; q < BP_WORDS; ++q) v[q] = (q <= BP_SUMF) ? 0.0 : INFINITY;
}

__device__ inline void words_fold(double* v, const double* p) {
    v[BP_SUMDF] = v[BP_SUMDF] + p[BP_SUMDF];
    v[BP_SUMF] = v[BP_SUMF] + p[BP_SUMF];
#pragma unroll
    for (int q = BP_CR; q < BP_WORDS; ++q) v[q] = cr_min(v[q], p[q]);
}

// Number of fold groups (also the number of finalize workgroups of a single simulation).
__device__ __host__ inline int fold_groups(int C, int S, int nseg) {
    const int64_t nb = (int64_t)C * nseg, nq = (int64_t)C * S;
    return (int)(((nb > nq ? nb : nq) + UB - 1) / UB);
}

// Moments of (cell, species) pair q: segment partials in segment order.
__device__ inline void fold_moments(const UpdateArgs& a, int nseg, int64_t q) {
    constexpr int LB = 8;
    const int64_t cc = q / a.S, sp = q % a.S;
    double num = 0.0, en = 0.0;
    for (int g0 = 0; g0 < nseg; g0 += LB) {
        double pn[LB], pe[LB];
#pragma unroll
        for (int u = 0; u < LB; ++u) {
            const int g = g0 + u;
            const double* mp = a.mom_part + (((cc * nseg + g) * a.S) + sp) * 2;
            pn[u] = g < nseg ? mp[0] : 0.0;
            pe[u] = g < nseg ? mp[1] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < LB; ++u)
            if (g0 + u < nseg) {
                num = num + pn[u];
                en = en + pe[u];
            }
    }
    a.moments[q * 2 + 0] = num;
    a.moments[q * 2 + 1] = en;
}

// Fold group g: moments of pairs [UB g, UB g + UB) and fold partial g of the block partials.
__device__ inline void fold_group(const UpdateArgs& a, int nseg, int g, double* red) {
    const int tid = threadIdx.x;
    const int64_t q = (int64_t)g * UB + tid;
    if (q < (int64_t)a.C * a.S) fold_moments(a, nseg, q);
    double v[BP_WORDS];
    words_identity(v);
    if (q < (int64_t)a.C * nseg) {
        double p[BP_WORDS];
#pragma unroll
        for (int w = 0; w < BP_WORDS; ++w) p[w] = a.block_part[q * BP_WORDS + w];
        words_fold(v, p);
    }
    const double r = wg_reduce_words(v, red);
    double* fold = a.block_part + (int64_t)a.C * nseg * BP_WORDS;
    if (tid < BP_WORDS) fold[(int64_t)g * BP_WORDS + tid] = r;
}

// Reduce the fold partials and write the status record of the step.
__device__ inline void fold_status(const UpdateArgs& a, int nseg, int ngroups, double* red) {
    const int tid = threadIdx.x;
    const double* fold = a.block_part + (int64_t)a.C * nseg * BP_WORDS;
    double v[BP_WORDS];
    words_identity(v);
    for (int g = tid; g < ngroups; g += UB) {
        double p[BP_WORDS];
#pragma unroll
        for (int w = 0; w < BP_WORDS; ++w) p[w] = fold[(int64_t)g * BP_WORDS + w];
        words_fold(v, p);
    }
    const double r = wg_reduce_words(v, red);
    if (tid < BP_WORDS) red[tid] = r;
    __syncthreads();
    if (tid == 0) {
        double w[BP_WORDS];
        for (int q = 0; q < BP_WORDS; ++q) w[q] = red[q];
        double* st = a.status + (int64_t)(a.step_index % a.ring) * STATUS_WORDS;
        const bool cr_eval = (a.verbose >= 1 && a.verbose <= 3) && (w[BP_SUMF] != 0.0);
        st[ST_STEP] = (double)a.step_index;
        st[ST_FINITE] = isfinite(w[BP_SUMDF]) ? 1.0 : 0.0;
        st[ST_CR_EVAL] = cr_eval ? 1.0 : 0.0;
        st[ST_CR] = cr_eval ? -w[BP_CR] : 0.0;
        st[ST_CR_BIN] = (cr_eval && a.verbose == 3 && a.binary) ? -w[BP_CR_BIN] : 0.0;
        st[ST_CR_EMI] = (cr_eval && a.verbose == 3 && a.emission) ? -w[BP_CR_EMI] : 0.0;
        st[ST_CR_FLUX] = (cr_eval && a.verbose == 3) ? -w[BP_CR_FLUX] : 0.0;
        st[ST_CR_MOM] = (cr_eval && a.verbose == 3) ? -w[BP_CR_MOM] : 0.0;
        st[ST_CR_SPACE] = (cr_eval && a.verbose == 3) ? -w[BP_CR_SPACE] : 0.0;
        st[ST_SUMF] = w[BP_SUMF];
        st[ST_DT] = a.dt;
        st[ST_TSTART] = a.t_start;
        st[ST_SAVE] = (double)a.save;
        st[ST_SUMDF] = w[BP_SUMDF];
    }
}

// Small grids (at most FINALIZE_SINGLE_GROUPS fold groups): one workgroup folds the per-block
// partials (thread-strided, then the workgroup pattern) into the moments of every (cell, species)
// and the status record of the step.  Loads are issued in batches ahead of the ordered adds.
__device__ void finalize_single(const UpdateArgs& a, int nseg, double* red,
                                bool moments = true) {
    constexpr int LB = 8;
    const int tid = threadIdx.x;
    const int64_t nblocks = (int64_t)a.C * nseg;
    for (int64_t q = tid; moments && q < (int64_t)a.C * a.S; q += UB) {
        const int64_t cc = q / a.S, sp = q % a.S;
        double num = 0.0, en = 0.0;
        for (int g0 = 0; g0 < nseg; g0 += LB) {
            double pn[LB], pe[LB];
#pragma unroll
            for (int u = 0; u < LB; ++u) {
                const int g = g0 + u;
                const double* mp = a.mom_part + (((cc * nseg + g) * a.S) + sp) * 2;
                pn[u] = g < nseg ? mp[0] : 0.0;
                pe[u] = g < nseg ? mp[1] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < LB; ++u)
                if (g0 + u < nseg) {
                    num = num + pn[u];
                    en = en + pe[u];
                }
        }
        a.moments[q * 2 + 0] = num;
        a.moments[q * 2 + 1] = en;
    }
    double v[BP_WORDS];
#pragma unroll
    for (int q = 0; q < BP_WORDS; ++q) v[q] = (q <= BP_SUMF) ? 0.0 : INFINITY;
    constexpr int BB = 4;  // blocks per batch (BB * BP_WORDS loads in flight)
    for (int64_t b0 = tid; b0 < nblocks; b0 += (int64_t)UB * BB) {
        double p[BB][BP_WORDS];
#pragma unroll
        for (int u = 0; u < BB; ++u) {
            const int64_t bb = b0 + (int64_t)u * UB;
#pragma unroll
            for (int q = 0; q < BP_WORDS; ++q)
                p[u][q] = bb < nblocks ? a.block_part[bb * BP_WORDS + q] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < BB; ++u) {
            if (b0 + (int64_t)u * UB >= nblocks) break;
            v[BP_SUMDF] = v[BP_SUMDF] + p[u][BP_SUMDF];
            v[BP_SUMF] = v[BP_SUMF] + p[u][BP_SUMF];
#pragma unroll
            for (int q = BP_CR; q < BP_WORDS; ++q) v[q] = cr_min(v[q], p[u][q]);
        }
    }
#pragma unroll
    for (int q = 0; q < BP_WORDS; ++q) wave_put(v[q], q, q <= BP_SUMF, red);
    __syncthreads();
    const double r = tid < BP_WORDS ? wave_combine(tid, tid <= BP_SUMF, red) : 0.0;
    __syncthreads();
    if (tid < BP_WORDS) red[tid] = r;
    __syncthreads();
    if (tid == 0) {
        double w[BP_WORDS];
        for (int q = 0; q < BP_WORDS; ++q) w[q] = red[q];
        double* st = a.status + (int64_t)(a.step_index % a.ring) * STATUS_WORDS;
        const bool cr_eval = (a.verbose >= 1 && a.verbose <= 3) && (w[BP_SUMF] != 0.0);
        st[ST_STEP] = (double)a.step_index;
        st[ST_FINITE] = isfinite(w[BP_SUMDF]) ? 1.0 : 0.0;
        st[ST_CR_EVAL] = cr_eval ? 1.0 : 0.0;
        st[ST_CR] = cr_eval ? -w[BP_CR] : 0.0;
        st[ST_CR_BIN] = (cr_eval && a.verbose == 3 && a.binary) ? -w[BP_CR_BIN] : 0.0;
        st[ST_CR_EMI] = (cr_eval && a.verbose == 3 && a.emission) ? -w[BP_CR_EMI] : 0.0;
        st[ST_CR_FLUX] = (cr_eval && a.verbose == 3) ? -w[BP_CR_FLUX] : 0.0;
        st[ST_CR_MOM] = (cr_eval && a.verbose == 3) ? -w[BP_CR_MOM] : 0.0;
        st[ST_CR_SPACE] = (cr_eval && a.verbose == 3) ? -w[BP_CR_SPACE] : 0.0;
        st[ST_SUMF] = w[BP_SUMF];
        st[ST_DT] = a.dt;
        st[ST_TSTART] = a.t_start;
        st[ST_SAVE] = (double)a.save;
        st[ST_SUMDF] = w[BP_SUMDF];
    }
}

// Whether any entry of cell c's f is non-zero (the Domain all-zero test), over the workgroup.
template <int NT = UB>
__device__ int cell_nonzero(const UpdateArgs& a, int c, int* ired) {
    const int64_t base = (int64_t)c * a.N;
    int nz = 0;
    for (int k = threadIdx.x; k < a.N; k += NT) nz |= (a.f[base + k] != 0.0);
    return block_or<NT>(nz, ired);
}

// ----------------------------------------------------------------------- ForwardEuler ------
// Flux row r spread over the four lanes of its state: lane q loads ELL slots q, q + 4, ...
// (ELL_LANE of them, so rows up to 4 ELL_LANE wide in one pass), gathers their x and forms the
// products; ell_sum then adds the products in slot order (ascending column) on lane q = 0, the
// order and operations of ell_row.  Padding slots (column -1) contribute nothing.
constexpr int ELL_LANE = 2;

struct EllLane {
    int32_t c[ELL_LANE];
    double v[ELL_LANE];
};

__device__ inline void ell_issue(const UpdateArgs& a, int64_t r, int q, bool mine, int s0,
                                 EllLane& L) {
#pragma unroll
    for (int j = 0; j < ELL_LANE; ++j) {
        const int e = s0 + q + 4 * j;
        const bool in = mine && e < a.ell_width;
        L.c[j] = in ? a.ell_col[(int64_t)e * a.n + r] : -1;
        L.v[j] = in ? a.ell_val[(int64_t)e * a.n + r] : 0.0;
    }
}

// The gathered f of the lane's slots.
__device__ inline void ell_gather(const UpdateArgs& a, const EllLane& L, double* x) {
#pragma unroll
    for (int j = 0; j < ELL_LANE; ++j) x[j] = L.c[j] >= 0 ? a.f[L.c[j]] : 0.0;
}

// Adds the products of slots [s0, s0 + 4 ELL_LANE) to acc on lane q = 0 of each state (x: the
// gathered f of the lane's slots).
__device__ inline double ell_sum_x(const EllLane& L, const double* x, int s0, int width,
                                   double acc) {
    double p[ELL_LANE];
#pragma unroll
    for (int j = 0; j < ELL_LANE; ++j) p[j] = L.c[j] >= 0 ? L.v[j] * x[j] : 0.0;
#pragma unroll
    for (int j = 0; j < ELL_LANE; ++j)
#pragma unroll
        for (int qq = 0; qq < 4; ++qq) {
            const double pv = quad_lane_rt(p[j], qq);
            const int cv = quad_lane_rt(L.c[j], qq);
            if (s0 + 4 * j + qq < width && cv >= 0) acc = acc + pv;
        }
    return acc;
}

__device__ inline double ell_sum(const UpdateArgs& a, const EllLane& L, int s0, double acc) {
    double x[ELL_LANE];
    ell_gather(a, L, x);
    return ell_sum_x(L, x, s0, a.ell_width, acc);
}

__device__ void euler_update_body(const UpdateArgs& a, int c, int seg, int nseg, double* red,
                                  int* ired) {
    const int tid = threadIdx.x, N = a.N;
    const int k0 = seg * UPDATE_STATES, k1 = min(N, k0 + UPDATE_STATES);
    const int64_t base = (int64_t)c * N;
    const int k = k0 + (tid >> 2), q = tid & 3;
    const bool mine = k < k1;
    const int64_t r = base + k;
    // Every load that does not depend on another is issued before the first wait: the state's
    // own values (with its emission term), its flux row's columns and values (over the four
    // lanes of the state), and, with the moment-rate correction, its gain and loss sums; the
    // workgroup barriers below (Domain test, coefficients) wait only for LDS, so these stay in
    // flight across them.
    double fk = 0.0, iap = 0.0, dinj = 0.0, dE = 0.0;
    if (mine && q == 0) {
        fk = a.f[r];
        iap = a.inv_ap[r];
        dinj = a.df_inj[r];
        if (a.emission) dE = a.demi[r];
    }
    const StateMeta meta = state_meta(a, k, mine && q == 0);
    EllLane ell;
    ell_issue(a, r, q, mine, 0, ell);
    double gs = 0.0, ls = 0.0;
    if (a.gsum && mine && q == 0) {
        gs = a.gsum[r];
        ls = a.part[r];
    }
    const int any_nz = (a.binary && a.active) ? cell_nonzero(a, c, ired) : 1;
    // update_Big_Bin! evaluates a cell when Domain is nothing, the cell is in Domain, or its f
    // is all zero (src/SteppingMethods.jl:369); otherwise df_Bin keeps its zero initial value.
    const bool active = a.binary && (a.active == nullptr || a.active[c] || !any_nz);

    double sumdf = 0.0, sumf = 0.0, cr = INFINITY, crb = INFINITY, cre = INFINITY,
           crf = INFINITY, fnew = 0.0;
    const bool want_cr = a.verbose >= 1 && a.verbose <= 3;
    __shared__ double coef_lds[2 * DTX_MAX_SPECIES];
    const double* coef = a.gsum ? load_correction(a, c, k0, k1, coef_lds) : nullptr;
    double dB = 0.0;
    if (a.binary) {
        if (a.gsum) {
            // the state's gain sum rescaled by (1 + alpha + beta E_k) of its species, plus its
            // loss sum (binary_term's corrected form, on the sums loaded above)
            double g = 0.0;
            if (mine && active && q == 0) {
                const double* cf = coef + 2 * meta.sp;
                g = gs * (1.0 + (cf[0] + cf[1] * meta.ew)) + ls;
            }
            g = g + xor_lane<1>(g);
            g = g + xor_lane<2>(g);
            dB = g * ((mine && active) ? a.vol[c] : 0.0);
        } else {
            dB = binary_term(a, c, k, q, mine && active, coef);
        }
    }
    double dF = ell_sum(a, ell, 0, 0.0);
    for (int s0 = 4 * ELL_LANE; s0 < a.ell_width; s0 += 4 * ELL_LANE) {  // wider rows
        EllLane more;
        ell_issue(a, r, q, mine, s0, more);
        dF = ell_sum(a, more, s0, dF);
    }
    if (mine && q == 0) {
        sumf = fk;
        double df = -dF;
        if (a.emission) df = df + dE;
        if (a.binary) df = df + dB;
        df = df * (iap * a.s);
        sumdf = df;
        if (want_cr) {
            cr = cr_min(cr, df / fk);
            if (a.verbose == 3) {
                if (a.binary) crb = cr_min(crb, dB / fk);
                if (a.emission) cre = cr_min(cre, dE / fk);
                crf = cr_min(crf, (-dF) / fk);
            }
        }
        double x = fk + (df + dinj * a.s);
        x = strong_zero(x, x >= THR_1F_10);
        x = strong_zero(x, x != INFINITY);
        a.f_out[r] = x;
        fnew = x;
        publish(a, c, k, x);
    }
    finish_block(a, c, seg, nseg, mine && q == 0, meta.sp, meta.ew, fnew, red, sumdf, sumf, cr,
                 crb, cre, crf, INFINITY, INFINITY);
}

// Grid (C, nseg).  With batched finalisation, workgroup (0, 0) also records the step's scalars.
__global__ __launch_bounds__(UB) void euler_update_kernel(UpdateArgs a) {
    __shared__ double red[(UB / 64) * RED_VALUES];
    __shared__ int ired[UB / 64];
    if (a.fin_rec && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
        double* rec = a.fin_rec + (int64_t)a.fin_slot * FIN_WORDS;
        rec[0] = (double)a.step_index;
        rec[1] = a.t_start;
        rec[2] = a.dt;
        rec[3] = (double)a.save;
        rec[4] = (double)a.verbose;
    }
    euler_update_body(a, blockIdx.x, blockIdx.y, gridDim.y, red, ired);
}

// Batched finalisation: workgroup i finalises the pending step in ring slot (first + i) mod
// FIN_BATCH from its record and block partials, with the arithmetic of the step's own
// finalize_single; only the last step's moments are written (the moments are those of the
// current state).
__global__ __launch_bounds__(UB) void finalize_batch_kernel(UpdateArgs base, int nseg,
                                                            int64_t bp_words, int first,
                                                            int count) {
    __shared__ double red[(UB / 64) * RED_VALUES];
    const int i = blockIdx.x;
    const int slot = (first + i) % FIN_BATCH;
    const double* rec = base.fin_rec + (int64_t)slot * FIN_WORDS;
    UpdateArgs a = base;
    a.block_part = base.block_part + (int64_t)slot * bp_words;
    a.step_index = (int32_t)rec[0];
    a.t_start = rec[1];
    a.dt = rec[2];
    a.save = (int32_t)rec[3];
    a.verbose = (int32_t)rec[4];
    finalize_single(a, nseg, red, i == count - 1);
}

// The two levels are two launches: every fold group in parallel, then the final reduction
// (the kernel boundary makes the fold partials visible to it; a device-wide fence inside one
// launch would write back the L2 of every XCD and cost more than the second launch).
__global__ __launch_bounds__(UB) void finalize_fold_kernel(UpdateArgs a, int nseg) {
    __shared__ double red[(UB / 64) * RED_VALUES];
    fold_group(a, nseg, blockIdx.x, red);
}

__global__ __launch_bounds__(UB) void finalize_status_kernel(UpdateArgs a, int nseg, int ngroups) {
    __shared__ double red[(UB / 64) * RED_VALUES];
    fold_status(a, nseg, ngroups, red);
}

// Both levels in one workgroup, with the arithmetic of the two launches: wave w reduces its
// 64-block slice of each fold group by the butterfly of wave_put, the four wave results of a
// group are combined in wave order as wave_combine does, and the fold partials then go through
// fold_status.  Groups are taken SG at a time, so their loads are in flight together.
constexpr int SG = 4;

__device__ inline void finalize_serial(const UpdateArgs& a, int nseg, double* red, double* wres) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int ng = fold_groups(a.C, a.S, nseg);
    const int64_t nblocks = (int64_t)a.C * nseg;
    for (int64_t q = tid; q < (int64_t)a.C * a.S; q += UB) fold_moments(a, nseg, q);
    double* fold = a.block_part + nblocks * BP_WORDS;
    for (int g0 = 0; g0 < ng; g0 += SG) {
        double p[SG][BP_WORDS];
#pragma unroll
        for (int u = 0; u < SG; ++u) {
            const int64_t b = (int64_t)(g0 + u) * UB + tid;
            words_identity(p[u]);
            if (g0 + u < ng && b < nblocks) {
                double x[BP_WORDS];
#pragma unroll
                for (int w = 0; w < BP_WORDS; ++w) x[w] = a.block_part[b * BP_WORDS + w];
                words_fold(p[u], x);
            }
        }
#pragma unroll
        for (int u = 0; u < SG; ++u) wave_reduce_batch(p[u], 0x3u);
        if (lane == 0)
#pragma unroll
            for (int u = 0; u < SG; ++u)
#pragma unroll
                for (int q = 0; q < BP_WORDS; ++q)
                    wres[(u * (UB / 64) + wave) * BP_WORDS + q] = p[u][q];
        __syncthreads();
        if (tid < SG * BP_WORDS && g0 + tid / BP_WORDS < ng) {
            const int u = tid / BP_WORDS, q = tid % BP_WORDS;
            double r = wres[(u * (UB / 64)) * BP_WORDS + q];
            for (int w = 1; w < UB / 64; ++w) {
                const double x = wres[(u * (UB / 64) + w) * BP_WORDS + q];
                r = q <= BP_SUMF ? r + x : cr_min(r, x);
            }
            fold[(int64_t)(g0 + u) * BP_WORDS + q] = r;
        }
        __syncthreads();
    }
    fold_status(a, nseg, ng, red);
}

// Grids with at most this many fold groups are finalised by one workgroup (finalize_single):
// one short launch instead of two.
constexpr int FINALIZE_SINGLE_GROUPS = 16;

__global__ __launch_bounds__(UB) void finalize_single_kernel(UpdateArgs a, int nseg) {
    __shared__ double red[(UB / 64) * RED_VALUES];
    finalize_single(a, nseg, red);
}

// Member of a grouped launch that owns block b: the last g with start[g] <= b.
__device__ inline int group_member(const int32_t* start, int count, int b) {
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Grouped forms for several simulations in one launch (serving): block b of the update grid
// belongs to the simulation g with block_start[g] <= b < block_start[g+1]; the finalize grid has
// one block per simulation.
__global__ __launch_bounds__(UB) void euler_update_group_kernel(const UpdateArgs* table,
                                                                const int32_t* block_start,
                                                                int count) {
    __shared__ double red[(UB / 64) * RED_VALUES];
    __shared__ int ired[UB / 64];
    const int b = blockIdx.x;
    const int lo = group_member(block_start, count, b);
    const UpdateArgs a = table[lo];
    const int nseg = (a.N + UPDATE_STATES - 1) / UPDATE_STATES;
    const int local = b - block_start[lo];
    euler_update_body(a, local / nseg, local % nseg, nseg, red, ired);
}

// One workgroup per simulation, with the arithmetic the simulation gets when stepped alone
// (finalize_single for small grids, the serial form of the two-level fold otherwise), so its
// status and moments do not depend on whether it was stepped alone or in a group.
__global__ __launch_bounds__(UB) void finalize_group_kernel(const UpdateArgs* table) {
    __shared__ double red[(UB / 64) * RED_VALUES];
    __shared__ double wres[SG * (UB / 64) * BP_WORDS];
    const UpdateArgs a = table[blockIdx.x];
    const int nseg = (a.N + UPDATE_STATES - 1) / UPDATE_STATES;
    if (fold_groups(a.C, a.S, nseg) <= FINALIZE_SINGLE_GROUPS) finalize_single(a, nseg, red);
    else finalize_serial(a, nseg, red, wres);
}

// ------------------------------------------------------------- ForwardSymplecticEuler ------
// Momentum half (src/SteppingMethods.jl:137-160): P flux, emission (demi, from the emission
// GEMV) and binary terms from f, f_Momentum = f + df_Momentum clamped at 1f-10.
__global__ __launch_bounds__(UB) void symplectic_momentum_kernel(UpdateArgs a) {
    __shared__ int ired[UB / 64];
    const int c = blockIdx.x, tid = threadIdx.x, N = a.N;
    const int k0 = blockIdx.y * UPDATE_STATES, k1 = min(N, k0 + UPDATE_STATES);
    const int64_t base = (int64_t)c * N;
    const int any_nz = (a.binary && a.active) ? cell_nonzero(a, c, ired) : 1;
    const bool active = a.binary && (a.active == nullptr || a.active[c] || !any_nz);
    const int k = k0 + (tid >> 2), q = tid & 3;
    const bool mine = k < k1;
    __shared__ double coef_lds[2 * DTX_MAX_SPECIES];
    const double* coef = a.gsum ? load_correction(a, c, k0, k1, coef_lds) : nullptr;
    const double dB = a.binary ? binary_term(a, c, k, q, mine && active, coef) : 0.0;
    if (mine && q == 0) {
        const int64_t r = base + k;
        const double dP = a.ell_width ? ell_row(a.ell_col, a.ell_val, a.ell_width, a.n, r, a.f)
                                      : 0.0;
        double dm = -dP;
        if (a.emission) dm = dm + a.demi[r];
        if (a.binary) {
            dm = dm + dB;
            a.dbin[r] = dB;
        }
        dm = dm * (a.inv_ap[r] * a.s);
        a.dpflux[r] = dP;
        a.df_mom[r] = dm;
        const double x = a.f[r] + dm;
        a.f_mom[r] = strong_zero(x, x >= THR_1F_10);
    }
}

// Space half and final update (src/SteppingMethods.jl:162-237).  The ELL fields hold X_Flux.
__global__ __launch_bounds__(UB) void symplectic_space_kernel(UpdateArgs a) {
    __shared__ double red[(UB / 64) * RED_VALUES];
    const int c = blockIdx.x, tid = threadIdx.x, N = a.N;
    const int k0 = blockIdx.y * UPDATE_STATES, k1 = min(N, k0 + UPDATE_STATES);
    const int64_t base = (int64_t)c * N;
    double sumf = 0.0, cr = INFINITY, crb = INFINITY, cre = INFINITY, crf = INFINITY;
    double crm = INFINITY, crs = INFINITY, fnew = 0.0;
    const bool want_cr = a.verbose >= 1 && a.verbose <= 3;
    const int k = k0 + (tid >> 2);
    const StateMeta meta = state_meta(a, k, k < k1 && (tid & 3) == 0);
    if (k < k1 && (tid & 3) == 0) {
        const int64_t r = base + k;
        const double f0 = a.f[r];
        sumf = f0;
        const double fm = a.f_mom[r];
        const double dX = a.ell_width ? ell_row(a.ell_col, a.ell_val, a.ell_width, a.n, r,
                                                a.f_mom)
                                      : 0.0;
        const double dS = ((-dX) * a.inv_ap[r]) * a.s;
        double fs = fm + dS;
        fs = strong_zero(fs, fs >= THR_1F_20);
        if (want_cr) {
            const double dm = a.df_mom[r];
            cr = cr_min(cr, (dS + dm) / f0);
            if (a.verbose == 3) {
                if (a.binary) crb = cr_min(crb, ((a.dbin[r] * a.inv_ap[r]) * a.s) / f0);
                if (a.emission) cre = cr_min(cre, ((a.demi[r] * a.inv_ap[r]) * a.s) / f0);
                double rp = (((-a.dpflux[r]) * a.inv_ap[r]) * a.s) / f0;
                if (isnan(rp)) rp = 0.0;  // replace!(df_tmp, NaN => 0.0)
                crf = cr_min(crf, rp);
                crm = cr_min(crm, dm / f0);
                crs = cr_min(crs, dS / fm);
            }
        }
        double x = fs + a.df_inj[r] * a.s;
        x = strong_zero(x, x >= THR_1F_20) * jl_sign(x);
        x = strong_zero(x, x != INFINITY);
        a.f_out[r] = x;
        fnew = x;
        publish(a, c, k, x);
    }
    finish_block(a, c, blockIdx.y, gridDim.y, k < k1 && (tid & 3) == 0, meta.sp, meta.ew, fnew,
                 red, 0.0, sumf, cr, crb, cre, crf, crm, crs);
}

// --------------------------------------------------------- FP64 moment-rate correction -----
// The reduced-precision gain contraction carries a relative error of its arithmetic in every
// term; the exact FP64 number and energy rates of the gain term, R_N = sum_ij W_N[i,j] f_i f_j
// and R_E likewise, come from the moment-rate slots.  Per (cell c, output species s) the gain
// sums of the species are rescaled per state by (1 + alpha + beta E_k) with (alpha, beta) such
// that their number and energy sums equal R_N and R_E (number only when no energy weights are
// set), and the loss sums are added unchanged.  One launch sums the slots of every state and
// forms per-workgroup moment partials (many workgroups, the slot reads in flight together);
// each update workgroup then derives the coefficients of its species (correction_coef) and
// applies them (binary_term).

// Sum of slots b + q, b + q + tps, ... below e of element idx (in order; FP32 slots if f32).
// FP32 slots: past the last slot a load repeats slot s0 (valid, unused), so the loads of a
// batch carry no branch and stay in flight together (a conditional load followed by its
// conversion waits for that load alone).
__device__ inline double slot_sum(const CorrectArgs& a, int64_t idx, int q, int b, int e,
                                  bool f32) {
    constexpr int LB = 8;
    const int t = a.tps;
    double sum = 0.0;
    for (int s0 = b + q; s0 < e; s0 += t * LB) {
        if (f32) {
            float v[LB];
#pragma unroll
            for (int u = 0; u < LB; ++u) {
                const int sl = s0 + t * u < e ? s0 + t * u : s0;
                v[u] = ((const float*)(a.part + (int64_t)sl * a.part_stride))[idx];
            }
#pragma unroll
            for (int u = 0; u < LB; ++u)
                if (s0 + t * u < e) sum = sum + (double)v[u];
        } else {
            double v[LB];
#pragma unroll
            for (int u = 0; u < LB; ++u)
                v[u] = s0 + t * u < e ? a.part[(int64_t)(s0 + t * u) * a.part_stride + idx] : 0.0;
#pragma unroll
            for (int u = 0; u < LB; ++u)
                if (s0 + t * u < e) sum = sum + v[u];
        }
    }
    return sum;
}

// Grid (C, S, ceil(max species states / (256 / tps))): tps threads per state q = 0 .. tps - 1
// (4, or 16 for small launches), each summing slots q, q + tps, ... of the gain slots (FP32 sums
// for wide split-BF16 items) and of the loss slots; the tps sums are combined by a butterfly in
// a fixed order ((s0 + s1) + (s2 + s3) for four).  Gain sums go to gsum, loss sums
// to lsum, and the workgroup's number and energy moments of its gain sums (sum g, sum E g,
// sum E^2 g) to cpart[c][s][z].
__device__ inline void correct_sums_body(const CorrectArgs& a, int c, int s, int z, double* red) {
    const int tid = threadIdx.x;
    const int q = tid & (a.tps - 1);
    const int k0 = a.species_off[s], ns = a.species_off[s + 1] - k0;
    const int kb = z * (UB / a.tps);
    if (kb >= ns) return;  // the whole workgroup
    const int kk = kb + tid / a.tps;
    const bool valid = kk < ns;
    const int64_t idx = (int64_t)c * a.N + k0 + (valid ? kk : 0);
    const int gs = a.gslots[s], nsl = a.nslots[s];
    double g = slot_sum(a, idx, q, 0, gs, a.gain_f32 != 0);
    double l = slot_sum(a, idx, q, gs, nsl, false);
    for (int m = 1; m < a.tps; m <<= 1) {
        g = g + xor_lane_rt(g, m);
        l = l + xor_lane_rt(l, m);
    }
    const bool mine = valid && q == 0;
    if (mine) {
        a.gsum[idx] = g;
        a.lsum[idx] = l;
    }
    const double e = (mine && a.eweight) ? a.eweight[k0 + kk] : 0.0;
    const double gm = mine ? g : 0.0;
    double v[WB] = {gm, e * gm, (e * e) * gm, 0.0, 0.0, 0.0, 0.0, 0.0};
    wave_reduce_batch(v, 0xFFu);
    if ((tid & 63) == 0)
#pragma unroll
        for (int i = 0; i < 3; ++i) red[(tid >> 6) * RED_VALUES + i] = v[i];
    __syncthreads();
    if (tid < 3)
        a.cpart[(((int64_t)c * a.S + s) * a.zb + z) * 3 + tid] = wave_combine(tid, true, red);
}

__global__ __launch_bounds__(UB) void correct_sums_kernel(CorrectArgs a) {
    __shared__ double red[(UB / 64) * RED_VALUES];
    correct_sums_body(a, blockIdx.x, blockIdx.y, blockIdx.z, red);
}

hipError_t launch_gain_correct(const CorrectArgs& a, hipStream_t stream) {
    correct_sums_kernel<<<dim3(a.C, a.S, a.zb), UB, 0, stream>>>(a);
    return hipGetLastError();
}

// Grouped slot sums (serving): member g owns blocks [start[g], start[g+1]), one per
// (cell, species, z) of its own grid in the order of correct_sums_kernel's grid.
__global__ __launch_bounds__(UB) void correct_sums_group_kernel(const CorrectArgs* table,
                                                                const int32_t* start, int count) {
    __shared__ double red[(UB / 64) * RED_VALUES];
    const int g = group_member(start, count, blockIdx.x);
    const CorrectArgs a = table[g];
    const int local = blockIdx.x - start[g];
    const int z = local % a.zb, s = (local / a.zb) % a.S, c = local / (a.zb * a.S);
    correct_sums_body(a, c, s, z, red);
}

hipError_t launch_gain_correct_group(const CorrectArgs* table, const int32_t* start, int count,
                                     int total_blocks, hipStream_t stream) {
    if (count <= 0 || total_blocks <= 0) return hipSuccess;
    correct_sums_group_kernel<<<total_blocks, UB, 0, stream>>>(table, start, count);
    return hipGetLastError();
}

// Grouped coefficient launch: member g owns blocks [start[g], start[g+1]) of UB / 64 waves,
// one wave per (cell, species) as in correct_coef_kernel (members without a coefficient array
// own no blocks).
__global__ __launch_bounds__(UB) void correct_coef_group_kernel(const UpdateArgs* table,
                                                                const int32_t* start, int count) {
    const int g = group_member(start, count, blockIdx.x);
    const UpdateArgs& a = table[g];
    const int64_t p = (int64_t)(blockIdx.x - start[g]) * (UB / 64) + (threadIdx.x >> 6);
    if (!a.coef || p >= (int64_t)a.C * a.S) return;  // the whole wave
    correction_coef(a, (int)(p / a.S), (int)(p % a.S), a.coef + p * 2);
}

hipError_t launch_coef_group(const UpdateArgs* table, const int32_t* start, int count,
                             int total_blocks, hipStream_t stream) {
    if (count <= 0 || total_blocks <= 0) return hipSuccess;
    correct_coef_group_kernel<<<total_blocks, UB, 0, stream>>>(table, start, count);
    return hipGetLastError();
}

// ------------------------------------------ ForwardEuler with the correction fused in -----
// One workgroup of FUSED_THREADS per (cell c, species s), for handles whose species all start
// on a segment boundary (a multiple of UPDATE_STATES) and hold at most FUSED_SEGMENTS segments,
// with four threads per state in the slot sums.  The workgroup is FUSED_GROUPS groups of UB
// threads; pass p gives group G the species' segment z = FUSED_GROUPS p + G, with the thread
// layout of both launches it replaces (state z 64 + t / 4, thread q = t % 4).
//   phase 1: each state's gain and loss sums (correct_sums_body's order) and the segment's
//            gain-moment partials (into LDS instead of cpart);
//   phase 2: wave 0 solves the species' (alpha, beta) from the same partials and the
//            moment-rate item sums (correction_coef's order);
//   phase 3: each segment's update as euler_update_body evaluates it, with the sums from
//            registers and its block partials written per segment.
// Every result is bitwise that of correct_sums_kernel + euler_update_kernel.  The state's own
// loads and its flux row are issued at the start, so they are in flight with the slot loads,
// and the row's gathers right after the slot sums.  The chains of a workgroup do not overlap
// those of the next round on its CU, so the runtime takes this form for grids of at most one
// workgroup per CU (at transport_1d's 128 cells x 2 species it replaced 25.8 us of three
// launches by 22 us; at 2048 workgroups it was slower than them).
constexpr int FUSED_THREADS = 1024;
constexpr int FUSED_GROUPS = FUSED_THREADS / UB;
constexpr int FUSED_PASSES = 4;
constexpr int FUSED_SEGMENTS = FUSED_GROUPS * FUSED_PASSES;  // 1024 states per species

struct StateLoads {
    double fk, iap, dinj, dE;
    EllLane ell;
    double x[ELL_LANE];  // gathered f of the row's first slots (ell_gather)
};

__device__ inline void state_loads(const UpdateArgs& a, int64_t r, int q, bool mine,
                                   StateLoads& L) {
    L.fk = L.iap = L.dinj = L.dE = 0.0;
    if (mine && q == 0) {
        L.fk = a.f[r];
        L.iap = a.inv_ap[r];
        L.dinj = a.df_inj[r];
        if (a.emission) L.dE = a.demi[r];
    }
    ell_issue(a, r, q, mine, 0, L.ell);
}

__global__ __launch_bounds__(FUSED_THREADS) void euler_update_fused_kernel(UpdateArgs a,
                                                                           CorrectArgs ca) {
    __shared__ double red[(FUSED_THREADS / 64) * RED_VALUES];
    __shared__ double cpl[FUSED_SEGMENTS * 3];
    __shared__ double coefl[2];
    __shared__ int ired[FUSED_THREADS / 64];
    const int c = blockIdx.x, s = blockIdx.y;
    const int tid = threadIdx.x, G = tid >> 8, t = tid & (UB - 1), q = t & 3, N = a.N;
    double* rg = red + G * (UB / 64) * RED_VALUES;
    const int k0s = a.species_off[s], ns = a.species_off[s + 1] - k0s;
    const int nz = (ns + UPDATE_STATES - 1) / UPDATE_STATES;
    const int seg0 = k0s / UPDATE_STATES, nseg = (N + UPDATE_STATES - 1) / UPDATE_STATES;
    if (a.fin_rec && c == 0 && s == 0 && tid == 0) {
        double* rec = a.fin_rec + (int64_t)a.fin_slot * FIN_WORDS;
        rec[0] = (double)a.step_index;
        rec[1] = a.t_start;
        rec[2] = a.dt;
        rec[3] = (double)a.save;
        rec[4] = (double)a.verbose;
    }
    // wave 0: the moment-rate item sums, issued before the slot loads
    double cv[WB] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (tid < 64) mom_rate_sums(a, c, s, tid, cv);
    // the first pass's state loads and flux row
    StateLoads ld;
    {
        const int kk = G * UPDATE_STATES + (t >> 2);
        const bool mine = G < nz && kk < ns;
        state_loads(a, (int64_t)c * N + k0s + kk, q, mine, ld);
    }
    // phase 1: slot sums and gain-moment partials per segment
    const int gsl = ca.gslots[s], nsl = ca.nslots[s];
    double gv[FUSED_PASSES], lv[FUSED_PASSES];
#pragma unroll
    for (int p = 0; p < FUSED_PASSES; ++p) {
        gv[p] = lv[p] = 0.0;
        if (p * FUSED_GROUPS >= nz) break;  // the whole workgroup
        const int z = p * FUSED_GROUPS + G;
        const bool live = z < nz;
        const int kk = z * UPDATE_STATES + (t >> 2);
        const bool valid = live && kk < ns;
        const int64_t idx = (int64_t)c * N + k0s + (valid ? kk : 0);
        double g = 0.0, l = 0.0;
        if (live) {
            g = slot_sum(ca, idx, q, 0, gsl, ca.gain_f32 != 0);
            l = slot_sum(ca, idx, q, gsl, nsl, false);
        }
#pragma unroll
        for (int m = 1; m < 4; m <<= 1) {
            g = g + xor_lane_rt(g, m);
            l = l + xor_lane_rt(l, m);
        }
        gv[p] = g;
        lv[p] = l;
        // the first pass's flux gathers: in flight across the barriers up to phase 3
        if (p == 0) ell_gather(a, ld.ell, ld.x);
        const bool mine = valid && q == 0;
        const double e = (mine && ca.eweight) ? ca.eweight[k0s + kk] : 0.0;
        const double gm = mine ? g : 0.0;
        double v[WB] = {gm, e * gm, (e * e) * gm, 0.0, 0.0, 0.0, 0.0, 0.0};
        wave_reduce_batch(v, 0xFFu);
        if ((t & 63) == 0)
#pragma unroll
            for (int i = 0; i < 3; ++i) rg[(t >> 6) * RED_VALUES + i] = v[i];
        __syncthreads();
        if (live && t < 3) cpl[z * 3 + t] = wave_combine(t, true, rg);
        __syncthreads();
    }
    // phase 2: (alpha, beta) of the species
    if (tid < 64) {
        cpart_sums(cpl, nz, tid, cv);
        coef_solve(a, cv, tid, coefl);
    }
    const bool need_cell = a.binary && a.active;
    const int any_nz = need_cell ? cell_nonzero<FUSED_THREADS>(a, c, ired) : 1;
    if (!need_cell) __syncthreads();  // coefl (cell_nonzero ends with barriers)
    const bool active = a.binary && (a.active == nullptr || a.active[c] || !any_nz);
    const bool want_cr = a.verbose >= 1 && a.verbose <= 3;
    const double cf0 = coefl[0], cf1 = coefl[1];
    // phase 3: the update of each segment
#pragma unroll
    for (int p = 0; p < FUSED_PASSES; ++p) {
        if (p * FUSED_GROUPS >= nz) break;  // the whole workgroup
        const int z = p * FUSED_GROUPS + G;
        const bool live = z < nz;
        const int kk = z * UPDATE_STATES + (t >> 2);
        const bool mine = live && kk < ns;
        const int k = k0s + kk;
        const int64_t r = (int64_t)c * N + k;
        if (p > 0) {
            state_loads(a, r, q, mine, ld);
            ell_gather(a, ld.ell, ld.x);
        }
        double sumdf = 0.0, sumf = 0.0, cr = INFINITY, crb = INFINITY, cre = INFINITY,
               crf = INFINITY, fnew = 0.0;
        double dB = 0.0;
        if (a.binary) {
            double g = 0.0;
            if (mine && active && q == 0) {
                const double e = a.eweight ? a.eweight[k] : 0.0;
                g = gv[p] * (1.0 + (cf0 + cf1 * e)) + lv[p];
            }
            g = g + xor_lane<1>(g);
            g = g + xor_lane<2>(g);
            dB = g * ((mine && active) ? a.vol[c] : 0.0);
        }
        double dF = ell_sum_x(ld.ell, ld.x, 0, a.ell_width, 0.0);
        for (int s0 = 4 * ELL_LANE; s0 < a.ell_width; s0 += 4 * ELL_LANE) {  // wider rows
            EllLane more;
            ell_issue(a, r, q, mine, s0, more);
            dF = ell_sum(a, more, s0, dF);
        }
        if (mine && q == 0) {
            const double fk = ld.fk;
            sumf = fk;
            const double dE = ld.dE;
            double df = -dF;
            if (a.emission) df = df + dE;
            if (a.binary) df = df + dB;
            df = df * (ld.iap * a.s);
            sumdf = df;
            if (want_cr) {
                cr = cr_min(cr, df / fk);
                if (a.verbose == 3) {
                    if (a.binary) crb = cr_min(crb, dB / fk);
                    if (a.emission) cre = cr_min(cre, dE / fk);
                    crf = cr_min(crf, (-dF) / fk);
                }
            }
            double x = fk + (df + ld.dinj * a.s);
            x = strong_zero(x, x >= THR_1F_10);
            x = strong_zero(x, x != INFINITY);
            a.f_out[r] = x;
            fnew = x;
            publish(a, c, k, x);
        }
        const StateMeta meta = state_meta(a, k, mine && q == 0);
        finish_block(a, c, seg0 + (live ? z : 0), nseg, mine && q == 0, meta.sp, meta.ew, fnew, rg,
                     sumdf, sumf, cr, crb, cre, crf, INFINITY, INFINITY, t, live);
        __syncthreads();  // rg is reused by the next pass
    }
}

bool fused_update_fits(int N, int S, const int32_t* species_off, int tps) {
    if (tps != 4) return false;
    for (int s = 0; s < S; ++s) {
        if (species_off[s] % UPDATE_STATES != 0) return false;
        const int nz = (species_off[s + 1] - species_off[s] + UPDATE_STATES - 1) / UPDATE_STATES;
        if (nz > FUSED_SEGMENTS) return false;
    }
    return species_off[S] == N;
}

hipError_t launch_euler_update_fused(const UpdateArgs& a, const CorrectArgs& ca, bool finalize,
                                     hipStream_t stream) {
    euler_update_fused_kernel<<<dim3(a.C, a.S), FUSED_THREADS, 0, stream>>>(a, ca);
    hipError_t e = hipGetLastError();
    return (e != hipSuccess || !finalize) ? e : launch_finalize(a, stream);
}

__global__ void publish_state_kernel(UpdateArgs a) {
    const int64_t total = (int64_t)a.N * a.C;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x)
        publish(a, (int)(i / a.N), (int)(i % a.N), a.f[i]);
}

hipError_t launch_publish_state(const double* f, int N, int C, int ldc, const int32_t* pad_index,
                                double* fT, float* fT32, uint16_t* fhi, uint16_t* flo,
                                hipStream_t stream) {
    UpdateArgs a{};
    a.f = f;
    a.N = N;
    a.C = C;
    a.ldc = ldc;
    a.ldb = ldc / 16;
    a.pad_index = pad_index;
    a.fT = fT;
    a.fT32 = fT32;
    a.fhi = fhi;
    a.flo = flo;
    const int64_t total = (int64_t)N * C;
    const int blocks = (int)std::min<int64_t>((total + 255) / 256, 8192);
    publish_state_kernel<<<blocks, 256, 0, stream>>>(a);
    return hipGetLastError();
}

__global__ void fill_zero_kernel(double* p, int64_t count) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

static dim3 update_grid(const UpdateArgs& a) {
    return dim3(a.C, (a.N + UPDATE_STATES - 1) / UPDATE_STATES);
}

bool finalize_is_single(const UpdateArgs& a) {
    return fold_groups(a.C, a.S, (int)update_grid(a).y) <= FINALIZE_SINGLE_GROUPS;
}

hipError_t launch_finalize(const UpdateArgs& a, hipStream_t stream) {
    const int nseg = (int)update_grid(a).y, ng = fold_groups(a.C, a.S, nseg);
    if (ng <= FINALIZE_SINGLE_GROUPS) {
        finalize_single_kernel<<<1, UB, 0, stream>>>(a, nseg);
        return hipGetLastError();
    }
    finalize_fold_kernel<<<ng, UB, 0, stream>>>(a, nseg);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    finalize_status_kernel<<<1, UB, 0, stream>>>(a, nseg, ng);
    return hipGetLastError();
}

static hipError_t launch_coef(const UpdateArgs& a, hipStream_t stream) {
    if (!a.gsum || !a.coef) return hipSuccess;
    const int64_t pairs = (int64_t)a.C * a.S;
    correct_coef_kernel<<<(unsigned)((pairs + UB / 64 - 1) / (UB / 64)), UB, 0, stream>>>(a);
    return hipGetLastError();
}

hipError_t launch_euler_update(const UpdateArgs& a, hipStream_t stream) {
    hipError_t e0 = launch_coef(a, stream);
    if (e0 != hipSuccess) return e0;
    euler_update_kernel<<<update_grid(a), UB, 0, stream>>>(a);
    hipError_t e = hipGetLastError();
    return e != hipSuccess ? e : launch_finalize(a, stream);
}

hipError_t launch_euler_update_only(const UpdateArgs& a, hipStream_t stream) {
    hipError_t e0 = launch_coef(a, stream);
    if (e0 != hipSuccess) return e0;
    euler_update_kernel<<<update_grid(a), UB, 0, stream>>>(a);
    return hipGetLastError();
}

hipError_t launch_finalize_batch(const UpdateArgs& base, int64_t bp_words, int first, int count,
                                 hipStream_t stream) {
    if (count <= 0) return hipSuccess;
    finalize_batch_kernel<<<count, UB, 0, stream>>>(base, (int)update_grid(base).y, bp_words,
                                                    first, count);
    return hipGetLastError();
}

hipError_t launch_symplectic_momentum(const UpdateArgs& a, hipStream_t stream) {
    hipError_t e0 = launch_coef(a, stream);
    if (e0 != hipSuccess) return e0;
    symplectic_momentum_kernel<<<update_grid(a), UB, 0, stream>>>(a);
    return hipGetLastError();
}

hipError_t launch_symplectic_space(const UpdateArgs& a, hipStream_t stream) {
    symplectic_space_kernel<<<update_grid(a), UB, 0, stream>>>(a);
    hipError_t e = hipGetLastError();
    return e != hipSuccess ? e : launch_finalize(a, stream);
}

hipError_t launch_euler_update_group(const UpdateArgs* table, const int32_t* block_start,
                                    int count, int total_blocks, hipStream_t stream) {
    if (count <= 0) return hipSuccess;
    euler_update_group_kernel<<<total_blocks, UB, 0, stream>>>(table, block_start, count);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    finalize_group_kernel<<<count, UB, 0, stream>>>(table);
    return hipGetLastError();
}

hipError_t launch_fill_zero(double* p, int64_t count, hipStream_t stream) {
    if (count <= 0) return hipSuccess;
    int64_t blocks = (count + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    fill_zero_kernel<<<(int)blocks, 256, 0, stream>>>(p, count);
    return hipGetLastError();
}

}  // namespace dtx
