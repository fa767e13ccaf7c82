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

namespace dtx {

__device__ constexpr double THR_1F_10 = (double)1e-10f;  // Julia Float32 literal 1f-10
__device__ constexpr double THR_1F_20 = (double)1e-20f;  // Julia Float32 literal 1f-20

constexpr int UB = 256;  // threads per update block

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

// Block reductions: a butterfly over the 64 lanes of each wave (__shfl_xor), then the four wave
// results combined in wave order.  The pattern is fixed, so every reduction is deterministic.
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) v = v + __shfl_xor(v, m, 64);
    return v;
}

__device__ inline double wave_min(double v) {
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) v = cr_min(v, __shfl_xor(v, m, 64));
    return v;
}

__device__ double block_sum(double v, double* red) {
    v = wave_sum(v);
    const int tid = threadIdx.x;
    if ((tid & 63) == 0) red[tid >> 6] = v;
    __syncthreads();
    double r = red[0];
    for (int w = 1; w < UB / 64; ++w) r = r + red[w];
    __syncthreads();
    return r;
}

__device__ double block_min(double v, double* red) {
    v = wave_min(v);
    const int tid = threadIdx.x;
    if ((tid & 63) == 0) red[tid >> 6] = v;
    __syncthreads();
    double r = red[0];
    for (int w = 1; w < UB / 64; ++w) r = cr_min(r, red[w]);
    __syncthreads();
    return r;
}

__device__ int block_or(int v, int* red) {
    const int tid = threadIdx.x;
    v = __any(v) ? 1 : 0;
    if ((tid & 63) == 0) red[tid >> 6] = v;
    __syncthreads();
    int r = 0;
    for (int w = 0; w < UB / 64; ++w) r |= red[w];
    __syncthreads();
    return r;
}

// SpMV row: sum over ELL slots in ascending column order, separate multiply and add.
__device__ inline double ell_row(const int32_t* col, const double* val, int width, int64_t n,
                                 int64_t r, const double* x) {
    double acc = 0.0;
    for (int s = 0; s < width; ++s) {
        const int32_t c = col[(int64_t)s * n + r];
        if (c < 0) break;
        acc = acc + val[(int64_t)s * n + r] * x[c];
    }
    return acc;
}

// Binary term of state k of cell c (unscaled by invAp): the contraction's partial slots (gain
// channels and the FP64 loss channels) times Vol.  The four threads q = 0..3 of the state sum
// slots q, q+4, ...; the four sums are combined as (s0 + s1) + (s2 + s3) by a butterfly, so every
// lane holds the same, order-fixed result.
__device__ inline double binary_term(const UpdateArgs& a, int c, int k, int q, bool valid) {
    const int64_t idx = (int64_t)c * a.N + k;
    double g = 0.0;
    if (valid) {
        const int ns = a.nslot[k];
        int s = q;
        for (; s + 12 < ns; s += 16) {  // four loads in flight per thread
            double v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = a.part[(int64_t)(s + 4 * u) * a.part_stride + idx];
#pragma unroll
            for (int u = 0; u < 4; ++u) g = g + v[u];
        }
        for (; s < ns; s += 4) g = g + a.part[(int64_t)s * a.part_stride + idx];
    }
    g = g + __shfl_xor(g, 1, 64);
    g = g + __shfl_xor(g, 2, 64);
    return g * (valid ? a.vol[c] : 0.0);
}

// Copies of a new state value for the next contraction: FP64 state-major (padded states), and
// the FP32 copy or the BF16 hi/lo split when that precision is in use.
__device__ inline void publish(const UpdateArgs& a, int c, int k, double x) {
    const int kp = a.pad_index[k];
    a.fT[(int64_t)kp * a.ldc + c] = x;
    if (a.fT32) a.fT32[(int64_t)kp * a.ldc + c] = (float)x;
    if (a.fhi) {
        uint16_t hi, lo;
        split_bf16(x, &hi, &lo);
        a.fhi[(int64_t)c * a.ldb + kp] = hi;
        a.flo[(int64_t)c * a.ldb + kp] = lo;
    }
}

__device__ inline double emission_term(const UpdateArgs& a, int c, int k, const double* fc) {
    const double* E = a.emi + (int64_t)c * a.N * a.N;
    double acc = 0.0;
    for (int j = 0; j < a.N; ++j) acc = acc + E[k + (int64_t)a.N * j] * fc[j];
    return acc;
}

// Per-species moment partials of this block's new states (fn, LDS, indexed by local state) and
// the per-block partials (sums and Cr minima).  finalize_step_kernel folds them in block order.
__device__ void finish_block(const UpdateArgs& a, int c, int seg, int nseg, const double* fn,
                             int k0, int k1, double* red, double sumdf, double sumf, double cr,
                             double crb, double cre, double crf, double crm, double crs) {
    const int tid = threadIdx.x;
    const int64_t b = (int64_t)c * nseg + seg;
    for (int s = 0; s < a.S; ++s) {
        const int lo = max(k0, a.species_off[s]), hi = min(k1, a.species_off[s + 1]);
        double num = 0.0, en = 0.0;
        if (lo < hi) {
            const int k = k0 + (tid >> 2);
            if ((tid & 3) == 0 && k >= lo && k < hi) {
                num = fn[tid >> 2];
                if (a.eweight) en = fn[tid >> 2] * a.eweight[k];
            }
            num = block_sum(num, red);
            en = block_sum(en, red);
        }
        if (tid == 0) {
            a.mom_part[(b * a.S + s) * 2 + 0] = num;
            a.mom_part[(b * a.S + s) * 2 + 1] = en;
        }
    }
    sumdf = block_sum(sumdf, red);
    sumf = block_sum(sumf, red);
    if (a.verbose >= 1 && a.verbose <= 3) {
        cr = block_min(cr, red);
        if (a.verbose == 3) {
            crb = block_min(crb, red);
            cre = block_min(cre, red);
            crf = block_min(crf, red);
            crm = block_min(crm, red);
            crs = block_min(crs, red);
        }
    }
    if (tid == 0) {
        double* bp = a.block_part + b * BP_WORDS;
        bp[BP_SUMDF] = sumdf;
        bp[BP_SUMF] = sumf;
        bp[BP_CR] = cr;
        bp[BP_CR_BIN] = crb;
        bp[BP_CR_EMI] = cre;
        bp[BP_CR_FLUX] = crf;
        bp[BP_CR_MOM] = crm;
        bp[BP_CR_SPACE] = crs;
    }
}

// One workgroup: folds the per-block partials of the step (block order) into the moments of
// every (cell, species) and the status record of the step.
__device__ void finalize_step_body(const UpdateArgs& a, int nseg, double* red) {
    const int tid = threadIdx.x;
    const int64_t nblocks = (int64_t)a.C * nseg;
    for (int64_t q = tid; q < (int64_t)a.C * a.S; q += UB) {
        const int64_t cc = q / a.S, sp = q % a.S;
        double num = 0.0, en = 0.0;
        for (int g = 0; g < nseg; ++g) {
            const double* mp = a.mom_part + (((cc * nseg + g) * a.S) + sp) * 2;
            num = num + mp[0];
            en = en + mp[1];
        }
        a.moments[q * 2 + 0] = num;
        a.moments[q * 2 + 1] = en;
    }
    double v[BP_WORDS];
    for (int q = 0; q < BP_WORDS; ++q) v[q] = (q <= BP_SUMF) ? 0.0 : INFINITY;
    for (int64_t bb = tid; bb < nblocks; bb += UB) {
        const double* bp = a.block_part + bb * BP_WORDS;
        v[BP_SUMDF] = v[BP_SUMDF] + bp[BP_SUMDF];
        v[BP_SUMF] = v[BP_SUMF] + bp[BP_SUMF];
        for (int q = BP_CR; q < BP_WORDS; ++q) v[q] = cr_min(v[q], bp[q]);
    }
    v[BP_SUMDF] = block_sum(v[BP_SUMDF], red);
    v[BP_SUMF] = block_sum(v[BP_SUMF], red);
    for (int q = BP_CR; q < BP_WORDS; ++q) v[q] = block_min(v[q], red);
    if (tid == 0) {
        double* st = a.status + (int64_t)(a.step_index % a.ring) * STATUS_WORDS;
        const bool cr_eval = (a.verbose >= 1 && a.verbose <= 3) && (v[BP_SUMF] != 0.0);
        st[ST_STEP] = (double)a.step_index;
        st[ST_FINITE] = isfinite(v[BP_SUMDF]) ? 1.0 : 0.0;
        st[ST_CR_EVAL] = cr_eval ? 1.0 : 0.0;
        st[ST_CR] = cr_eval ? -v[BP_CR] : 0.0;
        st[ST_CR_BIN] = (cr_eval && a.verbose == 3 && a.binary) ? -v[BP_CR_BIN] : 0.0;
        st[ST_CR_EMI] = (cr_eval && a.verbose == 3 && a.emission) ? -v[BP_CR_EMI] : 0.0;
        st[ST_CR_FLUX] = (cr_eval && a.verbose == 3) ? -v[BP_CR_FLUX] : 0.0;
        st[ST_CR_MOM] = (cr_eval && a.verbose == 3) ? -v[BP_CR_MOM] : 0.0;
        st[ST_CR_SPACE] = (cr_eval && a.verbose == 3) ? -v[BP_CR_SPACE] : 0.0;
        st[ST_SUMF] = v[BP_SUMF];
        st[ST_DT] = a.dt;
        st[ST_TSTART] = a.t_start;
        st[ST_SAVE] = (double)a.save;
        st[ST_SUMDF] = v[BP_SUMDF];
    }
}

// Load the whole cell into LDS (needed by the emission GEMV and the Domain all-zero test) and
// return whether any entry is non-zero.
__device__ int load_cell(const UpdateArgs& a, int c, double* fc, int* ired) {
    const int64_t base = (int64_t)c * a.N;
    int nz = 0;
    for (int k = threadIdx.x; k < a.N; k += UB) {
        const double v = a.f[base + k];
        fc[k] = v;
        nz |= (v != 0.0);
    }
    return block_or(nz, ired);  // also orders the fc writes before use
}

// ----------------------------------------------------------------------- ForwardEuler ------
__device__ void euler_update_body(const UpdateArgs& a, int c, int seg, int nseg, double* smem,
                                  double* red, int* ired) {
    double* fn = smem;                    // new f of this block's states
    double* fc = smem + UPDATE_STATES;    // f of the whole cell (emission / Domain only)
    const int tid = threadIdx.x, N = a.N;
    const int k0 = seg * UPDATE_STATES, k1 = min(N, k0 + UPDATE_STATES);
    const int64_t base = (int64_t)c * N;
    const bool need_cell = a.emission || (a.binary && a.active);
    const int any_nz = need_cell ? load_cell(a, c, fc, ired) : 1;
    // update_Big_Bin! evaluates a cell when Domain is nothing, the cell is in Domain, or its f
    // is all zero (src/SteppingMethods.jl:369); otherwise df_Bin keeps its zero initial value.
    const bool active = a.binary && (a.active == nullptr || a.active[c] || !any_nz);

    double sumdf = 0.0, sumf = 0.0, cr = INFINITY, crb = INFINITY, cre = INFINITY,
           crf = INFINITY;
    const bool want_cr = a.verbose >= 1 && a.verbose <= 3;
    const int k = k0 + (tid >> 2), q = tid & 3;
    const bool mine = k < k1;
    const double dB = a.binary ? binary_term(a, c, k, q, mine && active) : 0.0;
    if (mine && q == 0) {
        const int64_t r = base + k;
        const double fk = a.f[r];
        sumf = fk;
        const double dF = a.ell_width ? ell_row(a.ell_col, a.ell_val, a.ell_width, a.n, r, a.f)
                                      : 0.0;
        double df = -dF;
        double dE = 0.0;
        if (a.emission) {
            dE = emission_term(a, c, k, fc);
            df = df + dE;
        }
        if (a.binary) df = df + dB;
        df = df * (a.inv_ap[r] * a.s);
        sumdf = df;
        if (want_cr) {
            cr = cr_min(cr, df / fk);
            if (a.verbose == 3) {
                if (a.binary) crb = cr_min(crb, dB / fk);
                if (a.emission) cre = cr_min(cre, dE / fk);
                crf = cr_min(crf, (-dF) / fk);
            }
        }
        double x = fk + (df + a.df_inj[r] * a.s);
        x = strong_zero(x, x >= THR_1F_10);
        x = strong_zero(x, x != INFINITY);
        a.f_out[r] = x;
        fn[tid >> 2] = x;
        publish(a, c, k, x);
    }
    __syncthreads();
    finish_block(a, c, seg, nseg, fn, k0, k1, red, sumdf, sumf, cr, crb, cre, crf, INFINITY,
                 INFINITY);
}

__global__ __launch_bounds__(UB) void euler_update_kernel(UpdateArgs a) {
    extern __shared__ double smem[];
    __shared__ double red[UB / 64];
    __shared__ int ired[UB / 64];
    euler_update_body(a, blockIdx.x, blockIdx.y, gridDim.y, smem, red, ired);
}

__global__ __launch_bounds__(UB) void finalize_step_kernel(UpdateArgs a, int nseg) {
    __shared__ double red[UB / 64];
    finalize_step_body(a, nseg, red);
}

// Grouped forms for several simulations in one launch (serving): block b of the update grid
// belongs to the simulation g with block_start[g] <= b < block_start[g+1]; the finalize grid has
// one block per simulation.
__global__ __launch_bounds__(UB) void euler_update_group_kernel(const UpdateArgs* table,
                                                                const int32_t* block_start,
                                                                int count) {
    extern __shared__ double smem[];
    __shared__ double red[UB / 64];
    __shared__ int ired[UB / 64];
    const int b = blockIdx.x;
    int lo = 0, hi = count - 1;
    while (lo < hi) {  // last g with block_start[g] <= b
        const int mid = (lo + hi + 1) >> 1;
        if (block_start[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    const UpdateArgs a = table[lo];
    const int nseg = (a.N + UPDATE_STATES - 1) / UPDATE_STATES;
    const int local = b - block_start[lo];
    euler_update_body(a, local / nseg, local % nseg, nseg, smem, red, ired);
}

__global__ __launch_bounds__(UB) void finalize_group_kernel(const UpdateArgs* table) {
    __shared__ double red[UB / 64];
    const UpdateArgs a = table[blockIdx.x];
    finalize_step_body(a, (a.N + UPDATE_STATES - 1) / UPDATE_STATES, red);
}

// ------------------------------------------------------------- ForwardSymplecticEuler ------
// Momentum half (src/SteppingMethods.jl:137-160): P flux, emission and binary terms from f,
// f_Momentum = f + df_Momentum clamped at 1f-10.
__global__ __launch_bounds__(UB) void symplectic_momentum_kernel(UpdateArgs a) {
    extern __shared__ double smem[];
    double* fc = smem + UPDATE_STATES;
    __shared__ int ired[UB / 64];
    const int c = blockIdx.x, tid = threadIdx.x, N = a.N;
    const int k0 = blockIdx.y * UPDATE_STATES, k1 = min(N, k0 + UPDATE_STATES);
    const int64_t base = (int64_t)c * N;
    const bool need_cell = a.emission || (a.binary && a.active);
    const int any_nz = need_cell ? load_cell(a, c, fc, ired) : 1;
    const bool active = a.binary && (a.active == nullptr || a.active[c] || !any_nz);
    const int k = k0 + (tid >> 2), q = tid & 3;
    const bool mine = k < k1;
    const double dB = a.binary ? binary_term(a, c, k, q, mine && active) : 0.0;
    if (mine && q == 0) {
        const int64_t r = base + k;
        const double dP = a.ell_width ? ell_row(a.ell_col, a.ell_val, a.ell_width, a.n, r, a.f)
                                      : 0.0;
        double dm = -dP;
        if (a.emission) {
            const double dE = emission_term(a, c, k, fc);
            dm = dm + dE;
            a.demi[r] = dE;
        }
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
    extern __shared__ double smem[];
    double* fn = smem;
    __shared__ double red[UB / 64];
    const int c = blockIdx.x, tid = threadIdx.x, N = a.N;
    const int k0 = blockIdx.y * UPDATE_STATES, k1 = min(N, k0 + UPDATE_STATES);
    const int64_t base = (int64_t)c * N;
    double sumf = 0.0, cr = INFINITY, crb = INFINITY, cre = INFINITY, crf = INFINITY;
    double crm = INFINITY, crs = INFINITY;
    const bool want_cr = a.verbose >= 1 && a.verbose <= 3;
    const int k = k0 + (tid >> 2);
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
        fn[tid >> 2] = x;
        publish(a, c, k, x);
    }
    __syncthreads();
    finish_block(a, c, blockIdx.y, gridDim.y, fn, k0, k1, red, 0.0, sumf, cr, crb, cre, crf, crm,
                 crs);
}

__global__ void publish_state_kernel(UpdateArgs a) {
    const int64_t total = (int64_t)a.N * a.C;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x)
        publish(a, (int)(i / a.N), (int)(i % a.N), a.f[i]);
}

hipError_t launch_publish_state(const double* f, int N, int C, int ldc, const int32_t* pad_index,
                                int Np, double* fT, float* fT32, uint16_t* fhi, uint16_t* flo,
                                hipStream_t stream) {
    UpdateArgs a{};
    a.f = f;
    a.N = N;
    a.C = C;
    a.ldc = ldc;
    a.ldb = Np;
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

static size_t update_smem(const UpdateArgs& a) {
    return (size_t)(UPDATE_STATES + a.N) * sizeof(double);
}

static dim3 update_grid(const UpdateArgs& a) {
    return dim3(a.C, (a.N + UPDATE_STATES - 1) / UPDATE_STATES);
}

static hipError_t launch_finalize(const UpdateArgs& a, hipStream_t stream) {
    finalize_step_kernel<<<1, UB, 0, stream>>>(a, (int)update_grid(a).y);
    return hipGetLastError();
}

hipError_t launch_euler_update(const UpdateArgs& a, hipStream_t stream) {
    euler_update_kernel<<<update_grid(a), UB, update_smem(a), stream>>>(a);
    hipError_t e = hipGetLastError();
    return e != hipSuccess ? e : launch_finalize(a, stream);
}

hipError_t launch_symplectic_momentum(const UpdateArgs& a, hipStream_t stream) {
    symplectic_momentum_kernel<<<update_grid(a), UB, update_smem(a), stream>>>(a);
    return hipGetLastError();
}

hipError_t launch_symplectic_space(const UpdateArgs& a, hipStream_t stream) {
    symplectic_space_kernel<<<update_grid(a), UB, update_smem(a), stream>>>(a);
    hipError_t e = hipGetLastError();
    return e != hipSuccess ? e : launch_finalize(a, stream);
}

hipError_t launch_euler_update_group(const UpdateArgs* table, const int32_t* block_start,
                                    int count, int total_blocks, int max_N, hipStream_t stream) {
    if (count <= 0) return hipSuccess;
    euler_update_group_kernel<<<total_blocks, UB, (size_t)(UPDATE_STATES + max_N) * 8, stream>>>(
        table, block_start, count);
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
