// kernels_update.hip -- fused per-cell update kernels for gfx950 (HBM-bound).
//
// One workgroup per spatial cell.  For every state of the cell it evaluates, in FP64 and with
// the reference's operation order (this file is built with -ffp-contract=off, so no FMA):
//   ForwardEuler, src/SteppingMethods.jl:31-101
//     dF  = F f                           (ELL rows, columns ascending = CSC column order)
//     df  = -dF ; df = df + dE ; df = df + dB
//     df  = df * (invAp * s)
//     f   = f + (df + df_Inj * s) ; f = f*(f >= 1f-10) ; f = f*(f != Inf)   (Bool strong zero)
//   plus Cr / Cr_Bin / Cr_Emi / Cr_Flux, isfinite(sum(df)), sum(f) and the per-(cell, species)
//   number and energy moments of the new f.
// dE is the cell's emission block times f (ascending column order) and dB the binary term:
// the sum of the contraction's partial slots plus f_k * (Loss f)_k, times Vol.
// Cross-cell reductions go to per-block partials that the last block folds in block order, so
// every reduction is deterministic.
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

__device__ double block_sum(double v, double* red) {
    const int tid = threadIdx.x;
    red[tid] = v;
    __syncthreads();
    for (int s = UB / 2; s > 0; s >>= 1) {
        if (tid < s) red[tid] = red[tid] + red[tid + s];
        __syncthreads();
    }
    double r = red[0];
    __syncthreads();
    return r;
}

__device__ double block_min(double v, double* red) {
    const int tid = threadIdx.x;
    red[tid] = v;
    __syncthreads();
    for (int s = UB / 2; s > 0; s >>= 1) {
        if (tid < s) red[tid] = cr_min(red[tid], red[tid + s]);
        __syncthreads();
    }
    double r = red[0];
    __syncthreads();
    return r;
}

__device__ int block_or(int v, int* red) {
    const int tid = threadIdx.x;
    red[tid] = v;
    __syncthreads();
    for (int s = UB / 2; s > 0; s >>= 1) {
        if (tid < s) red[tid] |= red[tid + s];
        __syncthreads();
    }
    int r = red[0];
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

// Binary term of state k of cell c (unscaled by invAp): (gain partials + f_k (Loss f)_k) * Vol.
__device__ inline double binary_term(const UpdateArgs& a, int c, int k, const double* fc) {
    const int64_t idx = (int64_t)c * a.N + k;
    double g = 0.0;
    const int ns = a.nslot[k];
    for (int s = 0; s < ns; ++s) g = g + a.part[(int64_t)s * a.part_stride + idx];
    if (a.loss) {
        double l = 0.0;
        for (int j = 0; j < a.N; ++j) l = l + a.loss[k + (int64_t)a.N * j] * fc[j];
        g = g + fc[k] * l;
    }
    return g * a.vol[c];
}

__device__ inline double emission_term(const UpdateArgs& a, int c, int k, const double* fc) {
    const double* E = a.emi + (int64_t)c * a.N * a.N;
    double acc = 0.0;
    for (int j = 0; j < a.N; ++j) acc = acc + E[k + (int64_t)a.N * j] * fc[j];
    return acc;
}

// Per-species moments of the new state (in LDS) and the per-block partials; the last block to
// finish folds all partials (block order) into the status record of this step.
__device__ void finish_block(const UpdateArgs& a, const double* fn, double* red,
                             double sumdf, double sumf, double cr, double crb, double cre,
                             double crf, double crm, double crs) {
    const int c = blockIdx.x, tid = threadIdx.x;
    for (int s = 0; s < a.S; ++s) {
        const int k0 = a.species_off[s], k1 = a.species_off[s + 1];
        double num = 0.0, en = 0.0;
        for (int k = k0 + tid; k < k1; k += UB) {
            num = num + fn[k];
            if (a.eweight) en = en + fn[k] * a.eweight[k];
        }
        num = block_sum(num, red);
        en = block_sum(en, red);
        if (tid == 0) {
            a.moments[((int64_t)c * a.S + s) * 2 + 0] = num;
            a.moments[((int64_t)c * a.S + s) * 2 + 1] = en;
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
    __shared__ int is_last;
    if (tid == 0) {
        double* bp = a.block_part + (int64_t)c * BP_WORDS;
        bp[BP_SUMDF] = sumdf;
        bp[BP_SUMF] = sumf;
        bp[BP_CR] = cr;
        bp[BP_CR_BIN] = crb;
        bp[BP_CR_EMI] = cre;
        bp[BP_CR_FLUX] = crf;
        bp[BP_CR_MOM] = crm;
        bp[BP_CR_SPACE] = crs;
        __threadfence();
        const unsigned prev = atomicAdd(a.counter, 1u);
        is_last = (prev == gridDim.x - 1);
    }
    __syncthreads();
    if (!is_last) return;
    __threadfence();
    double v[BP_WORDS];
    for (int q = 0; q < BP_WORDS; ++q) v[q] = (q <= BP_SUMF) ? 0.0 : INFINITY;
    for (int b = tid; b < (int)gridDim.x; b += UB) {
        const volatile double* bp = a.block_part + (int64_t)b * BP_WORDS;
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
        *a.counter = 0u;  // ready for the next launch
    }
}

// ----------------------------------------------------------------------- ForwardEuler ------
__global__ __launch_bounds__(UB) void euler_update_kernel(UpdateArgs a) {
    extern __shared__ double smem[];
    double* fc = smem;            // f of this cell
    double* fn = smem + a.N;      // new f of this cell
    __shared__ double red[UB];
    __shared__ int ired[UB];
    const int c = blockIdx.x, tid = threadIdx.x, N = a.N;
    const int64_t base = (int64_t)c * N;

    double sumf = 0.0;
    int nz = 0;
    for (int k = tid; k < N; k += UB) {
        const double v = a.f[base + k];
        fc[k] = v;
        sumf = sumf + v;
        nz |= (v != 0.0);
    }
    const int any_nz = block_or(nz, ired);  // also orders the fc writes before use
    // update_Big_Bin! evaluates a cell when Domain is nothing, the cell is in Domain, or its f
    // is all zero (src/SteppingMethods.jl:369); otherwise df_Bin keeps its zero initial value.
    const bool active = a.binary && (a.active == nullptr || a.active[c] || !any_nz);

    double sumdf = 0.0, cr = INFINITY, crb = INFINITY, cre = INFINITY, crf = INFINITY;
    const bool want_cr = a.verbose >= 1 && a.verbose <= 3;
    for (int k = tid; k < N; k += UB) {
        const int64_t r = base + k;
        const double fk = fc[k];
        const double dF = a.ell_width ? ell_row(a.ell_col, a.ell_val, a.ell_width, a.n, r, a.f)
                                      : 0.0;
        double df = -dF;
        double dE = 0.0, dB = 0.0;
        if (a.emission) {
            dE = emission_term(a, c, k, fc);
            df = df + dE;
        }
        if (a.binary) {
            dB = active ? binary_term(a, c, k, fc) : 0.0;
            df = df + dB;
        }
        df = df * (a.inv_ap[r] * a.s);
        sumdf = sumdf + df;
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
        fn[k] = x;
        a.fT[(int64_t)k * a.ldc + c] = x;
    }
    __syncthreads();
    finish_block(a, fn, red, sumdf, sumf, cr, crb, cre, crf, INFINITY, INFINITY);
}

// ------------------------------------------------------------- ForwardSymplecticEuler ------
// Momentum half (src/SteppingMethods.jl:137-160): P flux, emission and binary terms from f,
// f_Momentum = f + df_Momentum clamped at 1f-10.
__global__ __launch_bounds__(UB) void symplectic_momentum_kernel(UpdateArgs a) {
    extern __shared__ double smem[];
    double* fc = smem;
    __shared__ int ired[UB];
    const int c = blockIdx.x, tid = threadIdx.x, N = a.N;
    const int64_t base = (int64_t)c * N;
    int nz = 0;
    for (int k = tid; k < N; k += UB) {
        const double v = a.f[base + k];
        fc[k] = v;
        nz |= (v != 0.0);
    }
    const int any_nz = block_or(nz, ired);
    const bool active = a.binary && (a.active == nullptr || a.active[c] || !any_nz);
    for (int k = tid; k < N; k += UB) {
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
            const double dB = active ? binary_term(a, c, k, fc) : 0.0;
            dm = dm + dB;
            a.dbin[r] = dB;
        }
        dm = dm * (a.inv_ap[r] * a.s);
        a.dpflux[r] = dP;
        a.df_mom[r] = dm;
        const double x = fc[k] + dm;
        a.f_mom[r] = strong_zero(x, x >= THR_1F_10);
    }
}

// Space half and final update (src/SteppingMethods.jl:162-237).  The ELL fields hold X_Flux.
__global__ __launch_bounds__(UB) void symplectic_space_kernel(UpdateArgs a) {
    extern __shared__ double smem[];
    double* fn = smem;
    __shared__ double red[UB];
    const int c = blockIdx.x, tid = threadIdx.x, N = a.N;
    const int64_t base = (int64_t)c * N;
    double sumf = 0.0, cr = INFINITY, crb = INFINITY, cre = INFINITY, crf = INFINITY;
    double crm = INFINITY, crs = INFINITY;
    const bool want_cr = a.verbose >= 1 && a.verbose <= 3;
    for (int k = tid; k < N; k += UB) {
        const int64_t r = base + k;
        const double f0 = a.f[r];
        sumf = sumf + f0;
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
                if (rp < crf) crf = rp;
                crm = cr_min(crm, dm / f0);
                crs = cr_min(crs, dS / fm);
            }
        }
        double x = fs + a.df_inj[r] * a.s;
        x = strong_zero(x, x >= THR_1F_20) * jl_sign(x);
        x = strong_zero(x, x != INFINITY);
        a.f_out[r] = x;
        fn[k] = x;
        a.fT[(int64_t)k * a.ldc + c] = x;
    }
    __syncthreads();
    finish_block(a, fn, red, 0.0, sumf, cr, crb, cre, crf, crm, crs);
}

__global__ void fill_zero_kernel(double* p, int64_t count) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

static size_t update_smem(const UpdateArgs& a, int arrays) {
    return (size_t)arrays * a.N * sizeof(double);
}

hipError_t launch_euler_update(const UpdateArgs& a, hipStream_t stream) {
    euler_update_kernel<<<a.C, UB, update_smem(a, 2), stream>>>(a);
    return hipGetLastError();
}

hipError_t launch_symplectic_momentum(const UpdateArgs& a, hipStream_t stream) {
    symplectic_momentum_kernel<<<a.C, UB, update_smem(a, 1), stream>>>(a);
    return hipGetLastError();
}

hipError_t launch_symplectic_space(const UpdateArgs& a, hipStream_t stream) {
    symplectic_space_kernel<<<a.C, UB, update_smem(a, 1), stream>>>(a);
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
