// Collision-matrix ingest: one binary interaction's gain and loss arrays (as DiplodocusCollisions
// stores them) added to a handle's binary operator, the terms Fill_M_Bin! writes into M_Bin
// (src/Collisions/BuildBinaryMatrices.jl:81-284), without forming the N^2 x N array.
//
// Every M_Bin entry M[k, i, j] (row (i-1) N + k, column j) that a gain or loss entry produces goes
// where dtx_set_binary_csc puts a stored entry: to the FP64 loss matrix at (k, the other index)
// when k == i or k == j, else to the pair (min(i, j), max(i, j)) of channel (species of k; species
// of the pair), row k.  Entries are summed in the reference's loop order for each k.
//
// Host code only.  The work is split over threads by the px of the output state (k), so each
// thread owns its accumulators and the sums do not depend on the thread count.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <string>
#include <thread>
#include <vector>

#include "dtx_internal.h"

using namespace dtx;

namespace {

// Grid of one species: px, py, pz counts, first state in a cell, u and phi widths.
struct SpeciesGrid {
    int X, Y, Z;
    int64_t off;
    const double* dpy;
    const double* dpz;
    int64_t local(int px, int py, int pz) const { return px + (int64_t)X * (py + (int64_t)Y * pz); }
    int64_t size() const { return (int64_t)X * Y * Z; }
};

// Sequential left-to-right sum (Julia's sum(dpy) of a short vector).
double vsum(const double* v, int n) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s = s + v[i];
    return s;
}

// Where an M_Bin entry M[k, i, j] lands (k, i, j global states of one cell, 0-based).
struct Target {
    BinaryHostView ch;  // channel (species of k; species of min(i,j), of max(i,j))
    int64_t off_out, off_lo, off_hi;
    double* loss;
    int64_t N;
    bool* loss_used;  // set when a loss entry was written (one flag per thread)

    void add(int64_t k, int64_t i, int64_t j, double v) const {
        if (k == i || k == j) {
            const int64_t other = (k == i) ? j : i;
            loss[k + N * other] = loss[k + N * other] + v;
            *loss_used = true;
            return;
        }
        const int64_t lo = std::min(i, j), hi = std::max(i, j);
        const int a = (int)(lo - off_lo), b = (int)(hi - off_hi);
        const int64_t p = ch.tri ? (int64_t)a * ch.N1 - (int64_t)a * (a - 1) / 2 + (b - a)
                                 : (int64_t)a * ch.N2 + b;
        double& dst = ch.S[p * ch.No + (k - off_out)];
        dst = dst + v;
    }
};

// Runs body(x0, x1, used) over [0, X) split in contiguous ranges, one per thread.
template <typename F>
void split_px(int X, bool* any_loss, F body) {
    const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
    const int T = std::max(1, std::min({X, hw, 16}));
    std::vector<char> used((size_t)T, 0);
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t) {
        const int x0 = (int)((int64_t)X * t / T), x1 = (int)((int64_t)X * (t + 1) / T);
        th.emplace_back([&, t, x0, x1] {
            bool u = false;
            body(x0, x1, &u);
            used[(size_t)t] = u;
        });
    }
    for (auto& x : th) x.join();
    for (char u : used) *any_loss = *any_loss || u;
}

// GainMatrix_to_M_Bin! (BuildBinaryMatrices.jl:114-204) for output species g3 and incoming g1,
// g2: val * w / 2 at M[a, b, c] and M[a, c, b].
void gain_terms(const double* G, double scale, const SpeciesGrid& g3, const SpeciesGrid& g1,
                const SpeciesGrid& g2, int mode, bool sparse, const Target& proto,
                bool* any_loss) {
    const int64_t s1 = g3.X + 2, s2 = s1 * g3.Y, s3 = s2 * g3.Z;  // strides of py3, pz3, px1
    const int64_t s4 = s3 * g1.X, s5 = s4 * g1.Y, s6 = s5 * g1.Z;  // py1, pz1, px2
    const int64_t s7 = s6 * g2.X, s8 = s7 * g2.Y;                  // py2, pz2
    auto at = [&](int px3, int py3, int pz3, int px1, int py1, int pz1, int px2, int py2,
                  int pz2) {  // the scaled entry for bin px3 (the array's bin px3 + 1)
        return G[(px3 + 1) + s1 * py3 + s2 * pz3 + s3 * px1 + s4 * py1 + s5 * pz1 + s6 * px2 +
                 s7 * py2 + s8 * pz2] * scale;
    };
    const double w_iso = 1.0 / (vsum(g1.dpz, g1.Z) * vsum(g2.dpz, g2.Z) * vsum(g3.dpz, g3.Z) *
                                vsum(g1.dpy, g1.Y) * vsum(g2.dpy, g2.Y) * vsum(g3.dpy, g3.Y));
    const double w_axi = 1.0 / (vsum(g1.dpz, g1.Z) * vsum(g2.dpz, g2.Z) * vsum(g3.dpz, g3.Z));
    split_px(g3.X, any_loss, [&](int x0, int x1, bool* used) {
        Target tg = proto;
        tg.loss_used = used;
        auto push = [&](int px3, int py3, int pz3, int px1, int py1, int pz1, int px2, int py2,
                        int pz2, double v) {
            const int64_t a = g3.off + g3.local(px3, py3, pz3);
            const int64_t b = g1.off + g1.local(px1, py1, pz1);
            const int64_t c = g2.off + g2.local(px2, py2, pz2);
            tg.add(a, b, c, v);
            tg.add(a, c, b, v);
        };
        std::vector<double> iso_val, axi_val;
        for (int px2 = 0; px2 < g2.X; ++px2)
            for (int px1 = 0; px1 < g1.X; ++px1) {
                if (mode == DTX_BIN_ISO) {  // average over u and phi of all three
                    iso_val.assign((size_t)(x1 - x0), 0.0);
                    for (int px3 = x0; px3 < x1; ++px3) {
                        double val = 0.0;
                        for (int py1 = 0; py1 < g1.Y; ++py1)
                            for (int pz1 = 0; pz1 < g1.Z; ++pz1)
                                for (int py2 = 0; py2 < g2.Y; ++py2)
                                    for (int pz2 = 0; pz2 < g2.Z; ++pz2)
                                        for (int py3 = 0; py3 < g3.Y; ++py3)
                                            for (int pz3 = 0; pz3 < g3.Z; ++pz3)
                                                val += at(px3, py3, pz3, px1, py1, pz1, px2, py2,
                                                          pz2) * g1.dpz[pz1] * g2.dpz[pz2] *
                                                       g3.dpz[pz3] * g1.dpy[py1] * g2.dpy[py2] *
                                                       g3.dpy[py3];
                        iso_val[(size_t)(px3 - x0)] = val;
                    }
                }
                for (int py1 = 0; py1 < g1.Y; ++py1)
                    for (int py2 = 0; py2 < g2.Y; ++py2)
                        for (int py3 = 0; py3 < g3.Y; ++py3) {
                            if (mode == DTX_BIN_AXI) {  // average over phi of all three
                                axi_val.assign((size_t)(x1 - x0), 0.0);
                                for (int px3 = x0; px3 < x1; ++px3) {
                                    double val = 0.0;
                                    for (int pz1 = 0; pz1 < g1.Z; ++pz1)
                                        for (int pz2 = 0; pz2 < g2.Z; ++pz2)
                                            for (int pz3 = 0; pz3 < g3.Z; ++pz3)
                                                val += at(px3, py3, pz3, px1, py1, pz1, px2, py2,
                                                          pz2) * g1.dpz[pz1] * g2.dpz[pz2] *
                                                       g3.dpz[pz3];
                                    axi_val[(size_t)(px3 - x0)] = val;
                                }
                            }
                            for (int pz1 = 0; pz1 < g1.Z; ++pz1)
                                for (int pz2 = 0; pz2 < g2.Z; ++pz2)
                                    for (int pz3 = 0; pz3 < g3.Z; ++pz3)
                                        for (int px3 = x0; px3 < x1; ++px3) {
                                            double val, w;
                                            if (mode == DTX_BIN_ISO) {
                                                val = iso_val[(size_t)(px3 - x0)];
                                                w = w_iso;
                                            } else if (mode == DTX_BIN_AXI) {
                                                val = axi_val[(size_t)(px3 - x0)];
                                                w = w_axi;
                                            } else {
                                                val = at(px3, py3, pz3, px1, py1, pz1, px2, py2,
                                                         pz2);
                                                w = 1.0;
                                                if (sparse) {  // largest entry over the px3 bins
                                                    double gmax = -INFINITY;
                                                    for (int q = -1; q <= g3.X; ++q)
                                                        gmax = std::max(
                                                            gmax, at(q, py3, pz3, px1, py1, pz1,
                                                                     px2, py2, pz2));
                                                    if (val < gmax * 1e-16) continue;
                                                }
                                            }
                                            if (val == 0.0) continue;
                                            push(px3, py3, pz3, px1, py1, pz1, px2, py2, pz2,
                                                 val * w / 2);
                                        }
                        }
            }
    });
}

// LossMatrix_to_M_Bin! (BuildBinaryMatrices.jl:206-284) for loss species g1 against g2:
// -val * w / 2 at M[a, a, c] and M[a, c, a].
void loss_terms(const double* L, double scale, const SpeciesGrid& g1, const SpeciesGrid& g2,
                int mode, const Target& proto, bool* any_loss) {
    const int64_t s1 = g1.X, s2 = s1 * g1.Y, s3 = s2 * g1.Z;  // py1, pz1, px2
    const int64_t s4 = s3 * g2.X, s5 = s4 * g2.Y;              // py2, pz2
    auto at = [&](int px1, int py1, int pz1, int px2, int py2, int pz2) {
        return L[px1 + s1 * py1 + s2 * pz1 + s3 * px2 + s4 * py2 + s5 * pz2] * scale;
    };
    const double w_iso = 1.0 / (vsum(g1.dpz, g1.Z) * vsum(g2.dpz, g2.Z) * vsum(g1.dpy, g1.Y) *
                                vsum(g2.dpy, g2.Y));
    const double w_axi = 1.0 / (vsum(g1.dpz, g1.Z) * vsum(g2.dpz, g2.Z));
    split_px(g1.X, any_loss, [&](int x0, int x1, bool* used) {
        Target tg = proto;
        tg.loss_used = used;
        for (int px2 = 0; px2 < g2.X; ++px2)
            for (int px1 = x0; px1 < x1; ++px1) {
                double iso = 0.0;
                if (mode == DTX_BIN_ISO)
                    for (int py2 = 0; py2 < g2.Y; ++py2)
                        for (int py1 = 0; py1 < g1.Y; ++py1)
                            for (int pz2 = 0; pz2 < g2.Z; ++pz2)
                                for (int pz1 = 0; pz1 < g1.Z; ++pz1)
                                    iso += at(px1, py1, pz1, px2, py2, pz2) * g1.dpz[pz1] *
                                           g2.dpz[pz2] * g1.dpy[py1] * g2.dpy[py2];
                for (int py2 = 0; py2 < g2.Y; ++py2)
                    for (int py1 = 0; py1 < g1.Y; ++py1) {
                        double axi = 0.0;
                        if (mode == DTX_BIN_AXI)
                            for (int pz2 = 0; pz2 < g2.Z; ++pz2)
                                for (int pz1 = 0; pz1 < g1.Z; ++pz1)
                                    axi += at(px1, py1, pz1, px2, py2, pz2) * g1.dpz[pz1] *
                                           g2.dpz[pz2];
                        for (int pz2 = 0; pz2 < g2.Z; ++pz2)
                            for (int pz1 = 0; pz1 < g1.Z; ++pz1) {
                                double val, w;
                                if (mode == DTX_BIN_ISO) {
                                    val = iso;
                                    w = w_iso;
                                } else if (mode == DTX_BIN_AXI) {
                                    val = axi;
                                    w = w_axi;
                                } else {
                                    val = at(px1, py1, pz1, px2, py2, pz2);
                                    w = 1.0;
                                }
                                if (val == 0.0) continue;
                                const int64_t a = g1.off + g1.local(px1, py1, pz1);
                                const int64_t c = g2.off + g2.local(px2, py2, pz2);
                                tg.add(a, a, c, -(val * w / 2));
                                tg.add(a, c, a, -(val * w / 2));
                            }
                    }
            }
    });
}

}  // namespace

extern "C" int dtx_add_binary_interaction(dtx_handle* h, const dtx_binary_interaction* x) {
    if (!h) return DTX_ERR_INVALID;
    if (!x) return handle_fail(h, DTX_ERR_INVALID, "dtx_add_binary_interaction: null descriptor");
    const dtx_config& cfg = *handle_config(h);
    const int S = cfg.num_species;
    for (int q = 0; q < 4; ++q)
        if (x->species[q] < 0 || x->species[q] >= S)
            return handle_fail(h, DTX_ERR_INVALID,
                               "dtx_add_binary_interaction: species index out of range");
    if (x->mode != DTX_BIN_ANI && x->mode != DTX_BIN_AXI && x->mode != DTX_BIN_ISO)
        return handle_fail(h, DTX_ERR_INVALID, "dtx_add_binary_interaction: unknown mode");
    if (!x->gain3 || !x->gain4 || !x->loss1 || !x->loss2 || !x->dpy || !x->dpz)
        return handle_fail(h, DTX_ERR_INVALID, "dtx_add_binary_interaction: null array");
    int rc = binary_host_begin(h, "dtx_add_binary_interaction");
    if (rc) return rc;
    const int32_t* off = handle_species_offsets(h);
    std::vector<SpeciesGrid> g((size_t)S);
    int64_t oy = 0, oz = 0;
    for (int s = 0; s < S; ++s) {
        g[(size_t)s] = SpeciesGrid{cfg.px_num[s], cfg.py_num[s], cfg.pz_num[s], off[s],
                                   x->dpy + oy, x->dpz + oz};
        oy += cfg.py_num[s];
        oz += cfg.pz_num[s];
    }
    const int n1 = x->species[0], n2 = x->species[1], n3 = x->species[2], n4 = x->species[3];
    const int lo = std::min(n1, n2), hi = std::max(n1, n2);
    auto target = [&](int s_out) {  // s_out < 0: loss terms only (no channel)
        Target t;
        t.ch = s_out >= 0 ? binary_host_channel(h, s_out, lo, hi) : BinaryHostView{};
        t.off_out = s_out >= 0 ? off[s_out] : 0;
        t.off_lo = off[lo];
        t.off_hi = off[hi];
        t.loss = binary_host_loss(h);
        t.N = off[S];
        t.loss_used = nullptr;
        return t;
    };
    bool any_loss = false;
    const bool sparse = x->sparse != 0;
    // Fill_M_Bin! (BuildBinaryMatrices.jl:104-107): both gains, then both losses
    gain_terms(x->gain3, x->gain_scale, g[(size_t)n3], g[(size_t)n1], g[(size_t)n2], x->mode,
               sparse, target(n3), &any_loss);
    gain_terms(x->gain4, x->gain_scale, g[(size_t)n4], g[(size_t)n1], g[(size_t)n2], x->mode,
               sparse, target(n4), &any_loss);
    // loss entries have k == i or k == j: they reach the loss matrix only
    loss_terms(x->loss1, x->loss_scale, g[(size_t)n1], g[(size_t)n2], x->mode, target(-1),
               &any_loss);
    loss_terms(x->loss2, x->loss_scale, g[(size_t)n2], g[(size_t)n1], x->mode, target(-1),
               &any_loss);
    binary_host_end(h, any_loss);
    return DTX_OK;
}
