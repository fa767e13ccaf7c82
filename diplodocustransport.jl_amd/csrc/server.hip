// server.hip -- request loop with continuous batching (dtx_server_* of include/dtx.h).
//
// A scheduler thread owns the device work.  Requests wait in a FIFO; the head is admitted when
// a slot is free and its HBM footprint fits the budget next to the running requests.  All
// running requests advance together, one ForwardEuler step each per iteration, through the
// grouped contraction and update launches (runtime.hip); a request that reaches its step count
// is retired at once (final f and moments copied out, memory released) and the next waiting
// requests take its place in the following iteration.  Every request keeps its own work
// decomposition, so its result is bit-identical whether it runs alone or in a mix.
//
// Requests are synthetic workloads built from (shape, seed): species 0 electron, 1 photon,
// 2.. positron; log p grid on [1e-3, 1e3], uniform u, phi on [0, 2 pi]; Maxwell-Juttner f0 at
// theta with 1% seeded noise; lognormal gain tensors generated on the device; x advection plus
// synchrotron cooling when with_flux and x_num > 1, built on the device; dt = 0.1 / (fastest
// expected loss rate).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "dtx_internal.h"

using namespace dtx;

namespace {

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct SpeciesGrid {
    double mass, charge;
    int p, u, h;
    std::vector<double> pr, ur, hr;
};

SpeciesGrid species_grid(const dtx_config& c, int s) {
    SpeciesGrid g;
    g.mass = (s == 1) ? 0.0 : 1.0;
    g.charge = (s == 1) ? 0.0 : (s == 0 ? -1.0 : 1.0);
    g.p = c.px_num[s];
    g.u = c.py_num[s];
    g.h = c.pz_num[s];
    for (int i = 0; i <= g.p; ++i) g.pr.push_back(std::pow(10.0, -3.0 + 6.0 * i / g.p));
    for (int i = 0; i <= g.u; ++i) g.ur.push_back(-1.0 + 2.0 * i / g.u);
    for (int i = 0; i <= g.h; ++i) g.hr.push_back(2.0 * M_PI * i / g.h);
    return g;
}

double h_uniform(uint64_t key, uint64_t idx) {
    return ((double)(splitmix_key(key, idx) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

double h_normal(uint64_t key, uint64_t idx) {
    const double u1 = h_uniform(key, 2 * idx), u2 = h_uniform(key, 2 * idx + 1);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
}

// Maxwell-Juttner bin integrals (10-sample trapezium in p, the u and phi factors are exact),
// renormalised to one particle; p fastest, then u, then phi.
std::vector<double> maxwell_juttner(const SpeciesGrid& g, double theta) {
    std::vector<double> fp(g.p);
    for (int i = 0; i < g.p; ++i) {
        double acc = 0.0;
        for (int q = 0; q < 10; ++q) {
            const double x = g.pr[i] + (g.pr[i + 1] - g.pr[i]) * q / 9.0;
            const double E = std::sqrt(g.mass * g.mass + x * x);
            const double dis = (g.mass != 0.0) ? std::exp((g.mass - E) / theta) : std::exp(-x / theta);
            acc += ((q == 0 || q == 9) ? 0.5 : 1.0) * dis * x * x;
        }
        fp[i] = acc * (g.pr[i + 1] - g.pr[i]) / 9.0;
    }
    std::vector<double> f((size_t)g.p * g.u * g.h);
    double tot = 0.0;
    for (int hh = 0; hh < g.h; ++hh)
        for (int uu = 0; uu < g.u; ++uu)
            for (int i = 0; i < g.p; ++i) {
                const double v = fp[i] * (g.ur[uu + 1] - g.ur[uu]) * (g.hr[hh + 1] - g.hr[hh]);
                f[(size_t)i + (size_t)g.p * (uu + (size_t)g.u * hh)] = v;
                tot += v;
            }
    for (double& v : f) v /= tot;
    return f;
}

// Synchrotron coefficient K that gives the fastest-cooling bin a cooling time of about 100
// steps: the largest outflow coefficient I(p0) / (dp du) of the massive charged species at K = 1.
double sync_coefficient(const std::vector<SpeciesGrid>& sg, double dt) {
    double cmax = 0.0;
    for (const SpeciesGrid& g : sg) {
        if (g.mass == 0.0 || g.charge == 0.0) continue;
        for (int u = 0; u < g.u; ++u)
            for (int p = 1; p < g.p; ++p) {
                const double u0 = g.ur[u], u1 = g.ur[u + 1];
                const double ang = (3 * u1 - u1 * u1 * u1) - (3 * u0 - u0 * u0 * u0);
                const double pp = g.pr[p];
                const double I = 1.0 / (3 * g.mass) * pp * std::sqrt(g.mass * g.mass + pp * pp) * ang;
                cmax = std::max(cmax, I / ((g.pr[p + 1] - g.pr[p]) * (u1 - u0)));
            }
    }
    return cmax > 0.0 ? 1.0 / (100.0 * dt * cmax) : 0.0;
}

// Periodic upwind x advection and synchrotron cooling in p (closed boundaries), F = X + P, with
// 1/Ap and Vol, built on the device (dtx_synth_flux): x cells of width 1 / x_num on [0, 1].
int set_flux(dtx_handle* h, const dtx_config& c, const std::vector<SpeciesGrid>& sg, double dt) {
    dtx_flux_synth fs{};
    std::vector<double> xr(c.x_num + 1), yr(c.y_num + 1), zr(c.z_num + 1), pr, ur, hr;
    for (int i = 0; i <= c.x_num; ++i) xr[i] = (double)i / c.x_num;
    for (int i = 0; i <= c.y_num; ++i) yr[i] = (double)i / c.y_num;
    for (int i = 0; i <= c.z_num; ++i) zr[i] = (double)i / c.z_num;
    for (size_t s = 0; s < sg.size(); ++s) {
        pr.insert(pr.end(), sg[s].pr.begin(), sg[s].pr.end());
        ur.insert(ur.end(), sg[s].ur.begin(), sg[s].ur.end());
        hr.insert(hr.end(), sg[s].hr.begin(), sg[s].hr.end());
        fs.mass[s] = sg[s].mass;
        fs.charge[s] = sg[s].charge;
    }
    fs.xr = xr.data();
    fs.yr = yr.data();
    fs.zr = zr.data();
    fs.pr = pr.data();
    fs.ur = ur.data();
    fs.hr = hr.data();
    fs.sync_K = sync_coefficient(sg, dt);
    fs.axes = 1;
    fs.periodic = 1;
    return dtx_synth_flux(h, &fs);
}

struct Active {
    int64_t id;
    dtx_handle* h = nullptr;
    int64_t steps = 0, done = 0;
    double dt = 0.0;
    int64_t bytes = 0;
};

struct Record {
    dtx_request req;
    dtx_request_result res;
    std::vector<double> f;
};

}  // namespace

struct dtx_server {
    dtx_server_config cfg;
    Group* group = nullptr;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<int64_t> pending;
    std::map<int64_t, Record> records;
    std::vector<Active> active;
    int64_t next_id = 0;
    int64_t used_bytes = 0;
    bool stop = false;
    int64_t unfinished = 0;
    std::thread worker;
    std::string err;
};

static int64_t div_up(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Device footprint of a request: packed gain channels in the contraction precision, FP64 loss
// channels, partial-sum slots (bounded), state copies.
static int64_t estimate_bytes(const dtx_config& c, int precision) {
    const int S = c.num_species;
    std::vector<int64_t> ns(S);
    int64_t N = 0, Np = 0;
    for (int s = 0; s < S; ++s) {
        ns[s] = (int64_t)c.px_num[s] * c.py_num[s] * c.pz_num[s];
        N += ns[s];
        Np += div_up(ns[s], SPECIES_ALIGN) * SPECIES_ALIGN;
    }
    const int64_t C = (int64_t)c.x_num * c.y_num * c.z_num;
    const int64_t eb = precision == DTX_PREC_FP64 ? 8 : 4;  // BF16X3 stores hi + lo
    int64_t gain = 0;
    for (int a = 0; a < S; ++a)
        for (int b = a; b < S; ++b) {
            int64_t chunks = 0;
            if (a == b)
                for (int64_t i = 0; i < ns[a]; ++i) chunks += div_up(ns[a] - i / KC * KC, KC);
            else
                chunks = ns[a] * div_up(ns[b], KC);
            const int nout = (a == b) ? 1 : 2;
            const int outs[2] = {a, b};
            for (int q = 0; q < nout; ++q) gain += div_up(ns[outs[q]], TK) * chunks * KC * TK * eb;
        }
    int64_t loss = 0;
    for (int s = 0; s < S; ++s) loss += div_up(ns[s], TK) * (Np / KC) * KC * TK * 8;
    const int64_t part = 160 * C * N * 8;
    const int64_t state = C * N * 8 * 8 + Np * div_up(C, 16) * 16 * 12;
    return gain + loss + part + state + ((int64_t)N * N * 8);
}

static int build_request(dtx_server* sv, const dtx_request& rq, dtx_handle** out, double* dt_out) {
    dtx_config c = rq.cfg;
    c.precision = sv->cfg.precision;
    c.method = DTX_METHOD_FORWARD_EULER;
    c.device = sv->cfg.device;
    if (!(c.dt0 > 0.0)) c.dt0 = 1.0;
    dtx_handle* h = nullptr;
    int rc = dtx_create(&c, &h);
    if (rc) {
        sv->err = dtx_last_error(nullptr);
        return rc;
    }
    // grouped steps add the contraction's slots directly (no per-request correction launch)
    rc = dtx_set_moment_correction(h, 0);
    if (rc) {
        sv->err = dtx_last_error(h);
        dtx_destroy(h);
        return rc;
    }
    const int S = c.num_species;
    std::vector<SpeciesGrid> sg;
    std::vector<int> off;
    int N = 0;
    for (int s = 0; s < S; ++s) {
        sg.push_back(species_grid(c, s));
        off.push_back(N);
        N += sg.back().p * sg.back().u * sg.back().h;
    }
    const int C = c.x_num * c.y_num * c.z_num;
    const int64_t n = (int64_t)N * C;
    // f0: Maxwell-Juttner per species, 1% noise per state
    std::vector<double> base((size_t)N);
    for (int s = 0; s < S; ++s) {
        std::vector<double> b = maxwell_juttner(sg[s], rq.theta > 0 ? rq.theta : 1.0);
        std::copy(b.begin(), b.end(), base.begin() + off[s]);
    }
    std::vector<double> f0((size_t)n), ap((size_t)n), vol((size_t)C), inj((size_t)n, 0.0);
    const uint64_t nkey = splitmix_key(rq.seed, 1000);
    for (int64_t i = 0; i < n; ++i)
        f0[(size_t)i] = base[(size_t)(i % N)] * (1.0 + 0.01 * h_normal(nkey, (uint64_t)i));
    const double dx = 1.0 / c.x_num, dy = 1.0 / c.y_num, dz = 1.0 / c.z_num;
    for (int cc = 0; cc < C; ++cc) {
        vol[(size_t)cc] = c.dt0 * dx * dy * dz;
        for (int k = 0; k < N; ++k) ap[(size_t)cc * N + k] = dx * dy * dz;
    }
    // dt = 0.1 / fastest loss rate with the expected loss matrix L[i,j] = No e^(1/2)
    double worst = 0.0;
    for (int cc = 0; cc < C; ++cc) {
        std::vector<double> tot(S, 0.0);
        for (int s = 0; s < S; ++s)
            for (int k = 0; k < sg[s].p * sg[s].u * sg[s].h; ++k)
                tot[s] += f0[(size_t)cc * N + off[s] + k];
        for (int a = 0; a < S; ++a) {
            double rate = 0.0;
            const double na = sg[a].p * sg[a].u * sg[a].h;
            for (int x = 0; x < S; ++x)
                for (int y = x; y < S; ++y) {
                    if (a == x && x == y) rate += na * std::exp(0.5) * tot[x];
                    else if (a == x) rate += na * std::exp(0.5) * tot[y];
                    else if (a == y) rate += na * std::exp(0.5) * tot[x];
                }
            worst = std::max(worst, rate * vol[(size_t)cc] / ap[(size_t)cc * N]);
        }
    }
    const double dt = 0.1 / worst;
    {
        const bool flux = rq.with_flux && c.x_num > 1;
        if (flux) {
            rc = set_flux(h, c, sg, dt);
            if (rc) goto fail;
        }
        rc = dtx_synth_binary(h, rq.seed, 1.0);
        if (rc) goto fail;
        // with the flux built on the device, 1/Ap and Vol are already there
        rc = dtx_set_vectors(h, flux ? nullptr : ap.data(), flux ? nullptr : vol.data(),
                             f0.data(), inj.data());
    }
    if (rc) goto fail;
    {
        std::vector<double> w((size_t)N);
        for (int s = 0; s < S; ++s) {
            const SpeciesGrid& g = sg[s];
            for (int k = 0; k < g.p * g.u * g.h; ++k) {
                const int p = k % g.p;
                const double p0 = g.pr[p], p1 = g.pr[p + 1], m = g.mass;
                w[(size_t)off[s] + k] =
                    (m != 0.0) ? (p0 * std::sqrt(m * m + p0 * p0) - p1 * std::sqrt(m * m + p1 * p1) +
                                  m * m * (std::asinh(p0 / m) - std::asinh(p1 / m))) /
                                     (2 * (p0 - p1))
                               : (p0 + p1) / 2;
            }
        }
        rc = dtx_set_moment_weights(h, w.data());
        if (rc) goto fail;
    }
    *out = h;
    *dt_out = dt;
    return DTX_OK;
fail:
    sv->err = dtx_last_error(h);
    dtx_destroy(h);
    return rc;
}

static void retire(dtx_server* sv, Active& a, int status) {
    Record rec;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        rec = sv->records[a.id];
    }
    Record& r = rec;
    r.res.steps_done = a.done;
    if (status == 2 && a.h) {
        const int64_t n = (int64_t)r.req.cfg.x_num * r.req.cfg.y_num * r.req.cfg.z_num;
        int64_t N = 0;
        for (int s = 0; s < r.req.cfg.num_species; ++s)
            N += (int64_t)r.req.cfg.px_num[s] * r.req.cfg.py_num[s] * r.req.cfg.pz_num[s];
        r.f.assign((size_t)(n * N), 0.0);
        std::vector<double> num((size_t)(n * r.req.cfg.num_species)), en(num.size());
        if (dtx_get_f(a.h, r.f.data()) || dtx_moments(a.h, num.data(), en.data())) status = -3;
        double tn = 0, te = 0;
        for (size_t i = 0; i < num.size(); ++i) {
            tn += num[i];
            te += en[i];
        }
        r.res.number_total = tn;
        r.res.energy_total = te;
    }
    if (a.h) dtx_destroy(a.h);
    r.res.status = status;
    r.res.finish_ms = now_ms();
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->records[a.id] = std::move(r);
    sv->used_bytes -= a.bytes;
    sv->unfinished--;
    sv->cv.notify_all();
}

static void scheduler(dtx_server* sv) {
    (void)hipSetDevice(sv->cfg.device);
    std::vector<double> dts;
    std::vector<dtx_handle*> hs;
    bool changed = true;
    while (true) {
        // admission: FIFO head, while a slot and the HBM budget allow
        while (true) {
            int64_t id = -1;
            dtx_request rq;
            {
                std::unique_lock<std::mutex> lk(sv->mu);
                if (sv->stop) return;
                if (sv->pending.empty() || (int)sv->active.size() >= sv->cfg.max_active) {
                    if (sv->active.empty())
                        sv->cv.wait_for(lk, std::chrono::milliseconds(2));
                    break;
                }
                id = sv->pending.front();
                rq = sv->records[id].req;
                const int64_t est = estimate_bytes(rq.cfg, sv->cfg.precision);
                if (sv->used_bytes + est > sv->cfg.hbm_budget && !sv->active.empty()) break;
                sv->pending.pop_front();
                sv->used_bytes += est;
                sv->records[id].res.status = 1;
                sv->records[id].res.start_ms = now_ms();
            }
            Active a;
            a.id = id;
            a.steps = rq.steps;
            a.bytes = estimate_bytes(rq.cfg, sv->cfg.precision);
            if (build_request(sv, rq, &a.h, &a.dt) != DTX_OK) {
                a.h = nullptr;
                retire(sv, a, -1);
                continue;
            }
            sv->active.push_back(a);
            changed = true;
        }
        if (sv->active.empty()) continue;
        if (changed) {
            hs.clear();
            dts.clear();
            for (auto& a : sv->active) {
                hs.push_back(a.h);
                dts.push_back(a.dt);
            }
            if (group_set(sv->group, hs.data(), (int)hs.size()) != DTX_OK) {
                sv->err = group_error(sv->group);
                for (auto& a : sv->active) retire(sv, a, -2);
                sv->active.clear();
                continue;
            }
            changed = false;
        }
        if (group_step(sv->group, dts.data(), 0) != DTX_OK) {
            sv->err = group_error(sv->group);
            for (auto& a : sv->active) retire(sv, a, -2);
            sv->active.clear();
            changed = true;
            continue;
        }
        bool finished = false;
        for (auto& a : sv->active) finished |= (++a.done >= a.steps);
        if (finished) {
            (void)group_sync(sv->group);
            std::vector<Active> keep;
            for (auto& a : sv->active) {
                if (a.done >= a.steps) retire(sv, a, 2);
                else keep.push_back(a);
            }
            sv->active.swap(keep);
            changed = true;
        }
    }
}

extern "C" {

int dtx_server_create(const dtx_server_config* cfg, dtx_server** out) {
    if (!cfg || !out) return DTX_ERR_INVALID;
    *out = nullptr;
    if (cfg->max_active < 1 || cfg->hbm_budget <= 0) return DTX_ERR_INVALID;
    dtx_server* sv = new dtx_server();
    sv->cfg = *cfg;
    sv->group = group_create(cfg->device);
    if (!sv->group) {
        delete sv;
        return DTX_ERR_HIP;
    }
    sv->worker = std::thread(scheduler, sv);
    *out = sv;
    return DTX_OK;
}

int dtx_server_submit(dtx_server* sv, const dtx_request* req, int64_t* id) {
    if (!sv || !req || !id) return DTX_ERR_INVALID;
    if (req->cfg.num_species < 1 || req->cfg.num_species > DTX_MAX_SPECIES || req->steps < 1)
        return DTX_ERR_INVALID;
    std::lock_guard<std::mutex> lk(sv->mu);
    const int64_t rid = sv->next_id++;
    Record r;
    r.req = *req;
    memset(&r.res, 0, sizeof r.res);
    r.res.id = rid;
    r.res.status = 0;
    r.res.submit_ms = now_ms();
    sv->records[rid] = r;
    sv->pending.push_back(rid);
    sv->unfinished++;
    sv->cv.notify_all();
    *id = rid;
    return DTX_OK;
}

int dtx_server_poll(dtx_server* sv, int64_t id, dtx_request_result* out) {
    if (!sv || !out) return DTX_ERR_INVALID;
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->records.find(id);
    if (it == sv->records.end()) return DTX_ERR_INVALID;
    *out = it->second.res;
    return DTX_OK;
}

int dtx_server_fetch(dtx_server* sv, int64_t id, double* f, int64_t n) {
    if (!sv || !f) return DTX_ERR_INVALID;
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->records.find(id);
    if (it == sv->records.end() || it->second.res.status != 2) return DTX_ERR_STATE;
    if ((int64_t)it->second.f.size() != n) return DTX_ERR_INVALID;
    std::copy(it->second.f.begin(), it->second.f.end(), f);
    return DTX_OK;
}

int dtx_server_drain(dtx_server* sv) {
    if (!sv) return DTX_ERR_INVALID;
    std::unique_lock<std::mutex> lk(sv->mu);
    sv->cv.wait(lk, [&] { return sv->unfinished == 0; });
    return DTX_OK;
}

void dtx_server_destroy(dtx_server* sv) {
    if (!sv) return;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stop = true;
        sv->cv.notify_all();
    }
    if (sv->worker.joinable()) sv->worker.join();
    (void)group_sync(sv->group);
    for (auto& a : sv->active)
        if (a.h) dtx_destroy(a.h);
    group_destroy(sv->group);
    delete sv;
}

const char* dtx_server_last_error(const dtx_server* sv) { return sv ? sv->err.c_str() : ""; }

}  // extern "C"
