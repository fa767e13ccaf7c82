// server.hip -- request loop with continuous batching (dtx_server_* of include/dtx.h).
//
// Three threads share the work of a server:
//  - the builder takes requests FIFO: the head is admitted when a slot is free and its HBM
//    footprint fits the budget next to the requests already admitted, then its simulation is
//    built (synthetic requests) or prepared (caller-built handles) on its own stream;
//  - the stepping thread owns the device work: all admitted, built requests advance together,
//    one ForwardEuler step each per iteration, through the grouped contraction, moment-rate
//    correction and update launches (runtime.hip).  A request that reaches its step count
//    leaves the group at once; the event recorded after its last step is handed on, and the
//    next built requests take its place in the following iteration.  This thread never waits
//    for the device: membership changes enqueue their tables on the group stream;
//  - the reaper waits for a finished request's event, copies its final f and moments out and
//    releases its memory (hipFree synchronises the device, so it stays off the stepping thread).
// Every request keeps its own work decomposition and arithmetic, so its result is bit-identical
// alone, in any mix, and stepped through dtx_step on the handle dtx_request_create builds.
//
// Synthetic requests are built from (shape, seed): species 0 electron, 1 photon, 2.. positron;
// log p grid on [1e-3, 1e3], uniform u, phi on [0, 2 pi]; Maxwell-Juttner f0 at theta with 1%
// seeded noise; lognormal gain tensors generated on the device; x advection plus synchrotron
// cooling when with_flux and x_num > 1, built on the device; dt = 0.1 / (fastest expected loss
// rate), rounded down to 12 significant bits.
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
    hipEvent_t finished = nullptr;  // recorded on the group stream after the last step
};

struct Record {
    dtx_request req;
    dtx_handle* handle = nullptr;   // caller-built simulation (dtx_server_submit_handle)
    double dt = 0.0;                // its step
    dtx_request_result res;
    std::vector<double> f;
};

}  // namespace

struct dtx_server {
    dtx_server_config cfg;
    Group* group = nullptr;
    std::mutex mu;
    std::condition_variable cv;     // any state change (submit, built, retired, stop)
    std::deque<int64_t> pending;    // submitted, not admitted
    std::deque<Active> built;       // admitted and built, not yet stepping
    std::deque<Active> retiring;    // finished, waiting for the reaper
    std::map<int64_t, Record> records;
    int64_t next_id = 0;
    int64_t used_bytes = 0;         // admitted requests (building, stepping, retiring)
    int n_admitted = 0;             // admitted and not yet retired
    bool stop = false;
    int64_t unfinished = 0;
    std::thread builder, stepper, reaper;
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
    // moment-rate matrices of the reduced-precision modes: two FP64 N1 x N2 per gain channel
    int64_t moment = 0;
    if (precision != DTX_PREC_FP64)
        for (int a = 0; a < S; ++a)
            for (int b = a; b < S; ++b) moment += 2 * 2 * div_up(ns[a], TK) * TK * div_up(ns[b], KC) * KC * 8;
    const int64_t part = 160 * C * N * 8;
    const int64_t state = C * N * 8 * 12 + Np * div_up(C, 16) * 16 * 12;
    return gain + loss + moment + part + state + ((int64_t)N * N * 8);
}

// The simulation of a synthetic request as a handle ready for grouped steps, and its dt.
static int build_request(const dtx_server_config& sc, const dtx_request& rq, dtx_handle** out,
                         double* dt_out, std::string* err) {
    dtx_config c = rq.cfg;
    c.precision = sc.precision;
    c.method = DTX_METHOD_FORWARD_EULER;
    c.device = sc.device;
    if (!(c.dt0 > 0.0)) c.dt0 = 1.0;
    dtx_handle* h = nullptr;
    int rc = dtx_create(&c, &h);
    if (rc) {
        *err = dtx_last_error(nullptr);
        return rc;
    }
    rc = dtx_set_moment_correction(h, sc.no_moment_correction ? 0 : 1);
    if (rc) {
        *err = dtx_last_error(h);
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
    double dt = 0.1 / worst;
    {  // 12 significant bits: k dt is exact, as for the benchmark workloads
        int e2 = 0;
        const double m2 = std::frexp(dt, &e2);
        dt = std::ldexp(std::floor(m2 * 4096.0) / 4096.0, e2);
    }
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
    rc = group_prepare(h);
    if (rc) goto fail;
    *out = h;
    *dt_out = dt;
    return DTX_OK;
fail:
    *err = dtx_last_error(h);
    dtx_destroy(h);
    return rc;
}

// Builder thread: FIFO admission by slot and HBM budget, then build / prepare.
static void builder(dtx_server* sv) {
    (void)hipSetDevice(sv->cfg.device);
    while (true) {
        int64_t id = -1;
        Record rec;
        int64_t est = 0;
        {
            std::unique_lock<std::mutex> lk(sv->mu);
            while (true) {
                if (sv->stop) return;
                if (!sv->pending.empty() && sv->n_admitted < sv->cfg.max_active) {
                    const Record& r = sv->records[sv->pending.front()];
                    est = r.handle ? dtx_device_bytes(r.handle)
                                   : estimate_bytes(r.req.cfg, sv->cfg.precision);
                    // the head waits for the budget unless nothing else holds memory
                    if (sv->used_bytes + est <= sv->cfg.hbm_budget || sv->n_admitted == 0) break;
                }
                sv->cv.wait(lk);
            }
            id = sv->pending.front();
            sv->pending.pop_front();
            sv->used_bytes += est;
            sv->n_admitted++;
            Record& r = sv->records[id];
            r.res.status = 1;
            r.res.start_ms = now_ms();
            rec.req = r.req;
            rec.handle = r.handle;
            rec.dt = r.dt;
        }
        Active a;
        a.id = id;
        a.steps = rec.req.steps;
        a.bytes = est;
        std::string err;
        int rc;
        if (rec.handle) {
            a.h = rec.handle;
            a.dt = rec.dt;
            rc = group_prepare(a.h);
            if (rc) err = dtx_last_error(a.h);
        } else {
            rc = build_request(sv->cfg, rec.req, &a.h, &a.dt, &err);
            if (rc) a.h = nullptr;
        }
        if (!rc && hipEventCreateWithFlags(&a.finished, hipEventDisableTiming) != hipSuccess) {
            rc = DTX_ERR_HIP;
            err = "event creation failed";
        }
        std::lock_guard<std::mutex> lk(sv->mu);
        if (rc) {
            sv->err = err;
            if (a.h) dtx_destroy(a.h);
            Record& r = sv->records[id];
            r.handle = nullptr;
            r.res.status = -1;
            r.res.finish_ms = now_ms();
            sv->used_bytes -= est;
            sv->n_admitted--;
            sv->unfinished--;
        } else {
            sv->built.push_back(a);
        }
        sv->cv.notify_all();
    }
}

// Reaper thread: final f and moments of finished requests, then their memory.
static void reaper(dtx_server* sv) {
    (void)hipSetDevice(sv->cfg.device);
    while (true) {
        Active a;
        {
            std::unique_lock<std::mutex> lk(sv->mu);
            sv->cv.wait(lk, [&] { return sv->stop || !sv->retiring.empty(); });
            if (sv->retiring.empty()) return;  // stop with nothing left
            a = sv->retiring.front();
            sv->retiring.pop_front();
        }
        int status = a.done >= a.steps ? 2 : -2;
        std::vector<double> f;
        double tn = 0, te = 0;
        if (a.h && hipEventSynchronize(a.finished) != hipSuccess) status = -3;
        if (status == 2) {
            const dtx_config& c = *handle_config(a.h);
            int64_t N = 0;
            for (int q = 0; q < c.num_species; ++q)
                N += (int64_t)c.px_num[q] * c.py_num[q] * c.pz_num[q];
            const int64_t C = (int64_t)c.x_num * c.y_num * c.z_num;
            const int64_t n = N * C, cs = C * c.num_species;
            f.assign((size_t)n, 0.0);
            std::vector<double> num((size_t)cs), en((size_t)cs);
            if (dtx_get_f(a.h, f.data()) || dtx_moments(a.h, num.data(), en.data())) status = -3;
            for (size_t i = 0; i < num.size(); ++i) {
                tn += num[i];
                te += en[i];
            }
        }
        if (a.finished) (void)hipEventDestroy(a.finished);
        if (a.h) dtx_destroy(a.h);
        std::lock_guard<std::mutex> lk(sv->mu);
        Record& r = sv->records[a.id];
        r.handle = nullptr;
        r.f = std::move(f);
        r.res.steps_done = a.done;
        r.res.number_total = tn;
        r.res.energy_total = te;
        r.res.status = status;
        r.res.finish_ms = now_ms();
        sv->used_bytes -= a.bytes;
        sv->n_admitted--;
        sv->unfinished--;
        sv->cv.notify_all();
    }
}

// Stepping thread: grouped steps of every built request; never waits for the device.
static void stepper(dtx_server* sv) {
    (void)hipSetDevice(sv->cfg.device);
    std::vector<Active> active;
    std::vector<double> dts;
    std::vector<dtx_handle*> hs;
    bool changed = false;
    auto retire_all = [&](int status_steps_ok) {
        std::lock_guard<std::mutex> lk(sv->mu);
        for (auto& a : active) {
            if (!status_steps_ok) a.done = -1;
            (void)hipEventRecord(a.finished, group_stream(sv->group));
            sv->retiring.push_back(a);
        }
        active.clear();
        sv->cv.notify_all();
    };
    while (true) {
        {
            std::unique_lock<std::mutex> lk(sv->mu);
            if (active.empty())
                sv->cv.wait(lk, [&] { return sv->stop || !sv->built.empty(); });
            if (sv->stop) break;
            while (!sv->built.empty()) {
                active.push_back(sv->built.front());
                sv->built.pop_front();
                changed = true;
            }
        }
        if (changed) {
            hs.clear();
            dts.clear();
            for (auto& a : active) {
                hs.push_back(a.h);
                dts.push_back(a.dt);
            }
            if (group_set(sv->group, hs.data(), (int)hs.size()) != DTX_OK) {
                {
                    std::lock_guard<std::mutex> lk(sv->mu);
                    sv->err = group_error(sv->group);
                }
                retire_all(0);
                continue;
            }
            changed = false;
        }
        if (group_step(sv->group, dts.data(), 0) != DTX_OK) {
            {
                std::lock_guard<std::mutex> lk(sv->mu);
                sv->err = group_error(sv->group);
            }
            retire_all(0);
            continue;
        }
        bool finished = false;
        for (auto& a : active) finished |= (++a.done >= a.steps);
        if (finished) {
            std::vector<Active> keep;
            std::lock_guard<std::mutex> lk(sv->mu);
            for (auto& a : active) {
                if (a.done >= a.steps) {
                    (void)hipEventRecord(a.finished, group_stream(sv->group));
                    sv->retiring.push_back(a);
                } else {
                    keep.push_back(a);
                }
            }
            active.swap(keep);
            changed = true;
            sv->cv.notify_all();
        }
    }
    // stopped: what is still stepping goes to the reaper as failed
    for (auto& a : active) a.done = -1;
    std::lock_guard<std::mutex> lk(sv->mu);
    for (auto& a : active) {
        (void)hipEventRecord(a.finished, group_stream(sv->group));
        sv->retiring.push_back(a);
    }
    sv->cv.notify_all();
}

extern "C" {

int dtx_request_create(const dtx_request* req, const dtx_server_config* cfg, dtx_handle** out,
                       double* dt) {
    if (!req || !cfg || !out || !dt) return DTX_ERR_INVALID;
    *out = nullptr;
    if (req->cfg.num_species < 1 || req->cfg.num_species > DTX_MAX_SPECIES)
        return DTX_ERR_INVALID;
    std::string err;
    return build_request(*cfg, *req, out, dt, &err);
}

int dtx_server_create(const dtx_server_config* cfg, dtx_server** out) {
    if (!cfg || !out) return DTX_ERR_INVALID;
    *out = nullptr;
    if (cfg->max_active < 1 || cfg->hbm_budget <= 0) return DTX_ERR_INVALID;
    if (cfg->precision < DTX_PREC_FP64 || cfg->precision > DTX_PREC_BF16X3) return DTX_ERR_INVALID;
    dtx_server* sv = new dtx_server();
    sv->cfg = *cfg;
    sv->group = group_create(cfg->device);
    if (!sv->group) {
        delete sv;
        return DTX_ERR_HIP;
    }
    sv->builder = std::thread(builder, sv);
    sv->stepper = std::thread(stepper, sv);
    sv->reaper = std::thread(reaper, sv);
    *out = sv;
    return DTX_OK;
}

static int64_t enqueue_record(dtx_server* sv, Record&& r) {
    std::lock_guard<std::mutex> lk(sv->mu);
    const int64_t rid = sv->next_id++;
    memset(&r.res, 0, sizeof r.res);
    r.res.id = rid;
    r.res.status = 0;
    r.res.submit_ms = now_ms();
    sv->records[rid] = std::move(r);
    sv->pending.push_back(rid);
    sv->unfinished++;
    sv->cv.notify_all();
    return rid;
}

int dtx_server_submit(dtx_server* sv, const dtx_request* req, int64_t* id) {
    if (!sv || !req || !id) return DTX_ERR_INVALID;
    if (req->cfg.num_species < 1 || req->cfg.num_species > DTX_MAX_SPECIES || req->steps < 1)
        return DTX_ERR_INVALID;
    Record r;
    r.req = *req;
    *id = enqueue_record(sv, std::move(r));
    return DTX_OK;
}

int dtx_server_submit_handle(dtx_server* sv, dtx_handle* h, int64_t steps, double dt,
                             int64_t* id) {
    if (!sv || !h || !id || steps < 1 || !(dt > 0.0) || !std::isfinite(dt))
        return DTX_ERR_INVALID;
    const dtx_config& hc = *handle_config(h);
    if (hc.precision != sv->cfg.precision || hc.method != DTX_METHOD_FORWARD_EULER ||
        !handle_has_vectors(h))
        return DTX_ERR_INVALID;
    Record r;
    memset(&r.req, 0, sizeof r.req);
    r.req.cfg = hc;
    r.req.steps = steps;
    r.handle = h;
    r.dt = dt;
    *id = enqueue_record(sv, std::move(r));
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
    if (sv->builder.joinable()) sv->builder.join();
    if (sv->stepper.joinable()) sv->stepper.join();
    if (sv->reaper.joinable()) sv->reaper.join();
    (void)group_sync(sv->group);
    // whatever the builder left built or the reaper left waiting
    for (auto* q : {&sv->built, &sv->retiring})
        for (auto& a : *q) {
            if (a.finished) (void)hipEventDestroy(a.finished);
            if (a.h) dtx_destroy(a.h);
        }
    for (auto& kv : sv->records)
        if (kv.second.handle && kv.second.res.status == 0) dtx_destroy(kv.second.handle);
    group_destroy(sv->group);
    delete sv;
}

const char* dtx_server_last_error(const dtx_server* sv) { return sv ? sv->err.c_str() : ""; }

}  // extern "C"
