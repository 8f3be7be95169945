// Serving loop: implicit-step requests of ensemble members on meshes of
// several refinement levels, continuously batched on one GPU.
//
// No counterpart in the reference (which steps one BarVortex at a time, e.g.
// examples/test_barvortex.cpp:441-470); built on the slot API of
// phg_barvortex_* so that every request is exactly BarVortex::calc of its
// member.
//
// * One engine per mesh level: a BarVortex object with S slots, its own HIP
//   stream and its own host thread, so the levels' kernels overlap on the
//   device.
// * Admission by HBM footprint: each engine's slot count is the largest of
//   8, 4, 2, 1 whose estimated device bytes (matrices, Krylov bases, slot
//   vectors) fit the budget left after the member states.
// * Continuous batching: between two fixed-point passes an engine retires
//   the members whose step converged and admits released requests into the
//   free slots (FIFO by arrival; a member's next request waits for its
//   previous one), so requests that arrive while others run join them.
// * Member states stay resident on the device between requests.
// * Optional sub-steps per level: a request runs k steps of tau / k in its
//   slot, for levels where the fixed point does not contract at tau.
// Requests are released at their arrival times (open loop, wall clock), and
// latency is finish - arrival.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>

#include "phg_internal.h"

using namespace phg;

namespace {

struct Member {
    int level = 0, engine = -1;
    DevBuf<double> state;
    bool busy = false;
};

struct Engine {
    int level = 0, S = 0, V = 0, substeps = 1;
    phg_mesh_t mesh = nullptr;
    phg_barvortex_t bv = nullptr;
    hipStream_t stream = nullptr;
    double occupancy_sum = 0.0;
    long passes = 0;
    ~Engine()
    {
        if (bv) phg_barvortex_destroy(bv);
        if (mesh) phg_mesh_destroy(mesh);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

double now_s(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct phg_server_s {
    phg_serve_config cfg{};
    int device = 0;
    std::vector<std::unique_ptr<Engine>> engines;
    std::vector<std::unique_ptr<Member>> members;
    size_t engine_bytes = 0, member_bytes = 0;
};

namespace {

int engine_of_level(const phg_server_s* S, int level)
{
    for (size_t e = 0; e < S->engines.size(); ++e)
        if (S->engines[e]->level == level) return (int)e;
    return -1;
}

struct RunShared {
    const int* member;
    const double* arrival;
    double* start;
    double* finish;
    phg_bv_info* infos;
    std::chrono::steady_clock::time_point t0;
    std::mutex err_mu;
    int rc = PHG_OK;
    std::string err;
};

void engine_loop(phg_server_s* Sv, int e, std::vector<int> reqs, RunShared* R)
{
    (void)hipSetDevice(Sv->device);
    Engine& E = *Sv->engines[e];
    const int S = E.S;
    std::vector<int> slot_req(S, -1), state(S, 0), slot_left(S, 0);
    std::vector<phg_bv_info> inf(S);
    std::deque<int> waiting;
    size_t next = 0, left = reqs.size();
    auto fail = [&](int rc) {
        std::lock_guard<std::mutex> g(R->err_mu);
        if (R->rc == PHG_OK) {
            R->rc = rc;
            R->err = phg_last_error();
        }
    };
    while (left > 0) {
        const double t = now_s(R->t0);
        while (next < reqs.size() && R->arrival[reqs[next]] <= t) waiting.push_back(reqs[next++]);
        // admit into free slots, FIFO among members that are not stepping
        for (auto it = waiting.begin(); it != waiting.end();) {
            int free_slot = -1;
            for (int k = 0; k < S && free_slot < 0; ++k)
                if (slot_req[k] < 0) free_slot = k;
            if (free_slot < 0) break;
            const int r = *it;
            Member& M = *Sv->members[R->member[r]];
            if (M.busy) {
                ++it;
                continue;
            }
            int rc = phg_barvortex_load(E.bv, free_slot, M.state.p, E.stream);
            if (rc) {
                fail(rc);
                return;
            }
            M.busy = true;
            slot_req[free_slot] = r;
            slot_left[free_slot] = E.substeps;
            R->start[r] = now_s(R->t0);
            it = waiting.erase(it);
        }
        int active = 0;
        for (int k = 0; k < S; ++k) active += slot_req[k] >= 0;
        if (active == 0) {
            if (next < reqs.size()) {
                const double dt = R->arrival[reqs[next]] - now_s(R->t0);
                if (dt > 0) std::this_thread::sleep_for(std::chrono::duration<double>(dt));
            }
            continue;
        }
        int rc = phg_barvortex_iterate(E.bv, &Sv->cfg.opts, state.data(), inf.data(), E.stream);
        if (rc) {
            fail(rc);
            return;
        }
        E.occupancy_sum += active;
        E.passes += 1;
        for (int k = 0; k < S; ++k) {
            if (slot_req[k] < 0 || state[k] != 2) continue;
            const int r = slot_req[k];
            Member& M = *Sv->members[R->member[r]];
            rc = phg_barvortex_store(E.bv, k, M.state.p, E.stream);
            if (rc) {
                fail(rc);
                return;
            }
            if (--slot_left[k] > 0 && !inf[k].diverged) {
                // next sub-step of the request: the slot restarts from the new state
                rc = phg_barvortex_load(E.bv, k, M.state.p, E.stream);
                if (rc) {
                    fail(rc);
                    return;
                }
                continue;
            }
            if (hipStreamSynchronize(E.stream) != hipSuccess) {
                set_error("serve: stream error");
                fail(PHG_ERR_HIP);
                return;
            }
            R->finish[r] = now_s(R->t0);
            if (R->infos) R->infos[r] = inf[k];
            if (Sv->cfg.verbose)
                fprintf(stderr,
                        "[serve] L%d request %d member %d: %d passes (rel %.1e), "
                        "arrived %.3fs admitted %.3fs done %.3fs, engine passes %ld\n",
                        E.level, r, R->member[r], inf[k].iterations, inf[k].rel_change,
                        R->arrival[r], R->start[r], R->finish[r], E.passes);
            M.busy = false;
            slot_req[k] = -1;
            --left;
        }
        {
            std::lock_guard<std::mutex> g(R->err_mu);
            if (R->rc != PHG_OK) return;
        }
    }
}

double percentile(std::vector<double> v, double q)
{
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    size_t k = (size_t)std::ceil(q * v.size());
    if (k < 1) k = 1;
    return v[std::min(k, v.size()) - 1];
}

}  // namespace

extern "C" {

void phg_serve_default_config(phg_serve_config* c)
{
    memset(c, 0, sizeof *c);
    c->n_levels = 4;
    for (int i = 0; i < 4; ++i) c->levels[i] = 5 + i;
    c->hbm_budget = (size_t)200 << 30;
    c->coriolis = 2.0;
    const double omega = 7.292e-5;  // the time unit is 1 / Earth's angular velocity
    c->conf.tau = 600.0 * omega;
    c->conf.sigma = 1.6e-2;
    c->conf.mu = 8e-5;
    c->conf.k1 = 1.0;
    c->conf.k2 = 1.0;
    c->conf.theta = 0.5;
    phg_bv_default_opts(&c->opts);
    c->opts.max_iters = 100;  // a step not done in 100 passes returns unconverged
}

int phg_serve_create(const phg_serve_config* cfg, phg_server_t* out)
{
    if (!cfg || !out || cfg->n_levels < 1 || cfg->n_levels > PHG_SERVE_MAXLEVELS) {
        set_error("phg_serve_create: bad argument");
        return PHG_ERR_ARG;
    }
    std::unique_ptr<phg_server_s> Sv(new phg_server_s());
    Sv->cfg = *cfg;
    PHG_HIP(hipGetDevice(&Sv->device));
    size_t used = 0;
    for (int i = 0; i < cfg->n_levels; ++i) {
        std::unique_ptr<Engine> E(new Engine());
        E->level = cfg->levels[i];
        int rc = phg_mesh_sphere(E->level, &E->mesh);
        if (rc) return rc;
        E->V = E->mesh->h.n_points;
        int S = cfg->slots[i];
        if (S <= 0) {
            // admission: the largest slot count whose footprint fits the budget
            S = 8;
            while (S > 1 && used + phg_barvortex_estimate_bytes(E->V, S) > cfg->hbm_budget) S /= 2;
        }
        const size_t need = phg_barvortex_estimate_bytes(E->V, S);
        if (used + need > cfg->hbm_budget) {
            set_error("phg_serve_create: HBM budget too small for level " +
                      std::to_string(E->level));
            return PHG_ERR_NOMEM;
        }
        used += need;
        E->S = S;
        E->substeps = cfg->substeps[i] > 1 ? cfg->substeps[i] : 1;
        PHG_HIP(hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking));
        // lh = coriolis * sin(latitude) (proj of conf.coriolis, barvortex_impl.h:121)
        std::vector<double> lh(E->V);
        for (int p = 0; p < E->V; ++p) {
            const double z = E->mesh->h.pr[3 * p + 2];
            lh[p] = cfg->coriolis * std::max(-1.0, std::min(1.0, z));
        }
        DevBuf<double> d_lh;
        PHG_HIP(d_lh.upload(lh.data(), lh.size()));
        phg_bv_conf conf = cfg->conf;
        conf.tau = cfg->conf.tau / E->substeps;
        rc = phg_barvortex_create(E->mesh, &conf, d_lh.p, S, nullptr, &E->bv);
        if (rc) return rc;
        Sv->engines.push_back(std::move(E));
    }
    Sv->engine_bytes = used;
    *out = Sv.release();
    return PHG_OK;
}

void phg_serve_destroy(phg_server_t Sv) { delete Sv; }

int phg_serve_slots(phg_server_t Sv, int level)
{
    const int e = Sv ? engine_of_level(Sv, level) : -1;
    return e < 0 ? 0 : Sv->engines[e]->S;
}

int phg_serve_add_member(phg_server_t Sv, int level, const double* h_u, int* member_id)
{
    const int e = Sv ? engine_of_level(Sv, level) : -1;
    if (e < 0 || !h_u || !member_id) {
        set_error("phg_serve_add_member: bad argument (level not served?)");
        return PHG_ERR_ARG;
    }
    const size_t V = Sv->engines[e]->V;
    if (Sv->engine_bytes + Sv->member_bytes + V * sizeof(double) > Sv->cfg.hbm_budget) {
        set_error("phg_serve_add_member: HBM budget exhausted");
        return PHG_ERR_NOMEM;
    }
    std::unique_ptr<Member> M(new Member());
    M->level = level;
    M->engine = e;
    PHG_HIP(M->state.upload(h_u, V));
    Sv->member_bytes += V * sizeof(double);
    *member_id = (int)Sv->members.size();
    Sv->members.push_back(std::move(M));
    return PHG_OK;
}

int phg_serve_member_state(phg_server_t Sv, int member, double* h_u)
{
    if (!Sv || member < 0 || member >= (int)Sv->members.size() || !h_u) {
        set_error("phg_serve_member_state: bad argument");
        return PHG_ERR_ARG;
    }
    const Member& M = *Sv->members[member];
    PHG_HIP(hipMemcpy(h_u, M.state.p, M.state.n * sizeof(double), hipMemcpyDeviceToHost));
    return PHG_OK;
}

int phg_serve_run(phg_server_t Sv, int n, const int* member, const double* arrival_s,
                  double* start_s, double* finish_s, phg_bv_info* infos, phg_serve_stats* stats)
{
    if (!Sv || n < 0 || (n && (!member || !arrival_s || !start_s || !finish_s))) {
        set_error("phg_serve_run: bad argument");
        return PHG_ERR_ARG;
    }
    for (int r = 0; r < n; ++r)
        if (member[r] < 0 || member[r] >= (int)Sv->members.size()) {
            set_error("phg_serve_run: unknown member");
            return PHG_ERR_ARG;
        }
    std::vector<std::vector<int>> per(Sv->engines.size());
    std::vector<int> order(n);
    for (int r = 0; r < n; ++r) order[r] = r;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return arrival_s[a] < arrival_s[b]; });
    for (int r : order) per[Sv->members[member[r]]->engine].push_back(r);
    for (auto& E : Sv->engines) {
        E->occupancy_sum = 0.0;
        E->passes = 0;
    }
    RunShared R;
    R.member = member;
    R.arrival = arrival_s;
    R.start = start_s;
    R.finish = finish_s;
    R.infos = infos;
    for (int r = 0; r < n; ++r) start_s[r] = finish_s[r] = -1.0;
    R.t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> th;
    for (size_t e = 0; e < Sv->engines.size(); ++e)
        if (!per[e].empty()) th.emplace_back(engine_loop, Sv, (int)e, per[e], &R);
    for (auto& t : th) t.join();
    const double wall = now_s(R.t0);
    if (R.rc != PHG_OK) {
        set_error("phg_serve_run: " + R.err);
        return R.rc;
    }
    if (stats) {
        std::vector<double> lat, svc;
        double first = 1e300, last = 0.0;
        for (int r = 0; r < n; ++r) {
            lat.push_back(finish_s[r] - arrival_s[r]);
            svc.push_back(finish_s[r] - start_s[r]);
            first = std::min(first, arrival_s[r]);
            last = std::max(last, finish_s[r]);
        }
        memset(stats, 0, sizeof *stats);
        stats->completed = n;
        stats->wall_s = wall;
        stats->throughput = n > 0 && last > first ? n / (last - first) : 0.0;
        stats->p50_latency_s = percentile(lat, 0.5);
        stats->p99_latency_s = percentile(lat, 0.99);
        double m = 0.0;
        for (double x : lat) m += x;
        stats->mean_latency_s = n ? m / n : 0.0;
        stats->p50_service_s = percentile(svc, 0.5);
        double occ = 0.0;
        long passes = 0;
        for (auto& E : Sv->engines) {
            occ += E->occupancy_sum;
            passes += E->passes;
        }
        stats->mean_occupancy = passes ? occ / passes : 0.0;
        stats->passes = passes;
        stats->device_bytes = Sv->engine_bytes + Sv->member_bytes;
    }
    return PHG_OK;
}

}  // extern "C"
