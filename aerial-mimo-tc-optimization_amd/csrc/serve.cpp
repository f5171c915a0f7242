// serve.cpp -- slot serving loop (pe_serve_run): PUSCH slot jobs from many cells arrive as a Poisson process,
// are admitted against a device-memory budget, and are launched in continuous batches on several streams.
//
// The reference measures N cells launched together and checks one 0.5 ms deadline
// (multicell_evidence.cu:497-566); it has no arrival process. Here one host thread runs the loop:
//   1. completions: a batch whose event has completed releases its jobs' device buffers; each job's latency is
//      completion time minus its arrival time;
//   2. arrivals: jobs whose arrival time has passed join a FIFO;
//   3. admission and launch: whenever a stream is free, the FIFO head is admitted while its buffers (C,
//      noise_inv, LLR) fit the device arena (the HBM budget), up to max_batch_jobs; the batch runs as one
//      launch of that stream's slot executor (MMSE, fused apply + demap; eagerly with grids
//      sized to the batch, or as a replay of its captured graph).
// Shedding (shed_late): 1 drops a job whose deadline has passed before launch; 2 (deadline-aware admission)
// drops a job whose predicted completion -- its age plus the predicted latency of the batch it would join --
// is past its deadline (the model is described at LatencyModel). Under overload this serves the jobs that
// can still make their deadline instead of serving every job late.
// The report carries the reference's online contention indicator (pmu_contention.cu:230-245): the mean batch
// latency of the first 5 batches divided by the isolated one-job latency.
// Jobs that arrive while batches are on the GPU are admitted as soon as a stream and memory free up, so the
// batches overlap on the device. Each job's channel estimate, MMSE coefficients and samples stay
// device-resident; the coefficients are computed once per job and applied to all of its data symbols.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <map>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "pe_internal.hpp"
#include "pusch_eq.h"
#include "serve_host.hpp"
#include "slot_exec.hpp"

namespace {

using pe::serve_host::Arena;
using pe::serve_host::LatencyModel;

pe_status serve_fail(pe_status s, const std::string &m) {
    pe::set_error(m.c_str());
    return s;
}

struct JobBuf {
    size_t off = 0, bytes = 0;
};

// Coefficient bytes of a job: complex64, or binary16 pairs plus one exponent per subcarrier.
size_t coef_bytes(const pe_cell_desc &c) {
    const size_t nsc = (size_t)c.nSC;
    const size_t b = c.prec == PE_PREC_F16_C16 ? nsc * c.nLayer * c.nAnt * 4 + nsc * 4 : nsc * c.nLayer * c.nAnt * 8;
    return (b + 255) & ~size_t(255);
}

size_t job_bytes(const pe_cell_desc &c, bool llr_in_arena) {
    const size_t nsc = (size_t)c.nSC;
    size_t b = coef_bytes(c) + ((nsc * c.nLayer * 4 + 255) & ~size_t(255));
    if (llr_in_arena) b += (nsc * c.nLayer * c.nSym * 2 * c.nPamBits * 4 + 255) & ~size_t(255);
    return b;
}

double nearest_rank(std::vector<double> v, double p) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    long i = (long)std::ceil(p / 100.0 * (double)v.size()) - 1;
    i = std::max(0L, std::min((long)v.size() - 1, i));
    return v[(size_t)i];
}

}  // namespace

extern "C" pe_status pe_serve_run(const pe_serve_config *cfg, pe_serve_report *rep, double *job_latency_s,
                                  int *job_cell) {
    using clock = std::chrono::steady_clock;
    if (!cfg || !rep) return serve_fail(PE_ERR_INVALID, "null config or report");
    if (!cfg->cells || cfg->n_cells < 1) return serve_fail(PE_ERR_INVALID, "no cells");
    if (cfg->n_jobs < 1) return serve_fail(PE_ERR_INVALID, "n_jobs < 1");
    if (cfg->max_batch_jobs < 1 || cfg->max_inflight < 1 || cfg->max_inflight > 8)
        return serve_fail(PE_ERR_INVALID, "max_batch_jobs >= 1 and 1 <= max_inflight <= 8 required");
    if (!(cfg->deadline_s > 0.0)) return serve_fail(PE_ERR_INVALID, "deadline_s must be positive");
    if (cfg->shed_late < 0 || cfg->shed_late > 2) return serve_fail(PE_ERR_INVALID, "shed_late must be 0, 1 or 2");
    const bool llr_in_arena = cfg->job_llr == nullptr;
    std::vector<pe_cell_desc> pop(cfg->cells, cfg->cells + cfg->n_cells);
    size_t largest = 0;
    std::string err;
    for (int i = 0; i < cfg->n_cells; ++i) {
        pe_cell_desc c = pop[i];
        if (!c.H || !c.sigma2 || !c.Y_rx)
            return serve_fail(PE_ERR_INVALID, "cell " + std::to_string(i) + ": H, sigma2 and Y_rx are required");
        c.C = reinterpret_cast<pe_c64 *>(16);  // job buffers come from the arena; validate shapes only
        c.noise_inv = reinterpret_cast<float *>(16);
        c.llr = reinterpret_cast<float *>(16);
        pe_status s = pe::check_cell_desc(c, i, err);
        if (s != PE_OK) return serve_fail(s, err);
        largest = std::max(largest, job_bytes(c, llr_in_arena));
    }
    if (largest > cfg->hbm_budget_bytes)
        return serve_fail(PE_ERR_CAPACITY, "hbm_budget_bytes is smaller than one job of the largest cell (" +
                                               std::to_string(largest) + " bytes)");
    const int nj = cfg->n_jobs;
    for (int j = 0; cfg->job_cells && j < nj; ++j)
        if (cfg->job_cells[j] < 0 || cfg->job_cells[j] >= cfg->n_cells)
            return serve_fail(PE_ERR_INVALID, "job " + std::to_string(j) + ": cell index out of range");
    if ((cfg->job_H || cfg->job_sigma2 || cfg->job_Y) && !cfg->job_cells)
        return serve_fail(PE_ERR_INVALID, "per-job inputs need job_cells (their shapes follow the job's cell)");

    // ---- arrival process (virtual times from the start of the run) ------------------------------------
    std::vector<double> arrival(nj);
    std::vector<int> cellof(nj);
    {
        std::mt19937_64 rng(cfg->seed);
        std::exponential_distribution<double> ex(cfg->arrival_rate_hz > 0 ? cfg->arrival_rate_hz : 1.0);
        std::uniform_int_distribution<int> pick(0, cfg->n_cells - 1);
        double t = 0.0;
        for (int j = 0; j < nj; ++j) {
            if (cfg->arrival_rate_hz > 0) t += ex(rng);
            arrival[j] = cfg->arrival_rate_hz > 0 ? t : 0.0;
            const int drawn = pick(rng);  // drawn either way: the arrival times do not depend on job_cells
            cellof[j] = cfg->job_cells ? cfg->job_cells[j] : drawn;
        }
    }
    // per-job inputs and destinations: present and aligned as the tile kernels need
    for (int j = 0; j < nj; ++j) {
        const pe_cell_desc &c = pop[cellof[j]];
        const struct {
            const void *p;
            bool given;
            size_t a;
            const char *name;
        } req[] = {{cfg->job_llr ? cfg->job_llr[j] : nullptr, cfg->job_llr != nullptr,
                    pe::llr_align(c.nPamBits, false), "job_llr"},
                   {cfg->job_H ? cfg->job_H[j] : nullptr, cfg->job_H != nullptr, pe::h_align(c.nLayer), "job_H"},
                   {cfg->job_sigma2 ? cfg->job_sigma2[j] : nullptr, cfg->job_sigma2 != nullptr, 4, "job_sigma2"},
                   {cfg->job_Y ? cfg->job_Y[j] : nullptr, cfg->job_Y != nullptr, 16, "job_Y"}};
        for (const auto &r : req)
            if (r.given && (!r.p || !pe::is_aligned(r.p, r.a)))
                return serve_fail(PE_ERR_INVALID, "job " + std::to_string(j) + ": " + r.name + " is null or not " +
                                                      std::to_string(r.a) + "-byte aligned");
    }

    // ---- device state ------------------------------------------------------------------------------------
    unsigned char *arena_dev = nullptr;
    hipError_t e = hipMalloc(&arena_dev, cfg->hbm_budget_bytes);
    if (e != hipSuccess) return serve_fail(PE_ERR_NOMEM, std::string("arena: ") + hipGetErrorString(e));
    const int K = cfg->max_inflight;
    std::vector<std::unique_ptr<pe::SlotExec>> ex(K);
    std::vector<hipStream_t> st(K, nullptr);
    std::vector<hipEvent_t> ev(K, nullptr);
    auto cleanup = [&]() {
        for (int k = 0; k < K; ++k) {
            if (st[k]) (void)hipStreamSynchronize(st[k]);
            if (ev[k]) (void)hipEventDestroy(ev[k]);
            if (st[k]) (void)hipStreamDestroy(st[k]);
            ex[k].reset();
        }
        (void)hipFree(arena_dev);
    };
    for (int k = 0; k < K; ++k) {
        ex[k].reset(new pe::SlotExec());
        // eager batches form the coefficients on chip from solve records (per-job C is arena memory nobody else
        // reads); graph replays keep the coefficient path, whose striding apply kernel needs fewer registers
        const unsigned xf = PE_MMSE_UNBIASED | (cfg->use_graph ? 0u : PE_SLOT_SOLVE_RECORDS);
        pe_status s = ex[k]->init(pop.data(), (int)pop.size(), cfg->max_batch_jobs, xf, false, err);
        if (s == PE_OK && hipStreamCreateWithFlags(&st[k], hipStreamNonBlocking) != hipSuccess) s = PE_ERR_DEVICE;
        if (s == PE_OK && hipEventCreateWithFlags(&ev[k], hipEventDisableTiming) != hipSuccess) s = PE_ERR_DEVICE;
        if (s != PE_OK) {
            cleanup();
            return serve_fail(s, err.empty() ? "stream/event creation failed" : err);
        }
    }
    Arena arena(cfg->hbm_budget_bytes);
    std::vector<JobBuf> buf(nj);
    // job j's descriptor; its C / noise_inv (and LLRs when in_arena) in the arena block b
    auto make_desc = [&](int j, const JobBuf &b, bool in_arena) {
        pe_cell_desc d = pop[cellof[j]];
        if (cfg->job_H) d.H = cfg->job_H[j];
        if (cfg->job_sigma2) d.sigma2 = cfg->job_sigma2[j];
        if (cfg->job_Y) d.Y_rx = cfg->job_Y[j];
        const size_t nsc = (size_t)d.nSC;
        unsigned char *p = arena_dev + b.off;
        d.C = reinterpret_cast<pe_c64 *>(p);
        p += coef_bytes(d);
        d.noise_inv = reinterpret_cast<float *>(p);
        p += (nsc * d.nLayer * 4 + 255) & ~size_t(255);
        d.llr = in_arena ? reinterpret_cast<float *>(p) : cfg->job_llr[j];
        return d;
    };
    // graph replay (capacity grids) for batches of at least a quarter of max_batch_jobs; smaller batches are
    // launched eagerly with grids sized to their tiles, so a lone job does not dispatch a full batch's grid
    auto submit = [&](int k, const std::vector<pe_cell_desc> &cells, bool force_graph) -> pe_status {
        pe_status s = ex[k]->stage(cells.data(), (int)cells.size(), err);
        if (s == PE_OK) s = ex[k]->upload(st[k], err);
        const bool graph = cfg->use_graph && (force_graph || 4 * (int)cells.size() >= cfg->max_batch_jobs);
        if (s == PE_OK) s = graph ? ex[k]->graph_launch(st[k], err) : ex[k]->launch(st[k], err, true);
        if (s == PE_OK && hipEventRecord(ev[k], st[k]) != hipSuccess) {
            s = PE_ERR_DEVICE;
            err = "event record failed";
        }
        return s;
    };

    // warm-up, not measured: one batch of the first job's cell on every stream, as a graph (captured here)
    // and eagerly (every kernel of the executor runs, at least one workgroup each). Then the isolated latency
    // of one job (eager, idle GPU; median of 5) seeds the latency model.
    LatencyModel model;
    {
        // the warm-up job's LLRs stay in the arena: no caller buffer is written for a job that may be shed
        JobBuf b;
        b.bytes = job_bytes(pop[cellof[0]], true);
        if (!arena.alloc(b.bytes, b.off)) {
            cleanup();
            return serve_fail(PE_ERR_CAPACITY, "arena cannot hold one job");
        }
        pe_cell_desc d = make_desc(0, b, true);
        std::vector<pe_cell_desc> one(1, d);
        for (int k = 0; k < K; ++k) {
            pe_status s = submit(k, one, true);
            if (s == PE_OK && hipStreamSynchronize(st[k]) != hipSuccess) s = PE_ERR_DEVICE;
            if (s == PE_OK) s = submit(k, one, false);
            if (s == PE_OK && hipStreamSynchronize(st[k]) != hipSuccess) s = PE_ERR_DEVICE;
            if (s != PE_OK) {
                cleanup();
                return serve_fail(s, err);
            }
        }
        std::vector<double> iso;
        for (int r = 0; r < 5; ++r) {
            const auto a = clock::now();
            pe_status s = submit(0, one, false);
            while (s == PE_OK && hipEventQuery(ev[0]) == hipErrorNotReady) {
            }
            if (s == PE_OK && hipEventSynchronize(ev[0]) != hipSuccess) s = PE_ERR_DEVICE;
            if (s != PE_OK) {
                cleanup();
                return serve_fail(s, err);
            }
            iso.push_back(std::chrono::duration<double>(clock::now() - a).count());
        }
        model.iso = nearest_rank(iso, 50.0);
        arena.release(b.off, b.bytes);
    }

    // ---- loop ----------------------------------------------------------------------------------------------
    std::vector<double> lat(nj, -1.0);
    std::vector<std::vector<int>> batch_jobs(K);
    std::vector<bool> busy(K, false);
    std::vector<double> batch_t0(K, 0.0), batch_sc(K, 0.0);  // per stream: launch time, subcarriers
    std::vector<int> batch_inflight(K, 0);                   // per stream: other batches in flight at launch
    std::vector<double> first_batches;  // latencies of the first 5 batches (contention indicator)
    std::deque<int> ready;
    int next = 0, finished = 0, shed = 0, batches = 0, max_inflight_seen = 0;
    long batched_jobs = 0;
    std::vector<pe_cell_desc> cells;
    const auto t0 = clock::now();
    auto now_s = [&]() { return std::chrono::duration<double>(clock::now() - t0).count(); };
    pe_status status = PE_OK;
    while (finished + shed < nj) {
        // 1. completions
        int inflight = 0;
        for (int k = 0; k < K; ++k) {
            if (!busy[k]) continue;
            const hipError_t q = hipEventQuery(ev[k]);
            if (q == hipErrorNotReady) {
                ++inflight;
                continue;
            }
            if (q != hipSuccess) {
                status = PE_ERR_DEVICE;
                err = std::string("batch failed: ") + hipGetErrorString(q);
                break;
            }
            const double t = now_s();
            model.observe(t - batch_t0[k], batch_sc[k], batch_inflight[k]);
            if (first_batches.size() < 5) first_batches.push_back(t - batch_t0[k]);
            bool batch_late = false;
            for (int j : batch_jobs[k]) {
                lat[j] = t - arrival[j];
                batch_late = batch_late || lat[j] > cfg->deadline_s;
                arena.release(buf[j].off, buf[j].bytes);
                ++finished;
            }
            model.outcome(batch_late, cfg->deadline_s);
            batch_jobs[k].clear();
            busy[k] = false;
        }
        if (status != PE_OK) break;
        // 2. arrivals
        const double now = now_s();
        while (next < nj && arrival[next] <= now) ready.push_back(next++);
        // 3. admission and launch on free streams
        for (int k = 0; k < K && !ready.empty(); ++k) {
            if (busy[k]) continue;
            // job ages at this admission (a batch submitted on a previous stream in this pass took host time)
            const double t_adm = now_s();
            cells.clear();
            double sc_batch = 0.0;
            // deadline-aware admission predicts with the batch this launch will have if nothing more is shed: the
            // jobs admitted so far plus the queue's head up to max_batch_jobs (a batch only grows while it is
            // formed; a shed job leaves it, and the prediction is made again for what remains)
            int n_busy = 0;
            for (int kk = 0; kk < K; ++kk) n_busy += busy[kk] ? 1 : 0;
            auto prospective = [&]() {
                double sc = sc_batch;
                int m = (int)cells.size();
                for (auto it = ready.begin(); it != ready.end() && m < cfg->max_batch_jobs; ++it, ++m)
                    sc += pop[cellof[*it]].nSC;
                return model.predict(sc, n_busy);  // no batch on the GPU: the fit without the margin
            };
            double pred = cfg->shed_late == 2 ? prospective() : 0.0;
            while (!ready.empty() && (int)cells.size() < cfg->max_batch_jobs) {
                const int j = ready.front();
                const double nsc_j = pop[cellof[j]].nSC;
                const bool late = cfg->shed_late == 1 && t_adm - arrival[j] > cfg->deadline_s;
                const bool predicted_late = cfg->shed_late == 2 && t_adm - arrival[j] + pred > cfg->deadline_s;
                if (late || predicted_late) {
                    ready.pop_front();
                    ++shed;
                    if (predicted_late) pred = prospective();
                    continue;
                }
                JobBuf b;
                b.bytes = job_bytes(pop[cellof[j]], llr_in_arena);
                if (!arena.alloc(b.bytes, b.off)) break;  // budget exhausted: wait for completions (FIFO order)
                buf[j] = b;
                ready.pop_front();
                cells.push_back(make_desc(j, b, llr_in_arena));
                batch_jobs[k].push_back(j);
                sc_batch += nsc_j;
            }
            if (cells.empty()) {
                if (ready.empty()) continue;  // everything at the head was shed: the next stream may take more
                break;                        // budget exhausted
            }
            batch_t0[k] = now_s();
            batch_sc[k] = sc_batch;
            batch_inflight[k] = 0;
            for (int kk = 0; kk < K; ++kk) batch_inflight[k] += busy[kk] ? 1 : 0;
            status = submit(k, cells, false);
            if (status != PE_OK) break;
            busy[k] = true;
            ++batches;
            batched_jobs += (long)cells.size();
            ++inflight;
        }
        if (status != PE_OK) break;
        max_inflight_seen = std::max(max_inflight_seen, inflight);
    }
    const double wall = now_s();
    cleanup();
    if (status != PE_OK) return serve_fail(status, err);

    std::vector<double> served;
    int missed = shed, late = 0;
    double sum = 0.0, mx = 0.0;
    long sc_done = 0;
    for (int j = 0; j < nj; ++j) {
        if (job_latency_s) job_latency_s[j] = lat[j];
        if (job_cell) job_cell[j] = cellof[j];
        if (lat[j] < 0) continue;
        served.push_back(lat[j]);
        sum += lat[j];
        mx = std::max(mx, lat[j]);
        if (lat[j] > cfg->deadline_s) {
            ++missed;
            ++late;
        }
        sc_done += pop[cellof[j]].nSC;
    }
    rep->n_jobs = nj;
    rep->n_done = finished;
    rep->n_missed = missed;
    rep->n_shed = shed;
    rep->n_batches = batches;
    rep->wall_s = wall;
    rep->p50_s = nearest_rank(served, 50.0);
    rep->p99_s = nearest_rank(served, 99.0);
    rep->max_s = mx;
    rep->mean_s = served.empty() ? 0.0 : sum / (double)served.size();
    rep->miss_rate = (double)missed / (double)nj;
    rep->jobs_per_s = wall > 0 ? finished / wall : 0.0;
    rep->subcarriers_per_s = wall > 0 ? (double)sc_done / wall : 0.0;
    rep->mean_batch_jobs = batches ? (double)batched_jobs / batches : 0.0;
    rep->max_inflight_seen = max_inflight_seen;
    rep->offered_rate_hz = cfg->arrival_rate_hz;
    rep->iso_job_s = model.iso;
    double fb = 0.0;
    for (double v : first_batches) fb += v;
    rep->contention_index = first_batches.empty() || model.iso <= 0 ? 0.0 : fb / first_batches.size() / model.iso;
    rep->model_s_per_sc = model.slope();
    rep->n_late = late;
    rep->model_margin_s = model.margin;
    pe::set_error("");
    return PE_OK;
}
