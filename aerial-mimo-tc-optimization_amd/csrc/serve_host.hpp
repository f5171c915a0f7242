// serve_host.hpp -- host-only parts of the serving loop (no HIP): the arena allocator and the batch-latency model
// of deadline-aware admission. Used by serve.cpp and by the CPU-only checks under AddressSanitizer.
#pragma once
#include <algorithm>
#include <cstddef>
#include <deque>
#include <iterator>
#include <map>
#include <utility>
#include <vector>

namespace pe {
namespace serve_host {

// First-fit allocator over one device arena (offsets only; 256-byte granules).
class Arena {
public:
    explicit Arena(size_t bytes) { free_[0] = bytes; }
    bool alloc(size_t n, size_t &off) {
        n = (n + 255) & ~size_t(255);
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= n) {
                off = it->first;
                const size_t rest = it->second - n;
                free_.erase(it);
                if (rest) free_[off + n] = rest;
                return true;
            }
        return false;
    }
    void release(size_t off, size_t n) {
        n = (n + 255) & ~size_t(255);
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }

    size_t free_bytes() const {
        size_t s = 0;
        for (const auto &b : free_) s += b.second;
        return s;
    }
    size_t free_blocks() const { return free_.size(); }

private:
    std::map<size_t, size_t> free_;  // offset -> bytes of each free block, disjoint and never adjacent
};

// Latency of a batch as a function of its size: the latencies (launch to observed completion) of the last 16
// completed batches fitted as latency = a + b * subcarriers by a Theil-Sen line (b: median of the pairwise slopes,
// a: median of latency - b * subcarriers), robust to the occasional batch delayed by the host. It predicts only
// once 8 batches are in the window (a line through two or three batches of similar size extrapolates wildly).
struct LatencyFit {
    static constexpr size_t kWindow = 16, kMinPoints = 8;
    std::deque<std::pair<double, double>> recent;  // (subcarriers, latency)
    double a = 0.0, b = 0.0;
    bool ready = false;
    double at(double n_sc) const { return a + b * n_sc; }
    static double median(std::vector<double> &v) {
        std::nth_element(v.begin(), v.begin() + v.size() / 2, v.end());
        return v[v.size() / 2];
    }
    void observe(double latency, double n_sc) {
        if (n_sc <= 0) return;
        recent.emplace_back(n_sc, latency);
        if (recent.size() > kWindow) recent.pop_front();
        if (recent.size() < kMinPoints) return;
        std::vector<double> v;
        for (size_t i = 0; i < recent.size(); ++i)
            for (size_t j = i + 1; j < recent.size(); ++j)
                if (recent[j].first != recent[i].first)
                    v.push_back((recent[j].second - recent[i].second) / (recent[j].first - recent[i].first));
        b = v.empty() ? 0.0 : std::max(0.0, median(v));
        v.clear();
        for (const auto &r : recent) v.push_back(r.second - b * r.first);
        a = std::max(0.0, median(v));
        ready = true;
    }
};

// Batch-latency model of deadline-aware admission. A batch's latency depends on its size and on how many other
// batches share the GPU while it runs, so batches are fitted separately by the number of batches already in flight
// when they were launched (and once pooled, the fallback while a class has too few batches). A batch is
// predicted to take its class's fit plus, when other batches are in flight, a margin calibrated from observed
// lateness once per completed batch: a batch with a job past its deadline raises the margin by kUp, a batch
// without lowers it by kDown (floor 0, cap one deadline). The margin settles where kDown / (kUp + kDown) = 0.5 % of
// batches hold a late job, so under overload the served jobs' tail stays inside the deadline; when nothing is late
// it falls back to zero within about two hundred batches, so below capacity a job is shed only when the fit itself
// puts it past the deadline. Measured alternatives: feedback per job instead of per batch let one late batch of
// many jobs raise the margin to the cap (28 % shed at 90 % load); gating shedding on sustained lateness let the
// margin reach the cap while disengaged and oscillate under overload. Step sizes (64-cell serving, 10,000 jobs per
// load): kUp = 100 us kept p99 at 0.46-0.48 ms for 100-120 % load with 6-18 % of jobs shed or late, against
// 0.48-0.51 ms at 50 us (arrivals change faster than the smaller steps follow); 200 us shed 0.2 % at 90 % load, and
// a 0.2 % target (kDown = kUp / 499) shed up to 23 % at 110 %; raising the margin from zero only after 3 late
// batches among the last 100 (to ignore isolated late batches below capacity) let lateness come in bursts each time
// it had decayed to zero: p99 0.47-0.55 against 0.45-0.49 ms at 100-150 % load, with no gain at 90 % (0-7 jobs shed
// in four runs either way). Until a fit is ready nothing is predicted late.
struct LatencyModel {
    static constexpr int kClasses = 9;  // 0 .. 8 batches in flight at launch
    static constexpr double kUp = 100e-6, kDown = 100e-6 / 199.0;  // seconds per batch with / without a late job
    double iso = 0.0;  // one-job batch latency at warm-up (seconds), reported
    LatencyFit by_load[kClasses], pooled;
    double margin = 0.0;
    const LatencyFit *fit_for(int inflight) const {
        const LatencyFit &f = by_load[std::min(inflight, kClasses - 1)];
        return f.ready ? &f : pooled.ready ? &pooled : nullptr;
    }
    // predicted latency of a batch of n_sc subcarriers launched with `inflight` other batches on the GPU
    double predict(double n_sc, int inflight) const {
        const LatencyFit *f = fit_for(inflight);
        return f ? f->at(n_sc) + (inflight > 0 ? margin : 0.0) : 0.0;
    }
    void observe(double latency, double n_sc, int inflight_at_launch) {
        by_load[std::min(inflight_at_launch, kClasses - 1)].observe(latency, n_sc);
        pooled.observe(latency, n_sc);
    }
    void outcome(bool late, double cap) { margin = std::min(cap, std::max(0.0, margin + (late ? kUp : -kDown))); }
    double slope() const { return pooled.b; }
};

}  // namespace serve_host
}  // namespace pe
