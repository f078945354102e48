// Continuous-batching render server over one slot pool.
//
// Requests (a frame each: camera, resolution, samples, depth cap, seed, film)
// queue FIFO and are admitted while the pool has room: admission is by
// path-state memory, i.e. slots (354 B each), in contiguous ranges of
// kPoolAlign slots. A scheduler thread runs the wavefront in blocks of
// iterations; between blocks it retires requests whose slots have all gone
// idle, admits queued requests into the freed slots and applies
// cancellations, so new requests join paths already in flight instead of
// waiting for a frame boundary. The reference renders one frame per
// PTRenderer::render call (libSLR/Renderer/PTRenderer.cpp:26-101) and has no
// serving layer; each request here produces exactly what slrh_render_batch
// produces for that frame with the same seed and slots per pixel.
#include "server.h"

#include <algorithm>
#include <cmath>

namespace slrh {

namespace {
double nowSeconds() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
}  // namespace

int Server::start(DeviceScene* ds, uint64_t pathBytes, int traceVariant, std::string* err) {
    uint64_t slots = pathBytes / kSlotBytes;
    slots = std::min<uint64_t>(slots, 1u << 26);
    slots -= slots % kPoolAlign;
    if (slots < kPoolAlign) { if (err) *err = "path-state budget below one slot block"; return -5; }
    int r = poolOpen(ds, (uint32_t)slots, &pool_, err);
    if (r) return r;
    traceVariant_ = traceVariant;
    blockUsed_.assign(poolCapacity(pool_) / kPoolAlign, false);
    for (uint32_t i = 0; i < kPoolMaxJobs; ++i) freeJobs_.push_back(kPoolMaxJobs - 1 - i);
    thread_ = std::thread([this]() { loop(); });
    return 0;
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_.notify_all();
    if (thread_.joinable()) thread_.join();
    if (pool_) poolClose(pool_);
}

int Server::submit(const JobSpec& spec, uint32_t planes, uint64_t* ticket, std::string* err) {
    const uint64_t pixels = (uint64_t)spec.width * spec.height;
    if (pixels == 0 || spec.spp == 0 || !spec.film) { if (err) *err = "empty request"; return -1; }
    const uint64_t need = (pixels * planesFor(spec, planes) + kPoolAlign - 1) / kPoolAlign * kPoolAlign;
    if (need > poolCapacity(pool_)) { if (err) *err = "request larger than the path-state budget"; return -5; }
    std::lock_guard<std::mutex> lk(mu_);
    if (failed_) { if (err) *err = failure_; return -4; }
    Request q;
    q.ticket = ++lastTicket_;
    q.spec = spec;
    q.planes = planes;
    q.submitted = nowSeconds();
    requests_[q.ticket] = q;
    queue_.push_back(q.ticket);
    *ticket = q.ticket;
    cv_.notify_all();
    return 0;
}

int Server::state(uint64_t ticket, int* st) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = requests_.find(ticket);
    if (it == requests_.end()) return -1;
    *st = it->second.state;
    return 0;
}

int Server::wait(uint64_t ticket, int timeoutMs, int* st) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = requests_.find(ticket);
    if (it == requests_.end()) return -1;
    auto settled = [&]() { return it->second.state >= kDone; };
    if (timeoutMs < 0) cv_.wait(lk, settled);
    else cv_.wait_for(lk, std::chrono::milliseconds(timeoutMs), settled);
    *st = it->second.state;
    return 0;
}

int Server::cancel(uint64_t ticket) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = requests_.find(ticket);
    if (it == requests_.end()) return -1;
    Request& q = it->second;
    if (q.state == kQueued) {
        queue_.erase(std::remove(queue_.begin(), queue_.end(), ticket), queue_.end());
        q.state = kCancelled;
        cv_.notify_all();
    } else if (q.state == kRunning) {
        q.cancelRequested = true;   // applied by the scheduler between blocks
    }
    return 0;
}

void Server::stats(double* out) {
    std::lock_guard<std::mutex> lk(mu_);
    std::vector<double> lat = latencies_;
    std::sort(lat.begin(), lat.end());
    auto pct = [&](double p) {
        if (lat.empty()) return 0.0;
        const size_t k = std::min(lat.size() - 1, (size_t)std::ceil(p * lat.size()) - (p > 0 ? 1 : 0));
        return lat[k] * 1e3;
    };
    double mean = 0;
    for (double x : lat) mean += x;
    uint32_t queued = 0, running = 0;
    for (auto& kv : requests_) {
        queued += kv.second.state == kQueued;
        running += kv.second.state == kRunning;
    }
    out[0] = (double)lat.size();
    out[1] = queued;
    out[2] = running;
    out[3] = pct(0.5);
    out[4] = pct(0.99);
    out[5] = lat.empty() ? 0.0 : mean / lat.size() * 1e3;
    out[6] = samplesDone_;
    out[7] = busySeconds_;
}

// First-fit range of free slot blocks; returns false if none is long enough.
bool Server::allocBlocks(uint32_t nblocks, uint32_t* first) {
    uint32_t run = 0;
    for (uint32_t b = 0; b < blockUsed_.size(); ++b) {
        run = blockUsed_[b] ? 0 : run + 1;
        if (run == nblocks) {
            *first = b + 1 - nblocks;
            for (uint32_t k = *first; k <= b; ++k) blockUsed_[k] = true;
            return true;
        }
    }
    return false;
}

void Server::freeBlocks(uint32_t first, uint32_t nblocks) {
    for (uint32_t k = first; k < first + nblocks; ++k) blockUsed_[k] = false;
}

// Slots per pixel of a request: the caller's count, or with 0 a count that
// depends only on the request and the pool (never on the load, so a seeded
// request renders the same film whenever it is admitted): about
// kServeSlotsPerRequest live paths, at most its spp and 16, and no more
// than the whole pool holds.
uint32_t Server::planesFor(const JobSpec& spec, uint32_t requested) const {
    const uint64_t pixels = (uint64_t)spec.width * spec.height;
    uint64_t planes = requested;
    if (planes == 0) {
        planes = std::max<uint64_t>(1, std::min<uint64_t>({kServeSlotsPerRequest / pixels, spec.spp, 16}));
        const uint64_t fit = std::max<uint64_t>(1, poolCapacity(pool_) / pixels);
        planes = std::min(planes, fit);
    }
    return (uint32_t)std::min<uint64_t>(planes, spec.spp);
}

// Admits queued requests in order while they fit (holding mu_). A request
// whose film a running request still accumulates into waits for it: slots
// add into their film cells without atomics.
bool Server::admitQueued(std::string* err, int* rc) {
    bool any = false;
    while (!queue_.empty() && !freeJobs_.empty()) {
        Request& q = requests_[queue_.front()];
        const uint64_t pixels = (uint64_t)q.spec.width * q.spec.height;
        bool filmBusy = false;
        for (uint64_t t : running_) filmBusy = filmBusy || requests_[t].spec.film == q.spec.film;
        if (filmBusy) break;   // FIFO: wait until that request is done
        const uint32_t planes = planesFor(q.spec, q.planes);
        uint32_t first = 0;
        const uint32_t nblocks = (uint32_t)((planes * pixels + kPoolAlign - 1) / kPoolAlign);
        if (!allocBlocks(nblocks, &first)) break;   // FIFO: wait for room
        const uint32_t job = freeJobs_.back();
        freeJobs_.pop_back();
        int r = poolAdmit(pool_, job, q.spec, planes, first * kPoolAlign, err);
        if (r) { *rc = r; return any; }
        q.state = kRunning;
        q.job = job;
        q.blockFirst = first;
        q.blockCount = nblocks;
        q.admitted = nowSeconds();
        running_.push_back(q.ticket);
        queue_.pop_front();
        any = true;
    }
    return any;
}

void Server::fail(const std::string& msg) {
    std::lock_guard<std::mutex> lk(mu_);
    failed_ = true;
    failure_ = msg;
    for (auto& kv : requests_)
        if (kv.second.state < kDone) kv.second.state = kFailed;
    queue_.clear();
    running_.clear();
    cv_.notify_all();
}

void Server::loop() {
    std::string err;
    if (poolBindThread(pool_)) { fail("cannot select the pool's device"); return; }
    for (;;) {
        bool changed = false;
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [&]() { return stop_ || !queue_.empty() || !running_.empty(); });
            if (stop_) break;
            // cancellations of running requests
            for (size_t i = 0; i < running_.size();) {
                Request& q = requests_[running_[i]];
                if (q.cancelRequested) {
                    int r = poolCancel(pool_, q.job, &err);
                    if (r) { lk.unlock(); fail(err); return; }
                    freeBlocks(q.blockFirst, q.blockCount);
                    freeJobs_.push_back(q.job);
                    q.state = kCancelled;
                    running_.erase(running_.begin() + i);
                    changed = true;
                    cv_.notify_all();
                } else {
                    ++i;
                }
            }
            int rc = 0;
            changed |= admitQueued(&err, &rc);
            if (rc) { lk.unlock(); fail(err); return; }
            if (running_.empty()) continue;
        }
        const double t0 = nowSeconds();
        int r = 0;
        if (changed) r = poolRecount(pool_, &err);
        if (!r) r = poolRunBlock(pool_, traceVariant_, &err);
        if (r) { fail(err); return; }
        const uint32_t* active = poolActiveCounts(pool_);
        std::vector<uint64_t> finished;
        {
            std::lock_guard<std::mutex> lk(mu_);
            for (uint64_t t : running_)
                if (active[requests_[t].job] == 0) finished.push_back(t);
        }
        for (uint64_t t : finished) {
            uint32_t job;
            {
                std::lock_guard<std::mutex> lk(mu_);
                job = requests_[t].job;
            }
            if ((r = poolFinish(pool_, job, &err))) { fail(err); return; }
        }
        double samples = 0, rays = 0;
        if (!finished.empty() && (r = poolSamples(pool_, &samples, &rays, &err))) { fail(err); return; }
        {
            std::lock_guard<std::mutex> lk(mu_);
            const double t1 = nowSeconds();
            busySeconds_ += t1 - t0;
            if (!finished.empty()) samplesDone_ = samples;
            for (uint64_t t : finished) {
                Request& q = requests_[t];
                q.state = kDone;
                q.done = t1;
                latencies_.push_back(q.done - q.submitted);
                freeBlocks(q.blockFirst, q.blockCount);
                freeJobs_.push_back(q.job);
                running_.erase(std::remove(running_.begin(), running_.end(), t), running_.end());
            }
            if (!finished.empty()) cv_.notify_all();
        }
    }
}

}  // namespace slrh
