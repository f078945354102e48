// Continuous-batching render server (see server.cpp).
#pragma once

#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "engine.h"

namespace slrh {

class Server {
public:
    enum State { kQueued = 0, kRunning = 1, kDone = 2, kCancelled = 3, kFailed = 4 };
    static constexpr uint64_t kSlotBytes = 320 + 4 * 8 + 2;   // slot record, queue entries, code and flag bytes
    // live paths a request without a fixed slots-per-pixel count aims for
    static constexpr uint64_t kServeSlotsPerRequest = 2u << 20;

    ~Server();
    int start(DeviceScene* ds, uint64_t pathBytes, int traceVariant, std::string* err);
    // planes: slots per pixel, 0 to choose at admission
    int submit(const JobSpec& spec, uint32_t planes, uint64_t* ticket, std::string* err);
    int state(uint64_t ticket, int* st);
    int wait(uint64_t ticket, int timeoutMs, int* st);
    int cancel(uint64_t ticket);
    // completed, queued, running, p50 ms, p99 ms, mean ms, samples filed, busy seconds
    void stats(double* out);

private:
    struct Request {
        uint64_t ticket = 0;
        JobSpec spec;
        uint32_t planes = 0, job = 0, blockFirst = 0, blockCount = 0;
        int state = kQueued;
        bool cancelRequested = false;
        double submitted = 0, admitted = 0, done = 0;
    };
    void loop();
    bool admitQueued(std::string* err, int* rc);
    uint32_t planesFor(const JobSpec& spec, uint32_t requested) const;
    bool allocBlocks(uint32_t nblocks, uint32_t* first);
    void freeBlocks(uint32_t first, uint32_t nblocks);
    void fail(const std::string& msg);

    SlotPool* pool_ = nullptr;
    int traceVariant_ = 0;
    std::thread thread_;
    std::mutex mu_;
    std::condition_variable cv_;
    bool stop_ = false, failed_ = false;
    std::string failure_;
    uint64_t lastTicket_ = 0;
    std::map<uint64_t, Request> requests_;
    std::deque<uint64_t> queue_;
    std::vector<uint64_t> running_;
    std::vector<bool> blockUsed_;
    std::vector<uint32_t> freeJobs_;
    std::vector<double> latencies_;
    double samplesDone_ = 0, busySeconds_ = 0;
};

}  // namespace slrh
