// Serving runtime: step requests of different cloths arrive at any time, are
// admitted against a device-memory budget, stay resident on the device for
// their steps, and are solved together with whatever else is resident
// (continuous batching at the granularity of solver chunks).
//
// Threads: submit() enqueues and returns at once.  Builder threads parse
// each request's configuration, builds its model on the host, waits for
// budget and a state slot, and uploads it (a Batch of one cloth).  A device
// thread owns the GPU stream: it starts each resident request's next step
// (element and assembly kernels, solver reset), runs the PCG launches of all
// solving requests as one multi-batch launch sequence, reads back every
// request's solver state in one copy, finishes the steps whose solve is done
// (position update), and retires requests that have done all their steps
// (positions copied to the host, device memory returned to the budget).
// A request's arithmetic does not depend on what else is resident: its
// results are bit-identical alone and in any mix.
#pragma once
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "batch.h"

namespace bw
{

struct ServerOptions
{
    size_t budget_bytes = (size_t)200 << 30; // device memory for resident requests
    int max_resident = 256;                  // state slots
    bool fp32 = true;                        // "mixed" (true) or "fp64" requests
    int builders = 4;                        // host threads building requests
    int spmv_lpr = 1;                        // SpMV lanes per row (fastest on serving mixes)
    // SpMV directions staged in LDS (1) or gathered (0); BW_SPMV_STAGE
    // overrides.  Staged: 42.1 against 38.9 cloth-steps/s on the serving mix
    // of 64 requests (half Delaunay cloths), same requests, same device
    int spmv_stage = 1;
    // matrix loads with the nontemporal hint; BW_SPMV_NT overrides.  The
    // resident requests' matrices together exceed the Infinity Cache: 48.7-49.5
    // against 46.1-46.4 cloth-steps/s on the serving mix
    int spmv_nt = 1;
};

struct RequestInfo
{
    int status = 0;        // BW_OK, or the error of a failed request
    bool done = false;     // all steps taken (or failed)
    int nv = 0;            // vertices
    int steps = 0;         // steps asked
    int steps_done = 0;
    long long iterations = 0; // PCG iterations over all steps
    double max_residual = 0;  // worst final relative residual over the steps
    // seconds on the server's clock (steady, from server creation)
    double t_submit = 0, t_admit = 0, t_first = 0, t_done = 0;
    size_t device_bytes = 0;
    size_t admitted_bytes = 0; // the estimate admission reserved
    std::string error;
    std::vector<double> pos; // final positions (3 nv), once done
};

struct ServerCounters
{
    int queued = 0;    // submitted, not yet resident
    int resident = 0;  // on the device
    int completed = 0;
    size_t bytes_in_use = 0;
    long long chunks = 0;    // multi-batch launch chunks run
    long long launches = 0;  // solver launches in them
    int max_resident_seen = 0;
};

class Server
{
public:
    explicit Server(const ServerOptions &opt);
    ~Server();
    Server(const Server &) = delete;
    Server &operator=(const Server &) = delete;

    // returns the request id (> 0)
    long long submit(const std::string &json, const double *x0, int nx0, int nsteps);
    // true when the request is done (or failed) within timeout_s
    bool wait(long long id, double timeout_s);
    // copy of a request's record (positions included once done)
    RequestInfo info(long long id);
    void release(long long id);
    ServerCounters counters();
    double now() const;

private:
    struct Job
    {
        long long id;
        std::string json;
        std::vector<double> x0;
        int nsteps;
    };
    struct Resident
    {
        long long id;
        int slot;
        std::unique_ptr<Batch> batch;
        int nsteps = 0, steps_done = 0;
        bool solving = false;
        long long launched = 0;   // solver launches this step
        long long predicted = 0;  // launches the previous step needed
        double t_step0 = 0;
    };

    void builder_loop();
    void device_loop();
    int run_round(std::vector<std::unique_ptr<Resident>> &res);
    // final positions to the host (when error_code is BW_OK), device memory
    // and slot back, waiters woken
    void retire(Resident &r, int error_code, const std::string &error);
    void upload_table(const std::vector<std::unique_ptr<Resident>> &res);

    ServerOptions opt_;
    hipStream_t stream_ = nullptr;
    std::chrono::steady_clock::time_point t0_;

    // device tables
    CgState *states_dev_ = nullptr, *states_host_ = nullptr; // 2 per slot
    DeviceBatch *table_dev_ = nullptr, *table_host_ = nullptr; // 1 per slot
    int2 *spmv_list_dev_ = nullptr, *spmv_list_host_ = nullptr;
    int2 *upd_list_dev_ = nullptr, *upd_list_host_ = nullptr;
    int list_cap_ = 0; // entries of each list
    int2 *grp_list_dev_ = nullptr, *grp_list_host_ = nullptr; // (slot, group) of large cloths
    int grp_cap_ = 0;

    std::mutex mu_;
    std::mutex capture_mu_; // graph capture vs. device allocations on the builder thread
    std::condition_variable cv_jobs_, cv_ready_, cv_done_, cv_budget_;
    std::deque<Job> jobs_;
    std::deque<std::unique_ptr<Resident>> ready_;
    std::map<long long, RequestInfo> info_;
    std::vector<int> free_slots_;
    size_t bytes_in_use_ = 0;
    long long next_id_ = 1;
    bool stop_ = false;
    ServerCounters counters_;

    struct GraphEntry
    {
        int length, spmv_grid, grp_grid, halo;
        hipGraphExec_t exec;
        unsigned long long last_use;
    };
    std::vector<GraphEntry> graphs_;
    unsigned long long graph_clock_ = 0;
    // staged SpMV: LDS entries per workgroup (the longest column list among
    // the resident requests, rounded up to a power of two; 0: gathered)
    int halo_ = 0;
    void run_chunk(int length, int spmv_grid, int grp_grid);
    void launch_chain(int length, int spmv_grid, int grp_grid);

    std::vector<std::thread> builders_;
    std::thread device_;
};

} // namespace bw
