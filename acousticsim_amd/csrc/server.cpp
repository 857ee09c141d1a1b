#include "server.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <stdexcept>

#include "status.h"

namespace bw
{

namespace
{
void check(hipError_t e, const char *what)
{
    if (e != hipSuccess)
        throw std::runtime_error(std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
}

// Device bytes a one-cloth Batch will allocate, from the model alone (an
// upper bound: the ELL slots are counted at the widest row).
size_t estimate_bytes(const ClothModel &m)
{
    int wmax = 0;
    for (int v = 0; v < m.nv; v++)
        wmax = std::max(wmax, m.row_ptr[v + 1] - m.row_ptr[v]);
    const size_t slices = ((size_t)(m.nv + SLICE - 1) / SLICE + BW_SB - 1) / BW_SB * BW_SB;
    const size_t rows = slices * SLICE;
    const size_t entries = rows * (size_t)wmax;
    const size_t nie = m.interior.size();
    size_t b = entries * (4 + 72 + 36);                       // col, W64, W32
    b += rows * (3 * 8 * 5 + 3 * 4 + 3 * 8 * 5 + 8 + 4 * 3); // x..P64, P32, solver vectors, mass, row ints
    // slice ints, slice_info, partials, launch list entries (one per slice
    // and one per solver workgroup)
    b += slices * (3 * 4 + 16 + 2 * PARTS * 8 + sizeof(SolverEntry)) +
         slices / BW_SB * sizeof(SolverEntry);
    b += COL_PAD * 4;
    b += (size_t)m.nv * 4 * 4 + (size_t)m.nnzb() * 4 * 3;
    b += ((size_t)9 * m.nt + 16 * nie) * 4 + ((size_t)3 * m.nt + 4 * nie) * 4 * 2;
    b += (size_t)m.nt * (3 * 4 + 4 + 12 * 4 + 4 + 25 * 8) + nie * (4 * 4 + 8 + 25 * 8);
    return b + (1 << 20);
}

// list launch grids: 64 * {1, 1.5, 2, 3, 4, 6, ...}
int bucket(int n)
{
    int b = 64;
    while (b < n)
        b = (b & (b - 1)) == 0 ? b / 2 * 3 : b / 3 * 4;
    return b;
}
} // namespace

Server::Server(const ServerOptions &opt) : opt_(opt), t0_(std::chrono::steady_clock::now())
{
    if (opt_.max_resident < 1 || opt_.max_resident > 65536)
        throw std::runtime_error("config: max_resident must be in [1, 65536]");
    if (const char *e = std::getenv("BW_SPMV_STAGE"))
        opt_.spmv_stage = std::atoi(e) != 0;
    if (const char *e = std::getenv("BW_SPMV_NT"))
        opt_.spmv_nt = std::atoi(e) != 0;
    check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    const size_t ns = 2 * (size_t)opt_.max_resident;
    check(hipMalloc((void **)&states_dev_, ns * sizeof(CgState)), "hipMalloc");
    check(hipHostMalloc((void **)&states_host_, ns * sizeof(CgState)), "hipHostMalloc");
    check(hipMalloc((void **)&table_dev_, opt_.max_resident * sizeof(DeviceBatch)), "hipMalloc");
    check(hipHostMalloc((void **)&table_host_, opt_.max_resident * sizeof(DeviceBatch)),
          "hipHostMalloc");
    check(hipMemset(states_dev_, 0, ns * sizeof(CgState)), "hipMemset");
    check(hipStreamSynchronize(nullptr), "hipStreamSynchronize"); // before the server stream
    for (int s = opt_.max_resident - 1; s >= 0; s--)
        free_slots_.push_back(s);
    for (int k = 0; k < std::max(1, opt_.builders); k++)
        builders_.emplace_back([this] { builder_loop(); });
    device_ = std::thread([this] { device_loop(); });
}

Server::~Server()
{
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_jobs_.notify_all();
    cv_ready_.notify_all();
    cv_budget_.notify_all();
    cv_done_.notify_all();
    for (auto &t : builders_)
        if (t.joinable())
            t.join();
    if (device_.joinable())
        device_.join();
    ready_.clear();
    for (auto &g : graphs_)
        (void)hipGraphExecDestroy(g.exec);
    (void)hipStreamSynchronize(stream_);
    (void)hipFree(states_dev_);
    (void)hipFree(table_dev_);
    (void)hipFree(spmv_list_dev_);
    (void)hipFree(upd_list_dev_);
    (void)hipHostFree(states_host_);
    (void)hipHostFree(table_host_);
    (void)hipHostFree(spmv_list_host_);
    (void)hipHostFree(upd_list_host_);
    (void)hipFree(grp_list_dev_);
    (void)hipHostFree(grp_list_host_);
    (void)hipStreamDestroy(stream_);
}

double Server::now() const
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
}

long long Server::submit(const std::string &json, const double *x0, int nx0, int nsteps)
{
    if (nsteps < 1)
        throw std::runtime_error("config: a request needs at least one step");
    Job job;
    job.json = json;
    if (x0 && nx0 > 0)
        job.x0.assign(x0, x0 + nx0);
    job.nsteps = nsteps;
    std::lock_guard<std::mutex> lk(mu_);
    if (stop_)
        throw std::runtime_error("server: stopped");
    job.id = next_id_++;
    RequestInfo &ri = info_[job.id];
    ri.steps = nsteps;
    ri.t_submit = now();
    counters_.queued++;
    jobs_.push_back(std::move(job));
    cv_jobs_.notify_one();
    return jobs_.back().id;
}

bool Server::wait(long long id, double timeout_s)
{
    std::unique_lock<std::mutex> lk(mu_);
    auto it = info_.find(id);
    if (it == info_.end())
        throw std::runtime_error("config: unknown request id");
    return cv_done_.wait_for(lk, std::chrono::duration<double>(timeout_s), [&] {
        auto j = info_.find(id);
        return j == info_.end() || j->second.done;
    });
}

RequestInfo Server::info(long long id)
{
    std::lock_guard<std::mutex> lk(mu_);
    auto it = info_.find(id);
    if (it == info_.end())
        throw std::runtime_error("config: unknown request id");
    return it->second;
}

void Server::release(long long id)
{
    std::lock_guard<std::mutex> lk(mu_);
    auto it = info_.find(id);
    if (it == info_.end())
        throw std::runtime_error("config: unknown request id");
    if (!it->second.done)
        throw std::runtime_error("server: request still running");
    info_.erase(it);
}

ServerCounters Server::counters()
{
    std::lock_guard<std::mutex> lk(mu_);
    ServerCounters c = counters_;
    c.bytes_in_use = bytes_in_use_;
    return c;
}

void Server::builder_loop()
{
    for (;;)
    {
        Job job;
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_jobs_.wait(lk, [&] { return stop_ || !jobs_.empty(); });
            if (stop_)
                return;
            job = std::move(jobs_.front());
            jobs_.pop_front();
        }
        int slot = -1;
        size_t reserved = 0;
        try
        {
            ClothModel m = ClothModel::from_json_text(job.json);
            if (m.cfg.solve_mode == SolveMode::Reference || m.cfg.gpu_form)
                throw std::runtime_error("config: served requests use the tolerance solver "
                                         "(not \"reference\" mode or \"fem_gpu\")");
            if ((m.cfg.precision == Precision::Mixed) != opt_.fp32)
                throw std::runtime_error(std::string("config: this server runs \"") +
                                         (opt_.fp32 ? "mixed" : "fp64") + "\" requests");
            if (!job.x0.empty() && job.x0.size() != 3 * (size_t)m.nv)
                throw std::runtime_error("config: x0 must hold 3 doubles per vertex");
            const size_t est = estimate_bytes(m);
            if (est > opt_.budget_bytes)
                throw std::runtime_error("capacity: request needs more device memory than the budget");
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_budget_.wait(lk, [&] {
                    return stop_ || (bytes_in_use_ + est <= opt_.budget_bytes && !free_slots_.empty());
                });
                if (stop_)
                    return;
                slot = free_slots_.back();
                free_slots_.pop_back();
                bytes_in_use_ += est;
                reserved = est;
            }
            const int nv = m.nv;
            BatchOptions bo;
            bo.stream = stream_;
            bo.cg_states = states_dev_ + 2 * (size_t)slot;
            bo.shared_launches = true;
            std::vector<ClothModel> ms;
            ms.push_back(std::move(m));
            auto r = std::make_unique<Resident>();
            r->id = job.id;
            r->slot = slot;
            r->nsteps = job.nsteps;
            {
                // allocations and copies on this thread would invalidate a
                // graph the device thread is capturing
                std::lock_guard<std::mutex> cap(capture_mu_);
                r->batch = std::make_unique<Batch>(std::move(ms), bo);
                if (!job.x0.empty())
                {
                    r->batch->set_pos(0, job.x0.data());
                    r->batch->set_pre_pos(0, job.x0.data());
                }
            }
            const size_t actual = r->batch->device_bytes();
            {
                std::lock_guard<std::mutex> lk(mu_);
                bytes_in_use_ = bytes_in_use_ - reserved + actual;
                reserved = actual;
                RequestInfo &ri = info_[job.id];
                ri.t_admit = now();
                ri.nv = nv;
                ri.device_bytes = actual;
                ri.admitted_bytes = est;
                counters_.queued--;
                counters_.resident++;
                counters_.max_resident_seen =
                    std::max(counters_.max_resident_seen, counters_.resident);
                ready_.push_back(std::move(r));
            }
            cv_ready_.notify_one();
        }
        catch (...)
        {
            std::string msg;
            const int code = status_of_current_exception(msg);
            std::lock_guard<std::mutex> lk(mu_);
            if (slot >= 0)
            {
                free_slots_.push_back(slot);
                bytes_in_use_ -= reserved;
            }
            RequestInfo &ri = info_[job.id];
            ri.status = code;
            ri.error = msg;
            ri.done = true;
            ri.t_done = now();
            counters_.queued--;
            cv_done_.notify_all();
            cv_budget_.notify_all();
        }
    }
}

void Server::upload_table(const std::vector<std::unique_ptr<Resident>> &res)
{
    for (auto &r : res)
        table_host_[r->slot] = r->batch->device();
    // the whole table: slots are few and small
    int top = 0;
    for (auto &r : res)
        top = std::max(top, r->slot + 1);
    if (top > 0)
        check(hipMemcpyAsync(table_dev_, table_host_, top * sizeof(DeviceBatch),
                             hipMemcpyHostToDevice, stream_),
              "hipMemcpyAsync table");
    // one SpMV form per launch: staged when every resident request has its
    // column lists
    int hmax = 0;
    bool all = opt_.spmv_stage != 0;
    for (auto &r : res)
    {
        all = all && r->batch->halo_max() > 0;
        hmax = std::max(hmax, r->batch->halo_lds());
    }
    halo_ = 0;
    if (all && hmax > 0)
        for (halo_ = 64; halo_ < hmax;)
            halo_ *= 2;
}

void Server::launch_chain(int length, int spmv_grid, int grp_grid)
{
    const int ugrid = (spmv_grid + BW_SB - 1) / BW_SB;
    for (int k = 0; k < length; k++)
        launch_cg_step_multi(table_dev_, spmv_list_dev_, spmv_grid, upd_list_dev_, ugrid,
                             grp_list_dev_, grp_grid, opt_.fp32, k & 1, opt_.spmv_lpr, halo_,
                             opt_.spmv_nt != 0, 1, stream_);
}

void Server::run_chunk(int length, int spmv_grid, int grp_grid)
{
    // BW_GRAPHS=0: direct launches (profilers that do not follow graphs)
    static const bool graphs = [] {
        const char *g = std::getenv("BW_GRAPHS");
        return !(g && std::string(g) == "0");
    }();
    if (!graphs)
    {
        launch_chain(length, spmv_grid, grp_grid);
        check(hipGetLastError(), "solver launch");
        return;
    }
    // graphs of power-of-two lengths, cached by (length, grids)
    for (int piece = 1 << 30; piece >= 32; piece >>= 1)
    {
        if (!(length & piece))
            continue;
        ++graph_clock_;
        hipGraphExec_t exec = nullptr;
        for (auto &g : graphs_)
            if (g.length == piece && g.spmv_grid == spmv_grid && g.grp_grid == grp_grid &&
                g.halo == halo_)
            {
                exec = g.exec;
                g.last_use = graph_clock_;
            }
        if (!exec)
        {
            if (graphs_.size() >= 128)
            {
                auto lru = std::min_element(graphs_.begin(), graphs_.end(),
                                            [](const GraphEntry &a, const GraphEntry &b) {
                                                return a.last_use < b.last_use;
                                            });
                check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
                (void)hipGraphExecDestroy(lru->exec);
                graphs_.erase(lru);
            }
            hipGraph_t graph;
            {
                std::lock_guard<std::mutex> cap(capture_mu_);
                check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal),
                      "BeginCapture");
                launch_chain(piece, spmv_grid, grp_grid);
                check(hipStreamEndCapture(stream_, &graph), "EndCapture");
            }
            check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "GraphInstantiate");
            (void)hipGraphDestroy(graph);
            graphs_.push_back({piece, spmv_grid, grp_grid, halo_, exec, graph_clock_});
        }
        check(hipGraphLaunch(exec, stream_), "GraphLaunch");
    }
    if (length & 31)
        launch_chain(length & 31, spmv_grid, grp_grid);
}

void Server::retire(Resident &r, int error_code, const std::string &error)
{
    std::vector<double> pos;
    int code = error_code;
    std::string msg = error;
    if (code == BW_OK)
    {
        try
        {
            pos.resize(3 * (size_t)r.batch->model(0).nv);
            r.batch->get_pos(0, pos.data());
        }
        catch (...)
        {
            code = status_of_current_exception(msg);
        }
    }
    const size_t bytes = r.batch->device_bytes();
    r.batch.reset();
    std::lock_guard<std::mutex> lk(mu_);
    RequestInfo &ri = info_[r.id];
    ri.status = code;
    ri.error = msg;
    ri.pos = std::move(pos);
    ri.done = true;
    ri.t_done = now();
    free_slots_.push_back(r.slot);
    bytes_in_use_ -= bytes;
    counters_.resident--;
    counters_.completed++;
    cv_done_.notify_all();
    cv_budget_.notify_all();
}

int Server::run_round(std::vector<std::unique_ptr<Resident>> &res)
{
    // start the next step of every resident request that is not solving
    for (auto &r : res)
        if (!r->solving)
        {
            r->batch->begin_step();
            r->solving = true;
            r->launched = 0;
            r->t_step0 = now();
            if (r->steps_done == 0)
            {
                std::lock_guard<std::mutex> lk(mu_);
                info_[r->id].t_first = r->t_step0;
            }
        }
    // lists of every solving request's slices and solver workgroups
    int n = 0, nb = 0, ng = 0;
    for (auto &r : res)
    {
        n += r->batch->device().n_slices;
        nb += r->batch->device().n_blocks_wg;
        ng += r->batch->device().n_groups;
    }
    const int grid = bucket(n), ugrid = (grid + BW_SB - 1) / BW_SB;
    const int ggrid = ng > 0 ? bucket(ng) : 0;
    if (ggrid > grp_cap_)
    {
        check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
        for (auto &g : graphs_)
            (void)hipGraphExecDestroy(g.exec);
        graphs_.clear();
        (void)hipFree(grp_list_dev_);
        (void)hipHostFree(grp_list_host_);
        grp_cap_ = std::max(ggrid, 2 * grp_cap_);
        check(hipMalloc((void **)&grp_list_dev_, grp_cap_ * sizeof(int2)), "hipMalloc");
        check(hipHostMalloc((void **)&grp_list_host_, grp_cap_ * sizeof(int2)), "hipHostMalloc");
    }
    if (grid > list_cap_)
    {
        // grow the lists; graphs hold the old addresses
        check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
        for (auto &g : graphs_)
            (void)hipGraphExecDestroy(g.exec);
        graphs_.clear();
        (void)hipFree(spmv_list_dev_);
        (void)hipFree(upd_list_dev_);
        (void)hipHostFree(spmv_list_host_);
        (void)hipHostFree(upd_list_host_);
        list_cap_ = std::max(grid, 2 * list_cap_);
        check(hipMalloc((void **)&spmv_list_dev_, list_cap_ * sizeof(int2)), "hipMalloc");
        check(hipMalloc((void **)&upd_list_dev_, list_cap_ * sizeof(int2)), "hipMalloc");
        check(hipHostMalloc((void **)&spmv_list_host_, list_cap_ * sizeof(int2)), "hipHostMalloc");
        check(hipHostMalloc((void **)&upd_list_host_, list_cap_ * sizeof(int2)), "hipHostMalloc");
    }
    int k = 0, kb = 0;
    for (auto &r : res)
    {
        const DeviceBatch &d = r->batch->device();
        for (int s = 0; s < d.n_slices; s++)
            spmv_list_host_[k++] = make_int2(r->slot, s);
        for (int b = 0; b < d.n_blocks_wg; b++)
            upd_list_host_[kb++] = make_int2(r->slot, b);
    }
    for (; k < grid; k++)
        spmv_list_host_[k] = make_int2(-1, 0);
    for (; kb < ugrid; kb++)
        upd_list_host_[kb] = make_int2(-1, 0);
    int kg = 0;
    for (auto &r : res)
        for (int g = 0; g < r->batch->device().n_groups; g++)
            grp_list_host_[kg++] = make_int2(r->slot, g);
    for (; kg < ggrid; kg++)
        grp_list_host_[kg] = make_int2(-1, 0);
    if (ggrid > 0)
        check(hipMemcpyAsync(grp_list_dev_, grp_list_host_, ggrid * sizeof(int2),
                             hipMemcpyHostToDevice, stream_),
              "hipMemcpyAsync list");
    check(hipMemcpyAsync(spmv_list_dev_, spmv_list_host_, grid * sizeof(int2),
                         hipMemcpyHostToDevice, stream_),
          "hipMemcpyAsync list");
    check(hipMemcpyAsync(upd_list_dev_, upd_list_host_, ugrid * sizeof(int2),
                         hipMemcpyHostToDevice, stream_),
          "hipMemcpyAsync list");
    // run until the request expected to finish first should be done
    long long want = 1LL << 40;
    for (auto &r : res)
        want = std::min(want, r->predicted > r->launched ? r->predicted - r->launched : 64LL);
    const int length = (int)std::min<long long>(1024, std::max<long long>(32, (want + 31) / 32 * 32));
    run_chunk(length, grid, ggrid);
    for (auto &r : res)
        r->launched += length;
    // every slot's states in one copy
    int top = 0;
    for (auto &r : res)
        top = std::max(top, r->slot + 1);
    check(hipMemcpyAsync(states_host_, states_dev_, 2 * (size_t)top * sizeof(CgState),
                         hipMemcpyDeviceToHost, stream_),
          "hipMemcpyAsync states");
    check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    {
        std::lock_guard<std::mutex> lk(mu_);
        counters_.chunks++;
        counters_.launches += length;
    }
    // finish the steps whose solve is done
    std::vector<Resident *> finished, failed;
    for (auto &r : res)
    {
        const CgState &st = states_host_[2 * r->slot];
        if (!st.done)
        {
            const long long cap = 4LL * r->batch->model(0).cfg.solver_max_iters + 64;
            if (r->launched > cap)
                failed.push_back(r.get());
            continue;
        }
        try
        {
            r->batch->end_step(states_host_ + 2 * r->slot, 1e6 * (now() - r->t_step0));
        }
        catch (...)
        {
            failed.push_back(r.get());
            continue;
        }
        r->predicted = 2LL * (st.total + st.updates) + 2;
        r->solving = false;
        r->steps_done++;
        {
            std::lock_guard<std::mutex> lk(mu_);
            RequestInfo &ri = info_[r->id];
            ri.steps_done = r->steps_done;
            ri.iterations += st.total;
            const auto &s = r->batch->stats(0);
            ri.max_residual = std::max(ri.max_residual, s.rel_residual);
        }
        if (r->steps_done == r->nsteps)
            finished.push_back(r.get());
    }
    if (!finished.empty() || !failed.empty())
    {
        check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
        for (Resident *r : finished)
            retire(*r, BW_OK, "");
        for (Resident *r : failed)
            retire(*r, BW_ERR_NUMERIC, "step: solver did not terminate or produced NaN/Inf");
        res.erase(std::remove_if(res.begin(), res.end(),
                                 [](const std::unique_ptr<Resident> &r) { return !r->batch; }),
                  res.end());
    }
    return (int)res.size();
}

void Server::device_loop()
{
    std::vector<std::unique_ptr<Resident>> res;
    for (;;)
    {
        bool changed = false;
        {
            std::unique_lock<std::mutex> lk(mu_);
            if (res.empty())
                cv_ready_.wait(lk, [&] { return stop_ || !ready_.empty(); });
            if (stop_)
                break;
            while (!ready_.empty())
            {
                res.push_back(std::move(ready_.front()));
                ready_.pop_front();
                changed = true;
            }
        }
        try
        {
            if (changed)
                upload_table(res);
            run_round(res);
        }
        catch (...)
        {
            // a device error ends every resident request
            std::string msg;
            const int code = status_of_current_exception(msg);
            for (auto &r : res)
                if (r->batch)
                    retire(*r, code == BW_OK ? BW_ERR_STATE : code, msg);
            res.clear();
        }
    }
    for (auto &r : res)
        r->batch.reset();
}

} // namespace bw
