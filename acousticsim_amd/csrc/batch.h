// Device-resident batch of independent cloths and the host driver of one
// implicit substep (cBaraffCloth::UpdatePos, BaraffCloth.cpp:90-163).
#pragma once
#include <memory>
#include <string>
#include <vector>

#include "device.h"
#include "model.h"

namespace bw
{

struct StepStats
{
    int pcg_iterations = 0;
    int outer_iterations = 0;
    double rel_residual = 0;
    double step_us = 0;
};

class Batch
{
public:
    explicit Batch(std::vector<ClothModel> models);
    ~Batch();
    Batch(const Batch &) = delete;
    Batch &operator=(const Batch &) = delete;

    int size() const { return (int)models_.size(); }
    const ClothModel &model(int i) const { return models_[i]; }
    // pins changed on the host model: re-upload the pinned mask
    void refresh_pins(int i, const std::vector<int> &pins);

    void step(int nsteps);
    // assemble W and rhs at the current state (no solve) and copy them out in
    // the model's block order
    void assemble_system(int i, double *W, double *rhs);
    void get_pos(int i, double *x) const;
    void get_pre_pos(int i, double *x) const;
    void set_pos(int i, const double *x);
    void set_pre_pos(int i, const double *x);
    const StepStats &stats(int i) const { return stats_[i]; }
    void set_graphs(bool on) { use_graphs_ = on && !graphs_disabled_; }
    // mean microseconds per SpMV launch over reps launches; algorithmic bytes
    double time_spmv(int reps, double *alg_bytes);

    // device layout (for tests and tools)
    const DeviceBatch &device() const { return d_; }
    int row_base(int i) const { return row_base_[i]; }

private:
    void build_layout();
    void upload_state(int i, const double *src, double *dst_dev);
    void download(int i, const double *src_dev, double *dst) const;
    void read_states();
    bool all(bool (*pred)(const InstState &)) const;
    void run_chunk(bool fp32, int iters);

    std::vector<ClothModel> models_;
    std::vector<int> row_base_, slice_base_, nslices_;
    std::vector<StepStats> stats_;
    DeviceBatch d_;
    hipStream_t stream_ = nullptr;
    InstState *host_state_ = nullptr; // pinned
    bool mixed_ = false, any_ref_ = false;
    bool use_graphs_ = true;
    bool graphs_disabled_ = false; // BW_GRAPHS=0 in the environment
    int last_iters_ = 0;
    // cached chunk graphs: key (fp32, iterations)
    struct GraphEntry
    {
        bool fp32;
        int iters;
        hipGraphExec_t exec;
    };
    std::vector<GraphEntry> graphs_;
    std::vector<void *> allocations_;
    template <typename T> T *dalloc(size_t n);
};

} // namespace bw
