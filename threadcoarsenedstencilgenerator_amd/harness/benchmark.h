// Benchmark harness over the tcsg C ABI: the MI355X counterpart of the
// reference's class stencilBenchmark (hip/benchmark.h:18-124).
#pragma once
#include <string>
#include <utility>
#include <vector>

#include "tcsg.h"

namespace tcsg {

struct BenchConfig {
    std::string neighbourhood, alignment, coeff, orientation;
    unsigned dimension = 0, cl_size = 0, range = 0;
    unsigned xc = 1, yc = 1, zc = 1;
    unsigned size[3] = {0, 0, 1};
    double max_time = 0.0;
    std::string kernel_path;       // generated code object
    bool reference_sweep = false;  // bound the sweep by the occupancy-optimal size (reference behaviour)
};

struct ShapeResult {
    tcsg_dim3 block;
    std::vector<double> durations;  // per-step host-synchronised durations [s]
    double event_ms_per_step = 0.0; // hipEvent time per step, back-to-back launches
};

class StencilBenchmark {
  public:
    explicit StencilBenchmark(const BenchConfig &c);  // checkParameters + setTypeAndSize
    ~StencilBenchmark();
    void init_gpu();                                   // initGPU
    void init_benchmark();                             // initBenchmark
    void run();                                        // Benchmark2D / Benchmark3D
    void single_run(unsigned bx, unsigned by, unsigned bz);  // singleRun
    void print_results() const;                        // printResults

  private:
    void bench_shape(tcsg_dim3 block);
    BenchConfig c_;
    tcsg_desc desc_{};
    unsigned stencil_size_ = 0, flops_per_update_ = 0;
    tcsg_kernel *kernel_ = nullptr;
    tcsg_geometry geom_{};
    void *d0_ = nullptr, *d1_ = nullptr;
    std::string gpu_name_;
    unsigned max_threads_ = 0, warp_ = 64, max_dims_[3] = {1024, 1024, 1024};
    std::vector<ShapeResult> results_;
};

} // namespace tcsg
