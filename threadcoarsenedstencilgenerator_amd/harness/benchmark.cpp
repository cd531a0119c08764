#include "benchmark.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <sstream>

#include "primes.h"

namespace tcsg {

[[noreturn]] static void die(const std::string &m)
{
    std::cerr << m << " Exiting!" << std::endl;
    std::exit(EXIT_FAILURE);
}

static void check(int rc, const char *what)
{
    if (rc != TCSG_OK) {
        std::cerr << what << ": " << tcsg_status_string(rc) << " (hip " << tcsg_last_hip_error() << ")"
                  << std::endl;
        std::exit(EXIT_FAILURE);
    }
}

// Parameter checks of hip/benchmark.cpp:428-489 and the stencil size / flop
// count of :62-81.
StencilBenchmark::StencilBenchmark(const BenchConfig &c) : c_(c)
{
    if (c.neighbourhood != "Box" && c.neighbourhood != "Star") die("Unknown stencil-neighbourhood: " + c.neighbourhood + ".");
    if (c.dimension != 2 && c.dimension != 3) die("Unknown dimension.");
    if (c.alignment != "None" && c.alignment != "Write" && c.alignment != "Read") die("Unknown alignment: " + c.alignment + ".");
    if (c.alignment != "None" && c.cl_size == 0) die("size of alignment is 0, but alignment is given.");
    if (c.coeff != "Jacobi" && c.coeff != "Variable") die("Unknown coefficients: " + c.coeff + ".");
    if (c.orientation != "Row" && c.orientation != "Column") die("Unknown orientation: " + c.orientation + ".");
    if (c.range == 0) die("Stencil range must be greater than zero!");
    if (!c.xc || !c.yc || (c.dimension == 3 && !c.zc)) die("Coarsening factors have to be greater than zero!");
    if (!c.size[0] || !c.size[1] || (c.dimension == 3 && !c.size[2])) die("Grid dimensions have to be greater than zero!");
    if (c.max_time < 0) die("Time for iterations must be non-negative.");
    desc_.dimension = c.dimension;
    desc_.neighbourhood = c.neighbourhood == "Box" ? TCSG_BOX : TCSG_STAR;
    desc_.orientation = c.orientation == "Row" ? TCSG_ROW : TCSG_COLUMN;
    desc_.coeff = c.coeff == "Jacobi" ? TCSG_JACOBI : TCSG_VARIABLE;
    desc_.range = c.range;
    desc_.xc = c.xc;
    desc_.yc = c.yc;
    desc_.zc = c.dimension == 3 ? c.zc : 1;
    desc_.alignment = c.alignment == "None" ? TCSG_ALIGN_NONE : c.alignment == "Read" ? TCSG_ALIGN_READ : TCSG_ALIGN_WRITE;
    desc_.cl_size = c.alignment == "None" ? 0 : c.cl_size;
    const unsigned w = 2 * c.range + 1;
    stencil_size_ = c.neighbourhood == "Box" ? (c.dimension == 2 ? w * w : w * w * w) : 2 * c.dimension * c.range + 1;
    flops_per_update_ = c.coeff == "Jacobi" ? stencil_size_ : 2 * stencil_size_ - 1;
}

StencilBenchmark::~StencilBenchmark()
{
    if (d0_) (void)hipFree(d0_);
    if (d1_) (void)hipFree(d1_);
    tcsg_kernel_free(kernel_);
}

void StencilBenchmark::init_gpu()
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) die("No gpu found.");
    if (hipSetDevice(0) != hipSuccess) die("hipSetDevice failed.");
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, 0) != hipSuccess) die("hipGetDeviceProperties failed.");
    for (int a = 0; a < 3; ++a) max_dims_[a] = p.maxThreadsDim[a];
    gpu_name_ = p.name;
    // the marketing name can be empty where the driver's name table is missing: the ISA name then
    if (gpu_name_.empty()) gpu_name_ = p.gcnArchName;
    warp_ = p.warpSize;
}

void StencilBenchmark::init_benchmark()
{
    // load the generated kernel and check it was generated for these parameters
    check(tcsg_kernel_load(&kernel_, c_.kernel_path.c_str(), nullptr), "loading kernel");
    tcsg_desc got;
    check(tcsg_kernel_desc(kernel_, &got), "kernel descriptor");
    desc_.dtype = got.dtype;
    desc_.variant = got.variant;
    if (std::memcmp(&got, &desc_, sizeof(desc_)) != 0)
        die("Generated kernel " + c_.kernel_path + " was built for different stencil parameters.");
    if (desc_.coeff == TCSG_VARIABLE) {
        // isotropic test mask as in hip/benchmark.cpp:120-129, or a seeded one
        std::vector<double> mask(stencil_size_, 1.0 / stencil_size_);
        if (const char *s = std::getenv("TCSG_MASK_SEED")) {
            const uint64_t seed = std::strtoull(s, nullptr, 10);
            double sum = 0.0;
            for (unsigned k = 0; k < stencil_size_; ++k) {
                uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ull;
                z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
                z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
                z ^= z >> 31;
                mask[k] = (double)(z >> 11) * 0x1.0p-52 - 1.0;
                sum += std::fabs(mask[k]);
            }
            for (auto &m : mask) m /= sum;
        }
        check(tcsg_set_mask(kernel_, mask.data(), stencil_size_), "copying mask");
    }
    int occ = 0, legal = 0;
    check(tcsg_max_block_size(kernel_, &occ, &legal), "block size query");
    max_threads_ = c_.reference_sweep ? occ : legal;
    tcsg_dim3 size{c_.size[0], c_.size[1], c_.dimension == 3 ? c_.size[2] : 1u};
    // the kernel's own geometry: truncated as the reference's, or whole for a
    // kernel generated with remainder handling
    check(tcsg_kernel_geometry(kernel_, size, &geom_), "geometry");
    const size_t es = desc_.dtype == TCSG_F32C ? 4 : 8;
    size_t free_b = 0, total_b = 0;
    (void)hipMemGetInfo(&free_b, &total_b);
    if (2 * es * geom_.n_points > free_b) {
        std::ostringstream m;
        m << "Requested grid size is too big for the GPU (needed/free/total: " << 2 * es * geom_.n_points << " / "
          << free_b << " / " << total_b << ").";
        die(m.str());
    }
    if (hipMalloc(&d0_, es * geom_.n_points) != hipSuccess || hipMalloc(&d1_, es * geom_.n_points) != hipSuccess)
        die("hipMalloc failed.");
    int mode = TCSG_INIT_REFERENCE;
    if (const char *s = std::getenv("TCSG_INIT")) {
        const std::string v = s;
        mode = v == "u01" ? TCSG_INIT_U01 : v == "upm1" ? TCSG_INIT_UPM1 : v == "normal" ? TCSG_INIT_NORMAL : TCSG_INIT_REFERENCE;
    }
    check(tcsg_launch_init(kernel_, &geom_, mode, 1, d0_, nullptr), "init");
    check(tcsg_launch_init(kernel_, &geom_, mode, 1, d1_, nullptr), "init");
    if (hipDeviceSynchronize() != hipSuccess) die("init failed.");
}

void StencilBenchmark::bench_shape(tcsg_dim3 block)
{
    if (tcsg_check_block(kernel_, block) != TCSG_OK) {
        std::cerr << "kernel failed for blockSize: " << block.x << "\t" << block.y << "\t" << block.z << std::endl;
        return;
    }
    const tcsg_dim3 blocks = tcsg_grid_blocks(&geom_, block);
    std::vector<double> dur(1 << 16);
    size_t n = 0;
    int rc = tcsg_benchmark_update_grid(kernel_, &geom_, block, blocks, c_.max_time, d0_, d1_, dur.data(), dur.size(), &n);
    if (rc != TCSG_OK) {
        std::cerr << "kernel failed for blockSize: " << block.x << "\t" << block.y << "\t" << block.z << std::endl;
        return;
    }
    dur.resize(std::min(n, dur.size()));
    ShapeResult r{block, dur, 0.0};
    float ms = 0.f;
    const int steps = (int)std::max<size_t>(1, std::min<size_t>(n, 64));
    if (tcsg_time_steps(kernel_, &geom_, block, blocks, d0_, d1_, 1, steps, nullptr, &ms) == TCSG_OK)
        r.event_ms_per_step = ms / steps;
    results_.push_back(r);
}

// Benchmark2D / Benchmark3D: hip/benchmark.cpp:221-249, 285-316.
void StencilBenchmark::run()
{
    for (unsigned warps = 1; warps <= max_threads_ / warp_; ++warps) {
        const unsigned T = warps * warp_;
        if (c_.dimension == 2) {
            for (auto s : shapes2d(T)) {
                if (s[0] > max_dims_[0] || s[1] > max_dims_[1]) continue;
                bench_shape({s[0], s[1], 1});
            }
        } else {
            for (auto s : shapes3d(T)) {
                if (s[0] > max_dims_[0] || s[1] > max_dims_[1] || s[2] > max_dims_[2]) continue;
                bench_shape({s[0], s[1], s[2]});
            }
        }
    }
}

// singleRun: hip/benchmark.cpp:251-281, 318-349.
void StencilBenchmark::single_run(unsigned bx, unsigned by, unsigned bz)
{
    if (!bx || !by || !bz) die("One of the dimensions of the requested threadblock is 0.");
    if (bx > max_dims_[0] || by > max_dims_[1] || bz > max_dims_[2]) die("Requested threadblock dimensions bigger than maximal possible.");
    if ((unsigned long)bx * by * bz > max_threads_) die("Threadblock is too big for this kernel.");
    bench_shape({bx, by, bz});
}

// printResults: hip/benchmark.cpp:352-426 (same header and columns; with
// TCSG_EXTRA_COLUMNS set, event-timed GLup/s and GB/s are appended).
void StencilBenchmark::print_results() const
{
    std::stringstream out;
    out << std::scientific;
    const bool extra = std::getenv("TCSG_EXTRA_COLUMNS") != nullptr;
    out << "#GPU: " << gpu_name_ << "\n";
    out << "#grid size excluding border: x " << geom_.elements.x << " y " << geom_.elements.y;
    if (c_.dimension == 3) out << " z " << geom_.elements.z;
    out << "\n#coarsening factors: x " << c_.xc << " y " << c_.yc;
    if (c_.dimension == 3) out << " z " << c_.zc;
    out << "\n#alignment: " << c_.alignment << "\n";
    if (c_.alignment != "None") out << "#aligned to " << c_.cl_size * sizeof(double) << " byte\n";
    out << "#orientation: " << c_.orientation << "\n#coefficients: " << c_.coeff << "\n#neighbourhood: "
        << c_.neighbourhood << "\n#range: " << c_.range << "\n#dimension: " << c_.dimension << "\n";
    out << "#xthreads\tythreads\tzthreads\tPerf [Flop/s]\tPerf dev [Flop/s]\tPerf [Lup/s]\tPerf dev [Lup/s]\titerations\ttotal time[s]";
    if (extra) out << "\tevent Perf [Lup/s]\tevent [GB/s]";
    out << "\n";
    const double pts = (double)geom_.elements.x * geom_.elements.y * (c_.dimension == 3 ? geom_.elements.z : 1);
    const double bytes_per_lup = desc_.dtype == TCSG_F32C ? 8.0 : 16.0;
    for (const auto &r : results_) {
        double mean = 0.0, total = 0.0;
        for (double t : r.durations) { mean += pts / t; total += t; }
        mean /= r.durations.size();
        double sd = 0.0;
        for (double t : r.durations) sd += (pts / t - mean) * (pts / t - mean);
        sd = std::sqrt(sd / r.durations.size());
        out << r.block.x << "\t" << r.block.y << "\t" << (c_.dimension == 3 ? r.block.z : 0) << "\t"
            << mean * flops_per_update_ << "\t" << sd * flops_per_update_ << "\t" << mean << "\t" << sd << "\t"
            << r.durations.size() << "\t" << total;
        if (extra) {
            const double lups = r.event_ms_per_step > 0 ? pts / (r.event_ms_per_step * 1e-3) : 0.0;
            out << "\t" << lups << "\t" << lups * bytes_per_lup * 1e-9;
        }
        out << "\n";
    }
    std::cout << out.rdbuf();
}

} // namespace tcsg
