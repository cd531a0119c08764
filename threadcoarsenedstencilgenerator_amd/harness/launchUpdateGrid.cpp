// The reference's generated launch interface (hip/generateLaunchUpdateGrid.py)
// re-expressed over libtcsg: one process-wide kernel, as in the reference,
// where kernels.cu is compiled into the harness.
#include "launchUpdateGrid.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>

#include "tcsg.h"

namespace {
tcsg_kernel *g_kernel = nullptr;

tcsg_kernel *kernel()
{
    if (!g_kernel) {
        const char *p = std::getenv("TCSG_KERNEL");
        const int rc = tcsg_kernel_load(&g_kernel, p ? p : "kernels.co", nullptr);
        if (rc != TCSG_OK) {
            std::fprintf(stderr, "GPUassert: %s (loading %s)\n", tcsg_status_string(rc), p ? p : "kernels.co");
            std::exit(EXIT_FAILURE);
        }
    }
    return g_kernel;
}

// the geometry a padded gridSize and rightOffset describe for this kernel
bool geometry_of(const tcsg_desc &d, dim3 gs, unsigned ro, tcsg_geometry &g)
{
    const unsigned lo = (d.alignment == TCSG_ALIGN_WRITE && d.range % d.cl_size) ? d.cl_size - d.range % d.cl_size : 0;
    if (gs.x < lo + ro) return false;
    tcsg_dim3 user{gs.x - lo - ro, gs.y, gs.z};
    if (tcsg_kernel_geometry(kernel(), user, &g) != TCSG_OK) return false;
    return g.size.x == gs.x && g.size.y == gs.y && (d.dimension == 2 || g.size.z == gs.z) && g.right_offset == ro;
}

void report(int rc, int line)
{
    if (rc != TCSG_OK) std::fprintf(stderr, "GPUassert: %s %s %d\n", tcsg_status_string(rc), __FILE__, line);
}
} // namespace

// hip/generateLaunchUpdateGrid.py:103-147: sanity check of the compiled-in
// parameters (throws hipErrorInvalidValue), then the doubling timing loop.
std::vector<double> benchmarkUpdateGrid(std::string coeffType, std::string stencilType, unsigned int stencilRange,
                                        unsigned int dimension, std::string alignment, double maxTime,
                                        dim3 gridSize, dim3 blockSize, dim3 gridBlocks, double *d_array0,
                                        double *d_array1, unsigned int rightOffset)
{
    tcsg_desc d;
    tcsg_kernel_desc(kernel(), &d);
    const char *al[] = {"None", "Read", "Write"};
    bool ok = coeffType == (d.coeff == TCSG_JACOBI ? "Jacobi" : "Variable") &&
              stencilType == (d.neighbourhood == TCSG_BOX ? "Box" : "Star") && stencilRange == d.range &&
              dimension == d.dimension && alignment == al[d.alignment] && d.dtype == TCSG_F64;
#ifdef TCSG_EXPECT_RANGE
    // parameters fixed by generateLaunchUpdateGrid.py (hip/generateLaunchUpdateGrid.py:106-119)
    ok = ok && coeffType == TCSG_EXPECT_COEFF && stencilType == TCSG_EXPECT_NEIGHBOURHOOD &&
         stencilRange == TCSG_EXPECT_RANGE && dimension == TCSG_EXPECT_DIMENSION && alignment == TCSG_EXPECT_ALIGNMENT;
#endif
    tcsg_geometry g;
    if (!ok || !geometry_of(d, gridSize, rightOffset, g)) throw(hipErrorInvalidValue);
    std::vector<double> dur(1 << 16);
    size_t n = 0;
    const int rc = tcsg_benchmark_update_grid(kernel(), &g, {blockSize.x, blockSize.y, blockSize.z},
                                              {gridBlocks.x, gridBlocks.y, gridBlocks.z}, maxTime, d_array0,
                                              d_array1, dur.data(), dur.size(), &n);
    if (rc == TCSG_ERR_BLOCK_SHAPE || rc == TCSG_ERR_ARG) throw(hipErrorInvalidValue);
    report(rc, __LINE__);
    dur.resize(rc == TCSG_OK ? std::min(n, dur.size()) : 0);
    return dur;
}

// hip/generateLaunchUpdateGrid.py:150-158 (the reference init values; the
// padding is zeroed too).
void launchInit(dim3 gridSize, dim3 blockSize, dim3 gridBlocks, double *d_array, unsigned int rightOffset)
{
    (void)blockSize;
    (void)gridBlocks;
    tcsg_desc d;
    tcsg_kernel_desc(kernel(), &d);
    tcsg_geometry g;
    if (!geometry_of(d, gridSize, rightOffset, g)) { report(TCSG_ERR_ARG, __LINE__); return; }
    report(tcsg_launch_init(kernel(), &g, TCSG_INIT_REFERENCE, 0, d_array, nullptr), __LINE__);
    report(hipDeviceSynchronize() == hipSuccess ? TCSG_OK : TCSG_ERR_HIP, __LINE__);
}

// hip/generateLaunchUpdateGrid.py:84-89
int maxPotentialBlockSize()
{
    int occ = 0;
    report(tcsg_max_block_size(kernel(), &occ, nullptr), __LINE__);
    return occ;
}

// hip/generateLaunchUpdateGrid.py:77-82
void copyMask(double *mask, unsigned int maskSize)
{
    report(tcsg_set_mask(kernel(), mask, maskSize), __LINE__);
}
