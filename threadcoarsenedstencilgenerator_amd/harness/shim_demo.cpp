// A harness written only against the reference's launchUpdateGrid.h (the way
// hip/benchmark.cpp uses it), linked with liblaunchUpdateGrid.so.
//   shim_demo X Y [Z]   with $TCSG_KERNEL naming the generated kernel
// Prints the occupancy-optimal block size and the per-step durations of a
// 0-second benchmark (a single step) on a (64,2,1) block.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "launchUpdateGrid.h"

int main(int argc, char **argv)
{
    if (argc < 7) {
        std::fprintf(stderr, "usage: shim_demo Box|Star Jacobi|Variable range dim alignment X Y [Z]\n");
        return 1;
    }
    const std::string nb = argv[1], coeff = argv[2], align = argv[5];
    const unsigned r = std::atoi(argv[3]), dim = std::atoi(argv[4]);
    dim3 grid(std::atoi(argv[6]), std::atoi(argv[7]), dim == 3 ? std::atoi(argv[8]) : 1);
    const unsigned n = nb == "Box" ? (dim == 2 ? (2 * r + 1) * (2 * r + 1) : (2 * r + 1) * (2 * r + 1) * (2 * r + 1))
                                   : 2 * dim * r + 1;
    std::vector<double> mask(n, 1.0 / n);
    copyMask(mask.data(), n);
    const int mb = maxPotentialBlockSize();
    const size_t N = (size_t)grid.x * grid.y * grid.z;
    double *d0 = nullptr, *d1 = nullptr;
    if (hipMalloc(&d0, N * 8) != hipSuccess || hipMalloc(&d1, N * 8) != hipSuccess) return 2;
    launchInit(grid, dim3(64, 1, 1), dim3(1, 1, 1), d0, 0);
    launchInit(grid, dim3(64, 1, 1), dim3(1, 1, 1), d1, 0);
    dim3 block(64, 2, 1), blocks((grid.x + 63) / 64, (grid.y + 1) / 2, grid.z);
    std::vector<double> dur;
    try {
        dur = benchmarkUpdateGrid(coeff, nb, r, dim, align, 0.0, grid, block, blocks, d0, d1, 0);
    } catch (hipError_t e) {
        std::printf("sanity check failed (%d)\n", (int)e);
        return 3;
    }
    std::printf("maxPotentialBlockSize %d steps %zu first %.3e\n", mb, dur.size(), dur.empty() ? 0.0 : dur[0]);
    (void)hipFree(d0);
    (void)hipFree(d1);
    return dur.size() == 1 ? 0 : 4;
}
