// Drives a generated kernels.cu through both of the reference's launch paths, in one
// translation unit with the launchUpdateGrid.cu that generateLaunchUpdateGrid.py prints (which
// compiles kernels.cu in; the build puts both on the include path):
//   unit    the four launchUpdateGrid.h entry points of that unit, set up as hip/benchmark.cpp
//           sets up its grids: copyMask with the weights 1/n, launchInit on both buffers, then
//           `steps` calls of benchmarkUpdateGrid with maxTime 0 (one launch each);
//   plugin  the kernel file's own symbols, used the way the reference's generated unit uses them
//           (hip/generateLaunchUpdateGrid.py:77-82,92-100,150-158): the weights 1/n copied to
//           d_mask, init launched over the grid with the reference's init blocks, then `steps`
//           launches of updateGrid.
// Buffers alternate between steps; the final one is written to out.bin.
//   compiled_in_check unit|plugin nb coeff range dim alignment xdim ydim zdim rightOffset
//                     bx by bz tx ty tz steps out.bin
// (xdim.. are the padded allocation extents, tx.. the threads per axis.)  Exit status 3 when
// benchmarkUpdateGrid throws (e.g. a block shape the kernel cannot run), 5 when a plugin launch
// fails.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "launchUpdateGrid.cu"

namespace {

void plugin_init(dim3 grid, unsigned ro, double *d)
{
#if TCSG_PLUGIN_DIM == 3
    const dim3 b(16, 16, 4);
    const dim3 n((grid.x + b.x - 1) / b.x, (grid.y + b.y - 1) / b.y, (grid.z + b.z - 1) / b.z);
#if TCSG_PLUGIN_RIGHT_OFFSET
    hipLaunchKernelGGL(init, n, b, 0, 0, d, grid.x, grid.y, grid.z, ro);
#else
    (void)ro;
    hipLaunchKernelGGL(init, n, b, 0, 0, d, grid.x, grid.y, grid.z);
#endif
#else
    const dim3 b(32, 16);
    const dim3 n((grid.x + b.x - 1) / b.x, (grid.y + b.y - 1) / b.y);
#if TCSG_PLUGIN_RIGHT_OFFSET
    hipLaunchKernelGGL(init, n, b, 0, 0, d, grid.x, grid.y, ro);
#else
    (void)ro;
    hipLaunchKernelGGL(init, n, b, 0, 0, d, grid.x, grid.y);
#endif
#endif
}

void plugin_step(dim3 grid, unsigned ro, dim3 blocks, dim3 block, double *in, double *out)
{
#if TCSG_PLUGIN_DIM == 3
#if TCSG_PLUGIN_RIGHT_OFFSET
    hipLaunchKernelGGL(updateGrid, blocks, block, 0, 0, grid.x, grid.y, grid.z, in, out, ro);
#else
    (void)ro;
    hipLaunchKernelGGL(updateGrid, blocks, block, 0, 0, grid.x, grid.y, grid.z, in, out);
#endif
#else
#if TCSG_PLUGIN_RIGHT_OFFSET
    hipLaunchKernelGGL(updateGrid, blocks, block, 0, 0, grid.x, grid.y, in, out, ro);
#else
    (void)ro;
    hipLaunchKernelGGL(updateGrid, blocks, block, 0, 0, grid.x, grid.y, in, out);
#endif
#endif
}

} // namespace

int main(int argc, char **argv)
{
    if (argc != 19) {
        std::fprintf(stderr, "usage: compiled_in_check unit|plugin nb coeff range dim alignment xdim ydim zdim "
                             "rightOffset bx by bz tx ty tz steps out.bin\n");
        return 1;
    }
    const std::string mode = argv[1], nb = argv[2], coeff = argv[3], align = argv[6];
    const unsigned r = std::atoi(argv[4]), dim = std::atoi(argv[5]);
    const dim3 grid(std::atoi(argv[7]), std::atoi(argv[8]), std::atoi(argv[9]));
    const unsigned ro = std::atoi(argv[10]);
    const dim3 block(std::atoi(argv[11]), std::atoi(argv[12]), std::atoi(argv[13]));
    const dim3 threads(std::atoi(argv[14]), std::atoi(argv[15]), std::atoi(argv[16]));
    const int steps = std::atoi(argv[17]);
    if (mode != "unit" && mode != "plugin") return 1;
    const unsigned w = 2 * r + 1;
    const unsigned n = nb == "Box" ? (dim == 2 ? w * w : w * w * w) : 2 * dim * r + 1;
    std::vector<double> mask(n, 1.0 / n);
    const size_t N = (size_t)grid.x * grid.y * grid.z;
    double *d[2] = {nullptr, nullptr};
    if (hipMalloc(&d[0], N * 8) != hipSuccess || hipMalloc(&d[1], N * 8) != hipSuccess) return 2;
    const dim3 blocks((threads.x + block.x - 1) / block.x, (threads.y + block.y - 1) / block.y,
                      (threads.z + block.z - 1) / block.z);
    if (mode == "unit") {
        copyMask(mask.data(), n);
        std::printf("maxPotentialBlockSize %d\n", maxPotentialBlockSize());
        launchInit(grid, dim3(16, 16, 4), dim3(1, 1, 1), d[0], ro);
        launchInit(grid, dim3(16, 16, 4), dim3(1, 1, 1), d[1], ro);
        for (int s = 0; s < steps; ++s) {
            try {
                const std::vector<double> dur = benchmarkUpdateGrid(coeff, nb, r, dim, align, 0.0, grid, block, blocks,
                                                                    d[s & 1], d[(s + 1) & 1], ro);
                if (dur.size() != 1) return 4;
            } catch (hipError_t e) {
                std::printf("benchmarkUpdateGrid threw %d\n", (int)e);
                return 3;
            }
        }
    } else {
        if (n != TCSG_PLUGIN_STENCIL_SIZE) return 1;
#if TCSG_PLUGIN_VARIABLE
        if (hipMemcpyToSymbol(HIP_SYMBOL(d_mask), mask.data(), n * sizeof(double)) != hipSuccess) return 5;
#endif
        // the reference's init leaves row padding as allocated; clear it so whole grids compare
        if (hipMemset(d[0], 0, N * 8) != hipSuccess || hipMemset(d[1], 0, N * 8) != hipSuccess) return 2;
        plugin_init(grid, ro, d[0]);
        plugin_init(grid, ro, d[1]);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) return 5;
        for (int s = 0; s < steps; ++s) {
            plugin_step(grid, ro, blocks, block, d[s & 1], d[(s + 1) & 1]);
            if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) return 5;
        }
        std::printf("plugin launches done\n");
    }
    std::vector<double> host(N);
    if (hipMemcpy(host.data(), d[steps & 1], N * 8, hipMemcpyDeviceToHost) != hipSuccess) return 2;
    FILE *f = std::fopen(argv[18], "wb");
    if (!f || std::fwrite(host.data(), 8, N, f) != N) return 2;
    std::fclose(f);
    (void)hipFree(d[0]);
    (void)hipFree(d[1]);
    std::printf("ok %d steps\n", steps);
    return 0;
}
