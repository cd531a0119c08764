// A harness written only against launchUpdateGrid.h, linked with the launch unit that
// generateLaunchUpdateGrid.py prints (the generated kernels.cu compiled in, the reference's build
// layout).  It does what hip/benchmark.cpp does before its sweep -- copyMask with the weights 1/n,
// launchInit on both buffers -- then `steps` calls of benchmarkUpdateGrid with maxTime 0 (one
// launch each, buffers alternating) and writes the final buffer to a file.
//   compiled_in_check nb coeff range dim alignment xdim ydim zdim rightOffset
//                     bx by bz tx ty tz steps out.bin
// (xdim.. are the padded allocation extents, tx.. the threads per axis.)  Exit status 3 when
// benchmarkUpdateGrid throws (e.g. a block shape the kernel cannot run).
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "launchUpdateGrid.h"

int main(int argc, char **argv)
{
    if (argc != 18) {
        std::fprintf(stderr, "usage: compiled_in_check nb coeff range dim alignment xdim ydim zdim rightOffset "
                             "bx by bz tx ty tz steps out.bin\n");
        return 1;
    }
    const std::string nb = argv[1], coeff = argv[2], align = argv[5];
    const unsigned r = std::atoi(argv[3]), dim = std::atoi(argv[4]);
    const dim3 grid(std::atoi(argv[6]), std::atoi(argv[7]), std::atoi(argv[8]));
    const unsigned ro = std::atoi(argv[9]);
    const dim3 block(std::atoi(argv[10]), std::atoi(argv[11]), std::atoi(argv[12]));
    const dim3 threads(std::atoi(argv[13]), std::atoi(argv[14]), std::atoi(argv[15]));
    const int steps = std::atoi(argv[16]);
    const unsigned w = 2 * r + 1;
    const unsigned n = nb == "Box" ? (dim == 2 ? w * w : w * w * w) : 2 * dim * r + 1;
    std::vector<double> mask(n, 1.0 / n);
    copyMask(mask.data(), n);
    std::printf("maxPotentialBlockSize %d\n", maxPotentialBlockSize());
    const size_t N = (size_t)grid.x * grid.y * grid.z;
    double *d[2] = {nullptr, nullptr};
    if (hipMalloc(&d[0], N * 8) != hipSuccess || hipMalloc(&d[1], N * 8) != hipSuccess) return 2;
    launchInit(grid, dim3(16, 16, 4), dim3(1, 1, 1), d[0], ro);
    launchInit(grid, dim3(16, 16, 4), dim3(1, 1, 1), d[1], ro);
    const dim3 blocks((threads.x + block.x - 1) / block.x, (threads.y + block.y - 1) / block.y,
                      (threads.z + block.z - 1) / block.z);
    for (int s = 0; s < steps; ++s) {
        try {
            const std::vector<double> dur =
                benchmarkUpdateGrid(coeff, nb, r, dim, align, 0.0, grid, block, blocks, d[s & 1], d[(s + 1) & 1], ro);
            if (dur.size() != 1) return 4;
        } catch (hipError_t e) {
            std::printf("benchmarkUpdateGrid threw %d\n", (int)e);
            return 3;
        }
    }
    std::vector<double> host(N);
    if (hipMemcpy(host.data(), d[steps & 1], N * 8, hipMemcpyDeviceToHost) != hipSuccess) return 2;
    FILE *f = std::fopen(argv[17], "wb");
    if (!f || std::fwrite(host.data(), 8, N, f) != N) return 2;
    std::fclose(f);
    (void)hipFree(d[0]);
    (void)hipFree(d[1]);
    std::printf("ok %d steps\n", steps);
    return 0;
}
