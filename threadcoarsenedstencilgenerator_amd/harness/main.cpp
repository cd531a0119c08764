// `bench`: same positional argv as the reference's bench (hip/main.cpp:7-147).
//   sweep:   neighbourhood dimension range coeff orientation alignment CLsize maxTime xC yC [zC] X Y [Z]
//   single:  ... the same, then block sizes x y [z]
// The generated code object is taken from $TCSG_KERNEL (default ./kernels.co).
// Unlike the reference, single-block mode also reads the grid sizes.
#include <cstdlib>
#include <iostream>
#include <string>

#include "benchmark.h"

static int usage()
{
    std::cerr << "Unknown arguments, the relevant arguments have to match the generated kernel code.\n"
                 "Usage:\n\t\tstencil neighbourhood ('Box' or 'Star')\n\t\tdimension ('2D' or '3D')\n"
                 "\t\tstencil range > 0\n\t\tcoefficients type ('Jacobi' or 'Variable')\n"
                 "\t\torientation\t('Column' or 'Row')\n\t\talignment\t('None', 'Write' or 'Read')\n"
                 "\t\tsize of cache line in units of grid elements\n"
                 "\t\tmaximal runtime for single threadblock size in seconds\n"
                 "\t\tcoarsening factors (x y z[only for 3D])\n\n\t\tgrid size (x y z[only for 3D])\n\n"
                 "optional for run of single threadblock size:\n\t\tthread block size (x y z[only for 3D])\n";
    return EXIT_FAILURE;
}

int main(int argc, char **argv)
{
    if (argc < 13 || argc > 18) return usage();
    tcsg::BenchConfig c;
    c.neighbourhood = argv[1];
    const std::string dim = argv[2];
    if (dim != "2D" && dim != "3D") { std::cerr << "Unknown dimension: " << dim << std::endl; return usage(); }
    c.dimension = dim == "2D" ? 2 : 3;
    c.range = std::atoi(argv[3]);
    c.coeff = argv[4];
    c.orientation = argv[5];
    c.alignment = argv[6];
    c.cl_size = std::atoi(argv[7]);
    c.max_time = std::atof(argv[8]);
    c.xc = std::atoi(argv[9]);
    c.yc = std::atoi(argv[10]);
    int a = 11;
    if (c.dimension == 3) c.zc = std::atoi(argv[a++]);
    const int nsize = c.dimension;
    const int want_sweep = a + nsize, want_single = a + 2 * nsize;
    if (argc != want_sweep && argc != want_single) {
        std::cerr << "Dimension and amount of arguments mismatch!" << std::endl;
        return usage();
    }
    for (int i = 0; i < nsize; ++i) c.size[i] = std::atoi(argv[a++]);
    const char *kp = std::getenv("TCSG_KERNEL");
    c.kernel_path = kp ? kp : "kernels.co";
    c.reference_sweep = std::getenv("TCSG_REFERENCE_SWEEP") != nullptr;
    tcsg::StencilBenchmark b(c);
    b.init_gpu();
    b.init_benchmark();
    if (argc == want_sweep) {
        b.run();
    } else {
        unsigned t[3] = {1, 1, 1};
        for (int i = 0; i < nsize; ++i) t[i] = std::atoi(argv[a++]);
        b.single_run(t[0], t[1], t[2]);
    }
    b.print_results();
    return EXIT_SUCCESS;
}
