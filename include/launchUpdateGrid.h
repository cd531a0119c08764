// launchUpdateGrid.h -- source-compatible replacement for the reference's
// hip/launchUpdateGrid.h:13-18 (same four C++ signatures), implemented over
// the tcsg C ABI by threadcoarsenedstencilgenerator_amd/harness/launchUpdateGrid.cpp
// (liblaunchUpdateGrid.so).  A harness written against the reference header
// links against it unchanged.  The generated kernel is the code object named
// by $TCSG_KERNEL (default ./kernels.co), loaded on first use.
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include <vector>

std::vector<double> benchmarkUpdateGrid(std::string coeffType, std::string stencilType, unsigned int stencilRange,
                                        unsigned int dimension, std::string alignment, double maxTime,
                                        dim3 gridSize, dim3 blockSize, dim3 gridBlocks, double *d_array0,
                                        double *d_array1, unsigned int rightOffset);
void launchInit(dim3 gridSize, dim3 blockSize, dim3 gridBlocks, double *d_array, unsigned int rightOffset);

// helper functions
int maxPotentialBlockSize();
void copyMask(double *mask, unsigned int maskSize);
