// bifhip.hpp -- reference-named C++ entry points over the C ABI.
//
// A caller written against the reference's Library/hostLibrary.cuh
// (KiShiVi/CUDABifurcation3.0) compiles unchanged against this header:
// same function names, argument order and types, `void` returns.  As in
// the reference, a device error prints "GPUassert: <msg> <file> <line>" and
// exits with the error code (Library/cudaMacros.cu:3-9), and a file that
// cannot be opened prints "\nOutput file open error\n" and exits with 1
// (Library/hostLibrary.cu:134-138).  Link with -lbifhip.
#pragma once

#include <cstdio>
#include <cstdlib>

#include "bifhip.h"

namespace bifhip_detail {
inline void check(int rc, const char* file, int line)
{
    if (rc == BIFHIP_OK) return;
    if (rc == BIFHIP_E_FILE) {
        printf("\nOutput file open error\n");
        exit(1);
    }
    fprintf(stderr, "GPUassert: %s %s %d\n", bifhip_last_error(), file, line);
    exit(rc > 0 ? rc : 1);
}
}  // namespace bifhip_detail

#define BIFHIP_CHECK(call) bifhip_detail::check((call), __FILE__, __LINE__)

// Library/hostLibrary.cuh:29-42
inline void bifurcation1D(const double tMax, const int nPts, const double h,
                          const int amountOfInitialConditions, const double* initialConditions,
                          const double* ranges, const int* indicesOfMutVars, const int writableVar,
                          const double maxValue, const double transientTime, const double* values,
                          const int amountOfValues, const int preScaller)
{
    BIFHIP_CHECK(bifhip_bifurcation1D(tMax, nPts, h, amountOfInitialConditions, initialConditions,
                                      ranges, indicesOfMutVars, writableVar, maxValue,
                                      transientTime, values, amountOfValues, preScaller));
}

// Library/hostLibrary.cuh:129-147
inline void bifurcation2DKDE(const double tMax, const int nPts, const double h,
                             const int amountOfInitialConditions, const double* initialConditions,
                             const double* ranges, const int* indicesOfMutVars,
                             const int writableVar, const double maxValue,
                             const int maxAmountOfPeaks, const double transientTime,
                             const double* values, const int amountOfValues,
                             const int preScaller, const int kdeSampling,
                             const double kdeSamplesInterval1, const double kdeSamplesInterval2,
                             const double kdeSamplesSmooth)
{
    BIFHIP_CHECK(bifhip_bifurcation2DKDE(tMax, nPts, h, amountOfInitialConditions,
                                         initialConditions, ranges, indicesOfMutVars, writableVar,
                                         maxValue, maxAmountOfPeaks, transientTime, values,
                                         amountOfValues, preScaller, kdeSampling,
                                         kdeSamplesInterval1, kdeSamplesInterval2,
                                         kdeSamplesSmooth));
}

// Library/hostLibrary.cuh:171-186
inline void bifurcation2DDBSCAN(const double tMax, const int nPts, const double h,
                                const int amountOfInitialConditions,
                                const double* initialConditions, const double* ranges,
                                const int* indicesOfMutVars, const int writableVar,
                                const double maxValue, const int maxAmountOfPeaks,
                                const double transientTime, const double* values,
                                const int amountOfValues, const int preScaller, const double eps)
{
    BIFHIP_CHECK(bifhip_bifurcation2DDBSCAN(tMax, nPts, h, amountOfInitialConditions,
                                            initialConditions, ranges, indicesOfMutVars,
                                            writableVar, maxValue, maxAmountOfPeaks,
                                            transientTime, values, amountOfValues, preScaller,
                                            eps));
}

// Library/hostLibrary.cuh:89-102
inline void bifurcation1DIC(const double tMax, const int nPts, const double h,
                            const int amountOfInitialConditions, const double* initialConditions,
                            const double* ranges, const int* indicesOfMutVars,
                            const int writableVar, const double maxValue,
                            const double transientTime, const double* values,
                            const int amountOfValues, const int preScaller)
{
    BIFHIP_CHECK(bifhip_bifurcation1DIC(tMax, nPts, h, amountOfInitialConditions,
                                        initialConditions, ranges, indicesOfMutVars, writableVar,
                                        maxValue, transientTime, values, amountOfValues,
                                        preScaller));
}

// Library/hostLibrary.cuh:254-272
inline void bifurcation2DKDEIC(const double tMax, const int nPts, const double h,
                               const int amountOfInitialConditions,
                               const double* initialConditions, const double* ranges,
                               const int* indicesOfMutVars, const int writableVar,
                               const double maxValue, const int maxAmountOfPeaks,
                               const double transientTime, const double* values,
                               const int amountOfValues, const int preScaller,
                               const int kdeSampling, const double kdeSamplesInterval1,
                               const double kdeSamplesInterval2, const double kdeSamplesSmooth)
{
    BIFHIP_CHECK(bifhip_bifurcation2DKDEIC(tMax, nPts, h, amountOfInitialConditions,
                                           initialConditions, ranges, indicesOfMutVars,
                                           writableVar, maxValue, maxAmountOfPeaks,
                                           transientTime, values, amountOfValues, preScaller,
                                           kdeSampling, kdeSamplesInterval1,
                                           kdeSamplesInterval2, kdeSamplesSmooth));
}

// Library/hostLibrary.cuh:296-311
inline void bifurcation2DDBSCANIC(const double tMax, const int nPts, const double h,
                                  const int amountOfInitialConditions,
                                  const double* initialConditions, const double* ranges,
                                  const int* indicesOfMutVars, const int writableVar,
                                  const double maxValue, const int maxAmountOfPeaks,
                                  const double transientTime, const double* values,
                                  const int amountOfValues, const int preScaller,
                                  const double eps)
{
    BIFHIP_CHECK(bifhip_bifurcation2DDBSCANIC(tMax, nPts, h, amountOfInitialConditions,
                                              initialConditions, ranges, indicesOfMutVars,
                                              writableVar, maxValue, maxAmountOfPeaks,
                                              transientTime, values, amountOfValues, preScaller,
                                              eps));
}

// Library/hostLibrary.cuh:59-67
inline void modelingOneSystemDenis(const double tMax, const double h, const double hSpecial,
                                   const int amountOfInitialConditions,
                                   const double* initialConditions, const int writableVar,
                                   const double* values, const int amountOfValues)
{
    BIFHIP_CHECK(bifhip_modelingOneSystemDenis(tMax, h, hSpecial, amountOfInitialConditions,
                                               initialConditions, writableVar, values,
                                               amountOfValues));
}

// Library/hostLibrary.cuh:211-227
inline void calculationOfPeriodicityByOstrovsky(
    const double tMax, const int nPts, const double h, const int amountOfInitialConditions,
    const double* initialConditions, const double* ranges, const int* indicesOfMutVars,
    const int writableVar, const double maxValue, const int maxAmountOfPeaks,
    const double transientTime, const double* values, const int amountOfValues,
    const int preScaller, const double eps, const double ostrovskyThreshold)
{
    BIFHIP_CHECK(bifhip_calculationOfPeriodicityByOstrovsky(
        tMax, nPts, h, amountOfInitialConditions, initialConditions, ranges, indicesOfMutVars,
        writableVar, maxValue, maxAmountOfPeaks, transientTime, values, amountOfValues,
        preScaller, eps, ostrovskyThreshold));
}

#define BIFHIP_LYAP_WRAPPER(NAME)                                                              \
    inline void NAME(const double tMax, const double NT, const int nPts, const double h,       \
                     const double eps, const double* initialConditions,                        \
                     const int amountOfInitialConditions, const double* ranges,                \
                     const int* indicesOfMutVars, const int writableVar, const double maxValue, \
                     const double transientTime, const double* values,                         \
                     const int amountOfValues)                                                 \
    {                                                                                          \
        BIFHIP_CHECK(bifhip_##NAME(tMax, NT, nPts, h, eps, initialConditions,                  \
                                   amountOfInitialConditions, ranges, indicesOfMutVars,        \
                                   writableVar, maxValue, transientTime, values,               \
                                   amountOfValues));                                           \
    }

BIFHIP_LYAP_WRAPPER(LLE1D)  // Library/hostLibrary.cuh:334-348
BIFHIP_LYAP_WRAPPER(LLE2D)  // Library/hostLibrary.cuh:408-422
BIFHIP_LYAP_WRAPPER(LS1D)   // Library/hostLibrary.cuh:482-496
BIFHIP_LYAP_WRAPPER(LS2D)   // Library/hostLibrary.cuh:520-534
BIFHIP_LYAP_WRAPPER(LLE1DIC)  // Library/hostLibrary.cuh:371-385
BIFHIP_LYAP_WRAPPER(LLE2DIC)  // Library/hostLibrary.cuh:445-459

#undef BIFHIP_LYAP_WRAPPER
