/*
 * Reference-built checker -- TEST INFRASTRUCTURE ONLY.
 *
 * Compiles the reference's own arithmetic from /root/reference, unmodified:
 *   common/include/gpu/interface.h            RESET/COMPUTE_SEMBLANCE, REDUCE_SEMBLANCE_STACK
 *   opencl/src/semblance/kernel/oct/common.cl computeDisplacedMidpoint, computeTime
 * (the OpenCL file is plain C; <math.h> supplies the float overloads of sqrt/fabs that
 * OpenCL C has built in).  The per-thread loop around them is written here because
 * the kernel files need a CUDA or OpenCL compiler; it follows
 * cmp/linear_search.cu:55-76, zocrs/linear_search.cu:74-100 and
 * oct/linear_search.cu:75-106.  Built with -ffp-contract=off, like the oracle.
 */
#include <math.h>

#include "common/include/gpu/interface.h"
#include "opencl/src/semblance/kernel/oct/common.cl"

extern "C" void ref_eval(int model, const float *samples, const float *midpoint,
                         const float *halfoffset, const float *halfoffsetSquared,
                         unsigned int traceCount, unsigned int samplesPerTrace,
                         float dtInSeconds, int tauIndexDisplacement, int windowSize,
                         float m0, float h0, float apm, unsigned int sampleIndex,
                         const float *prm, float *semblanceOut, float *stackOut,
                         unsigned int *usedOut, unsigned int *notUsedOut)
{
    float semblance = 0;
    float stack = 0;
    float t0 = static_cast<float>(sampleIndex) * dtInSeconds;
    float numeratorComponents[MAX_WINDOW_SIZE];
    float denominatorSum = 0;
    float linearSum = 0;
    RESET_SEMBLANCE_NUM_COMP(numeratorComponents, MAX_WINDOW_SIZE);
    unsigned int usedCount = 0, notUsedCount = 0;

    for (unsigned int traceIndex = 0; traceIndex < traceCount; traceIndex++) {
        const float *traceSamples = samples + traceIndex * samplesPerTrace;
        if (model == CMP) {
            float c = prm[0];
            float h_sq = halfoffsetSquared[traceIndex];
            float t = sqrt(t0 * t0 + c * h_sq);
            COMPUTE_SEMBLANCE(t, dtInSeconds, samplesPerTrace, tauIndexDisplacement, windowSize,
                              numeratorComponents, linearSum, denominatorSum, usedCount);
        } else if (model == ZOCRS) {
            float c = prm[0], a = prm[1], b = prm[2];
            float m = midpoint[traceIndex];
            float h_sq = halfoffsetSquared[traceIndex];
            float dm = m - m0;
            float tmp = t0 + a * dm;
            tmp = tmp * tmp + b * dm * dm + c * h_sq;
            if (tmp >= 0) {
                float t = sqrt(tmp);
                COMPUTE_SEMBLANCE(t, dtInSeconds, samplesPerTrace, tauIndexDisplacement, windowSize,
                                  numeratorComponents, linearSum, denominatorSum, usedCount);
            }
        } else {
            float c = prm[0], slope = prm[1];
            float m = midpoint[traceIndex];
            float h = halfoffset[traceIndex];
            float mh, t;
            enum gpu_error_code errorCode = computeDisplacedMidpoint(h, h0, t0, m0, c, slope, &mh);
            if (errorCode != NO_ERROR || fabs(m - mh) > apm) {
                notUsedCount++;
                continue;
            }
            errorCode = computeTime(h, h0, t0, m0, m, mh, c, slope, &t);
            if (errorCode == NO_ERROR) {
                COMPUTE_SEMBLANCE(t, dtInSeconds, samplesPerTrace, tauIndexDisplacement, windowSize,
                                  numeratorComponents, linearSum, denominatorSum, usedCount);
            }
        }
    }
    REDUCE_SEMBLANCE_STACK(numeratorComponents, linearSum, denominatorSum, windowSize, usedCount,
                           semblance, stack);
    *semblanceOut = semblance;
    *stackOut = stack;
    *usedOut = usedCount;
    *notUsedOut = notUsedCount;
}

extern "C" int ref_oct_displaced_midpoint(float h, float h0, float t0, float m0, float c,
                                          float slope, float *mh)
{
    return (int)computeDisplacedMidpoint(h, h0, t0, m0, c, slope, mh);
}

extern "C" int ref_oct_time(float h, float h0, float t0, float m0, float m, float mh, float c,
                            float slope, float *out)
{
    return (int)computeTime(h, h0, t0, m0, m, mh, c, slope, out);
}
