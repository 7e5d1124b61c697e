/*
 * CPU oracle, libstdc++ part -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The OCT grid-search trace selection draws 1024 (velocity, slope) pairs from a
 * default-constructed std::default_random_engine through
 * std::uniform_real_distribution<float>, all velocities first and then all slopes,
 * from one engine, converting velocities to c = 4/v^2
 * (cuda/src/semblance/algorithm/CudaLinearSearchAlgorithm.cu:268-292).  The values
 * are whatever libstdc++ produces, so this file uses <random> directly.
 */
#include "oracle.h"

#include <random>

extern "C" void orc_oct_selection_samples(float vmin, float vmax, float smin, float smax,
                                          int count, float *out)
{
    std::default_random_engine generator;
    const float lo[2] = {vmin, smin};
    const float hi[2] = {vmax, smax};
    for (int prm = 0; prm < 2; prm++) {
        std::uniform_real_distribution<float> dist(lo[prm], hi[prm]);
        for (int idx = 0; idx < count; idx++) {
            float r = dist(generator);
            if (prm == 0) {
                r = 4.0f / (r * r);
            }
            out[idx * 2 + prm] = r;
        }
    }
}
