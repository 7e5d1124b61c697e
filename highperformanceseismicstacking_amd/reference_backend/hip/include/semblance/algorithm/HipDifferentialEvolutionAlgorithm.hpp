// Differential evolution on the HIP backend (replaces CudaDifferentialEvolutionAlgorithm,
// cuda/src/semblance/algorithm/CudaDifferentialEvolutionAlgorithm.cu).  The reference's
// driver (DifferentialEvolutionAlgorithm::computeSemblanceAndParametersForMidpoint) calls
// startAllPopulations, then computeSemblanceAtGpuForMidpoint once for the start
// population and once per later generation, with mutate / crossover / advanceGeneration
// around it.  hss_de_semblance runs a range of generations with the operators fused, so
// here each computeSemblanceAtGpuForMidpoint runs the next generation (the first: the
// start population and its fitness), the operator calls only keep the count, and
// selectBestIndividuals reads the result the last generation wrote.  The populations live
// in the library's workspace, not in the reference's x / u / v / fx / fu containers.
#pragma once

#include "common/include/semblance/algorithm/DifferentialEvolutionAlgorithm.hpp"
#include "hip/include/semblance/algorithm/HipMidpointBatch.hpp"

#include <cstdint>

class HipDifferentialEvolutionAlgorithm : public DifferentialEvolutionAlgorithm {
    public:
        HipDifferentialEvolutionAlgorithm(
            shared_ptr<Traveltime> traveltime,
            shared_ptr<DeviceContext> context,
            DataContainerBuilder* dataBuilder,
            unsigned int threadCount,
            unsigned int generation,
            unsigned int individualsPerPopulation
        );
        ~HipDifferentialEvolutionAlgorithm() override;

        void computeSemblanceAtGpuForMidpoint(float m0) override;
        void selectTracesToBeUsedForMidpoint(float m0) override;

        void setupRandomSeedArray() override;
        void startAllPopulations() override;
        void mutateAllPopulations() override;
        void crossoverPopulationIndividuals() override;
        void advanceGeneration() override;
        void selectBestIndividuals(vector<float>& resultArrays) override;

    private:
        HipMidpointBatch batch;
        uint64_t seed = 0;
        uint32_t midpointStream = 0;   // Philox stream of the current midpoint
        unsigned int nextGeneration = 0;
        void* workspace = nullptr;
        uint64_t workspaceBytes = 0;
        float* result = nullptr;       // [(2 + np)][ns] written by the last generation
};
