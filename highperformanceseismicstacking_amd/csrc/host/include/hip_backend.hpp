// The HIP backend: subclasses of the common drivers that run the gfx950 kernels of
// libhss.so through the C ABI (include/hss.h).  They stand where the reference's Cuda*
// classes stand (cuda/include/semblance/algorithm/*.hpp, CudaComputeAlgorithmBuilder).
//
// Differences in how the work is laid out (results are the same):
//  * the whole line -- samples in the chosen storage, midpoints, half offsets -- is
//    uploaded once in setUp(); a midpoint's traces are a range (CMP, ZO-CRS) or an index
//    list (OCT) into it, instead of one host-to-device copy per selected trace
//    (ComputeAlgorithm.cpp:107-153);
//  * a grid search or a whole differential evolution is one launch per batch of
//    midpoints, the best-candidate pick fused in;
//  * computeSemblanceAndParametersForMidpoints() processes any number of midpoints per
//    launch; the per-midpoint interface stays available and gives the same values.
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "algorithm.hpp"
#include "hss.h"

// Settings of the HIP backend that the reference has no flag for (see parser.hpp).
struct BackendOptions {
    int variant = HSS_FAST;
    int storage = HSS_F32;
    unsigned int batchSize = 0;  // midpoints per launch in the runner; 0 = all of them
    bool seedGiven = false;
    uint64_t seed = 0;
    unsigned int polishSteps = 0;  // Gauss-Newton steps applied to every winner (hss_polish)
    static BackendOptions &get();
};

// The line resident on the device.
class HipLine {
  public:
    void upload(const Gather &gather, int storage, hipStream_t stream);
    const hss_traces &traces() const { return view; }
    const std::vector<float> &midpoints() const { return mid; }
    const std::vector<float> &halfoffsets() const { return hoff; }

  private:
    DeviceBuffer dSamples, dScale, dMid, dHoff, dHoff2;
    std::vector<float> mid, hoff;
    hss_traces view{};
};

// Which traces each midpoint of a batch uses, on the host and on the device.
class MidpointBatch {
  public:
    void clear();
    void addRange(float m0, int begin, int count);
    void addIndexList(float m0, const std::vector<int32_t> &traces);
    void upload(hipStream_t stream);
    hss_midpoints view(float h0, float apm) const;
    size_t size() const { return m0.size(); }
    int count(size_t g) const { return cnt[g]; }
    // Keeps the midpoints at the given positions (ascending), in that order; upload() again.
    void retain(const std::vector<size_t> &positions);

  private:
    std::vector<float> m0;
    std::vector<int32_t> beg, cnt, index, ord;
    bool indexed = false;
    DeviceBuffer dBeg, dCnt, dM0, dIndex, dOrder;
};

// Launch order of midpoints with the given work (selected-trace counts) that gives each
// of the chip's 8 XCDs the same share (hss_midpoints::order).
std::vector<int32_t> xcd_balanced_order(const std::vector<int32_t> &work);

// Replaces the search result in dRes ([n][(2+np)][ns] on the device) by its Gauss-Newton
// polish inside the traveltime's parameter bounds when --polish-steps is set.
void polish_results(const HipLine &line, const hss_midpoints &mids, const Traveltime &tt,
                    DeviceBuffer &dRes, hipStream_t stream);

// Trace selection on the sorted line: CMP m == m0 and ZO-CRS |m0 - m| <= apm are ranges
// (found by binary search with the same float comparisons as cmp/common.cu:13 and
// zocrs/common.cu:14); OCT runs hss_select_traces and turns the masks into index lists.
struct TraceRange {
    int begin, count;
};
TraceRange selectRange(const std::vector<float> &sortedMidpoints, traveltime_t model, float m0, float apm);

class HipLinearSearchAlgorithm : public LinearSearchAlgorithm {
  public:
    using LinearSearchAlgorithm::LinearSearchAlgorithm;
    void computeSemblanceAtGpuForMidpoint(float m0) override;
    void selectTracesToBeUsedForMidpoint(float m0) override;
    void initializeParameters() override;
    void computeSemblanceAndParametersForMidpoints(const std::vector<float> &m0s,
                                                   std::vector<MidpointOutput> &out) override;

  protected:
    void copyGatherDataToDevice() override;
    float getNotUsedCount() override { return lastNotUsed; }

  private:
    void selectBatch(const std::vector<float> &m0s);
    hipStream_t stream() const;
    HipLine line;
    MidpointBatch batch;
    DeviceBuffer dOctPairs, dMask, dNotUsed;
    float lastNotUsed = 0;
};

class HipDifferentialEvolutionAlgorithm : public DifferentialEvolutionAlgorithm {
  public:
    using DifferentialEvolutionAlgorithm::DifferentialEvolutionAlgorithm;
    // fused: every generation in one launch
    void computeSemblanceAndParametersForMidpoint(float m0) override;
    void computeSemblanceAtGpuForMidpoint(float m0) override;
    void selectTracesToBeUsedForMidpoint(float m0) override;
    void computeSemblanceAndParametersForMidpoints(const std::vector<float> &m0s,
                                                   std::vector<MidpointOutput> &out) override;
    // Operator steps of the reference's interface.  One generation launch does mutation,
    // crossover, evaluation and selection, so mutate/crossover/advance only sequence it:
    // startAllPopulations() rewinds to generation 0, each computeSemblanceAtGpuForMidpoint()
    // runs the next generation, selectBestIndividuals() reads the final pick.
    void setupRandomSeedArray() override;
    void startAllPopulations() override { nextGeneration = 0; }
    void mutateAllPopulations() override {}
    void crossoverPopulationIndividuals() override {}
    void advanceGeneration() override {}
    void selectBestIndividuals(std::vector<float> &resultArrays) override;
    uint64_t getSeed() const { return seed; }

  protected:
    void copyGatherDataToDevice() override;
    float getNotUsedCount() override { return lastNotUsed; }

  private:
    void selectBatch(const std::vector<float> &m0s, const std::vector<uint32_t> &streams);
    void runGenerations(const hss_midpoints &mids, uint32_t streamBase, int g0, int g1,
                        float *dResult);
    void bounds(float lo[3], float hi[3]) const;
    hipStream_t stream() const;
    uint32_t streamOf(float m0) const;
    HipLine line;
    MidpointBatch batch;
    DeviceBuffer dWorkspace, dMask, dNotUsed;
    std::map<float, uint32_t> cdpOrdinal;
    uint64_t seed = 0;
    int nextGeneration = 0;
    float lastNotUsed = 0;
};

class HipStretchFreeAlgorithm : public StretchFreeAlgorithm {
  public:
    using StretchFreeAlgorithm::StretchFreeAlgorithm;
    void computeSemblanceAtGpuForMidpoint(float m0) override;
    void selectTracesToBeUsedForMidpoint(float m0) override;
    void computeSemblanceAndParametersForMidpoints(const std::vector<float> &m0s,
                                                   std::vector<MidpointOutput> &out) override;

  protected:
    void copyGatherDataToDevice() override;
    float getNotUsedCount() override { return lastNotUsed; }

  private:
    void selectBatch(const std::vector<float> &m0s);
    void uploadParameters(const std::vector<float> &m0s);
    hipStream_t stream() const;
    HipLine line;
    MidpointBatch batch;
    DeviceBuffer dVparams, dMask, dNotUsed;
    float lastNotUsed = 0;
};

class HipComputeAlgorithmBuilder : public ComputeAlgorithmBuilder {
  public:
    static ComputeAlgorithmBuilder *getInstance();
    LinearSearchAlgorithm *buildLinearSearchAlgorithm(std::shared_ptr<Traveltime> traveltime,
                                                      std::shared_ptr<DeviceContext> context,
                                                      unsigned int threadCount,
                                                      const std::vector<int> &discretization) override;
    DifferentialEvolutionAlgorithm *buildDifferentialEvolutionAlgorithm(
        std::shared_ptr<Traveltime> traveltime, std::shared_ptr<DeviceContext> context,
        unsigned int threadCount, unsigned int generation, unsigned int individuals) override;
    StretchFreeAlgorithm *buildStretchFreeAlgorithm(std::shared_ptr<Traveltime> traveltime,
                                                    std::shared_ptr<DeviceContext> context,
                                                    unsigned int threadCount,
                                                    const std::vector<std::string> &files) override;
};
