// The per-midpoint search drivers: the reference's ComputeAlgorithm hierarchy and
// ComputeAlgorithmBuilder (common/include/semblance/algorithm/*.hpp), re-implemented.
//
// A backend subclasses LinearSearchAlgorithm / DifferentialEvolutionAlgorithm /
// StretchFreeAlgorithm and provides the device steps (trace selection, semblance
// evaluation, DE operators); the common classes own the per-midpoint sequence, the
// grid discretisation and the statistics.  Beyond the reference, a backend may override
// computeSemblanceAndParametersForMidpoints() to process many midpoints per launch; the
// default runs the per-midpoint path for each.
#pragma once

#include <array>
#include <chrono>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "data.hpp"
#include "model.hpp"
#include "output.hpp"
#include "traveltime.hpp"

#define DEFAULT_THREAD_COUNT 64

using Seconds = std::chrono::duration<double>;

// Adds the wall time of `expr` to `acc` (execution/Utils.hpp:5-9).
#define MEASURE_EXEC_TIME(acc, expr)                                           \
    do {                                                                       \
        auto hss_t0_ = std::chrono::steady_clock::now();                       \
        (expr);                                                                \
        (acc) += std::chrono::steady_clock::now() - hss_t0_;                   \
    } while (0)

constexpr unsigned kStatisticCount = static_cast<unsigned>(StatisticResult::CNT);

// One midpoint's outcome of a batched computation.
struct MidpointOutput {
    float m0 = 0;
    std::vector<float> results;                        // numberOfResults x ns
    std::array<float, kStatisticCount> statistics{};   // indexed by StatisticResult
};

class ComputeAlgorithm {
  public:
    ComputeAlgorithm(const std::string &name, std::shared_ptr<Traveltime> traveltime,
                     std::shared_ptr<DeviceContext> context, DataContainerBuilder *dataBuilder,
                     unsigned int threadCount = DEFAULT_THREAD_COUNT);
    virtual ~ComputeAlgorithm() = default;

    bool isSetUp() const { return isSet; }
    const std::string &getAlgorithmName() const { return algorithmName; }
    const std::vector<float> &getComputedResults() const { return computedResults; }
    float getStatisticalResult(StatisticResult r) const;
    // Fills the statistics of the last midpoint (ComputeAlgorithm.cpp:71-105):
    // interpolations/s = w * (used - notUsed) / kernel time, efficiency = 1 - notUsed/used.
    void saveStatisticalResults(float totalUsedTracesCount, Seconds totalExecutionTime,
                                Seconds selectionExecutionTime);

    // The reference compiles OpenCL sources given by --kernel-path here; HIP kernels are
    // built ahead of time for gfx950, so the path is accepted and ignored.
    virtual void compileKernels(const std::string &deviceKernelSourcePath);

    virtual void computeSemblanceAndParametersForMidpoint(float m0) = 0;
    virtual void computeSemblanceAtGpuForMidpoint(float m0) = 0;
    virtual void selectTracesToBeUsedForMidpoint(float m0) = 0;
    virtual void setUp() = 0;
    virtual const std::string toString() const = 0;

    virtual void computeSemblanceAndParametersForMidpoints(const std::vector<float> &m0s,
                                                           std::vector<MidpointOutput> &out);

  protected:
    // Uploads what the backend keeps resident for the whole line.
    virtual void copyGatherDataToDevice() = 0;
    // OCT trace rejections counted by the last midpoint's kernels.
    virtual float getNotUsedCount() { return 0.0f; }

    bool isSet = false;
    DataContainerBuilder *dataFactory;
    std::string algorithmName;
    std::shared_ptr<DeviceContext> deviceContext;
    std::shared_ptr<Traveltime> traveltime;
    unsigned int threadCount;
    std::unique_ptr<DataContainer> deviceParameterArray, deviceResultArray;
    std::unordered_map<int, float> computedStatisticalResults;
    unsigned int filteredTracesCount = 0;
    std::vector<float> computedResults;
};

class LinearSearchAlgorithm : public ComputeAlgorithm {
  public:
    LinearSearchAlgorithm(std::shared_ptr<Traveltime> traveltime, std::shared_ptr<DeviceContext> context,
                          DataContainerBuilder *dataBuilder, unsigned int threadCount);
    void computeSemblanceAndParametersForMidpoint(float m0) override;
    void setUp() override;
    const std::string toString() const override;

    // value of parameter p at flat grid index `iteration` (LinearSearchAlgorithm.cpp:59-62)
    float getParameterValueAt(unsigned int iteration, unsigned int p) const;
    unsigned int getTotalNumberOfParameters() const;
    void setDiscretizationDivisorForParameter(unsigned int p, unsigned int d);
    void setDiscretizationGranularityForParameter(unsigned int p, unsigned int g);
    float getDiscretizationStep(unsigned int p) const { return discretizationStep.at(p); }
    unsigned int getDiscretizationGranularity(unsigned int p) const { return discretizationGranularity.at(p); }

    virtual void initializeParameters() = 0;

  protected:
    void setupArrays();
    // step = (upper - lower) / granularity in float: the upper bound is excluded
    // (LinearSearchAlgorithm.cpp:129-141)
    void setupDiscretizationSteps();

    std::vector<unsigned int> discretizationGranularity, discretizationDivisor;
    std::vector<float> discretizationStep;
};

class DifferentialEvolutionAlgorithm : public ComputeAlgorithm {
  public:
    DifferentialEvolutionAlgorithm(std::shared_ptr<Traveltime> traveltime,
                                   std::shared_ptr<DeviceContext> context,
                                   DataContainerBuilder *dataBuilder, unsigned int threadCount,
                                   unsigned int generations, unsigned int individualsPerPopulation);
    // The reference's generation loop (DifferentialEvolutionAlgorithm.cpp:21-81) over the
    // operator steps below.
    void computeSemblanceAndParametersForMidpoint(float m0) override;
    void setUp() override;
    const std::string toString() const override;

    unsigned int getGenerations() const { return generations; }
    unsigned int getIndividualsPerPopulation() const { return individualsPerPopulation; }

    virtual void setupRandomSeedArray() = 0;
    virtual void startAllPopulations() = 0;
    virtual void mutateAllPopulations() = 0;
    virtual void crossoverPopulationIndividuals() = 0;
    virtual void advanceGeneration() = 0;
    virtual void selectBestIndividuals(std::vector<float> &resultArrays) = 0;

  protected:
    unsigned int generations, individualsPerPopulation;
};

class StretchFreeAlgorithm : public ComputeAlgorithm {
  public:
    StretchFreeAlgorithm(std::shared_ptr<Traveltime> traveltime, std::shared_ptr<DeviceContext> context,
                         DataContainerBuilder *dataBuilder, unsigned int threadCount,
                         const std::vector<std::string> &parameterFiles);
    void computeSemblanceAndParametersForMidpoint(float m0) override;
    void setUp() override;
    const std::string toString() const override;
    // One SU trace per CDP per file, in ascending midpoint order
    // (StretchFreeAlgorithm.cpp:101-125); kept on the host as [np][ns] per midpoint.
    void readNonStretchedFreeParameterFromFile();

    static constexpr int NMAX = 10;  // shifts n in [-NMAX, NMAX] (StretchFreeAlgorithm.cpp:9)

  protected:
    unsigned int getTotalNumberOfParameters() const { return 2 * NMAX + 1; }
    std::vector<std::string> parameterFileArray;
    std::map<float, std::vector<float>> nonStretchFreeParameters;
    std::vector<float> nArray;
};

class ComputeAlgorithmBuilder {
  public:
    virtual ~ComputeAlgorithmBuilder() = default;
    virtual LinearSearchAlgorithm *buildLinearSearchAlgorithm(
        std::shared_ptr<Traveltime> traveltime, std::shared_ptr<DeviceContext> context,
        unsigned int threadCount, const std::vector<int> &discretizationArray) = 0;
    virtual DifferentialEvolutionAlgorithm *buildDifferentialEvolutionAlgorithm(
        std::shared_ptr<Traveltime> traveltime, std::shared_ptr<DeviceContext> context,
        unsigned int threadCount, unsigned int generation, unsigned int individualsPerPopulation) = 0;
    virtual StretchFreeAlgorithm *buildStretchFreeAlgorithm(
        std::shared_ptr<Traveltime> traveltime, std::shared_ptr<DeviceContext> context,
        unsigned int threadCount, const std::vector<std::string> &parameterFileArray) = 0;
};
