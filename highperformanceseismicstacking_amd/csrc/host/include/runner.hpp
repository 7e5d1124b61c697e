// The single-host driver of the three programs: the reference's SingleHostRunner
// (common/include/execution/SingleHostRunner.hpp, SingleHostRunner.cpp:43-196),
// re-implemented.
//
// One worker thread per device, each with its own DeviceContext and ComputeAlgorithm,
// pulls midpoints from a shared queue in ascending midpoint order and stores results
// under a mutex; results are written after all workers finish.  A worker takes up to
// --batch-size midpoints at a time (0 = the whole queue) and hands them to the
// backend's batched path.  Exit codes follow the reference: std::invalid_argument
// prints the message and the help and returns 0, std::runtime_error returns -ENODEV,
// any other exception -1.
#pragma once

#include <memory>
#include <mutex>
#include <queue>

#include "algorithm.hpp"
#include "output.hpp"
#include "parser.hpp"

class SingleHostRunner {
  public:
    SingleHostRunner(Parser *parser, ComputeAlgorithmBuilder *algorithmBuilder,
                     DeviceContextBuilder *deviceContextBuilder);
    virtual ~SingleHostRunner() = default;
    int main(int argc, const char *argv[]);
    ComputeAlgorithm *getComputeAlgorithm();
    ResultSet *getResultSet() { return resultSet.get(); }
    std::queue<float> &getMidpointQueue() { return midpointQueue; }
    std::mutex &getDeviceMutex() { return deviceMutex; }
    std::mutex &getQueueMutex() { return queueMutex; }
    std::mutex &getResultSetMutex() { return resultSetMutex; }
    static void workerThread(SingleHostRunner *ref);
    virtual unsigned int getNumOfDevices() const = 0;

  protected:
    // the steps of main(): read the line and queue its midpoints; run one worker per
    // device; write the result files; map the exception being handled to an exit status
    void loadLine();
    void runWorkers(unsigned int devices);
    void writeResults(Dumper &dumper);
    int exitStatusOfFailure() const;

    Parser *parser;
    std::mutex resultSetMutex, queueMutex, deviceMutex;
    std::queue<float> midpointQueue;
    ComputeAlgorithmBuilder *algorithmBuilder;
    DeviceContextBuilder *deviceContextBuilder;
    std::shared_ptr<Traveltime> traveltime;
    std::unique_ptr<ResultSet> resultSet;
    unsigned int deviceIndex = 0;
    std::exception_ptr workerError;
};

class HipSingleHostRunner : public SingleHostRunner {
  public:
    explicit HipSingleHostRunner(Parser *parser);
    // hipGetDeviceCount, throwing std::runtime_error when there is none
    // (CudaSingleHostRunner.cu:18-29)
    unsigned int getNumOfDevices() const override;
};
