// Per-midpoint tasks: the reference's SPITS module (SpitzJobManager, SpitzWorker,
// SpitzCommitter; common/src/execution/Spitz*.cpp) as a plain C++ interface, without the
// SPITS runtime (libspits is not part of the reference tree and is not needed here).
//
// A job is one command line of one of the three programs and the line it reads.  Its
// job manager hands out one task per CDP, in ascending midpoint order, each a byte record
// holding the midpoint (SpitzJobManager.cpp:17-42).  A worker owns one device and one
// compute algorithm; it turns a task into a result record -- the midpoint, every result
// array of that midpoint (numberOfResults x ns floats, ComputeAlgorithm::
// getComputedResults order) and the statistics (StatisticResult order) -- as
// SpitzWorker::run does (SpitzWorker.cpp:10-40).  The committer stores result records in
// any order and finally writes the same files as the single-host programs
// (SpitzCommitter.cpp:40-99).  Records are native-endian float32 arrays.
//
// Workers may run on other threads or other processes (a record is self-contained); the
// job's next-task and commit calls are serialised by the job.  The line is held by the
// process-wide Gather, as in the reference, so a process opens one job at a time.
#pragma once

#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "algorithm.hpp"
#include "output.hpp"
#include "parser.hpp"

class TaskJob {
  public:
    // program 0: single_host_linear_search, 1: single_host_de, 2: single_host_stretch_free.
    // Parses the command line (std::invalid_argument on a usage error), builds the
    // traveltime, reads the line and creates the output directory.
    TaskJob(int program, int argc, const char **argv);

    size_t taskBytes() const { return sizeof(float); }
    size_t resultBytes() const;

    // The next task record (false when every CDP has been handed out).
    bool nextTask(std::vector<uint8_t> &task);
    // Stores one result record (std::invalid_argument when its length or midpoint is not
    // one of this job's).
    void commitTask(const uint8_t *record, size_t length);
    // Writes the result files; returns the output directory.
    std::string commitJob();

    Parser *getParser() const { return parser; }
    std::shared_ptr<Traveltime> getTraveltime() const { return traveltime; }
    size_t committedCount() const { return committed; }
    size_t taskCount() const { return midpoints.size(); }

  private:
    Parser *parser;
    std::shared_ptr<Traveltime> traveltime;
    std::unique_ptr<Dumper> dumper;
    std::unique_ptr<ResultSet> resultSet;
    std::vector<float> midpoints;  // ascending
    size_t next = 0, committed = 0;
    unsigned int samples = 0;
    std::mutex mu;
};

class TaskWorker {
  public:
    // A compute algorithm of the job's program on device `device` (set up on the first
    // task, as SpitzWorker::run does).
    TaskWorker(TaskJob &job, unsigned int device);
    // Computes the task's midpoint; writes its result record.
    void run(const uint8_t *task, size_t length, std::vector<uint8_t> &result);

  private:
    TaskJob &job;
    std::unique_ptr<ComputeAlgorithm> algorithm;
};
