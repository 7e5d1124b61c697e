// Per-midpoint results, statistics and the output files: the reference's ResultSet /
// MidpointResult / StatisticalMidpointResult (common/include/semblance/result/*.hpp) and
// Dumper (common/include/output/Dumper.hpp), re-implemented.
//
// Output layout kept from the reference (Dumper.cpp:13-150): a directory
// <output>/<input stem>_<method>_<traveltime word>_<YYYYMMDD_HHMM>/ holding one SU file per
// result (coherence.su, stack.su, velocity.su, ...; one trace per CDP in ascending
// midpoint order, the CDP header followed by ns float32), one CSV per statistic
// ("<m0>,<value>" lines in default stream formatting), gather.txt and traveltime.txt.
#pragma once

#include <map>
#include <string>
#include <unordered_map>
#include <vector>

#include "model.hpp"
#include "traveltime.hpp"

enum class StatisticResult {
    EFFICIENCY,
    INTR_PER_SEC,
    SELECTED_TRACES,
    TOTAL_SELECTION_KERNEL_EXECUTION_TIME,
    TOTAL_KERNEL_EXECUTION_TIME,
    CNT
};

// file names of the statistics (StatisticResult.cpp:5-11)
const std::string &statisticName(StatisticResult r);

class MidpointResult {
  public:
    const std::vector<float> &get(float m0) const;
    void save(float m0, std::vector<float>::const_iterator begin, std::vector<float>::const_iterator end);

  private:
    std::map<float, std::vector<float>> values;
};

class StatisticalMidpointResult {
  public:
    float getStatisticalResultForMidpoint(float m0) const;
    void setStatisticalResultForMidpoint(float m0, float value) { values[m0] = value; }
    // mean over all midpoints of the non-zero values (StatisticalMidpointResult.cpp:17-32)
    float getAverageOfAllMidpoints() const;

  private:
    std::map<float, float> values;
};

class ResultSet {
  public:
    ResultSet(unsigned int numberOfResults, unsigned int samples)
        : resultArray(numberOfResults), samplesPerResult(samples) {}
    const MidpointResult &getArrayForResult(unsigned int i) const { return resultArray.at(i); }
    const StatisticalMidpointResult &get(StatisticResult r) const;
    // array holds numberOfResults consecutive blocks of `samples` floats
    void setAllResultsForMidpoint(float m0, const std::vector<float> &array);
    void setStatisticalResultForMidpoint(float m0, StatisticResult r, float value) {
        statistics[static_cast<int>(r)].setStatisticalResultForMidpoint(m0, value);
    }

  private:
    std::vector<MidpointResult> resultArray;
    std::unordered_map<int, StatisticalMidpointResult> statistics;
    unsigned int samplesPerResult;
};

class Dumper {
  public:
    Dumper(const std::string &path, const std::string &dataFile, const std::string &computeMethod,
           const std::string &traveltime);
    void createDir() const;
    void dumpGatherParameters(const std::string &file) const;
    void dumpResult(const std::string &resultName, const MidpointResult &result) const;
    void dumpStatisticalResult(const std::string &name, const StatisticalMidpointResult &r) const;
    void dumpTraveltime(const Traveltime *model) const;
    void dumpAlgorithm(const std::string &summary) const;
    const std::string &getOutputDirectoryPath() const { return outputDirectoryPath; }

  private:
    std::string outputDirectoryPath;
};

// The files a finished run leaves (SingleHostRunner.cpp:150-170, SpitzCommitter.cpp:
// 65-80): gather.txt with the input path, traveltime.txt, one SU file per result and one
// CSV per statistic, in that order, into the dumper's (already created) directory.
void dumpResultSet(const Dumper &dumper, const std::string &inputPath, const Traveltime &model,
                   const ResultSet &results);
