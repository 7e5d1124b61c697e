// Command line of the three programs: the reference's Parser hierarchy
// (common/include/parser/*.hpp) with the same flags, built on a small option table
// instead of Boost.program_options (absent here).
//
// Kept from the reference (Parser.cpp:13-51 and the subclasses):
//   required --aph --apm --azimuth --input --output --traveltime --lower-bounds...
//   --upper-bounds...; optional --ain --bpctg --v0 --h0 --tau --lowest-midpoint
//   --highest-midpoint --kernel-path --thread-count (64) --verbose (0..4) --help;
//   linear search adds --granularity... (required), DE --generations and
//   --population-size (required), stretch-free --stack-datafiles... (required).
//   Only long options; a vector option takes every following token up to the next
//   "--" option (so "-0.001" is a value); "--name=value" is accepted; any parse error
//   becomes std::invalid_argument("Unable to parse input arguments: ...").
// Added for this backend (no reference counterpart): --variant exact|fast,
//   --storage f32|f16|bf16, --batch-size N (midpoints per launch, 0 = all),
//   --seed S (differential evolution; default: the clock, as the reference).
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "algorithm.hpp"
#include "logger.hpp"
#include "traveltime.hpp"

class OptionTable {
  public:
    enum Kind { FLOAT, UINT, STRING, FLOATS, INTS, STRINGS, FLAG };
    struct Option {
        Kind kind;
        bool required;
        std::string defaultValue;
        std::string description;
    };
    void add(const std::string &name, Kind kind, const std::string &description,
             bool required = false, const std::string &defaultValue = "");
    // name -> the raw tokens given for it
    std::map<std::string, std::vector<std::string>> parse(int argc, const char *argv[]) const;
    void print(std::ostream &out) const;
    const Option &at(const std::string &name) const { return options.at(name); }

  private:
    std::map<std::string, Option> options;
};

class Parser {
  public:
    Parser();
    virtual ~Parser() = default;

    void parseArguments(int argc, const char *argv[]);
    const std::string getFilename() const;
    unsigned int getGpuThreadCount() const;
    const std::string getInputFilePath() const;
    const std::string getOutputDirectory() const;
    LogLevel getLogVerbosity() const;
    virtual Traveltime *parseTraveltime() const;
    void printHelp() const;
    void readGather() const;
    static traveltime_t traveltimeFromString(const std::string &model);

    virtual ComputeAlgorithm *parseComputeAlgorithm(ComputeAlgorithmBuilder *builder,
                                                    std::shared_ptr<DeviceContext> deviceContext,
                                                    std::shared_ptr<Traveltime> traveltime) const = 0;
    virtual const std::string getParserType() const = 0;

  protected:
    bool has(const std::string &name) const { return argumentMap.count(name) > 0; }
    float getFloat(const std::string &name) const;
    unsigned int getUInt(const std::string &name) const;
    std::string getString(const std::string &name) const;
    std::vector<float> getFloats(const std::string &name) const;
    std::vector<int> getInts(const std::string &name) const;
    std::vector<std::string> getStrings(const std::string &name) const;

    OptionTable arguments;
    std::map<std::string, std::vector<std::string>> argumentMap;
};

class LinearSearchParser : public Parser {
  public:
    LinearSearchParser();
    static Parser *getInstance();
    ComputeAlgorithm *parseComputeAlgorithm(ComputeAlgorithmBuilder *builder,
                                            std::shared_ptr<DeviceContext> deviceContext,
                                            std::shared_ptr<Traveltime> traveltime) const override;
    const std::string getParserType() const override { return "greedy"; }
};

class DifferentialEvolutionParser : public Parser {
  public:
    DifferentialEvolutionParser();
    static Parser *getInstance();
    ComputeAlgorithm *parseComputeAlgorithm(ComputeAlgorithmBuilder *builder,
                                            std::shared_ptr<DeviceContext> deviceContext,
                                            std::shared_ptr<Traveltime> traveltime) const override;
    const std::string getParserType() const override { return "de"; }
};

class StretchFreeParser : public Parser {
  public:
    StretchFreeParser();
    static Parser *getInstance();
    Traveltime *parseTraveltime() const override;
    ComputeAlgorithm *parseComputeAlgorithm(ComputeAlgorithmBuilder *builder,
                                            std::shared_ptr<DeviceContext> deviceContext,
                                            std::shared_ptr<Traveltime> traveltime) const override;
    const std::string getParserType() const override { return "stretch_free"; }
};
