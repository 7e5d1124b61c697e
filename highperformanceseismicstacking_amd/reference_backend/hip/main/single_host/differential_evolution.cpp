#include "common/include/parser/DifferentialEvolutionParser.hpp"
#include "hip/include/execution/HipSingleHostRunner.hpp"

int main(int argc, char* argv[]) {
    HipSingleHostRunner runner(DifferentialEvolutionParser::getInstance());
    return runner.main(argc, const_cast<const char**>(argv));
}
