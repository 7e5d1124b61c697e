// single_host_de: the reference's program of the same name
// (cuda/main/single_host/differential_evolution.cpp) on the HIP backend.
#include "parser.hpp"
#include "runner.hpp"

int main(int argc, char *argv[]) {
    HipSingleHostRunner runner(DifferentialEvolutionParser::getInstance());
    return runner.main(argc, const_cast<const char **>(argv));
}
