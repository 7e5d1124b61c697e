// single_host_linear_search: the reference's program of the same name
// (cuda/main/single_host/linear_search.cpp) on the HIP backend.
#include "parser.hpp"
#include "runner.hpp"

int main(int argc, char *argv[]) {
    HipSingleHostRunner runner(LinearSearchParser::getInstance());
    return runner.main(argc, const_cast<const char **>(argv));
}
