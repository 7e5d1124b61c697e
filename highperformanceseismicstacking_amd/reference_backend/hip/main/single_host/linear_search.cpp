#include "common/include/parser/LinearSearchParser.hpp"
#include "hip/include/execution/HipSingleHostRunner.hpp"

int main(int argc, char* argv[]) {
    HipSingleHostRunner runner(LinearSearchParser::getInstance());
    return runner.main(argc, const_cast<const char**>(argv));
}
