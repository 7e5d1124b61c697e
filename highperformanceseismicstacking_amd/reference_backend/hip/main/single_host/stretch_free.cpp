#include "common/include/parser/StretchFreeParser.hpp"
#include "hip/include/execution/HipSingleHostRunner.hpp"

int main(int argc, char* argv[]) {
    HipSingleHostRunner runner(StretchFreeParser::getInstance());
    return runner.main(argc, const_cast<const char**>(argv));
}
