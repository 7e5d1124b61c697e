// single_host_stretch_free: the reference's program of the same name
// (cuda/main/single_host/stretch_free.cpp) on the HIP backend.
#include "parser.hpp"
#include "runner.hpp"

int main(int argc, char *argv[]) {
    HipSingleHostRunner runner(StretchFreeParser::getInstance());
    return runner.main(argc, const_cast<const char **>(argv));
}
