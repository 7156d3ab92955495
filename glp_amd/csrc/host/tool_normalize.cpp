// normalize: the command-line form of tools/normalize.cpp:350-365.
//   normalize graph.bin         -> ./sorted_graph.bin
#include <glp.h>
#include <glp_tools.h>

#include <cstdio>
#include <string>

int main(int argc, char** argv) {
    if (argc != 2) {
        std::printf("usage: ./normalize filename\n");
        return 0;
    }
    const std::string filename(argv[1]);
    // output name exactly as tools/normalize.cpp:327 derives it
    const std::string out = std::string("sorted_") +
                            filename.substr(filename.find_last_of("/") + 1,
                                            filename.find_last_of(".") - filename.find_last_of("/")) +
                            std::string("bin");
    if (glp_tool_normalize(filename.c_str(), out.c_str()) != GLP_OK) {
        std::fprintf(stderr, "normalize: %s\n", glp_last_error());
        return 1;
    }
    return 0;
}
