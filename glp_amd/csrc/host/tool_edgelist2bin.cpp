// edgelist2bin: the command-line form of tools/edgelist2bin.cpp:115-130.
//   edgelist2bin graph.txt      -> ./graph.bin
#include <glp.h>
#include <glp_tools.h>

#include <cstdio>
#include <string>

int main(int argc, char** argv) {
    if (argc != 2) {
        std::printf("usage: ./edgelist2bin filename\n");
        return 0;
    }
    const std::string filename(argv[1]);
    // output name exactly as tools/edgelist2bin.cpp:54 derives it
    const std::string out = filename.substr(filename.find_last_of("/") + 1,
                                            filename.find_last_of(".") - filename.find_last_of("/")) +
                            std::string("bin");
    if (glp_tool_edgelist2bin(filename.c_str(), out.c_str()) != GLP_OK) {
        std::fprintf(stderr, "edgelist2bin: %s\n", glp_last_error());
        return 1;
    }
    return 0;
}
