// direct_edgelist2bin: the command-line form of
// tools/direct_edgelist2bin.cpp:208-222 (its main converts a .bin).
//   direct_edgelist2bin graph.bin [id_cutoff]   -> ./sorted_direct_graph.bin
// id_cutoff defaults to the reference's hard-coded 5013066; 0 keeps all edges.
#include <glp.h>
#include <glp_tools.h>

#include <cstdio>
#include <cstdlib>
#include <string>

int main(int argc, char** argv) {
    if (argc != 2 && argc != 3) {
        std::printf("usage: ./direct_edgelist2bin filename\n");
        return 0;
    }
    const std::string filename(argv[1]);
    const int32_t cutoff = argc == 3 ? (int32_t)std::atoi(argv[2]) : 5013066;
    // output name exactly as tools/direct_edgelist2bin.cpp:182 derives it
    const std::string out = std::string("sorted_direct_") +
                            filename.substr(filename.find_last_of("/") + 1,
                                            filename.find_last_of(".") - filename.find_last_of("/")) +
                            std::string("bin");
    if (glp_tool_direct_edgelist2bin(filename.c_str(), out.c_str(), cutoff) != GLP_OK) {
        std::fprintf(stderr, "direct_edgelist2bin: %s\n", glp_last_error());
        return 1;
    }
    return 0;
}
