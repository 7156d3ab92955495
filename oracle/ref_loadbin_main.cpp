// Driver for the reference's own CSR loader (header/graph.h, compiled
// unchanged from the reference tree by oracle/Makefile into
// oracle/_ref/ref_loadbin).  Test infrastructure only.
//
//   ref_loadbin <in.bin> <directed 0|1> <out.csr>
// writes int32 n, int32 m, offsets[n+1], neighbors[m] (little endian).
#include "graph.h"

#include <cstdio>
#include <cstdlib>

int main(int argc, char** argv) {
    if (argc != 4) {
        fprintf(stderr, "usage: ref_loadbin in.bin directed out.csr\n");
        return 2;
    }
    CSRGraph<int32_t, int32_t> g(argv[1], 0, atoi(argv[2]) != 0);
    FILE* f = fopen(argv[3], "wb");
    if (!f) return 1;
    int32_t n = g.n, m = g.m;
    fwrite(&n, 4, 1, f);
    fwrite(&m, 4, 1, f);
    fwrite(g.offsets.data(), 4, (size_t)n + 1, f);
    if (m) fwrite(g.neighbors.data(), 4, (size_t)m, f);
    fclose(f);
    return 0;
}
