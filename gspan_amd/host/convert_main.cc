// convert -- text graph to binary CSR, with the reference's interface (convert.cc:6-16):
//   convert <input.txt> <output.bin>
#include <cstdio>
#include <stdexcept>

#include "graph.h"

int main(int argc, char* argv[]) {
    if (argc < 3) {
        std::printf("usage: %s input_name output_name\n", argv[0]);
        return 0;
    }
    try {
        Graph g = Graph::from_txt(argv[1]);
        CSRGraph csr(g);
        csr.save(argv[2]);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "convert: %s\n", e.what());
        return 1;
    }
    return 0;
}
