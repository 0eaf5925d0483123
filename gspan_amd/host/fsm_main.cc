// fsm -- command line with the reference's interface (fsm.cu:11-43):
//   fsm <input_graph.bin> <min_support> <k> [--mode=plain|batched] [--mni=gpu|omp-literal]
// prints the pattern lines and "load: %lfs prepare: %lfs fsm: %lfs".
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>

#include "graph.h"
#include "gspan.h"
#include "gspan_amd.h"

int main(int argc, char* argv[]) {
    if (argc < 4) {
        std::printf("usage: %s input_graph min_support k [--mode=plain|batched] "
                    "[--mni=gpu|omp-literal]\n", argv[0]);
        return argc < 3 ? 0 : 1;
    }
    int mode = GS_MODE_BATCHED, mni = GS_MNI_GPU;
    for (int i = 4; i < argc; ++i) {
        if (!std::strcmp(argv[i], "--mode=plain")) mode = GS_MODE_PLAIN;
        else if (!std::strcmp(argv[i], "--mode=batched")) mode = GS_MODE_BATCHED;
        else if (!std::strcmp(argv[i], "--mni=gpu")) mni = GS_MNI_GPU;
        else if (!std::strcmp(argv[i], "--mni=omp-literal")) mni = GS_MNI_OMP_LITERAL;
        else {
            std::fprintf(stderr, "unknown option %s\n", argv[i]);
            return 1;
        }
    }
    using clk = std::chrono::steady_clock;
    try {
        auto t1 = clk::now();
        CSRGraph g;
        g.load(argv[1]);
        int min_sup = std::atoi(argv[2]);
        int k = std::atoi(argv[3]);
        auto t2 = clk::now();
        auto t3 = clk::now();
        fsm(g, k, min_sup, mode, mni);
        auto t4 = clk::now();
        auto us = [](clk::duration d) {
            return std::chrono::duration_cast<std::chrono::microseconds>(d).count() * 1e-6;
        };
        std::printf("load: %lfs prepare: %lfs fsm: %lfs\n", us(t2 - t1), us(t3 - t2), us(t4 - t3));
    } catch (const std::exception& e) {
        std::fprintf(stderr, "fsm: %s\n", e.what());
        return 1;
    }
    return 0;
}
