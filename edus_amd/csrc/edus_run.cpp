// Drop-in executable: `edus_run input.txt [outdir]` reads the reference's input
// file and writes the reference's Output/*.txt files (under outdir, default the
// current directory), advancing the k-grid on the GPU.  Replaces main_MPI.cpp.
#include <chrono>
#include <cstdio>
#include <cstdlib>

#include "../../include/edus_host.h"

int main(int argc, char **argv) {
    if (argc < 2) {
        std::printf("usage: %s <filename> [outdir] [max_steps]\n", argv[0]);
        return 1;
    }
    const char *outdir = argc > 2 ? argv[2] : ".";
    long long max_steps = argc > 3 ? std::atoll(argv[3]) : -1;
    std::printf("Starting read input...\n");
    auto t0 = std::chrono::steady_clock::now();
    edus_setup *s = nullptr;
    if (edus_setup_from_input(argv[1], 0, &s) != 0) {
        std::printf("%s\n", edus_host_last_error());
        return 1;
    }
    edus_setup_info info;
    edus_setup_get_info(s, &info);
    std::printf("Input file read: %d x %d x %d k points, %d bands, %lld steps, dt %.6f au\n", info.nk0, info.nk1, info.nk2, info.nb,
                (long long)info.itfi + 1, info.dt);
    auto t1 = std::chrono::steady_clock::now();
    long long n = edus_run(s, outdir, max_steps, 1);
    auto t2 = std::chrono::steady_clock::now();
    if (n < 0) {
        std::printf("%s\n", edus_host_last_error());
        edus_setup_free(s);
        return 1;
    }
    double setup_s = std::chrono::duration<double>(t1 - t0).count(), loop_s = std::chrono::duration<double>(t2 - t1).count();
    std::printf("Setup %.3f s; time loop %.3f s for %lld steps (%.3e k-point*steps/s)\n", setup_s, loop_s, n,
                (double)info.nk * n / loop_s);
    edus_setup_free(s);
    return 0;
}
