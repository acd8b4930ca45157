// amcfd <deck.k> : run a whole simulation through the C ABI.
//
// Same invocation as the reference executable (src/main.cpp:25-45): one
// argument, the keyword deck; --version / -v print the library version.
// Prints one JSON line per output interval and a summary line at the end.

#include <chrono>
#include <cstdio>
#include <cstring>

#include "../../include/amcfd.h"

int main(int argc, char **argv) {
    if (argc != 2) {
        std::printf("Missing parameter: input file name\n");
        return 1;
    }
    if (!std::strcmp(argv[1], "--version") || !std::strcmp(argv[1], "-v")) {
        std::printf("%s\n", amcfd_version());
        return 0;
    }
    amcfd_deck *deck = nullptr;
    int st = amcfd_deck_load(argv[1], &deck);
    if (st) {
        std::fprintf(stderr, "ERROR: %s\n", amcfd_last_error());
        return 1;
    }
    const amcfd_params *p = amcfd_deck_params(deck);
    amcfd_solver *s = nullptr;
    st = amcfd_create(p, AMCFD_MODE_FP64, &s);
    if (st) {
        std::fprintf(stderr, "ERROR: %s\n", amcfd_last_error());
        amcfd_deck_free(deck);
        return 1;
    }
    const long long cells = (long long)(p->ni - 2) * (p->nj - 2) * (p->nk_active0 - 2);
    const double outtime = p->outtime;
    double outtrack = 0.0;
    int steps = 0, finished = 0;
    auto t0 = std::chrono::steady_clock::now();
    amcfd_scalars sc;
    while (!finished) {
        st = amcfd_step(s, &finished);
        if (st < 0) {
            std::fprintf(stderr, "ERROR at step %d: %s\n", steps + 1, amcfd_last_error());
            amcfd_destroy(s);
            amcfd_deck_free(deck);
            return 2;
        }
        ++steps;
        outtrack += p->delt;
        if (outtrack >= outtime || finished) {
            outtrack = 0.0;
            amcfd_get_scalars(s, &sc);
            std::printf("{\"step\": %d, \"time\": %.9g, \"iters\": %d, \"tpeak\": %.6f, \"length\": %.6e, "
                        "\"depth\": %.6e, \"width\": %.6e, \"ratio\": %.6f}\n",
                        sc.step, sc.timet, sc.itertot, sc.tpeak, sc.alen, sc.depth, sc.width, sc.ratio);
        }
    }
    double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    amcfd_get_scalars(s, &sc);
    std::printf("{\"steps\": %d, \"seconds\": %.3f, \"cell_steps_per_s\": %.6g, \"itertot\": %d}\n", steps, secs,
                (double)cells * steps / secs, sc.itertot);
    amcfd_destroy(s);
    amcfd_deck_free(deck);
    return 0;
}
