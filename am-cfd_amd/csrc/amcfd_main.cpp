// amcfd <deck.k> [--mode fp64|scan|mixed] : run a whole simulation through the C ABI.
//
// Same invocation and the same output files as the reference executable
// (src/main.cpp:25-241): one argument, the keyword deck (--version / -v print
// the library version).  Output follows the reference's sequence: the VTS file
// <dir>/<id>_0.vts after set-up, then at every output interval (transient runs)
// a residual block in <id>_Residuals.txt and the next VTS file, and at the end
// one more of each plus the final melt-pool geometry when the deck asks for
// it; <id>.pvd is rewritten after every VTS file.  <dir> is the deck's
// *DIRECTORY (default "Results"), <id> its *KEYWORD_ID.  With
// *TRACK_ENERGY_BASED_FREE_SURFACE the free surface is evaluated before each
// of those outputs (monitor in <id>_FreeSurfaceMonitor.txt); with
// *OUTPUT_SECTION every step may append an extracted section, and the end
// writes one more.  Also prints one JSON line per output and a summary line.

#include <sys/stat.h>

#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/amcfd.h"

static int fail(const char *what) {
    std::fprintf(stderr, "ERROR (%s): %s\n", what, amcfd_last_error());
    return 2;
}

int main(int argc, char **argv) {
    if (argc < 2) {
        std::printf("Missing parameter: input file name\n");
        return 1;
    }
    if (!std::strcmp(argv[1], "--version") || !std::strcmp(argv[1], "-v")) {
        std::printf("%s\n", amcfd_version());
        return 0;
    }
    int mode = AMCFD_MODE_FP64;
    for (int a = 2; a + 1 < argc; a += 2)
        if (!std::strcmp(argv[a], "--mode"))
            mode = !std::strcmp(argv[a + 1], "scan")    ? AMCFD_MODE_SCAN
                   : !std::strcmp(argv[a + 1], "mixed") ? AMCFD_MODE_MIXED
                                                        : AMCFD_MODE_FP64;

    amcfd_deck *deck = nullptr;
    if (amcfd_deck_load(argv[1], &deck)) return fail("deck");
    const amcfd_params *p = amcfd_deck_params(deck);
    amcfd_solver *s = nullptr;
    if (amcfd_create(p, mode, &s)) {
        amcfd_deck_free(deck);
        return fail("create");
    }

    // output names (Mesh.h:43-49 defaults)
    const char *dir_c = amcfd_deck_string(deck, "directory", 0);
    const char *id_c = amcfd_deck_string(deck, "keyword_id", 0);
    const char *geom_c = amcfd_deck_string(deck, "final_geometry", 0);
    const std::string dir = dir_c ? dir_c : "Results", id = id_c ? id_c : "";
    const std::string folder = dir + "/", pvd = id + ".pvd", resid = id + "_Residuals.txt";
    const std::string monitor = id + "_FreeSurfaceMonitor.txt";
    std::vector<const char *> names;
    for (int q = 0; amcfd_deck_string(deck, "scalar_out", q); ++q) names.push_back(amcfd_deck_string(deck, "scalar_out", q));
    if (amcfd_deck_string(deck, "vector_out", 0))
        std::fprintf(stderr, "note: *VECTOR_OUT outputs are not written\n");
    mkdir(dir.c_str(), 0777);

    std::vector<std::string> files;
    std::vector<double> times;
    int outct = 0;
    amcfd_scalars sc;
    auto visual_out = [&]() -> int {
        const std::string file = folder + id + "_" + std::to_string(outct) + ".vts";
        if (amcfd_write_vts(s, file.c_str(), names.data(), (int)names.size())) return fail("vts");
        if (amcfd_get_scalars(s, &sc)) return fail("scalars");
        files.push_back(file);
        times.push_back(sc.timet);
        std::vector<const char *> fp;
        for (auto &f : files) fp.push_back(f.c_str());
        if (amcfd_write_pvd(pvd.c_str(), (int)files.size(), times.data(), fp.data())) return fail("pvd");
        return 0;
    };
    auto report = [&]() {
        std::printf("{\"step\": %d, \"time\": %.9g, \"iters\": %d, \"tpeak\": %.6f, \"length\": %.6e, "
                    "\"depth\": %.6e, \"width\": %.6e, \"ratio\": %.6f}\n",
                    sc.step, sc.timet, sc.itertot, sc.tpeak, sc.alen, sc.depth, sc.width, sc.ratio);
    };

    int rc = visual_out();
    const long long cells = (long long)(p->ni - 2) * (p->nj - 2) * (p->nk_active0 - 2);
    double outtrack = 0.0;
    int steps = 0, finished = 0;
    auto t0 = std::chrono::steady_clock::now();
    while (!rc && !finished) {
        if (amcfd_step(s, &finished) < 0) {
            rc = fail("step");
            break;
        }
        ++steps;
        if (amcfd_get_scalars(s, &sc)) {
            rc = fail("scalars");
            break;
        }
        outtrack += sc.delt;
        int written = 0;
        if (p->outputsection && amcfd_output_section(s, id.c_str(), 0, &written)) {   // main.cpp:155-156
            rc = fail("section");
            break;
        }
        if (!p->steady && outtrack >= sc.outtime) {
            ++outct;
            if (p->energyfreesurface && amcfd_free_surface(s, monitor.c_str())) { rc = fail("free surface"); break; }
            if (amcfd_append_residuals(s, resid.c_str())) { rc = fail("residuals"); break; }
            if ((rc = visual_out())) break;
            outtrack = 0.0;
            report();
        }
    }
    if (!rc) {   // main.cpp:204-225
        int written = 0;
        if (p->energyfreesurface && amcfd_free_surface(s, monitor.c_str())) rc = fail("free surface");
        if (!rc && p->outputsection && amcfd_output_section(s, id.c_str(), 1, &written)) rc = fail("section");
    }
    if (!rc) {
        ++outct;
        if (amcfd_append_residuals(s, resid.c_str())) rc = fail("residuals");
        else rc = visual_out();
        if (!rc && geom_c && amcfd_append_final_geometry(s, geom_c)) rc = fail("geometry");
        report();
    }
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("{\"steps\": %d, \"seconds\": %.3f, \"cell_steps_per_s\": %.6g, \"itertot\": %d}\n", steps, secs,
                (double)cells * steps / secs, sc.itertot);
    amcfd_destroy(s);
    amcfd_deck_free(deck);
    return rc;
}
