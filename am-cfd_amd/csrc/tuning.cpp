// The mixed mode's launch-shape table, written by am-cfd_amd/autotune.py and
// read when a mixed-mode solver is created.
//
// The captured SIMPLE loop has shapes that change its speed and not its
// results: the iterations per graph launch (the scalars are read once per
// launch), the workgroups that stride over the melt-pool box, and whether a
// box variable's passes run in one persistent launch or one launch per half
// pass, the line batch of the enthalpy's x-line passes (consecutive lines
// per workgroup), and whether the next iteration's assembly runs beside the
// flow part.  Their best values depend on the grid size, so the tuner
// measures each candidate on representative grids and writes one row per
// size class:
//     max_cells  graph_block  box_blocks  [box_fused  [line_batch  [overlap]]]
// ('#' starts a comment; box_fused 0 or 1, default 1; line_batch 1..64,
// default 1; overlap 0 or 1, default 0).  A solver takes the first row whose
// max_cells is at least its interior cell count; with no row (or no table)
// it keeps the built-in 8, 256, 1, 1, 0.  The overlap is off by default
// because it pays on some grids only: measured 5% faster on the 5.12 M-cell
// track, 2-8% slower on sweep instances of 0.3-2 M cells and on the
// 9.7 M-cell multilayer deck.  The table is the file named by $AMCFD_TUNING,
// else tuning.txt beside libamcfd.so.

#include <dlfcn.h>

#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "host.h"

namespace amcfd {

namespace {

struct TuneRow {
    long long max_cells;
    int graph_block, box_blocks, box_fused, line_batch, overlap;   // overlap -1: not given
};

// rows in file order; a malformed line is an error (the whole table is refused)
int parse_table(const char *path, std::vector<TuneRow> &rows) {
    FILE *f = std::fopen(path, "r");
    if (!f) return set_error(AMCFD_E_IO, std::string("cannot open the tuning table ") + path);
    char line[512];
    int lineno = 0, st = AMCFD_OK;
    while (std::fgets(line, sizeof line, f)) {
        ++lineno;
        char *hash = std::strchr(line, '#');
        if (hash) *hash = '\0';
        long long cells;
        int gb, bb, bf = 1, lb = 1, ov = -1;
        char extra;
        bool blank = true;
        for (const char *c = line; *c; ++c) blank = blank && std::isspace((unsigned char)*c);
        if (blank) continue;   // blank or comment
        const int n = std::sscanf(line, "%lld %d %d %d %d %d %c", &cells, &gb, &bb, &bf, &lb, &ov, &extra);
        if (n < 3 || n > 6 || cells < 1 || gb < 1 || gb > 64 || bb < 1 || bb > 4096 || (bf != 0 && bf != 1) ||
            lb < 1 || lb > 64 || (n == 6 && ov != 0 && ov != 1)) {
            st = set_error(AMCFD_E_PARSE, std::string(path) + ":" + std::to_string(lineno) +
                                              ": expected 'max_cells graph_block(1..64) box_blocks(1..4096) "
                                              "[box_fused(0|1) [line_batch(1..64) [overlap(0|1)]]]'");
            break;
        }
        rows.push_back(TuneRow{cells, gb, bb, bf, lb, n == 6 ? ov : -1});
    }
    std::fclose(f);
    return st;
}

bool choose(const std::vector<TuneRow> &rows, long long cells, int *gb, int *bb, int *bf, int *lb, int *ov) {
    for (const TuneRow &r : rows)
        if (cells <= r.max_cells) {
            *gb = r.graph_block;
            *bb = r.box_blocks;
            *bf = r.box_fused;
            *lb = r.line_batch;
            if (r.overlap >= 0) *ov = r.overlap;
            return true;
        }
    return false;
}

// the shapes a solver of `cells` interior cells starts from without a table row
void built_in(int *gb, int *bb, int *bf, int *lb, int *ov) {
    *gb = 8;
    *bb = 256;
    *bf = 1;
    *lb = 1;
    *ov = 0;
}

std::string default_path() {
    const char *env = std::getenv("AMCFD_TUNING");
    if (env && *env) return env;
    Dl_info info;
    if (dladdr((void *)&default_path, &info) && info.dli_fname) {
        std::string so = info.dli_fname;
        const size_t slash = so.rfind('/');
        return (slash == std::string::npos ? std::string(".") : so.substr(0, slash)) + "/tuning.txt";
    }
    return "tuning.txt";
}

std::once_flag table_once;
std::vector<TuneRow> table;   // the library's table, loaded once

void load_default() {
    const std::string p = default_path();
    FILE *f = std::fopen(p.c_str(), "r");
    if (!f) return;   // no table: the built-in shapes
    std::fclose(f);
    std::vector<TuneRow> rows;
    if (parse_table(p.c_str(), rows) == AMCFD_OK) table = rows;
}

}  // namespace

void tuning_lookup(long long cells, int *graph_block, int *box_blocks, int *box_fused, int *line_batch,
                   int *overlap) {
    std::call_once(table_once, load_default);
    built_in(graph_block, box_blocks, box_fused, line_batch, overlap);
    choose(table, cells, graph_block, box_blocks, box_fused, line_batch, overlap);
}

}  // namespace amcfd

extern "C" {

int amcfd_tuning_choice(const char *path, long long cells, int *graph_block, int *box_blocks, int *box_fused,
                        int *line_batch, int *overlap) {
    using namespace amcfd;
    if (!graph_block || !box_blocks || !box_fused || !line_batch || !overlap)
        return set_error(AMCFD_E_ARG, "null argument");
    if (cells < 1) return set_error(AMCFD_E_ARG, "cells must be positive");
    if (!path) {
        tuning_lookup(cells, graph_block, box_blocks, box_fused, line_batch, overlap);
        return AMCFD_OK;
    }
    std::vector<TuneRow> rows;
    const int st = parse_table(path, rows);
    if (st) return st;
    built_in(graph_block, box_blocks, box_fused, line_batch, overlap);
    choose(rows, cells, graph_block, box_blocks, box_fused, line_batch, overlap);
    return AMCFD_OK;
}

}  // extern "C"
