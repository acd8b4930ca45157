// Keyword-deck reader for the AM-CFD input format.
//
// Grammar and semantics follow the reference parser:
//   Mesh::getDomainInfo       src/sources/Mesh.cpp:39-435  (keywords, blocks)
//   Mesh::assignParameters    src/sources/Mesh.cpp:440-470 (*PARAMETERS)
//   Mesh::readMeshFile        src/sources/Mesh.cpp:475-550 (*X/Y/Z_GRID_PARAMETERS)
//   Mesh::getVariableTimeSteps src/sources/Mesh.cpp:634-653
//   DomainManager::assignInputValues src/sources/DomainManager.cpp:29-263
//   DomainManager::getToolpath / getHeatSourceParameters  DomainManager.cpp:268-312
//
// Where the reference crashes or reads out of bounds (blank line, missing
// values, short toolpath rows) this reader returns AMCFD_E_PARSE instead.
// Output keywords are kept for the output writers (output.cpp); other
// output-only keywords (*OUTPUT_TEC, *CELL_OUT) are accepted and ignored.

#include "internal.h"

#include <cmath>
#include <fstream>
#include <map>
#include <numeric>
#include <sstream>
#include <string>
#include <vector>

namespace amcfd {

namespace {

// every whitespace-separated leading number on a line (istream_iterator<double>
// stops at the first token that is not a number, as in the reference)
std::vector<double> numbers(const std::string &line) {
    std::istringstream in(line);
    std::vector<double> v;
    double d;
    while (in >> d) v.push_back(d);
    return v;
}

std::vector<std::string> words(const std::string &line) {
    std::istringstream in(line);
    std::vector<std::string> v;
    std::string w;
    while (in >> w) v.push_back(w);
    return v;
}

int read_table(const std::string &path, size_t min_cols, std::vector<std::vector<double>> &rows,
               const char *what) {
    std::ifstream f(path);
    if (!f.is_open()) return set_error(AMCFD_E_IO, std::string("cannot open ") + what + " file: " + path);
    std::string line;
    while (std::getline(f, line)) {
        std::vector<double> r = numbers(line);
        if (r.empty()) continue;
        if (r.size() < min_cols)
            return set_error(AMCFD_E_PARSE, std::string(what) + " row has fewer than " +
                                                std::to_string(min_cols) + " values: " + line);
        rows.push_back(r);
    }
    return AMCFD_OK;
}

enum Block {
    B_NONE = -1, B_TRACK = 3, B_MAT, B_RELAX, B_TEMP, B_MESH, B_ID, B_PARAMS, B_ENERGY, B_DT,
    B_TERM, B_GAUSS, B_MOMENTUM, B_VECOUT, B_SCALOUT, B_CELLOUT, B_DIR, B_TEC, B_EVAP, B_SPECIES,
    B_TOOL, B_MOVING, B_FREESURF, B_DISC, B_POWDER, B_VOLHEAT, B_GEOM, B_SECTION, B_HEATTABLE,
    B_DTVAR
};

}  // namespace

int Deck::load(const std::string &path) {
    std::ifstream f(path);
    if (!f.is_open()) return set_error(AMCFD_E_IO, "cannot open deck: " + path);

    static const std::map<std::string, int> kw = {
        {"*SINGLE_TRACK_TOOL", B_TRACK}, {"*MATERIAL_PROPERTIES", B_MAT},
        {"*UNDER_RELAXATION", B_RELAX}, {"*BOUNDARY_TEMPERATURES", B_TEMP},
        {"*MESH_FILE", B_MESH}, {"*KEYWORD_ID", B_ID}, {"*PARAMETERS", B_PARAMS},
        {"*BOUNDARY_ENERGY_EQUATION", B_ENERGY}, {"*CONTROL_TIMESTEP", B_DT},
        {"*CONTROL_TERMINATION", B_TERM}, {"*GAUSS_LASER_PARAMETERS", B_GAUSS},
        {"*BOUNDARY_MOMENTUM_EQUATION", B_MOMENTUM}, {"*VECTOR_OUT", B_VECOUT},
        {"*SCALAR_OUT", B_SCALOUT}, {"*CELL_OUT", B_CELLOUT}, {"*DIRECTORY", B_DIR},
        {"*OUTPUT_TEC", B_TEC}, {"*BOUNDARY_EVAPORATION", B_EVAP}, {"*SOLVE_SPECIES", B_SPECIES},
        {"*TOOLPATH_FILE", B_TOOL}, {"*ACTIVATE_MOVING_MESH", B_MOVING},
        {"*TRACK_ENERGY_BASED_FREE_SURFACE", B_FREESURF}, {"*SELECT_DISCRETIZATION", B_DISC},
        {"*POWDER_BED_MODE", B_POWDER}, {"*VOLUMETRIC_HEAT_SOURCE", B_VOLHEAT},
        {"*OUTPUT_FINAL_GEOMETRY", B_GEOM}, {"*OUTPUT_SECTION", B_SECTION},
        {"*VOLUMETRIC_HEAT_SOURCE_TABLE", B_HEATTABLE}, {"*CONTROL_TIMESTEP_VARIABLE", B_DTVAR}};

    std::vector<double> track, mat, relax, temps, energy, gauss, momentum, evap, powder, volheat, species, freesurf,
        section;
    std::map<std::string, double> params;
    bool have_mesh = false, have_dt = false;
    int block = B_NONE, lineno = 0;
    std::string line;
    while (std::getline(f, line)) {
        ++lineno;
        if (line.empty())
            return set_error(AMCFD_E_PARSE, path + ":" + std::to_string(lineno) + ": empty line");
        if (line[0] == '$') continue;
        if (line == "*FLUID_FLOW_OFF") { flags.navierstokes = 0; continue; }
        // a flag line that, like *FLUID_FLOW_OFF, leaves the current block open (Mesh.cpp:53-62)
        if (line == "*CALCULATE_SOLIDIFICATION_PARAMETERS") { flags.solidification = 1; continue; }
        auto it = kw.find(line);
        if (it != kw.end()) {
            block = it->second;
            switch (block) {
            case B_GAUSS: flags.gaussheatsource = 1; break;
            case B_EVAP: flags.isevap = 1; break;
            case B_TOOL: flags.udtoolpath = 1; break;
            case B_POWDER: flags.powderbed = 1; break;
            case B_TEC: tecplot = 1; break;
            case B_VOLHEAT: flags.volheatsource = 1; break;
            case B_HEATTABLE: flags.volheatsource = 1; flags.volheatsourcetable = 1; break;
            case B_SPECIES: flags.species = 1; break;
            case B_MOVING: flags.movingmesh = 1; break;
            case B_FREESURF: flags.energyfreesurface = 1; break;
            case B_SECTION: flags.outputsection = 1; break;
            default: break;
            }
            continue;
        }
        if (line[0] == '*') { block = B_NONE; continue; }

        std::vector<double> v;
        std::vector<std::string> w;
        switch (block) {
        case B_TRACK: v = numbers(line); track.insert(track.end(), v.begin(), v.end()); break;
        case B_MAT: v = numbers(line); mat.insert(mat.end(), v.begin(), v.end()); break;
        case B_RELAX: v = numbers(line); relax.insert(relax.end(), v.begin(), v.end()); break;
        case B_TEMP: v = numbers(line); temps.insert(temps.end(), v.begin(), v.end()); break;
        case B_ENERGY: v = numbers(line); energy.insert(energy.end(), v.begin(), v.end()); break;
        case B_GAUSS: v = numbers(line); gauss.insert(gauss.end(), v.begin(), v.end()); break;
        case B_MOMENTUM: v = numbers(line); momentum.insert(momentum.end(), v.begin(), v.end()); break;
        case B_EVAP: v = numbers(line); evap.insert(evap.end(), v.begin(), v.end()); break;
        case B_POWDER: v = numbers(line); powder.insert(powder.end(), v.begin(), v.end()); break;
        case B_VOLHEAT: v = numbers(line); volheat.insert(volheat.end(), v.begin(), v.end()); break;
        case B_SPECIES: v = numbers(line); species.insert(species.end(), v.begin(), v.end()); break;
        case B_FREESURF: v = numbers(line); freesurf.insert(freesurf.end(), v.begin(), v.end()); break;
        case B_SECTION: v = numbers(line); section.insert(section.end(), v.begin(), v.end()); break;
        case B_MOVING:   // each line sets both 1-based zone indices, Mesh.cpp:364-369
            v = numbers(line);
            if (v.size() < 2)
                return set_error(AMCFD_E_PARSE, path + ":" + std::to_string(lineno) + ": *ACTIVATE_MOVING_MESH needs two zone indices");
            p.xmoveindex = (int)v[0] - 1;
            p.ymoveindex = (int)v[1] - 1;
            break;
        case B_MESH:
            w = words(line);
            if (!w.empty()) { mesh_file = w[0]; have_mesh = true; }
            break;
        case B_ID: w = words(line); if (!w.empty()) keyword_id = w[0]; break;
        case B_PARAMS:
            w = words(line);
            if (w.size() < 2)
                return set_error(AMCFD_E_PARSE, path + ":" + std::to_string(lineno) + ": *PARAMETERS needs name value");
            params[w[0]] = std::stod(w[1]);
            break;
        case B_DT:
            v = numbers(line);
            if (v.size() < 2)
                return set_error(AMCFD_E_PARSE, path + ":" + std::to_string(lineno) + ": *CONTROL_TIMESTEP needs dt outtime");
            p.delt = v[0];
            p.outtime = v[1];
            dt_rows.clear();
            have_dt = true;
            break;
        case B_TERM:
            v = numbers(line);
            if (v.size() < 3)
                return set_error(AMCFD_E_PARSE, path + ":" + std::to_string(lineno) + ": *CONTROL_TERMINATION needs 3 values");
            p.finaltime = v[0];
            p.maxit = v[1];
            p.nonlintol = v[2];
            break;
        case B_TOOL: w = words(line); if (!w.empty()) tool_file = w[0]; break;
        case B_DISC:
            w = words(line);
            if (!w.empty() && w[0] == "power_law") p.discretization = AMCFD_DISC_POWERLAW;
            else if (!w.empty() && w[0] == "exponential_law") p.discretization = AMCFD_DISC_EXPONENTIAL;
            break;
        case B_HEATTABLE: w = words(line); if (!w.empty()) heat_file = w[0]; break;
        case B_DTVAR: {
            w = words(line);
            if (w.empty()) break;
            dt_rows.clear();
            int st = read_table(w[0], 3, dt_rows, "variable time-step");
            if (st) return st;
            if (dt_rows.empty()) return set_error(AMCFD_E_PARSE, "empty variable time-step table: " + w[0]);
            p.delt = dt_rows[0][1];
            p.outtime = dt_rows[0][2];
            have_dt = true;
            break;
        }
        case B_SCALOUT: w = words(line); if (!w.empty()) scalar_out.push_back(w[0]); break;
        case B_VECOUT: w = words(line); if (!w.empty()) vector_out.push_back(w[0]); break;
        case B_DIR: w = words(line); if (!w.empty()) directory = w[0]; break;
        case B_GEOM: w = words(line); if (!w.empty()) geometry_out = w[0]; break;
        default: break;  // other output-only blocks
        }
    }

    if (!have_mesh)
        return set_error(AMCFD_E_PARSE, "need to specify mesh file name in " + path + " with keyword \"*MESH_FILE\"");
    if (!have_dt) return set_error(AMCFD_E_PARSE, "no *CONTROL_TIMESTEP or *CONTROL_TIMESTEP_VARIABLE");

    // *PARAMETERS, Mesh.cpp:440-470 (unset names leave the value at zero)
    auto param = [&](const char *n) { auto q = params.find(n); return q == params.end() ? 0.0 : q->second; };
    p.grav = param("accelGravity");
    p.sigm = param("radBoltzmann");
    p.tempamb = param("ambientTemp");
    p.pressamb = param("ambientPres");
    p.gasconst = param("gasConstant");

    int st = read_mesh();
    if (st) return st;

    // DomainManager::assignInputValues
    auto need = [&](const std::vector<double> &v, size_t n, const char *k) -> int {
        if (v.size() < n)
            return set_error(AMCFD_E_PARSE, std::string(k) + " needs " + std::to_string(n) + " values");
        return AMCFD_OK;
    };
    if (!flags.udtoolpath) {
        if ((st = need(track, 4, "*SINGLE_TRACK_TOOL"))) return st;
        p.lasertime = track[0]; p.xstart = track[1]; p.ystart = track[2]; p.scanvel = track[3];
    }
    if ((st = need(mat, 15, "*MATERIAL_PROPERTIES"))) return st;
    p.dens = mat[0]; p.denl = mat[1]; p.viscos = mat[2]; p.tsolid = mat[3]; p.tliquid = mat[4];
    p.hsmelt = mat[5]; p.hlfriz = mat[6]; p.acpa = mat[7]; p.acpb = mat[8]; p.acpl = mat[9];
    p.thconsa = mat[10]; p.thconsb = mat[11]; p.thconsc = mat[12]; p.thconla = mat[13]; p.thconlb = mat[14];
    if ((st = need(relax, 5, "*UNDER_RELAXATION"))) return st;
    p.urfu = relax[0]; p.urfv = relax[1]; p.urfw = relax[2]; p.urfp = relax[3]; p.urfh = relax[4];
    if ((st = need(energy, 5, "*BOUNDARY_ENERGY_EQUATION"))) return st;
    p.htci = energy[0]; p.htcj = energy[1]; p.htck1 = energy[2]; p.htckn = energy[3]; p.emiss = energy[4];
    if ((st = need(temps, 5, "*BOUNDARY_TEMPERATURES"))) return st;
    p.tempwest = temps[0]; p.tempeast = temps[1]; p.tempnorth = temps[2]; p.tempbottom = temps[3];
    p.temppreheat = temps[4];
    if ((st = need(momentum, 4, "*BOUNDARY_MOMENTUM_EQUATION"))) return st;
    p.beta = momentum[0]; p.dgdtp = momentum[1]; p.darcyKo = momentum[2]; p.a_oxygen = momentum[3];
    p.alaseta = 0.0;          // DomainManager.h:34-35 defaults
    p.alasrb = 1.0e-6;
    if (flags.gaussheatsource) {
        if ((st = need(gauss, 4, "*GAUSS_LASER_PARAMETERS"))) return st;
        p.alaspow = gauss[0]; p.alaseta = gauss[1]; p.alasrb = gauss[2]; p.alasfact = gauss[3];
    }
    if (flags.isevap) {
        if ((st = need(evap, 4, "*BOUNDARY_EVAPORATION"))) return st;
        p.evaplatent = evap[0]; p.evaptemp = evap[1]; p.evapfact = evap[2]; p.molarmass = evap[3];
    }
    if (flags.energyfreesurface) {   // DomainManager.cpp:133-144
        if ((st = need(freesurf, 8, "*TRACK_ENERGY_BASED_FREE_SURFACE"))) return st;
        p.initlam1 = freesurf[0]; p.initlam2 = freesurf[1]; p.surftens = freesurf[2]; p.maxiterenergy = freesurf[3];
        p.tolenergy = freesurf[4]; p.maxiterbisect = freesurf[5]; p.tolbisect = freesurf[6]; p.urfree = freesurf[7];
    }
    if (flags.outputsection) {   // DomainManager.cpp:189-198
        if ((st = need(section, 7, "*OUTPUT_SECTION"))) return st;
        p.gxmin = section[0]; p.gymin = section[1]; p.gzmin = section[2];
        p.gxmax = section[3]; p.gymax = section[4]; p.gzmax = section[5]; p.toutsect = section[6];
    }
    if (flags.species) {   // DomainManager.cpp:123-130
        if ((st = need(species, 4, "*SOLVE_SPECIES"))) return st;
        p.initconcentration = species[0]; p.massdiff = species[1]; p.kp = species[2]; p.urfc = species[3];
    }
    if (flags.powderbed) {
        if ((st = need(powder, 8, "*POWDER_BED_MODE"))) return st;
        p.numlayers = (int)powder[0]; p.ncvzlayer = (int)powder[1]; p.layerheight = powder[2];
        p.powderden = powder[3]; p.powdercpa = powder[4]; p.powdercpb = powder[5];
        p.powderthcona = powder[6]; p.powderthconb = powder[7];
        // build layers are appended to the last z zone, DomainManager.cpp:159-161
        p.zzone[p.nzz - 1] += p.numlayers * p.layerheight;
        p.ncvz[p.nzz - 1] += p.numlayers * p.ncvzlayer;
    }
    if (flags.volheatsource) {
        if (!flags.volheatsourcetable) {
            if ((st = need(volheat, 5, "*VOLUMETRIC_HEAT_SOURCE"))) return st;
            p.avolpow = volheat[0]; p.apowseta = volheat[1]; p.heatthick = volheat[2];
            p.heatrb = volheat[3]; p.avolfact = volheat[4];
        } else {
            if ((st = read_table(heat_file, 6, heat_rows, "heat-source table"))) return st;
            if (heat_rows.empty()) return set_error(AMCFD_E_PARSE, "empty heat-source table: " + heat_file);
            p.avolpow = heat_rows[0][1]; p.apowseta = heat_rows[0][2]; p.heatthick = heat_rows[0][3];
            p.heatrb = heat_rows[0][4]; p.avolfact = heat_rows[0][5];
        }
    }
    if (flags.udtoolpath) {
        if ((st = read_table(tool_file, 6, tool_rows, "toolpath"))) return st;
        if (tool_rows.size() < 2) return set_error(AMCFD_E_PARSE, "toolpath needs at least two rows: " + tool_file);
    }
    flags.steady = p.delt < 100.0 ? 0 : 1;
    if (flags.movingmesh) {
        // generateGrid shifts xzone[ix-1]/xzone[ix+1] and, until the pool length
        // is steady, yzone[ix-1]/yzone[ix+1] (with the x index), afterwards
        // yzone[iy-1]/yzone[iy+1] (CFDSolverManager.cpp:102-118)
        const int ix = p.xmoveindex, iy = p.ymoveindex;
        if (ix < 1 || ix > p.nzx - 2 || iy < 1 || iy > p.nzy - 2 || ix > p.nzy - 2)
            return set_error(AMCFD_E_PARSE, "*ACTIVATE_MOVING_MESH zone indices need a zone on either side in x and y "
                                            "(x index " + std::to_string(ix + 1) + ", y index " +
                                            std::to_string(iy + 1) + ")");
    }

    double sx = 0, sy = 0, sz = 0;
    for (int z = 0; z < p.nzx; ++z) sx += p.ncvx[z];
    for (int z = 0; z < p.nzy; ++z) sy += p.ncvy[z];
    for (int z = 0; z < p.nzz; ++z) sz += p.ncvz[z];
    // std::accumulate with an int initial value truncates each partial sum
    int isx = 0, isy = 0, isz = 0;
    for (int z = 0; z < p.nzx; ++z) isx = (int)(isx + p.ncvx[z]);
    for (int z = 0; z < p.nzy; ++z) isy = (int)(isy + p.ncvy[z]);
    for (int z = 0; z < p.nzz; ++z) isz = (int)(isz + p.ncvz[z]);
    (void)sx; (void)sy; (void)sz;
    p.ni = isx + 2;
    p.nj = isy + 2;
    p.nk = isz + 2;
    p.nk_active0 = flags.powderbed ? p.nk - (p.numlayers - 1) * p.ncvzlayer : p.nk;
    if (p.ni < 4 || p.nj < 4 || p.nk < 4 || p.nk_active0 < 4)
        return set_error(AMCFD_E_PARSE, "grid needs at least two cells per direction");
    double cpavg = (p.acpa * p.tsolid + p.acpb + p.acpl) * 0.5;   // DomainManager.cpp:252-254
    p.hlcal = p.hsmelt + cpavg * (p.tliquid - p.tsolid);
    p.hlatnt = p.hlfriz - p.hlcal;

    finalize_tables();
    return AMCFD_OK;
}

int Deck::read_mesh() {
    std::ifstream f(mesh_file);
    if (!f.is_open()) return set_error(AMCFD_E_IO, "specified mesh was not found, searched in: " + mesh_file);
    std::vector<double> g[3];
    int opt = -1, lineno = 0;
    std::string line;
    while (std::getline(f, line)) {
        ++lineno;
        if (line.empty())
            return set_error(AMCFD_E_PARSE, mesh_file + ":" + std::to_string(lineno) + ": empty line");
        if (line[0] == '$') continue;
        if (line == "*X_GRID_PARAMETERS") { opt = 0; continue; }
        if (line == "*Y_GRID_PARAMETERS") { opt = 1; continue; }
        if (line == "*Z_GRID_PARAMETERS") { opt = 2; continue; }
        if (line[0] == '*') { opt = -1; continue; }
        if (opt >= 0) {
            std::vector<double> v = numbers(line);
            g[opt].insert(g[opt].end(), v.begin(), v.end());
        }
    }
    int *nz[3] = {&p.nzx, &p.nzy, &p.nzz};
    double *zone[3] = {p.xzone, p.yzone, p.zzone}, *ncv[3] = {p.ncvx, p.ncvy, p.ncvz},
           *pw[3] = {p.powrx, p.powry, p.powrz};
    for (int a = 0; a < 3; ++a) {
        if (g[a].empty()) return set_error(AMCFD_E_PARSE, "mesh file lacks a grid block: " + mesh_file);
        int n = (int)g[a][0];
        if (n < 1 || n > AMCFD_MAX_ZONES || (int)g[a].size() < 1 + 3 * n)
            return set_error(AMCFD_E_PARSE, "bad grid zone block in " + mesh_file);
        *nz[a] = n;
        for (int z = 0; z < n; ++z) {
            zone[a][z] = g[a][1 + z];
            ncv[a][z] = g[a][1 + n + z];
            pw[a][z] = g[a][1 + 2 * n + z];
        }
    }
    return AMCFD_OK;
}

void Deck::finalize_tables() {
    flat_tool.clear();
    for (auto &r : tool_rows) flat_tool.insert(flat_tool.end(), r.begin(), r.begin() + 6);
    flat_dt.clear();
    for (auto &r : dt_rows) flat_dt.insert(flat_dt.end(), r.begin(), r.begin() + 3);
    flat_heat.clear();
    for (auto &r : heat_rows) flat_heat.insert(flat_heat.end(), r.begin(), r.begin() + 6);
    p.n_toolpath = (int)tool_rows.size();
    p.toolpath = flat_tool.empty() ? nullptr : flat_tool.data();
    p.n_dt_table = (int)dt_rows.size();
    p.dt_table = flat_dt.empty() ? nullptr : flat_dt.data();
    p.n_heat_table = (int)heat_rows.size();
    p.heat_table = flat_heat.empty() ? nullptr : flat_heat.data();
    p.navierstokes = flags.navierstokes;
    p.steady = flags.steady;
    p.isevap = flags.isevap;
    p.powderbed = flags.powderbed;
    p.udtoolpath = flags.udtoolpath;
    p.movingmesh = flags.movingmesh;
    p.species = flags.species;
    p.energyfreesurface = flags.energyfreesurface;
    p.outputsection = flags.outputsection;
    p.solidification = flags.solidification;
    p.gaussheatsource = flags.gaussheatsource;
    p.volheatsource = flags.volheatsource;
    p.volheatsourcetable = flags.volheatsourcetable;
}

}  // namespace amcfd
