// Output files in the reference's formats, written from the device-resident
// solver's fields, so that a run of the CLI leaves the same files as the
// reference executable:
//   <dir>/<id>_<n>.vts   vtsBinWriter::visualOut, vtsBinWriter.cpp:50-139,
//                        144-257, 373-417, 497-526 (StructuredGrid, appended
//                        raw little-endian data)
//   <id>.pvd             vtsBinWriter::initializePVD/appendPVD/closePVD,
//                        vtsBinWriter.cpp:420-468
//   <id>_Residuals.txt   PrintManager::outputResiduals, PrintManager.cpp:380-533
//   final geometry       PrintManager::outputFinalGeometry, PrintManager.cpp:658-677
// The byte layout (XML text, offsets, record markers, iostream number
// formats) is the reference's, since readers of those files depend on it.
// Velocities are written as the reference writes them: averaged from the
// staggered faces to the cell centres and zeroed in solid cells
// (vtsBinWriter.cpp:497-526).  Vector outputs are not written: the
// reference's vector path reads three values per point from a scalar field
// (past its end).

#include "host.h"

#include <cmath>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

namespace amcfd {

namespace {

// the names the reference's solution registry holds for this run
// (CFDSolverManager.cpp:528-560, PrintManager.cpp:283-285,
// FreeSurfaceManager.cpp:64-65)
bool scalar_output_known(const std::string &n, const amcfd_params &p) {
    static const char *base[] = {"pp", "vis", "rho", "uvel", "vvel", "wvel", "temp", "diff", "hnot", "fracl",
                                 "fraclnot", "pressure", "enthalpy", "sourceinput", "auvel", "avvel", "awvel"};
    for (const char *b : base)
        if (n == b) return true;
    if (p.powderbed && (n == "csfrac" || n == "solfrac" || n == "mpPhase" || n == "mpLiqPhase")) return true;
    if (p.species && (n == "avgconcentration" || n == "concentration")) return true;
    if (p.solidification && (n == "rate" || n == "grad")) return true;
    if (n == "tempr" || n == "zr") return true;   // FreeSurfaceManager.cpp:64-65, registered for every deck
    return false;
}

// face velocities to cell centres, zero in solid cells (transferVelNodes)
void centre_velocities(Solver &s, std::vector<double> &au, std::vector<double> &av, std::vector<double> &aw,
                       int &st) {
    int ni, nj, nk;
    s.dims(&ni, &nj, &nk);
    const size_t nn = (size_t)ni * nj * nk;
    std::vector<double> u(nn), v(nn), w(nn), t(nn);
    st = s.get_field("uvel", u.data(), nn);
    if (!st) st = s.get_field("vvel", v.data(), nn);
    if (!st) st = s.get_field("wvel", w.data(), nn);
    if (!st) st = s.get_field("temp", t.data(), nn);
    if (st) return;
    au.assign(nn, 0.0);
    av.assign(nn, 0.0);
    aw.assign(nn, 0.0);
    const double ts = s.params().tsolid;
    const int nkm1 = s.active_nk() - 1;
    for (int k = 1; k < nkm1; ++k)
        for (int j = 1; j < nj - 1; ++j)
            for (int i = 1; i < ni - 1; ++i) {
                const size_t c = ((size_t)k * nj + j) * ni + i;
                au[c] = (u[c] + u[c + 1]) * 0.5;
                av[c] = (v[c] + v[c + ni]) * 0.5;
                aw[c] = (w[c] + w[c + (size_t)ni * nj]) * 0.5;
                if (t[c] <= ts) {
                    au[c] = 0.0;
                    av[c] = 0.0;
                    aw[c] = 0.0;
                }
            }
}

}  // namespace

int write_vts(Solver &s, const char *path, const char *const *names, int nnames) {
    int ni, nj, nk;
    s.dims(&ni, &nj, &nk);
    const amcfd_params &p = s.params();
    const int gridz = nk;   // the allocated height (before powder layers are hidden)
    const int npoints = ni * nj * (gridz - 1);
    std::vector<std::string> out;
    for (int q = 0; q < nnames; ++q) {
        std::string n = names[q] ? names[q] : "";
        if (!scalar_output_known(n, p)) continue;   // the reference lists these as unsupported
        if (n == "uvel") n = "auvel";
        if (n == "vvel") n = "avvel";
        if (n == "wvel") n = "awvel";
        out.push_back(n);
    }
    const size_t nn = (size_t)ni * nj * nk;
    std::vector<double> au, av, aw;
    int st = 0;
    centre_velocities(s, au, av, aw, st);
    if (st) return st;

    std::ofstream f(path, std::ios::out | std::ios::binary);
    if (!f) return set_error(AMCFD_E_IO, std::string("cannot write ") + path);
    int off = 0;
    f << "<?xml version=\"1.0\"?> ";
    f << "<VTKFile type=\"StructuredGrid\" version=\"2.1\" byte_order=\"LittleEndian\"> ";
    f << "<StructuredGrid WholeExtent=\"" << 0 << " " << ni - 1 << " " << 0 << " " << nj - 1 << " " << 0 << " "
      << gridz - 2 << "\"> ";
    f << "<Piece Extent=\"" << 0 << " " << ni - 1 << " " << 0 << " " << nj - 1 << " " << 0 << " " << gridz - 2
      << "\"> ";
    f << "<Points> ";
    f << "<DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"appended\" offset = \"" << off << "\" /> ";
    off += (int)(sizeof(double) * npoints * 3 + sizeof(int));
    f << "</Points> ";
    f << "<PointData> ";
    for (const std::string &n : out) {
        f << "<DataArray type=\"Float64\" Name=\"" << n << "\" format=\"appended\" offset=\"" << off << "\" />  ";
        off += (int)(sizeof(double) * npoints + sizeof(int));
    }
    if (p.powderbed) {
        f << "<DataArray type=\"UInt8\" Name=\"vtkGhostType\" format=\"appended\" offset=\"" << off << "\" />  ";
        off += (int)(sizeof(unsigned char) * npoints + sizeof(int));
    }
    f << "</PointData> ";
    f << "<CellData> ";
    f << "</CellData> ";
    f << "</Piece> ";
    f << "</StructuredGrid> ";
    f << "<AppendedData encoding=\"raw\">_";

    const HostGrid &g = s.grid();
    std::vector<double> buf(nn);
    // with the free surface the point heights are the shifted ones (vtsBinWriter.cpp:163-170)
    if (p.energyfreesurface && (st = s.get_field("zr", buf.data(), nn))) return st;
    int bytes = (int)(sizeof(double) * npoints * 3);
    f.write((const char *)&bytes, sizeof(int));
    for (int k = 0; k < gridz - 1; ++k)
        for (int j = 0; j < nj; ++j)
            for (int i = 0; i < ni; ++i) {
                f.write((const char *)&g.x[i], sizeof(double));
                f.write((const char *)&g.y[j], sizeof(double));
                const double *zp = p.energyfreesurface ? &buf[((size_t)k * nj + j) * ni + i] : &g.z[k];
                f.write((const char *)zp, sizeof(double));
            }
    for (const std::string &n : out) {
        const double *src;
        if (n == "auvel") src = au.data();
        else if (n == "avvel") src = av.data();
        else if (n == "awvel") src = aw.data();
        else {
            if ((st = s.get_field(n.c_str(), buf.data(), nn))) return st;
            src = buf.data();
        }
        bytes = (int)(sizeof(double) * npoints);
        f.write((const char *)&bytes, sizeof(int));
        f.write((const char *)src, sizeof(double) * (size_t)npoints);
    }
    if (p.powderbed) {
        bytes = (int)(sizeof(unsigned char) * npoints);
        f.write((const char *)&bytes, sizeof(int));
        const int nkm1 = s.active_nk() - 1;
        const unsigned char visible = 0, hidden = 2;
        for (int k = 0; k < gridz - 1; ++k)
            for (int j = 0; j < nj; ++j)
                for (int i = 0; i < ni; ++i) f.write((const char *)(k < nkm1 ? &visible : &hidden), 1);
    }
    f << "\n";
    f << "</AppendedData> </VTKFile>";
    f.close();
    if (!f) return set_error(AMCFD_E_IO, std::string("write failed: ") + path);
    return AMCFD_OK;
}

int write_pvd(const char *path, int n, const double *times, const char *const *files) {
    std::ofstream f(path, std::ios::out | std::ios::binary);
    if (!f) return set_error(AMCFD_E_IO, std::string("cannot write ") + path);
    f << "<?xml version=\"1.0\"?>\n";
    f << "<VTKFile type=\"Collection\" version=\"2.1\" byte_order=\"LittleEndian\">\n";
    f << "  " << "<Collection>\n";
    for (int q = 0; q < n; ++q)
        f << "   " << "<DataSet timestep=\"" << times[q] << "\" group=\"\" part=\"0\" file=\"" << files[q]
          << "\"/>\n";
    f << "  " << "</Collection>\n";
    f << "</VTKFile>\n";
    f.close();
    if (!f) return set_error(AMCFD_E_IO, std::string("write failed: ") + path);
    return AMCFD_OK;
}

int append_residuals(Solver &s, const char *path) {
    int ni, nj, nk;
    s.dims(&ni, &nj, &nk);
    const size_t nn = (size_t)ni * nj * nk;
    const amcfd_params &p = s.params();
    amcfd_scalars sc;
    int st = s.get_scalars(&sc);
    if (st) return st;
    const DevScalars &d = s.host_scalars();
    const int nkm2 = s.active_nk() - 2;
    std::vector<double> t(nn);
    if ((st = s.get_field("temp", t.data(), nn))) return st;

    // peak surface velocities over the melt-pool box (PrintManager.cpp:405-417)
    double umax = 0.0, vmax = 0.0, wmax = 0.0;
    if (sc.tpeak > p.tsolid) {
        std::vector<double> u(nn), v(nn), w(nn);
        if ((st = s.get_field("uvel", u.data(), nn)) || (st = s.get_field("vvel", v.data(), nn)) ||
            (st = s.get_field("wvel", w.data(), nn)))
            return st;
        for (int j = sc.jstat; j <= sc.jend; ++j)
            for (int i = sc.istat + 1; i <= sc.iend - 1; ++i) {
                const size_t c = ((size_t)nkm2 * nj + j) * ni + i;
                umax = std::max(umax, std::fabs(u[c]));
                vmax = std::max(vmax, std::fabs(v[c]));
                wmax = std::max(wmax, std::fabs(w[c]));
            }
    }
    // cooling rates along the track's surface line (PrintManager.cpp:419-455);
    // where the reference leaves a value unassigned this writes NaN
    const double nan = std::numeric_limits<double>::quiet_NaN();
    double xso = nan, tso = nan, xsob = nan, tsob = nan, xliq = nan, tliq = nan;
    const std::vector<double> &x = s.grid().x;
    const size_t line = ((size_t)nkm2 * nj + sc.jstart) * ni;
    for (int i = 1; i < ni - 1; ++i)
        if (t[line + i] > p.tsolid - 290) {
            tsob = t[line + i - 1];
            xsob = x[i - 1];
            break;
        }
    for (int i = 1; i < ni - 1; ++i)
        if (t[line + i] > p.tsolid) {
            tso = t[line + i - 1];
            xso = x[i - 1];
            break;
        }
    for (int i = 1; i < ni - 1; ++i)
        if (t[line + i] > p.tliquid) {
            tliq = t[line + i];
            xliq = x[i];
            break;
        }
    const double crs = (tso - tsob) / (xso - xsob) * sc.scanvel;
    const double crl = (tliq - tso) / (xliq - xso) * sc.scanvel;
    const double resorc = sc.resorc;   // 0 without species (the reference's member starts at zero)

    std::ofstream f(path, std::ios::out | std::ios::app);
    if (!f) return set_error(AMCFD_E_IO, std::string("cannot write ") + path);
    using std::fixed;
    using std::scientific;
    using std::setprecision;
    using std::setw;
    const char *rule = "====================================================================================================\n";
    if (p.steady) {
        f << rule;
        f << "  iter    res_enth   res_mass   res_u      res_v      res_w      res_concen\n";
        f << setprecision(2) << fixed << setw(8) << sc.niter << scientific << setw(15) << sc.resorh << setw(11)
          << sc.resorm << setw(11) << sc.resoru << setw(11) << sc.resorv << setw(11) << sc.resorw << setw(11)
          << resorc << " \n";
    } else {
        f << rule;
        f << "  time        iter  tot_iter    res_enth   res_mass   res_u      res_v      res_w      res_concen\n";
        f << setprecision(2) << scientific << setw(10) << sc.timet << fixed << setw(6) << sc.niter << setw(9)
          << sc.itertot << scientific << setw(15) << sc.resorh << setw(11) << sc.resorm << setw(11) << sc.resoru
          << setw(11) << sc.resorv << setw(11) << sc.resorw << setw(11) << resorc << " \n";
    }
    f << "  Tmax       umax       vmax       wmax        length      depth       width\n";
    f << setprecision(2) << fixed << setw(10) << sc.tpeak << scientific << setw(11) << umax << setw(11) << vmax
      << setw(11) << wmax << setw(12) << sc.alen << setw(12) << sc.depth << setw(12) << sc.width << "\n";
    f << "  north    south     top   toploss   bottom    west      east    hout    accu    hin      ratio\n";
    f << setprecision(1) << fixed << setw(6) << d.fluxnorth << setw(9) << d.fluxsouth << setw(9) << d.fluxtop
      << setw(9) << sc.ahtoploss << setw(9) << d.fluxbottom << setw(9) << d.fluxwest << setw(9) << d.fluxeast
      << setw(9) << sc.heatout << setw(8) << sc.accul << setw(8) << sc.heatinlaser << setw(8) << sc.ratio
      << (p.steady ? "\n" : "\n\n");
    if (!p.steady) {
        f << setprecision(4) << scientific << "  Cooling rate from solidus to 290K below solidus:     " << setw(14)
          << crs << "\n"
          << "  Cooling rate from liquidus to solidus:               " << setw(14) << crl << "\n";
    }
    f << rule;
    f.close();
    if (!f) return set_error(AMCFD_E_IO, std::string("write failed: ") + path);
    return AMCFD_OK;
}

int append_final_geometry(Solver &s, const char *path) {
    amcfd_scalars sc;
    int st = s.get_scalars(&sc);
    if (st) return st;
    std::ofstream f(path, std::ios::out | std::ios::app);
    if (!f) return set_error(AMCFD_E_IO, std::string("cannot write ") + path);
    using std::fixed;
    using std::scientific;
    using std::setprecision;
    using std::setw;
    f << "\n This is the final melt pool geometry for the given process parameters\n";
    f << "===================================================================\n";
    f << "  Laser Power    Laser Speed      length      depth      width\n";
    f << setprecision(2) << fixed << setw(8) << s.params().alaspow << setw(13) << sc.scanvel << scientific
      << setw(21) << sc.alen << setw(12) << sc.depth << setw(11) << sc.width << "\n";
    f << "===================================================================\n";
    f.close();
    if (!f) return set_error(AMCFD_E_IO, std::string("write failed: ") + path);
    return AMCFD_OK;
}

// the free-surface monitor block, FreeSurfaceManager.cpp:197-206
int append_free_surface_monitor(Solver &s, const char *path) {
    amcfd_scalars sc;
    int st = s.get_scalars(&sc);
    if (st) return st;
    const Solver::FsMonitor &m = s.fs_monitor_;
    std::ofstream f(path, std::ios::out | std::ios::app);
    if (!f) return set_error(AMCFD_E_IO, std::string("cannot write ") + path);
    using std::scientific;
    using std::setprecision;
    using std::setw;
    const char *rule = "====================================================================================================\n";
    f << rule;
    f << "  time         lambda        res_bisect    min_fi         max_fi        droplet_volume   iter_bisect\n";
    f << setprecision(4) << scientific << setw(12) << sc.timet << setw(14) << m.xmid << setw(14) << m.fmid
      << setw(14) << m.fimin << setw(14) << m.fimax << setw(14) << m.vd << setw(9) << m.niter << "\n";
    f << rule;
    f.close();
    if (!f) return set_error(AMCFD_E_IO, std::string("write failed: ") + path);
    return AMCFD_OK;
}

// C-style conversion of a double to int as x86-64 performs it (cvttsd2si):
// out-of-range and NaN values give INT_MIN
static int x86_int(double v) {
    if (!(v > -2147483649.0 && v < 2147483648.0)) return std::numeric_limits<int>::min();
    return (int)v;
}

// PrintManager::outputSection, PrintManager.cpp:538-624: temperatures in a
// box of nodes, appended to <base>_ExtractedSection<layer>.txt.  The layer
// number divides by the powder layer height, which is zero without a powder
// bed (the reference's file is then ...Section-2147483648.txt).  The stream
// keeps its scientific format from the first rows on, so the time of every
// block after the first one with rows is printed in scientific notation.
int write_section(Solver &s, const char *base, bool &sci) {
    int ni, nj, nk;
    s.dims(&ni, &nj, &nk);
    const int gridz = nk;
    const amcfd_params &p = s.params();
    const HostGrid &g = s.grid();
    amcfd_scalars sc;
    int st = s.get_scalars(&sc);
    if (st) return st;
    const int nkm1 = s.active_nk() - 1;
    const size_t nn = (size_t)ni * nj * nk;
    std::vector<double> t(nn);
    if ((st = s.get_field("temp", t.data(), nn))) return st;
    auto inx = [&](int i) { return g.x[i] >= p.gxmin && g.x[i] <= p.gxmax; };
    auto iny = [&](int j) { return g.y[j] >= p.gymin && g.y[j] <= p.gymax; };
    auto inz = [&](int k) { return g.z[k] >= p.gzmin && g.z[k] <= p.gzmax; };
    int outx = 0, outy = 0, outz = 0;
    for (int i = 0; i < ni; ++i) outx += inx(i);
    for (int j = 0; j < nj; ++j) outy += iny(j);
    for (int k = 0; k < gridz - 1; ++k) outz += inz(k);
    // baseheightk is set for powder beds only (PrintManager.cpp:72); zero otherwise
    const int base_k = p.powderbed ? p.nk_active0 - 2 - p.ncvzlayer : 0;
    const double tol = 1e-9;
    const int layer = x86_int((g.z[nkm1] - g.z[base_k] + tol) / p.layerheight);
    const std::string path = std::string(base) + "_ExtractedSection" + std::to_string(layer) + ".txt";
    std::ofstream f(path, std::ios::out | std::ios::app);
    if (!f) return set_error(AMCFD_E_IO, "cannot write " + path);
    using std::scientific;
    using std::setprecision;
    using std::setw;
    if (sci) f << setprecision(6) << scientific;
    f << "t " << "ix " "jy " << "kz " << "\n" << sc.timet << " " << outx << " " << outy << " " << outz << "\n";
    f << setw(15) << "x" << setw(15) << "y" << setw(15) << "z" << setw(15) << "tn" << "\n";
    for (int k = 0; k < gridz - 1; ++k)
        for (int j = 0; j < nj; ++j)
            for (int i = 0; i < ni; ++i) {
                if (!(inx(i) && iny(j) && inz(k))) continue;
                const double v = (k < nkm1) ? t[((size_t)k * nj + j) * ni + i] : 0.0;
                f << setprecision(6) << scientific << setw(15) << g.x[i] << setw(15) << g.y[j] << setw(15) << g.z[k]
                  << setw(15) << v << "\n";
                sci = true;
            }
    f << "\n";
    f.close();
    if (!f) return set_error(AMCFD_E_IO, "write failed: " + path);
    return AMCFD_OK;
}

}  // namespace amcfd
