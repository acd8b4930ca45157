// Test-only driver around the reference solver classes (compiled from the
// reference sources in place by oracle/Makefile; nothing here ships).  The
// energy-based free surface runs at output times as in main.cpp, so its
// monitor file (<id>_FreeSurfaceMonitor.txt) lands in the working directory.
//
// It reproduces the call order of the reference executable
//   src/main.cpp:53-74      construct Mesh / managers
//   src/sources/PrintManager.cpp:49-74   init order (assignInputValues,
//                                        generateGrid, initializeSolver, ...)
//   src/main.cpp:123-162    per-step order (dt, timet, updateLaser,
//                           generateGrid when the mesh moves, solve,
//                           updateSolutions)
// without any of the output writers, and dumps per-step scalars and (on
// request) full fields as raw little-endian float64 so tests can compare the
// C restatement and the GPU path against the reference program order.
//
// usage: ref_harness <deck.k> <outdir> [fields_every=0] [max_steps=-1]
//        [dump_first_iter_step=-1]
//
// Run with OMP_NUM_THREADS=1: only then is the reference deterministic
// (its j-parallel line sweeps race otherwise).

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <map>
#include <fstream>
#include <iostream>
#include <sstream>
#include <algorithm>
#include <cmath>
#include <omp.h>

// expose private members of the reference classes to this harness only
#define private public
#include <Mesh.h>
#include <DomainManager.h>
#include <SolutionVariableManager.h>
#include <BoundCondManager.h>
#include <CFDSolverManager.h>
#include <FreeSurfaceManager.h>
#undef private

static void write_raw(const std::string &path, const double *p, size_t n) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) { std::perror(path.c_str()); std::exit(2); }
    std::fwrite(p, sizeof(double), n, f);
    std::fclose(f);
}

int main(int argc, char **argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s deck.k outdir [fields_every] [max_steps]\n", argv[0]);
        return 1;
    }
    std::string deck = argv[1];
    std::string outdir = argv[2];
    int fields_every = argc > 3 ? std::atoi(argv[3]) : 0;
    int max_steps = argc > 4 ? std::atoi(argv[4]) : -1;

    Mesh meshObj(deck);
    meshObj.getDomainInfo();
    meshObj.assignParameters();
    if (!meshObj.foundMesh_) { std::fprintf(stderr, "no *MESH_FILE\n"); return 3; }
    meshObj.readMeshFile();

    DomainManager *dm = new DomainManager(&meshObj);
    SolutionVariableManager *sv = new SolutionVariableManager();
    BoundCondManager *bc = new BoundCondManager(&meshObj, dm, sv);
    CFDSolverManager *cfd = new CFDSolverManager(&meshObj, dm, sv, bc);
    FreeSurfaceManager *fs = new FreeSurfaceManager(&meshObj, dm, sv, bc);

    // PrintManager::initialization order
    dm->assignInputValues();
    if (meshObj.udtoolpath_) dm->getToolpath();
    cfd->generateGrid();
    cfd->init_ = false;
    cfd->initializeSolver();
    bc->initializeBoundaries();
    fs->initializeFreeSurface();
    if (meshObj.powderbed_) {
        dm->nk_ -= (dm->numlayers_ - 1) * dm->ncvzlayer_;
        dm->nkm1_ = dm->nk_ - 1;
        dm->nkm2_ = dm->nk_ - 2;
    }
    if (dm->steady_) bc->updateLaser();

    const int ni = dm->ni_, nj = dm->nj_;
    // nk may grow in powder mode; dump the full allocated extent
    int nk_full = dm->nk_;
    if (meshObj.powderbed_) nk_full += (dm->numlayers_ - 1) * dm->ncvzlayer_;
    const size_t nn = (size_t)ni * nj * nk_full;

    {
        std::ofstream g(outdir + "/grid.txt");
        g.precision(17);
        g << ni << " " << nj << " " << nk_full << "\n";
        g << "x"; for (int i = 0; i < ni; ++i) g << " " << cfd->x_[i]; g << "\n";
        g << "xu"; for (int i = 0; i < ni; ++i) g << " " << cfd->xu_[i]; g << "\n";
        g << "y"; for (int j = 0; j < nj; ++j) g << " " << cfd->y_[j]; g << "\n";
        g << "yv"; for (int j = 0; j < nj; ++j) g << " " << cfd->yv_[j]; g << "\n";
        g << "z"; for (int k = 0; k < nk_full; ++k) g << " " << cfd->z_[k]; g << "\n";
        g << "zw"; for (int k = 0; k < nk_full; ++k) g << " " << cfd->zw_[k]; g << "\n";
        g << "hlcal " << dm->hlcal_ << "\nhlatnt " << dm->hlatnt_ << "\n";
        g << "steady " << dm->steady_ << "\n";
    }

    FILE *sc = std::fopen((outdir + "/scalars.csv").c_str(), "w");
    std::fprintf(sc, "step,timet,niter,itertot,resorh,resorm,resoru,resorv,resorw,resorc,tpeak,alen,depth,width,"
                     "ratio,heatinlaser,heatout,accul,heatvol,ahtoploss,beamposx,beamposy,powerindicator,"
                     "istat,iend,jstat,jend,kstat,jstart,istart,nk\n");

    const char *names[] = {"enthalpy", "temp", "fracl", "uvel", "vvel", "wvel", "pressure", "pp",
                           "rho", "diff", "vis", "hnot", "fraclnot", "unot", "vnot", "wnot",
                           "sourceinput", "dux", "dvy", "dwz"};
    auto dump_fields = [&](int step) {
        const double *ptr[] = {cfd->enthalpy_, cfd->temp_, cfd->fracl_, cfd->uvel_, cfd->vvel_,
                               cfd->wvel_, cfd->pressure_, cfd->pp_, cfd->rho_, cfd->diff_, cfd->vis_,
                               cfd->hnot_, cfd->fraclnot_, cfd->unot_, cfd->vnot_, cfd->wnot_,
                               cfd->sourceinput_, cfd->dux_, cfd->dvy_, cfd->dwz_};
        for (int f = 0; f < 20; ++f) {
            char buf[512];
            std::snprintf(buf, sizeof buf, "%s/f%05d_%s.bin", outdir.c_str(), step, names[f]);
            write_raw(buf, ptr[f], nn);
        }
        // species and G/R fields, present only with their keywords
        const char *xnames[] = {"avgconcentration", "avgconcentrationnot", "concentration", "grad", "rate",
                                "zr", "tempr"};
        const double *xptr[] = {cfd->avgconcentration_, cfd->avgconcentrationnot_, cfd->concentration_,
                                cfd->grad_, cfd->rate_, fs->zr_, fs->tempr_};
        for (int f = 0; f < 7; ++f) {
            if (f < 3 ? !meshObj.species_ : f < 5 ? !meshObj.solidificationparams_ : !meshObj.energyfreesurface_)
                continue;
            char buf[512];
            std::snprintf(buf, sizeof buf, "%s/f%05d_%s.bin", outdir.c_str(), step, xnames[f]);
            write_raw(buf, xptr[f], nn);
        }
        // powder history fields, written on powder-bed decks
        const char *pnames[] = {"csfrac", "solfrac", "mpPhase", "mpLiqPhase"};
        const double *pptr[] = {cfd->csfrac_, cfd->solfrac_, cfd->mpPhase, cfd->mpLiqPhase};
        for (int f = 0; meshObj.powderbed_ && f < 4; ++f) {
            char buf[512];
            std::snprintf(buf, sizeof buf, "%s/f%05d_%s.bin", outdir.c_str(), step, pnames[f]);
            write_raw(buf, pptr[f], nn);
        }
    };
    if (fields_every > 0) dump_fields(0);

    int variableTimeStepState = 0;
    double outtrack = 0.0;
    dm->timet_ = 0.0;
    double simtime = meshObj.finaltime_;
    int step = 0;
    double t0 = omp_get_wtime();
    while (dm->timet_ < simtime) {
        if (max_steps >= 0 && step >= max_steps) break;
        if (meshObj.variableTimeStepFileName_.compare("FALSE")) {
            if ((dm->timet_ > meshObj.variableTimeStepValues[variableTimeStepState][0]) &&
                (variableTimeStepState < (int)meshObj.variableTimeStepValues.size() - 1))
                variableTimeStepState += 1;
            meshObj.delt_ = meshObj.variableTimeStepValues[variableTimeStepState][1];
            meshObj.outtime_ = meshObj.variableTimeStepValues[variableTimeStepState][2];
        }
        dm->timet_ += meshObj.delt_;
        outtrack += meshObj.delt_;
        if (!dm->steady_) bc->updateLaser();
        if (meshObj.movingmesh_) cfd->generateGrid();   // main.cpp:144-145
        cfd->iterativeLoopSolve();
        if (meshObj.solidificationparams_) cfd->getSolidificationParameters();   // main.cpp:151-152
        if (!dm->steady_) {
            cfd->updateSolutions();
            // the free surface runs at output times (main.cpp:167-176)
            if (outtrack >= meshObj.outtime_) {
                if (meshObj.energyfreesurface_) fs->execute();
                outtrack = 0.0;
            }
        }
        ++step;
        std::fprintf(sc,
                     "%d,%.17g,%d,%d,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,"
                     "%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%d,%d,%d,%d,%d,%d,%d,%d,%d\n",
                     step, dm->timet_, cfd->niter_, cfd->itertot_, cfd->resorh_, cfd->resorm_,
                     cfd->resoru_, cfd->resorv_, cfd->resorw_, cfd->resorc_, bc->tpeak_, bc->alen_, bc->depth_,
                     bc->width_, bc->ratio_, bc->heatinlaser_, bc->heatout_, bc->accul_, bc->heatvol_,
                     bc->ahtoploss_, bc->beamposx_, bc->beamposy_, bc->powerindicator_, bc->istat_,
                     bc->iend_, bc->jstat_, bc->jend_, bc->kstat_, bc->jstart_, bc->istart_, dm->nk_);
        std::fflush(sc);
        if (meshObj.movingmesh_) {   // the regenerated x/y node coordinates of this step
            FILE *gs = std::fopen((outdir + "/grid_steps.txt").c_str(), step == 1 ? "w" : "a");
            std::fprintf(gs, "%d x", step);
            for (int i = 0; i < ni; ++i) std::fprintf(gs, " %.17g", cfd->x_[i]);
            std::fprintf(gs, " y");
            for (int j = 0; j < nj; ++j) std::fprintf(gs, " %.17g", cfd->y_[j]);
            std::fprintf(gs, "\n");
            std::fclose(gs);
        }
        if (fields_every > 0 && step % fields_every == 0) dump_fields(step);
    }
    double t1 = omp_get_wtime();
    std::fclose(sc);
    long long dom = (long long)(ni - 2) * (nj - 2) * (dm->nk_ - 2);
    std::printf("{\"steps\": %d, \"compute_s\": %.6f, \"interior_cells\": %lld, \"threads\": %d, "
                "\"cell_steps_per_s\": %.6g, \"itertot\": %d}\n",
                step, t1 - t0, dom, omp_get_max_threads(), (double)dom * step / (t1 - t0), cfd->itertot_);
    return 0;
}
