// Test-only driver around the reference's solver classes and its output layer
// (PrintManager, vtsBinWriter, FreeSurfaceManager), compiled from the
// reference sources in place by oracle/Makefile; nothing here ships.
//
// It follows the reference executable's sequence (src/main.cpp:53-241): set
// up through PrintManager::initialization, write the initial VTS file, then
// per step dt / timet / updateLaser / generateGrid (moving mesh) /
// iterativeLoopSolve / getSolidificationParameters / outputSection /
// updateSolutions and, every output interval, the free surface (execute),
// outputResiduals + visualOut; at the end the free surface, a last section,
// outputResiduals + visualOut and outputFinalGeometry.  The files it leaves in
// the working directory (<dir>/<id>_<n>.vts, <id>.pvd, <id>_Residuals.txt,
// the final-geometry file, <id>_FreeSurfaceMonitor.txt, the extracted
// sections) are the reference's own output for the deck, for comparison with
// the MI355X CLI.
//
// usage: ref_outputs <deck.k>      (run with OMP_NUM_THREADS=1)

#include <chrono>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <string>

#define private public
#include <Mesh.h>
#include <DomainManager.h>
#include <SolutionVariableManager.h>
#include <BoundCondManager.h>
#include <CFDSolverManager.h>
#include <FreeSurfaceManager.h>
#include <PrintManager.h>
#include <vtsBinWriter.h>
#undef private

int main(int argc, char **argv) {
    if (argc != 2) {
        std::fprintf(stderr, "usage: %s deck.k\n", argv[0]);
        return 1;
    }
    std::string inputFile = argv[1];
    Mesh meshObj(inputFile);
    meshObj.getDomainInfo();
    meshObj.assignParameters();
    meshObj.outputInputs(argv[1]);

    DomainManager *domainMgr = new DomainManager(&meshObj);
    SolutionVariableManager *solVarMgr = new SolutionVariableManager();
    BoundCondManager *bcMgr = new BoundCondManager(&meshObj, domainMgr, solVarMgr);
    CFDSolverManager *cfdSolvMgr = new CFDSolverManager(&meshObj, domainMgr, solVarMgr, bcMgr);
    FreeSurfaceManager *freesurMgr = new FreeSurfaceManager(&meshObj, domainMgr, solVarMgr, bcMgr);

    int outCt = 0;
    std::string directoryName = meshObj.directoryName_;
    std::string folderName = directoryName + "/";
    std::string extName = ".vts";
    std::string outFile = folderName + meshObj.outFileName_ + "_" + std::to_string(outCt) + extName;
    PrintManager *printMgr = new vtsBinWriter(&meshObj, domainMgr, solVarMgr, bcMgr, cfdSolvMgr, freesurMgr,
                                              outFile, directoryName, meshObj.outFileName_);
    printMgr->initialization();
    if (domainMgr->steady_) bcMgr->updateLaser();

    int variableTimeStepState = 0;
    double outtrack = 0.0;
    domainMgr->timet_ = 0.0;
    double simtime = meshObj.finaltime_;
    printMgr->scanOutputData();
    printMgr->visualOut();

    while (domainMgr->timet_ < simtime) {
        if (meshObj.variableTimeStepFileName_.compare("FALSE")) {
            if ((domainMgr->timet_ > meshObj.variableTimeStepValues[variableTimeStepState][0]) &&
                (variableTimeStepState < (int)meshObj.variableTimeStepValues.size() - 1))
                variableTimeStepState += 1;
            meshObj.delt_ = meshObj.variableTimeStepValues[variableTimeStepState][1];
            meshObj.outtime_ = meshObj.variableTimeStepValues[variableTimeStepState][2];
        }
        domainMgr->timet_ += meshObj.delt_;
        outtrack += meshObj.delt_;
        if (!domainMgr->steady_) bcMgr->updateLaser();
        if (meshObj.movingmesh_) cfdSolvMgr->generateGrid();
        cfdSolvMgr->iterativeLoopSolve();
        if (meshObj.solidificationparams_) cfdSolvMgr->getSolidificationParameters();
        if (meshObj.outputsection_) printMgr->outputSection();
        if (!domainMgr->steady_) {
            cfdSolvMgr->updateSolutions();
            if (outtrack >= meshObj.outtime_) {
                outCt++;
                outFile = folderName + meshObj.outFileName_ + "_" + std::to_string(outCt) + extName;
                if (meshObj.energyfreesurface_) freesurMgr->execute();
                printMgr->outputResiduals();
                printMgr->visualOut();
                outtrack = 0.0;
            }
        }
    }
    domainMgr->finish_ = true;
    if (meshObj.energyfreesurface_) freesurMgr->execute();
    if (meshObj.outputsection_) printMgr->outputSection();
    outCt++;
    outFile = folderName + meshObj.outFileName_ + "_" + std::to_string(outCt) + extName;
    printMgr->outputResiduals();
    printMgr->visualOut();
    if (meshObj.outputgeom_) printMgr->outputFinalGeometry();
    std::printf("{\"outputs\": %d}\n", outCt + 1);
    return 0;
}
