// Input/output side of the C++ host layer: the public part of the reference's
// IO class (IO.h:36-131). Screen/export-file status lines follow IO.cpp:132-198;
// the VTK writers of the reference are not part of this library.
#ifndef HYBIRD_AMD_IO_H
#define HYBIRD_AMD_IO_H

#include <ctime>
#include <fstream>
#include <string>

#include <sys/stat.h>
#include <sys/types.h>

#include "DEM.h"
#include "LB.h"
#include "hybird_types.h"

class IO {
public:
    string problemNameString;
    struct tm* now = nullptr;
    string lbmCfgName;
    unsigned int currentTimeStep = 0, maximumTimeSteps = 0;
    double maxTime = 0.0;
    double screenExpTime = 0.0, stateExpTime = 0.0, fluidExpTime = 0.0, partExpTime = 0.0, objectExpTime = 0.0,
           recycleExpTime = 0.0, outputExpTime = 0.0;
    unsigned int lastScreenExp = 0, lastFluidExp = 0, lastPartExp = 0, lastObjectExp = 0, lastOutputExp = 0,
                 lastRecycleExp = 0;
    bool problem = false;
    bool demSolve = false, lbmSolve = false;
    double realTime = 0.0;
    string workDirectory, partDirectory, fluidDirectory;
    string exportFileName;
    ofstream exportFile;
    int saveCount = 0;

    void initialize();
    void outputStep(const double& demTime, const LB& lb, const DEM& dem);
    void outputFinal();
};

#endif  // HYBIRD_AMD_IO_H
