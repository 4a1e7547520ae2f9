// Input/output side of the C++ host layer: the public part of the reference's
// IO class (IO.h:36-131). Status lines follow IO.cpp:132-198; the fluid (.vti),
// particle (.vtu) and recycle (.dat) files follow IO.cpp:250-535, 698-831, and the
// MercuryDPM files (.data, .restart, .fstat) DEM.cpp:2280-2565, with
// the reference's names and number formatting, from views of the device state.
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
    string fluidFileFormat, partFileFormat, objectFileFormat;
    ofstream exportFile;
    int saveCount = 0;
    // MercuryDPM-format files (IO.cpp:41-48), written by DEM::discreteElementStep
    string dataFileName, statisticFileName, restartFileName;
    // status-line side files (IO.cpp:23-39, 105-123)
    string energyFileName, maxSpeedFileName, forceFileName, plasticityFileName;
    // precision the .restart stream starts a file with: the reference's stream
    // keeps the setprecision(10) of its first file for the later ones
    int restartPrecision = 6;

    void initialize();
    void outputStep(const double& demTime, const LB& lb, const DEM& dem);
    void outputFinal();
    // writers (IO.cpp:289-535, 698-831)
    void exportParaviewFluid(const LB& lb, const DEM& dem, const string& fluidFile) const;
    void exportParaviewParticles(const DEM& dem, const string& particleFile) const;
    void exportRecycleParticles(const DEM& dem) const;
    void exportParaviewObjects(const DEM& dem, const string& objectFile) const;
};

#endif  // HYBIRD_AMD_IO_H
