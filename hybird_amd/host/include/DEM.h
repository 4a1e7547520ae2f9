// Discrete-element side of the C++ host layer: the reference's DEM class
// interface (DEM.h:30-176) over the C ABI of libhybird_amd.so.
#ifndef HYBIRD_AMD_DEM_H
#define HYBIRD_AMD_DEM_H

#include <string>
#include <vector>

#include "hybird_amd.h"
#include "hybird_types.h"

class IO;

class DEM {
public:
    energy particleEnergy;
    double demInitialRepeat = 0.0;
    double deltat = 1.0;
    double demTime = 0.0;
    unsigned int demTimeStep = 0;
    tVect demF;
    unsigned int multiStep = 1;
    double criticalRatio = 0.1;
    wallList walls;
    cylinderList cylinders;
    pbcList pbcs;
    elmtList elmts;
    particleList particles;
    objectList objects;
    unsigned int stdParticles = 0;
    bool newNeighborList = false;
    double drumSpeed = 0.0;
    unsigned int stdObjects = 0;

    DEM();
    ~DEM();
    DEM(const DEM&) = delete;
    DEM& operator=(const DEM&) = delete;

    // ---- the reference's entry points (DEM.h:137-139)
    void discreteElementStep(IO& io);
    void discreteElementGet(GetPot& lbmCfgFile, GetPot& command_line);
    void discreteElementInit(unsIntList& boundary, unsIntList& lbSize, measureUnits& unit, tVect& lbF);

    // Copies element state (positions, velocities, forces) and particleEnergy from the device.
    void syncHost();
    std::vector<double> field(const char* name) const;
    std::vector<unsigned int> indexField(const char* name) const;  // u32 views ("p_idx", ...)
    hyb_dem* handle() const { return h_; }

private:
    // MercuryDPM-format files of the saving sub-steps of the last device step
    void exportSaved(IO& io);
    int saveCountSet_ = 0;
    hyb_config cfg_;
    std::vector<double> records_;
    std::vector<double> objectRecords_;  // HYB_OBJECT_RECORD per object
    hyb_dem* h_ = nullptr;
};

#endif  // HYBIRD_AMD_DEM_H
