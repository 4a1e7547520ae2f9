// Lattice-Boltzmann side of the C++ host layer: the reference's LB class
// interface (LB.h:32-213) over the C ABI of libhybird_amd.so. The lattice lives
// on the GPU; the host views below copy it on request.
#ifndef HYBIRD_AMD_LB_H
#define HYBIRD_AMD_LB_H

#include <string>
#include <vector>

#include "hybird_amd.h"
#include "hybird_types.h"

class LB {
public:
    // ---- parameters (lattice units after latticeBoltzmannGet)
    double initDynVisc = 1.0;
    double yieldStress = 1.0, plasticVisc = 1.0;
    double turbConst = 0.0;
    double slipCoefficient = 0.0;
    double initDensity = 1.0;
    tVect initVelocity;
    tVect lbF;
    unsigned int totNodes = 0;
    double totalMass = 0.0;
    unsIntList boundary;
    bool freeSurface = false;
    bool forceField = false;
    bool nonNewtonian = false;
    bool turbulenceOn = false;
    unsigned int lbmInitialRepeat = 0;
    unsigned int time = 0;
    unsIntList lbSize;
    measureUnits unit;
    // problem-specific inputs read by the driver (hybird.cpp:199-219)
    double maxVisc = 0.0, maxPlasticVisc = 0.0, maxYieldStress = 0.0, maxShearRate = 0.0;
    unsigned int viscSteps = 0, shearRateSteps = 0;
    double fluidMass = 0.0;
    double avalanchePosit = 0.0;

    LB();
    ~LB();
    LB(const LB&) = delete;
    LB& operator=(const LB&) = delete;

    // ---- the reference's entry points (LB.h:155-161)
    void LBShow() const;
    void latticeDefinition();
    void latticeBoltzmannGet(GetPot& lbmCfgFile, GetPot& command_line);
    void latticeBolzmannInit(cylinderList& cylinders, wallList& walls, particleList& particles, objectList& objects);
    void latticeBolzmannStep(elmtList& elmts, particleList& particles, wallList& walls);
    void latticeBoltzmannCouplingStep(bool& newNeighborList, elmtList& elmts, particleList& particles);
    void latticeBoltzmannFreeSurfaceStep();

    // ---- node coordinates (LB.cpp:2314-2369)
    unsigned int getX(const unsigned int& index) const { return index % lbSize[0]; }
    unsigned int getY(const unsigned int& index) const { return (index / lbSize[0]) % lbSize[1]; }
    unsigned int getZ(const unsigned int& index) const { return index / (lbSize[0] * lbSize[1]); }
    tVect setPosition(const unsigned int& index) const { return tVect(getX(index), getY(index), getZ(index)); }

    // ---- host views of the device state (names and layouts of hyb_lb_get)
    std::vector<double> field(const char* name) const;
    std::vector<unsigned char> nodeTypes() const;
    std::vector<unsigned char> nodeFlags(const char* name) const;  // u8 views ("pflag")
    std::vector<unsigned int> solidIndices() const;
    double totalFluidMass() const;
    // max |u|^2 over active nodes, active nodes, active nodes near the maximum viscosity
    std::vector<double> statusValues() const;
    hyb_lb* handle() const { return h_; }
    const hyb_config& config() const { return cfg_; }

private:
    hyb_config cfg_;
    hyb_lb* h_ = nullptr;
    bool inited_ = false;
    hyb_dem* dem_ = nullptr;     // DEM handle coupled at init (for the state file)
    std::string restartFile_;    // extension: state file restored after init
    std::string saveStateFile_;  // extension: state file written when the run ends
    bool fluidExport_ = false;   // fluid files requested (fluidExpTime > 0)
};

#endif  // HYBIRD_AMD_LB_H
