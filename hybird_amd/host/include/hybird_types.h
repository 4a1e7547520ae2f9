// Value types, lists and parameter macros of the C++ host layer above the C ABI
// of libhybird_amd.so. Names follow the reference's headers (vector.h, elmt.h,
// macros.h) so that code written against the reference's LB / DEM / IO
// classes compiles unchanged against LB.h, DEM.h and IO.h here.
#ifndef HYBIRD_AMD_TYPES_H
#define HYBIRD_AMD_TYPES_H

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "getpot.h"

using namespace std;

// ---- parameter macros (behaviour of macros.h:5-32)
#define INFO __FILE__ << ":" << __LINE__ << " "
#define ASSERT(cond)                                                                   \
    if (!(cond)) {                                                                     \
        cout << INFO << "ERROR: Condition " #cond " not fulfilled!" << std::endl;      \
        exit(1);                                                                       \
    }
#define REQUEST(cond) \
    if (!(cond)) { cout << INFO << "WARNING: Condition " #cond " not fulfilled!" << std::endl; }
// Reads a parameter from the configuration file; "-name value" on the command
// line overrides it; a missing parameter keeps the default with a warning.
// The command-line reader must be in scope as `command_line`.
#define BASE_PARSE_CFG(cfg, type, var, name, def, def_typed)                                      \
    if (!cfg.have_variable(name) && !command_line.search("-" name))                               \
        cout << INFO << "WARNING: Parameter " name " is missing! Using default: " << def << std::endl; \
    type var = cfg(name, def_typed);                                                              \
    if (command_line.search("-" name)) var = command_line.next(var);
#define MAKE_NAME_PARSE_CFG(cfg, type, var, name, def) BASE_PARSE_CFG(cfg, type, var, name, def, static_cast<type>(def))
#define MAKE_PARSE_CFG(cfg, type, var, def) MAKE_NAME_PARSE_CFG(cfg, type, var, #var, def)
#define PARSE_CLASS_MEMBER(cfg, var, name, def) BASE_PARSE_CFG(cfg, {}, var, name, def, def)

// Problem identifiers, in the reference's order (vector.h:21): the numeric
// values are printed by the driver and select the hard-coded set-ups.
enum ProblemName { NONE, SHEARCELL, SPLASH, AVALANCHE, DRUM, BOX, NET, DIFF, BARRIER, demChute };
extern ProblemName problemName;

enum ContactModel { LINEAR, HERTZIAN };

typedef std::vector<int> intList;
typedef std::vector<unsigned int> unsIntList;
typedef std::vector<double> doubleList;

// ---- 3-vector and quaternion (component-wise arithmetic)
class tVect {
public:
    double x, y, z;
    tVect() : x(0.0), y(0.0), z(0.0) {}
    tVect(double a, double b, double c) : x(a), y(b), z(c) {}
    tVect operator+(const tVect& o) const { return tVect(x + o.x, y + o.y, z + o.z); }
    tVect operator-(const tVect& o) const { return tVect(x - o.x, y - o.y, z - o.z); }
    tVect operator*(double s) const { return tVect(x * s, y * s, z * s); }
    tVect operator/(double s) const { return tVect(x / s, y / s, z / s); }
    tVect& operator+=(const tVect& o) { x += o.x; y += o.y; z += o.z; return *this; }
    tVect& operator-=(const tVect& o) { x -= o.x; y -= o.y; z -= o.z; return *this; }
    tVect& operator*=(double s) { x *= s; y *= s; z *= s; return *this; }
    tVect& operator/=(double s) { x /= s; y /= s; z /= s; return *this; }
    double dot(const tVect& o) const { return x * o.x + y * o.y + z * o.z; }
    double norm2() const { return x * x + y * y + z * z; }
    double norm() const { return std::sqrt(norm2()); }
    void reset() { x = y = z = 0.0; }
    void show() const { cout << "(" << x << ", " << y << ", " << z << ")"; }
};
inline tVect operator*(double s, const tVect& v) { return v * s; }

class tQuat {
public:
    double q0, q1, q2, q3;
    tQuat() : q0(1.0), q1(0.0), q2(0.0), q3(0.0) {}
    tQuat(double a, double b, double c, double d) : q0(a), q1(b), q2(c), q3(d) {}
};

// Unit system (node.cpp:476-488): base units and the composites derived from them.
// Particle energies (utils.h:165-190); DEM::particleEnergy is refreshed from the
// device by DEM::syncHost.
class energy {
public:
    double kin = 0.0, rotKin = 0.0, trKin = 0.0, grav = 0.0, elastic = 0.0, total = 0.0;
    void reset() { kin = rotKin = trKin = grav = elastic = total = 0.0; }
    void updateTotal() {
        kin = rotKin + trKin;
        total = rotKin + trKin + grav + elastic;
    }
};

class measureUnits {
public:
    double Length = 1.0, Time = 1.0, Density = 1.0;
    double Volume = 1.0, Speed = 1.0, Accel = 1.0, AngVel = 1.0, KinVisc = 1.0, DynVisc = 1.0, Force = 1.0,
           Torque = 1.0, Mass = 1.0, Stress = 1.0, Pressure = 1.0;
    void setComposite() {
        Volume = Length * Length * Length;
        Speed = Length / Time;
        Accel = Length / Time / Time;
        AngVel = 1.0 / Time;
        KinVisc = Length * Length / Time;
        DynVisc = Density * Length * Length / Time;
        Force = Density * Length * Length * Length * Length / Time / Time;
        Torque = Density * Length * Length * Length * Length * Length / Time / Time;
        Mass = Density * Length * Length * Length;
        Stress = Density * Length * Length / Time / Time;
        Pressure = Density * Length * Length / Time / Time;
    }
};

// ---- DEM objects. The device holds the state; these host records are filled
// from it on request (DEM::syncHost).
class elmt {
public:
    unsigned int index = 0, size = 1;
    double radius = 0.0, m = 0.0;
    tVect x0, x1, w0, FHydro, MHydro, FParticle, MParticle, FWall, MWall, FGrav;
    tQuat q0, q1;
    double fluidVolume = 0.0;
    intList components;  // its standard particles (elmt.h)
};
class particle {
public:
    unsigned int particleIndex = 0, clusterIndex = 0, protoIndex = 0;
    double r = 0.0;
    tVect x0, x1, radiusVec;
    bool isGhost = false;
};
class wall {
public:
    unsigned int index = 0;
    tVect n, p;
    bool moving = false, slip = false;
    tVect FHydro, FParticle;
};
class cylinder {
public:
    unsigned int index = 0;
};
class object {
public:
    unsigned int index = 0, originalIndex = 0;
    double r = 0.0;
    tVect x0, x1;
};
class pbc {
public:
    tVect v, p1, p2;
};

class DEM;
// Lists of DEM objects carry the DEM that owns them, so that the lattice calls
// that receive them (LB::latticeBolzmannInit/Step/CouplingStep) reach its device state.
template <class T>
class ownedList : public std::vector<T> {
public:
    DEM* owner = nullptr;
};
typedef ownedList<elmt> elmtList;
typedef ownedList<particle> particleList;
typedef ownedList<wall> wallList;
typedef ownedList<cylinder> cylinderList;
typedef ownedList<object> objectList;
typedef ownedList<pbc> pbcList;

// Calls into libhybird_amd.so report failures like the reference's fatal
// checks: print the location and the library's message, then exit(1).
void hybird_check(int rc, const char* what, const char* file, int line);
#define HYB_CALL(expr) hybird_check((expr), #expr, __FILE__, __LINE__)

#endif  // HYBIRD_AMD_TYPES_H
