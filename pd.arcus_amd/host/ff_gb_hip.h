// ff_gb_hip.h -- C++ host mirror of PD's forcefield-component interface for
// the GPU path. FF_GeneralizedBorn_Still_HIP has the public knobs and the
// virtual methods of Physics::FF_GeneralizedBorn_Still
// (forcefields/gbff.h:120-200 over forcefields/forcefield.h:93-222) and
// calls the C ABI of include/pdgpu.h. WorkSpaceView carries the parts of
// PD's WorkSpace the component reads and writes (positions, per-atom
// parameters, covalent partners with equilibrium distances, forces and
// the Hamiltonian), in PD's units.
#pragma once
#include <string>
#include <utility>
#include <vector>

#include "../../include/pdgpu.h"

namespace pdarcus {

struct Hamiltonian { // workspace/hamiltonian.h:23-90 (the fields the path touches)
    double epot = 0, epot_vdw = 0, epot_elec = 0, epot_pol = 0, epot_pol_self = 0, epot_pol_cross = 0;
    double epot_bond = 0, epot_angle = 0, epot_torsion = 0;
    double InternalVirial = 0;
    void zero() { *this = Hamiltonian(); }
};

struct AtomParams { // system/fundamentals.h:160-180
    double charge = 0, radius = 0, epsilon = 0, mass = 0;
    double gb_radius = 0; // Still solvation type radius of the atom's FFType
    int ir = 0;           // residue index
};

struct WorkSpaceView {
    std::vector<double> pos;   // cur.atom[].p, A, AoS
    std::vector<double> force; // cur.atom[].f, N, AoS
    std::vector<AtomParams> atom;
    std::vector<std::pair<int, int>> cov12, cov13, cov14; // covalent partners (i<j)
    std::vector<double> d12, d13;                         // equilibrium distances
    Hamiltonian ene;
    size_t nAtoms() const { return atom.size(); }
    void zeroForces() { // workspace/workspace.cpp:283-292
        ene.zero();
        force.assign(3 * atom.size(), 0.0);
    }
};

class ForcefieldBase { // forcefields/forcefield.h:93-222
public:
    explicit ForcefieldBase(WorkSpaceView &ws) : wspace(ws) {}
    virtual ~ForcefieldBase() {}
    virtual void setup() = 0;
    virtual void calcForces() = 0;
    virtual void calcEnergies() { calcForces(); }
    virtual void info() const {}
    virtual void infoLine() const {}
    virtual void infoLineHeader() const {}
    virtual unsigned long calcCheckSum() { return 0; }
    int ensuresetup() {
        if (needsetup) { setup(); needsetup = false; }
        return 0;
    }
    bool Active = true;
    int OutputLevel = 1;
    std::string name, ShortName;
    double epot = 0;

protected:
    WorkSpaceView &wspace;
    bool needsetup = true;
};

class FF_GeneralizedBorn_Still_HIP : public ForcefieldBase {
public:
    explicit FF_GeneralizedBorn_Still_HIP(WorkSpaceView &ws, int precision = PDGPU_FP64);
    ~FF_GeneralizedBorn_Still_HIP() override;

    // FF_NonBonded knobs (forcefields/nonbonded.h:55-126)
    double Cutoff, InnerCutoff, VdwCutoff, VdwInnerCutoff, Vdw14Scaling, Elec14Scaling;
    bool AutoScaling14; // true: take the 1-4 scalings given to setScaling14()
    bool EnergySwitch, ForceSwitch, DoVdw, DoElec;
    // FF_GeneralizedBorn_Still knobs (forcefields/gbff.h:127-133)
    bool FastMode;
    double DielectricSolvent, DielectricSolute, GbsaStillCutoff, ExpApproxThreshold, DielectricOffset,
        BornRadiusOffset;
    double P1, P2, P3, P4, P5;
    double Padding; // NeighbourList padding
    int Precision;  // PDGPU_FP64 or PDGPU_MIXED

    double epot_pol = 0, epot_pol_self = 0, epot_pol_cross = 0;

    void setScaling14(double vdw, double elec) { ffVdw14 = vdw; ffElec14 = elec; }
    void setup() override;
    void calcForces() override;
    void calcEnergies() override;
    void calcNewList(); // NeighbourList::calcNewList for this component's atoms
    void info() const override;
    void infoLine() const override;
    void infoLineHeader() const override;
    unsigned long calcCheckSum() override;
    const std::vector<double> &bornRadii();
    const char *lastError() const;

private:
    void fillOptions(pdgpu_options &o) const;
    void accumulate(const pdgpu_energies &e, const std::vector<double> &f, bool force_path);
    pdgpu_ctx *ctx = nullptr;
    int id = -1;
    double ffVdw14 = 1.0, ffElec14 = 1.0;
    std::vector<double> born;
};

// FF_Bonded's term lists as its setup assembles them from the parameter set
// (forcefields/ffbonded.cpp:36-486), in PD's units: see pdgpu_bonded.
struct BondedTerms {
    std::vector<int32_t> bond_ij, angle_iaj, torsion_iabj, torsion_terms, improper_iabj, improper_terms;
    std::vector<double> bond_lk, angle_tk, torsion_vn, torsion_n, torsion_gamma, improper_vn, improper_n,
        improper_gamma;
};

// Physics::FF_Bonded (forcefields/ffbonded.h:103-214) on the device.
class FF_Bonded_HIP : public ForcefieldBase {
public:
    FF_Bonded_HIP(WorkSpaceView &ws, BondedTerms terms, int precision = PDGPU_FP64);
    ~FF_Bonded_HIP() override;

    bool DoBonds = true, DoAngles = true, DoTorsions = true, DoImpropers = true;
    int Precision;

    void setup() override;
    void calcForces() override;
    void calcEnergies() override { calcForces(); } // ffbonded.cpp:804-807
    void info() const override;
    void infoLine() const override;
    void infoLineHeader() const override;
    double getEBond() const { return epot_bond; }
    double getEAngle() const { return epot_angle; }
    double getETorsion() const { return epot_torsion; }
    double getEImproper() const { return epot_improper; }

private:
    BondedTerms terms;
    pdgpu_ctx *ctx = nullptr;
    int id = -1;
    double epot_bond = 0, epot_angle = 0, epot_torsion = 0, epot_improper = 0;
};

} // namespace pdarcus
