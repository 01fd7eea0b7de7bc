// ref_driver.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A thin C interface over the reference's own SpringNetwork, compiled from the
// reference sources where they lie (see oracle/Makefile, target `ref`) into
// oracle/_ref/libbsref.so.  It is used to produce golden vectors for the
// oracle and, optionally, as the CPU baseline.  Nothing here re-implements
// physics: every number comes from the reference code.
//
//  - ref_create: SpringNetwork::addParticle / addSpring / setup
//    (src/spn/SpringNetwork.cpp:452-527,646-661).
//  - ref_compute_forces: idleRun + computeForces (SpringNetwork.cpp:244-249).
//  - ref_step: computeStep (SpringNetwork.cpp:251-268).
//  - ref_cells: the reference nsearch::NeighborSearch cell list
//    (src/nsearch.hpp:277-293,342-390), read through a subclass.
//  - ref_pairs: the processed pairs, recorded by a ForceField subclass whose
//    energy callbacks see particle ids encoded in a per-particle property;
//    the pair decisions are made by the reference's own Particle::add*Force.
//  - ref_add_rigid_body: RigidBodiesManager::InitRigidBodies
//    (src/rigidbody/RigidBodiesManager.cpp:11-43); the manager's collection is
//    global, so one reference object with bodies may exist at a time.
//  - ref_set_grid: a PotentialGrid filled cell by cell, then
//    PotentialGrid::compute_gradient (src/grid/PotentialGrid.cpp:20-60), as
//    the DX reader does after parsing (src/IO/OpenDXReader.cpp:146-173).
//  - ref_insertion: InsertionVector::computeVector/Angle/RollAngle
//    (src/InsertionVector.cpp:18-78).
//  - ref_springs_from_cutoff: topology::Topology::add_springs_from_cutoff
//    (src/topology/Topology.hpp:105-138), or with topology ids
//    add_springs_between_topologies_from_cutoff (Topology.hpp:140-180), on
//    particles at the given positions, after the given springs were added
//    with Topology::add_spring; all
//    springs read back as (first, second, float(equilibrium),
//    float(stiffness)) in the reference's own order.
//  - ref_set_trajectories: the pdbtrajectory / csvsampling settings applied
//    to the next ref_create, so that the reference's own writers
//    (io::modern::TrajectoryManager, SpringNetwork.cpp:768-778) run inside
//    its computeStep.  Its XTC writer is not driven: XTCTrajectoryWriter
//    never opens its XDR handle (XTCTrajectoryWriter.hpp:19-20, .cpp:13-31).
//  - ref_xtc_write / ref_xtc_read: frames through the reference's vendored
//    xdrfile write_xtc / read_xtc (xdrfile_xtc.c:89-119).
#include "SpringNetwork.h"
#include "configuration/Configuration.hpp"
#include "forcefield/ForceField.h"
#include "InsertionVector.h"
#include "grid/PotentialGrid.hpp"
#include "rigidbody/RigidBodiesManager.h"
#include "topology/Topology.hpp"
#include "xdrfile.h"
#include "xdrfile_xtc.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/bs_api.h"

using biospring::spn::Particle;
using biospring::spn::SpringNetwork;

namespace
{

const char * steric_mode_name(int m)
{
    switch (m)
    {
    case BS_STERIC_AMBER_12_6: return "lennard-jones-12-6Amber";
    case BS_STERIC_LEWITT_8_6: return "lennard-jones-8-6Lewitt";
    case BS_STERIC_ZACHARIAS_8_6: return "lennard-jones-8-6Zacharias";
    default: return "linear";
    }
}

struct TrajConf
{
    std::string pdb, csv;
    size_t pdbf = 100, csvf = 100;
} g_traj;

biospring::configuration::Configuration make_config(const bs_params * p)
{
    auto c = biospring::configuration::defaultConfiguration();
    if (!g_traj.pdb.empty())
    {
        c.pdbtraj.enable = true;
        c.pdbtraj.path = g_traj.pdb;
        c.pdbtraj.frequency = g_traj.pdbf;
    }
    if (!g_traj.csv.empty())
    {
        c.csvsample.enable = true;
        c.csvsample.path = g_traj.csv;
        c.csvsample.frequency = g_traj.csvf;
    }
    c.sim.nbsteps = -1;
    c.sim.timestep = p->timestep;
    c.sim.neighborskin = p->neighborskin;
    c.spring.enable = p->spring_enable != 0;
    c.spring.scale = p->spring_scale;
    c.steric.enable = p->steric_enable != 0;
    c.steric.mode = steric_mode_name(p->steric_mode);
    c.steric.cutoff = p->steric_cutoff;
    c.steric.gridscale = p->steric_gridscale;
    c.electrostatic.enable = p->electrostatic_enable != 0;
    c.electrostatic.cutoff = p->electrostatic_cutoff;
    c.electrostatic.scale = p->electrostatic_scale;
    c.electrostatic.dielectric = p->electrostatic_dielectric;
    c.viscosity.enable = p->viscosity_enable != 0;
    c.viscosity.value = p->viscosity_value;
    c.imp.enable = p->imp_enable != 0;
    c.imp.scale = p->imp_scale;
    c.hydrophobicity.enable = p->hydrophobicity_enable != 0;
    c.hydrophobicity.cutoff = p->hydrophobicity_cutoff;
    c.hydrophobicity.scale = p->hydrophobicity_scale;
    c.rigidbody.enable = p->rigidbody_enable != 0;
    c.probe.enable = p->probe_enable != 0;
    c.probe.enableelectrostatic = p->probe_enable_electrostatic != 0;
    c.probe.enablesteric = p->probe_enable_steric != 0;
    c.probe.x = p->probe_x;
    c.probe.y = p->probe_y;
    c.probe.z = p->probe_z;
    c.probe.mass = p->probe_mass;
    c.probe.epsilon = p->probe_epsilon;
    c.probe.radius = p->probe_radius;
    c.probe.charge = p->probe_charge;
    return c;
}

// Records the (j, i) ids of each processed pair; ids are encoded as
// epsilon (steric), |charge|-1 (electrostatic), |hydrophobicity|-1.
class RecordingForceField : public biospring::forcefield::ForceField
{
  public:
    mutable std::mutex mu;
    mutable std::vector<uint32_t> ij;
    int kind = BS_NS_STERIC;

    float computeStericEnergy(float, float, float e1, float e2, float) const override
    {
        if (kind == BS_NS_STERIC)
            push((uint32_t)e1, (uint32_t)e2);
        return 0.0f;
    }
    float computeStericForceModule(float, float, float, float, float) const override { return 0.0f; }
    float computeElectrostaticEnergy(float c1, float c2, float) const override
    {
        if (kind == BS_NS_ELECTROSTATIC)
            push((uint32_t)(std::fabs(c1) - 1.0f), (uint32_t)(std::fabs(c2) - 1.0f));
        return 0.0f;
    }
    float computeElectrostaticForceModule(float, float, float) const override { return 0.0f; }
    float computeHydrophobicityEnergy(float h1, float h2, float) const override
    {
        if (kind == BS_NS_HYDROPHOBIC)
            push((uint32_t)(std::fabs(h1) - 1.0f), (uint32_t)(std::fabs(h2) - 1.0f));
        return 0.0f;
    }
    float computeHydrophobicityForceModule(float, float, float) const override { return 0.0f; }
    float computeSpringEnergy(float, float, float) const override { return 0.0f; }
    float computeSpringForceModule(float, float, float) const override { return 0.0f; }
    float computeIMPEnergy(float, float, float, float, float) const override { return 0.0f; }
    Vector3f computeIMPForceVector(float, float, float, float, float) const override { return Vector3f(); }

  private:
    void push(uint32_t a, uint32_t b) const
    {
        std::lock_guard<std::mutex> lk(mu);
        ij.push_back(std::min(a, b));
        ij.push_back(std::max(a, b));
    }
};

class RefNet : public SpringNetwork
{
  public:
    void swap_forcefield(std::unique_ptr<biospring::forcefield::ForceField> ff) { _ff = std::move(ff); }
    std::vector<Particle> & particles() { return _particles; }
    // Grid options are switched on after setup(): setup would otherwise read
    // the grids from DX files (SpringNetwork.cpp:712-747).
    void enable_grid(int kind, double scale)
    {
        if (kind == BS_GRID_POTENTIAL)
        {
            _config.potentialgrid.enable = true;
            _config.potentialgrid.scale = scale;
            _ff->setForceFieldScale(scale);
        }
        else
        {
            _config.densitygrid.enable = true;
            _config.densitygrid.scale = scale;
        }
    }
};

using Container = std::vector<Particle>;

struct ExposedSearch : public biospring::nsearch::NeighborSearch<Container>
{
    using Base = biospring::nsearch::NeighborSearch<Container>;
    ExposedSearch(const Container & c, float cutoff, std::vector<size_t> inc, float skin)
        : Base(c, cutoff, std::move(inc), skin)
    {
    }
    ExposedSearch(const Container & c, float cutoff, float skin) : Base(c, cutoff, skin) {}
    int64_t cell_of(size_t i) const { return (int64_t)_compute_cell(biospring::concepts::locatable::get_position(_system->at(i))); }
    void dims(int64_t d[3]) const
    {
        d[0] = (int64_t)_ncells_x;
        d[1] = (int64_t)_ncells_y;
        d[2] = (int64_t)_ncells_z;
    }
    void box(double b[6]) const
    {
        for (int k = 0; k < 6; ++k)
            b[k] = _box.boundaries()[k];
    }
    const std::vector<size_t> * cell(size_t id) const
    {
        auto it = _cells.find(id);
        return it == _cells.end() ? nullptr : &it->second;
    }
};

struct RefSys
{
    bs_params p;
    int64_t n = 0; // particles given; the probe, when enabled, follows them
    std::unique_ptr<RefNet> net;
};

RefNet * build_net(const bs_params * p, int64_t n, const float * x3, const float * v3, const float * mass,
                   const float * charge, const float * radius, const float * eps, const float * sasa,
                   const float * transfer, const float * hydro, const uint8_t * flags, int64_t m,
                   const uint32_t * id1, const uint32_t * id2, const float * l0, const float * k)
{
    auto * net = new RefNet();
    for (int64_t i = 0; i < n; ++i)
    {
        Particle q;
        q.setPosition(Vector3f(x3[3 * i], x3[3 * i + 1], x3[3 * i + 2]));
        if (v3)
            q.setVelocity(Vector3f(v3[3 * i], v3[3 * i + 1], v3[3 * i + 2]));
        if (mass)
            q.setMass(mass[i]);
        if (charge)
            q.setCharge(charge[i]);
        if (radius)
            q.setRadius(radius[i]);
        if (eps)
            q.setEpsilon(eps[i]);
        if (sasa)
            q.setSolventAccessibilitySurface(sasa[i]);
        if (transfer)
            q.setTransferEnergyByAccessibleSurface(transfer[i]);
        if (hydro)
            q.setHydrophobicity(hydro[i]);
        q.setStatic(flags ? (flags[i] & BS_FLAG_STATIC) != 0 : false);
        net->addParticle(q);
    }
    for (int64_t s = 0; s < m; ++s)
        net->addSpring(id1[s], id2[s], l0[s], k[s]);
    net->setup(make_config(p));
    auto * ff = net->getForceField();
    ff->setImpDoubleMembraneUpperMembOffset(p->imp_upper_offset);
    ff->setImpDoubleMembraneLowerMembOffset(p->imp_lower_offset);
    ff->setImpDoubleMembraneUpperMembTubeCurv(p->imp_upper_curvature);
    ff->setImpDoubleMembraneLowerMembTubeCurv(p->imp_lower_curvature);
    return net;
}

void read_energies(const SpringNetwork & net, float * E6)
{
    E6[BS_E_SPRING] = net.getSpringEnergy();
    E6[BS_E_ELECTROSTATIC] = net.getElectrostaticEnergy();
    E6[BS_E_STERIC] = net.getStericEnergy();
    E6[BS_E_KINETIC] = net.getKineticEnergy();
    E6[BS_E_IMP] = net.getIMPEnergy();
    E6[BS_E_HYDROPHOBIC] = net.getHydrophobicEnergy();
}

} // namespace

extern "C" {

void ref_set_threads(int n)
{
#ifdef _OPENMP
    omp_set_num_threads(n);
#else
    (void)n;
#endif
}

void * ref_create(const bs_params * p, int64_t n, const float * x3, const float * v3, const float * mass,
                  const float * charge, const float * radius, const float * eps, const float * sasa,
                  const float * transfer, const float * hydro, const uint8_t * flags, int64_t m,
                  const uint32_t * id1, const uint32_t * id2, const float * l0, const float * k)
{
    auto * s = new RefSys();
    s->p = *p;
    s->n = n;
    s->net.reset(build_net(p, n, x3, v3, mass, charge, radius, eps, sasa, transfer, hydro, flags, m, id1, id2, l0, k));
    return s;
}

void ref_destroy(void * h)
{
    biospring::rigidbody::RigidBodiesManager::CleanRigidBodies();
    delete static_cast<RefSys *>(h);
}

// Bodies are made rigid through the reference manager; returns the number of
// bodies in its (global) collection.
int ref_add_rigid_body(void * h, int64_t n, const int64_t * ids)
{
    auto * s = static_cast<RefSys *>(h);
    std::vector<unsigned> v;
    for (int64_t t = 0; t < n; ++t)
        v.push_back((unsigned)ids[t]);
    biospring::rigidbody::RigidBodiesManager::InitRigidBodies(s->net.get(), v);
    return (int)biospring::rigidbody::RigidBodiesManager::getCollection().size();
}

// Body state: position, velocity, omega (3 each); the reference keeps L private.
int ref_rigid_state(int b, double * out9)
{
    const auto & col = biospring::rigidbody::RigidBodiesManager::getCollection();
    if (b < 0 || (size_t)b >= col.size())
        return -1;
    const Vector3f X = col[b]->getPos(), w = col[b]->getOmega();
    out9[0] = X.getX();
    out9[1] = X.getY();
    out9[2] = X.getZ();
    out9[6] = w.getX();
    out9[7] = w.getY();
    out9[8] = w.getZ();
    out9[3] = out9[4] = out9[5] = 0.0;
    return 0;
}

// Grid `kind` with origin, shape and cell size, scalar values z fastest.
void ref_set_grid(void * h, int kind, const double * origin, const int64_t * shape, const double * cell,
                  const float * scalar, double scale)
{
    auto * s = static_cast<RefSys *>(h);
    RefNet & net = *s->net;
    biospring::grid::PotentialGrid & g = kind == BS_GRID_POTENTIAL ? net.getPotentialGrid() : net.getDensityGrid();
    std::array<double, 6> box{origin[0], origin[1], origin[2], origin[0] + shape[0] * cell[0],
                              origin[1] + shape[1] * cell[1], origin[2] + shape[2] * cell[2]};
    g.reshape(box, {cell[0], cell[1], cell[2]});
    int64_t t = 0;
    for (int i = 0; i < shape[0]; ++i)
        for (int j = 0; j < shape[1]; ++j)
            for (int k = 0; k < shape[2]; ++k)
                g.at(biospring::grid::discrete_coordinates(i, j, k)).scalar = scalar[t++];
    g.compute_gradient();
    net.enable_grid(kind, scale);
}

void ref_set_burying(void * h, const float * burying)
{
    auto * s = static_cast<RefSys *>(h);
    for (unsigned i = 0; i < (unsigned)s->n; ++i)
        s->net->getParticle(i).setBurying(burying[i]);
}

// angle, roll angle, insertion depth, vector
void ref_insertion(void * h, int64_t i1, int64_t i2, double * out6)
{
    auto * s = static_cast<RefSys *>(h);
    RefNet & net = *s->net;
    InsertionVector iv(net, net.getParticle((unsigned)i1), net.getParticle((unsigned)i2));
    iv.computeVector();
    iv.computeAngle();
    iv.computeRollAngle();
    out6[0] = iv.getAngle();
    out6[1] = iv.getRollAngle();
    out6[2] = iv.getInsertionDepth();
    const Vector3f v = iv.getVector();
    out6[3] = v.getX();
    out6[4] = v.getY();
    out6[5] = v.getZ();
}

// idleRun + computeForces; forces of every particle (static ones too, as the
// reference accumulates them) and the six energies.
void ref_compute_forces(void * h, const float * ext3, float * F3, float * E6)
{
    auto * s = static_cast<RefSys *>(h);
    RefNet & net = *s->net;
    net.idleRun();
    if (ext3)
        for (unsigned i = 0; i < (unsigned)s->n; ++i)
        {
            float f[3] = {ext3[3 * i], ext3[3 * i + 1], ext3[3 * i + 2]};
            net.setForce(i, f);
        }
    net.computeForces();
    for (unsigned i = 0; i < (unsigned)s->n; ++i)
    {
        const Vector3f f = net.getParticle(i).getForce();
        F3[3 * i] = f.getX();
        F3[3 * i + 1] = f.getY();
        F3[3 * i + 2] = f.getZ();
    }
    read_energies(net, E6);
}

// nsteps of computeStep, each preceded by setForce(ext) when ext3 != NULL
// (the interactor path: forces are added before computeForces).
int ref_step(void * h, int nsteps, const float * ext3, float * E6)
{
    auto * s = static_cast<RefSys *>(h);
    RefNet & net = *s->net;
    for (int t = 0; t < nsteps; ++t)
    {
        if (ext3)
        {
            // computeStep begins with idleRun, which resets energies and adds
            // interactor forces; setForce before it is equivalent because the
            // particle forces were reset at the end of the previous step.
            for (unsigned i = 0; i < (unsigned)s->n; ++i)
            {
                float f[3] = {ext3[3 * i], ext3[3 * i + 1], ext3[3 * i + 2]};
                net.setForce(i, f);
            }
        }
        net.computeStep();
    }
    if (E6)
        read_energies(net, E6);
    return 0;
}

void ref_get_state(void * h, float * x3, float * v3)
{
    auto * s = static_cast<RefSys *>(h);
    RefNet & net = *s->net;
    for (unsigned i = 0; i < (unsigned)s->n; ++i)
    {
        const Particle & q = net.getParticle(i);
        if (x3)
        {
            x3[3 * i] = q.getX();
            x3[3 * i + 1] = q.getY();
            x3[3 * i + 2] = q.getZ();
        }
        if (v3)
        {
            const Vector3f v = q.getVelocity();
            v3[3 * i] = v.getX();
            v3[3 * i + 1] = v.getY();
            v3[3 * i + 2] = v.getZ();
        }
    }
}

// The probe particle (SpringNetwork::getProbeParticle, SpringNetwork.h:277):
// position, velocity, then the force of its last step (reset at the start of
// the probe block and then accumulated, SpringNetwork.cpp:158-177).
int ref_get_probe(void * h, float * out9)
{
    auto * s = static_cast<RefSys *>(h);
    if (!s->net->isProbeEnabled())
        return -1;
    const Particle & q = s->net->getProbeParticle();
    const Vector3f x = q.getPosition(), v = q.getVelocity(), f = q.getForce();
    const Vector3f * vs[3] = {&x, &v, &f};
    for (int k = 0; k < 3; ++k)
    {
        out9[3 * k] = vs[k]->getX();
        out9[3 * k + 1] = vs[k]->getY();
        out9[3 * k + 2] = vs[k]->getZ();
    }
    return 0;
}

// Cell list of search `kind` built by the reference NeighborSearch over the
// network's current particles.  Returns the number of cells; cell_of[n]
// (-1 for non-members), start[ncells+1], members[nmem], dims[3], box[6].
int64_t ref_cells(void * h, int kind, float cutoff, float skin, int64_t * cell_of, int64_t * start,
                  int64_t * members, int64_t * dims, double * box, int64_t * nmem)
{
    auto * s = static_cast<RefSys *>(h);
    RefNet & net = *s->net;
    const Container & parts = net.getParticles();
    // the probe, when enabled, is the last particle: in the box, not in the
    // cells (SpringNetwork::_excludeProbeFromNeighborSearch)
    const size_t np = (size_t)s->n;
    std::vector<size_t> inc;
    for (size_t i = 0; i < np; ++i)
    {
        const bool member = kind == BS_NS_ELECTROSTATIC ? parts[i].isCharged()
                            : kind == BS_NS_HYDROPHOBIC ? parts[i].isHydrophobic()
                                                        : true;
        if (member)
            inc.push_back(i);
    }
    std::unique_ptr<ExposedSearch> ns;
    if (kind == BS_NS_STERIC)
        ns.reset(new ExposedSearch(parts, cutoff, skin));
    else
        ns.reset(new ExposedSearch(parts, cutoff, inc, skin));
    if (parts.size() > np)
        ns->exclude_index(np);
    int64_t d[3];
    ns->dims(d);
    const int64_t ncells = d[0] * d[1] * d[2];
    if (dims)
        std::memcpy(dims, d, sizeof d);
    if (box)
        ns->box(box);
    std::vector<int64_t> member_flag(np, 0);
    for (size_t i : inc)
        member_flag[i] = 1;
    if (cell_of)
        for (size_t i = 0; i < np; ++i)
            cell_of[i] = member_flag[i] ? ns->cell_of(i) : -1;
    int64_t w = 0;
    for (int64_t c = 0; c < ncells; ++c)
    {
        if (start)
            start[c] = w;
        const auto * v = ns->cell((size_t)c);
        if (v)
            for (size_t i : *v)
            {
                if (members)
                    members[w] = (int64_t)i;
                ++w;
            }
    }
    if (start)
        start[ncells] = w;
    if (nmem)
        *nmem = w;
    return ncells;
}

// Processed pairs of `kind` at the current positions, recorded from the
// reference's own pair loops.  Builds a copy of the system whose per-particle
// property for that kind encodes the particle id (membership unchanged).
int64_t ref_pairs(const bs_params * p, int kind, int64_t n, const float * x3, const float * charge,
                  const float * hydro, const uint8_t * flags, int64_t m, const uint32_t * id1,
                  const uint32_t * id2, const float * l0, const float * k, uint32_t * ij, int64_t cap)
{
    std::vector<float> enc_eps(n), enc_q(n), enc_h(n);
    for (int64_t i = 0; i < n; ++i)
    {
        enc_eps[i] = (float)i;
        const float q = charge ? charge[i] : 0.0f;
        const float hh = hydro ? hydro[i] : 0.0f;
        enc_q[i] = q != 0.0f ? (float)(i + 1) : 0.0f;
        enc_h[i] = hh != 0.0f ? (float)(i + 1) : 0.0f;
    }
    bs_params q = *p;
    q.viscosity_enable = 0;
    q.imp_enable = 0;
    // the probe stays in the container (and so in the search box) but its
    // own interactions, which are not neighbour pairs, are not recorded
    q.probe_enable_steric = 0;
    q.probe_enable_electrostatic = 0;
    std::unique_ptr<RefNet> net(build_net(&q, n, x3, nullptr, nullptr, enc_q.data(), nullptr, enc_eps.data(), nullptr,
                                          nullptr, enc_h.data(), flags, m, id1, id2, l0, k));
    auto * rec = new RecordingForceField();
    rec->kind = kind;
    net->swap_forcefield(std::unique_ptr<biospring::forcefield::ForceField>(rec));
    net->idleRun();
    net->computeParticleForces();
    std::vector<std::pair<uint32_t, uint32_t>> v;
    for (size_t t = 0; t + 1 < rec->ij.size(); t += 2)
        v.emplace_back(rec->ij[t], rec->ij[t + 1]);
    std::sort(v.begin(), v.end());
    const int64_t cnt = (int64_t)v.size();
    if (ij)
        for (int64_t t = 0; t < cnt && t < cap; ++t)
        {
            ij[2 * t] = v[t].first;
            ij[2 * t + 1] = v[t].second;
        }
    return cnt;
}

// Springs of Topology::add_springs_from_cutoff: returns the count; with cap
// >= count, writes id1/id2 (indices in the given order) and the rest lengths
// as SpringNetwork::addSpring receives them (float).
int64_t ref_springs_from_cutoff(int64_t n, const float * x3, const uint8_t * flags, const int32_t * groups, int64_t m0,
                                const uint32_t * e1, const uint32_t * e2, const float * el0, const float * ek,
                                double cutoff, double stiffness, int64_t cap, uint32_t * id1, uint32_t * id2,
                                float * l0, float * k)
{
    using biospring::topology::Topology;
    using biospring::topology::Particle;
    std::vector<Particle> ps((size_t)n);
    for (int64_t i = 0; i < n; ++i)
    {
        ps[(size_t)i].set_position(Vector3f(x3[3 * i], x3[3 * i + 1], x3[3 * i + 2]));
        ps[(size_t)i].properties().set_static(flags && (flags[i] & BS_FLAG_STATIC));
        if (groups)
            ps[(size_t)i].properties().set_topology_id((size_t)groups[i]);
    }
    Topology top;
    top.add_particles(ps);
    std::unordered_map<int64_t, uint32_t> index; // the collection copies particles with new ids
    for (int64_t i = 0; i < n; ++i)
        index[(int64_t)top.get_particle((size_t)i).unique_id()] = (uint32_t)i;
    for (int64_t s = 0; s < m0; ++s) // springs present before, e.g. from CONECT records
        top.add_spring((size_t)e1[s], (size_t)e2[s], (double)el0[s], (double)ek[s]);
    if (groups)
        top.add_springs_between_topologies_from_cutoff(cutoff, stiffness);
    else
        top.add_springs_from_cutoff(cutoff, stiffness);
    const int64_t m = (int64_t)top.number_of_springs();
    if (cap >= m)
        for (int64_t s = 0; s < m; ++s)
        {
            const auto & sp = top.get_spring((size_t)s);
            id1[s] = index.at((int64_t)sp.first().unique_id());
            id2[s] = index.at((int64_t)sp.second().unique_id());
            l0[s] = (float)sp.equilibrium();
            k[s] = (float)sp.stiffness();
        }
    return m;
}

void ref_set_trajectories(const char * pdb, int64_t pdbf, const char * csv, int64_t csvf)
{
    g_traj.pdb = pdb ? pdb : "";
    g_traj.csv = csv ? csv : "";
    g_traj.pdbf = (size_t)pdbf;
    g_traj.csvf = (size_t)csvf;
}

// Frames k = 0..nframes-1 with step = time = k and a unit box, as the
// reference's XTC writer means to write them.
int ref_xtc_write(const char * path, int nframes, int natoms, const float * x3nm, float prec)
{
    XDRFILE * xd = xdrfile_open(path, "w");
    if (!xd)
        return -1;
    matrix box = {{1.0f, 0.0f, 0.0f}, {0.0f, 1.0f, 0.0f}, {0.0f, 0.0f, 1.0f}};
    std::vector<float> buf((size_t)3 * natoms);
    for (int k = 0; k < nframes; ++k)
    {
        std::memcpy(buf.data(), x3nm + (size_t)k * 3 * natoms, sizeof(float) * 3 * natoms);
        if (write_xtc(xd, natoms, k, (float)k, box, reinterpret_cast<rvec *>(buf.data()), prec) != exdrOK)
        {
            xdrfile_close(xd);
            return -2;
        }
    }
    xdrfile_close(xd);
    return 0;
}

// Reads up to maxframes frames; returns the number read.
int ref_xtc_read(const char * path, int natoms, int maxframes, float * x3nm, int * steps, float * times,
                 float * box9, float * precs)
{
    XDRFILE * xd = xdrfile_open(path, "r");
    if (!xd)
        return -1;
    int k = 0;
    for (; k < maxframes; ++k)
    {
        matrix box;
        int step = 0;
        float t = 0.0f, pr = 0.0f;
        if (read_xtc(xd, natoms, &step, &t, box, reinterpret_cast<rvec *>(x3nm + (size_t)k * 3 * natoms), &pr) !=
            exdrOK)
            break;
        steps[k] = step;
        times[k] = t;
        precs[k] = pr;
        std::memcpy(box9 + 9 * k, box, sizeof box);
    }
    xdrfile_close(xd);
    return k;
}

} // extern "C"
