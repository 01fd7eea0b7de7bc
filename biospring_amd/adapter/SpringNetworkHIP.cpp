// SpringNetworkHIP.cpp -- see SpringNetworkHIP.h.
//
// Built on the reference side, against its headers (src/, src/spn/) and the
// engine's include/bs_api.h, and linked with the reference's objects and the
// engine library (libbiospring_amd.so).
#include "SpringNetworkHIP.h"

#include "logging.h"
#include "rigidbody/RigidBodiesManager.h"

#include <stdexcept>
#include <string>

namespace biospring
{
namespace spn
{

SpringNetworkHIP::~SpringNetworkHIP()
{
    if (_h)
        bs_destroy(_h);
}

void SpringNetworkHIP::_check(int rc, const char * what) const
{
    if (rc == BS_OK)
        return;
    if (rc == BS_ERR_NONFINITE) // the reference's fatal condition (SpringNetwork.cpp:222-229)
        logging::die("Found non-finite position for particle %d.", (int)bs_last_nonfinite_particle(_h));
    throw std::runtime_error(std::string("SpringNetworkHIP: ") + what + ": " + bs_last_error(_h));
}

// The probe, when enabled, is the last of _particles (SpringNetwork::_setupProbe,
// SpringNetwork.cpp:763-764); the engine keeps it apart (bs_get_probe).
unsigned SpringNetworkHIP::_numNetworkParticles() const
{
    const unsigned n = getNumberOfParticles();
    return isProbeEnabled() && _probeparticule.getId() >= 0 && n > 0 ? n - 1 : n;
}

bs_params SpringNetworkHIP::engineParams() const
{
    bs_params p;
    bs_default_params(&p);
    p.timestep = getTimeStep();
    p.neighborskin = getNeighborSkin();
    p.spring_enable = isSpringEnabled();
    p.spring_scale = _config.spring.scale;
    p.steric_enable = isStericEnabled();
    // SpringNetwork::_setupForceField (SpringNetwork.cpp:690-710): unknown
    // modes fall back to linear
    const std::string & mode = _config.steric.mode;
    p.steric_mode = mode == "lennard-jones-8-6Lewitt"      ? BS_STERIC_LEWITT_8_6
                    : mode == "lennard-jones-8-6Zacharias" ? BS_STERIC_ZACHARIAS_8_6
                    : mode == "lennard-jones-12-6Amber"    ? BS_STERIC_AMBER_12_6
                                                           : BS_STERIC_LINEAR;
    p.steric_cutoff = getStericCutoff();
    p.steric_gridscale = getGridScale();
    p.electrostatic_enable = isElectrostaticEnabled();
    p.electrostatic_cutoff = getElectrostaticCutoff();
    p.electrostatic_scale = _config.electrostatic.scale;
    p.electrostatic_dielectric = _config.electrostatic.dielectric;
    p.dielectric_model = BS_DIELECTRIC_CONSTANT;
    p.viscosity_enable = isViscosityEnabled();
    p.viscosity_value = getViscosity();
    p.imp_enable = isIMPEnabled();
    p.imp_scale = _config.imp.scale;
    if (const forcefield::ForceField * ff = getForceField())
    { // the membrane set by the interactor's custom data (CustomData.cpp:61-108)
        p.imp_upper_offset = ff->getImpDoubleMembraneUpperMembOffset();
        p.imp_lower_offset = ff->getImpDoubleMembraneLowerMembOffset();
        p.imp_upper_curvature = ff->getImpDoubleMembraneUpperMembTubeCurv();
        p.imp_lower_curvature = ff->getImpDoubleMembraneLowerMembTubeCurv();
    }
    p.hydrophobicity_enable = isHydrophobicityEnabled();
    p.hydrophobicity_cutoff = getHydrophobicCutoff();
    p.hydrophobicity_scale = _config.hydrophobicity.scale;
    p.rigidbody_enable = isRigidBodyEnabled();
    p.potentialgrid_enable = isElectrostaticFieldEnabled();
    p.potentialgrid_scale = _config.potentialgrid.scale;
    p.densitygrid_enable = isDensityGridEnabled();
    p.densitygrid_scale = getDensityGridScale();
    p.probe_enable = isProbeEnabled();
    p.probe_enable_electrostatic = isProbeElectrostaticEnabled();
    p.probe_enable_steric = isProbeStericEnabled();
    p.probe_x = _config.probe.x;
    p.probe_y = _config.probe.y;
    p.probe_z = _config.probe.z;
    p.probe_mass = _config.probe.mass;
    p.probe_epsilon = _config.probe.epsilon;
    p.probe_radius = _config.probe.radius;
    p.probe_charge = _config.probe.charge;
    return p;
}

// A grid as the engine takes it: origin, shape, cell size and the scalar per
// cell, z fastest (the engine forms the gradient as
// PotentialGrid::compute_gradient does).
static int upload_grid(bs_handle * h, int kind, const grid::PotentialGrid & g)
{
    const auto o = g.origin();
    const auto & sh = g.shape();
    const auto & cs = g.cell_size();
    const double origin[3] = {o[0], o[1], o[2]}, cell[3] = {cs[0], cs[1], cs[2]};
    const int64_t shape[3] = {(int64_t)sh[0], (int64_t)sh[1], (int64_t)sh[2]};
    std::vector<float> scalar((size_t)(shape[0] * shape[1] * shape[2]));
    size_t t = 0;
    for (int i = 0; i < (int)sh[0]; ++i)
        for (int j = 0; j < (int)sh[1]; ++j)
            for (int k = 0; k < (int)sh[2]; ++k)
                scalar[t++] = g.at(grid::discrete_coordinates(i, j, k)).scalar;
    return bs_set_grid(h, kind, origin, shape, cell, scalar.data());
}

void SpringNetworkHIP::initRun()
{
    SpringNetwork::initRun(); // timers, iteration count, interactor threads
    if (_h)
    {
        bs_destroy(_h);
        _h = nullptr;
    }
    const bs_params p = engineParams();
    _check(bs_create(&p, &_h), "bs_create");
    _n = _numNetworkParticles();
    const size_t n = _n;
    std::vector<float> x(3 * n), v(3 * n), mass(n), q(n), r(n), eps(n), sasa(n), tr(n), hyd(n), bur(n);
    std::vector<uint8_t> fl(n);
    for (size_t i = 0; i < n; ++i)
    {
        const Particle & pp = _particles[i];
        const Vector3f vel = pp.getVelocity();
        x[3 * i] = pp.getX();
        x[3 * i + 1] = pp.getY();
        x[3 * i + 2] = pp.getZ();
        v[3 * i] = vel.getX();
        v[3 * i + 1] = vel.getY();
        v[3 * i + 2] = vel.getZ();
        mass[i] = pp.getMass();
        q[i] = pp.getCharge();
        r[i] = pp.getRadius();
        eps[i] = pp.getEpsilon();
        sasa[i] = pp.getSolventAccessibilitySurface();
        tr[i] = pp.getTransferEnergyByAccessibleSurface();
        hyd[i] = pp.getHydrophobicity();
        bur[i] = pp.getBurying();
        fl[i] = pp.isStatic() ? BS_FLAG_STATIC : 0u;
    }
    _check(bs_set_particles(_h, (int64_t)n, x.data(), v.data(), mass.data(), q.data(), r.data(), eps.data(),
                            sasa.data(), tr.data(), hyd.data(), fl.data()),
           "bs_set_particles");
    const std::vector<Spring> & springs = getSprings();
    std::vector<uint32_t> a(springs.size()), b(springs.size());
    std::vector<float> l0(springs.size()), k(springs.size());
    for (size_t s = 0; s < springs.size(); ++s)
    {
        a[s] = (uint32_t)springs[s].getParticle1().getId();
        b[s] = (uint32_t)springs[s].getParticle2().getId();
        l0[s] = springs[s].getEquilibrium();
        k[s] = springs[s].getStiffness();
    }
    _check(bs_set_springs(_h, (int64_t)springs.size(), a.data(), b.data(), l0.data(), k.data()), "bs_set_springs");
    _check(bs_set_burying(_h, bur.data()), "bs_set_burying");
    if (isElectrostaticFieldEnabled())
        _check(upload_grid(_h, BS_GRID_POTENTIAL, getPotentialGrid()), "bs_set_grid (potential)");
    if (isDensityGridEnabled())
        _check(upload_grid(_h, BS_GRID_DENSITY, getDensityGrid()), "bs_set_grid (density)");
    // bodies made before the run (RigidBodiesManager::InitRigidBodies)
    for (const rigidbody::RigidBody * body : rigidbody::RigidBodiesManager::getCollection())
    {
        const std::vector<unsigned> ids = body->getParticlesIds();
        std::vector<int64_t> id64(ids.begin(), ids.end());
        _check(bs_add_rigid_body(_h, (int64_t)id64.size(), id64.data()), "bs_add_rigid_body");
    }
    _check(bs_setup(_h), "bs_setup");
    _ext.assign(3 * n, 0.0f);
    _extSet = false;
    _xValid = false;
}

// idleRun's readers of the host particles: trajectory writers, the insertion
// vector and constraints.
bool SpringNetworkHIP::_hostReadsPositions() const
{
    return isPDBTrajectoryWriterEnabled() || isXTCTrajectoryWriterEnabled() || isCSVTrajectoryWriterEnabled() ||
           isInsertionVectorEnabled() || isConstraintEnabled();
}

void SpringNetworkHIP::_readEnergies()
{
    double e[BS_NUM_ENERGIES];
    _check(bs_get_energies(_h, e), "bs_get_energies");
    _energies.spring = (float)e[BS_E_SPRING];
    _energies.electrostatic = (float)e[BS_E_ELECTROSTATIC];
    _energies.steric = (float)e[BS_E_STERIC];
    _energies.kinetic = (float)e[BS_E_KINETIC];
    _energies.imp = (float)e[BS_E_IMP];
    _energies.hydrophobic = (float)e[BS_E_HYDROPHOBIC];
}

// SpringNetwork::computeStep (SpringNetwork.cpp:251-268): idleRun on the
// host, then forces, rigid bodies and the integration as one engine step.
void SpringNetworkHIP::computeStep()
{
    if (!_h)
        initRun();
    if (_hostReadsPositions())
        syncParticlesFromDevice(); // frames are written from the host particles
    idleRun();                     // interactors call getParticlePosition / setForce here
    _meanConstraintsDistances = 0.0;
    const int rc = bs_step(_h, 1, _extSet ? _ext.data() : nullptr);
    _check(rc, "bs_step");
    if (_extSet)
    { // setForce adds to the force of one step (the particle's force is reset after it)
        std::fill(_ext.begin(), _ext.end(), 0.0f);
        _extSet = false;
    }
    _xValid = false;
    _readEnergies();
    if (isInsertionVectorEnabled())
    {
        syncParticlesFromDevice();
        _updateInsertionVector();
    }
}

// SpringNetwork::computeForces (SpringNetwork.cpp:244-249): the forces and
// energies of the current state, into the host particles.
void SpringNetworkHIP::computeForces()
{
    if (!_h)
        initRun();
    std::vector<float> f(3 * (size_t)_n);
    double e[BS_NUM_ENERGIES];
    _check(bs_compute_forces(_h, _extSet ? _ext.data() : nullptr, f.data(), e), "bs_compute_forces");
    for (unsigned i = 0; i < _n; ++i)
        _particles[i].setForce(Vector3f(f[3 * i], f[3 * i + 1], f[3 * i + 2]));
    _energies.spring = (float)e[BS_E_SPRING];
    _energies.electrostatic = (float)e[BS_E_ELECTROSTATIC];
    _energies.steric = (float)e[BS_E_STERIC];
    _energies.imp = (float)e[BS_E_IMP];
    _energies.hydrophobic = (float)e[BS_E_HYDROPHOBIC];
}

void SpringNetworkHIP::endRun()
{
    if (_h)
        syncParticlesFromDevice();
    SpringNetwork::endRun();
}

void SpringNetworkHIP::syncParticlesFromDevice()
{
    if (!_h)
        return;
    std::vector<float> x(3 * (size_t)_n), v(3 * (size_t)_n);
    _check(bs_get_state(_h, x.data(), v.data()), "bs_get_state");
    for (unsigned i = 0; i < _n; ++i)
    {
        _particles[i].setPosition(Vector3f(x[3 * i], x[3 * i + 1], x[3 * i + 2]));
        _particles[i].setVelocity(Vector3f(v[3 * i], v[3 * i + 1], v[3 * i + 2]));
    }
    if (isProbeEnabled() && _probeparticule.getId() >= 0)
    { // SpringNetwork::_syncProbeParticle
        float px[3], pv[3];
        _check(bs_get_probe(_h, px, pv, nullptr), "bs_get_probe");
        _probeparticule.setPosition(Vector3f(px[0], px[1], px[2]));
        _probeparticule.setVelocity(Vector3f(pv[0], pv[1], pv[2]));
        _particles[(size_t)_probeparticule.getId()] = _probeparticule;
        _particles[(size_t)_probeparticule.getId()].setSpringNetwork(this);
    }
    _x = x;
    _xValid = true;
}

// SpringNetwork::getParticlePosition (SpringNetwork.cpp:430-435), from the
// engine's state (one download per step, on the first call).
void SpringNetworkHIP::getParticlePosition(unsigned i, float position[3]) const
{
    if (!_h)
    {
        SpringNetwork::getParticlePosition(i, position);
        return;
    }
    if (i >= _n)
    {
        if (isProbeParticle(i))
        {
            _check(bs_get_probe(_h, position, nullptr, nullptr), "bs_get_probe");
            return;
        }
        throw std::out_of_range("SpringNetworkHIP::getParticlePosition: particle index out of range");
    }
    if (!_xValid)
    {
        _x.resize(3 * (size_t)_n);
        _check(bs_get_state(_h, _x.data(), nullptr), "bs_get_state");
        _xValid = true;
    }
    position[0] = _x[3 * i];
    position[1] = _x[3 * i + 1];
    position[2] = _x[3 * i + 2];
}

// SpringNetwork::setForce (SpringNetwork.cpp:436-442): added to particle i's
// force of the next step, before the force evaluation.
void SpringNetworkHIP::setForce(unsigned i, float force[3])
{
    if (!_h)
    {
        SpringNetwork::setForce(i, force);
        return;
    }
    if (i >= _n)
    {
        if (isProbeParticle(i))
            return; // the probe's force is reset before its own terms (SpringNetwork.cpp:158)
        throw std::out_of_range("SpringNetworkHIP::setForce: particle index out of range");
    }
    _ext[3 * i] += force[0];
    _ext[3 * i + 1] += force[1];
    _ext[3 * i + 2] += force[2];
    _extSet = true;
}

} // namespace spn
} // namespace biospring
