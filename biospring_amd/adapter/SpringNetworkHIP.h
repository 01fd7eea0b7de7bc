// SpringNetworkHIP.h -- BioSpring's SpringNetwork with its step on an MI355X.
//
// The reference's extension point for another compute backend is a subclass
// of biospring::spn::SpringNetwork that overrides its virtual step methods
// (src/spn/SpringNetwork.h:303-332,395-397), as its OpenCL backend does
// (src/opencl/SpringNetworkOpenCL.h:84).  This subclass forwards them to the
// engine's C ABI (include/bs_api.h):
//
//   initRun                 SpringNetwork::initRun, then the network, the
//                           parameters (Configuration + ForceField), rigid
//                           bodies, grids and the insertion vector are copied
//                           to the engine and set up (bs_set_particles,
//                           bs_set_springs, bs_set_params, bs_add_rigid_body,
//                           bs_set_grid, bs_setup)
//   computeStep             idleRun (interactors, trajectory writers, energy
//                           reset, iteration count) on the host, then one
//                           engine step (bs_step) with the external forces
//                           the interactors set; energies come back every
//                           step, positions when something on the host reads
//                           them
//   computeForces           bs_compute_forces: forces and energies at the
//                           current state into the host particles
//   getParticlePosition     the engine's current position of particle i
//   setForce                an external force for the next step (added like
//                           SpringNetwork::setForce, before the forces)
//   endRun                  the final state back into the host particles
//
// Errors keep the reference's conventions: a non-finite position ends the
// program through logging::die with the reference's message
// (SpringNetwork.cpp:222-229); any other engine failure throws
// std::runtime_error with bs_last_error's text.
#pragma once

#include "SpringNetwork.h"

#include "bs_api.h"

#include <vector>

namespace biospring
{
namespace spn
{

class SpringNetworkHIP : public SpringNetwork
{
  public:
    SpringNetworkHIP() = default;
    ~SpringNetworkHIP() override;

    void initRun() override;
    void computeStep() override;
    void computeForces() override;
    void endRun() override;

    void getParticlePosition(unsigned i, float position[3]) const override;
    void setForce(unsigned i, float force[3]) override;

    // Copies the engine's positions and velocities into the host particles
    // (the trajectory writers, the insertion vector and rigid bodies read
    // them there).
    void syncParticlesFromDevice();

    // The engine's handle (nullptr before initRun).
    bs_handle * handle() const { return _h; }

    // The parameter block initRun gives the engine, from the configuration
    // and force field of this network.
    bs_params engineParams() const;

  private:
    void _check(int rc, const char * what) const;
    void _readEnergies();
    bool _hostReadsPositions() const;
    unsigned _numNetworkParticles() const;

    bs_handle * _h = nullptr;
    unsigned _n = 0;
    std::vector<float> _ext;
    bool _extSet = false;
    mutable std::vector<float> _x;
    mutable bool _xValid = false;
};

} // namespace spn
} // namespace biospring
