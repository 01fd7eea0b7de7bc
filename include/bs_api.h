/*
 * bs_api.h -- C ABI of the MI355X force-and-integrate engine for BioSpring
 * spring networks.
 *
 * The engine replaces the per-step path of biospring::spn::SpringNetwork
 * (src/spn/SpringNetwork.cpp:63-268): spring forces, the non-bonded pair
 * forces found through the cell-list search (src/nsearch.hpp:155-436),
 * viscosity, IMPALA membrane forces and the viscous semi-implicit Euler
 * integrator.  A host-side subclass of SpringNetwork forwards its virtual
 * step methods here; every entry point takes plain pointers and sizes, never
 * throws, and reports failure through its return code plus bs_last_error().
 *
 * Units are the reference's: Angstrom, Dalton, femtosecond, kJ/mol, e
 * (src/forcefield/constants.hpp:4-27).  Arrays of 3-vectors are packed
 * x0,y0,z0,x1,... in float (positions, velocities, forces).
 *
 * Threading: calls on one handle must be serialised by the caller (the
 * reference's main-thread model, src/spn/SpringNetwork.cpp:344-365).
 * Different handles are independent.
 */
#ifndef BIOSPRING_AMD_BS_API_H
#define BIOSPRING_AMD_BS_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- return codes ------------------------------------------------------ */
#define BS_OK             0
#define BS_ERR_ARG        1  /* bad argument: std::invalid_argument / out_of_range in the reference */
#define BS_ERR_NONFINITE  2  /* non-finite position: logging::die in SpringNetwork.cpp:222-229 */
#define BS_ERR_HIP        3  /* HIP runtime failure */
#define BS_ERR_STATE      4  /* call out of order (e.g. step before setup), capacity exceeded */
#define BS_ERR_NOGPU      5  /* no usable gfx950 device */

/* ---- steric modes (SpringNetwork::_setupForceField, SpringNetwork.cpp:690-710) */
#define BS_STERIC_LINEAR      0  /* "linear" and any unknown mode */
#define BS_STERIC_AMBER_12_6  1  /* "lennard-jones-12-6Amber" */
#define BS_STERIC_LEWITT_8_6  2  /* "lennard-jones-8-6Lewitt" */
#define BS_STERIC_ZACHARIAS_8_6 3 /* "lennard-jones-8-6Zacharias" */

#define BS_PAIR_CLUSTERS 0  /* bs_params.pair_layout */
#define BS_PAIR_ROWS     1
#define BS_PAIR_QUEUES   2

/* ---- dielectric models ------------------------------------------------- */
#define BS_DIELECTRIC_CONSTANT 0  /* reference: electrostatic.hpp:16-49 */
#define BS_DIELECTRIC_DISTANCE 1  /* extension: eps(r) = eps * r (r in Angstrom); no reference counterpart */

/* ---- per-particle flag bits (bs_set_particles) -------------------------- */
#define BS_FLAG_STATIC  0x01u    /* Particle::isStatic (Particle.h:57) */

/* ---- energy slots, FP64 (SpringNetwork::Energies, SpringNetwork.h:51-69) */
#define BS_E_SPRING        0
#define BS_E_ELECTROSTATIC 1
#define BS_E_STERIC        2
#define BS_E_KINETIC       3
#define BS_E_IMP           4
#define BS_E_HYDROPHOBIC   5
#define BS_NUM_ENERGIES    6

/* ---- neighbour-search kinds (SpringNetwork::NeighborSearch, SpringNetwork.h:79-100) */
#define BS_NS_STERIC        0  /* all particles, steric.cutoff */
#define BS_NS_ELECTROSTATIC 1  /* charged particles, electrostatic.cutoff */
#define BS_NS_HYDROPHOBIC   2  /* hydrophobic particles, hydrophobicity.cutoff */

/* ---- grids (SpringNetwork::_grids, SpringNetwork.h:45-49) ---------------- */
#define BS_GRID_POTENTIAL   0  /* electrostatic potential, potentialgrid.* */
#define BS_GRID_DENSITY     1  /* density, densitygrid.* */

/*
 * Simulation parameters.  Field names and defaults follow the reference's
 * Configuration (src/configuration/Configuration.hpp:153-209); the msp key
 * is given beside each field.  bs_default_params() fills the defaults.
 */
typedef struct bs_params {
    float timestep;               /* simulation.timestep (fs), default 0.01 */
    float neighborskin;           /* simulation.neighborskin (A), default 0 */

    int32_t spring_enable;        /* spring.enable, default 0 */
    float spring_scale;           /* spring.scale, default 1 */

    int32_t steric_enable;        /* steric.enable, default 0 */
    int32_t steric_mode;          /* steric.mode -> BS_STERIC_*, default linear */
    float steric_cutoff;          /* steric.cutoff (A), default 1 */
    float steric_gridscale;       /* steric.gridscale, default 1 */

    int32_t electrostatic_enable; /* electrostatic.enable, default 0 */
    float electrostatic_cutoff;   /* electrostatic.cutoff (A), default 16 */
    float electrostatic_scale;    /* electrostatic.scale, default 1 */
    float electrostatic_dielectric; /* electrostatic.dielectric, default 1 */
    int32_t dielectric_model;     /* BS_DIELECTRIC_*, default constant */

    int32_t viscosity_enable;     /* viscosity.enable, default 0 */
    float viscosity_value;        /* viscosity.value (Da/fs), default 1 */

    int32_t imp_enable;           /* imp.enable, default 0 */
    float imp_scale;              /* imp.scale, default 1 */
    float imp_upper_offset;       /* ForceField::_impuppermebraneoffset (A), default 0 */
    float imp_lower_offset;       /* ForceField::_implowermembraneoffset (A), default 0 */
    float imp_upper_curvature;    /* ForceField::_uppermembtubecurv (1/A), default 0 */
    float imp_lower_curvature;    /* ForceField::_lowermembtubecurv (1/A), default 0 */

    int32_t hydrophobicity_enable; /* hydrophobicity.enable, default 0 */
    float hydrophobicity_cutoff;  /* hydrophobicity.cutoff (A), default 15 */
    float hydrophobicity_scale;   /* hydrophobicity.scale, default 1 */

    /* Engine options with no reference counterpart. */
    float list_skin;              /* Verlet-list margin (A) added to the largest cutoff; the list is
                                     rebuilt whenever a particle drifts more than list_skin/2.  Pair
                                     sets are exact for any value >= 0.  Default 1.0 */
    int32_t rebuild_interval;     /* additionally rebuild every k steps (0 = only on drift). Default 0 */
    int32_t compensated_positions; /* keep a float low word per coordinate (two-sum), default 0 */

    /* Reference options read by the step beyond the force-field block above. */
    int32_t rigidbody_enable;     /* rigidbody.enable: aggregate and solve the bodies of
                                     bs_add_rigid_body each step (SpringNetwork.cpp:184-192,261-262), default 0 */
    int32_t potentialgrid_enable; /* potentialgrid.enable: electrostatic field force of grid
                                     BS_GRID_POTENTIAL (Particle.cpp:161-189); needs electrostatic_enable, default 0 */
    float potentialgrid_scale;    /* potentialgrid.scale (ForceField::_forcefieldscale), default 1 */
    int32_t densitygrid_enable;   /* densitygrid.enable: density field force of grid BS_GRID_DENSITY
                                     (Particle.cpp:136-159), default 0 */
    float densitygrid_scale;      /* densitygrid.scale, default 1 */

    /* Engine option: bs_step replays a captured HIP graph of this many steps
       (rounded up to even; the last step of a call, which also sums the
       energies, is launched directly).  0 launches every kernel directly.
       Results are identical either way.  Default 0 */
    int32_t graph_steps;

    /* Engine option: layout the pair kernels read.  BS_PAIR_QUEUES (the
       default): the cluster lists read in chunks of at most 128 entries,
       each entry's j data loaded once into LDS and queued for the
       particles that see it, so that lanes evaluate listed pairs at the
       cluster layout's load count.  BS_PAIR_CLUSTERS: the cluster lists in
       search order, 8 particles x 8 entries per wave, one load of j's data
       serving 8 lanes; lanes of entries a particle does not see stay idle.
       BS_PAIR_ROWS: the cluster lists expanded at each rebuild into one list
       per particle, every lane evaluating a listed pair at one gather per
       pair (slower where gathers, not arithmetic, bound the kernel).  Forces
       agree to FP64 rounding in every layout. */
    int32_t pair_layout;

    /* The probe particle (probe.*, Configuration.hpp:186-195): one extra
       particle placed at setup (SpringNetwork::_setupProbe,
       SpringNetwork.cpp:749-766) that interacts with every dynamic particle,
       with no cutoff and no spring exclusion, through the steric and/or
       Coulomb terms of the force field (Particle::addStericProbeForce /
       addElectrostaticProbeForce, Particle.cpp:434-465), is never a
       neighbour in the pair searches (SpringNetwork.cpp:816-824,
       Particle.cpp:200,280,379), and is moved once per step by the viscous
       Euler step (SpringNetwork.cpp:156-182).  Its state: bs_get_probe. */
    int32_t probe_enable;               /* probe.enable, default 0 */
    int32_t probe_enable_electrostatic; /* probe.enableelectrostatic, default 0 */
    int32_t probe_enable_steric;        /* probe.enablesteric, default 0 */
    float probe_x;                      /* probe.x (A), default 1 */
    float probe_y;                      /* probe.y (A), default 1 */
    float probe_z;                      /* probe.z (A), default 1 */
    float probe_mass;                   /* probe.mass (Da), default 1 */
    float probe_epsilon;                /* probe.epsilon (kJ/mol), default 1 */
    float probe_radius;                 /* probe.radius (A), default 1 */
    float probe_charge;                 /* probe.charge (e), default 0 */
} bs_params;

typedef struct bs_handle bs_handle;

/* ---- lifetime ------------------------------------------------------------ */

/* Fills *p with the reference defaults (Configuration.hpp:153-209). */
void bs_default_params(bs_params *p);

/* Creates a simulation on the current HIP device.  Replaces constructing a
 * SpringNetwork (SpringNetwork.h:102-112). */
int bs_create(const bs_params *p, bs_handle **out);

/* Frees every device buffer and the stream; h may be NULL. */
int bs_destroy(bs_handle *h);

/* Message of the last failing call on h (static text if h is NULL). */
const char *bs_last_error(const bs_handle *h);

/* Replaces Topology::add_springs_from_cutoff (src/topology/Topology.hpp:105-138)
 * followed by Topology::to_spring_network: one spring for each pair of the
 * loaded particles (at their current positions) closer than `cutoff` (A),
 * unless both are static, with rest length the FP64 distance of the float
 * positions rounded to float (Spring.hpp:47-56, then the float of
 * SpringNetwork::addSpring) and the given stiffness.  The pair test is the
 * reference's: FP64 distance < the cutoff as a float (NeighborSearch keeps
 * it as a float, nsearch.hpp:57,230); cutoffs under 1e-6 (nsearch.hpp:66-67),
 * NaN or infinite -> BS_ERR_ARG.  Built on the GPU.  Springs already loaded
 * (bs_set_springs, an earlier call) stay as they are, and their pairs get no
 * new spring, as the reference's add_spring refuses an existing pair
 * (SpringCollection.hpp:138-151,226-233); bs_set_springs with m = 0 clears
 * them first.  New springs follow the loaded ones, ordered by (first,
 * second) original id, first < second; the reference orders them by first
 * id, then by its cell scan.  With independent systems declared
 * (bs_set_systems), no spring joins two systems.  The table stays on the
 * device, where bs_setup builds the spring rows from it; a host copy is made
 * only when asked for (bs_get_springs, bs_save_netcdf, PDB CONECT records).
 * *count_out (nullable) = number of springs added. */
int bs_springs_from_cutoff(bs_handle *h, double cutoff, float stiffness, int64_t *count_out);

/* Replaces Topology::add_springs_between_topologies_from_cutoff
 * (src/topology/Topology.hpp:140-180), the spring step of merging
 * topologies (mergespn-cli.cpp:51): as bs_springs_from_cutoff, but only for
 * pairs whose particles belong to different groups (group[i], n entries: the
 * topology id of each particle). */
int bs_springs_between_groups_from_cutoff(bs_handle *h, double cutoff, float stiffness, const int32_t *group,
                                          int64_t *count_out);

/* The loaded springs (bs_set_springs or bs_springs_from_cutoff order);
 * every pointer nullable, bs_num_springs entries each. */
int bs_get_springs(const bs_handle *h, uint32_t *id1, uint32_t *id2, float *equilibrium, float *stiffness);

/* Number of particles / springs currently loaded. */
int64_t bs_num_particles(const bs_handle *h);
int64_t bs_num_springs(const bs_handle *h);

/* ---- state upload ------------------------------------------------------ */

/* Replaces SpringNetwork::addParticle (SpringNetwork.cpp:498-527) for a whole
 * system.  Particle ids are the array indices.  Any pointer except x3 may be
 * NULL, which selects the ParticleProperty defaults (ParticleProperty.h:20-25):
 * v=0, mass 1, charge 0, radius 1, epsilon 0, sasa 0, transfer 0,
 * hydrophobicity 0, flags 0 (dynamic).  isCharged = (charge != 0) and
 * isHydrophobic = (hydrophobicity != 0) as in ParticleProperty.cpp:8-35.
 * Must precede bs_set_springs, like the reference's addParticle. */
int bs_set_particles(bs_handle *h, int64_t n, const float *x3, const float *v3,
                     const float *mass, const float *charge, const float *radius,
                     const float *epsilon, const float *sasa, const float *transfer,
                     const float *hydrophobicity, const uint8_t *flags);

/* Replaces SpringNetwork::addSpring (SpringNetwork.cpp:452-486) for a whole
 * set: spring s links id1[s] and id2[s] with rest length equilibrium[s] (A)
 * and stiffness[s] (kJ/mol/A^2).  Ids out of range or equal -> BS_ERR_ARG;
 * a repeated pair is ignored, like the reference's isInSpringNeighbors test.
 * The force acts along x[id2]-x[id1] exactly as Spring::computeForce
 * (Spring.cpp:20-34).  Replaces the loaded springs (m = 0 clears them); the
 * current positions and velocities are kept.  A rejected call changes
 * nothing. */
int bs_set_springs(bs_handle *h, int64_t m, const uint32_t *id1, const uint32_t *id2,
                   const float *equilibrium, const float *stiffness);

/* Replaces SpringNetwork::setParams via setup(conf) (SpringNetwork.cpp:646-661).
 * Takes effect at the next step boundary. */
int bs_set_params(bs_handle *h, const bs_params *p);

/* Builds device structures (spring CSR, exclusions, neighbour lists).
 * Replaces SpringNetwork::setup (SpringNetwork.cpp:646-661). */
int bs_setup(bs_handle *h);

/* ---- step ---------------------------------------------------------------- */

/* Advances nsteps steps of SpringNetwork::computeStep (SpringNetwork.cpp:251-268):
 * forces (springs, pairs, viscosity, IMPALA), viscous Euler integration of every
 * dynamic particle, neighbour upkeep.  ext_force3 (nullable, 3n floats) is added
 * to the particle forces on every step before the force evaluation, like
 * SpringNetwork::setForce from Interactor::syncSystemStateData (Interactor.cpp:23-31).
 * Returns BS_ERR_NONFINITE if a position became non-finite; the offending
 * particle is then reported by bs_last_nonfinite_particle. */
int bs_step(bs_handle *h, int32_t nsteps, const float *ext_force3);

/* SpringNetwork::computeForces (SpringNetwork.cpp:244-249) plus the force
 * part of computeParticleForces, without integrating: writes the total force
 * on every particle (3n floats, nullable; static particles get 0) and the
 * energies (BS_NUM_ENERGIES doubles, nullable; kinetic slot = 0). */
int bs_compute_forces(bs_handle *h, const float *ext_force3, float *force3, double *energies);

/* ---- readback -------------------------------------------------------------- */

/* Positions and velocities (3n floats each, either nullable).  Replaces
 * SpringNetwork::getParticlePosition (SpringNetwork.cpp:430-436) for all
 * particles at once.  Synchronising. */
int bs_get_state(bs_handle *h, float *x3, float *v3);

/* Overwrites positions/velocities of all particles (nullable each). */
int bs_set_state(bs_handle *h, const float *x3, const float *v3);

/* Energies of the last step (SpringNetwork::getSpringEnergy ... getters,
 * SpringNetwork.h:157-162), accumulated in FP64. */
int bs_get_energies(bs_handle *h, double *energies);

/* Index of the first particle (lowest id) found with a non-finite position by
 * the last bs_step, or -1. */
int64_t bs_last_nonfinite_particle(const bs_handle *h);

/* Counters: [0] steps done, [1] neighbour-list rebuilds, [2] max list length,
 * [3] list capacity, [4] directed spring edges, [5] pairs in the last list. */
int bs_get_counters(const bs_handle *h, int64_t *out6);

/* ---- profiling --------------------------------------------------------------- */

/* Kernel classes timed with HIP events on the handle's stream. */
#define BS_PROF_REBUILD 0  /* neighbour-upkeep chain (box, binning, list build) */
#define BS_PROF_PAIR    1  /* non-bonded pair kernels */
#define BS_PROF_SPRING  2  /* spring gather + integrator (or force epilogue) */
#define BS_PROF_ADJ_PAIR   3  /* bs_adjoint: pair part of H w and dL/dq */
#define BS_PROF_ADJ_SPRING 4  /* bs_adjoint: spring part, parameter gradients, reverse update */
#define BS_PROF_TOPO_COUNT 5  /* bs_springs_from_cutoff: binning and the count pass */
#define BS_PROF_TOPO_FILL  6  /* bs_springs_from_cutoff: the fill pass (rows written and sorted) */
#define BS_PROF_NUM     7

/* Turns per-kernel-class event timing of bs_step and bs_adjoint on (1) or off
 * (0) and resets the totals.  Counterpart of the reference's wall-clock timeit::Profiler
 * (src/timeit.hpp:16-168; "Average framerate", SpringNetwork.cpp:305-308). */
int bs_profile(bs_handle *h, int32_t enable);

/* Accumulated milliseconds and launch counts per BS_PROF_* class. */
int bs_get_profile(const bs_handle *h, double *ms, int64_t *launches);

/* ---- parity hooks ------------------------------------------------------------- */

/* Rebuilds the reference cell list of search `kind` (BS_NS_*) on the GPU at the
 * current positions, with the reference definitions: bounding box over all
 * particles in FP64 (box.hpp:37-69), edge = cutoff + neighborskin,
 * n_a = size_t(ceil(L_a/edge) + 1), cell = size_t((x - min)/edge) clamped
 * (nsearch.hpp:277-293,342-390), id = cx + cy*nx + cz*nx*ny.
 * Then copies out (each nullable):
 *   cell_of_particle[n]  cell id per particle, -1 when not in this search;
 *   cell_start[ncells+1] CSR offsets into sorted_idx;
 *   sorted_idx[members]  particle ids per cell, ascending as in the reference;
 *   dims[3], box[6] (min xyz, max xyz).
 * *ncells and *nmembers receive the sizes; call first with NULL arrays. */
int bs_get_cells(bs_handle *h, int32_t kind, int64_t *ncells, int64_t *nmembers,
                 int64_t *cell_of_particle, int32_t *cell_start, int32_t *sorted_idx,
                 int64_t dims[3], double box[6]);

/* Processed non-bonded pairs of search `kind` at the current positions: the
 * pairs the reference's Particle::add{Steric,Electrostatic,Hydrophobicity}Force
 * (Particle.cpp:191-330,368-429) evaluates -- i dynamic, j static or j > i,
 * not spring-bonded when springs are enabled, FP64 distance < cutoff
 * (nsearch.hpp:230) and FP32 distance < cutoff (Particle.cpp:400-402), plus
 * distance != 0 for electrostatic/hydrophobic.  Written as (min,max) id pairs
 * sorted ascending.  *count receives the number; ij may be NULL to query it. */
int bs_get_pairs(bs_handle *h, int32_t kind, uint32_t *ij, int64_t cap, int64_t *count);

/* ---- the probe particle ------------------------------------------------------ */

/* The probe particle (probe_enable): SpringNetwork::getProbeParticle
 * (SpringNetwork.h:277).  Position, velocity and the force of the last
 * evaluation (its pair forces and viscosity, SpringNetwork.cpp:158-177),
 * 3 floats each, every pointer nullable.  The first setup after
 * bs_set_particles places the probe at probe_x/y/z at rest
 * (SpringNetwork::_setupProbe, SpringNetwork.cpp:749-766); every bs_step step
 * then moves it once, after the particles' forces are taken
 * (SpringNetwork.cpp:156-182).  bs_compute_forces includes the probe's pair
 * forces on the particles and its pair energies, and leaves the probe in
 * place (the reference's computeForces also moves it).  Not with a batch of
 * systems (BS_ERR_ARG at setup) nor in the adjoint (BS_ERR_ARG).
 * BS_ERR_STATE when the probe is not enabled. */
int bs_get_probe(bs_handle *h, float *x3, float *v3, float *f3);

/* ---- rigid bodies ------------------------------------------------------------- */

/* Makes particles ids[0..n) one rigid body: RigidBodiesManager::InitRigidBodies
 * (RigidBodiesManager.cpp:11-43).  As there, nothing happens when one of the
 * ids already belongs to a body, and n = 0 removes every body.  The body's
 * mass, barycentre and body-frame offsets are taken from the positions at the
 * next bs_setup / step (RigidBody::init, RigidBody.cpp:57-124; isotropic
 * inertia 2/5 M).  Springs between two rigid particles are no longer evaluated
 * (RigidBody.cpp:42-53).  Each step with rigidbody_enable sums the force and
 * torque of the body's dynamic particles (RigidBody.cpp:567-580), advances the
 * body (RigidBody.cpp:430-484) and places its particles
 * (RigidBody.cpp:600-635); with rigidbody_enable = 0 the particles are only
 * placed.  The reference places by the particle's index in the dynamic list
 * (SpringNetwork.cpp:216-218), which is the body index for the single body
 * SpringNetwork::run creates; this engine uses the index in the body, which
 * is also right for several bodies. */
int bs_add_rigid_body(bs_handle *h, int64_t n, const int64_t *ids);

/* Number of rigid bodies. */
int32_t bs_num_rigid_bodies(const bs_handle *h);

/* State of body b: position X, velocity, angular momentum L, angular velocity
 * omega (3 each) and orientation quaternion (x, y, z, w) -> out16
 * (RigidBody.h:156-183). */
int bs_get_rigid_body(bs_handle *h, int32_t b, double *out16);

/* ---- field grids -------------------------------------------------------------- */

/* Sets grid `kind` (BS_GRID_*) from its scalar values, shape[0]*shape[1]*shape[2]
 * floats with z fastest, as PotentialGrid holds them after opendx::readGrid
 * (OpenDXReader.cpp:146-173); the force vector of each cell is the central
 * difference gradient times -k_B (PotentialGrid::compute_gradient,
 * PotentialGrid.cpp:20-60).  Each step, for the potential grid (with
 * electrostatic_enable and potentialgrid_enable) F += vector * q * scale and
 * E_elec += scalar * q * k_B * N_A / 1000 * scale (Particle.cpp:161-189); for
 * the density grid (densitygrid_enable) F += vector * burying * scale
 * (Particle.cpp:136-159).  The cell is the nearest lower one, no
 * interpolation; a particle off the grid gets nothing
 * (GridCoordinatesSystem.hpp:85-93,206-218).  scalar = NULL removes the grid. */
int bs_set_grid(bs_handle *h, int32_t kind, const double origin[3], const int64_t shape[3],
                const double cell[3], const float *scalar);

/* Per-particle burying factor of the density field (ParticleProperty::_burying,
 * ParticleProperty.h:60-61; default 1).  NULL restores the default. */
int bs_set_burying(bs_handle *h, const float *burying);

/* ---- interactive updates -------------------------------------------------------- */

/* Replaces InteractorFreeSASA::syncParticleStateData
 * (src/interactor/freesasa/InteractorFreeSASA.cpp:229-247): new solvent
 * accessible surfaces for all particles (n floats), used by IMPALA from the
 * next step on, without a re-setup.  External forces from an interactor
 * (InteractorMDDriver::syncParticleStateData, InteractorMDDriver.cpp:309-321)
 * are the ext_force3 argument of bs_step. */
int bs_set_sasa(bs_handle *h, const float *sasa);

/* Replaces the IMPALA custom data of the MDDriver bridge
 * (src/interactor/mddriver/CustomData.cpp:61-108: "dmou", "dmol", "dmtc"):
 * new double-membrane offsets and tube curvatures from the next step on,
 * without a re-setup; a NaN keeps that value; ignored while IMPALA is off,
 * as there. */
int bs_set_membrane(bs_handle *h, float upper_offset, float lower_offset, float upper_curvature,
                    float lower_curvature);

/* ---- topology files ------------------------------------------------------------ */

/* Replaces io::readTopology of a .nc file (NetCDFReader::read,
 * src/IO/NetCDFReader.cpp:8-88) followed by Topology::to_spring_network
 * (Topology.hpp:245-296): bs_set_particles and bs_set_springs from the
 * file's particle_number / spring_number dimensions and its variables
 * particleids, resids, resnames[4], particlenames[5], chainnames[4],
 * dynamicstate (0 = static), coordinates[3], radii, mass (0 -> 1), epsilon,
 * charges, optional surfaceaccessibility, hydrophobicityscale (the IMPALA
 * transfer energy) and hydrophobicity, and springs[2], springsstiffness,
 * springsequilibrium; names, residues and chains become the PDB labels.
 * Values of any numeric type are converted as the netCDF library's getVar
 * does.  Classic, 64-bit-offset and 64-bit-data files are read by the
 * library itself; NetCDF-4 (HDF5) files -> BS_ERR_ARG.  A missing array,
 * wrong shape, missing particle dimension, spring index out of range, self
 * spring or repeated spring -> BS_ERR_ARG with the reference's message.
 * warnings (nullable, warnings_len bytes): the reader's warning lines. */
int bs_load_netcdf(bs_handle *h, const char *path, char *warnings, int64_t warnings_len);

/* Replaces NetCDFWriter::writeBinary (src/IO/NetCDFWriter.cpp:25-150) of
 * the loaded network: the reference's dimensions, variables and attributes
 * (coordinates at the current positions, particleids = the particle index,
 * the properties, dynamicstate, the PDB labels cut to 5/4/4 characters,
 * springs, springsstiffness, springsequilibrium, nbspringsperparticle),
 * written in the 64-bit-offset classic layout, which bs_load_netcdf and the
 * netCDF library read.  A file that cannot be written -> BS_ERR_ARG. */
int bs_save_netcdf(bs_handle *h, const char *path);

/* The per-particle properties as loaded (bs_set_particles order and
 * defaults); every pointer nullable, bs_num_particles entries each. */
int bs_get_particle_properties(const bs_handle *h, float *mass, float *charge, float *radius, float *epsilon,
                               float *sasa, float *transfer, float *hydrophobicity, uint8_t *flags);

/* ---- trajectory output ------------------------------------------------------- */

#define BS_TRAJ_PDB 0  /* io::modern::PDBTrajectoryWriter (PDBTrajectoryWriter.cpp:10-33) */
#define BS_TRAJ_XTC 1  /* io::modern::XTCTrajectoryWriter (XTCTrajectoryWriter.cpp:13-31) */
#define BS_TRAJ_CSV 2  /* io::modern::CSVTrajectoryWriter (CSVTrajectoryWriter.cpp:11-60) */

/* Replaces SpringNetwork::_setupTrajectories (SpringNetwork.cpp:768-778) for
 * one writer: the pdbtrajectory / xtctrajectory / csvsampling settings
 * (Setting.hpp:160-170) as kind, path and frequency (>= 1; the reference's
 * default is 100).  bs_step then writes a frame at the start of every step
 * whose iteration count (steps since bs_setup) the frequency divides, as
 * TrajectoryManager::write_step from SpringNetwork::idleRun
 * (TrajectoryManager.cpp:9-16, SpringNetwork.cpp:344-362): the positions at
 * that point and the energies of the step before.  The positions are packed
 * on the GPU and copied into pinned host slots; a writer thread formats and
 * writes them while the stream goes on stepping, and graph replays go on
 * between frames.  PDB: MODEL n / ATOM records (pdb_format.hpp:17-50) /
 * CONECT records of the springs in the first model / ENDMDL.  XTC: positions
 * in nm, step = time = the writer's frame count, unit box, precision 1000.
 * CSV: header at once, then one row per frame; the energy columns follow the
 * enabled terms of the current parameters, the insertion columns
 * bs_set_insertion_vector.  Labels, springs, parameters and the insertion
 * vector are taken as they are when the writer is added.  A file that cannot
 * be opened -> BS_ERR_ARG.  bs_set_particles closes every writer. */
int bs_add_trajectory(bs_handle *h, int32_t kind, const char *path, int64_t frequency);

/* Replaces SpringNetwork::writeNextStepNow (SpringNetwork.cpp:408): a frame
 * of the current state to every writer, whatever its frequency. */
int bs_write_frame_now(bs_handle *h);

/* Waits until every frame taken so far is written; BS_ERR_ARG if a write
 * failed. */
int bs_flush_trajectories(bs_handle *h);

/* Flushes, closes the files and removes the writers (also done by bs_destroy). */
int bs_close_trajectories(bs_handle *h);

/* The PDB fields of spn::Particle (Particle.h:36-41: name, residue name,
 * chain, residue id; ParticleProperty.h:20-21: occupancy, temperature factor,
 * electron charge) and the element name, n entries each; any pointer may be
 * NULL for the reference's defaults ("" or 0; the electron charge defaults to
 * (int)charge, ParticleProperty.cpp:8-16).  Used by PDB writers added
 * afterwards. */
int bs_set_particle_labels(bs_handle *h, const char *const *name, const char *const *resname,
                           const char *const *chain, const int32_t *resid, const float *occupancy,
                           const float *tempfactor, const char *const *element, const int32_t *electron_charge);

/* Replaces SpringNetwork::setInsertionVector (the ivector setting): CSV
 * writers added afterwards carry the insertion angle and depth of particles
 * i1, i2 (InsertionVector.cpp:18-40; depth = the centroid's z).  Both < 0
 * turns the columns off.  Ids out of range -> BS_ERR_ARG. */
int bs_set_insertion_vector(bs_handle *h, int64_t i1, int64_t i2);

/* ---- insertion vector --------------------------------------------------------- */

/* InsertionVector::computeVector / computeAngle / computeRollAngle
 * (InsertionVector.cpp:18-78) for particles i1, i2 at the current positions,
 * with the centroid of all particles (measure.hpp:121-140) reduced on the GPU:
 * out6 = angle (deg), roll angle (deg), insertion depth (centroid z),
 * vector p2 - p1 (3).  Replaces SpringNetwork::_updateInsertionVector
 * (SpringNetwork.cpp:52-59), which the reference runs every step; here it
 * runs when asked for. */
int bs_get_insertion_vector(bs_handle *h, int64_t i1, int64_t i2, double *out6);

/* ---- batches of independent systems ------------------------------------------ */

/* Declares the particles as nsys independent systems: system s is particles
 * [offsets[s], offsets[s+1]) (offsets[0] = 0, offsets[nsys] = n).  Springs and
 * rigid bodies must stay inside one system.  Systems never interact; each is
 * stepped exactly as alone (the parameters are shared), in one launch
 * sequence over all of them.  Per-system energies: bs_get_system_energies.
 * No reference counterpart (the reference runs one network per process);
 * serves many small networks on one GPU. */
int bs_set_systems(bs_handle *h, int32_t nsys, const int64_t *offsets);

/* Number of systems (1 unless bs_set_systems declared more). */
int32_t bs_num_systems(const bs_handle *h);

/* The six energies of every system after the last bs_step / bs_compute_forces,
 * out[6 * s + BS_E_*]. */
int bs_get_system_energies(bs_handle *h, double *out);

/* ---- adjoint ------------------------------------------------------------------- */

/* Gradients of the trajectory loss
 *   L = (1/n) sum_i |x_T,i - y_i|^2
 * after nsteps steps from the current state (which advances, as bs_step) with
 * respect to each spring's stiffness and rest length (dk, dl0: one value per
 * spring kept by bs_set_springs, in its order) and each particle's charge
 * (dq[n]).  target3 = y (3n floats, required); ext_force3 as in bs_step; every
 * output is nullable.  The reverse sweep
 *   a = lv + dt lx,  w = dt/m a,  lx += H(x_s) w,  lv = (1 - gamma dt/m) a
 * runs in FP64 over the FP32 trajectory, H being the Jacobian of the forces
 * (springs, the four steric modes, Coulomb with both dielectric models;
 * viscosity enters through gamma), with the pair set of each state fixed as
 * the forward step found it.  States are saved every checkpoint_every steps
 * and segments recomputed, so the work is about two forward passes plus the
 * reverse sweep and the memory n * (32 bytes per checkpoint + 16 bytes per
 * step of a segment).  checkpoint_every <= 0 chooses: the whole trajectory as
 * one segment, with no recomputation (one forward pass plus the sweep), when
 * n * (32 + 16 nsteps) bytes fit in half the free device memory, else
 * ceil(sqrt(nsteps)).  BS_ERR_ARG with IMPALA, the
 * pseudo-hydrophobic term, rigid bodies, field grids or compensated
 * positions.  No reference counterpart (the reference has no gradients). */
int bs_adjoint(bs_handle *h, int32_t nsteps, int32_t checkpoint_every, const float *target3,
               const float *ext_force3, double *loss, double *dk, double *dl0, double *dq);

/* The adjoint in two calls, for any loss of the final state.
 * bs_adjoint_forward runs nsteps steps from the current state (which
 * advances, as bs_step) and records them: checkpoints every checkpoint_every
 * steps (<= 0: ceil(sqrt(nsteps))) and the last segment.  The caller reads
 * x_T, v_T (bs_get_state) and forms dL/dx_T, dL/dv_T.
 * bs_adjoint_backward(dl_dx3, dl_dv3) (3n doubles each, either nullable, not
 * both) runs the reverse sweep of bs_adjoint from that seed and writes dk,
 * dl0, dq (nullable) as bs_adjoint does.  The state is back at x_T after it
 * and the record stays valid, so several seeds can be swept; any call that
 * changes the state, the parameters, the external force or the launch
 * configuration drops it (BS_ERR_STATE on a later backward).
 * No reference counterpart. */
int bs_adjoint_forward(bs_handle *h, int32_t nsteps, int32_t checkpoint_every, const float *ext_force3);
int bs_adjoint_backward(bs_handle *h, const double *dl_dx3, const double *dl_dv3, double *dk, double *dl0,
                        double *dq);

/* bs_adjoint_backward that allocates nothing: the same arguments, results
 * and errors, with the sweep's device buffers (about 192 bytes per particle
 * plus 40 per spring when dk or dl0 is wanted) taken from a fixed 64 MiB
 * array that the library holds on each GPU until the process ends.  Larger
 * systems are refused with BS_ERR_ARG and a message giving the size needed.
 * Calls on one GPU share that array: calls that overlap on one GPU race.
 * BS_ERR_STATE while profiling is on (bs_profile creates events).
 * No counterpart in the reference; it is for callers that cannot allocate
 * device memory (fitting loops under a no-allocation rule, serving loops). */
int bs_adjoint_backward_noalloc(bs_handle *h, const double *dl_dx3, const double *dl_dv3, double *dk,
                                double *dl0, double *dq);

/* ---- tuning ---------------------------------------------------------------------- */

/* Times `steps` steps of a copy of the current system for each launch
 * configuration and keeps the fastest for h; h's state is not advanced.
 * First the pair layout (BS_PAIR_CLUSTERS or BS_PAIR_QUEUES; a handle set to
 * BS_PAIR_ROWS keeps it), then lanes per spring row (4/8/16/32) x direct
 * launches or a 16-step graph.  The chosen values go to *group_out,
 * *graph_out, *layout_out (nullable).  No reference counterpart. */
int bs_autotune(bs_handle *h, int32_t steps, int32_t *group_out, int32_t *graph_out, int32_t *layout_out);

#ifdef __cplusplus
}
#endif

#endif /* BIOSPRING_AMD_BS_API_H */
