// Default parameters: the reference's defaultConfiguration()
// (src/configuration/Configuration.hpp:153-209) plus the engine options.
#ifndef BIOSPRING_AMD_PARAMS_DEFAULTS_H
#define BIOSPRING_AMD_PARAMS_DEFAULTS_H

#include <string.h>

#include "../../include/bs_api.h"

static inline void bs_params_fill_defaults(bs_params *p)
{
    memset(p, 0, sizeof *p);
    p->timestep = 0.01f;          /* sim.timestep */
    p->neighborskin = 0.0f;       /* sim.neighborskin */
    p->spring_enable = 0;
    p->spring_scale = 1.0f;
    p->steric_enable = 0;
    p->steric_mode = BS_STERIC_LINEAR;
    p->steric_cutoff = 1.0f;
    p->steric_gridscale = 1.0f;
    p->electrostatic_enable = 0;
    p->electrostatic_cutoff = 16.0f;
    p->electrostatic_scale = 1.0f;
    p->electrostatic_dielectric = 1.0f;
    p->dielectric_model = BS_DIELECTRIC_CONSTANT;
    p->viscosity_enable = 0;
    p->viscosity_value = 1.0f;
    p->imp_enable = 0;
    p->imp_scale = 1.0f;
    p->hydrophobicity_enable = 0;
    p->hydrophobicity_cutoff = 15.0f;
    p->hydrophobicity_scale = 1.0f;
    p->list_skin = 1.0f;
    p->rebuild_interval = 0;
    p->compensated_positions = 0;
    p->rigidbody_enable = 0;      /* rigidbody.enable */
    p->potentialgrid_enable = 0;  /* GridSetting defaults, Setting.hpp:192-202 */
    p->potentialgrid_scale = 1.0f;
    p->densitygrid_enable = 0;
    p->densitygrid_scale = 1.0f;
    p->graph_steps = 0;
    p->pair_layout = BS_PAIR_QUEUES;
    p->probe_enable = 0;          /* probe.*, Configuration.hpp:186-195 */
    p->probe_enable_electrostatic = 0;
    p->probe_enable_steric = 0;
    p->probe_x = 1.0f;
    p->probe_y = 1.0f;
    p->probe_z = 1.0f;
    p->probe_mass = 1.0f;
    p->probe_epsilon = 1.0f;
    p->probe_radius = 1.0f;
    p->probe_charge = 0.0f;
}

#endif
