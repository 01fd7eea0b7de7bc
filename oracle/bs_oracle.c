/*
 * bs_oracle.c -- CPU restatement of BioSpring's per-step force-and-integrate
 * path.  TEST INFRASTRUCTURE ONLY: the checker for the HIP engine.  Only the
 * tests, the smoke check and the benchmark's CPU-baseline leg load it; the
 * engine never calls into it.
 *
 * Two arithmetic modes share one enumeration of particles, springs and pairs:
 *
 *   ORC_F32  float-faithful: every float/double rounding step, loop order and
 *            serial reduction order of the reference CPU path, so that it can
 *            be compared bit-for-bit with the reference itself (oracle/_ref).
 *   ORC_F64  the same algorithm in double precision: the FP64 oracle the GPU
 *            forces and energies are held to (1e-5 relative).
 *
 * Pair decisions always follow the reference's two-stage test on float
 * positions (nsearch.hpp:230 then Particle.cpp:400-402), in both modes, so the
 * FP64 oracle evaluates exactly the reference's pair set.
 *
 * Reference files followed (paths relative to the BioSpring source root):
 *   src/spn/SpringNetwork.cpp:63-241      step order, serial reductions
 *   src/spn/Spring.cpp:20-39              spring force
 *   src/spn/Particle.cpp:70-128,191-330,368-429  pair loops, viscosity, Euler
 *   src/forcefield/ForceField.cpp:15-106  scales, IMPALA sign
 *   src/forcefield/energy/{spring,steric,electrostatic,hydrophobic,imp}.hpp energy / force-module formulas
 *   src/forcefield/constants.hpp:36-118   unit constants
 *   src/forcefield/CombinationRules.hpp:18-31
 *   src/nsearch.hpp:155-436, src/box.hpp:37-69, src/measure.hpp:163-196
 *   src/Vector3f.h:45-144, src/Vector3f.cpp:9-11
 *
 * Parity status of each part is pinned by tests/test_oracle_*.py (golden
 * vectors from the compiled reference, closed forms, finite differences,
 * brute-force pair sets).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/bs_api.h"

#define ORC_F32 0
#define ORC_F64 1

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ---------------------------------------------------------------------------
 * Unit constants, evaluated in double in the same order as constants.hpp.
 * ------------------------------------------------------------------------- */
static const double K_KJ_TO_J = 1.0E3;
static const double K_J_TO_KJ = 1.0E-3;
static const double K_E_TO_C = 1.60217653E-19;
static const double K_A_TO_M = 1.0E-10;
static const double K_M_TO_A = 1.0E10;
static const double K_DA_TO_KG = 1.66054E-27;
static const double K_S_TO_FS = 1.0E15;
static const double K_FS_TO_S = 1.0E-15;
static const double K_PI = 3.1415927;
static const double K_EPS0 = 8.854187817E-12;
static const double K_NA = 6.02214169E23;
static const double K_MIN_ELEC = 0.001;
static const double K_MIN_VDW = 0.1;
static const float K_IMP_ALPHA = 1.99f;
static const float K_IMP_Z0 = 15.75f;
static const float K_IMP_ALIP = -0.018f;

static double c_coulomb(void) { return 1.0 / (4.0 * K_PI * K_EPS0); }
static double c_newton(void) { return K_M_TO_A * (1.0 / K_DA_TO_KG) * K_FS_TO_S * K_FS_TO_S; }
static double c_spring(void) { return K_KJ_TO_J / K_NA * K_M_TO_A * c_newton(); }
static double c_elec(void) { return c_newton() * (K_E_TO_C * K_E_TO_C) / (K_EPS0 * K_A_TO_M * K_A_TO_M); }
static double c_kinetic(void)
{
    return (K_A_TO_M * K_S_TO_FS) * (K_A_TO_M * K_S_TO_FS) * K_NA * K_DA_TO_KG * K_J_TO_KJ;
}

/* exported for tests: [spring, elec, kinetic, coulomb] */
void orc_constants(double out[4])
{
    out[0] = c_spring();
    out[1] = c_elec();
    out[2] = c_kinetic();
    out[3] = c_coulomb();
}

/* ---------------------------------------------------------------------------
 * Geometry primitives with the reference rounding (Vector3f.cpp:9-11 ->
 * measure.hpp:188-196; Vector3f.h:45-60).
 * ------------------------------------------------------------------------- */
static inline float norm_f(float x, float y, float z)
{
    const double dx = x, dy = y, dz = z;
    if (fabs(dx) < 1e-40 && fabs(dy) < 1e-40 && fabs(dz) < 1e-40)
        return 0.0f;
    return (float)sqrt(dx * dx + dy * dy + dz * dz);
}

static inline void normalize_f(float v[3])
{
    const float nrm = norm_f(v[0], v[1], v[2]);
    if ((double)nrm > 1.0E-40) {
        v[0] /= nrm;
        v[1] /= nrm;
        v[2] /= nrm;
    } else {
        v[0] = v[1] = v[2] = 0.0f;
    }
}

static inline double norm_d(double x, double y, double z)
{
    if (fabs(x) < 1e-40 && fabs(y) < 1e-40 && fabs(z) < 1e-40)
        return 0.0;
    return sqrt(x * x + y * y + z * z);
}

static inline void normalize_d(double v[3])
{
    const double nrm = norm_d(v[0], v[1], v[2]);
    if (nrm > 1.0E-40) {
        v[0] /= nrm;
        v[1] /= nrm;
        v[2] /= nrm;
    } else {
        v[0] = v[1] = v[2] = 0.0;
    }
}

/* ---------------------------------------------------------------------------
 * Force-field formulas, float-faithful (energy/{spring,steric,electrostatic,hydrophobic,imp}.hpp).  The steric functions
 * are compiled in the reference TU that sees `using namespace std`
 * (reduce/ParticleGroup.h:9 before forcefield headers in SpringNetwork.cpp),
 * so their pow(float, float) is powf; ForceField.cpp's IMPALA and
 * hydrophobic formulas call the double exp/pow.
 * ------------------------------------------------------------------------- */
static inline float spring_energy_f(float d, float k, float l0)
{
    const float dv = d - l0; /* spring.hpp:17-21 */
    return (float)(0.5 * k * dv * dv);
}

static inline float spring_fm_f(float d, float k, float l0)
{
    float fm = k * (d - l0); /* spring.hpp:24-29 */
    fm = (float)(fm * c_spring());
    return fm;
}

static float steric_energy_f(int mode, float ri, float rj, float ei, float ej, float d)
{
    switch (mode) {
    case BS_STERIC_AMBER_12_6: { /* steric.hpp:60-72 */
        if (d < K_MIN_VDW)
            return 0.0f;
        const float eij = (float)sqrt((double)(ei * ej));
        const float rij = (float)sqrt((double)(ri * rj));
        const float rep = eij * powf(rij / d, 12.0f);
        const float att = (float)(-eij * 2.0 * powf(rij / d, 6.0f));
        return rep + att;
    }
    case BS_STERIC_LEWITT_8_6: { /* steric.hpp:94-106 */
        if (d < K_MIN_VDW)
            return 0.0f;
        const float eij = (float)sqrt((double)(ei * ej));
        const float rij = (float)sqrt((double)(ri * rj));
        const float rep = (float)(eij * 3.0 * powf(rij / d, 8.0f));
        const float att = (float)(-eij * 4.0 * powf(rij / d, 6.0f));
        return rep + att;
    }
    case BS_STERIC_ZACHARIAS_8_6: { /* steric.hpp:129-141 */
        if (d < K_MIN_VDW)
            return 0.0f;
        const float eij = ei * ej;
        const float rij = ri * rj;
        const float rep = eij * powf(rij / d, 8.0f);
        const float att = -eij * powf(rij / d, 6.0f);
        return rep + att;
    }
    default: { /* linear, steric.hpp:27-36 */
        const float eq = ri + rj;
        const float dv = d - eq;
        if (dv > 0)
            return 0.0f;
        return (float)(0.5 * 1.0f * dv * dv);
    }
    }
}

static float steric_fm_f(int mode, float ri, float rj, float ei, float ej, float d)
{
    switch (mode) {
    case BS_STERIC_AMBER_12_6: { /* steric.hpp:74-87 */
        if (d < K_MIN_VDW)
            return 0.0f;
        const float eij = (float)sqrt((double)(ei * ej));
        const float rij = (float)sqrt((double)(ri * rj));
        const float rep = (float)(-eij * 12.0 * (powf(rij, 12.0f) / powf(d, 13.0f)));
        const float att = (float)(eij * 2.0 * 6.0 * (powf(rij, 6.0f) / powf(d, 7.0f)));
        const float fm = rep + att;
        return (float)(fm * c_spring());
    }
    case BS_STERIC_LEWITT_8_6: { /* steric.hpp:108-122 */
        if (d < K_MIN_VDW)
            return 0.0f;
        const float eij = (float)sqrt((double)(ei * ej));
        const float rij = (float)sqrt((double)(ri * rj));
        const float rep = (float)(-eij * 3.0 * 8.0 * (powf(rij, 8.0f) / powf(d, 9.0f)));
        const float att = (float)(eij * 4.0 * 6.0 * (powf(rij, 6.0f) / powf(d, 7.0f)));
        const float fm = rep + att;
        return (float)(fm * c_spring());
    }
    case BS_STERIC_ZACHARIAS_8_6: { /* steric.hpp:143-157 */
        if (d < K_MIN_VDW)
            return 0.0f;
        const float eij = ei * ej;
        const float rij = ri * rj;
        const float rep = (float)(-eij * 8.0 * (powf(rij, 8.0f) / powf(d, 9.0f)));
        const float att = (float)(eij * 6.0 * (powf(rij, 6.0f) / powf(d, 7.0f)));
        const float fm = rep + att;
        return (float)(fm * c_spring());
    }
    default: { /* linear, steric.hpp:39-49 */
        const float eq = ri + rj;
        const float dv = d - eq;
        if (dv > 0)
            return 0.0f;
        const float fm = -1.0f * fabsf(dv);
        return (float)(fm * c_spring());
    }
    }
}

/* Coulomb, electrostatic.hpp:16-49.  With the distance-dependent dielectric
 * extension the relative permittivity becomes dielectric * d (d in A): the
 * energy uses it in place of dielectric, and the force module, dE/dd, gains a
 * factor 2. */
static inline float elec_energy_f(float q1, float q2, float d, float diel, int ddd)
{
    if (d < K_MIN_ELEC)
        return 0.0f;
    const double k = c_coulomb();
    const double c1 = q1 * K_E_TO_C;
    const double c2 = q2 * K_E_TO_C;
    const float eps = ddd ? diel * d : diel;
    const float dm = (float)(d * K_A_TO_M);
    double e = k * (c1 * c2) / (eps * dm);
    e = e * K_NA;
    e = e * 1.0E-3;
    return (float)e;
}

static inline float elec_fm_f(float q1, float q2, float d, float diel, int ddd)
{
    if (d < K_MIN_ELEC)
        return 0.0f;
    const float eps = ddd ? diel * d : diel;
    float fm = (float)(-(q1 * q2) / (4.0 * K_PI * eps * d * d));
    fm = (float)(fm * c_elec());
    if (ddd)
        fm = 2.0f * fm;
    return fm;
}

/* hydrophobic.hpp:20-39 (energy carries N_A*1e-3, the force does not). */
static inline float hydro_energy_f(float h1, float h2, float d)
{
    double e = -(h1 * h2) * exp((double)-d);
    e = e * K_NA;
    e = e * 1.0E-3;
    return (float)e;
}

static inline float hydro_fm_f(float h1, float h2, float d)
{
    float fm = (float)((h1 * h2) * exp((double)-d));
    fm = (float)(fm * c_spring());
    return fm;
}

/* IMPALA, imp.hpp:29-159, float-faithful. */
typedef struct {
    float up_off, lo_off, up_curv, lo_curv;
} imp_geom;

static inline int sgn_f(float c) { return (c > 0.0f) - (c < 0.0f); }

static void imp_zprime_f(const imp_geom *g, float x, float y, float z, float *z_up, float *z_lo,
                         float vup[3], float vlo[3], int *s_up, int *s_lo, float *r_up, float *r_lo)
{
    const float rup = g->up_curv == 0.0f ? 1000000.0f : fabsf(1 / g->up_curv);
    const float rlo = g->lo_curv == 0.0f ? 1000000.0f : fabsf(1 / g->lo_curv);
    const int sup = sgn_f(g->up_curv);
    const int slo = sgn_f(g->lo_curv);
    const float cup = g->up_off - sup * rup;
    const float clo = -g->lo_off - slo * rlo;
    vup[0] = x - 0.0f;
    vup[1] = y - y;
    vup[2] = z - cup;
    vlo[0] = x - 0.0f;
    vlo[1] = y - y;
    vlo[2] = z - clo;
    const float nup = norm_f(vup[0], vup[1], vup[2]);
    const float nlo = norm_f(vlo[0], vlo[1], vlo[2]);
    *z_up = sup == 0 ? z : (z > cup ? sup * nup + g->up_off - rup : -sup * nup + g->up_off - rup);
    *z_lo = slo == 0 ? z : (z > clo ? slo * nlo - g->lo_off - rlo : -slo * nlo - g->lo_off - rlo);
    *s_up = sup;
    *s_lo = slo;
    *r_up = rup;
    *r_lo = rlo;
}

static inline int imp_is_flat(const imp_geom *g)
{
    return g->up_off == 0.0f && g->lo_off == 0.0f && g->up_curv == 0.0f && g->lo_curv == 0.0f;
}

static float imp_energy_f(const imp_geom *g, float x, float y, float z, float sasa, float tr)
{
    float zu, zl, vu[3], vl[3], ru, rl;
    int su, sl;
    imp_zprime_f(g, x, y, z, &zu, &zl, vu, vl, &su, &sl, &ru, &rl);
    const double cz_up = 0.5 - 1.0 / (1.0 + exp((double)(K_IMP_ALPHA * (fabsf(zu - g->up_off) - K_IMP_Z0))));
    const double cz_lo = 0.5 - 1.0 / (1.0 + exp((double)(K_IMP_ALPHA * (fabsf(zl + g->lo_off) - K_IMP_Z0))));
    const double hu = -sasa * tr * cz_up;
    const double hl = -sasa * tr * cz_lo;
    const double lu = K_IMP_ALIP * sasa * cz_up;
    const double ll = K_IMP_ALIP * sasa * cz_lo;
    if (imp_is_flat(g))
        return (float)(hu + lu);
    return (float)(hu + lu + hl + ll);
}

static inline double imp_expo_f(float z, float off) { return exp((double)(K_IMP_ALPHA * (fabsf(z + off) - K_IMP_Z0))); }

/* Returns the vector imp_force_vector() (before the ForceField sign/scale). */
static void imp_force_f(const imp_geom *g, float x, float y, float z, float sasa, float tr, float out[3])
{
    float zu, zl, vu[3], vl[3], ru, rl;
    int su, sl;
    imp_zprime_f(g, x, y, z, &zu, &zl, vu, vl, &su, &sl, &ru, &rl);
    double dcu = (K_IMP_ALPHA * (zu - g->up_off) * imp_expo_f(zu, -g->up_off)) /
                 (pow(imp_expo_f(zu, -g->up_off) + 1, 2.0) * fabsf(zu - g->up_off));
    double dcl = (K_IMP_ALPHA * (zl + g->lo_off) * imp_expo_f(zl, g->lo_off)) /
                 (pow(imp_expo_f(zl, g->lo_off) + 1, 2.0) * fabsf(zl + g->lo_off));
    if (isnan(dcu) || !isfinite(dcu))
        dcu = 0.0;
    if (isnan(dcl) || !isfinite(dcl))
        dcl = 0.0;
    float du[3] = {0.0f, 0.0f, 1.0f}, dl[3] = {0.0f, 0.0f, 1.0f};
    if (su != 0)
        memcpy(du, vu, sizeof du);
    if (sl != 0)
        memcpy(dl, vl, sizeof dl);
    normalize_f(du);
    normalize_f(dl);
    const double hu = -sasa * tr * dcu;
    const double hl = -sasa * tr * dcl;
    const double lu = K_IMP_ALIP * sasa * dcu;
    const double ll = K_IMP_ALIP * sasa * dcl;
    const float mu = (float)((hu + lu) * c_spring());
    const float ml = (float)((hl + ll) * c_spring());
    const float fu[3] = {du[0] * mu, du[1] * mu, du[2] * mu};
    const float fl[3] = {dl[0] * ml, dl[1] * ml, dl[2] * ml};
    if (imp_is_flat(g)) {
        out[0] = fu[0];
        out[1] = fu[1];
        out[2] = fu[2];
    } else {
        out[0] = fu[0] + fl[0];
        out[1] = fu[1] + fl[1];
        out[2] = fu[2] + fl[2];
    }
}

/* ---------------------------------------------------------------------------
 * The same formulas in double precision (FP64 oracle).
 * ------------------------------------------------------------------------- */
static double steric_energy_d(int mode, double ri, double rj, double ei, double ej, double d)
{
    switch (mode) {
    case BS_STERIC_AMBER_12_6: {
        if (d < K_MIN_VDW)
            return 0.0;
        const double eij = sqrt(ei * ej), s = sqrt(ri * rj) / d;
        const double s6 = s * s * s * s * s * s;
        return eij * s6 * s6 - eij * 2.0 * s6;
    }
    case BS_STERIC_LEWITT_8_6: {
        if (d < K_MIN_VDW)
            return 0.0;
        const double eij = sqrt(ei * ej), s = sqrt(ri * rj) / d;
        const double s2 = s * s, s6 = s2 * s2 * s2;
        return eij * 3.0 * s6 * s2 - eij * 4.0 * s6;
    }
    case BS_STERIC_ZACHARIAS_8_6: {
        if (d < K_MIN_VDW)
            return 0.0;
        const double eij = ei * ej, s = ri * rj / d;
        const double s2 = s * s, s6 = s2 * s2 * s2;
        return eij * s6 * s2 - eij * s6;
    }
    default: {
        const double dv = d - (ri + rj);
        if (dv > 0)
            return 0.0;
        return 0.5 * dv * dv;
    }
    }
}

static double steric_fm_d(int mode, double ri, double rj, double ei, double ej, double d)
{
    switch (mode) {
    case BS_STERIC_AMBER_12_6: {
        if (d < K_MIN_VDW)
            return 0.0;
        const double eij = sqrt(ei * ej), s = sqrt(ri * rj) / d;
        const double s6 = s * s * s * s * s * s;
        return (-12.0 * eij * s6 * s6 + 12.0 * eij * s6) / d * c_spring();
    }
    case BS_STERIC_LEWITT_8_6: {
        if (d < K_MIN_VDW)
            return 0.0;
        const double eij = sqrt(ei * ej), s = sqrt(ri * rj) / d;
        const double s2 = s * s, s6 = s2 * s2 * s2;
        return (-24.0 * eij * s6 * s2 + 24.0 * eij * s6) / d * c_spring();
    }
    case BS_STERIC_ZACHARIAS_8_6: {
        if (d < K_MIN_VDW)
            return 0.0;
        const double eij = ei * ej, s = ri * rj / d;
        const double s2 = s * s, s6 = s2 * s2 * s2;
        return (-8.0 * eij * s6 * s2 + 6.0 * eij * s6) / d * c_spring();
    }
    default: {
        const double dv = d - (ri + rj);
        if (dv > 0)
            return 0.0;
        return -fabs(dv) * c_spring();
    }
    }
}

static inline double elec_energy_d(double q1, double q2, double d, double diel, int ddd)
{
    if (d < K_MIN_ELEC)
        return 0.0;
    const double eps = ddd ? diel * d : diel;
    return c_coulomb() * (q1 * K_E_TO_C) * (q2 * K_E_TO_C) / (eps * d * K_A_TO_M) * K_NA * 1.0E-3;
}

static inline double elec_fm_d(double q1, double q2, double d, double diel, int ddd)
{
    if (d < K_MIN_ELEC)
        return 0.0;
    const double eps = ddd ? diel * d : diel;
    const double fm = -(q1 * q2) / (4.0 * K_PI * eps * d * d) * c_elec();
    return ddd ? 2.0 * fm : fm;
}

static inline double hydro_energy_d(double h1, double h2, double d) { return -(h1 * h2) * exp(-d) * K_NA * 1.0E-3; }
static inline double hydro_fm_d(double h1, double h2, double d) { return (h1 * h2) * exp(-d) * c_spring(); }

static void imp_zprime_d(const imp_geom *g, double x, double y, double z, double *z_up, double *z_lo,
                         double vup[3], double vlo[3], int *s_up, int *s_lo)
{
    const double rup = g->up_curv == 0.0f ? 1000000.0 : fabs(1.0 / g->up_curv);
    const double rlo = g->lo_curv == 0.0f ? 1000000.0 : fabs(1.0 / g->lo_curv);
    const int sup = sgn_f(g->up_curv), slo = sgn_f(g->lo_curv);
    const double cup = g->up_off - sup * rup;
    const double clo = -(double)g->lo_off - slo * rlo;
    vup[0] = x;
    vup[1] = 0.0;
    vup[2] = z - cup;
    vlo[0] = x;
    vlo[1] = 0.0;
    vlo[2] = z - clo;
    const double nup = norm_d(vup[0], vup[1], vup[2]);
    const double nlo = norm_d(vlo[0], vlo[1], vlo[2]);
    *z_up = sup == 0 ? z : (z > cup ? sup * nup + g->up_off - rup : -sup * nup + g->up_off - rup);
    *z_lo = slo == 0 ? z : (z > clo ? slo * nlo - g->lo_off - rlo : -slo * nlo - g->lo_off - rlo);
    *s_up = sup;
    *s_lo = slo;
    (void)y;
}

static double imp_energy_d(const imp_geom *g, double x, double y, double z, double sasa, double tr)
{
    double zu, zl, vu[3], vl[3];
    int su, sl;
    imp_zprime_d(g, x, y, z, &zu, &zl, vu, vl, &su, &sl);
    const double a = K_IMP_ALPHA, z0 = K_IMP_Z0, alip = K_IMP_ALIP;
    const double cu = 0.5 - 1.0 / (1.0 + exp(a * (fabs(zu - g->up_off) - z0)));
    const double cl = 0.5 - 1.0 / (1.0 + exp(a * (fabs(zl + g->lo_off) - z0)));
    const double eu = (-sasa * tr + alip * sasa) * cu;
    const double el = (-sasa * tr + alip * sasa) * cl;
    return imp_is_flat(g) ? eu : eu + el;
}

static void imp_force_d(const imp_geom *g, double x, double y, double z, double sasa, double tr, double out[3])
{
    double zu, zl, vu[3], vl[3];
    int su, sl;
    imp_zprime_d(g, x, y, z, &zu, &zl, vu, vl, &su, &sl);
    const double a = K_IMP_ALPHA, z0 = K_IMP_Z0, alip = K_IMP_ALIP;
    const double wu = zu - g->up_off, wl = zl + g->lo_off;
    const double xu = exp(a * (fabs(wu) - z0)), xl = exp(a * (fabs(wl) - z0));
    double dcu = (a * wu * xu) / ((xu + 1) * (xu + 1) * fabs(wu));
    double dcl = (a * wl * xl) / ((xl + 1) * (xl + 1) * fabs(wl));
    if (!isfinite(dcu))
        dcu = 0.0;
    if (!isfinite(dcl))
        dcl = 0.0;
    double du[3] = {0.0, 0.0, 1.0}, dl[3] = {0.0, 0.0, 1.0};
    if (su != 0)
        memcpy(du, vu, sizeof du);
    if (sl != 0)
        memcpy(dl, vl, sizeof dl);
    normalize_d(du);
    normalize_d(dl);
    const double mu = (-sasa * tr + alip * sasa) * dcu * c_spring();
    const double ml = (-sasa * tr + alip * sasa) * dcl * c_spring();
    for (int c = 0; c < 3; ++c)
        out[c] = imp_is_flat(g) ? du[c] * mu : du[c] * mu + dl[c] * ml;
}

/* exported for closed-form / finite-difference tests of single terms. */
double orc_term_energy(int term, int mode, const double *a, int na)
{
    /* term: 0 spring(d,k,l0) 1 steric(mode_s,ri,rj,ei,ej,d) 2 elec(q1,q2,d,diel,ddd)
     *       3 hydro(h1,h2,d) 4 imp(x,y,z,sasa,tr,uo,lo,uc,lc) */
    (void)na;
    if (mode == ORC_F32) {
        switch (term) {
        case 0: return spring_energy_f((float)a[0], (float)a[1], (float)a[2]);
        case 1: return steric_energy_f((int)a[0], (float)a[1], (float)a[2], (float)a[3], (float)a[4], (float)a[5]);
        case 2: return elec_energy_f((float)a[0], (float)a[1], (float)a[2], (float)a[3], (int)a[4]);
        case 3: return hydro_energy_f((float)a[0], (float)a[1], (float)a[2]);
        case 4: {
            imp_geom g = {(float)a[5], (float)a[6], (float)a[7], (float)a[8]};
            return imp_energy_f(&g, (float)a[0], (float)a[1], (float)a[2], (float)a[3], (float)a[4]);
        }
        }
    } else {
        switch (term) {
        case 0: { const double dv = a[0] - a[2]; return 0.5 * a[1] * dv * dv; }
        case 1: return steric_energy_d((int)a[0], a[1], a[2], a[3], a[4], a[5]);
        case 2: return elec_energy_d(a[0], a[1], a[2], a[3], (int)a[4]);
        case 3: return hydro_energy_d(a[0], a[1], a[2]);
        case 4: {
            imp_geom g = {(float)a[5], (float)a[6], (float)a[7], (float)a[8]};
            return imp_energy_d(&g, a[0], a[1], a[2], a[3], a[4]);
        }
        }
    }
    return 0.0;
}

/* Force module (or IMPALA force vector into out3) of a single term. */
double orc_term_force(int term, int mode, const double *a, int na, double *out3)
{
    (void)na;
    if (mode == ORC_F32) {
        switch (term) {
        case 0: return spring_fm_f((float)a[0], (float)a[1], (float)a[2]);
        case 1: return steric_fm_f((int)a[0], (float)a[1], (float)a[2], (float)a[3], (float)a[4], (float)a[5]);
        case 2: return elec_fm_f((float)a[0], (float)a[1], (float)a[2], (float)a[3], (int)a[4]);
        case 3: return hydro_fm_f((float)a[0], (float)a[1], (float)a[2]);
        case 4: {
            imp_geom g = {(float)a[5], (float)a[6], (float)a[7], (float)a[8]};
            float f[3];
            imp_force_f(&g, (float)a[0], (float)a[1], (float)a[2], (float)a[3], (float)a[4], f);
            for (int c = 0; c < 3; ++c)
                out3[c] = f[c];
            return 0.0;
        }
        }
    } else {
        switch (term) {
        case 0: return a[1] * (a[0] - a[2]) * c_spring();
        case 1: return steric_fm_d((int)a[0], a[1], a[2], a[3], a[4], a[5]);
        case 2: return elec_fm_d(a[0], a[1], a[2], a[3], (int)a[4]);
        case 3: return hydro_fm_d(a[0], a[1], a[2]);
        case 4: {
            imp_geom g = {(float)a[5], (float)a[6], (float)a[7], (float)a[8]};
            imp_force_d(&g, a[0], a[1], a[2], a[3], a[4], out3);
            return 0.0;
        }
        }
    }
    return 0.0;
}

/* ---------------------------------------------------------------------------
 * System state.
 * ------------------------------------------------------------------------- */
typedef struct orc {
    bs_params p;
    int64_t n;
    float *xf, *vf, *ff;    /* float state, 3n */
    double *xd, *vd, *fd;   /* double state, 3n */
    float *xdec;            /* float positions used for pair decisions, 3n */
    float *mass, *charge, *radius, *eps, *sasa, *transfer, *hydro;
    uint8_t *isstatic, *charged, *hydrophobic;
    int64_t ns;             /* deduplicated springs, in addition order */
    uint32_t *s1, *s2;
    float *sl0, *sk;
    int64_t nds;            /* dynamic springs (not both static) */
    int64_t *dsp;
    int64_t ndyn;
    int64_t *dyn;
    int64_t *bptr;          /* bonded-partner CSR, partners sorted */
    uint32_t *bidx;
    float ef[BS_NUM_ENERGIES];
    double ed[BS_NUM_ENERGIES];
    int64_t nonfinite;
    /* rigid bodies (RigidBody.cpp, RigidBodiesManager.cpp) */
    int nbodies;
    struct rbody *bodies;
    int32_t *body_of;       /* n, -1 for flexible particles */
    int64_t *local_of;      /* n, index inside the particle's body */
    /* grids: [BS_GRID_POTENTIAL], [BS_GRID_DENSITY] (PotentialGrid.hpp) */
    struct ogrid *grid[2];
    float *burying;         /* ParticleProperty::_burying, default 1 */
    /* the probe particle (SpringNetwork.cpp:156-182,749-766): float and
     * double state and the force of the last evaluation; pdec is the float
     * rounding of the double position (pair decisions, like xdec) */
    float pxf[3], pvf[3], pff[3], pdec[3];
    double pxd[3], pvd[3], pfd[3];
} orc;

/* A dense grid as PotentialGrid holds it after loading: per cell the scalar
 * and its gradient vector (PotentialGrid::compute_gradient), x-major with z
 * fastest (DenseGrid::_data[i][j][k]). */
typedef struct ogrid {
    double origin[3], cell[3], bmax[3];
    int64_t shape[3];
    float *cf;    /* 4 per cell: scalar, vector (float-faithful gradient) */
    double *cd;   /* 4 per cell: scalar, vector (FP64 gradient) */
} ogrid;

/* One rigid body, RigidBody.h:156-183.  The F32 state follows the reference's
 * Vector3f/Quaternion (float) and Matrix/mass/dt (double) types; the F64 state
 * is the same algorithm in double. */
typedef struct rbody {
    int64_t n;
    int64_t *ids;           /* particle ids in body order (_particulesIds) */
    double M;               /* _mass, double sum of float masses */
    double iinv[9];         /* _ibodyinv */
    float X[3], v[3], a[3], L[3], w[3], q[4], F[3], T[3]; /* q = (x, y, z, w) */
    float *p0;              /* 3n, _p0 */
    double Xd[3], vd[3], Ld[3], wd[3], qd[4], Fd[3], Td[3];
    double *p0d;
} rbody;

static void potential_field_f(const orc *o, int64_t i, float *fi, float *ei);
static void potential_field_d(const orc *o, int64_t i, const float *pos, double *fi, double *ei);
static void density_field_f(const orc *o, int64_t i, float *fi);
static void density_field_d(const orc *o, int64_t i, const float *pos, double *fi);

typedef struct {
    double box[6];
    int64_t dims[3];
    float radius;           /* cell edge, float like NeighborSearch::_search_radius */
    float cutoff;
    int64_t ncells;
    int64_t *start;         /* ncells+1 */
    int64_t *members;       /* nmem, ids ascending within a cell */
    int64_t nmem;
    int64_t *cell_of;       /* n, -1 when not a member */
} cells_t;

static int cmp_u32(const void *a, const void *b)
{
    const uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return (x > y) - (x < y);
}

static int is_bonded(const orc *o, int64_t i, uint32_t j)
{
    int64_t lo = o->bptr[i], hi = o->bptr[i + 1];
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (o->bidx[mid] < j)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo < o->bptr[i + 1] && o->bidx[lo] == j;
}

/* SpringNetwork::_setupProbe (SpringNetwork.cpp:749-766): the probe starts at
 * probe.x/y/z at rest. */
static void probe_reset(orc *o)
{
    const float x[3] = {o->p.probe_x, o->p.probe_y, o->p.probe_z};
    for (int c = 0; c < 3; ++c) {
        o->pxf[c] = x[c];
        o->pxd[c] = x[c];
        o->pdec[c] = x[c];
        o->pvf[c] = o->pff[c] = 0.0f;
        o->pvd[c] = o->pfd[c] = 0.0;
    }
}

orc *orc_create(const bs_params *p, int64_t n, const float *x3, const float *v3, const float *mass,
                const float *charge, const float *radius, const float *eps, const float *sasa,
                const float *transfer, const float *hydro, const uint8_t *flags, int64_t m,
                const uint32_t *id1, const uint32_t *id2, const float *l0, const float *k)
{
    orc *o = (orc *)calloc(1, sizeof(orc));
    o->p = *p;
    o->n = n;
    probe_reset(o);
    o->xf = (float *)malloc(sizeof(float) * 3 * n);
    o->vf = (float *)calloc(3 * n, sizeof(float));
    o->ff = (float *)calloc(3 * n, sizeof(float));
    o->xd = (double *)malloc(sizeof(double) * 3 * n);
    o->vd = (double *)calloc(3 * n, sizeof(double));
    o->fd = (double *)calloc(3 * n, sizeof(double));
    o->xdec = (float *)malloc(sizeof(float) * 3 * n);
    o->mass = (float *)malloc(sizeof(float) * n);
    o->charge = (float *)malloc(sizeof(float) * n);
    o->radius = (float *)malloc(sizeof(float) * n);
    o->eps = (float *)malloc(sizeof(float) * n);
    o->sasa = (float *)malloc(sizeof(float) * n);
    o->transfer = (float *)malloc(sizeof(float) * n);
    o->hydro = (float *)malloc(sizeof(float) * n);
    o->isstatic = (uint8_t *)malloc(n);
    o->charged = (uint8_t *)malloc(n);
    o->hydrophobic = (uint8_t *)malloc(n);
    o->dyn = (int64_t *)malloc(sizeof(int64_t) * (n ? n : 1));
    o->ndyn = 0;
    o->body_of = (int32_t *)malloc(sizeof(int32_t) * (n ? n : 1));
    o->local_of = (int64_t *)malloc(sizeof(int64_t) * (n ? n : 1));
    o->burying = (float *)malloc(sizeof(float) * (n ? n : 1));
    for (int64_t i = 0; i < n; ++i) {
        o->body_of[i] = -1;
        o->local_of[i] = -1;
        o->burying[i] = 1.0f;
    }
    for (int64_t i = 0; i < n; ++i) {
        for (int c = 0; c < 3; ++c) {
            o->xf[3 * i + c] = x3[3 * i + c];
            o->xd[3 * i + c] = x3[3 * i + c];
            o->vf[3 * i + c] = v3 ? v3[3 * i + c] : 0.0f;
            o->vd[3 * i + c] = o->vf[3 * i + c];
        }
        /* ParticleProperty defaults, ParticleProperty.h:20-25 */
        o->mass[i] = mass ? mass[i] : 1.0f;
        o->charge[i] = charge ? charge[i] : 0.0f;
        o->radius[i] = radius ? radius[i] : 1.0f;
        o->eps[i] = eps ? eps[i] : 0.0f;
        o->sasa[i] = sasa ? sasa[i] : 0.0f;
        o->transfer[i] = transfer ? transfer[i] : 0.0f;
        o->hydro[i] = hydro ? hydro[i] : 0.0f;
        o->isstatic[i] = flags ? (flags[i] & BS_FLAG_STATIC) != 0 : 0;
        o->charged[i] = o->charge[i] != 0.0f;   /* ParticleProperty.cpp:8-16 */
        o->hydrophobic[i] = o->hydro[i] != 0.0f; /* ParticleProperty.cpp:28-35 */
        if (!o->isstatic[i])
            o->dyn[o->ndyn++] = i;
    }
    /* Springs: dedup on the unordered pair like SpringNetwork::addSpring's
     * isInSpringNeighbors test (SpringNetwork.cpp:459-486). */
    int64_t *deg = (int64_t *)calloc(n + 1, sizeof(int64_t));
    uint32_t *tmp = (uint32_t *)malloc(sizeof(uint32_t) * 2 * (m ? m : 1));
    int64_t *tptr = (int64_t *)calloc(n + 1, sizeof(int64_t));
    for (int64_t s = 0; s < m; ++s) {
        deg[id1[s]]++;
        deg[id2[s]]++;
    }
    for (int64_t i = 0; i < n; ++i)
        tptr[i + 1] = tptr[i] + deg[i];
    int64_t *fill = (int64_t *)calloc(n + 1, sizeof(int64_t));
    for (int64_t s = 0; s < m; ++s) {
        tmp[tptr[id1[s]] + fill[id1[s]]++] = id2[s];
        tmp[tptr[id2[s]] + fill[id2[s]]++] = id1[s];
    }
    /* keep first occurrence per unordered pair: mark duplicates by sorting partners */
    o->bptr = (int64_t *)calloc(n + 1, sizeof(int64_t));
    o->bidx = (uint32_t *)malloc(sizeof(uint32_t) * 2 * (m ? m : 1));
    for (int64_t i = 0; i < n; ++i) {
        qsort(tmp + tptr[i], (size_t)(tptr[i + 1] - tptr[i]), sizeof(uint32_t), cmp_u32);
        int64_t w = o->bptr[i];
        for (int64_t e = tptr[i]; e < tptr[i + 1]; ++e)
            if (e == tptr[i] || tmp[e] != tmp[e - 1])
                o->bidx[w++] = tmp[e];
        o->bptr[i + 1] = w;
    }
    /* Spring list in addition order, skipping repeats of an already added pair. */
    o->s1 = (uint32_t *)malloc(sizeof(uint32_t) * (m ? m : 1));
    o->s2 = (uint32_t *)malloc(sizeof(uint32_t) * (m ? m : 1));
    o->sl0 = (float *)malloc(sizeof(float) * (m ? m : 1));
    o->sk = (float *)malloc(sizeof(float) * (m ? m : 1));
    o->dsp = (int64_t *)malloc(sizeof(int64_t) * (m ? m : 1));
    /* seen-set via per-particle used flags on the sorted partner list */
    uint8_t *used = (uint8_t *)calloc(o->bptr[n] ? o->bptr[n] : 1, 1);
    o->ns = 0;
    o->nds = 0;
    for (int64_t s = 0; s < m; ++s) {
        const uint32_t a = id1[s], b = id2[s];
        /* locate b in a's sorted partner list */
        int64_t lo = o->bptr[a], hi = o->bptr[a + 1];
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (o->bidx[mid] < b)
                lo = mid + 1;
            else
                hi = mid;
        }
        if (used[lo])
            continue;
        used[lo] = 1;
        int64_t lo2 = o->bptr[b], hi2 = o->bptr[b + 1];
        while (lo2 < hi2) {
            const int64_t mid = (lo2 + hi2) >> 1;
            if (o->bidx[mid] < a)
                lo2 = mid + 1;
            else
                hi2 = mid;
        }
        used[lo2] = 1;
        o->s1[o->ns] = a;
        o->s2[o->ns] = b;
        o->sl0[o->ns] = l0[s];
        o->sk[o->ns] = k[s];
        if (!(o->isstatic[a] && o->isstatic[b]))
            o->dsp[o->nds++] = o->ns;
        o->ns++;
    }
    free(used);
    free(deg);
    free(tmp);
    free(tptr);
    free(fill);
    o->nonfinite = -1;
    return o;
}

static void free_bodies(orc *o);

void orc_destroy(orc *o)
{
    if (!o)
        return;
    free(o->xf); free(o->vf); free(o->ff); free(o->xd); free(o->vd); free(o->fd); free(o->xdec);
    free(o->mass); free(o->charge); free(o->radius); free(o->eps); free(o->sasa); free(o->transfer);
    free(o->hydro); free(o->isstatic); free(o->charged); free(o->hydrophobic); free(o->dyn);
    free(o->bptr); free(o->bidx); free(o->s1); free(o->s2); free(o->sl0); free(o->sk); free(o->dsp);
    free_bodies(o);
    free(o->body_of); free(o->local_of); free(o->burying);
    for (int k = 0; k < 2; ++k)
        if (o->grid[k]) {
            free(o->grid[k]->cf);
            free(o->grid[k]->cd);
            free(o->grid[k]);
        }
    free(o);
}

int64_t orc_num_springs(const orc *o) { return o->ns; }

/* New parameters; a probe switched on here is placed as at setup. */
void orc_set_params(orc *o, const bs_params *p)
{
    const int had_probe = o->p.probe_enable;
    o->p = *p;
    if (p->probe_enable && !had_probe)
        probe_reset(o);
}

/* The probe's position, velocity and the force of the last evaluation. */
void orc_probe_state(const orc *o, int mode, double *out9)
{
    for (int c = 0; c < 3; ++c) {
        out9[c] = mode == ORC_F32 ? o->pxf[c] : o->pxd[c];
        out9[3 + c] = mode == ORC_F32 ? o->pvf[c] : o->pvd[c];
        out9[6 + c] = mode == ORC_F32 ? o->pff[c] : o->pfd[c];
    }
}

void orc_set_state(orc *o, const float *x3, const float *v3)
{
    for (int64_t i = 0; i < 3 * o->n; ++i) {
        if (x3) {
            o->xf[i] = x3[i];
            o->xd[i] = x3[i];
        }
        if (v3) {
            o->vf[i] = v3[i];
            o->vd[i] = v3[i];
        }
    }
}

void orc_set_state_d(orc *o, const double *x3, const double *v3)
{
    for (int64_t i = 0; i < 3 * o->n; ++i) {
        if (x3)
            o->xd[i] = x3[i];
        if (v3)
            o->vd[i] = v3[i];
    }
}

void orc_get_state(const orc *o, int mode, double *x3, double *v3)
{
    for (int64_t i = 0; i < 3 * o->n; ++i) {
        if (x3)
            x3[i] = mode == ORC_F32 ? (double)o->xf[i] : o->xd[i];
        if (v3)
            v3[i] = mode == ORC_F32 ? (double)o->vf[i] : o->vd[i];
    }
}

/* ---------------------------------------------------------------------------
 * Cell list, nsearch.hpp:342-390 (NeighborSearch::_build_grid).
 * ------------------------------------------------------------------------- */
static int member_of(const orc *o, int kind, int64_t i)
{
    if (kind == BS_NS_ELECTROSTATIC)
        return o->charged[i];
    if (kind == BS_NS_HYDROPHOBIC)
        return o->hydrophobic[i];
    return 1;
}

static float kind_cutoff(const orc *o, int kind)
{
    if (kind == BS_NS_ELECTROSTATIC)
        return o->p.electrostatic_cutoff;
    if (kind == BS_NS_HYDROPHOBIC)
        return o->p.hydrophobicity_cutoff;
    return o->p.steric_cutoff;
}

static inline int64_t cell_coord(double x, double mn, float r, int64_t na)
{
    int64_t c = (int64_t)(size_t)((x - mn) / r); /* nsearch.hpp:282-289 */
    return c < na - 1 ? c : na - 1;
}

static void build_cells(const orc *o, int kind, const float *pos, cells_t *cl)
{
    const int64_t n = o->n;
    /* Box(container) over every particle, box.hpp:37-69; the container holds
     * the probe too (it is appended to the particles, SpringNetwork.cpp:763-764,
     * and only left out of the cells, nsearch.hpp:363-365) */
    double mn[3], mx[3];
    for (int c = 0; c < 3; ++c)
        mn[c] = mx[c] = pos[c];
    for (int64_t i = 1; i < n; ++i)
        for (int c = 0; c < 3; ++c) {
            const double v = pos[3 * i + c];
            if (v < mn[c])
                mn[c] = v;
            if (v > mx[c])
                mx[c] = v;
        }
    if (o->p.probe_enable) {
        const float *pp = pos == o->xf ? o->pxf : o->pdec;
        for (int c = 0; c < 3; ++c) {
            if ((double)pp[c] < mn[c])
                mn[c] = pp[c];
            if ((double)pp[c] > mx[c])
                mx[c] = pp[c];
        }
    }
    cl->cutoff = kind_cutoff(o, kind);
    cl->radius = cl->cutoff + o->p.neighborskin;
    for (int c = 0; c < 3; ++c) {
        cl->box[c] = mn[c];
        cl->box[3 + c] = mx[c];
        cl->dims[c] = (int64_t)(size_t)(ceil((mx[c] - mn[c]) / cl->radius) + 1);
    }
    cl->ncells = cl->dims[0] * cl->dims[1] * cl->dims[2];
    cl->start = (int64_t *)calloc(cl->ncells + 1, sizeof(int64_t));
    cl->cell_of = (int64_t *)malloc(sizeof(int64_t) * (n ? n : 1));
    cl->nmem = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (!member_of(o, kind, i)) {
            cl->cell_of[i] = -1;
            continue;
        }
        const int64_t cx = cell_coord(pos[3 * i + 0], mn[0], cl->radius, cl->dims[0]);
        const int64_t cy = cell_coord(pos[3 * i + 1], mn[1], cl->radius, cl->dims[1]);
        const int64_t cz = cell_coord(pos[3 * i + 2], mn[2], cl->radius, cl->dims[2]);
        cl->cell_of[i] = cx + cy * cl->dims[0] + cz * cl->dims[0] * cl->dims[1];
        cl->start[cl->cell_of[i] + 1]++;
        cl->nmem++;
    }
    for (int64_t c = 0; c < cl->ncells; ++c)
        cl->start[c + 1] += cl->start[c];
    cl->members = (int64_t *)malloc(sizeof(int64_t) * (cl->nmem ? cl->nmem : 1));
    int64_t *cur = (int64_t *)malloc(sizeof(int64_t) * (cl->ncells ? cl->ncells : 1));
    memcpy(cur, cl->start, sizeof(int64_t) * cl->ncells);
    for (int64_t i = 0; i < n; ++i)
        if (cl->cell_of[i] >= 0)
            cl->members[cur[cl->cell_of[i]]++] = i; /* push_back in ascending id order */
    free(cur);
}

static void free_cells(cells_t *cl)
{
    free(cl->start);
    free(cl->members);
    free(cl->cell_of);
}

static inline int64_t query_cell(const cells_t *cl, const float *xi, int c)
{
    return cell_coord(xi[c], cl->box[c], cl->radius, cl->dims[c]);
}

/* Calls visit(j) for every candidate of nsearch::for_each_neighbor
 * (nsearch.hpp:214-234): 27 cells, x-outer z-inner, FP64 distance < cutoff. */
#define FOR_EACH_NEIGHBOR(cl, pos, i, J, BODY)                                                         \
    do {                                                                                            \
        const float *xi_ = (pos) + 3 * (i);                                                         \
        const int64_t qx_ = query_cell((cl), xi_, 0), qy_ = query_cell((cl), xi_, 1),               \
                      qz_ = query_cell((cl), xi_, 2);                                               \
        for (int dx_ = -1; dx_ <= 1; dx_++)                                                         \
            for (int dy_ = -1; dy_ <= 1; dy_++)                                                     \
                for (int dz_ = -1; dz_ <= 1; dz_++) {                                               \
                    const int64_t ax_ = qx_ + dx_, ay_ = qy_ + dy_, az_ = qz_ + dz_;                 \
                    if (ax_ < 0 || ax_ >= (cl)->dims[0] || ay_ < 0 || ay_ >= (cl)->dims[1] ||        \
                        az_ < 0 || az_ >= (cl)->dims[2])                                            \
                        continue;                                                                   \
                    const int64_t cid_ = ax_ + ay_ * (cl)->dims[0] + az_ * (cl)->dims[0] * (cl)->dims[1]; \
                    for (int64_t e_ = (cl)->start[cid_]; e_ < (cl)->start[cid_ + 1]; ++e_) {        \
                        const int64_t J = (cl)->members[e_];                                        \
                        if (J == (i))                                                               \
                            continue;                                                               \
                        const float *xj_ = (pos) + 3 * J;                                           \
                        const double ddx_ = (double)xi_[0] - (double)xj_[0];                        \
                        const double ddy_ = (double)xi_[1] - (double)xj_[1];                        \
                        const double ddz_ = (double)xi_[2] - (double)xj_[2];                        \
                        if (!(sqrt(ddx_ * ddx_ + ddy_ * ddy_ + ddz_ * ddz_) < (double)(cl)->cutoff)) \
                            continue;                                                               \
                        BODY                                                                        \
                    }                                                                               \
                }                                                                                   \
    } while (0)

/* Particle-level acceptance of a candidate (Particle.cpp:205-224,384-402):
 * returns 1 and the float displacement/distance when the pair is processed. */
static inline int pair_accept(const orc *o, int kind, int64_t i, int64_t j, const float *pos, float cutoff,
                              float *d_out)
{
    if (!o->isstatic[j] && j < i)
        return 0;
    if (o->p.spring_enable && is_bonded(o, i, (uint32_t)j))
        return 0;
    if (j == i)
        return 0;
    const float f0 = pos[3 * j + 0] - pos[3 * i + 0];
    const float f1 = pos[3 * j + 1] - pos[3 * i + 1];
    const float f2 = pos[3 * j + 2] - pos[3 * i + 2];
    const float d = norm_f(f0, f1, f2);
    if (!(d < cutoff))
        return 0;
    if (kind != BS_NS_STERIC && !(d != 0.0f))
        return 0;
    *d_out = d;
    return 1;
}

static int kind_enabled(const orc *o, int kind)
{
    if (kind == BS_NS_STERIC)
        return o->p.steric_enable;
    if (kind == BS_NS_ELECTROSTATIC)
        return o->p.electrostatic_enable;
    return o->p.hydrophobicity_enable;
}

static const float *decision_positions(orc *o, int mode)
{
    if (mode == ORC_F32)
        return o->xf;
    for (int64_t i = 0; i < 3 * o->n; ++i)
        o->xdec[i] = (float)o->xd[i];
    for (int c = 0; c < 3; ++c)
        o->pdec[c] = (float)o->pxd[c];
    return o->xdec;
}

/* ---------------------------------------------------------------------------
 * Pair enumeration for the parity hooks.
 * ------------------------------------------------------------------------- */
typedef struct {
    uint32_t *ij;
    int64_t cap, count;
} pairbuf;

static void pb_push(pairbuf *pb, uint32_t a, uint32_t b)
{
    if (pb->count >= pb->cap) {
        pb->cap = pb->cap ? 2 * pb->cap : 1024;
        pb->ij = (uint32_t *)realloc(pb->ij, sizeof(uint32_t) * 2 * pb->cap);
    }
    pb->ij[2 * pb->count] = a < b ? a : b;
    pb->ij[2 * pb->count + 1] = a < b ? b : a;
    pb->count++;
}

static int cmp_pair(const void *a, const void *b)
{
    const uint32_t *x = (const uint32_t *)a, *y = (const uint32_t *)b;
    if (x[0] != y[0])
        return (x[0] > y[0]) - (x[0] < y[0]);
    return (x[1] > y[1]) - (x[1] < y[1]);
}

/* method 0: through the reference cell list; method 1: O(N^2) brute force
 * (NeighborSearchO2, nsearch.hpp:74-102) with the same two-stage test.
 * Returns the count; copies up to cap pairs into ij (sorted (min,max)). */
int64_t orc_pairs(orc *o, int kind, int method, int mode, uint32_t *ij, int64_t cap)
{
    const float *pos = decision_positions(o, mode);
    const float cutoff = kind_cutoff(o, kind);
    pairbuf pb = {0, 0, 0};
    cells_t cl;
    if (method == 0)
        build_cells(o, kind, pos, &cl);
    for (int64_t t = 0; t < o->ndyn; ++t) {
        const int64_t i = o->dyn[t];
        if (kind == BS_NS_ELECTROSTATIC && !o->charged[i])
            continue;
        if (kind == BS_NS_HYDROPHOBIC && !o->hydrophobic[i])
            continue;
        if (method == 0) {
            FOR_EACH_NEIGHBOR(&cl, pos, i, j, {
                float d;
                if (pair_accept(o, kind, i, j, pos, cutoff, &d))
                    pb_push(&pb, (uint32_t)i, (uint32_t)j);
            });
        } else {
            for (int64_t j = 0; j < o->n; ++j) {
                if (j == i || !member_of(o, kind, j))
                    continue;
                const double ddx = (double)pos[3 * i] - (double)pos[3 * j];
                const double ddy = (double)pos[3 * i + 1] - (double)pos[3 * j + 1];
                const double ddz = (double)pos[3 * i + 2] - (double)pos[3 * j + 2];
                if (!(sqrt(ddx * ddx + ddy * ddy + ddz * ddz) < (double)cutoff))
                    continue;
                float d;
                if (pair_accept(o, kind, i, j, pos, cutoff, &d))
                    pb_push(&pb, (uint32_t)i, (uint32_t)j);
            }
        }
    }
    if (method == 0)
        free_cells(&cl);
    if (pb.count)
        qsort(pb.ij, (size_t)pb.count, 2 * sizeof(uint32_t), cmp_pair);
    const int64_t cnt = pb.count;
    if (ij)
        memcpy(ij, pb.ij, sizeof(uint32_t) * 2 * (cnt < cap ? cnt : cap));
    free(pb.ij);
    return cnt;
}

/* Reference cell list of `kind` at the float positions of `mode`. */
int64_t orc_cells(orc *o, int kind, int mode, int64_t *cell_of, int64_t *start, int64_t *members,
                  int64_t dims[3], double box[6], int64_t *nmem)
{
    const float *pos = decision_positions(o, mode);
    cells_t cl;
    build_cells(o, kind, pos, &cl);
    if (cell_of)
        memcpy(cell_of, cl.cell_of, sizeof(int64_t) * o->n);
    if (start)
        memcpy(start, cl.start, sizeof(int64_t) * (cl.ncells + 1));
    if (members)
        memcpy(members, cl.members, sizeof(int64_t) * cl.nmem);
    if (dims)
        memcpy(dims, cl.dims, sizeof cl.dims);
    if (box)
        memcpy(box, cl.box, sizeof cl.box);
    if (nmem)
        *nmem = cl.nmem;
    const int64_t nc = cl.ncells;
    free_cells(&cl);
    return nc;
}

/* ---------------------------------------------------------------------------
 * Forces, float-faithful (SpringNetwork::computeForces).
 * ------------------------------------------------------------------------- */
typedef struct {
    uint32_t target;
    float f[3];
    float e;
} defer_f;

typedef struct {
    defer_f *v;
    int64_t n, cap;
} dvec_f;

static void dvf_push(dvec_f *d, uint32_t t, const float f[3], float e)
{
    if (d->n >= d->cap) {
        d->cap = d->cap ? 2 * d->cap : 256;
        d->v = (defer_f *)realloc(d->v, sizeof(defer_f) * d->cap);
    }
    d->v[d->n].target = t;
    d->v[d->n].f[0] = -f[0];
    d->v[d->n].f[1] = -f[1];
    d->v[d->n].f[2] = -f[2];
    d->v[d->n].e = e;
    d->n++;
}

typedef struct {
    uint32_t target;
    double f[3];
    double e;
} defer_d;

typedef struct {
    defer_d *v;
    int64_t n, cap;
} dvec_d;

static void dvd_push(dvec_d *d, uint32_t t, const double f[3], double e)
{
    if (d->n >= d->cap) {
        d->cap = d->cap ? 2 * d->cap : 256;
        d->v = (defer_d *)realloc(d->v, sizeof(defer_d) * d->cap);
    }
    d->v[d->n].target = t;
    d->v[d->n].f[0] = -f[0];
    d->v[d->n].f[1] = -f[1];
    d->v[d->n].f[2] = -f[2];
    d->v[d->n].e = e;
    d->n++;
}

static int max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

static int thread_id(void)
{
#ifdef _OPENMP
    return omp_get_thread_num();
#else
    return 0;
#endif
}

/* One pair from particle i's side: float-faithful evaluation, Particle.cpp. */
static inline void pair_eval_f(const orc *o, int kind, int64_t i, int64_t j, float d, float *fi, float *ei,
                               dvec_f *def)
{
    const bs_params *p = &o->p;
    float f[3] = {o->xf[3 * j] - o->xf[3 * i], o->xf[3 * j + 1] - o->xf[3 * i + 1],
                  o->xf[3 * j + 2] - o->xf[3 * i + 2]};
    float e, fm;
    if (kind == BS_NS_STERIC) {
        e = p->steric_gridscale *
            steric_energy_f(p->steric_mode, o->radius[j], o->radius[i], o->eps[j], o->eps[i], d);
        fm = p->steric_gridscale *
             steric_fm_f(p->steric_mode, o->radius[j], o->radius[i], o->eps[j], o->eps[i], d);
    } else if (kind == BS_NS_ELECTROSTATIC) {
        e = p->electrostatic_scale *
            elec_energy_f(o->charge[j], o->charge[i], d, p->electrostatic_dielectric, p->dielectric_model);
        fm = p->electrostatic_scale *
             elec_fm_f(o->charge[j], o->charge[i], d, p->electrostatic_dielectric, p->dielectric_model);
    } else {
        e = p->hydrophobicity_scale * hydro_energy_f(o->hydro[j], o->hydro[i], d);
        fm = p->hydrophobicity_scale * hydro_fm_f(o->hydro[j], o->hydro[i], d);
    }
    normalize_f(f);
    f[0] = f[0] * fm;
    f[1] = f[1] * fm;
    f[2] = f[2] * fm;
    fi[0] = fi[0] + f[0];
    fi[1] = fi[1] + f[1];
    fi[2] = fi[2] + f[2];
    if (!o->isstatic[j]) {
        const float half = kind == BS_NS_ELECTROSTATIC ? 0.5f * e : e / 2.0f;
        *ei += half;
        dvf_push(def, (uint32_t)j, f, half);
    } else {
        *ei += e;
    }
}

static inline void pair_eval_d(const orc *o, int kind, int64_t i, int64_t j, double *fi, double *ei, dvec_d *def)
{
    const bs_params *p = &o->p;
    double f[3] = {o->xd[3 * j] - o->xd[3 * i], o->xd[3 * j + 1] - o->xd[3 * i + 1],
                   o->xd[3 * j + 2] - o->xd[3 * i + 2]};
    const double d = norm_d(f[0], f[1], f[2]);
    double e, fm;
    if (kind == BS_NS_STERIC) {
        e = p->steric_gridscale *
            steric_energy_d(p->steric_mode, o->radius[j], o->radius[i], o->eps[j], o->eps[i], d);
        fm = p->steric_gridscale * steric_fm_d(p->steric_mode, o->radius[j], o->radius[i], o->eps[j], o->eps[i], d);
    } else if (kind == BS_NS_ELECTROSTATIC) {
        e = p->electrostatic_scale *
            elec_energy_d(o->charge[j], o->charge[i], d, p->electrostatic_dielectric, p->dielectric_model);
        fm = p->electrostatic_scale *
             elec_fm_d(o->charge[j], o->charge[i], d, p->electrostatic_dielectric, p->dielectric_model);
    } else {
        e = p->hydrophobicity_scale * hydro_energy_d(o->hydro[j], o->hydro[i], d);
        fm = p->hydrophobicity_scale * hydro_fm_d(o->hydro[j], o->hydro[i], d);
    }
    normalize_d(f);
    for (int c = 0; c < 3; ++c) {
        f[c] *= fm;
        fi[c] += f[c];
    }
    if (!o->isstatic[j]) {
        *ei += 0.5 * e;
        dvd_push(def, (uint32_t)j, f, 0.5 * e);
    } else {
        *ei += e;
    }
}

/* ---------------------------------------------------------------------------
 * The probe particle, SpringNetwork.cpp:154-182.  After the pair terms and
 * the per-particle energy sums: for every dynamic particle in order, its
 * steric then its Coulomb interaction with the probe (Particle.cpp:434-465;
 * no cutoff, no spring exclusion, the pair's full energy to the totals), the
 * force on the particle and its opposite on the probe; then the probe's
 * viscosity.  The probe's Euler step (probe_integrate_*) follows in a step.
 * ------------------------------------------------------------------------- */
static void probe_forces_f(orc *o, float *F, float *e_ster, float *e_elec)
{
    const bs_params *p = &o->p;
    float *pf = o->pff;
    pf[0] = pf[1] = pf[2] = 0.0f; /* resetForce */
    for (int64_t t = 0; t < o->ndyn; ++t) {
        const int64_t i = o->dyn[t];
        for (int term = 0; term < 2; ++term) {
            if (term == 0 ? !p->probe_enable_steric : !p->probe_enable_electrostatic)
                continue;
            float f[3] = {o->pxf[0] - o->xf[3 * i], o->pxf[1] - o->xf[3 * i + 1], o->pxf[2] - o->xf[3 * i + 2]};
            const float d = norm_f(f[0], f[1], f[2]);
            float e, fm;
            if (term == 0) {
                e = p->steric_gridscale *
                    steric_energy_f(p->steric_mode, p->probe_radius, o->radius[i], p->probe_epsilon, o->eps[i], d);
                fm = p->steric_gridscale *
                     steric_fm_f(p->steric_mode, p->probe_radius, o->radius[i], p->probe_epsilon, o->eps[i], d);
            } else {
                e = p->electrostatic_scale * elec_energy_f(p->probe_charge, o->charge[i], d,
                                                           p->electrostatic_dielectric, p->dielectric_model);
                fm = p->electrostatic_scale * elec_fm_f(p->probe_charge, o->charge[i], d,
                                                        p->electrostatic_dielectric, p->dielectric_model);
            }
            normalize_f(f);
            for (int c = 0; c < 3; ++c) {
                f[c] = f[c] * fm;
                F[3 * i + c] = F[3 * i + c] + f[c];
                pf[c] = pf[c] + -f[c];
            }
            if (term == 0)
                *e_ster += e;
            else
                *e_elec += e;
        }
    }
    if (p->viscosity_enable)
        for (int c = 0; c < 3; ++c)
            pf[c] = pf[c] - o->pvf[c] * p->viscosity_value;
}

static void probe_forces_d(orc *o, double *F, double *e_ster, double *e_elec)
{
    const bs_params *p = &o->p;
    double *pf = o->pfd;
    pf[0] = pf[1] = pf[2] = 0.0;
    for (int64_t t = 0; t < o->ndyn; ++t) {
        const int64_t i = o->dyn[t];
        for (int term = 0; term < 2; ++term) {
            if (term == 0 ? !p->probe_enable_steric : !p->probe_enable_electrostatic)
                continue;
            double f[3] = {o->pxd[0] - o->xd[3 * i], o->pxd[1] - o->xd[3 * i + 1], o->pxd[2] - o->xd[3 * i + 2]};
            const double d = norm_d(f[0], f[1], f[2]);
            double e, fm;
            if (term == 0) {
                e = p->steric_gridscale *
                    steric_energy_d(p->steric_mode, p->probe_radius, o->radius[i], p->probe_epsilon, o->eps[i], d);
                fm = p->steric_gridscale *
                     steric_fm_d(p->steric_mode, p->probe_radius, o->radius[i], p->probe_epsilon, o->eps[i], d);
            } else {
                e = p->electrostatic_scale * elec_energy_d(p->probe_charge, o->charge[i], d,
                                                           p->electrostatic_dielectric, p->dielectric_model);
                fm = p->electrostatic_scale * elec_fm_d(p->probe_charge, o->charge[i], d,
                                                        p->electrostatic_dielectric, p->dielectric_model);
            }
            normalize_d(f);
            for (int c = 0; c < 3; ++c) {
                F[3 * i + c] += f[c] * fm;
                pf[c] -= f[c] * fm;
            }
            if (term == 0)
                *e_ster += e;
            else
                *e_elec += e;
        }
    }
    if (p->viscosity_enable)
        for (int c = 0; c < 3; ++c)
            pf[c] -= o->pvd[c] * (double)p->viscosity_value;
}

/* The probe's IntegrateEuler and its kinetic energy, added to the step's
 * before the particles' (SpringNetwork.cpp:179-180, Particle.cpp:109-128). */
static void probe_integrate_f(orc *o)
{
    const float dt = o->p.timestep, m = o->p.probe_mass;
    float *x = o->pxf, *v = o->pvf, *f = o->pff;
    if (m > 0.0f)
        for (int c = 0; c < 3; ++c)
            v[c] = v[c] + (f[c] / m) * dt;
    const float vit = norm_f(v[0], v[1], v[2]);
    const float ke = (float)(0.5f * m * (vit * vit) * c_kinetic());
    for (int c = 0; c < 3; ++c)
        x[c] = x[c] + v[c] * dt;
    o->ef[BS_E_KINETIC] += ke;
}

static void probe_integrate_d(orc *o)
{
    const double dt = o->p.timestep, m = o->p.probe_mass;
    double *x = o->pxd, *v = o->pvd, *f = o->pfd;
    if (m > 0.0)
        for (int c = 0; c < 3; ++c)
            v[c] += f[c] / m * dt;
    o->ed[BS_E_KINETIC] += 0.5 * m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) * c_kinetic();
    for (int c = 0; c < 3; ++c)
        x[c] += v[c] * dt;
}

/* computeSpringForces + computeParticleForces, float-faithful.  ext3 (nullable)
 * is added first (SpringNetwork::setForce in idleRun). */
static void forces_f32(orc *o, const float *ext3)
{
    const int64_t n = o->n;
    const bs_params *p = &o->p;
    float *F = o->ff;
    for (int64_t i = 0; i < 3 * n; ++i)
        F[i] = 0.0f; /* resetForce at the end of the previous step */
    if (ext3)
        for (int64_t i = 0; i < 3 * n; ++i)
            F[i] = F[i] + ext3[i];
    for (int k = 0; k < BS_NUM_ENERGIES; ++k)
        o->ef[k] = 0.0f;

    /* --- springs, SpringNetwork.cpp:63-90 --- */
    if (p->spring_enable) {
        float *scratch = (float *)malloc(sizeof(float) * 3 * (o->nds ? o->nds : 1));
        float *senergy = (float *)malloc(sizeof(float) * (o->nds ? o->nds : 1));
#pragma omp parallel for schedule(static)
        for (int64_t t = 0; t < o->nds; ++t) {
            const int64_t s = o->dsp[t];
            const uint32_t a = o->s1[s], b = o->s2[s];
            senergy[t] = 0.0f;
            scratch[3 * t] = scratch[3 * t + 1] = scratch[3 * t + 2] = 0.0f;
            if ((o->body_of[a] >= 0 && o->body_of[b] >= 0) || (o->isstatic[a] && o->isstatic[b]))
                continue; /* Spring.cpp:24-25 */
            float d[3] = {o->xf[3 * b] - o->xf[3 * a], o->xf[3 * b + 1] - o->xf[3 * a + 1],
                          o->xf[3 * b + 2] - o->xf[3 * a + 2]};
            const float len = norm_f(d[0], d[1], d[2]);
            senergy[t] = p->spring_scale * spring_energy_f(len, o->sk[s], o->sl0[s]);
            normalize_f(d);
            const float fm = p->spring_scale * spring_fm_f(len, o->sk[s], o->sl0[s]);
            scratch[3 * t] = d[0] * fm;
            scratch[3 * t + 1] = d[1] * fm;
            scratch[3 * t + 2] = d[2] * fm;
        }
        float es = 0.0f;
        for (int64_t t = 0; t < o->nds; ++t) {
            const int64_t s = o->dsp[t];
            const uint32_t a = o->s1[s], b = o->s2[s];
            for (int c = 0; c < 3; ++c) {
                F[3 * a + c] = F[3 * a + c] + scratch[3 * t + c];
                F[3 * b + c] = F[3 * b + c] + -scratch[3 * t + c];
            }
            es += senergy[t];
        }
        o->ef[BS_E_SPRING] = es;
        free(scratch);
        free(senergy);
    }

    /* --- particle forces, SpringNetwork.cpp:93-198 --- */
    const float *pos = o->xf;
    cells_t cls[3];
    int have[3] = {0, 0, 0};
    for (int kind = 0; kind < 3; ++kind) {
        if (!kind_enabled(o, kind))
            continue;
        if (kind != BS_NS_STERIC) {
            int any = 0;
            for (int64_t i = 0; i < n && !any; ++i)
                any = member_of(o, kind, i);
            if (!any)
                continue; /* no search object when the subset is empty (SpringNetwork.cpp:729-735) */
        }
        build_cells(o, kind, pos, &cls[kind]);
        have[kind] = 1;
    }
    const int nth = max_threads();
    dvec_f *defs = (dvec_f *)calloc(3 * nth, sizeof(dvec_f));
    float *pe = (float *)calloc(4 * (n ? n : 1), sizeof(float)); /* elec, steric, imp, hydro per particle */
    const imp_geom g = {p->imp_upper_offset, p->imp_lower_offset, p->imp_upper_curvature, p->imp_lower_curvature};
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < o->ndyn; ++t) {
        const int tid = thread_id();
        const int64_t i = o->dyn[t];
        float *fi = F + 3 * i;
        if (p->electrostatic_enable && o->charged[i] && have[BS_NS_ELECTROSTATIC]) {
            const cells_t *cl = &cls[BS_NS_ELECTROSTATIC];
            FOR_EACH_NEIGHBOR(cl, pos, i, j, {
                float d;
                if (pair_accept(o, BS_NS_ELECTROSTATIC, i, j, pos, cl->cutoff, &d))
                    pair_eval_f(o, BS_NS_ELECTROSTATIC, i, j, d, fi, &pe[4 * i + 0], &defs[3 * tid + 1]);
            });
        }
        if (p->electrostatic_enable && p->potentialgrid_enable && o->grid[BS_GRID_POTENTIAL])
            potential_field_f(o, i, fi, &pe[4 * i + 0]);
        if (p->densitygrid_enable && o->grid[BS_GRID_DENSITY])
            density_field_f(o, i, fi);
        if (p->steric_enable && have[BS_NS_STERIC]) {
            const cells_t *cl = &cls[BS_NS_STERIC];
            FOR_EACH_NEIGHBOR(cl, pos, i, j, {
                float d;
                if (pair_accept(o, BS_NS_STERIC, i, j, pos, cl->cutoff, &d))
                    pair_eval_f(o, BS_NS_STERIC, i, j, d, fi, &pe[4 * i + 1], &defs[3 * tid + 0]);
            });
        }
        if (p->viscosity_enable) { /* Particle.cpp:80-84 */
            const float gam = p->viscosity_value;
            for (int c = 0; c < 3; ++c)
                fi[c] = fi[c] - o->vf[3 * i + c] * gam;
        }
        if (p->imp_enable) { /* Particle.cpp:259-267, ForceField.cpp:62-86 */
            pe[4 * i + 2] = p->imp_scale * imp_energy_f(&g, pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], o->sasa[i],
                                                         o->transfer[i]);
            float fv[3];
            imp_force_f(&g, pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], o->sasa[i], o->transfer[i], fv);
            for (int c = 0; c < 3; ++c)
                fi[c] = fi[c] + (-fv[c]) * p->imp_scale;
        }
        if (p->hydrophobicity_enable && o->hydrophobic[i] && have[BS_NS_HYDROPHOBIC]) {
            const cells_t *cl = &cls[BS_NS_HYDROPHOBIC];
            FOR_EACH_NEIGHBOR(cl, pos, i, j, {
                float d;
                if (pair_accept(o, BS_NS_HYDROPHOBIC, i, j, pos, cl->cutoff, &d))
                    pair_eval_f(o, BS_NS_HYDROPHOBIC, i, j, d, fi, &pe[4 * i + 3], &defs[3 * tid + 2]);
            });
        }
    }
    /* deferred Newton-3 halves: steric, then electrostatic, then hydrophobic
     * (SpringNetwork.cpp:139-141); thread buffers concatenate in particle order
     * under schedule(static). */
    float e_elec = 0.0f, e_ster = 0.0f, e_imp = 0.0f, e_hyd = 0.0f;
    const int order[3] = {0, 1, 2};
    float *acc[3] = {&e_ster, &e_elec, &e_hyd};
    for (int q = 0; q < 3; ++q)
        for (int t = 0; t < nth; ++t) {
            const dvec_f *d = &defs[3 * t + order[q]];
            for (int64_t e = 0; e < d->n; ++e) {
                const defer_f *c = &d->v[e];
                for (int k = 0; k < 3; ++k)
                    F[3 * c->target + k] = F[3 * c->target + k] + c->f[k];
                *acc[q] += c->e;
            }
        }
    for (int64_t t = 0; t < o->ndyn; ++t) {
        const int64_t i = o->dyn[t];
        e_elec += pe[4 * i + 0];
        e_ster += pe[4 * i + 1];
        e_imp += pe[4 * i + 2];
        e_hyd += pe[4 * i + 3];
    }
    if (p->probe_enable)
        probe_forces_f(o, F, &e_ster, &e_elec);
    o->ef[BS_E_ELECTROSTATIC] = e_elec;
    o->ef[BS_E_STERIC] = e_ster;
    o->ef[BS_E_IMP] = e_imp;
    o->ef[BS_E_HYDROPHOBIC] = e_hyd;
    for (int t = 0; t < 3 * nth; ++t)
        free(defs[t].v);
    free(defs);
    free(pe);
    for (int kind = 0; kind < 3; ++kind)
        if (have[kind])
            free_cells(&cls[kind]);
}

/* The same in double precision.  Pair decisions use float-rounded positions. */
static void forces_f64(orc *o, const double *ext3)
{
    const int64_t n = o->n;
    const bs_params *p = &o->p;
    double *F = o->fd;
    for (int64_t i = 0; i < 3 * n; ++i)
        F[i] = ext3 ? ext3[i] : 0.0;
    for (int k = 0; k < BS_NUM_ENERGIES; ++k)
        o->ed[k] = 0.0;
    if (p->spring_enable) {
        double *scratch = (double *)malloc(sizeof(double) * 3 * (o->nds ? o->nds : 1));
        double *senergy = (double *)malloc(sizeof(double) * (o->nds ? o->nds : 1));
#pragma omp parallel for schedule(static)
        for (int64_t t = 0; t < o->nds; ++t) {
            const int64_t s = o->dsp[t];
            const uint32_t a = o->s1[s], b = o->s2[s];
            double d[3] = {o->xd[3 * b] - o->xd[3 * a], o->xd[3 * b + 1] - o->xd[3 * a + 1],
                           o->xd[3 * b + 2] - o->xd[3 * a + 2]};
            const double len = norm_d(d[0], d[1], d[2]);
            const double ext = len - (double)o->sl0[s];
            senergy[t] = p->spring_scale * 0.5 * o->sk[s] * ext * ext;
            normalize_d(d);
            const double fm = p->spring_scale * o->sk[s] * ext * c_spring();
            for (int c = 0; c < 3; ++c)
                scratch[3 * t + c] = d[c] * fm;
        }
        double es = 0.0;
        for (int64_t t = 0; t < o->nds; ++t) {
            const int64_t s = o->dsp[t];
            for (int c = 0; c < 3; ++c) {
                F[3 * o->s1[s] + c] += scratch[3 * t + c];
                F[3 * o->s2[s] + c] -= scratch[3 * t + c];
            }
            es += senergy[t];
        }
        o->ed[BS_E_SPRING] = es;
        free(scratch);
        free(senergy);
    }
    const float *pos = decision_positions(o, ORC_F64);
    cells_t cls[3];
    int have[3] = {0, 0, 0};
    for (int kind = 0; kind < 3; ++kind) {
        if (!kind_enabled(o, kind))
            continue;
        if (kind != BS_NS_STERIC) {
            int any = 0;
            for (int64_t i = 0; i < n && !any; ++i)
                any = member_of(o, kind, i);
            if (!any)
                continue;
        }
        build_cells(o, kind, pos, &cls[kind]);
        have[kind] = 1;
    }
    const int nth = max_threads();
    dvec_d *defs = (dvec_d *)calloc(3 * nth, sizeof(dvec_d));
    double *pe = (double *)calloc(4 * (n ? n : 1), sizeof(double));
    const imp_geom g = {p->imp_upper_offset, p->imp_lower_offset, p->imp_upper_curvature, p->imp_lower_curvature};
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < o->ndyn; ++t) {
        const int tid = thread_id();
        const int64_t i = o->dyn[t];
        double *fi = F + 3 * i;
        if (p->electrostatic_enable && o->charged[i] && have[BS_NS_ELECTROSTATIC]) {
            const cells_t *cl = &cls[BS_NS_ELECTROSTATIC];
            FOR_EACH_NEIGHBOR(cl, pos, i, j, {
                float d;
                if (pair_accept(o, BS_NS_ELECTROSTATIC, i, j, pos, cl->cutoff, &d))
                    pair_eval_d(o, BS_NS_ELECTROSTATIC, i, j, fi, &pe[4 * i + 0], &defs[3 * tid + 1]);
            });
        }
        if (p->electrostatic_enable && p->potentialgrid_enable && o->grid[BS_GRID_POTENTIAL])
            potential_field_d(o, i, pos, fi, &pe[4 * i + 0]);
        if (p->densitygrid_enable && o->grid[BS_GRID_DENSITY])
            density_field_d(o, i, pos, fi);
        if (p->steric_enable && have[BS_NS_STERIC]) {
            const cells_t *cl = &cls[BS_NS_STERIC];
            FOR_EACH_NEIGHBOR(cl, pos, i, j, {
                float d;
                if (pair_accept(o, BS_NS_STERIC, i, j, pos, cl->cutoff, &d))
                    pair_eval_d(o, BS_NS_STERIC, i, j, fi, &pe[4 * i + 1], &defs[3 * tid + 0]);
            });
        }
        if (p->viscosity_enable)
            for (int c = 0; c < 3; ++c)
                fi[c] -= o->vd[3 * i + c] * (double)p->viscosity_value;
        if (p->imp_enable) {
            pe[4 * i + 2] = p->imp_scale * imp_energy_d(&g, o->xd[3 * i], o->xd[3 * i + 1], o->xd[3 * i + 2],
                                                         o->sasa[i], o->transfer[i]);
            double fv[3];
            imp_force_d(&g, o->xd[3 * i], o->xd[3 * i + 1], o->xd[3 * i + 2], o->sasa[i], o->transfer[i], fv);
            for (int c = 0; c < 3; ++c)
                fi[c] -= fv[c] * p->imp_scale;
        }
        if (p->hydrophobicity_enable && o->hydrophobic[i] && have[BS_NS_HYDROPHOBIC]) {
            const cells_t *cl = &cls[BS_NS_HYDROPHOBIC];
            FOR_EACH_NEIGHBOR(cl, pos, i, j, {
                float d;
                if (pair_accept(o, BS_NS_HYDROPHOBIC, i, j, pos, cl->cutoff, &d))
                    pair_eval_d(o, BS_NS_HYDROPHOBIC, i, j, fi, &pe[4 * i + 3], &defs[3 * tid + 2]);
            });
        }
    }
    double e_elec = 0.0, e_ster = 0.0, e_imp = 0.0, e_hyd = 0.0;
    double *acc[3] = {&e_ster, &e_elec, &e_hyd};
    for (int q = 0; q < 3; ++q)
        for (int t = 0; t < nth; ++t) {
            const dvec_d *d = &defs[3 * t + q];
            for (int64_t e = 0; e < d->n; ++e) {
                const defer_d *c = &d->v[e];
                for (int k = 0; k < 3; ++k)
                    F[3 * c->target + k] += c->f[k];
                *acc[q] += c->e;
            }
        }
    for (int64_t t = 0; t < o->ndyn; ++t) {
        const int64_t i = o->dyn[t];
        e_elec += pe[4 * i + 0];
        e_ster += pe[4 * i + 1];
        e_imp += pe[4 * i + 2];
        e_hyd += pe[4 * i + 3];
    }
    if (p->probe_enable)
        probe_forces_d(o, F, &e_ster, &e_elec);
    o->ed[BS_E_ELECTROSTATIC] = e_elec;
    o->ed[BS_E_STERIC] = e_ster;
    o->ed[BS_E_IMP] = e_imp;
    o->ed[BS_E_HYDROPHOBIC] = e_hyd;
    for (int t = 0; t < 3 * nth; ++t)
        free(defs[t].v);
    free(defs);
    free(pe);
    for (int kind = 0; kind < 3; ++kind)
        if (have[kind])
            free_cells(&cls[kind]);
}

/* ---------------------------------------------------------------------------
 * Rigid bodies.
 *   aggregation  RigidBody::computeParticleForceAndTorque  RigidBody.cpp:567-580
 *   solve        RigidBody::solve                           RigidBody.cpp:430-484
 *   placement    RigidBody::integrateParticleVelocity       RigidBody.cpp:600-635
 *   setup        RigidBody::init / RigidBodiesManager        RigidBody.cpp:57-124, RigidBodiesManager.cpp:11-43
 *   quaternions  Quaternion.cpp:11-96, Quaternion.h:17-65
 * The placement uses the particle's index inside its body.  The reference
 * passes the index in the dynamic-particle list (SpringNetwork.cpp:216-218),
 * which coincides for the single body SpringNetwork::run creates and is wrong
 * for several bodies; the oracle follows the intended (local) index.
 * ------------------------------------------------------------------------- */

/* Quaternion product a*b, Quaternion::operator* (Quaternion.cpp:80-92). */
static void qmul_f(const float a[4], const float b[4], float o[4])
{
    float ax1[3], ax2[3], ax3[3];
    for (int c = 0; c < 3; ++c) {
        ax1[c] = b[c] * a[3];
        ax2[c] = a[c] * b[3];
    }
    ax3[0] = a[1] * b[2] - a[2] * b[1];
    ax3[1] = a[2] * b[0] - a[0] * b[2];
    ax3[2] = a[0] * b[1] - a[1] * b[0];
    const float dot = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
    const float w = a[3] * b[3] - dot;
    for (int c = 0; c < 3; ++c)
        o[c] = (ax1[c] + ax2[c]) + ax3[c];
    o[3] = w;
}

static void qmul_d(const double a[4], const double b[4], double o[4])
{
    double t[4];
    t[0] = b[0] * a[3] + a[0] * b[3] + (a[1] * b[2] - a[2] * b[1]);
    t[1] = b[1] * a[3] + a[1] * b[3] + (a[2] * b[0] - a[0] * b[2]);
    t[2] = b[2] * a[3] + a[2] * b[3] + (a[0] * b[1] - a[1] * b[0]);
    t[3] = a[3] * b[3] - (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]);
    memcpy(o, t, sizeof t);
}

/* Quaternion::normalize: only above a magnitude of 0.005. */
static void qnormalize_f(float q[4])
{
    const float mag = sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    if (mag > 0.005)
        for (int c = 0; c < 4; ++c)
            q[c] /= mag;
}

static void qnormalize_d(double q[4])
{
    const double mag = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    if (mag > 0.005)
        for (int c = 0; c < 4; ++c)
            q[c] /= mag;
}

/* Quaternion::inverse: conjugate / |q|^2, with |q| = sqrt(w^2 + x^2 + y^2 + z^2). */
static void qinverse_f(const float q[4], float o[4])
{
    float av = sqrtf(q[3] * q[3] + q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
    av *= av;
    av = 1 / av;
    o[3] = q[3] * av;
    for (int c = 0; c < 3; ++c)
        o[c] = -q[c] * av;
}

static void qinverse_d(const double q[4], double o[4])
{
    const double av = 1.0 / (q[3] * q[3] + q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
    o[3] = q[3] * av;
    for (int c = 0; c < 3; ++c)
        o[c] = -q[c] * av;
}

/* Quaternion::fromAxisAngle; the axis is used as given (not normalised). */
static void qaxis_f(const float ax[3], float angle, float q[4])
{
    const float s = (float)sin(angle / 2.0);
    const float c = (float)cos(angle / 2.0);
    q[0] = ax[0] * s;
    q[1] = ax[1] * s;
    q[2] = ax[2] * s;
    q[3] = c;
    qnormalize_f(q);
}

static void qaxis_d(const double ax[3], double angle, double q[4])
{
    const double s = sin(angle / 2.0), c = cos(angle / 2.0);
    q[0] = ax[0] * s;
    q[1] = ax[1] * s;
    q[2] = ax[2] * s;
    q[3] = c;
    qnormalize_d(q);
}

/* Springs evaluated each step: not both static (SpringNetwork.cpp:478-483)
 * and not both rigid (RigidBody::UpdateSpringsState, RigidBody.cpp:42-53);
 * removal keeps the order of the others (SpringNetwork.h:197). */
static void rebuild_dynamic_springs(orc *o)
{
    o->nds = 0;
    for (int64_t s = 0; s < o->ns; ++s) {
        const uint32_t a = o->s1[s], b = o->s2[s];
        if (o->isstatic[a] && o->isstatic[b])
            continue;
        if (o->body_of[a] >= 0 && o->body_of[b] >= 0)
            continue;
        o->dsp[o->nds++] = s;
    }
}

static void free_bodies(orc *o)
{
    for (int b = 0; b < o->nbodies; ++b) {
        free(o->bodies[b].ids);
        free(o->bodies[b].p0);
        free(o->bodies[b].p0d);
    }
    free(o->bodies);
    o->bodies = NULL;
    o->nbodies = 0;
    for (int64_t i = 0; i < o->n; ++i) {
        o->body_of[i] = -1;
        o->local_of[i] = -1;
    }
}

/* RigidBodiesManager::InitRigidBodies: a new body from `ids` unless one of
 * them already belongs to a body (then nothing happens); an empty list
 * removes every body.  Returns the number of bodies, or -1 for an id out of
 * range.  The body's state is initialised from the current F32 and F64
 * positions (RigidBody::init, RigidBody.cpp:57-124). */
int orc_add_rigid_body(orc *o, int64_t n, const int64_t *ids)
{
    if (n <= 0) {
        free_bodies(o);
        rebuild_dynamic_springs(o);
        return 0;
    }
    for (int64_t t = 0; t < n; ++t) {
        if (ids[t] < 0 || ids[t] >= o->n)
            return -1;
        if (o->body_of[ids[t]] >= 0)
            return o->nbodies;
    }
    o->bodies = (rbody *)realloc(o->bodies, sizeof(rbody) * (o->nbodies + 1));
    rbody *B = &o->bodies[o->nbodies];
    memset(B, 0, sizeof *B);
    B->n = n;
    B->ids = (int64_t *)malloc(sizeof(int64_t) * n);
    B->p0 = (float *)malloc(sizeof(float) * 3 * n);
    B->p0d = (double *)malloc(sizeof(double) * 3 * n);
    memcpy(B->ids, ids, sizeof(int64_t) * n);
    /* computeTotalMass, computeBarycentre (RigidBody.cpp:128-154) */
    double M = 0.0;
    for (int64_t t = 0; t < n; ++t)
        M += o->mass[ids[t]];
    B->M = M;
    float bary[3] = {0.0f, 0.0f, 0.0f};
    double tm = 0.0;
    for (int64_t t = 0; t < n; ++t) {
        const int64_t i = ids[t];
        for (int c = 0; c < 3; ++c)
            bary[c] = bary[c] + o->xf[3 * i + c] * o->mass[i];
        tm = tm + o->mass[i];
    }
    const float inv = (float)(1.0 / tm);
    for (int c = 0; c < 3; ++c)
        B->X[c] = bary[c] * inv;
    double bd[3] = {0.0, 0.0, 0.0};
    for (int64_t t = 0; t < n; ++t)
        for (int c = 0; c < 3; ++c)
            bd[c] += o->xd[3 * ids[t] + c] * o->mass[ids[t]];
    for (int c = 0; c < 3; ++c)
        B->Xd[c] = bd[c] / M;
    /* computeAllLocalPositions (RigidBody.cpp:166-176) */
    for (int64_t t = 0; t < n; ++t)
        for (int c = 0; c < 3; ++c) {
            B->p0[3 * t + c] = o->xf[3 * ids[t] + c] - B->X[c];
            B->p0d[3 * t + c] = o->xd[3 * ids[t] + c] - B->Xd[c];
        }
    /* I_body = 2/5 M dmax^2 Id with dmax = 1; its inverse (RigidBody.cpp:103-108) */
    memset(B->iinv, 0, sizeof B->iinv);
    const double ib = (2.0 / 5.0) * M * 1.0 * 1.0;
    B->iinv[0] = B->iinv[4] = B->iinv[8] = 1.0 / ib;
    B->q[3] = 1.0f;
    B->qd[3] = 1.0;
    for (int64_t t = 0; t < n; ++t) {
        o->body_of[ids[t]] = o->nbodies;
        o->local_of[ids[t]] = t;
    }
    o->nbodies++;
    rebuild_dynamic_springs(o);
    return o->nbodies;
}

int orc_num_bodies(const orc *o) { return o->nbodies; }

/* Body state: X, v, L, omega, q(x, y, z, w) -> 16 values. */
int orc_rigid_state(const orc *o, int mode, int b, double *out16)
{
    if (b < 0 || b >= o->nbodies)
        return -1;
    const rbody *B = &o->bodies[b];
    for (int c = 0; c < 3; ++c) {
        out16[c] = mode == ORC_F32 ? B->X[c] : B->Xd[c];
        out16[3 + c] = mode == ORC_F32 ? B->v[c] : B->vd[c];
        out16[6 + c] = mode == ORC_F32 ? B->L[c] : B->Ld[c];
        out16[9 + c] = mode == ORC_F32 ? B->w[c] : B->wd[c];
    }
    for (int c = 0; c < 4; ++c)
        out16[12 + c] = mode == ORC_F32 ? B->q[c] : B->qd[c];
    return 0;
}

/* computeParticleForceAndTorque over the dynamic rigid particles in
 * particle order (SpringNetwork.cpp:186-192), after every force is in. */
static void rigid_aggregate_f(orc *o)
{
    for (int64_t t = 0; t < o->ndyn; ++t) {
        const int64_t i = o->dyn[t];
        const int b = o->body_of[i];
        if (b < 0)
            continue;
        rbody *B = &o->bodies[b];
        const float *f = o->ff + 3 * i;
        float r[3];
        for (int c = 0; c < 3; ++c)
            r[c] = o->xf[3 * i + c] - B->X[c];
        const float tq[3] = {r[1] * f[2] - r[2] * f[1], r[2] * f[0] - r[0] * f[2], r[0] * f[1] - r[1] * f[0]};
        for (int c = 0; c < 3; ++c) {
            B->F[c] += f[c];
            B->T[c] += 0.0f + tq[c];
        }
    }
}

static void rigid_aggregate_d(orc *o)
{
    for (int64_t t = 0; t < o->ndyn; ++t) {
        const int64_t i = o->dyn[t];
        const int b = o->body_of[i];
        if (b < 0)
            continue;
        rbody *B = &o->bodies[b];
        const double *f = o->fd + 3 * i;
        double r[3];
        for (int c = 0; c < 3; ++c)
            r[c] = o->xd[3 * i + c] - B->Xd[c];
        B->Fd[0] += f[0];
        B->Fd[1] += f[1];
        B->Fd[2] += f[2];
        B->Td[0] += r[1] * f[2] - r[2] * f[1];
        B->Td[1] += r[2] * f[0] - r[0] * f[2];
        B->Td[2] += r[0] * f[1] - r[1] * f[0];
    }
}

/* RigidBody::solve without restraints, sampling or Monte Carlo. */
static void rigid_solve_f(orc *o)
{
    const double dt = o->p.timestep;
    const float fdt = (float)dt;
    for (int b = 0; b < o->nbodies; ++b) {
        rbody *B = &o->bodies[b];
        const float fm = (float)B->M;
        for (int c = 0; c < 3; ++c) {
            B->a[c] = B->F[c] / fm;
            B->v[c] += B->a[c] * fdt;
        }
        for (int c = 0; c < 3; ++c)
            B->X[c] += B->v[c] * fdt + ((B->a[c] * 0.5f) * fdt) * fdt;
        for (int c = 0; c < 3; ++c)
            B->L[c] += B->T[c] * fdt;
        for (int r = 0; r < 3; ++r)
            B->w[r] = (float)(B->iinv[3 * r] * B->L[0] + B->iinv[3 * r + 1] * B->L[1] + B->iinv[3 * r + 2] * B->L[2]);
        const float ang = (float)((double)norm_f(B->w[0], B->w[1], B->w[2]) * dt);
        float dq[4], nq[4];
        qaxis_f(B->w, ang, dq);
        qmul_f(dq, B->q, nq);
        qnormalize_f(nq);
        memcpy(B->q, nq, sizeof nq);
        memset(B->F, 0, sizeof B->F);
        memset(B->T, 0, sizeof B->T);
    }
}

static void rigid_solve_d(orc *o)
{
    const double dt = o->p.timestep;
    for (int b = 0; b < o->nbodies; ++b) {
        rbody *B = &o->bodies[b];
        double a[3];
        for (int c = 0; c < 3; ++c) {
            a[c] = B->Fd[c] / B->M;
            B->vd[c] += a[c] * dt;
        }
        for (int c = 0; c < 3; ++c)
            B->Xd[c] += B->vd[c] * dt + a[c] * 0.5 * dt * dt;
        for (int c = 0; c < 3; ++c)
            B->Ld[c] += B->Td[c] * dt;
        for (int r = 0; r < 3; ++r)
            B->wd[r] = B->iinv[3 * r] * B->Ld[0] + B->iinv[3 * r + 1] * B->Ld[1] + B->iinv[3 * r + 2] * B->Ld[2];
        const double ang = norm_d(B->wd[0], B->wd[1], B->wd[2]) * dt;
        double dq[4], nq[4];
        qaxis_d(B->wd, ang, dq);
        qmul_d(dq, B->qd, nq);
        qnormalize_d(nq);
        memcpy(B->qd, nq, sizeof nq);
        memset(B->Fd, 0, sizeof B->Fd);
        memset(B->Td, 0, sizeof B->Td);
    }
}

/* New position X + q (p0, 0) q^-1 and velocity (x_new - x_prev) / dt. */
static void rigid_place_f(orc *o, int64_t i)
{
    const rbody *B = &o->bodies[o->body_of[i]];
    const int64_t l = o->local_of[i];
    const float P[4] = {B->p0[3 * l], B->p0[3 * l + 1], B->p0[3 * l + 2], 0.0f};
    float t1[4], qi[4], r[4];
    qmul_f(B->q, P, t1);
    qinverse_f(B->q, qi);
    qmul_f(t1, qi, r);
    const float fdt = (float)o->p.timestep;
    for (int c = 0; c < 3; ++c) {
        const float prev = o->xf[3 * i + c];
        const float nx = B->X[c] + r[c];
        o->xf[3 * i + c] = nx;
        o->vf[3 * i + c] = (nx - prev) / fdt;
    }
}

static void rigid_place_d(orc *o, int64_t i)
{
    const rbody *B = &o->bodies[o->body_of[i]];
    const int64_t l = o->local_of[i];
    const double P[4] = {B->p0d[3 * l], B->p0d[3 * l + 1], B->p0d[3 * l + 2], 0.0};
    double t1[4], qi[4], r[4];
    qmul_d(B->qd, P, t1);
    qinverse_d(B->qd, qi);
    qmul_d(t1, qi, r);
    const double dt = o->p.timestep;
    for (int c = 0; c < 3; ++c) {
        const double prev = o->xd[3 * i + c];
        const double nx = B->Xd[c] + r[c];
        o->xd[3 * i + c] = nx;
        o->vd[3 * i + c] = (nx - prev) / dt;
    }
}

/* ---------------------------------------------------------------------------
 * Grids.  PotentialGrid::compute_gradient (PotentialGrid.cpp:15-58),
 * GridCoordinatesSystem::cell_coordinates / is_out_of_grid
 * (GridCoordinatesSystem.hpp:85-93,206-218), the field forces of
 * Particle.cpp:136-189 and electrostatic_field_energy
 * (energy/electrostatic_field.hpp:15-23).
 * ------------------------------------------------------------------------- */
static const double K_BOLTZ = 1.3806504E-23;

static double grad_scale(void) { return K_BOLTZ * K_M_TO_A * c_newton(); }

/* kind: BS_GRID_POTENTIAL or BS_GRID_DENSITY; scalar holds shape[0]*shape[1]*
 * shape[2] values, z fastest.  Returns 0, or -1 for a bad argument. */
int orc_set_grid(orc *o, int kind, const double *origin, const int64_t *shape, const double *cell,
                 const float *scalar)
{
    if (kind < 0 || kind > 1)
        return -1;
    if (o->grid[kind]) {
        free(o->grid[kind]->cf);
        free(o->grid[kind]->cd);
        free(o->grid[kind]);
        o->grid[kind] = NULL;
    }
    if (!scalar)
        return 0;
    for (int c = 0; c < 3; ++c)
        if (shape[c] <= 0 || !(cell[c] > 0.0))
            return -1;
    ogrid *g = (ogrid *)calloc(1, sizeof(ogrid));
    const int64_t nx = shape[0], ny = shape[1], nz = shape[2], nc = nx * ny * nz;
    for (int c = 0; c < 3; ++c) {
        g->origin[c] = origin[c];
        g->cell[c] = cell[c];
        g->shape[c] = shape[c];
        g->bmax[c] = origin[c] + (double)shape[c] * cell[c]; /* Box(origin, shape * cell) */
    }
    g->cf = (float *)malloc(sizeof(float) * 4 * nc);
    g->cd = (double *)malloc(sizeof(double) * 4 * nc);
    const float fscale = (float)grad_scale();
    const double dscale = grad_scale();
    const float d3[3] = {(float)cell[0], (float)cell[1], (float)cell[2]};
#define GIDX(i, j, k) (((i) * ny + (j)) * nz + (k))
    for (int64_t i = 0; i < nx; ++i)
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t k = 0; k < nz; ++k) {
                const int64_t id = GIDX(i, j, k);
                const float cur = scalar[id];
                float gf[3] = {0.0f, 0.0f, 0.0f};
                double gd[3] = {0.0, 0.0, 0.0};
                const int64_t nb[3][2] = {{i > 0 && i < nx - 1 ? GIDX(i - 1, j, k) : -1, i < nx - 1 ? GIDX(i + 1, j, k) : -1},
                                          {j > 0 && j < ny - 1 ? GIDX(i, j - 1, k) : -1, j < ny - 1 ? GIDX(i, j + 1, k) : -1},
                                          {k > 0 && k < nz - 1 ? GIDX(i, j, k - 1) : -1, k < nz - 1 ? GIDX(i, j, k + 1) : -1}};
                for (int a = 0; a < 3; ++a) {
                    if (nb[a][0] < 0)
                        continue;
                    const float prv = scalar[nb[a][0]], nxt = scalar[nb[a][1]];
                    gf[a] = (float)(((cur - prv) + (nxt - cur)) / (d3[a] * 2.0));
                    gd[a] = ((double)nxt - (double)prv) / (2.0 * cell[a]);
                }
                g->cf[4 * id] = cur;
                g->cd[4 * id] = cur;
                for (int a = 0; a < 3; ++a) {
                    g->cf[4 * id + 1 + a] = gf[a] * -fscale;
                    g->cd[4 * id + 1 + a] = gd[a] * -dscale;
                }
            }
#undef GIDX
    o->grid[kind] = g;
    return 0;
}

void orc_set_burying(orc *o, const float *burying)
{
    for (int64_t i = 0; i < o->n; ++i)
        o->burying[i] = burying ? burying[i] : 1.0f;
}

/* Cell index of a position, -1 when off the grid. */
static int64_t grid_cell(const ogrid *g, double x, double y, double z)
{
    const double p[3] = {x, y, z};
    for (int c = 0; c < 3; ++c)
        if (p[c] < g->origin[c] || p[c] > g->bmax[c] - 1e-6)
            return -1;
    int64_t q[3];
    for (int c = 0; c < 3; ++c) {
        q[c] = (int)((p[c] - g->origin[c]) / g->cell[c]);
        if (q[c] < 0 || q[c] >= g->shape[c])
            return -1; /* DenseGrid::at would throw */
    }
    return (q[0] * g->shape[1] + q[1]) * g->shape[2] + q[2];
}

/* Electrostatic field (potential grid) force and energy, float-faithful. */
static void potential_field_f(const orc *o, int64_t i, float *fi, float *ei)
{
    const ogrid *g = o->grid[BS_GRID_POTENTIAL];
    const float *x = o->xf + 3 * i;
    const int64_t c = grid_cell(g, x[0], x[1], x[2]);
    if (c < 0)
        return;
    const float s = o->p.potentialgrid_scale, q = o->charge[i];
    for (int a = 0; a < 3; ++a)
        fi[a] = fi[a] + (g->cf[4 * c + 1 + a] * q) * s;
    float e = g->cf[4 * c] * q;
    e = (float)(e * K_BOLTZ);
    e = (float)(e * K_J_TO_KJ);
    e = (float)(e * K_NA);
    *ei += s * e;
}

static void potential_field_d(const orc *o, int64_t i, const float *pos, double *fi, double *ei)
{
    const ogrid *g = o->grid[BS_GRID_POTENTIAL];
    const int64_t c = grid_cell(g, pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    if (c < 0)
        return;
    const double s = o->p.potentialgrid_scale, q = o->charge[i];
    for (int a = 0; a < 3; ++a)
        fi[a] += g->cd[4 * c + 1 + a] * q * s;
    *ei += s * g->cd[4 * c] * q * K_BOLTZ * K_J_TO_KJ * K_NA;
}

static void density_field_f(const orc *o, int64_t i, float *fi)
{
    const ogrid *g = o->grid[BS_GRID_DENSITY];
    const float *x = o->xf + 3 * i;
    const int64_t c = grid_cell(g, x[0], x[1], x[2]);
    if (c < 0)
        return;
    const float s = o->p.densitygrid_scale, bu = o->burying[i];
    for (int a = 0; a < 3; ++a)
        fi[a] = fi[a] + (g->cf[4 * c + 1 + a] * bu) * s;
}

static void density_field_d(const orc *o, int64_t i, const float *pos, double *fi)
{
    const ogrid *g = o->grid[BS_GRID_DENSITY];
    const int64_t c = grid_cell(g, pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    if (c < 0)
        return;
    const double s = o->p.densitygrid_scale, bu = o->burying[i];
    for (int a = 0; a < 3; ++a)
        fi[a] += g->cd[4 * c + 1 + a] * bu * s;
}

/* ---------------------------------------------------------------------------
 * Insertion vector, InsertionVector.cpp:18-78 with the centroid of every
 * particle (measure.hpp:121-140).  out6 = angle (deg), roll angle (deg),
 * insertion depth (centroid z), vector (3).
 * ------------------------------------------------------------------------- */
int orc_insertion(const orc *o, int mode, int64_t i1, int64_t i2, double *out6)
{
    if (i1 < 0 || i2 < 0 || i1 >= o->n || i2 >= o->n)
        return -1;
    double cs[3] = {0.0, 0.0, 0.0};
    for (int64_t i = 0; i < o->n; ++i)
        for (int c = 0; c < 3; ++c)
            cs[c] += mode == ORC_F32 ? (double)o->xf[3 * i + c] : o->xd[3 * i + c];
    const double f = o->n ? 1.0 / (double)o->n : 0.0;
    for (int c = 0; c < 3; ++c)
        cs[c] *= f;
    if (mode == ORC_F32) {
        const float *p1 = o->xf + 3 * i1, *p2 = o->xf + 3 * i2;
        float v[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
        float angle = acosf(-v[2] / norm_f(v[0], v[1], v[2]));
        angle = (float)(angle * (180.0 / M_PI));
        angle = (float)(angle - 90.0);
        const float bary[3] = {(float)cs[0], (float)cs[1], (float)cs[2]};
        const float n2sq = (float)pow(norm_f(v[0], v[1], v[2]), 2.0);
        float proj[3], n1[3], n3[3], ref[3], n2[3];
        for (int c = 0; c < 3; ++c) /* term-by-term products, as written */
            proj[c] = p1[c] + v[c] * (((bary[c] - p1[c]) * v[c]) / n2sq);
        for (int c = 0; c < 3; ++c)
            n1[c] = proj[c] - bary[c];
        normalize_f(n1);
        n3[0] = n1[1] * v[2] - n1[2] * v[1];
        n3[1] = n1[2] * v[0] - n1[0] * v[2];
        n3[2] = n1[0] * v[1] - n1[1] * v[0];
        normalize_f(n3);
        const float am = (float)(atan2f(n3[2], n1[2]) - M_PI);
        const float r = (float)sqrt(((double)bary[0] - proj[0]) * ((double)bary[0] - proj[0]) +
                                    ((double)bary[1] - proj[1]) * ((double)bary[1] - proj[1]) +
                                    ((double)bary[2] - proj[2]) * ((double)bary[2] - proj[2]));
        const float ca = cosf(am), sa = sinf(am);
        for (int c = 0; c < 3; ++c)
            ref[c] = (proj[c] + (n1[c] * r) * ca) + (n3[c] * r) * sa;
        for (int c = 0; c < 3; ++c)
            n2[c] = proj[c] - ref[c];
        normalize_f(n2);
        const float dot12 = n1[0] * n2[0] + n1[1] * n2[1] + n1[2] * n2[2];
        const float arad = acosf(dot12 / (norm_f(n1[0], n1[1], n1[2]) * norm_f(n2[0], n2[1], n2[2])));
        const float c12[3] = {n1[1] * n2[2] - n1[2] * n2[1], n1[2] * n2[0] - n1[0] * n2[2],
                              n1[0] * n2[1] - n1[1] * n2[0]};
        const float s = c12[0] * (p2[0] - proj[0]) + c12[1] * (p2[1] - proj[1]) + c12[2] * (p2[2] - proj[2]);
        const float adeg = (float)((arad * 180) / M_PI);
        out6[0] = angle;
        out6[1] = s > 0.0 ? 360 - adeg : adeg;
        out6[2] = (float)cs[2];
        for (int c = 0; c < 3; ++c)
            out6[3 + c] = v[c];
    } else {
        const double *p1 = o->xd + 3 * i1, *p2 = o->xd + 3 * i2;
        double v[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
        const double vn = norm_d(v[0], v[1], v[2]);
        out6[0] = acos(-v[2] / vn) * (180.0 / M_PI) - 90.0;
        double proj[3], n1[3], n3[3], ref[3], n2[3];
        for (int c = 0; c < 3; ++c)
            proj[c] = p1[c] + v[c] * (((cs[c] - p1[c]) * v[c]) / (vn * vn));
        for (int c = 0; c < 3; ++c)
            n1[c] = proj[c] - cs[c];
        normalize_d(n1);
        n3[0] = n1[1] * v[2] - n1[2] * v[1];
        n3[1] = n1[2] * v[0] - n1[0] * v[2];
        n3[2] = n1[0] * v[1] - n1[1] * v[0];
        normalize_d(n3);
        const double am = atan2(n3[2], n1[2]) - M_PI;
        const double r = norm_d(cs[0] - proj[0], cs[1] - proj[1], cs[2] - proj[2]);
        for (int c = 0; c < 3; ++c)
            ref[c] = proj[c] + n1[c] * r * cos(am) + n3[c] * r * sin(am);
        for (int c = 0; c < 3; ++c)
            n2[c] = proj[c] - ref[c];
        normalize_d(n2);
        const double dot12 = n1[0] * n2[0] + n1[1] * n2[1] + n1[2] * n2[2];
        const double arad = acos(dot12 / (norm_d(n1[0], n1[1], n1[2]) * norm_d(n2[0], n2[1], n2[2])));
        const double c12[3] = {n1[1] * n2[2] - n1[2] * n2[1], n1[2] * n2[0] - n1[0] * n2[2],
                               n1[0] * n2[1] - n1[1] * n2[0]};
        const double s = c12[0] * (p2[0] - proj[0]) + c12[1] * (p2[1] - proj[1]) + c12[2] * (p2[2] - proj[2]);
        const double adeg = arad * 180.0 / M_PI;
        out6[1] = s > 0.0 ? 360.0 - adeg : adeg;
        out6[2] = cs[2];
        for (int c = 0; c < 3; ++c)
            out6[3 + c] = v[c];
    }
    return 0;
}

/* ---------------------------------------------------------------------------
 * Integrator, SpringNetwork.cpp:202-241 and Particle.cpp:109-128.
 * ------------------------------------------------------------------------- */
static int integrate_f32(orc *o)
{
    const float dt = o->p.timestep;
    float ke = 0.0f;
    int64_t bad = -1;
    for (int64_t t = 0; t < o->ndyn; ++t) {
        const int64_t i = o->dyn[t];
        float *x = o->xf + 3 * i, *v = o->vf + 3 * i, *f = o->ff + 3 * i;
        const float m = o->mass[i];
        float kei = 0.0f; /* a rigid particle's kinetic energy is reset before it is summed */
        if (o->body_of[i] >= 0) {
            rigid_place_f(o, i);
        } else {
            if (m > 0.0f)
                for (int c = 0; c < 3; ++c)
                    v[c] = v[c] + (f[c] / m) * dt;
            const float vit = norm_f(v[0], v[1], v[2]);
            kei = (float)(0.5f * m * (vit * vit) * c_kinetic());
            for (int c = 0; c < 3; ++c)
                x[c] = x[c] + v[c] * dt;
        }
        if ((!isfinite(x[0]) || !isfinite(x[1]) || !isfinite(x[2])) && bad < 0)
            bad = i;
        ke += kei;
        f[0] = f[1] = f[2] = 0.0f;
    }
    o->ef[BS_E_KINETIC] += ke;
    o->nonfinite = bad;
    return bad < 0 ? BS_OK : BS_ERR_NONFINITE;
}

static int integrate_f64(orc *o)
{
    const double dt = o->p.timestep;
    double ke = 0.0;
    int64_t bad = -1;
    for (int64_t t = 0; t < o->ndyn; ++t) {
        const int64_t i = o->dyn[t];
        double *x = o->xd + 3 * i, *v = o->vd + 3 * i, *f = o->fd + 3 * i;
        const double m = o->mass[i];
        if (o->body_of[i] >= 0) {
            rigid_place_d(o, i);
        } else {
            if (m > 0.0)
                for (int c = 0; c < 3; ++c)
                    v[c] += f[c] / m * dt;
            ke += 0.5 * m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) * c_kinetic();
            for (int c = 0; c < 3; ++c)
                x[c] += v[c] * dt;
        }
        if ((!isfinite(x[0]) || !isfinite(x[1]) || !isfinite(x[2])) && bad < 0)
            bad = i;
        f[0] = f[1] = f[2] = 0.0;
    }
    o->ed[BS_E_KINETIC] += ke;
    o->nonfinite = bad;
    return bad < 0 ? BS_OK : BS_ERR_NONFINITE;
}

/* ---------------------------------------------------------------------------
 * Adjoint (no reference counterpart): gradients of the trajectory loss
 *   L = (1/n) sum_i |x_T,i - y_i|^2
 * after T FP64 steps with respect to spring stiffness k, rest length L0 and
 * particle charge q.  Reverse sweep of the semi-implicit Euler step
 * v' = v + dt/m (F(x) - gamma v), x' = x + dt v':
 *   a = lv + dt lx,  w = dt/m a,  lx += H(x) w,  lv = (1 - gamma dt/m) a,
 * with H = dF/dx (symmetric: the terms are conservative) and the parameter
 * gradients sum_i w_i . dF_i/dtheta.  Each pair term acts along d = x_j - x_i
 * with F_i = phi(r) d, phi = fm / r (fm: the force modules above), so
 *   H w |_i = phi dw + (phi'/r) (d . dw) d,  dw = w_j - w_i,  phi' = (fm' - phi)/r.
 * The pair set of each step is the forward's (frozen at each state).
 * Terms: springs, the four steric modes, Coulomb (both dielectric models),
 * viscosity.  IMPALA, the pseudo-hydrophobic term, grids and rigid bodies
 * are not differentiated (orc_adjoint returns NAN when they are on).
 * ------------------------------------------------------------------------- */
static double steric_dfm_d(int mode, double ri, double rj, double ei, double ej, double d)
{
    switch (mode) {
    case BS_STERIC_AMBER_12_6: {
        if (d < K_MIN_VDW)
            return 0.0;
        const double eij = sqrt(ei * ej), R = sqrt(ri * rj), s = R / d;
        const double s6 = s * s * s * s * s * s;
        return 12.0 * eij * (-7.0 * s6 + 13.0 * s6 * s6) / (d * d) * c_spring();
    }
    case BS_STERIC_LEWITT_8_6: {
        if (d < K_MIN_VDW)
            return 0.0;
        const double eij = sqrt(ei * ej), R = sqrt(ri * rj), s = R / d;
        const double s2 = s * s, s6 = s2 * s2 * s2;
        return 24.0 * eij * (-7.0 * s6 + 9.0 * s6 * s2) / (d * d) * c_spring();
    }
    case BS_STERIC_ZACHARIAS_8_6: {
        if (d < K_MIN_VDW)
            return 0.0;
        const double eij = ei * ej, s = ri * rj / d;
        const double s2 = s * s, s6 = s2 * s2 * s2;
        return eij * (-42.0 * s6 + 72.0 * s6 * s2) / (d * d) * c_spring();
    }
    default: {
        const double dv = d - (ri + rj);
        return dv > 0 ? 0.0 : c_spring();
    }
    }
}

static double elec_dfm_d(double q1, double q2, double d, double diel, int ddd)
{
    if (d < K_MIN_ELEC)
        return 0.0;
    const double kq = (q1 * q2) / (4.0 * K_PI * diel) * c_elec();
    return ddd ? 6.0 * kq / (d * d * d * d) : 2.0 * kq / (d * d * d);
}

/* One accepted pair of the adjoint sweep (see adjoint_terms). */
static void adjoint_pair(orc *o, int kind, int64_t i, int64_t j, const float *pos, float cutoff, const double *w,
                         double *hw, double *dq)
{
    const bs_params *p = &o->p;
    float dfl;
    if (!pair_accept(o, kind, i, j, pos, cutoff, &dfl))
        return;
    double d[3], dw[3];
    for (int c = 0; c < 3; ++c) {
        d[c] = o->xd[3 * j + c] - o->xd[3 * i + c];
        dw[c] = w[3 * j + c] - w[3 * i + c];
    }
    const double r = norm_d(d[0], d[1], d[2]);
    if (r == 0.0)
        return;
    double fm, dfm;
    if (kind == BS_NS_STERIC) {
        fm = p->steric_gridscale * steric_fm_d(p->steric_mode, o->radius[j], o->radius[i], o->eps[j], o->eps[i], r);
        dfm = p->steric_gridscale *
              steric_dfm_d(p->steric_mode, o->radius[j], o->radius[i], o->eps[j], o->eps[i], r);
    } else {
        fm = p->electrostatic_scale *
             elec_fm_d(o->charge[j], o->charge[i], r, p->electrostatic_dielectric, p->dielectric_model);
        dfm = p->electrostatic_scale *
              elec_dfm_d(o->charge[j], o->charge[i], r, p->electrostatic_dielectric, p->dielectric_model);
    }
    const double phi = fm / r, dphi = (dfm - phi) / r;
    const double dd = d[0] * dw[0] + d[1] * dw[1] + d[2] * dw[2];
    for (int c = 0; c < 3; ++c) {
        const double hv = phi * dw[c] + dphi / r * dd * d[c];
        hw[3 * i + c] += hv;
        if (!o->isstatic[j])
            hw[3 * j + c] -= hv;
    }
    if (kind == BS_NS_ELECTROSTATIC) {
        const double qq = (double)o->charge[i] * o->charge[j];
        if (qq != 0.0) {
            const double g = -(phi / qq) * dd; /* c (w_i - w_j) . d */
            dq[i] += g * o->charge[j];
            dq[j] += g * o->charge[i];
        }
    }
}

/* H w and the parameter gradients at the current F64 state; hw[3n] is added to. */
static void adjoint_terms(orc *o, const double *w, double *hw, double *dk, double *dl0, double *dq)
{
    const bs_params *p = &o->p;
    const int64_t n = o->n;
    if (p->spring_enable)
        for (int64_t t = 0; t < o->nds; ++t) {
            const int64_t s = o->dsp[t];
            const uint32_t a = o->s1[s], b = o->s2[s];
            double d[3], dw[3];
            for (int c = 0; c < 3; ++c) {
                d[c] = o->xd[3 * b + c] - o->xd[3 * a + c];
                dw[c] = w[3 * b + c] - w[3 * a + c];
            }
            const double r = norm_d(d[0], d[1], d[2]);
            if (r == 0.0)
                continue;
            const double C = c_spring() * p->spring_scale;
            const double fm = o->sk[s] * (r - o->sl0[s]) * C, dfm = o->sk[s] * C;
            const double phi = fm / r, dphi = (dfm - phi) / r;
            const double dd = d[0] * dw[0] + d[1] * dw[1] + d[2] * dw[2];
            for (int c = 0; c < 3; ++c) {
                const double hv = phi * dw[c] + dphi / r * dd * d[c];
                if (!o->isstatic[a])
                    hw[3 * a + c] += hv;
                if (!o->isstatic[b])
                    hw[3 * b + c] -= hv;
            }
            const double wu = -dd / r; /* (w_a - w_b) . d / r */
            dk[s] += wu * (r - o->sl0[s]) * C;
            dl0[s] += -wu * o->sk[s] * C;
        }
    const float *pos = decision_positions(o, ORC_F64);
    for (int kind = 0; kind < 2; ++kind) {
        if (!kind_enabled(o, kind))
            continue;
        int any = kind == BS_NS_STERIC;
        for (int64_t i = 0; i < n && !any; ++i)
            any = member_of(o, kind, i);
        if (!any)
            continue;
        cells_t cl;
        build_cells(o, kind, pos, &cl);
        for (int64_t t = 0; t < o->ndyn; ++t) {
            const int64_t i = o->dyn[t];
            if (kind == BS_NS_ELECTROSTATIC && !o->charged[i])
                continue;
            FOR_EACH_NEIGHBOR(&cl, pos, i, j, { adjoint_pair(o, kind, i, j, pos, cl.cutoff, w, hw, dq); });
        }
        free_cells(&cl);
    }
}

/* Runs nsteps FP64 steps from the current F64 state (which advances), then
 * the reverse sweep.  Returns the loss; dk, dl0 per spring (addition order),
 * dq per particle.  NAN if a term without an adjoint is enabled. */
double orc_adjoint(orc *o, int nsteps, const double *y3, double *dk, double *dl0, double *dq)
{
    const bs_params *p = &o->p;
    if (p->imp_enable || p->hydrophobicity_enable || o->nbodies > 0 || p->probe_enable ||
        (p->electrostatic_enable && p->potentialgrid_enable && o->grid[BS_GRID_POTENTIAL]) ||
        (p->densitygrid_enable && o->grid[BS_GRID_DENSITY]))
        return NAN;
    const int64_t n = o->n;
    double *xs = (double *)malloc(sizeof(double) * 3 * n * (size_t)(nsteps + 1));
    memcpy(xs, o->xd, sizeof(double) * 3 * n);
    for (int s = 0; s < nsteps; ++s) {
        forces_f64(o, NULL);
        integrate_f64(o);
        memcpy(xs + (size_t)(s + 1) * 3 * n, o->xd, sizeof(double) * 3 * n);
    }
    double loss = 0.0;
    for (int64_t k = 0; k < 3 * n; ++k) {
        const double e = o->xd[k] - y3[k];
        loss += e * e;
    }
    loss /= (double)n;
    double *lx = (double *)calloc(3 * n, sizeof(double)), *lv = (double *)calloc(3 * n, sizeof(double));
    double *w = (double *)calloc(3 * n, sizeof(double)), *hw = (double *)calloc(3 * n, sizeof(double));
    for (int64_t i = 0; i < n; ++i)
        if (!o->isstatic[i])
            for (int c = 0; c < 3; ++c)
                lx[3 * i + c] = 2.0 * (o->xd[3 * i + c] - y3[3 * i + c]) / (double)n;
    memset(dk, 0, sizeof(double) * o->ns);
    memset(dl0, 0, sizeof(double) * o->ns);
    memset(dq, 0, sizeof(double) * n);
    const double dt = p->timestep, gam = p->viscosity_enable ? p->viscosity_value : 0.0;
    for (int s = nsteps - 1; s >= 0; --s) {
        for (int64_t i = 0; i < n; ++i) {
            const double m = o->mass[i];
            for (int c = 0; c < 3; ++c) {
                const double a = lv[3 * i + c] + dt * lx[3 * i + c];
                w[3 * i + c] = (!o->isstatic[i] && m > 0.0) ? dt / m * a : 0.0;
                lv[3 * i + c] = a; /* finished below */
                hw[3 * i + c] = 0.0;
            }
        }
        memcpy(o->xd, xs + (size_t)s * 3 * n, sizeof(double) * 3 * n);
        adjoint_terms(o, w, hw, dk, dl0, dq);
        for (int64_t i = 0; i < n; ++i) {
            if (o->isstatic[i])
                continue;
            const double m = o->mass[i];
            for (int c = 0; c < 3; ++c) {
                lx[3 * i + c] += hw[3 * i + c];
                if (m > 0.0)
                    lv[3 * i + c] *= 1.0 - gam * dt / m;
            }
        }
    }
    memcpy(o->xd, xs + (size_t)nsteps * 3 * n, sizeof(double) * 3 * n);
    free(xs);
    free(lx);
    free(lv);
    free(w);
    free(hw);
    return loss;
}

/* Loss of nsteps FP64 steps from the given positions/velocities (the state
 * advances), for finite differences. */
double orc_loss(orc *o, int nsteps, const double *y3)
{
    if (o->p.probe_enable)
        return NAN;
    for (int s = 0; s < nsteps; ++s) {
        forces_f64(o, NULL);
        integrate_f64(o);
    }
    double loss = 0.0;
    for (int64_t k = 0; k < 3 * o->n; ++k) {
        const double e = o->xd[k] - y3[k];
        loss += e * e;
    }
    return loss / (double)o->n;
}

/* Replaces spring k / L0 (addition order of the deduplicated springs) and
 * charges (nullable each); charged flags are kept. */
void orc_set_params_arrays(orc *o, const float *k, const float *l0, const float *q)
{
    for (int64_t s = 0; s < o->ns; ++s) {
        if (k)
            o->sk[s] = k[s];
        if (l0)
            o->sl0[s] = l0[s];
    }
    if (q)
        for (int64_t i = 0; i < o->n; ++i)
            o->charge[i] = q[i];
}

/* H w at the current F64 state (forces' Jacobian times w), for tests. */
void orc_hessian_vec(orc *o, const double *w, double *hw)
{
    double *dk = (double *)calloc(o->ns ? o->ns : 1, sizeof(double));
    double *dl = (double *)calloc(o->ns ? o->ns : 1, sizeof(double));
    double *dq = (double *)calloc(o->n, sizeof(double));
    memset(hw, 0, sizeof(double) * 3 * o->n);
    adjoint_terms(o, w, hw, dk, dl, dq);
    free(dk);
    free(dl);
    free(dq);
}

/* ---------------------------------------------------------------------------
 * Public entry points.
 * ------------------------------------------------------------------------- */

/* Forces at the current state of `mode` into F3 (double, 3n) and energies
 * into E6 (kinetic = 0).  ext3 nullable. */
void orc_compute_forces(orc *o, int mode, const double *ext3, double *F3, double *E6)
{
    if (mode == ORC_F32) {
        float *ef = NULL;
        if (ext3) {
            ef = (float *)malloc(sizeof(float) * 3 * o->n);
            for (int64_t i = 0; i < 3 * o->n; ++i)
                ef[i] = (float)ext3[i];
        }
        forces_f32(o, ef);
        free(ef);
        for (int64_t i = 0; i < 3 * o->n; ++i)
            F3[i] = o->isstatic[i / 3] ? 0.0 : o->ff[i];
        for (int k = 0; k < BS_NUM_ENERGIES; ++k)
            E6[k] = o->ef[k];
    } else {
        forces_f64(o, ext3);
        for (int64_t i = 0; i < 3 * o->n; ++i)
            F3[i] = o->isstatic[i / 3] ? 0.0 : o->fd[i];
        for (int k = 0; k < BS_NUM_ENERGIES; ++k)
            E6[k] = o->ed[k];
    }
}

/* nsteps of SpringNetwork::computeStep.
 * Energies of the last step go to E6 (nullable). */
int orc_step(orc *o, int mode, int nsteps, const double *ext3, double *E6)
{
    int rc = BS_OK;
    float *ef = NULL;
    if (ext3 && mode == ORC_F32) {
        ef = (float *)malloc(sizeof(float) * 3 * o->n);
        for (int64_t i = 0; i < 3 * o->n; ++i)
            ef[i] = (float)ext3[i];
    }
    for (int s = 0; s < nsteps && rc == BS_OK; ++s) {
        if (mode == ORC_F32) {
            forces_f32(o, ef);
            if (o->p.probe_enable)
                probe_integrate_f(o);
            if (o->p.rigidbody_enable) {
                rigid_aggregate_f(o);
                rigid_solve_f(o);
            }
            rc = integrate_f32(o);
        } else {
            forces_f64(o, ext3);
            if (o->p.probe_enable)
                probe_integrate_d(o);
            if (o->p.rigidbody_enable) {
                rigid_aggregate_d(o);
                rigid_solve_d(o);
            }
            rc = integrate_f64(o);
        }
    }
    free(ef);
    if (E6)
        for (int k = 0; k < BS_NUM_ENERGIES; ++k)
            E6[k] = mode == ORC_F32 ? (double)o->ef[k] : o->ed[k];
    return rc;
}

int64_t orc_last_nonfinite(const orc *o) { return o->nonfinite; }

int orc_num_threads(void) { return max_threads(); }

/* ---------------------------------------------------------------------------
 * Spring network from a distance cutoff: topology::Topology::
 * add_springs_from_cutoff (src/topology/Topology.hpp:105-138).  One spring per
 * pair of particles closer than the cutoff, unless both are static
 * (Topology.hpp:121-122).  The test is NeighborSearch's
 * (src/nsearch.hpp:230): measure::distance (src/measure.hpp:162-168, the FP64
 * distance of the float positions) < the cutoff held as a float
 * (nsearch.hpp:57); the rest length is that distance (Spring.hpp:55-56),
 * kept as a float by SpringNetwork::addSpring.  Each pair is found once, from
 * its lower index, so the reference's SpringAlreadyExistsException path
 * never arises here.
 *
 * Written as a plain grid of cells of edge >= the cutoff scanned over the 27
 * cells around each particle; rows are sorted by the second index, so the
 * output is ordered by (first, second).  With groups (nullable), only pairs
 * of different groups are linked: add_springs_between_topologies_from_cutoff
 * (Topology.hpp:140-180), groups being the topology ids.  Returns the number
 * of springs and,
 * when id1 is not NULL, writes the rows that fit in cap entries (all of them
 * when cap >= that number); -1 for a cutoff the reference rejects
 * (nsearch.hpp:66-67).
 * ------------------------------------------------------------------------- */
static int cmp_row_j(const void *a, const void *b)
{
    const uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return (x > y) - (x < y);
}

int64_t orc_springs_from_cutoff(int64_t n, const float *x3, const uint8_t *flags, const int32_t *groups,
                                double cutoff, int64_t cap, uint32_t *id1, uint32_t *id2, float *l0)
{
    const float rcf = (float)cutoff;
    if (!(rcf >= 1e-6f) || !isfinite(rcf))
        return -1;
    if (n < 2)
        return 0;
    const double rc = (double)rcf;
    double lo[3], hi[3];
    for (int q = 0; q < 3; ++q)
        lo[q] = hi[q] = x3[q];
    for (int64_t i = 1; i < n; ++i)
        for (int q = 0; q < 3; ++q)
        {
            const double v = x3[3 * i + q];
            lo[q] = v < lo[q] ? v : lo[q];
            hi[q] = v > hi[q] ? v : hi[q];
        }
    /* the cell edge is 1.001 x the cutoff: rounding of the cell coordinate
     * can then never put two particles closer than the cutoff two cells apart */
    const double edge = rc * 1.001;
    int64_t dims[3], ncell = 1;
    for (int q = 0; q < 3; ++q)
    {
        double d = floor((hi[q] - lo[q]) / edge);
        dims[q] = d < 1.0 ? 1 : (d > 128.0 ? 128 : (int64_t)d);
        ncell *= dims[q];
    }
    int64_t *cell = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    int64_t *start = (int64_t *)calloc((size_t)ncell + 1, sizeof(int64_t));
    int64_t *member = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    for (int64_t i = 0; i < n; ++i)
    {
        int64_t c[3];
        for (int q = 0; q < 3; ++q)
        {
            const double e = fmax(edge, (hi[q] - lo[q]) / (double)dims[q]);
            const double u = ((double)x3[3 * i + q] - lo[q]) / e;
            c[q] = u <= 0.0 ? 0 : (u >= (double)(dims[q] - 1) ? dims[q] - 1 : (int64_t)u);
        }
        cell[i] = c[0] + dims[0] * (c[1] + dims[1] * c[2]);
        start[cell[i] + 1]++;
    }
    for (int64_t c = 0; c < ncell; ++c)
        start[c + 1] += start[c];
    {
        int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (size_t)ncell);
        memcpy(fill, start, sizeof(int64_t) * (size_t)ncell);
        for (int64_t i = 0; i < n; ++i)
            member[fill[cell[i]]++] = i;
        free(fill);
    }
    struct
    {
        uint32_t j;
        float l;
    } *row = NULL;
    int64_t rowcap = 0, m = 0;
    for (int64_t i = 0; i < n; ++i)
    {
        const int si = flags && (flags[i] & BS_FLAG_STATIC);
        const int64_t ci = cell[i];
        const int64_t c0 = ci % dims[0], c1 = (ci / dims[0]) % dims[1], c2 = ci / (dims[0] * dims[1]);
        int64_t r = 0;
        for (int64_t z = c2 - 1; z <= c2 + 1; ++z)
            for (int64_t y = c1 - 1; y <= c1 + 1; ++y)
                for (int64_t x = c0 - 1; x <= c0 + 1; ++x)
                {
                    if (x < 0 || y < 0 || z < 0 || x >= dims[0] || y >= dims[1] || z >= dims[2])
                        continue;
                    const int64_t cc = x + dims[0] * (y + dims[1] * z);
                    for (int64_t t = start[cc]; t < start[cc + 1]; ++t)
                    {
                        const int64_t j = member[t];
                        if (j <= i || (si && flags && (flags[j] & BS_FLAG_STATIC)))
                            continue;
                        if (groups && groups[i] == groups[j])
                            continue;
                        const double dx = (double)x3[3 * i] - (double)x3[3 * j];
                        const double dy = (double)x3[3 * i + 1] - (double)x3[3 * j + 1];
                        const double dz = (double)x3[3 * i + 2] - (double)x3[3 * j + 2];
                        const double d = sqrt(dx * dx + dy * dy + dz * dz);
                        if (!(d < rc))
                            continue;
                        if (r == rowcap)
                        {
                            rowcap = rowcap ? 2 * rowcap : 256;
                            row = realloc(row, sizeof(*row) * (size_t)rowcap);
                        }
                        row[r].j = (uint32_t)j;
                        row[r].l = (float)d;
                        ++r;
                    }
                }
        qsort(row, (size_t)r, sizeof(*row), cmp_row_j);
        if (cap >= 0 && id1 && m + r <= cap)
            for (int64_t t = 0; t < r; ++t)
            {
                id1[m + t] = (uint32_t)i;
                id2[m + t] = row[t].j;
                l0[m + t] = row[t].l;
            }
        m += r;
    }
    free(row);
    free(member);
    free(start);
    free(cell);
    return m;
}
