// bs_engine.hpp -- device data layout and kernel launchers of the engine.
//
// Layout in HBM (all arrays in the engine's internal particle order, a
// spatial Morton order fixed at setup; perm[k] = original id of internal k):
//   pos[2]   float4[N]  x, y, z, charge           ping-pong between steps
//   plo[2]   float4[N]  low words of x, y, z      compensated positions only
//   vel      float4[N]  vx, vy, vz, 0
//   fnb      double4[N] non-bonded force of the current step (FP64 sums)
//   jprm     float4[N]  sqrt(radius), sqrt(eps), hydrophobicity, flags (bits)
//   iprm     float4[N]  mass, sasa, transfer, burying
//   xref     float4[N]  positions at the last neighbour-list build
//   springs  CSR over particles, both directions, each row sorted by j:
//            sp_ptr int32[N+1], sp_j int32[E] (bit31 = j static), sp_k float[E],
//            sp_l0 float[E]
//   lists    per i-cluster of 8 particles: cl[c * cap + t] = j | mask << 24,
//            cln[c] entries, zero words up to the next multiple of 16
//   ex       static exclusions per i-cluster: ex_ptr[c], ex_w = j << 8 | mask
//   grids    per field, float4 per cell: gradient force vector (x, y, z) and
//            the scalar, x-major with z fastest (DenseGrid::_data[i][j][k])
//   rigid    per dynamic rigid particle ("slot", grouped by body in body
//            order): slot_p int32, slot_b int32, p0 double4 (body frame),
//            rx double4 (FP64 position); bodies RigidBody[nb]; frig double4[N]
//            total force of rigid particles in the current step
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace bs
{

// Particle flag bits kept in jprm.w (as raw int bits).
constexpr uint32_t PF_STATIC = 1u;
constexpr uint32_t PF_CHARGED = 2u;
constexpr uint32_t PF_HYDRO = 4u;
constexpr uint32_t PF_RIGID = 8u;

// Term bits of a neighbour list.
constexpr int TERM_STERIC = 1;
constexpr int TERM_ELEC = 2;
constexpr int TERM_HYDRO = 4;

// Subset filters.
constexpr uint32_t SUB_ALL = 0u;

struct Consts
{
    double c_spring;  // kJ/mol/A -> Da A fs^-2
    double c_elec;    // e^2/A^2 -> Da A fs^-2 (before 1/(4 pi eps))
    double c_kinetic; // Da A^2 fs^-2 -> kJ/mol
    double c_coulomb; // 1/(4 pi eps0)
    double e_coul;    // Coulomb energy prefactor for q1 q2 / d (d in A), kJ/mol
    double pi;
};

// Unit constants evaluated in double in the order of
// src/forcefield/constants.hpp:36-118.
inline Consts make_consts()
{
    const double kj_to_j = 1.0E3, j_to_kj = 1.0E-3, e2c = 1.60217653E-19, a2m = 1.0E-10, m2a = 1.0E10,
                 da2kg = 1.66054E-27, s2fs = 1.0E15, fs2s = 1.0E-15, pi = 3.1415927, eps0 = 8.854187817E-12,
                 na = 6.02214169E23;
    const double newton = m2a * (1.0 / da2kg) * fs2s * fs2s;
    Consts c;
    c.c_spring = kj_to_j / na * m2a * newton;
    c.c_elec = newton * (e2c * e2c) / (eps0 * a2m * a2m);
    c.c_kinetic = (a2m * s2fs) * (a2m * s2fs) * na * da2kg * j_to_kj;
    c.c_coulomb = 1.0 / (4.0 * pi * eps0);
    c.e_coul = c.c_coulomb * (e2c * e2c) / a2m * na * 1.0E-3;
    c.pi = pi;
    return c;
}

// Neighbour lists are per i-cluster (8 consecutive internal particles): entry
// t of cluster c is cl[c * cap + t] = j | (m << 24), where bit k of the 8-bit
// m marks the pair (8c + k, j) as within the list radius at the last build,
// valid and not spring-bonded.
constexpr int32_t CLUSTER = 8;
constexpr uint32_t CL_JMASK = 0x00FFFFFFu;

// The queue layout reads a cluster's `kept` entries (in search order) in
// nch = ceil(kept / QCHUNK) chunks of nearly equal size: chunk c holds
// q + (c < r) entries from offset c q + min(c, r) (q = kept / nch,
// r = kept mod nch).
constexpr int32_t QCHUNK = 128;
struct Chunks
{
    int32_t nch, q, r;
};
constexpr Chunks chunks_of(int32_t kept) // constexpr: callable on host and device
{
    Chunks d{0, 0, 0};
    d.nch = kept > 0 ? (kept + QCHUNK - 1) / QCHUNK : 0;
    d.q = kept > 0 ? kept / d.nch : 0;
    d.r = kept > 0 ? kept - d.q * d.nch : 0;
    return d;
}

// Parameters of one pair kernel launch.
struct PairArgs
{
    const float4 *pos;
    const float4 *jprm;
    const uint32_t *cl;  // [ncl * cap], zero-padded to a multiple of 16 entries
    const int32_t *cln;  // [ncl]
    double4 *fnb;       // non-bonded force, FP64
    double4 *epp;       // [N] per-particle steric, elec, hydro energy (energy steps)
    const uint32_t *rows; // nullable: per-particle lists (row layout), else the cluster lists
    const int32_t *rlen;  // [64 * row blocks] entries per particle
    int32_t rcap;         // entries per particle row
    int32_t queues;       // cluster lists read through per-particle LDS queues (k_pair_q)
    int32_t n, ncl, cap;
    int32_t accumulate; // add to fnb instead of overwriting
    int32_t energy;     // also write epp (the last step of a call only)
    const int32_t *status; // nullable: the call's status words (the kernel does nothing once halted)
    float rc, rc2lo, rc2hi;
    int32_t steric_mode;
    float steric_scale;
    float elec_scale, diel;
    double ke, kf;      // Coulomb energy / force prefactors incl. scale and 1/eps
    float hydro_scale;
    uint32_t i_mask, j_mask; // required flag bits (PF_CHARGED / PF_HYDRO) or 0
    Consts k;
};

// Rigid body state (RigidBody.h:156-183) in FP64; q = (x, y, z, w).
struct RigidBody
{
    double M;
    double iinv[9];
    double X[3], v[3], L[3], w[3], q[4];
};

// A field grid as the kernels read it (PotentialGrid after compute_gradient).
struct GridView
{
    const float4 *cells; // nullptr: field off
    double origin[3], bmax[3], cell[3];
    int32_t shape[3];
    float scale;
};

struct RigidArgs
{
    const float4 *pos;      // positions at the start of the step (charge in .w)
    float4 *pos_out;
    float4 *vel;
    const double4 *frig;    // [N] total force of each rigid particle
    double4 *rx;            // [nr] FP64 position of each slot
    const double4 *p0;      // [nr] offset in the body frame
    const int32_t *slot_p;  // [nr] internal particle index
    const int32_t *slot_b;  // [nr] body index
    const int32_t *bstart;  // [nb + 1] slot range of each body
    RigidBody *bodies;      // [nb]
    const float4 *xref;
    const int32_t *perm;
    int32_t *status;
    int32_t nr, nb;
    int32_t solve;          // rigidbody.enable: aggregate and solve before placing
    double dt;
    float half_skin2;
};

// The probe particle (SpringNetwork.cpp:156-182,749-766) on the device.
struct ProbeDev
{
    float4 x;     // position, charge
    float4 v;     // velocity
    double4 f;    // force of the last evaluation (its viscosity included)
    double e[4];  // steric, Coulomb, kinetic energy of the last evaluation
};

// The probe's interactions as the spring kernel's epilogue evaluates them:
// one pair with every dynamic particle, no cutoff, no spring exclusion.
struct ProbeArgs
{
    ProbeDev *state;       // nullptr: no probe
    double *part;          // [spring blocks * 8] block sums: force on the probe (x, y, z), steric E, Coulomb E
    int32_t steric, elec;  // probe terms on
    int32_t steric_mode;
    float steric_scale;
    float radius, eps;     // probe radius and epsilon in jprm's form (square roots for Amber / Lewitt)
    float mass;
    double ke, kf;         // Coulomb energy / force prefactors incl. scale and 1/dielectric
    int32_t ddd;           // eps(r) = eps r
    const int32_t *status; // the call's status words
};

struct SpringArgs
{
    const float4 *pos;
    const float4 *plo;   // nullable
    float4 *pos_out;     // integrate mode
    float4 *plo_out;     // nullable
    float4 *vel;
    const double4 *fnb;  // nullable (no pair terms)
    const float4 *fext;  // nullable
    const float4 *iprm;
    const float4 *jprm;
    const float4 *xref;
    const int32_t *sp_ptr;
    const int32_t *sp_j;
    const float *sp_k;   // null: every spring has stiffness k_uni (not read per edge)
    const float *sp_l0;
    float k_uni;
    const int32_t *perm;
    float4 *fout;        // forces-only mode
    double4 *frig;       // integrate mode: total force of rigid particles (not integrated here)
    GridView gpot, gden; // electrostatic potential field, density field
    double e_field;      // potential * charge -> kJ/mol (electrostatic_field.hpp:15-23)
    double4 *eps;        // [N] per-particle spring, kinetic, imp, field energy (energy steps)
    int32_t *status;     // the call's status words (bs_device.hpp)
    int32_t *steps_done; // integrate mode: incremented once per completed step
    int32_t n;
    int32_t spring_enable;
    float spring_scale;
    float dt;
    int32_t visc_enable;
    float visc;
    int32_t imp_enable;
    float imp_scale, imp_uo, imp_lo, imp_uc, imp_lc;
    float half_skin2;    // (list_skin/2)^2, conservative
    int32_t integrate;
    int32_t energy;      // also write eps (the last step of a call only)
    ProbeArgs probe;
    Consts k;
};

// Adjoint of the pair terms at one state: H w on i (each pair seen from both
// sides) and dL/dq.  Terms in FP32, sums in FP64.
struct AdjPairArgs
{
    const float4 *pos;
    const float4 *jprm;
    const uint32_t *cl;
    const int32_t *cln;
    const float4 *w;      // [N] adjoint weights dt/m a, FP32 (0 for static particles)
    double4 *hw;          // [N] H w of this list (overwritten, or added with accumulate)
    double *dq;           // [N] dL/dq accumulator
    int32_t n, ncl, cap, accumulate;
    int32_t queues;       // the lists read in chunks through per-particle LDS queues (k_adj_pair_q)
    float rc, rc2lo, rc2hi;
    float steric_scale;
    double kf;            // Coulomb force prefactor incl. scale and 1/eps
    uint32_t i_mask, j_mask;
    Consts k;
};

// Spring part of H w, the spring-parameter gradients and the reverse Euler
// update of the adjoint state, per particle row.
struct AdjSpringArgs
{
    const float4 *pos;
    const float4 *jprm;
    const float4 *iprm;
    const int32_t *sp_ptr;
    const int32_t *sp_j;
    const float *sp_k;    // null: every spring has stiffness k_uni
    float k_uni;
    const float *sp_l0;
    const float4 *w;      // [N] weights of this step (FP32)
    const double4 *hwp;   // nullable: pair part of H w
    double4 *lx, *lv;     // [N] adjoint of x and v (updated in place)
    float4 *w_next;       // [N] weights of the step before
    float *dk, *dl0;      // nullable: [slots] gradient sums since the last flush, per spring slot
    const int32_t *slot_ptr; // [N + 1] first slot of each row: a dynamic row's slots are its owned edges
                             // (partner static or of higher index), numbered in edge order
    int32_t n, spring_enable;
    float spring_scale;
    double dt, gamma;     // gamma = 0 without viscosity
    Consts k;
};

// Springs from a cutoff: count pass (offs == nullptr: counts[i] = springs of
// row i) or fill pass (row i at offs[i], sorted by j).
struct TopoArgs
{
    const float4 *pos;        // original order; .w = static flag bits
    const double *box;        // [6]
    const int32_t *cstart;    // [ncells + 1]
    const int32_t *cidx;      // members sorted by index within each cell
    const int32_t *offs;      // nullable: fill pass, row starts [n + 1]
    int32_t *counts;          // count pass: [n]
    int32_t *jout;            // fill pass: [springs]
    float *l0out;
    int32_t n;
    int32_t dims[3];
    float edge;               // minimum cell edge (>= the cutoff)
    double rc;                // the cutoff as the reference compares it: (double)(float)cutoff
    double r2max;             // smallest double whose correctly rounded sqrt is >= rc
    const int32_t *ex_off;    // nullable: springs already loaded, as rows of
    const int32_t *ex_j;      //   their lower index [n + 1], partners sorted
    int32_t max_row;          // fill pass: the longest row
    const int32_t *group;     // nullable: only pairs of different groups
};

struct ListArgs
{
    const float4 *pos;
    const float4 *shift;       // nullable: per-particle offset of its system for binning
    const float4 *jprm;
    const int32_t *cell_start; // [ncells + 1]
    const int32_t *cell_idx;   // sorted members
    const double *box;         // [6] current box (device)
    const int32_t *ex_ptr;     // [ncl + 1] static exclusions per i-cluster
    const uint32_t *ex_w;      // j << 8 | bonded mask, sorted by j
    uint32_t *cl;              // [ncl * cap] search order
    int32_t *cln;              // [ncl]
    int32_t *status;           // [2] overflow, [3] max count
    int32_t n, ncl, cap;
    int32_t dims[3];
    float cell_edge;           // minimum cell edge of the binning grid
    float radius;              // list radius
    float r2list;
    uint32_t i_mask, j_mask;
    int32_t exclude_bonded;
    const int32_t *gate;       // nullable: skip unless *gate != 0
};

// ---- launchers (bs_kernels.hip) ----
// shift (nullable): per-particle offset added to positions for binning, which
// keeps independent systems of a batch apart on the engine's grid.
void launch_box_reduce(const float4 *pos, const float4 *shift, int32_t n, double *partial, double *box,
                       int32_t nblocks, const int32_t *gate, hipStream_t s);
// box_part (nullable): the box is first reduced from box_nb partials (of
// launch_prestep) and stored in box; otherwise box is read.  xref (nullable): positions copied
// there as well.
void launch_cell_ids(const float4 *pos, const float4 *shift, const float4 *jprm, uint32_t mask, int32_t n,
                     double *box, const double *box_part, int32_t box_nb, float4 *xref, int32_t dims0,
                     int32_t dims1, int32_t dims2, double edge_exact, float min_edge, int32_t exact_mode,
                     int32_t *cell_of, int32_t *counts, const int32_t *gate, hipStream_t s);
void launch_scan(const int32_t *counts, int32_t *start, int32_t ncells, const int32_t *gate, hipStream_t s);
// drain: cursor is the counts array, left at zero; otherwise a zeroed cursor.
void launch_scatter(const int32_t *cell_of, int32_t n, const int32_t *start, int32_t *cursor, int32_t drain,
                    int32_t *tmp_idx, const int32_t *gate, hipStream_t s);
void launch_cell_sort(const int32_t *cell_of, int32_t n, const int32_t *start, const int32_t *tmp_idx,
                      const int32_t *key, int32_t *out_idx, const int32_t *gate, hipStream_t s);
void launch_list_build(const ListArgs &a, hipStream_t s);
// Row layout: the cluster lists expanded into one list per particle.  Blocks
// of 64 consecutive particles; entry t of particle i is
// rows[((i / 64) * rcap + t) * 64 + i % 64], in the order of the cluster list;
// rlen[i] entries (rcap >= cap).
void launch_rows(const uint32_t *cl, const int32_t *cln, int32_t ncl, int32_t cap, uint32_t *rows, int32_t *rlen,
                 int32_t rcap, const int32_t *gate, hipStream_t s);
void launch_pair(const PairArgs &a, int32_t elec_mode, int32_t hydro, int32_t nblocks, hipStream_t s);
void launch_spring(const SpringArgs &a, int32_t group, int32_t nblocks, hipStream_t s);
// Per-system energies from the per-particle ones: segments of contiguous
// particles (never crossing a system) are summed by one workgroup each, then
// every system's segments in order; out_sys[nsys * 6], out_total[6].
void launch_energies(const double4 *epp, const double4 *eps, const int32_t *seg_start, int32_t nseg,
                     const int32_t *sys_seg, int32_t nsys, double *seg_part, double *out_sys, double *out_total,
                     const ProbeDev *probe, hipStream_t s);
// The probe's force from the spring kernel's nparts block sums, its
// viscosity and (integrate) its Euler step and kinetic energy.
void launch_probe(const ProbeArgs &p, int32_t nparts, float dt, int32_t visc_enable, float visc, double c_kinetic,
                  int32_t integrate, hipStream_t s);
// Status words: [0] rebuild gate of the current step's list kernels, [1]
// lowest particle with a non-finite position (INT_MAX: none), [2] list
// overflow, [3] longest list, [4] the adjoint sweep's drift test of a
// segment's stored states, [5] stop, [6] steps completed in the pass, [7]
// drift flag (a rebuild at the next gate).  launch_prestep sets the gate from
// the drift flag and, at the start of a step (advance), from the rebuild
// interval; counters[0] counts steps, counters[1] rebuilds.  partial
// (nullable): it also writes nblocks (<= BLOCK) box partials of pos (+ shift)
// for launch_cell_ids, ungated.  log (nullable): the gate is also written
// there.
void launch_prestep(int32_t *status, int32_t *counters, int32_t interval, int32_t advance, const float4 *pos,
                    const float4 *shift, int32_t n, double *partial, int32_t nblocks, int32_t *log, hipStream_t s);
void launch_rigid(const RigidArgs &a, hipStream_t s);
void launch_adj_pair(const AdjPairArgs &a, int32_t steric_mode, int32_t elec_mode, int32_t nblocks, hipStream_t s);
void launch_adj_spring(const AdjSpringArgs &a, int32_t group, int32_t nblocks, hipStream_t s);
// lx = 2 (x - y) / n for dynamic particles, lv = 0, w = dt/m (dt lx); loss
// partials |x - y|^2 per block (nblocks of BLOCK).
void launch_adj_init(const float4 *pos, const float4 *target, const float4 *jprm, const float4 *iprm, int32_t n,
                     double dt, double4 *lx, double4 *lv, float4 *w, double *loss_part, int32_t nblocks,
                     hipStream_t s);
// A caller's seed (dL/dx_T, dL/dv_T in particle order, either null) into lx,
// lv and the first weights w, in internal order (perm: internal -> particle).
void launch_adj_seed(const double *dx3, const double *dv3, const int32_t *perm, const float4 *jprm,
                     const float4 *iprm, int32_t n, double dt, double4 *lx, double4 *lv, float4 *w, hipStream_t s);
// out[perm[t]] = in[t] for t < n.
void launch_unperm_f64(const double *in, const int32_t *perm, int32_t n, double *out, hipStream_t s);
// Spring-gradient flush: for each slot r (springs in the order of their
// owning CSR edge) dk[r] += dke[r] in FP64 and the FP32 sum is cleared;
// likewise dl0.
void launch_spring_grads(float *dke, float *dle, int64_t nslots, double *dk, double *dl0, hipStream_t s);
// Per spring from per slot: out[s] = slot_of[s] >= 0 ? v[slot_of[s]] : 0.
void launch_spring_out(const int32_t *slot_of, int64_t m, const double *dk, const double *dl0, double *outk,
                       double *outl, hipStream_t s);
void launch_topo_springs(const TopoArgs &a, hipStream_t s);
// Spring CSR on the device from the spring table (sa, sb: external ids; m
// springs; iperm: external -> internal).  Count pass (cnt zeroed: both ends
// of each spring), fill pass (cursor zeroed; 64-bit keys partner << 32 |
// spring << 1 | side at the rows' offsets ptr), then the rows sorted by key
// and written as partner (bit 31: static partner), stiffness, rest length.
void launch_edge_count(const uint32_t *sa, const uint32_t *sb, int64_t m, const int32_t *iperm, int32_t *cnt,
                       hipStream_t s);
void launch_edge_fill(const uint32_t *sa, const uint32_t *sb, int64_t m, const int32_t *iperm, const int32_t *ptr,
                      int32_t *cursor, uint64_t *key, hipStream_t s);
void launch_edge_sort(const int32_t *ptr, int32_t n, const uint64_t *key, const float4 *jprm, const float *sl,
                      const float *sk, int32_t *ej, float *ek, float *el, hipStream_t s);
// Exclusions of each i-cluster from its rows: merged (j << 8 | row bit, in
// j order) into merged[ptr[8c] ..) with the count of distinct j per cluster,
// then the distinct words (row bits of equal j or-ed) at ex_ptr[c].
void launch_ex_merge(const int32_t *ptr, const int32_t *ej, int32_t n, int32_t ncl, uint32_t *merged,
                     int32_t *ucount, hipStream_t s);
void launch_ex_fill(const int32_t *ptr, int32_t n, int32_t ncl, const uint32_t *merged, const int32_t *ex_ptr,
                    uint32_t *ex_w, hipStream_t s);
// Rows without the edges joining two rigid members (rigid[i] != 0): fptr
// null counts them into cnt; otherwise they are copied to fptr's offsets.
void launch_edge_keep(const int32_t *ptr, int32_t n, const int32_t *ej, const float *ek, const float *el,
                      const uint8_t *rigid, const int32_t *fptr, int32_t *cnt, int32_t *fej, float *fek, float *fel,
                      hipStream_t s);
// ida[t] = i for t in [offs[i], offs[i + 1]).
void launch_rows_expand(const int32_t *offs, int32_t n, uint32_t *ida, hipStream_t s);
void launch_pack_positions(const float4 *pos, const float4 *plo, const int32_t *perm, int32_t n, float *out3,
                           hipStream_t s);
// flag = 1 when some v[s] differs from v[0] (bitwise), s < m.
void launch_differs_f32(const float *v, int64_t m, int32_t *flag, hipStream_t s);
// flag = 1 when a particle of any of nstates stored states (states[k * n + i])
// is more than half the list skin from xref.
void launch_drift_states(const float4 *states, int32_t nstates, int32_t n, const float4 *xref, float half_skin2,
                         int32_t *flag, hipStream_t s);
// status[7] (the drift flag) = 1 when any particle is more than half the list skin from xref.
void launch_drift_check(const float4 *pos, const float4 *xref, int32_t n, float half_skin2, int32_t *status,
                        hipStream_t s);
void launch_centroid(const float4 *pos, int32_t n, double *partial, int32_t nblocks, double *out3, hipStream_t s);
void launch_fill_i32(int32_t *p, int32_t v, int64_t n, const int32_t *gate, hipStream_t s);
void launch_pairs_extract(const float4 *pos, const float4 *jprm, const uint32_t *cl, const int32_t *cln,
                          const int32_t *perm, int32_t n, int32_t ncl, int32_t lcap, float rc, float rc2lo,
                          float rc2hi, uint32_t i_mask, uint32_t j_mask, int32_t nonzero, uint32_t *out,
                          int32_t *count, int64_t cap, hipStream_t s);

constexpr int32_t BLOCK = 256;

} // namespace bs
