// bs_api.cpp -- C ABI of the engine (declared in include/bs_api.h).
//
// Host side: owns device buffers and one stream per handle, builds the spring
// CSR and the neighbour lists, and sequences the kernels of one step in the
// order of SpringNetwork::computeStep (src/spn/SpringNetwork.cpp:251-268).
#include "bs_engine.hpp"
#include "bs_netcdf.hpp"
#include "bs_traj.hpp"
#include "bs_params_defaults.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <type_traits>
#include <numeric>
#include <string>
#include <vector>

using namespace bs;

namespace
{

struct List
{
    int terms = 0;
    float rc = 0.0f;        // shared cutoff of the list's terms
    float radius = 0.0f;    // rc + list_skin
    uint32_t imask = 0, jmask = 0;
    int32_t dims[3] = {1, 1, 1};
    int32_t ncells = 1;
    int32_t cap = 0;          // entries per i-cluster
    uint32_t *cl = nullptr;   // [ncl * cap] j | mask << 24, in search order
    int32_t *cln = nullptr;   // [ncl]
    uint32_t *rows = nullptr; // row layout (BS_PAIR_ROWS): [row blocks * rcap * 64]
    int32_t *rlen = nullptr;  // [row blocks * 64]
    int32_t rcap = 0;
};

template <typename T> T *dalloc(size_t count, hipError_t &err)
{
    void *p = nullptr;
    if (count == 0)
        count = 1;
    err = hipMalloc(&p, count * sizeof(T));
    return static_cast<T *>(p);
}

} // namespace

// Device memory of bs_adjoint_backward_noalloc: a fixed array in the code
// object, one copy on each GPU the library runs on, held until the process
// ends.  In host code the array's name is a host-side stand-in, not a device
// address: its address on the current device comes from hipGetSymbolAddress
// (noalloc_setup, from bs_create) and only that address is used.  Every call
// on one GPU shares that GPU's copy; calls that overlap on one GPU race.
constexpr size_t NOALLOC_BYTES = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

namespace
{

constexpr int NOALLOC_MAXDEV = 64;
std::mutex g_noalloc_mu;
char *g_noalloc_ptr[NOALLOC_MAXDEV];
hipError_t g_noalloc_err[NOALLOC_MAXDEV];
bool g_noalloc_done[NOALLOC_MAXDEV];

// Once per device, with that device current: the address of its copy.
void noalloc_setup(int dev)
{
    if (dev < 0 || dev >= NOALLOC_MAXDEV)
        return;
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (g_noalloc_done[dev])
        return;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
    g_noalloc_ptr[dev] = e == hipSuccess ? static_cast<char *>(ptr) : nullptr;
    g_noalloc_err[dev] = e;
    g_noalloc_done[dev] = true;
}

} // namespace

struct bs_handle
{
    bs_params p{};
    std::string err;
    int device = 0; // current device at bs_create
    hipStream_t stream = nullptr;
    Consts k{};

    // host copies, original order
    int64_t n = 0;
    std::vector<float> x, v, mass, charge, radius, eps, sasa, transfer, hydro;
    std::vector<uint8_t> flags;
    // The spring table (original ids) in host memory (sp_host) and/or on the
    // device (sp_dev): construction leaves it on the device and setup reads
    // it there; the host copy is filled on demand (springs_host).
    std::vector<uint32_t> s1, s2;
    std::vector<float> sl0, sk;
    uint32_t *dsa = nullptr, *dsb = nullptr;
    float *dsl = nullptr, *dsk = nullptr;
    int64_t dsp_cap = 0, nsp = 0; // device capacity, springs
    bool sp_host = true, sp_dev = false;
    bool have_particles = false, have_springs = false, ready = false;

    // internal order
    std::vector<int32_t> perm, iperm;

    // rigid bodies (host copies; the state survives a re-setup)
    struct BodyHost
    {
        std::vector<int64_t> ids; // original ids, body order
        bool init = false;
        RigidBody st{};
        std::vector<double> p0;   // 3 per member, body frame
    };
    std::vector<BodyHost> bodies;
    std::vector<int32_t> body_of; // [n] body index or -1
    std::vector<float> burying;   // [n] ParticleProperty::_burying

    // field grids (PotentialGrid after compute_gradient), per BS_GRID_*
    struct GridHost
    {
        bool set = false;
        double origin[3] = {0, 0, 0}, cell[3] = {1, 1, 1};
        int64_t shape[3] = {0, 0, 0};
        std::vector<float4> cells; // gradient force vector, scalar
    };
    GridHost grids[2];

    // device
    float4 *pos[2] = {nullptr, nullptr}, *plo[2] = {nullptr, nullptr};
    float4 *vel = nullptr, *jprm = nullptr, *iprm = nullptr, *xref = nullptr, *fext = nullptr, *fout = nullptr;
    double4 *fnb = nullptr;
    int32_t *sp_ptr = nullptr, *sp_j = nullptr, *perm_d = nullptr;
    float *sp_k = nullptr, *sp_l0 = nullptr;
    bool k_uniform = false; // every spring has the stiffness k_uni: the kernels read no per-edge k
    float k_uni = 0.0f;
    int64_t nedges = 0;
    int32_t ncl = 0;                                  // i-clusters of 8
    // bs_adjoint_forward's record: checkpoints, the last segment's positions
    // and the state at x_T, valid while the state is unchanged
    struct Tape
    {
        bool ready = false;
        int32_t T = 0, K = 0, ncp = 0;
        int32_t seg_of = -1; // the segment whose positions seg holds
        int ext = 0;
        int64_t cap_cp = 0, cap_seg = 0, cap_log = 0;
        float4 *cp_pos = nullptr, *cp_vel = nullptr, *seg = nullptr, *fin_pos = nullptr, *fin_vel = nullptr;
        int32_t *cp_cnt = nullptr, *fin_cnt = nullptr;
        int32_t *gate_log = nullptr;  // [T] the rebuild gate of each recorded step
        std::vector<char> rebuilt;    // its host copy: a recomputation launches the list rebuild there only
    } tape;
    int32_t *gate_log = nullptr; // while recording: where the next step's gate is logged
    // bs_adjoint: springs with a dynamic end in the order of their owning CSR
    // edge ("slots"; the owning edge is in the dynamic end's row, or in the
    // lower index's when both ends move): slot_ptr[i] = first slot of row i,
    // slot_of[s] = slot of spring s or -1
    int32_t *slot_ptr = nullptr, *slot_of = nullptr;
    int64_t nslots = -1;
    int32_t *ex_ptr = nullptr;                        // [ncl + 1] static exclusions
    uint32_t *ex_w = nullptr;                         // j << 8 | bonded mask
    int32_t *cell_of = nullptr, *counts = nullptr, *cstart = nullptr, *ctmp = nullptr,
            *cidx = nullptr;
    int32_t cell_cap = 0;
    double *box_part = nullptr, *box = nullptr;
    int32_t box_blocks = 0;
    std::vector<List> lists;
    int32_t pair_blocks = 0;
    int32_t spring_blocks = 0;
    // energies: per particle on energy steps, summed per segment of at most
    // SEG contiguous particles of one system, then per system
    double4 *epp = nullptr, *ese = nullptr;
    int32_t *seg_start = nullptr, *sys_seg = nullptr;
    double *seg_part = nullptr, *sys_E = nullptr;
    int32_t nseg = 0;
    bool last_call_forces_only = false;
    // captured multi-step launch sequence (graph_steps)
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int32_t graph_len = 0;
    int graph_ext = 0;
    // independent systems (original-id offsets) and their binning offsets
    std::vector<int64_t> sys_off;
    float4 *shift = nullptr;
    RigidBody *bodies_d = nullptr;
    double4 *rx = nullptr, *rp0 = nullptr, *frig = nullptr;
    int32_t *slot_p = nullptr, *slot_b = nullptr, *bstart = nullptr;
    int32_t nr = 0;
    float4 *grid_d[2] = {nullptr, nullptr};
    double *cent_part = nullptr, *cent = nullptr;
    int32_t cent_blocks = 0;
    int32_t group = 16;
    double *energies_d = nullptr;
    int32_t *status = nullptr;   // [0] rebuild, [1] nonfinite id, [2] overflow, [3] max count, [5] stopped,
                                 // [6] steps completed in this call (bs_device.hpp)
    int32_t *counters = nullptr; // [0] step counter, [1] rebuilds
    double energies[BS_NUM_ENERGIES] = {0, 0, 0, 0, 0, 0};
    int cur = 0;
    int64_t steps_done = 0;
    int64_t nonfinite = -1;
    int ext_uploaded = 0;
    int32_t list_longest = 0; // longest neighbour list of the last overflowing build

    // event timing of bs_step (bs_profile)
    bool prof = false;
    std::vector<hipEvent_t> evpool;
    size_t evused = 0;
    struct Span
    {
        int cls;
        size_t a, b;
    };
    std::vector<Span> spans;
    double prof_ms[BS_PROF_NUM] = {};
    int64_t prof_n[BS_PROF_NUM] = {};

    // trajectory output (bs_add_trajectory): writers, pinned frame slots, and
    // the device buffer the positions are packed into
    std::unique_ptr<bs::TrajectorySet> traj;
    bs::TrajLabels labels;
    int64_t ivec1 = -1, ivec2 = -1;
    float *traj_stage = nullptr;
    std::vector<float *> traj_x;
    std::vector<double *> traj_e;
    std::vector<hipEvent_t> traj_ev;

    // the probe particle (probe.*): its state between setups (placed at
    // probe_x/y/z at rest by the first setup after bs_set_particles, as
    // SpringNetwork::_setupProbe), the device copy and the spring kernel's
    // block sums of its force and pair energies
    struct ProbeHost
    {
        bool placed = false;
        float x[3] = {0, 0, 0}, v[3] = {0, 0, 0}, f[3] = {0, 0, 0};
    } probe;
    ProbeDev *probe_d = nullptr;
    double *probe_part = nullptr;
};

namespace
{

int fail(bs_handle *h, int code, const char *fmt, ...)
{
    if (h)
    {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        h->err = buf;
    }
    return code;
}

#define HIPCHK(h, expr)                                                                                           \
    do                                                                                                            \
    {                                                                                                             \
        const hipError_t e_ = (expr);                                                                             \
        if (e_ != hipSuccess)                                                                                     \
            return fail((h), BS_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));                         \
    } while (0)

void drop_graph(bs_handle *h)
{
    if (h->graph_exec)
        (void)hipGraphExecDestroy(h->graph_exec);
    if (h->graph)
        (void)hipGraphDestroy(h->graph);
    h->graph_exec = nullptr;
    h->graph = nullptr;
    h->graph_len = 0;
}

constexpr int TRAJ_SLOTS = 8; // frames in flight between the stream and the file writers

// Writes out the queued frames, closes the files and frees the slots.
void traj_release(bs_handle *h)
{
    if (h->traj)
        h->traj->close();
    h->traj.reset();
    for (float *p : h->traj_x)
        (void)hipHostFree(p);
    for (double *p : h->traj_e)
        (void)hipHostFree(p);
    for (hipEvent_t e : h->traj_ev)
        (void)hipEventDestroy(e);
    h->traj_x.clear();
    h->traj_e.clear();
    h->traj_ev.clear();
    if (h->traj_stage)
    {
        (void)hipStreamSynchronize(h->stream);
        (void)hipFree(h->traj_stage);
        h->traj_stage = nullptr;
    }
}

void tape_free(bs_handle::Tape &tp)
{
    for (void *q : {(void *)tp.cp_pos, (void *)tp.cp_vel, (void *)tp.seg, (void *)tp.fin_pos, (void *)tp.fin_vel,
                    (void *)tp.cp_cnt, (void *)tp.fin_cnt, (void *)tp.gate_log})
        if (q)
            (void)hipFree(q);
    tp = bs_handle::Tape{};
}

void free_device(bs_handle *h)
{
    drop_graph(h);
    tape_free(h->tape);
    auto f = [](void *p) {
        if (p)
            (void)hipFree(p);
    };
    for (int b = 0; b < 2; ++b)
    {
        f(h->pos[b]);
        f(h->plo[b]);
        h->pos[b] = h->plo[b] = nullptr;
    }
    f(h->vel); f(h->fnb); f(h->jprm); f(h->iprm); f(h->xref); f(h->fext); f(h->fout);
    f(h->sp_ptr); f(h->sp_j); f(h->perm_d); f(h->sp_k); f(h->sp_l0);
    f(h->cell_of); f(h->counts); f(h->cstart); f(h->ctmp); f(h->cidx);
    f(h->box_part); f(h->box); f(h->energies_d); f(h->status); f(h->counters);
    f(h->epp); f(h->ese); f(h->seg_start); f(h->sys_seg); f(h->seg_part); f(h->sys_E); f(h->shift);
    h->epp = h->ese = nullptr;
    h->seg_start = h->sys_seg = nullptr;
    h->seg_part = h->sys_E = nullptr;
    h->shift = nullptr;
    h->nseg = 0;
    f(h->ex_ptr); f(h->ex_w); f(h->slot_ptr); f(h->slot_of);
    h->ex_ptr = nullptr;
    h->ex_w = nullptr;
    h->slot_ptr = h->slot_of = nullptr;
    h->nslots = -1;
    f(h->bodies_d); f(h->rx); f(h->rp0); f(h->frig); f(h->slot_p); f(h->slot_b); f(h->bstart);
    f(h->grid_d[0]); f(h->grid_d[1]); f(h->cent_part); f(h->cent);
    f(h->probe_d); f(h->probe_part);
    h->probe_d = nullptr;
    h->probe_part = nullptr;
    h->bodies_d = nullptr;
    h->rx = h->rp0 = h->frig = nullptr;
    h->slot_p = h->slot_b = h->bstart = nullptr;
    h->nr = 0;
    h->grid_d[0] = h->grid_d[1] = nullptr;
    h->cent_part = h->cent = nullptr;
    for (auto &l : h->lists)
    {
        f(l.cl);
        f(l.cln);
        f(l.rows);
        f(l.rlen);
    }
    h->lists.clear();
    h->vel = h->jprm = h->iprm = h->xref = h->fext = h->fout = nullptr;
    h->fnb = nullptr;
    h->sp_ptr = h->sp_j = h->perm_d = nullptr;
    h->sp_k = h->sp_l0 = nullptr;
    h->cell_of = h->counts = h->cstart = h->ctmp = h->cidx = nullptr;
    h->box_part = h->box = h->energies_d = nullptr;
    h->status = h->counters = nullptr;
    h->cell_cap = 0;
    h->nedges = 0;
    h->ready = false;
}

// Host-side guard run before any kernel sequence: every buffer the kernels
// touch must be allocated at the size they assume.
int check_device(bs_handle *h)
{
    const void *need[] = {h->pos[0], h->pos[1], h->vel, h->fnb, h->jprm, h->iprm, h->xref, h->fext, h->fout,
                          h->sp_ptr, h->sp_j, h->sp_k, h->sp_l0, h->perm_d, h->cell_of, h->ctmp, h->cidx,
                          h->counts, h->cstart, h->box, h->box_part, h->epp, h->ese, h->seg_start,
                          h->sys_seg, h->seg_part, h->sys_E, h->energies_d, h->status, h->counters, h->ex_ptr,
                          h->ex_w};
    for (const void *q : need)
        if (!q)
            return fail(h, BS_ERR_STATE, "internal: device buffer missing");
    if (h->p.compensated_positions && (!h->plo[0] || !h->plo[1]))
        return fail(h, BS_ERR_STATE, "internal: compensated-position buffers missing");
    if (!h->bodies.empty() && (!h->bodies_d || !h->frig || !h->bstart || (h->nr > 0 && (!h->rx || !h->rp0 ||
                                                                                        !h->slot_p || !h->slot_b))))
        return fail(h, BS_ERR_STATE, "internal: rigid-body buffers missing");
    for (int g = 0; g < 2; ++g)
        if (h->grids[g].set && !h->grid_d[g])
            return fail(h, BS_ERR_STATE, "internal: grid buffer missing");
    if (h->sys_off.size() > 2 && !h->shift)
        return fail(h, BS_ERR_STATE, "internal: system offsets missing");
    if (h->p.probe_enable && (!h->probe_d || !h->probe_part))
        return fail(h, BS_ERR_STATE, "internal: probe buffers missing");
    if (h->ncl != (int32_t)((h->n + CLUSTER - 1) / CLUSTER))
        return fail(h, BS_ERR_STATE, "internal: cluster count mismatch");
    for (const auto &l : h->lists)
        if (!l.cl || !l.cln || l.cap <= 0 || l.ncells > h->cell_cap ||
            (h->p.pair_layout == BS_PAIR_ROWS && (!l.rows || !l.rlen || l.rcap < l.cap)))
            return fail(h, BS_ERR_STATE, "internal: neighbour list buffers missing");
    return BS_OK;
}

inline uint64_t spread3(uint64_t v)
{
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}

// Internal order: Morton order of a 4 A grid; ties keep the original order.
void make_order(bs_handle *h)
{
    // Morton order of 4 A cells within each system; systems stay contiguous
    // in their given order, so each is one range of internal indices.
    const int64_t n = h->n;
    std::vector<std::pair<std::pair<int32_t, uint64_t>, int32_t>> key((size_t)n);
    const double cell = 4.0;
    const int32_t nsys = (int32_t)h->sys_off.size() - 1;
    for (int32_t sy = 0; sy < nsys; ++sy)
    {
        const int64_t b = h->sys_off[(size_t)sy], e = h->sys_off[(size_t)sy + 1];
        double lo[3] = {INFINITY, INFINITY, INFINITY};
        for (int64_t i = b; i < e; ++i)
            for (int c = 0; c < 3; ++c)
                lo[c] = std::min(lo[c], (double)h->x[3 * i + c]);
        for (int64_t i = b; i < e; ++i)
        {
            uint64_t q[3];
            for (int c = 0; c < 3; ++c)
            {
                double t = std::floor(((double)h->x[3 * i + c] - lo[c]) / cell);
                t = std::min(std::max(t, 0.0), 2097151.0);
                q[c] = (uint64_t)t;
            }
            key[(size_t)i] = {{sy, spread3(q[0]) | spread3(q[1]) << 1 | spread3(q[2]) << 2}, (int32_t)i};
        }
    }
    std::sort(key.begin(), key.end());
    h->perm.resize((size_t)n);
    h->iperm.resize((size_t)n);
    for (int64_t t = 0; t < n; ++t)
    {
        h->perm[(size_t)t] = key[(size_t)t].second;
        h->iperm[(size_t)key[(size_t)t].second] = (int32_t)t;
    }
}

// Binning offsets of a batch: system s is moved to site s of a cubic lattice
// whose spacing exceeds every system's extent plus twice the largest list
// radius and a 20 A margin, so no neighbour search reaches across systems.
// The offsets apply to binning only; forces use each system's own
// coordinates.  shift.w carries the system index (as raw bits).
std::vector<float4> system_shifts(const bs_handle *h, float max_radius)
{
    const int32_t nsys = (int32_t)h->sys_off.size() - 1;
    std::vector<std::array<double, 6>> bb((size_t)nsys);
    double ext = 0.0;
    for (int32_t sy = 0; sy < nsys; ++sy)
    {
        auto &b = bb[(size_t)sy];
        b = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
        for (int64_t i = h->sys_off[(size_t)sy]; i < h->sys_off[(size_t)sy + 1]; ++i)
            for (int c = 0; c < 3; ++c)
            {
                b[c] = std::min(b[c], (double)h->x[3 * i + c]);
                b[3 + c] = std::max(b[3 + c], (double)h->x[3 * i + c]);
            }
        for (int c = 0; c < 3; ++c)
            ext = std::max(ext, b[3 + c] - b[c]);
    }
    const double spacing = std::ceil(ext + 2.0 * max_radius + 20.0);
    const int32_t side = (int32_t)std::ceil(std::cbrt((double)nsys) - 1e-9);
    std::vector<float4> out((size_t)h->n);
    for (int32_t sy = 0; sy < nsys; ++sy)
    {
        const int32_t site[3] = {sy % side, (sy / side) % side, sy / (side * side)};
        float o[3];
        for (int c = 0; c < 3; ++c)
            o[c] = (float)(site[c] * spacing - bb[(size_t)sy][c]);
        uint32_t tag = (uint32_t)sy;
        float w;
        std::memcpy(&w, &tag, 4);
        for (int64_t i = h->sys_off[(size_t)sy]; i < h->sys_off[(size_t)sy + 1]; ++i)
            out[(size_t)h->iperm[(size_t)i]] = make_float4(o[0], o[1], o[2], w);
    }
    return out;
}

uint32_t particle_flags(const bs_handle *h, int64_t i)
{
    uint32_t f = 0;
    if (h->flags[(size_t)i] & BS_FLAG_STATIC)
        f |= PF_STATIC;
    if (h->charge[(size_t)i] != 0.0f)
        f |= PF_CHARGED;
    if (h->hydro[(size_t)i] != 0.0f)
        f |= PF_HYDRO;
    if (!h->body_of.empty() && h->body_of[(size_t)i] >= 0)
        f |= PF_RIGID;
    return f;
}

float4 jprm_of(const bs_handle *h, int64_t i)
{
    const bool sq = h->p.steric_mode == BS_STERIC_AMBER_12_6 || h->p.steric_mode == BS_STERIC_LEWITT_8_6;
    const float r = h->radius[(size_t)i], e = h->eps[(size_t)i];
    float4 q;
    q.x = sq ? sqrtf(r) : r;
    q.y = sq ? sqrtf(e) : e;
    q.z = h->hydro[(size_t)i];
    uint32_t fl = particle_flags(h, i);
    std::memcpy(&q.w, &fl, 4);
    return q;
}

int upload_positions(bs_handle *h)
{
    const int64_t n = h->n;
    std::vector<float4> pv((size_t)n), vv((size_t)n);
    for (int64_t t = 0; t < n; ++t)
    {
        const int64_t i = h->perm[(size_t)t];
        pv[(size_t)t] = make_float4(h->x[3 * i], h->x[3 * i + 1], h->x[3 * i + 2], h->charge[(size_t)i]);
        vv[(size_t)t] = make_float4(h->v[3 * i], h->v[3 * i + 1], h->v[3 * i + 2], 0.0f);
    }
    for (int b = 0; b < 2; ++b)
    {
        HIPCHK(h, hipMemcpyAsync(h->pos[b], pv.data(), sizeof(float4) * n, hipMemcpyHostToDevice, h->stream));
        if (h->plo[b])
            HIPCHK(h, hipMemsetAsync(h->plo[b], 0, sizeof(float4) * n, h->stream));
    }
    HIPCHK(h, hipMemcpyAsync(h->vel, vv.data(), sizeof(float4) * n, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return BS_OK;
}

void plan_lists(bs_handle *h)
{
    const bs_params &p = h->p;
    bool any_q = false, all_q = true, any_h = false, all_h = true, same_qh = true;
    for (int64_t i = 0; i < h->n; ++i)
    {
        const bool q = h->charge[(size_t)i] != 0.0f, hy = h->hydro[(size_t)i] != 0.0f;
        any_q |= q;
        all_q &= q;
        any_h |= hy;
        all_h &= hy;
        same_qh &= (q == hy);
    }
    h->lists.clear();
    int steric_idx = -1, elec_idx = -1;
    if (p.steric_enable)
    {
        List l;
        l.terms = TERM_STERIC;
        l.rc = p.steric_cutoff;
        h->lists.push_back(l);
        steric_idx = 0;
    }
    if (p.electrostatic_enable && any_q)
    {
        if (steric_idx >= 0 && p.electrostatic_cutoff == p.steric_cutoff && all_q)
            h->lists[(size_t)steric_idx].terms |= TERM_ELEC;
        else
        {
            List l;
            l.terms = TERM_ELEC;
            l.rc = p.electrostatic_cutoff;
            l.imask = l.jmask = PF_CHARGED;
            h->lists.push_back(l);
            elec_idx = (int)h->lists.size() - 1;
        }
    }
    if (p.hydrophobicity_enable && any_h)
    {
        if (steric_idx >= 0 && p.hydrophobicity_cutoff == p.steric_cutoff && all_h)
            h->lists[(size_t)steric_idx].terms |= TERM_HYDRO;
        else if (elec_idx >= 0 && p.hydrophobicity_cutoff == p.electrostatic_cutoff && same_qh)
            h->lists[(size_t)elec_idx].terms |= TERM_HYDRO;
        else
        {
            List l;
            l.terms = TERM_HYDRO;
            l.rc = p.hydrophobicity_cutoff;
            l.imask = l.jmask = PF_HYDRO;
            h->lists.push_back(l);
        }
    }
    for (auto &l : h->lists)
        l.radius = l.rc + std::max(p.list_skin, 0.0f);
}

// Engine grid dimensions of a list from the current host positions.
void plan_grid(bs_handle *h, List &l, const std::vector<float4> &shift)
{
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = 0; i < h->n; ++i)
    {
        float o[3] = {0.0f, 0.0f, 0.0f};
        if (!shift.empty())
        {
            const float4 q = shift[(size_t)h->iperm[(size_t)i]];
            o[0] = q.x;
            o[1] = q.y;
            o[2] = q.z;
        }
        for (int c = 0; c < 3; ++c)
        {
            const double v = (double)(h->x[3 * i + c] + o[c]);
            lo[c] = std::min(lo[c], v);
            hi[c] = std::max(hi[c], v);
        }
    }
    // Binning grid of edge >= radius / 2: a cluster's search covers its
    // bounding box grown by the radius, about 5-6 cells per axis.
    int64_t total = 1;
    for (int c = 0; c < 3; ++c)
    {
        int64_t d = (int64_t)std::floor((hi[c] - lo[c]) / (0.5 * l.radius));
        d = std::max<int64_t>(1, std::min<int64_t>(d, 1024));
        l.dims[c] = (int32_t)d;
        total *= d;
    }
    while (total > (1 << 22))
    { // keep the grid modest for sparse, spread-out systems
        for (int c = 0; c < 3; ++c)
            l.dims[c] = std::max(1, l.dims[c] / 2);
        total = (int64_t)l.dims[0] * l.dims[1] * l.dims[2];
    }
    l.ncells = (int32_t)total;
}

int ensure_cells(bs_handle *h, int32_t ncells)
{
    if (ncells <= h->cell_cap)
        return BS_OK;
    hipError_t e;
    auto f = [](void *p) {
        if (p)
            (void)hipFree(p);
    };
    f(h->counts);
    f(h->cstart);
    h->counts = dalloc<int32_t>((size_t)ncells, e);
    HIPCHK(h, e);
    h->cstart = dalloc<int32_t>((size_t)ncells + 1, e);
    HIPCHK(h, e);
    // every binning counts its cells back down to zero (launch_scatter's
    // drain), so the counts are zeroed only here
    launch_fill_i32(h->counts, 0, ncells, nullptr, h->stream);
    h->cell_cap = ncells;
    return BS_OK;
}

const int32_t *gate_of(bs_handle *h, bool gated) { return gated ? h->status : nullptr; }

// Engine cell grid + Verlet list of list l.  gated: skipped on the device
// unless the rebuild gate (status[0]) is set.
void build_list(bs_handle *h, List &l, bool gated, int32_t cap_override)
{
    const int32_t n = (int32_t)h->n;
    const int32_t *g = gate_of(h, gated);
    hipStream_t s = h->stream;
    // the box from the partials of launch_box_partial; the first list also
    // takes the reference positions of the drift test
    launch_cell_ids(h->pos[h->cur], h->shift, h->jprm, l.jmask, n, h->box, h->box_part, h->box_blocks,
                    &l == &h->lists.front() ? h->xref : nullptr, l.dims[0], l.dims[1], l.dims[2], 0.0,
                    0.5f * l.radius, 0, h->cell_of, h->counts, g, s);
    launch_scan(h->counts, h->cstart, l.ncells, g, s);
    launch_scatter(h->cell_of, n, h->cstart, h->counts, 1, h->ctmp, g, s);
    launch_cell_sort(h->cell_of, n, h->cstart, h->ctmp, nullptr, h->cidx, g, s);
    ListArgs a{};
    a.pos = h->pos[h->cur];
    a.shift = h->shift;
    a.jprm = h->jprm;
    a.cell_start = h->cstart;
    a.cell_idx = h->cidx;
    a.box = h->box;
    a.ex_ptr = h->ex_ptr;
    a.ex_w = h->ex_w;
    a.cl = l.cl;
    a.cln = l.cln;
    a.status = h->status;
    a.n = n;
    a.ncl = h->ncl;
    a.cap = cap_override >= 0 ? cap_override : l.cap;
    for (int c = 0; c < 3; ++c)
        a.dims[c] = l.dims[c];
    a.cell_edge = 0.5f * l.radius;
    a.radius = l.radius;
    a.r2list = (float)((double)l.radius * l.radius * (1.0 + 1e-5));
    a.i_mask = l.imask;
    a.j_mask = l.jmask;
    a.exclude_bonded = h->p.spring_enable ? 1 : 0;
    a.gate = g;
    launch_list_build(a, s);
    if (l.rows && cap_override < 0)
        launch_rows(l.cl, l.cln, h->ncl, l.cap, l.rows, l.rlen, l.rcap, g, s);
}

// The gate kernel: with lists, it also writes the box partials of the
// current positions for the chain's cell binning.
void launch_gate(bs_handle *h, int32_t interval, int32_t advance)
{
    const bool lists = !h->lists.empty();
    launch_prestep(h->status, h->counters, interval, advance, lists ? h->pos[h->cur] : nullptr, h->shift,
                   (int32_t)h->n, lists ? h->box_part : nullptr, h->box_blocks, advance ? h->gate_log : nullptr,
                   h->stream);
}

// The gated list rebuild.  prepped: the gate (and the box partials) were
// just set by a step's prestep; otherwise they are set here from the drift
// flag alone.
void rebuild_chain(bs_handle *h, bool gated, bool prepped = false)
{
    if (!prepped)
        launch_gate(h, 0, 0);
    for (auto &l : h->lists)
        build_list(h, l, gated, -1);
}

float rc2(float rc, double f) { return (float)((double)rc * rc * f); }

// energy: also sum this evaluation's energies (the last step of a call).
void launch_pairs(bs_handle *h, bool energy)
{
    const int32_t n = (int32_t)h->n;
    for (size_t li = 0; li < h->lists.size(); ++li)
    {
        const List &l = h->lists[li];
        PairArgs a{};
        a.pos = h->pos[h->cur];
        a.jprm = h->jprm;
        a.cl = l.cl;
        a.cln = l.cln;
        a.fnb = h->fnb;
        a.epp = h->epp;
        a.rows = l.rows;
        a.rlen = l.rlen;
        a.rcap = l.rcap;
        a.queues = h->p.pair_layout == BS_PAIR_QUEUES ? 1 : 0;
        a.n = n;
        a.ncl = h->ncl;
        a.cap = l.cap;
        a.accumulate = li > 0 ? 1 : 0;
        a.energy = energy ? 1 : 0;
        a.status = h->status;
        a.rc = l.rc;
        a.rc2lo = rc2(l.rc, 1.0 - 4e-6);
        a.rc2hi = rc2(l.rc, 1.0 + 4e-6);
        a.steric_mode = (l.terms & TERM_STERIC) ? h->p.steric_mode : -1;
        a.steric_scale = h->p.steric_gridscale;
        a.elec_scale = h->p.electrostatic_scale;
        a.diel = h->p.electrostatic_dielectric;
        // Coulomb prefactors (electrostatic.hpp:16-49): energy e_C q1 q2 / (eps d),
        // force module q1 q2 / (4 pi eps d^2) * C_elec, both times the scale
        a.ke = h->k.e_coul / a.diel * a.elec_scale;
        a.kf = h->k.c_elec / (4.0 * h->k.pi * a.diel) * a.elec_scale;
        a.hydro_scale = h->p.hydrophobicity_scale;
        a.i_mask = l.imask;
        a.j_mask = l.jmask;
        a.k = h->k;
        const int em = (l.terms & TERM_ELEC) ? (h->p.dielectric_model == BS_DIELECTRIC_DISTANCE ? 2 : 1) : 0;
        launch_pair(a, em, (l.terms & TERM_HYDRO) ? 1 : 0, h->pair_blocks, h->stream);
    }
}

GridView grid_view(const bs_handle *h, int kind, bool on, float scale)
{
    GridView g{};
    const auto &G = h->grids[kind];
    if (!on || !G.set)
        return g;
    g.cells = h->grid_d[kind];
    for (int c = 0; c < 3; ++c)
    {
        g.origin[c] = G.origin[c];
        g.cell[c] = G.cell[c];
        g.bmax[c] = G.origin[c] + (double)G.shape[c] * G.cell[c]; // Box(origin, shape * cell)
        g.shape[c] = (int32_t)G.shape[c];
    }
    g.scale = scale;
    return g;
}

ProbeArgs probe_args(const bs_handle *h)
{
    ProbeArgs a{};
    if (!h->p.probe_enable)
        return a;
    const bs_params &P = h->p;
    a.state = h->probe_d;
    a.part = h->probe_part;
    a.steric = P.probe_enable_steric;
    a.elec = P.probe_enable_electrostatic;
    a.steric_mode = P.steric_mode;
    a.steric_scale = P.steric_gridscale;
    const bool sq = P.steric_mode == BS_STERIC_AMBER_12_6 || P.steric_mode == BS_STERIC_LEWITT_8_6;
    a.radius = sq ? std::sqrt(P.probe_radius) : P.probe_radius; // as jprm holds the particles'
    a.eps = sq ? std::sqrt(P.probe_epsilon) : P.probe_epsilon;
    a.mass = P.probe_mass;
    a.ke = h->k.e_coul / P.electrostatic_dielectric * P.electrostatic_scale;
    a.kf = h->k.c_elec / (4.0 * h->k.pi * P.electrostatic_dielectric) * P.electrostatic_scale;
    a.ddd = P.dielectric_model == BS_DIELECTRIC_DISTANCE ? 1 : 0;
    a.status = h->status;
    return a;
}

void launch_springs(bs_handle *h, bool integrate, bool energy)
{
    SpringArgs a{};
    a.pos = h->pos[h->cur];
    a.plo = h->p.compensated_positions ? h->plo[h->cur] : nullptr;
    a.pos_out = h->pos[h->cur ^ 1];
    a.plo_out = h->p.compensated_positions ? h->plo[h->cur ^ 1] : nullptr;
    a.vel = h->vel;
    a.fnb = h->lists.empty() ? nullptr : h->fnb;
    a.fext = h->ext_uploaded ? h->fext : nullptr;
    a.iprm = h->iprm;
    a.jprm = h->jprm;
    a.xref = h->xref;
    a.sp_ptr = h->sp_ptr;
    a.sp_j = h->sp_j;
    a.sp_k = h->k_uniform ? nullptr : h->sp_k;
    a.k_uni = h->k_uni;
    a.sp_l0 = h->sp_l0;
    a.perm = h->perm_d;
    a.fout = h->fout;
    a.eps = h->ese;
    a.status = h->status;
    a.n = (int32_t)h->n;
    a.spring_enable = h->p.spring_enable;
    a.spring_scale = h->p.spring_scale;
    a.dt = h->p.timestep;
    a.visc_enable = h->p.viscosity_enable;
    a.visc = h->p.viscosity_value;
    a.imp_enable = h->p.imp_enable;
    a.imp_scale = h->p.imp_scale;
    a.imp_uo = h->p.imp_upper_offset;
    a.imp_lo = h->p.imp_lower_offset;
    a.imp_uc = h->p.imp_upper_curvature;
    a.imp_lc = h->p.imp_lower_curvature;
    const double hs = 0.5 * std::max(h->p.list_skin, 0.0f);
    a.half_skin2 = (float)(hs * hs * (1.0 - 1e-5));
    a.integrate = integrate ? 1 : 0;
    a.energy = energy ? 1 : 0;
    a.k = h->k;
    a.frig = h->frig;
    // the electrostatic field needs electrostatic.enable (SpringNetwork.cpp:107-113)
    a.gpot = grid_view(h, BS_GRID_POTENTIAL, h->p.electrostatic_enable && h->p.potentialgrid_enable,
                       h->p.potentialgrid_scale);
    a.gden = grid_view(h, BS_GRID_DENSITY, h->p.densitygrid_enable, h->p.densitygrid_scale);
    a.e_field = 1.3806504E-23 * 1.0E-3 * 6.02214169E23; // k_B, J -> kJ, N_A (constants.hpp:70-73)
    a.probe = probe_args(h);
    a.steps_done = h->status + 6;
    launch_spring(a, h->group, h->spring_blocks, h->stream);
}

// The probe's step after the spring kernel (integrate) or its force and pair
// energies only (a force evaluation).
void launch_probe_step(bs_handle *h, bool integrate)
{
    if (!h->p.probe_enable)
        return;
    launch_probe(probe_args(h), h->spring_blocks, h->p.timestep, h->p.viscosity_enable, h->p.viscosity_value,
                 h->k.c_kinetic, integrate ? 1 : 0, h->stream);
}

void launch_bodies(bs_handle *h)
{
    if (h->bodies.empty())
        return;
    RigidArgs a{};
    a.pos = h->pos[h->cur];
    a.pos_out = h->pos[h->cur ^ 1];
    a.vel = h->vel;
    a.frig = h->frig;
    a.rx = h->rx;
    a.p0 = h->rp0;
    a.slot_p = h->slot_p;
    a.slot_b = h->slot_b;
    a.bstart = h->bstart;
    a.bodies = h->bodies_d;
    a.xref = h->xref;
    a.perm = h->perm_d;
    a.status = h->status;
    a.nr = h->nr;
    a.nb = (int32_t)h->bodies.size();
    a.solve = h->p.rigidbody_enable;
    a.dt = (double)h->p.timestep;
    const double hs = 0.5 * std::max(h->p.list_skin, 0.0f);
    a.half_skin2 = (float)(hs * hs * (1.0 - 1e-5));
    launch_rigid(a, h->stream);
}

// Records an event on the handle's stream and returns its pool index.
size_t prof_mark(bs_handle *h)
{
    if (h->evused == h->evpool.size())
    {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess)
            return SIZE_MAX;
        h->evpool.push_back(e);
    }
    const size_t idx = h->evused++;
    if (hipEventRecord(h->evpool[idx], h->stream) != hipSuccess)
        return SIZE_MAX;
    return idx;
}

void prof_span(bs_handle *h, int cls, size_t a, size_t b)
{
    if (a != SIZE_MAX && b != SIZE_MAX)
        h->spans.push_back({cls, a, b});
}

// After the stream has been synchronised: folds the recorded spans in.
void prof_collect(bs_handle *h)
{
    for (const auto &s : h->spans)
    {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, h->evpool[s.a], h->evpool[s.b]) == hipSuccess)
        {
            h->prof_ms[s.cls] += ms;
            h->prof_n[s.cls] += 1;
        }
    }
    h->spans.clear();
    h->evused = 0;
}

// Bounds the event pool: folds in what is recorded once it holds 4096 events.
int prof_fold(bs_handle *h)
{
    if (h->prof && h->evused > 4096)
    {
        HIPCHK(h, hipStreamSynchronize(h->stream));
        prof_collect(h);
    }
    return BS_OK;
}

// The six energies of the last energy step, summed into energies_d.
void launch_energy_sum(bs_handle *h)
{
    launch_energies(h->lists.empty() ? nullptr : h->epp, h->ese, h->seg_start, h->nseg, h->sys_seg,
                    (int32_t)h->sys_off.size() - 1, h->seg_part, h->sys_E, h->energies_d,
                    h->p.probe_enable ? h->probe_d : nullptr, h->stream);
}

// Queues a trajectory frame of the current state at iteration nb: positions
// packed into particle order and copied to a pinned slot, with the energies
// of the last energy step (on the device when dev_energies, else the host
// copy of the last call's), an event marking the copies; the writer thread
// writes it when the event has passed.  all: to every writer.
int traj_capture(bs_handle *h, int64_t nb, bool dev_energies, bool all)
{
    const int64_t n = h->n;
    if (!h->traj_stage)
        HIPCHK(h, hipMalloc((void **)&h->traj_stage, sizeof(float) * 3 * n));
    const int s = h->traj->acquire();
    hipError_t e = hipSuccess;
    launch_pack_positions(h->pos[h->cur], h->p.compensated_positions ? h->plo[h->cur] : nullptr, h->perm_d,
                          (int32_t)n, h->traj_stage, h->stream);
    e = hipMemcpyAsync(h->traj_x[(size_t)s], h->traj_stage, sizeof(float) * 3 * n, hipMemcpyDeviceToHost,
                       h->stream);
    if (e == hipSuccess && dev_energies)
    {
        launch_energy_sum(h);
        e = hipMemcpyAsync(h->traj_e[(size_t)s], h->energies_d, sizeof(double) * BS_NUM_ENERGIES,
                           hipMemcpyDeviceToHost, h->stream);
    }
    else if (e == hipSuccess)
        std::memcpy(h->traj_e[(size_t)s], h->energies, sizeof(double) * BS_NUM_ENERGIES);
    int32_t *halt = reinterpret_cast<int32_t *>(h->traj_e[(size_t)s] + BS_NUM_ENERGIES);
    if (e == hipSuccess) // whether the call had stopped before this frame
        e = hipMemcpyAsync(halt, h->status, sizeof(int32_t) * 8, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess)
        e = hipEventRecord(h->traj_ev[(size_t)s], h->stream);
    if (e != hipSuccess)
    {
        h->traj->release(s);
        HIPCHK(h, e);
    }
    bs::TrajFrame f;
    f.nbiter = nb;
    f.x3 = h->traj_x[(size_t)s];
    f.energies = h->traj_e[(size_t)s];
    f.status = halt;
    const hipEvent_t ev = h->traj_ev[(size_t)s];
    h->traj->submit(
        s, f,
        [ev]() -> std::string {
            const hipError_t r = hipEventSynchronize(ev);
            return r == hipSuccess ? std::string() : std::string("frame copy failed: ") + hipGetErrorString(r);
        },
        all);
    return BS_OK;
}

// Ends a call's launches: sums the energies of its last energy step and reads
// the status words.  completed (nullable): steps the call completed before it
// stopped or ended.  overflow (nullable): set, instead of failing, when a list
// overflowed, for the caller to lengthen the lists and resume.
int finish_call(bs_handle *h, int64_t *completed = nullptr, bool *overflow = nullptr)
{
    launch_energy_sum(h);
    int32_t st[8];
    HIPCHK(h, hipMemcpyAsync(h->energies, h->energies_d, sizeof(h->energies), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(st, h->status, sizeof st, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    prof_collect(h);
    if (completed)
        *completed = st[6];
    if (st[2] != 0)
    {
        h->list_longest = st[3];
        if (overflow)
        {
            *overflow = true;
            return BS_OK;
        }
        return fail(h, BS_ERR_STATE, "neighbour list capacity exceeded (max %d > cap); rerun bs_setup", st[3]);
    }
    if (st[1] != INT_MAX)
    {
        h->nonfinite = st[1];
        return fail(h, BS_ERR_NONFINITE, "Found non-finite position for particle %d.", st[1]);
    }
    return BS_OK;
}

// Longer neighbour lists after an overflow: every list shorter than the
// longest list the failed build found, plus the setup's margin, is
// reallocated; the stop is cleared and the next step rebuilds the lists.
int regrow_lists(bs_handle *h)
{
    HIPCHK(h, hipStreamSynchronize(h->stream));
    drop_graph(h);
    h->tape.ready = false;
    const int32_t need = (std::max<int32_t>(16, (int32_t)std::ceil(h->list_longest * 1.25) + 64) + 15) / 16 * 16;
    hipError_t e;
    for (auto &l : h->lists)
    {
        if (l.cap >= need)
            continue;
        (void)hipFree(l.cl);
        l.cl = nullptr;
        l.cap = need;
        l.cl = dalloc<uint32_t>((size_t)h->ncl * l.cap, e);
        HIPCHK(h, e);
        if (l.rows)
        {
            (void)hipFree(l.rows);
            l.rows = nullptr;
            l.rcap = l.cap;
            l.rows = dalloc<uint32_t>((size_t)((h->n + 63) / 64) * 64 * l.rcap, e);
            HIPCHK(h, e);
        }
    }
    int32_t st[8];
    HIPCHK(h, hipMemcpy(st, h->status, sizeof st, hipMemcpyDeviceToHost));
    st[7] = 1; // rebuild at the next step
    st[2] = st[5] = st[6] = 0;
    HIPCHK(h, hipMemcpy(h->status, st, sizeof st, hipMemcpyHostToDevice));
    return BS_OK;
}

// One step: rebuild gate, (gated) list rebuild, pair forces, springs +
// integration, rigid bodies.  energy: also write the per-particle energies.
void launch_step(bs_handle *h, bool energy)
{
    const size_t e0 = h->prof ? prof_mark(h) : SIZE_MAX;
    launch_gate(h, h->p.rebuild_interval, 1);
    rebuild_chain(h, true, true);
    const size_t e1 = h->prof ? prof_mark(h) : SIZE_MAX;
    if (!h->lists.empty())
        launch_pairs(h, energy);
    const size_t e2 = h->prof ? prof_mark(h) : SIZE_MAX;
    launch_springs(h, true, energy);
    launch_bodies(h);
    launch_probe_step(h, true);
    if (h->prof)
    {
        const size_t e3 = prof_mark(h);
        prof_span(h, BS_PROF_REBUILD, e0, e1);
        if (!h->lists.empty())
            prof_span(h, BS_PROF_PAIR, e1, e2);
        prof_span(h, BS_PROF_SPRING, e2, e3);
    }
    h->cur ^= 1;
}

// A step whose list rebuild is known not to happen (a recomputation of a
// recorded step that did not rebuild): no gate, no gated chain.  The step
// counter is not advanced; the recomputation restores the counters.
void launch_step_unlisted(bs_handle *h)
{
    const size_t e1 = h->prof ? prof_mark(h) : SIZE_MAX;
    if (!h->lists.empty())
        launch_pairs(h, false);
    const size_t e2 = h->prof ? prof_mark(h) : SIZE_MAX;
    launch_springs(h, true, false);
    launch_bodies(h);
    launch_probe_step(h, true);
    if (h->prof)
    {
        if (!h->lists.empty())
            prof_span(h, BS_PROF_PAIR, e1, e2);
        prof_span(h, BS_PROF_SPRING, e2, prof_mark(h));
    }
    h->cur ^= 1;
}

// Captures G steps (G even, starting at cur = 0) into a HIP graph.  The step
// holds no host synchronisation and no memset nodes; rebuild decisions stay
// on the device.
int capture_graph(bs_handle *h, int32_t G)
{
    drop_graph(h);
    HIPCHK(h, hipStreamSynchronize(h->stream));
    // the trajectory writer thread waits on events of this stream: let it
    // finish first, as a wait on a capturing stream's event ends the capture
    if (h->traj)
        h->traj->drain();
    HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeRelaxed));
    const int cur0 = h->cur;
    for (int32_t s = 0; s < G; ++s)
        launch_step(h, false);
    h->cur = cur0;
    hipGraph_t g = nullptr;
    HIPCHK(h, hipStreamEndCapture(h->stream, &g));
    h->graph = g;
    HIPCHK(h, hipGraphInstantiate(&h->graph_exec, g, nullptr, nullptr, 0));
    h->graph_len = G;
    h->graph_ext = h->ext_uploaded;
    return BS_OK;
}

int upload_ext(bs_handle *h, const float *ext3)
{
    if (ext3)
        h->tape.ready = false; // a recorded forward pass used the force it had
    h->ext_uploaded = 0;
    if (!ext3)
        return BS_OK;
    std::vector<float4> fe((size_t)h->n);
    for (int64_t t = 0; t < h->n; ++t)
    {
        const int64_t i = h->perm[(size_t)t];
        fe[(size_t)t] = make_float4(ext3[3 * i], ext3[3 * i + 1], ext3[3 * i + 2], 0.0f);
    }
    HIPCHK(h, hipMemcpyAsync(h->fext, fe.data(), sizeof(float4) * h->n, hipMemcpyHostToDevice, h->stream));
    h->ext_uploaded = 1;
    return BS_OK;
}

// Device buffers of one call, released on every return path.
struct Scratch
{
    std::vector<void *> ptrs;
    Scratch() = default;
    Scratch(const Scratch &) = delete;
    Scratch &operator=(const Scratch &) = delete;
    ~Scratch()
    {
        for (void *q : ptrs)
            (void)hipFree(q);
    }
    template <typename T> T *get(size_t count, hipError_t &err)
    {
        T *q = dalloc<T>(count, err);
        if (q)
            ptrs.push_back(q);
        return q;
    }
};

// The host copy of the spring table, downloaded when the device holds the
// only copy (after a construction).
int springs_host(bs_handle *h)
{
    if (h->sp_host)
        return BS_OK;
    const size_t m = (size_t)h->nsp;
    h->s1.resize(m);
    h->s2.resize(m);
    h->sl0.resize(m);
    h->sk.resize(m);
    if (m)
    {
        HIPCHK(h, hipMemcpy(h->s1.data(), h->dsa, sizeof(uint32_t) * m, hipMemcpyDeviceToHost));
        HIPCHK(h, hipMemcpy(h->s2.data(), h->dsb, sizeof(uint32_t) * m, hipMemcpyDeviceToHost));
        HIPCHK(h, hipMemcpy(h->sl0.data(), h->dsl, sizeof(float) * m, hipMemcpyDeviceToHost));
        HIPCHK(h, hipMemcpy(h->sk.data(), h->dsk, sizeof(float) * m, hipMemcpyDeviceToHost));
    }
    h->sp_host = true;
    return BS_OK;
}

// Device table of capacity >= cap whose first keep springs are the current
// table's (device copy, or uploaded from the host copy).
int springs_device_reserve(bs_handle *h, int64_t cap, int64_t keep)
{
    if (keep > 0 && !h->sp_dev && !h->sp_host)
        return fail(h, BS_ERR_STATE, "internal: no copy of the spring table");
    hipError_t e = hipSuccess;
    if (cap > h->dsp_cap)
    {
        uint32_t *a = dalloc<uint32_t>((size_t)cap, e), *b = nullptr;
        float *l = nullptr, *k = nullptr;
        if (e == hipSuccess)
            b = dalloc<uint32_t>((size_t)cap, e);
        if (e == hipSuccess)
            l = dalloc<float>((size_t)cap, e);
        if (e == hipSuccess)
            k = dalloc<float>((size_t)cap, e);
        if (e == hipSuccess && keep > 0 && h->sp_dev)
        {
            const size_t kb = (size_t)keep;
            e = hipMemcpy(a, h->dsa, sizeof(uint32_t) * kb, hipMemcpyDeviceToDevice);
            if (e == hipSuccess)
                e = hipMemcpy(b, h->dsb, sizeof(uint32_t) * kb, hipMemcpyDeviceToDevice);
            if (e == hipSuccess)
                e = hipMemcpy(l, h->dsl, sizeof(float) * kb, hipMemcpyDeviceToDevice);
            if (e == hipSuccess)
                e = hipMemcpy(k, h->dsk, sizeof(float) * kb, hipMemcpyDeviceToDevice);
        }
        if (e != hipSuccess)
        {
            for (void *q : {(void *)a, (void *)b, (void *)l, (void *)k})
                if (q)
                    (void)hipFree(q);
            HIPCHK(h, e);
        }
        for (void *q : {(void *)h->dsa, (void *)h->dsb, (void *)h->dsl, (void *)h->dsk})
            if (q)
                (void)hipFree(q);
        h->dsa = a;
        h->dsb = b;
        h->dsl = l;
        h->dsk = k;
        h->dsp_cap = cap;
    }
    if (keep > 0 && !h->sp_dev)
    {
        const size_t kb = (size_t)keep;
        HIPCHK(h, hipMemcpy(h->dsa, h->s1.data(), sizeof(uint32_t) * kb, hipMemcpyHostToDevice));
        HIPCHK(h, hipMemcpy(h->dsb, h->s2.data(), sizeof(uint32_t) * kb, hipMemcpyHostToDevice));
        HIPCHK(h, hipMemcpy(h->dsl, h->sl0.data(), sizeof(float) * kb, hipMemcpyHostToDevice));
        HIPCHK(h, hipMemcpy(h->dsk, h->sk.data(), sizeof(float) * kb, hipMemcpyHostToDevice));
    }
    return BS_OK;
}

// The device copy of the spring table, uploaded when only the host has it.
int springs_device(bs_handle *h)
{
    if (h->sp_dev)
        return BS_OK;
    const int rc = springs_device_reserve(h, std::max<int64_t>(h->nsp, 1), h->nsp);
    if (rc)
        return rc;
    h->sp_dev = true;
    return BS_OK;
}

// The host copy was replaced: the device copy is stale.
void springs_host_changed(bs_handle *h)
{
    h->nsp = (int64_t)h->s1.size();
    h->sp_host = true;
    h->sp_dev = false;
}

// The CSR edge on which the adjoint sweep accumulates each spring's
// gradients: the edge of the dynamic row that owns it (a static partner, or
// else the larger index is the partner; the rule of k_adj_spring), -1 when
// both ends are static.  Built once per setup.
int build_spring_owners(bs_handle *h)
{
    const int64_t n = h->n, m = h->nsp, E = h->nedges;
    if (h->nslots >= 0 || m == 0)
        return BS_OK;
    const int hrc = springs_host(h);
    if (hrc)
        return hrc;
    std::vector<int32_t> hp((size_t)n + 1), hj((size_t)std::max<int64_t>(E, 1)), own((size_t)m, -1);
    HIPCHK(h, hipMemcpy(hp.data(), h->sp_ptr, sizeof(int32_t) * (n + 1), hipMemcpyDeviceToHost));
    if (E > 0)
        HIPCHK(h, hipMemcpy(hj.data(), h->sp_j, sizeof(int32_t) * E, hipMemcpyDeviceToHost));
    // a slot per owned edge: in a moving row, the edges to static partners
    // and to partners of higher index, numbered in edge order
    std::vector<int32_t> ord((size_t)std::max<int64_t>(E, 1), -1), sp((size_t)n + 1, 0);
    for (int64_t i = 0; i < n; ++i)
    {
        int32_t c = sp[(size_t)i];
        if (!(h->flags[(size_t)h->perm[(size_t)i]] & BS_FLAG_STATIC))
            for (int32_t t = hp[(size_t)i]; t < hp[(size_t)i + 1]; ++t)
                if (hj[(size_t)t] < 0 || hj[(size_t)t] > (int32_t)i)
                    ord[(size_t)t] = c++;
        sp[(size_t)i + 1] = c;
    }
    const int64_t ns = sp[(size_t)n];
    // each spring with a moving end finds its edge in the owning row (the
    // table holds no repeated pair: every way in drops repeats)
#pragma omp parallel for schedule(static, 4096)
    for (int64_t s = 0; s < m; ++s)
    {
        int32_t a = h->iperm[h->s1[(size_t)s]], b = h->iperm[h->s2[(size_t)s]];
        const bool sa = (h->flags[h->s1[(size_t)s]] & BS_FLAG_STATIC) != 0;
        const bool sb = (h->flags[h->s2[(size_t)s]] & BS_FLAG_STATIC) != 0;
        if ((sa && sb) || a == b)
            continue;
        if (sa || (!sb && b < a))
            std::swap(a, b); // row a owns the edge to b
        int32_t lo = hp[(size_t)a], hi = hp[(size_t)a + 1];
        while (lo < hi)
        {
            const int32_t mid = (lo + hi) >> 1;
            if ((hj[(size_t)mid] & 0x7fffffff) < b)
                lo = mid + 1;
            else
                hi = mid;
        }
        if (lo < hp[(size_t)a + 1] && (hj[(size_t)lo] & 0x7fffffff) == b)
            own[(size_t)s] = lo;
    }
    std::vector<int32_t> so((size_t)m, -1);
    std::vector<char> used((size_t)std::max<int64_t>(ns, 1), 0);
    for (int64_t s = 0; s < m; ++s)
    {
        const int32_t e = own[(size_t)s];
        if (e < 0)
            continue;
        const int32_t r = ord[(size_t)e];
        if (r < 0 || used[(size_t)r])
            return fail(h, BS_ERR_STATE, "internal: spring %lld has no free slot", (long long)s);
        used[(size_t)r] = 1;
        so[(size_t)s] = r;
    }
    hipError_t e = hipSuccess;
    h->slot_ptr = dalloc<int32_t>((size_t)n + 1, e);
    HIPCHK(h, e);
    h->slot_of = dalloc<int32_t>((size_t)m, e);
    HIPCHK(h, e);
    HIPCHK(h, hipMemcpy(h->slot_ptr, sp.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->slot_of, so.data(), sizeof(int32_t) * m, hipMemcpyHostToDevice));
    h->nslots = ns;
    return BS_OK;
}

void launch_adj_pairs(bs_handle *h, const float4 *w, double4 *hw, double *dq);

// Terms the adjoint covers; BS_ERR_ARG otherwise.
int adjoint_check(bs_handle *h)
{
    const bs_params &P = h->p;
    if (P.imp_enable || P.hydrophobicity_enable)
        return fail(h, BS_ERR_ARG, "bs_adjoint: IMPALA and the pseudo-hydrophobic term have no adjoint");
    if (!h->bodies.empty())
        return fail(h, BS_ERR_ARG, "bs_adjoint: rigid bodies have no adjoint");
    if ((P.electrostatic_enable && P.potentialgrid_enable && h->grids[BS_GRID_POTENTIAL].set) ||
        (P.densitygrid_enable && h->grids[BS_GRID_DENSITY].set))
        return fail(h, BS_ERR_ARG, "bs_adjoint: field grids have no adjoint");
    if (P.compensated_positions)
        return fail(h, BS_ERR_ARG, "bs_adjoint: needs compensated_positions = 0");
    if (P.probe_enable)
        return fail(h, BS_ERR_ARG, "bs_adjoint: the probe particle has no adjoint");
    return BS_OK;
}

// Forward pass of the adjoint: T steps from the current state, which
// advances as with bs_step.  The state is saved (and the lists rebuilt)
// every K steps, so that recomputing a segment repeats it bit for bit, and
// the last segment's positions are kept as they pass.  The record ("tape")
// stays in the handle until the state changes.
int adjoint_record(bs_handle *h, int32_t T, int32_t checkpoint_every)
{
    auto &tp = h->tape;
    tp.ready = false;
    if (h->nsp > 0)
    { // the backward calls find the spring gradient slots built
        const int rc = build_spring_owners(h);
        if (rc)
            return rc;
    }
    const int64_t n = h->n;
    int32_t K = checkpoint_every;
    if (K <= 0)
    {
        // automatic: the whole trajectory as one stored segment (no
        // recomputation) when it takes at most half the free device memory;
        // else segments of ceil(sqrt(T)) steps, the least memory
        size_t mfree = 0, mtotal = 0;
        const double whole = (double)n * (32.0 + 16.0 * (double)std::max(T, 1));
        if (hipMemGetInfo(&mfree, &mtotal) == hipSuccess && whole <= 0.5 * (double)mfree)
            K = std::max(T, 1);
        else
            K = (int32_t)std::ceil(std::sqrt((double)std::max(T, 1)));
    }
    K = std::max(1, std::min(K, std::max(T, 1)));
    const int32_t ncp = (T + K - 1) / K;
    const size_t s4 = sizeof(float4) * (size_t)n;
    hipStream_t st = h->stream;
    hipError_t e = hipSuccess;
    // buffers grow as needed and are reused
    const int64_t need_cp = (int64_t)std::max(ncp, 1), need_seg = K, need_log = std::max(T, 1);
    if (need_cp > tp.cap_cp || need_seg > tp.cap_seg || need_log > tp.cap_log || !tp.fin_pos)
    {
        tape_free(tp);
        tp.cp_pos = dalloc<float4>((size_t)(n * need_cp), e);
        HIPCHK(h, e);
        tp.cp_vel = dalloc<float4>((size_t)(n * need_cp), e);
        HIPCHK(h, e);
        tp.cp_cnt = dalloc<int32_t>((size_t)(4 * need_cp), e);
        HIPCHK(h, e);
        tp.seg = dalloc<float4>((size_t)(n * need_seg), e);
        HIPCHK(h, e);
        tp.fin_pos = dalloc<float4>((size_t)n, e);
        HIPCHK(h, e);
        tp.fin_vel = dalloc<float4>((size_t)n, e);
        HIPCHK(h, e);
        tp.fin_cnt = dalloc<int32_t>(4, e);
        HIPCHK(h, e);
        tp.gate_log = dalloc<int32_t>((size_t)need_log, e);
        HIPCHK(h, e);
        tp.cap_cp = need_cp;
        tp.cap_seg = need_seg;
        tp.cap_log = need_log;
    }
    launch_fill_i32(tp.gate_log, 0, need_log, nullptr, st);
    const int32_t s_last = (ncp - 1) * K;
    for (int32_t s = 0; s < T; ++s)
    {
        if (s >= s_last)
            HIPCHK(h, hipMemcpyAsync(tp.seg + (size_t)(s - s_last) * n, h->pos[h->cur], s4, hipMemcpyDeviceToDevice,
                                     st));
        if (s % K == 0)
        {
            const size_t c = (size_t)(s / K);
            HIPCHK(h, hipMemcpyAsync(tp.cp_pos + c * n, h->pos[h->cur], s4, hipMemcpyDeviceToDevice, st));
            HIPCHK(h, hipMemcpyAsync(tp.cp_vel + c * n, h->vel, s4, hipMemcpyDeviceToDevice, st));
            HIPCHK(h, hipMemcpyAsync(tp.cp_cnt + 4 * c, h->counters, 4 * sizeof(int32_t), hipMemcpyDeviceToDevice,
                                     st));
            launch_fill_i32(h->status + 7, 1, 1, nullptr, st);
        }
        const int rc = prof_fold(h);
        if (rc)
            return rc;
        h->gate_log = tp.gate_log + s;
        launch_step(h, s + 1 == T);
        h->gate_log = nullptr;
    }
    if (T > 0)
    {
        const int rc = finish_call(h);
        if (rc)
            return rc;
        h->steps_done += T;
        h->last_call_forces_only = false;
    }
    HIPCHK(h, hipMemcpyAsync(tp.fin_pos, h->pos[h->cur], s4, hipMemcpyDeviceToDevice, st));
    HIPCHK(h, hipMemcpyAsync(tp.fin_vel, h->vel, s4, hipMemcpyDeviceToDevice, st));
    HIPCHK(h, hipMemcpyAsync(tp.fin_cnt, h->counters, 4 * sizeof(int32_t), hipMemcpyDeviceToDevice, st));
    {
        std::vector<int32_t> g((size_t)std::max(T, 1));
        HIPCHK(h, hipMemcpyAsync(g.data(), tp.gate_log, sizeof(int32_t) * g.size(), hipMemcpyDeviceToHost, st));
        HIPCHK(h, hipStreamSynchronize(st));
        tp.rebuilt.assign((size_t)T, 0);
        for (int32_t s = 0; s < T; ++s)
            tp.rebuilt[(size_t)s] = g[(size_t)s] != 0;
    }
    tp.T = T;
    tp.K = K;
    tp.ncp = ncp;
    tp.seg_of = ncp - 1;
    tp.ext = h->ext_uploaded;
    tp.ready = true;
    return BS_OK;
}

// Device buffers of one reverse sweep, laid out in 256-byte steps in one
// block: the adjoint state lx, lv (FP64); the weights w of the step being
// reversed and of the next (FP32); the pair terms' H w; the spring gradients
// (FP32 per edge between flushes, FP64 per slot, per spring at the end);
// dL/dq in internal and in particle order; the caller's seed as uploaded.
struct AdjBufs
{
    double4 *lx = nullptr, *lv = nullptr, *hwp = nullptr;
    float4 *w[2] = {nullptr, nullptr};
    float *dke = nullptr, *dle = nullptr;
    double *dks = nullptr, *dls = nullptr, *outk = nullptr, *outl = nullptr;
    double *dqd = nullptr, *dqo = nullptr, *sx = nullptr, *sv = nullptr;
    bool springs = false;
};

// Lays the sweep's buffers out from base (nullptr: sizes only) and returns
// the bytes they take.  springs: spring gradients wanted (the slots must be
// built); seed: room for a caller's dL/dx, dL/dv.
size_t adj_carve(const bs_handle *h, bool springs, bool seed, char *base, AdjBufs &b)
{
    const size_t n = (size_t)h->n, ns = (size_t)std::max<int64_t>(h->nslots, 0), m = (size_t)h->nsp;
    size_t off = 0;
    auto take = [&](auto *&ptr, size_t count) {
        using T = std::remove_pointer_t<std::remove_reference_t<decltype(ptr)>>;
        off = (off + 255) & ~(size_t)255;
        ptr = base ? reinterpret_cast<T *>(base + off) : nullptr;
        off += sizeof(T) * std::max<size_t>(count, 1);
    };
    b = AdjBufs{};
    b.springs = springs;
    take(b.lx, n);
    take(b.lv, n);
    take(b.w[0], n);
    take(b.w[1], n);
    if (!h->lists.empty())
        take(b.hwp, n);
    if (springs)
    {
        take(b.dke, ns);
        take(b.dle, ns);
        take(b.dks, ns);
        take(b.dls, ns);
        take(b.outk, m);
        take(b.outl, m);
    }
    take(b.dqd, n);
    take(b.dqo, n);
    if (seed)
    {
        take(b.sx, 3 * n);
        take(b.sv, 3 * n);
    }
    return off;
}

// Reverse sweep over the recorded forward pass from the seeded state at x_T
// (b.lx, b.lv, b.w[0]):
//   a = lv + dt lx,  w = dt/m a,  lx += H(x_s) w,  lv = (1 - gamma dt/m) a,
// segment by segment (earlier segments recomputed from their checkpoints),
// then the state goes back to x_T; the record stays usable.  Allocates
// nothing: every buffer comes in b.
constexpr int32_t ADJ_SPRING_LANES = 32;

int adjoint_sweep(bs_handle *h, AdjBufs &b, double *dk, double *dl0, double *dq)
{
    auto &tp = h->tape;
    const bs_params &P = h->p;
    const int64_t n = h->n, m = h->nsp;
    const int32_t T = tp.T, K = tp.K, ncp = tp.ncp;
    const size_t s4 = sizeof(float4) * (size_t)n;
    hipStream_t st = h->stream;
    int rc = BS_OK;
    if (tp.ext != h->ext_uploaded)
        return fail(h, BS_ERR_STATE, "bs_adjoint: external force changed since the forward pass");
    const bool want_springs = b.springs && m > 0;
    if (want_springs && h->nslots < 0)
        return fail(h, BS_ERR_STATE, "internal: spring gradient slots missing");
    double4 *hwp = b.hwp;
    float4 *wb[2] = {b.w[0], b.w[1]};
    float *dke = b.dke, *dle = b.dle;
    double *dks = b.dks, *dls = b.dls, *outk = b.outk, *outl = b.outl, *dqd = b.dqd;
    const int64_t ns = std::max<int64_t>(h->nslots, 0);
    // cleared by a fill kernel (no memset nodes, should a caller capture)
    if (want_springs)
    {
        launch_fill_i32(reinterpret_cast<int32_t *>(dke), 0, std::max<int64_t>(ns, 1), nullptr, st);
        launch_fill_i32(reinterpret_cast<int32_t *>(dle), 0, std::max<int64_t>(ns, 1), nullptr, st);
        launch_fill_i32(reinterpret_cast<int32_t *>(dks), 0, 2 * std::max<int64_t>(ns, 1), nullptr, st);
        launch_fill_i32(reinterpret_cast<int32_t *>(dls), 0, 2 * std::max<int64_t>(ns, 1), nullptr, st);
    }
    launch_fill_i32(reinterpret_cast<int32_t *>(dqd), 0, 2 * n, nullptr, st);
    constexpr int32_t FLUSH = 64;
    int32_t since_flush = 0;
    const double dt = (double)P.timestep, gam = P.viscosity_enable ? (double)P.viscosity_value : 0.0;
    const double hs = 0.5 * std::max(P.list_skin, 0.0f);
    const float half_skin2 = (float)(hs * hs * (1.0 - 1e-5));
    int wc = 0;
    for (int32_t c = ncp - 1; c >= 0; --c)
    {
        const int32_t s0 = c * K, s1 = std::min(T, s0 + K);
        // a segment other than the one seg holds (every segment but the last
        // on the first sweep; all of them on later sweeps of the record) is
        // recomputed from its checkpoint, bit for bit as the forward pass
        const bool recompute = tp.seg_of != c;
        if (recompute)
        {
            tp.seg_of = -1; // partly overwritten until the segment is complete
            HIPCHK(h, hipMemcpyAsync(h->pos[h->cur], tp.cp_pos + (size_t)c * n, s4, hipMemcpyDeviceToDevice, st));
            HIPCHK(h, hipMemcpyAsync(h->vel, tp.cp_vel + (size_t)c * n, s4, hipMemcpyDeviceToDevice, st));
            HIPCHK(h, hipMemcpyAsync(h->counters, tp.cp_cnt + 4 * (size_t)c, 4 * sizeof(int32_t),
                                     hipMemcpyDeviceToDevice, st));
            launch_fill_i32(h->status + 7, 1, 1, nullptr, st);
        }
        // the recomputed steps rebuild the lists where the recorded steps
        // did (the same states give the same decisions), so a step that did
        // not is launched without its gate and gated rebuild
        for (int32_t s = s0; s < s1 && recompute; ++s)
        {
            HIPCHK(h, hipMemcpyAsync(tp.seg + (size_t)(s - s0) * n, h->pos[h->cur], s4, hipMemcpyDeviceToDevice, st));
            if (s + 1 < s1)
            {
                rc = prof_fold(h);
                if (rc)
                    return rc;
                if (tp.rebuilt[(size_t)s] || s == s0)
                {
                    launch_fill_i32(h->status + 7, 1, 1, nullptr, st);
                    launch_step(h, false);
                }
                else
                    launch_step_unlisted(h);
            }
        }
        if (recompute)
            tp.seg_of = c;
        // The reversed steps start from the lists the recomputation left
        // (rebuilt from its checkpoint like the forward pass, so the same
        // on every sweep).  The last segment is the one a first sweep takes
        // from the record, with whatever lists the record's last step
        // left: its lists are rebuilt at its first reversed state, on every
        // sweep, so that two sweeps of one record give the same bits.
        if (!recompute || c == ncp - 1)
            launch_fill_i32(h->status + 7, 1, 1, nullptr, st);
        // After the first reversed state, one test of all the segment's other
        // states against the lists' reference positions: when none drifted,
        // no rebuild would happen, so their drift tests and gated rebuild
        // launches are skipped (same lists, same bits) and the kernels read
        // the stored positions in place.
        bool still = false;
        struct PosRestore
        {
            bs_handle *h;
            int cur;
            float4 *p;
            ~PosRestore() { h->pos[cur] = p; }
        } restore{h, h->cur, h->pos[h->cur]};
        for (int32_t s = s1 - 1; s >= s0; --s)
        {
            float4 *xs = tp.seg + (size_t)(s - s0) * n;
            if (still)
                h->pos[h->cur] = xs;
            else
                HIPCHK(h, hipMemcpyAsync(restore.p, xs, s4, hipMemcpyDeviceToDevice, st));
            rc = prof_fold(h);
            if (rc)
                return rc;
            const size_t e0 = h->prof ? prof_mark(h) : SIZE_MAX;
            if (!still)
            {
                launch_drift_check(h->pos[h->cur], h->xref, (int32_t)n, half_skin2, h->status, st);
                rebuild_chain(h, true);
            }
            if (s == s1 - 1 && s > s0 && !h->lists.empty())
            {
                launch_fill_i32(h->status + 4, 0, 1, nullptr, st);
                launch_drift_states(tp.seg, s - s0, (int32_t)n, h->xref, half_skin2, h->status + 4, st);
                int32_t moved = 1;
                HIPCHK(h, hipMemcpyAsync(&moved, h->status + 4, sizeof moved, hipMemcpyDeviceToHost, st));
                HIPCHK(h, hipStreamSynchronize(st));
                still = moved == 0;
            }
            const size_t e1 = h->prof ? prof_mark(h) : SIZE_MAX;
            if (hwp)
                launch_adj_pairs(h, wb[wc], hwp, dqd);
            const size_t e2 = h->prof ? prof_mark(h) : SIZE_MAX;
            AdjSpringArgs a{};
            a.pos = h->pos[h->cur];
            a.jprm = h->jprm;
            a.iprm = h->iprm;
            a.sp_ptr = h->sp_ptr;
            a.sp_j = h->sp_j;
            a.sp_k = h->k_uniform ? nullptr : h->sp_k;
            a.k_uni = h->k_uni;
            a.sp_l0 = h->sp_l0;
            a.w = wb[wc];
            a.hwp = hwp;
            a.lx = b.lx;
            a.lv = b.lv;
            a.w_next = wb[wc ^ 1];
            a.dk = want_springs ? dke : nullptr;
            a.dl0 = want_springs ? dle : nullptr;
            a.slot_ptr = h->slot_ptr;
            a.n = (int32_t)n;
            a.spring_enable = P.spring_enable;
            a.spring_scale = P.spring_scale;
            a.dt = dt;
            a.gamma = gam;
            a.k = h->k;
            // 32 lanes a row: 0.265 ms against 0.273 with 16 and 0.357 with 8
            // at 100k beads (the forward's kernel prefers 16)
            launch_adj_spring(a, ADJ_SPRING_LANES, (int32_t)((n * ADJ_SPRING_LANES + BLOCK - 1) / BLOCK), st);
            if (want_springs && (++since_flush == FLUSH || s == 0))
            {
                launch_spring_grads(dke, dle, ns, dks, dls, st);
                since_flush = 0;
            }
            if (h->prof)
            {
                const size_t e3 = prof_mark(h);
                prof_span(h, BS_PROF_REBUILD, e0, e1);
                if (hwp)
                    prof_span(h, BS_PROF_ADJ_PAIR, e1, e2);
                prof_span(h, BS_PROF_ADJ_SPRING, e2, e3);
            }
            wc ^= 1;
        }
    }
    // back to x_T (lists rebuilt at the next step)
    HIPCHK(h, hipMemcpyAsync(h->pos[h->cur], tp.fin_pos, s4, hipMemcpyDeviceToDevice, st));
    HIPCHK(h, hipMemcpyAsync(h->vel, tp.fin_vel, s4, hipMemcpyDeviceToDevice, st));
    HIPCHK(h, hipMemcpyAsync(h->counters, tp.fin_cnt, 4 * sizeof(int32_t), hipMemcpyDeviceToDevice, st));
    launch_fill_i32(h->status + 7, 1, 1, nullptr, st);

    if (want_springs)
        launch_spring_out(h->slot_of, m, dks, dls, outk, outl, st);
    if (dq)
        launch_unperm_f64(dqd, h->perm_d, (int32_t)n, b.dqo, st);
    int32_t stv[4];
    HIPCHK(h, hipMemcpyAsync(stv, h->status, sizeof stv, hipMemcpyDeviceToHost, st));
    if (dq)
        HIPCHK(h, hipMemcpyAsync(dq, b.dqo, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    if (want_springs && dk)
        HIPCHK(h, hipMemcpyAsync(dk, outk, sizeof(double) * m, hipMemcpyDeviceToHost, st));
    if (want_springs && dl0)
        HIPCHK(h, hipMemcpyAsync(dl0, outl, sizeof(double) * m, hipMemcpyDeviceToHost, st));
    HIPCHK(h, hipStreamSynchronize(st));
    prof_collect(h);
    if (stv[2] != 0)
        return fail(h, BS_ERR_STATE, "neighbour list capacity exceeded (max %d > cap); rerun bs_setup", stv[3]);
    return BS_OK;
}

// H w of the pair terms at the positions in pos[cur] (lists valid there),
// into hw, and their dL/dq into dq.
void launch_adj_pairs(bs_handle *h, const float4 *w, double4 *hw, double *dq)
{
    for (size_t li = 0; li < h->lists.size(); ++li)
    {
        const List &l = h->lists[li];
        AdjPairArgs a{};
        a.pos = h->pos[h->cur];
        a.jprm = h->jprm;
        a.cl = l.cl;
        a.cln = l.cln;
        a.w = w;
        a.hw = hw;
        a.dq = dq;
        a.n = (int32_t)h->n;
        a.ncl = h->ncl;
        a.cap = l.cap;
        a.accumulate = li > 0 ? 1 : 0;
        a.queues = h->p.pair_layout == BS_PAIR_QUEUES ? 1 : 0;
        a.rc = l.rc;
        a.rc2lo = rc2(l.rc, 1.0 - 4e-6);
        a.rc2hi = rc2(l.rc, 1.0 + 4e-6);
        a.steric_scale = h->p.steric_gridscale;
        a.kf = h->k.c_elec / (4.0 * h->k.pi * (double)h->p.electrostatic_dielectric) *
               (double)h->p.electrostatic_scale;
        a.i_mask = l.imask;
        a.j_mask = l.jmask;
        a.k = h->k;
        const int32_t sm = (l.terms & TERM_STERIC) ? h->p.steric_mode : -1;
        const int32_t em =
            (l.terms & TERM_ELEC) ? (h->p.dielectric_model == BS_DIELECTRIC_DISTANCE ? 2 : 1) : 0;
        launch_adj_pair(a, sm, em, h->pair_blocks, h->stream);
    }
}

int do_setup(bs_handle *h)
{
    if (!h->have_particles)
        return fail(h, BS_ERR_STATE, "bs_setup: no particles");
    free_device(h);
    h->cur = 0;
    const int64_t n = h->n;
    if (n > (int64_t)CL_JMASK)
        return fail(h, BS_ERR_ARG, "bs_setup: more than %u particles", CL_JMASK);
    make_order(h);
    hipError_t e;
#define ALLOC(ptr, T, cnt)                                                                                        \
    do                                                                                                            \
    {                                                                                                             \
        ptr = dalloc<T>((size_t)(cnt), e);                                                                        \
        HIPCHK(h, e);                                                                                             \
    } while (0)
    for (int b = 0; b < 2; ++b)
    {
        ALLOC(h->pos[b], float4, n);
        if (h->p.compensated_positions)
            ALLOC(h->plo[b], float4, n);
    }
    ALLOC(h->vel, float4, n);
    ALLOC(h->fnb, double4, n);
    ALLOC(h->jprm, float4, n);
    ALLOC(h->iprm, float4, n);
    ALLOC(h->xref, float4, n);
    ALLOC(h->fext, float4, n);
    ALLOC(h->fout, float4, n);
    ALLOC(h->perm_d, int32_t, n);
    ALLOC(h->cell_of, int32_t, n);
    ALLOC(h->ctmp, int32_t, n);
    ALLOC(h->cidx, int32_t, n);
    ALLOC(h->status, int32_t, 8);
    ALLOC(h->counters, int32_t, 4);
    ALLOC(h->energies_d, double, BS_NUM_ENERGIES);
    h->box_blocks = (int32_t)std::min<int64_t>(BLOCK, (n + BLOCK - 1) / BLOCK); // one per final-reduce thread
    ALLOC(h->box_part, double, 6 * h->box_blocks);
    ALLOC(h->box, double, 6);

    // per-particle parameters
    std::vector<float4> jp((size_t)n), ip((size_t)n);
    for (int64_t t = 0; t < n; ++t)
    {
        const int64_t i = h->perm[(size_t)t];
        jp[(size_t)t] = jprm_of(h, i);
        ip[(size_t)t] = make_float4(h->mass[(size_t)i], h->sasa[(size_t)i], h->transfer[(size_t)i],
                                    h->burying.empty() ? 1.0f : h->burying[(size_t)i]);
    }
    HIPCHK(h, hipMemcpyAsync(h->jprm, jp.data(), sizeof(float4) * n, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->iprm, ip.data(), sizeof(float4) * n, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->perm_d, h->perm.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));

    // Spring CSR on the device from the device copy of the spring table:
    // both directions, rows sorted by partner (internal ids).  The full rows
    // give each i-cluster's exclusions; the force rows leave out springs
    // between two rigid particles (RigidBody::UpdateSpringsState,
    // RigidBody.cpp:42-53; Spring.cpp:24) while the exclusions keep every
    // spring.
    int rc = springs_device(h);
    if (rc)
        return rc;
    const int64_t m = h->nsp;
    if (m > (int64_t)(INT32_MAX / 2))
        return fail(h, BS_ERR_ARG, "bs_setup: more than 2^30 springs");
    hipStream_t st = h->stream;
    Scratch S;
    int32_t *iperm_d = S.get<int32_t>((size_t)n, e);
    HIPCHK(h, e);
    int32_t *cnt = S.get<int32_t>((size_t)n, e);
    HIPCHK(h, e);
    HIPCHK(h, hipMemcpyAsync(iperm_d, h->iperm.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, st));
    const bool any_rigid = !h->bodies.empty();
    int32_t *fptr = nullptr; // full rows
    if (any_rigid)
    {
        fptr = S.get<int32_t>((size_t)n + 1, e);
        HIPCHK(h, e);
    }
    else
    {
        ALLOC(h->sp_ptr, int32_t, n + 1);
        fptr = h->sp_ptr;
    }
    launch_fill_i32(cnt, 0, n, nullptr, st);
    launch_edge_count(h->dsa, h->dsb, m, iperm_d, cnt, st);
    launch_scan(cnt, fptr, (int32_t)n, nullptr, st);
    int32_t efull = 0;
    HIPCHK(h, hipMemcpyAsync(&efull, fptr + n, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HIPCHK(h, hipStreamSynchronize(st));
    uint64_t *key = S.get<uint64_t>((size_t)efull, e);
    HIPCHK(h, e);
    int32_t *fej = nullptr;
    float *fek = nullptr, *fel = nullptr;
    if (any_rigid)
    {
        fej = S.get<int32_t>((size_t)efull, e);
        HIPCHK(h, e);
        fek = S.get<float>((size_t)efull, e);
        HIPCHK(h, e);
        fel = S.get<float>((size_t)efull, e);
        HIPCHK(h, e);
    }
    else
    {
        ALLOC(h->sp_j, int32_t, efull);
        ALLOC(h->sp_k, float, efull);
        ALLOC(h->sp_l0, float, efull);
        fej = h->sp_j;
        fek = h->sp_k;
        fel = h->sp_l0;
    }
    launch_fill_i32(cnt, 0, n, nullptr, st);
    launch_edge_fill(h->dsa, h->dsb, m, iperm_d, fptr, cnt, key, st);
    launch_edge_sort(fptr, (int32_t)n, key, h->jprm, h->dsl, h->dsk, fej, fek, fel, st);
    // static exclusions per i-cluster: the union of its 8 rows as sorted
    // (j, bonded-mask) words; springs never change during a run
    {
        const int32_t ncl = (int32_t)((n + CLUSTER - 1) / CLUSTER);
        uint32_t *merged = S.get<uint32_t>((size_t)efull, e);
        HIPCHK(h, e);
        int32_t *ucnt = S.get<int32_t>((size_t)ncl, e);
        HIPCHK(h, e);
        launch_ex_merge(fptr, fej, (int32_t)n, ncl, merged, ucnt, st);
        ALLOC(h->ex_ptr, int32_t, ncl + 1);
        launch_scan(ucnt, h->ex_ptr, ncl, nullptr, st);
        int32_t nex = 0;
        HIPCHK(h, hipMemcpyAsync(&nex, h->ex_ptr + ncl, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        HIPCHK(h, hipStreamSynchronize(st));
        ALLOC(h->ex_w, uint32_t, nex);
        launch_ex_fill(fptr, (int32_t)n, ncl, merged, h->ex_ptr, h->ex_w, st);
    }
    h->nedges = efull;
    if (any_rigid)
    {
        std::vector<uint8_t> rg((size_t)n);
        for (int64_t t = 0; t < n; ++t)
            rg[(size_t)t] = h->body_of[(size_t)h->perm[(size_t)t]] >= 0 ? 1 : 0;
        uint8_t *rg_d = S.get<uint8_t>((size_t)n, e);
        HIPCHK(h, e);
        HIPCHK(h, hipMemcpyAsync(rg_d, rg.data(), (size_t)n, hipMemcpyHostToDevice, st));
        launch_edge_keep(fptr, (int32_t)n, fej, fek, fel, rg_d, nullptr, cnt, nullptr, nullptr, nullptr, st);
        ALLOC(h->sp_ptr, int32_t, n + 1);
        launch_scan(cnt, h->sp_ptr, (int32_t)n, nullptr, st);
        int32_t ekeep = 0;
        HIPCHK(h, hipMemcpyAsync(&ekeep, h->sp_ptr + n, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        HIPCHK(h, hipStreamSynchronize(st)); // also: rg outlives its copy
        ALLOC(h->sp_j, int32_t, ekeep);
        ALLOC(h->sp_k, float, ekeep);
        ALLOC(h->sp_l0, float, ekeep);
        launch_edge_keep(fptr, (int32_t)n, fej, fek, fel, rg_d, h->sp_ptr, nullptr, h->sp_j, h->sp_k, h->sp_l0, st);
        h->nedges = ekeep;
    }
    // one stiffness for every spring (a network built from a cutoff, an
    // elastic network model): the spring kernels read no per-edge k
    h->k_uniform = false;
    if (m > 0)
    {
        int32_t *diff = S.get<int32_t>(1, e);
        HIPCHK(h, e);
        launch_fill_i32(diff, 0, 1, nullptr, st);
        launch_differs_f32(h->dsk, m, diff, st);
        int32_t dv = 1;
        HIPCHK(h, hipMemcpyAsync(&dv, diff, sizeof dv, hipMemcpyDeviceToHost, st));
        HIPCHK(h, hipMemcpyAsync(&h->k_uni, h->dsk, sizeof(float), hipMemcpyDeviceToHost, st));
        HIPCHK(h, hipStreamSynchronize(st));
        h->k_uniform = dv == 0;
    }
    HIPCHK(h, hipStreamSynchronize(st)); // the scratch is freed on return
    rc = upload_positions(h);
    if (rc)
        return rc;

    // rigid bodies: state from the current positions on first use
    // (RigidBody::init, RigidBody.cpp:57-124), slots for their dynamic members
    if (!h->bodies.empty())
    {
        const int32_t nb = (int32_t)h->bodies.size();
        std::vector<RigidBody> st((size_t)nb);
        std::vector<int32_t> sp, sb, bs((size_t)nb + 1, 0);
        std::vector<double4> p0v, rxv;
        for (int32_t b = 0; b < nb; ++b)
        {
            auto &B = h->bodies[(size_t)b];
            if (!B.init)
            {
                RigidBody r{};
                double M = 0.0, X[3] = {0, 0, 0};
                for (int64_t id : B.ids)
                    M += h->mass[(size_t)id];
                for (int64_t id : B.ids)
                    for (int c = 0; c < 3; ++c)
                        X[c] += (double)h->x[3 * id + c] * h->mass[(size_t)id];
                r.M = M;
                for (int c = 0; c < 3; ++c)
                    r.X[c] = X[c] / M;
                const double ib = (2.0 / 5.0) * M; // I_body = 2/5 M dmax^2 with dmax = 1
                r.iinv[0] = r.iinv[4] = r.iinv[8] = 1.0 / ib;
                r.q[3] = 1.0;
                B.p0.resize(3 * B.ids.size());
                for (size_t t = 0; t < B.ids.size(); ++t)
                    for (int c = 0; c < 3; ++c)
                        B.p0[3 * t + c] = (double)h->x[3 * B.ids[t] + c] - r.X[c];
                B.st = r;
                B.init = true;
            }
            st[(size_t)b] = B.st;
            bs[(size_t)b] = (int32_t)sp.size();
            for (size_t t = 0; t < B.ids.size(); ++t)
            {
                const int64_t id = B.ids[t];
                if (h->flags[(size_t)id] & BS_FLAG_STATIC)
                    continue;
                sp.push_back(h->iperm[(size_t)id]);
                sb.push_back(b);
                p0v.push_back(make_double4(B.p0[3 * t], B.p0[3 * t + 1], B.p0[3 * t + 2], 0.0));
                rxv.push_back(make_double4(h->x[3 * id], h->x[3 * id + 1], h->x[3 * id + 2], 0.0));
            }
        }
        bs[(size_t)nb] = (int32_t)sp.size();
        h->nr = (int32_t)sp.size();
        ALLOC(h->bodies_d, RigidBody, nb);
        ALLOC(h->bstart, int32_t, nb + 1);
        ALLOC(h->frig, double4, n);
        ALLOC(h->slot_p, int32_t, h->nr);
        ALLOC(h->slot_b, int32_t, h->nr);
        ALLOC(h->rp0, double4, h->nr);
        ALLOC(h->rx, double4, h->nr);
        HIPCHK(h, hipMemcpy(h->bodies_d, st.data(), sizeof(RigidBody) * nb, hipMemcpyHostToDevice));
        HIPCHK(h, hipMemcpy(h->bstart, bs.data(), sizeof(int32_t) * (nb + 1), hipMemcpyHostToDevice));
        if (h->nr > 0)
        {
            HIPCHK(h, hipMemcpy(h->slot_p, sp.data(), sizeof(int32_t) * h->nr, hipMemcpyHostToDevice));
            HIPCHK(h, hipMemcpy(h->slot_b, sb.data(), sizeof(int32_t) * h->nr, hipMemcpyHostToDevice));
            HIPCHK(h, hipMemcpy(h->rp0, p0v.data(), sizeof(double4) * h->nr, hipMemcpyHostToDevice));
            HIPCHK(h, hipMemcpy(h->rx, rxv.data(), sizeof(double4) * h->nr, hipMemcpyHostToDevice));
        }
    }
    for (int g = 0; g < 2; ++g)
        if (h->grids[g].set)
        {
            const auto &G = h->grids[g];
            ALLOC(h->grid_d[g], float4, G.cells.size());
            HIPCHK(h, hipMemcpy(h->grid_d[g], G.cells.data(), sizeof(float4) * G.cells.size(), hipMemcpyHostToDevice));
        }
    h->cent_blocks = (int32_t)std::min<int64_t>(1024, (n + BLOCK - 1) / BLOCK);
    ALLOC(h->cent_part, double, 3 * h->cent_blocks);
    ALLOC(h->cent, double, 3);

    // energy segments: at most SEG contiguous particles of one system each
    {
        constexpr int64_t SEG = 2048;
        const int32_t nsys = (int32_t)h->sys_off.size() - 1;
        std::vector<int32_t> ss, sg((size_t)nsys + 1, 0);
        for (int32_t sy = 0; sy < nsys; ++sy)
        {
            sg[(size_t)sy] = (int32_t)ss.size();
            for (int64_t b = h->sys_off[(size_t)sy]; b < h->sys_off[(size_t)sy + 1]; b += SEG)
                ss.push_back((int32_t)b);
        }
        sg[(size_t)nsys] = (int32_t)ss.size();
        h->nseg = (int32_t)ss.size();
        ss.push_back((int32_t)n);
        ALLOC(h->seg_start, int32_t, ss.size());
        ALLOC(h->sys_seg, int32_t, sg.size());
        ALLOC(h->seg_part, double, 6 * std::max(h->nseg, 1));
        ALLOC(h->sys_E, double, 6 * nsys);
        HIPCHK(h, hipMemcpy(h->seg_start, ss.data(), sizeof(int32_t) * ss.size(), hipMemcpyHostToDevice));
        HIPCHK(h, hipMemcpy(h->sys_seg, sg.data(), sizeof(int32_t) * sg.size(), hipMemcpyHostToDevice));
    }

    // status / counters
    const int32_t st0[8] = {0, INT_MAX, 0, 0, 0, 0, 0, 1}, ct0[4] = {0, 0, 0, 0};
    HIPCHK(h, hipMemcpyAsync(h->status, st0, sizeof st0, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->counters, ct0, sizeof ct0, hipMemcpyHostToDevice, h->stream));

    // kernel geometry: pair kernels run one wave per i-cluster
    h->ncl = (int32_t)((n + CLUSTER - 1) / CLUSTER);
    h->pair_blocks = (h->ncl + BLOCK / 64 - 1) / (BLOCK / 64);
    h->spring_blocks = (int32_t)((n * h->group + BLOCK - 1) / BLOCK);
    ALLOC(h->epp, double4, n);
    ALLOC(h->ese, double4, n);

    // the probe particle: sized for the widest spring launch (32 lanes a row)
    if (h->p.probe_enable)
    {
        if (h->sys_off.size() > 2)
            return fail(h, BS_ERR_ARG, "bs_setup: the probe particle belongs to a single network, not to a batch of "
                                       "systems");
        if (!h->probe.placed)
        { // SpringNetwork::_setupProbe (SpringNetwork.cpp:749-766): at probe.x/y/z, at rest
            const float x[3] = {h->p.probe_x, h->p.probe_y, h->p.probe_z};
            for (int c = 0; c < 3; ++c)
            {
                h->probe.x[c] = x[c];
                h->probe.v[c] = h->probe.f[c] = 0.0f;
            }
            h->probe.placed = true;
        }
        ALLOC(h->probe_d, ProbeDev, 1);
        ALLOC(h->probe_part, double, 8 * ((n * 32 + BLOCK - 1) / BLOCK));
        ProbeDev pd{};
        pd.x = make_float4(h->probe.x[0], h->probe.x[1], h->probe.x[2], h->p.probe_charge);
        pd.v = make_float4(h->probe.v[0], h->probe.v[1], h->probe.v[2], 0.0f);
        pd.f = make_double4(h->probe.f[0], h->probe.f[1], h->probe.f[2], 0.0);
        HIPCHK(h, hipMemcpy(h->probe_d, &pd, sizeof pd, hipMemcpyHostToDevice));
    }

    // neighbour lists: a counting pass sizes the capacity
    plan_lists(h);
    std::vector<float4> shift;
    if (h->sys_off.size() > 2)
    {
        float maxr = 0.0f;
        for (const auto &l : h->lists)
            maxr = std::max(maxr, l.radius);
        shift = system_shifts(h, maxr);
        ALLOC(h->shift, float4, n);
        HIPCHK(h, hipMemcpy(h->shift, shift.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
    }
    int32_t maxcells = 1;
    for (auto &l : h->lists)
    {
        plan_grid(h, l, shift);
        maxcells = std::max(maxcells, l.ncells);
    }
    rc = ensure_cells(h, maxcells);
    if (rc)
        return rc;
    if (!h->lists.empty())
        launch_box_reduce(h->pos[h->cur], h->shift, (int32_t)n, h->box_part, h->box, h->box_blocks, nullptr,
                          h->stream);
    for (auto &l : h->lists)
    {
        ALLOC(l.cln, int32_t, h->ncl);
        launch_fill_i32(h->status + 3, 0, 1, nullptr, h->stream);
        build_list(h, l, false, 0);
        int32_t st[4];
        HIPCHK(h, hipMemcpyAsync(st, h->status, sizeof st, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        const int32_t mx = st[3];
        l.cap = (std::max<int32_t>(16, (int32_t)std::ceil(mx * 1.25) + 64) + 15) / 16 * 16;
        ALLOC(l.cl, uint32_t, (int64_t)h->ncl * l.cap);
        if (h->p.pair_layout == BS_PAIR_ROWS)
        {
            const int64_t nrb = (n + 63) / 64;
            l.rcap = l.cap; // a row is a subset of its cluster's list
            ALLOC(l.rows, uint32_t, nrb * 64 * l.rcap);
            ALLOC(l.rlen, int32_t, nrb * 64);
        }
        launch_fill_i32(h->status + 2, 0, 1, nullptr, h->stream);
    }
    const int32_t st1[8] = {0, INT_MAX, 0, 0, 0, 0, 0, 1};
    HIPCHK(h, hipMemcpyAsync(h->status, st1, sizeof st1, hipMemcpyHostToDevice, h->stream));
    rebuild_chain(h, true);
    HIPCHK(h, hipStreamSynchronize(h->stream));
#undef ALLOC
    h->ready = true;
    h->cur = 0;
    return BS_OK;
}

int ensure_ready(bs_handle *h)
{
    if (!h->ready)
    {
        const int rc = do_setup(h);
        if (rc)
            return rc;
    }
    return check_device(h);
}

int kind_term(int kind) { return kind == BS_NS_ELECTROSTATIC ? TERM_ELEC : kind == BS_NS_HYDROPHOBIC ? TERM_HYDRO : TERM_STERIC; }

// Before a re-setup: the device state (positions, velocities, body states)
// back into the host copies, then the device buffers released.
int pull_and_release(bs_handle *h)
{
    if (!h->ready)
        return BS_OK;
    int rc = bs_get_state(h, h->x.data(), h->v.data());
    if (rc)
        return rc;
    if (h->probe_d)
    {
        ProbeDev pd;
        HIPCHK(h, hipMemcpy(&pd, h->probe_d, sizeof pd, hipMemcpyDeviceToHost));
        const float x[3] = {pd.x.x, pd.x.y, pd.x.z}, v[3] = {pd.v.x, pd.v.y, pd.v.z};
        const double f[3] = {pd.f.x, pd.f.y, pd.f.z};
        for (int c = 0; c < 3; ++c)
        {
            h->probe.x[c] = x[c];
            h->probe.v[c] = v[c];
            h->probe.f[c] = (float)f[c];
        }
    }
    if (!h->bodies.empty() && h->bodies_d)
    {
        std::vector<RigidBody> st(h->bodies.size());
        HIPCHK(h, hipMemcpy(st.data(), h->bodies_d, sizeof(RigidBody) * st.size(), hipMemcpyDeviceToHost));
        for (size_t b = 0; b < st.size(); ++b)
            h->bodies[b].st = st[b];
    }
    free_device(h);
    return BS_OK;
}

} // namespace

// ===========================================================================
extern "C" {

void bs_default_params(bs_params *p)
{
    if (p)
        bs_params_fill_defaults(p);
}

int bs_create(const bs_params *p, bs_handle **out)
{
    if (!out)
        return BS_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return BS_ERR_NOGPU;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess)
        return BS_ERR_NOGPU;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess)
        return BS_ERR_NOGPU;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return BS_ERR_NOGPU;
    bs_handle *h = new (std::nothrow) bs_handle();
    if (!h)
        return BS_ERR_ARG;
    if (p)
        h->p = *p;
    else
        bs_params_fill_defaults(&h->p);
    h->k = make_consts();
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
    {
        delete h;
        return BS_ERR_HIP;
    }
    h->device = dev;
    noalloc_setup(dev);
    *out = h;
    return BS_OK;
}

int bs_destroy(bs_handle *h)
{
    if (!h)
        return BS_OK;
    if (h->stream)
        (void)hipStreamSynchronize(h->stream);
    traj_release(h);
    for (hipEvent_t e : h->evpool)
        (void)hipEventDestroy(e);
    free_device(h);
    for (void *q : {(void *)h->dsa, (void *)h->dsb, (void *)h->dsl, (void *)h->dsk})
        if (q)
            (void)hipFree(q);
    if (h->stream)
        (void)hipStreamDestroy(h->stream);
    delete h;
    return BS_OK;
}

const char *bs_last_error(const bs_handle *h)
{
    if (!h)
        return "no handle (bs_create failed or was not called)";
    return h->err.c_str();
}

int64_t bs_num_particles(const bs_handle *h) { return h ? h->n : 0; }
int64_t bs_num_springs(const bs_handle *h) { return h ? h->nsp : 0; }

int bs_set_particles(bs_handle *h, int64_t n, const float *x3, const float *v3, const float *mass,
                     const float *charge, const float *radius, const float *epsilon, const float *sasa,
                     const float *transfer, const float *hydrophobicity, const uint8_t *flags)
{
    if (!h)
        return BS_ERR_ARG;
    if (n <= 0 || !x3)
        return fail(h, BS_ERR_ARG, "bs_set_particles: empty particle list");
    traj_release(h); // writers follow one particle set
    h->labels = bs::TrajLabels{};
    h->ivec1 = h->ivec2 = -1;
    free_device(h);
    h->n = n;
    auto take = [n](std::vector<float> &dst, const float *src, int64_t per, float dflt) {
        dst.assign((size_t)(n * per), dflt);
        if (src)
            std::memcpy(dst.data(), src, sizeof(float) * n * per);
    };
    take(h->x, x3, 3, 0.0f);
    take(h->v, v3, 3, 0.0f);
    take(h->mass, mass, 1, 1.0f);
    take(h->charge, charge, 1, 0.0f);
    take(h->radius, radius, 1, 1.0f);
    take(h->eps, epsilon, 1, 0.0f);
    take(h->sasa, sasa, 1, 0.0f);
    take(h->transfer, transfer, 1, 0.0f);
    take(h->hydro, hydrophobicity, 1, 0.0f);
    h->flags.assign((size_t)n, 0);
    if (flags)
        std::memcpy(h->flags.data(), flags, (size_t)n);
    h->bodies.clear();
    h->body_of.assign((size_t)n, -1);
    h->burying.assign((size_t)n, 1.0f);
    h->sys_off = {0, n};
    h->s1.clear();
    h->s2.clear();
    h->sl0.clear();
    h->sk.clear();
    springs_host_changed(h);
    h->have_particles = true;
    h->steps_done = 0;
    h->nonfinite = -1;
    h->probe.placed = false; // the next setup places the probe anew
    return BS_OK;
}

int bs_set_springs(bs_handle *h, int64_t m, const uint32_t *id1, const uint32_t *id2, const float *equilibrium,
                   const float *stiffness)
{
    if (!h)
        return BS_ERR_ARG;
    if (!h->have_particles)
        return fail(h, BS_ERR_STATE, "bs_set_springs: particles must be set before springs");
    if (m < 0 || (m > 0 && (!id1 || !id2 || !equilibrium || !stiffness)))
        return fail(h, BS_ERR_ARG, "bs_set_springs: bad arrays");
    // SpringNetwork::addSpring: range and self checks, duplicates ignored.
    // The table is built aside, so that a rejected call leaves h as it was.
    std::vector<std::vector<uint32_t>> seen((size_t)h->n);
    std::vector<uint32_t> s1, s2;
    std::vector<float> l0v, kv;
    s1.reserve((size_t)m);
    s2.reserve((size_t)m);
    l0v.reserve((size_t)m);
    kv.reserve((size_t)m);
    for (int64_t s = 0; s < m; ++s)
    {
        const uint32_t a = id1[s], b = id2[s];
        if (a >= (uint64_t)h->n || b >= (uint64_t)h->n)
            return fail(h, BS_ERR_ARG, "SpringNetwork::addSpring: particle index out of range (spring %lld)",
                        (long long)s);
        if (a == b)
            return fail(h, BS_ERR_ARG, "SpringNetwork::addSpring: a spring requires two distinct particles");
        auto &sa = seen[a];
        if (std::find(sa.begin(), sa.end(), b) != sa.end())
            continue;
        sa.push_back(b);
        seen[b].push_back(a);
        s1.push_back(a);
        s2.push_back(b);
        l0v.push_back(equilibrium[s]);
        kv.push_back(stiffness[s]);
    }
    // the current (device) state stays: the step rebuilds around it
    int rc = pull_and_release(h);
    if (rc)
        return rc;
    h->s1.swap(s1);
    h->s2.swap(s2);
    h->sl0.swap(l0v);
    h->sk.swap(kv);
    springs_host_changed(h);
    h->have_springs = true;
    return BS_OK;
}

int bs_set_params(bs_handle *h, const bs_params *p)
{
    if (!h || !p)
        return BS_ERR_ARG;
    // keep the current state, then rebuild lazily with the new parameters
    int rc = pull_and_release(h);
    if (rc)
        return rc;
    h->p = *p;
    return BS_OK;
}

int bs_setup(bs_handle *h)
{
    if (!h)
        return BS_ERR_ARG;
    return do_setup(h);
}

int bs_step(bs_handle *h, int32_t nsteps, const float *ext_force3)
{
    if (!h)
        return BS_ERR_ARG;
    if (nsteps < 0)
        return fail(h, BS_ERR_ARG, "bs_step: negative step count");
    if (nsteps == 0)
        return BS_OK;
    int rc = ensure_ready(h);
    if (rc)
        return rc;
    rc = upload_ext(h, ext_force3);
    if (rc)
        return rc;
    h->tape.ready = false;
    // Trajectory frames are taken at the start of the steps whose iteration
    // count a writer's frequency divides (SpringNetwork::idleRun,
    // SpringNetwork.cpp:344-362): the state then, with the energies of the
    // step before, so that step computes its energies.
    const bool tr = h->traj && !h->traj->empty();
    const int32_t G = (h->p.graph_steps + 1) / 2 * 2;
    int64_t left = nsteps;
    for (int pass = 0;; ++pass)
    {
    launch_fill_i32(h->status + 6, 0, 1, nullptr, h->stream); // steps completed in this pass
    const bool graphs = G > 0 && !h->prof && left > G;
    int64_t s = 0;
    while (s < left)
    {
        const int64_t nb = h->steps_done + s;
        if (tr && h->traj->due(nb) && !(pass > 0 && s == 0))
        {
            rc = traj_capture(h, nb, s > 0, false);
            if (rc)
                return rc;
        }
        // replay a captured graph of G steps (from cur = 0 back to cur = 0)
        // when no frame falls inside it or right after it; the last step stays
        // outside, for the energies
        if (graphs && h->cur == 0 && left - 1 - s >= G && (!tr || h->traj->next_due(nb) > nb + G))
        {
            if (!h->graph_exec || h->graph_len != G || h->graph_ext != h->ext_uploaded)
            {
                rc = capture_graph(h, G);
                if (rc)
                    return rc;
            }
            HIPCHK(h, hipGraphLaunch(h->graph_exec, h->stream));
            s += G;
            continue;
        }
        const int frc = prof_fold(h);
        if (frc)
            return frc;
        launch_step(h, s + 1 == left || (tr && h->traj->due(nb + 1)));
        ++s;
    }
    int64_t done = 0;
    bool overflow = false;
    rc = finish_call(h, &done, &overflow);
    h->steps_done += done;
    h->last_call_forces_only = false;
    if (rc || !overflow)
        return rc;
    // A list overflowed at the start of step `done` of this pass, which
    // stopped there: lengthen the lists, rebuild at that state and go on.
    // Frames taken after the stop were not written and are taken again.
    if (pass >= 3)
        return fail(h, BS_ERR_STATE, "neighbour list capacity exceeded (max %d) after lengthening the lists",
                    h->list_longest);
    rc = regrow_lists(h);
    if (rc)
        return rc;
    left -= done;
    }
}

int bs_compute_forces(bs_handle *h, const float *ext_force3, float *force3, double *energies)
{
    if (!h)
        return BS_ERR_ARG;
    int rc = ensure_ready(h);
    if (rc)
        return rc;
    rc = upload_ext(h, ext_force3);
    if (rc)
        return rc;
    launch_fill_i32(h->status + 2, 0, 1, nullptr, h->stream);
    launch_fill_i32(h->status + 5, 0, 1, nullptr, h->stream);
    rebuild_chain(h, true);
    if (!h->lists.empty())
        launch_pairs(h, true);
    launch_springs(h, false, true);
    launch_probe_step(h, false);
    rc = finish_call(h);
    h->energies[BS_E_KINETIC] = 0.0;
    h->last_call_forces_only = true;
    if (rc)
        return rc;
    if (force3)
    {
        std::vector<float4> f((size_t)h->n);
        HIPCHK(h, hipMemcpy(f.data(), h->fout, sizeof(float4) * h->n, hipMemcpyDeviceToHost));
        for (int64_t t = 0; t < h->n; ++t)
        {
            const int64_t i = h->perm[(size_t)t];
            const bool st = (h->flags[(size_t)i] & BS_FLAG_STATIC) != 0;
            force3[3 * i] = st ? 0.0f : f[(size_t)t].x;
            force3[3 * i + 1] = st ? 0.0f : f[(size_t)t].y;
            force3[3 * i + 2] = st ? 0.0f : f[(size_t)t].z;
        }
    }
    if (energies)
        std::memcpy(energies, h->energies, sizeof(h->energies));
    return BS_OK;
}

int bs_get_state(bs_handle *h, float *x3, float *v3)
{
    if (!h)
        return BS_ERR_ARG;
    if (!h->ready)
    {
        if (x3)
            std::memcpy(x3, h->x.data(), sizeof(float) * 3 * h->n);
        if (v3)
            std::memcpy(v3, h->v.data(), sizeof(float) * 3 * h->n);
        return BS_OK;
    }
    std::vector<float4> pv((size_t)h->n), vv((size_t)h->n);
    HIPCHK(h, hipStreamSynchronize(h->stream));
    HIPCHK(h, hipMemcpy(pv.data(), h->pos[h->cur], sizeof(float4) * h->n, hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(vv.data(), h->vel, sizeof(float4) * h->n, hipMemcpyDeviceToHost));
    std::vector<float4> lv;
    if (h->p.compensated_positions)
    {
        lv.resize((size_t)h->n);
        HIPCHK(h, hipMemcpy(lv.data(), h->plo[h->cur], sizeof(float4) * h->n, hipMemcpyDeviceToHost));
    }
    for (int64_t t = 0; t < h->n; ++t)
    {
        const int64_t i = h->perm[(size_t)t];
        float4 p = pv[(size_t)t];
        if (!lv.empty())
        { // round hi + lo to the nearest float
            p.x = (float)((double)p.x + lv[(size_t)t].x);
            p.y = (float)((double)p.y + lv[(size_t)t].y);
            p.z = (float)((double)p.z + lv[(size_t)t].z);
        }
        if (x3)
        {
            x3[3 * i] = p.x;
            x3[3 * i + 1] = p.y;
            x3[3 * i + 2] = p.z;
        }
        if (v3)
        {
            v3[3 * i] = vv[(size_t)t].x;
            v3[3 * i + 1] = vv[(size_t)t].y;
            v3[3 * i + 2] = vv[(size_t)t].z;
        }
    }
    return BS_OK;
}

int bs_set_state(bs_handle *h, const float *x3, const float *v3)
{
    if (!h)
        return BS_ERR_ARG;
    if (!h->have_particles)
        return fail(h, BS_ERR_STATE, "bs_set_state: no particles");
    if (h->ready)
    {
        // pull the current device state so that a NULL argument keeps it
        int rc = bs_get_state(h, h->x.data(), h->v.data());
        if (rc)
            return rc;
    }
    if (x3)
        std::memcpy(h->x.data(), x3, sizeof(float) * 3 * h->n);
    if (v3)
        std::memcpy(h->v.data(), v3, sizeof(float) * 3 * h->n);
    h->tape.ready = false;
    if (!h->ready)
        return BS_OK;
    h->cur = 0;
    int rc = upload_positions(h);
    if (rc)
        return rc;
    // a new state: rebuild the lists, clear a stop and a non-finite report
    const int32_t st[8] = {0, INT_MAX, 0, 0, 0, 0, 0, 1};
    HIPCHK(h, hipMemcpy(h->status, st, sizeof st, hipMemcpyHostToDevice));
    h->nonfinite = -1;
    return BS_OK;
}

int bs_get_energies(bs_handle *h, double *energies)
{
    if (!h || !energies)
        return BS_ERR_ARG;
    std::memcpy(energies, h->energies, sizeof(h->energies));
    return BS_OK;
}

int64_t bs_last_nonfinite_particle(const bs_handle *h) { return h ? h->nonfinite : -1; }

int bs_get_counters(const bs_handle *h, int64_t *out6)
{
    if (!h || !out6)
        return BS_ERR_ARG;
    for (int k = 0; k < 6; ++k)
        out6[k] = 0;
    out6[0] = h->steps_done;
    out6[4] = h->nedges;
    if (h->ready)
    {
        int32_t st[4], ct[4];
        if (hipMemcpy(st, h->status, sizeof st, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(ct, h->counters, sizeof ct, hipMemcpyDeviceToHost) != hipSuccess)
            return BS_ERR_HIP;
        out6[1] = ct[1];
        out6[2] = st[3];
        int64_t cap = 0;
        for (const auto &l : h->lists)
            cap = std::max<int64_t>(cap, l.cap);
        out6[3] = cap;
    }
    return BS_OK;
}

int bs_profile(bs_handle *h, int32_t enable)
{
    if (!h)
        return BS_ERR_ARG;
    h->prof = enable != 0;
    h->spans.clear();
    h->evused = 0;
    for (int c = 0; c < BS_PROF_NUM; ++c)
    {
        h->prof_ms[c] = 0.0;
        h->prof_n[c] = 0;
    }
    return BS_OK;
}

int bs_get_profile(const bs_handle *h, double *ms, int64_t *launches)
{
    if (!h)
        return BS_ERR_ARG;
    for (int c = 0; c < BS_PROF_NUM; ++c)
    {
        if (ms)
            ms[c] = h->prof_ms[c];
        if (launches)
            launches[c] = h->prof_n[c];
    }
    return BS_OK;
}

int bs_get_cells(bs_handle *h, int32_t kind, int64_t *ncells_out, int64_t *nmembers_out, int64_t *cell_of_particle,
                 int32_t *cell_start, int32_t *sorted_idx, int64_t dims_out[3], double box_out[6])
{
    if (!h || kind < 0 || kind > 2)
        return BS_ERR_ARG;
    if (h->sys_off.size() > 2)
        return fail(h, BS_ERR_STATE, "bs_get_cells: the reference cell list is defined for a single system");
    int rc = ensure_ready(h);
    if (rc)
        return rc;
    const int32_t n = (int32_t)h->n;
    // box over every particle, FP64 (box.hpp:37-69)
    launch_box_reduce(h->pos[h->cur], nullptr, n, h->box_part, h->box, h->box_blocks, nullptr, h->stream);
    double bx[6];
    HIPCHK(h, hipMemcpyAsync(bx, h->box, sizeof bx, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (h->p.probe_enable)
    { // the probe is in the reference's container, so in its box, though never in a cell
      // (SpringNetwork.cpp:763-764,816-824; nsearch.hpp:342-366)
        ProbeDev pd;
        HIPCHK(h, hipMemcpy(&pd, h->probe_d, sizeof pd, hipMemcpyDeviceToHost));
        const double q[3] = {pd.x.x, pd.x.y, pd.x.z};
        for (int c = 0; c < 3; ++c)
        {
            bx[c] = std::min(bx[c], q[c]);
            bx[3 + c] = std::max(bx[3 + c], q[c]);
        }
        HIPCHK(h, hipMemcpy(h->box, bx, sizeof bx, hipMemcpyHostToDevice));
    }
    const float cutoff = kind == BS_NS_ELECTROSTATIC ? h->p.electrostatic_cutoff
                         : kind == BS_NS_HYDROPHOBIC ? h->p.hydrophobicity_cutoff
                                                     : h->p.steric_cutoff;
    const float r = cutoff + h->p.neighborskin; // NeighborSearch::_search_radius, float
    int64_t d[3];
    for (int c = 0; c < 3; ++c)
        d[c] = (int64_t)(size_t)(std::ceil((bx[3 + c] - bx[c]) / r) + 1); // nsearch.hpp:352-354
    const int64_t ncells = d[0] * d[1] * d[2];
    if (ncells > INT32_MAX - 1)
        return fail(h, BS_ERR_ARG, "bs_get_cells: %lld cells exceed the engine limit", (long long)ncells);
    rc = ensure_cells(h, (int32_t)ncells);
    if (rc)
        return rc;
    const uint32_t mask = kind == BS_NS_ELECTROSTATIC ? PF_CHARGED : kind == BS_NS_HYDROPHOBIC ? PF_HYDRO : 0u;
    hipStream_t s = h->stream;
    launch_cell_ids(h->pos[h->cur], nullptr, h->jprm, mask, n, h->box, nullptr, 0, nullptr, (int32_t)d[0],
                    (int32_t)d[1], (int32_t)d[2], (double)r, r, 1, h->cell_of, h->counts, nullptr, s);
    launch_scan(h->counts, h->cstart, (int32_t)ncells, nullptr, s);
    launch_scatter(h->cell_of, n, h->cstart, h->counts, 1, h->ctmp, nullptr, s);
    launch_cell_sort(h->cell_of, n, h->cstart, h->ctmp, h->perm_d, h->cidx, nullptr, s);
    std::vector<int32_t> cof((size_t)n), cst((size_t)ncells + 1), cid((size_t)n);
    HIPCHK(h, hipMemcpyAsync(cof.data(), h->cell_of, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
    HIPCHK(h, hipMemcpyAsync(cst.data(), h->cstart, sizeof(int32_t) * (ncells + 1), hipMemcpyDeviceToHost, s));
    HIPCHK(h, hipMemcpyAsync(cid.data(), h->cidx, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
    HIPCHK(h, hipStreamSynchronize(s));
    const int64_t nmem = cst[(size_t)ncells];
    if (ncells_out)
        *ncells_out = ncells;
    if (nmembers_out)
        *nmembers_out = nmem;
    if (dims_out)
        std::memcpy(dims_out, d, sizeof d);
    if (box_out)
        std::memcpy(box_out, bx, sizeof bx);
    if (cell_of_particle)
        for (int32_t t = 0; t < n; ++t)
            cell_of_particle[h->perm[(size_t)t]] = cof[(size_t)t];
    if (cell_start)
        std::memcpy(cell_start, cst.data(), sizeof(int32_t) * (ncells + 1));
    if (sorted_idx)
        for (int64_t t = 0; t < nmem; ++t)
            sorted_idx[t] = h->perm[(size_t)cid[(size_t)t]];
    return BS_OK;
}

int bs_get_pairs(bs_handle *h, int32_t kind, uint32_t *ij, int64_t cap, int64_t *count)
{
    if (!h || kind < 0 || kind > 2 || !count)
        return BS_ERR_ARG;
    int rc = ensure_ready(h);
    if (rc)
        return rc;
    *count = 0;
    const int term = kind_term(kind);
    const List *lst = nullptr;
    for (const auto &l : h->lists)
        if (l.terms & term)
            lst = &l;
    if (!lst)
        return BS_OK;
    rebuild_chain(h, true);
    const int32_t n = (int32_t)h->n;
    const uint32_t mask = kind == BS_NS_ELECTROSTATIC ? PF_CHARGED : kind == BS_NS_HYDROPHOBIC ? PF_HYDRO : 0u;
    const float rc_k = kind == BS_NS_ELECTROSTATIC ? h->p.electrostatic_cutoff
                       : kind == BS_NS_HYDROPHOBIC ? h->p.hydrophobicity_cutoff
                                                   : h->p.steric_cutoff;
    hipError_t e;
    int32_t *cnt = dalloc<int32_t>(1, e);
    HIPCHK(h, e);
    // pass 1 counts (writes nothing), pass 2 fills an exactly sized buffer
    int32_t c = 0;
    uint32_t *out = nullptr;
    hipError_t e1 = hipSuccess, e2 = hipSuccess;
    for (int pass = 0; pass < 2 && e1 == hipSuccess && e2 == hipSuccess; ++pass)
    {
        if (pass == 1)
        {
            if (!ij || c == 0)
                break;
            out = dalloc<uint32_t>(2 * (size_t)c, e1);
            if (e1 != hipSuccess)
                break;
        }
        launch_fill_i32(cnt, 0, 1, nullptr, h->stream);
        launch_pairs_extract(h->pos[h->cur], h->jprm, lst->cl, lst->cln, h->perm_d, n, h->ncl, lst->cap, rc_k,
                             rc2(rc_k, 1.0 - 4e-6), rc2(rc_k, 1.0 + 4e-6), mask, mask, kind != BS_NS_STERIC ? 1 : 0,
                             pass == 1 ? out : nullptr, cnt, pass == 1 ? c : 0, h->stream);
        int32_t got = 0;
        e1 = hipMemcpyAsync(&got, cnt, sizeof got, hipMemcpyDeviceToHost, h->stream);
        e2 = hipStreamSynchronize(h->stream);
        if (pass == 1 && got != c)
            e1 = hipErrorInvalidValue;
        c = got;
    }
    std::vector<uint32_t> v;
    if (e1 == hipSuccess && e2 == hipSuccess && !ij)
    { // count only
        (void)hipFree(cnt);
        *count = c;
        return BS_OK;
    }
    if (e1 == hipSuccess && e2 == hipSuccess && out)
    {
        v.resize(2 * (size_t)c);
        e1 = hipMemcpy(v.data(), out, sizeof(uint32_t) * v.size(), hipMemcpyDeviceToHost);
    }
    if (out)
        (void)hipFree(out);
    (void)hipFree(cnt);
    if (e1 != hipSuccess || e2 != hipSuccess)
        return fail(h, BS_ERR_HIP, "bs_get_pairs: extraction failed");
    const int64_t np = (int64_t)v.size() / 2;
    std::vector<uint64_t> keys((size_t)np);
    for (int64_t t = 0; t < np; ++t)
        keys[(size_t)t] = (uint64_t)v[2 * t] << 32 | v[2 * t + 1];
    std::sort(keys.begin(), keys.end());
    *count = np;
    if (ij)
        for (int64_t t = 0; t < np && t < cap; ++t)
        {
            ij[2 * t] = (uint32_t)(keys[(size_t)t] >> 32);
            ij[2 * t + 1] = (uint32_t)keys[(size_t)t];
        }
    return BS_OK;
}

int bs_add_rigid_body(bs_handle *h, int64_t n, const int64_t *ids)
{
    if (!h)
        return BS_ERR_ARG;
    if (!h->have_particles)
        return fail(h, BS_ERR_STATE, "bs_add_rigid_body: particles must be set first");
    if (n < 0 || (n > 0 && !ids))
        return fail(h, BS_ERR_ARG, "bs_add_rigid_body: bad id list");
    std::vector<int64_t> v(ids, ids + n);
    for (int64_t id : v)
        if (id < 0 || id >= h->n)
            return fail(h, BS_ERR_ARG, "bs_add_rigid_body: particle index %lld out of range", (long long)id);
    {
        std::vector<int64_t> u(v);
        std::sort(u.begin(), u.end());
        if (std::adjacent_find(u.begin(), u.end()) != u.end())
            return fail(h, BS_ERR_ARG, "bs_add_rigid_body: repeated particle index");
    }
    int rc = pull_and_release(h);
    if (rc)
        return rc;
    if (n == 0)
    { // RigidBodiesManager::InitRigidBodies with no ids clears the collection
        h->bodies.clear();
        h->body_of.assign((size_t)h->n, -1);
        return BS_OK;
    }
    for (int64_t id : v)
        if (h->body_of[(size_t)id] >= 0)
            return BS_OK; // already rigid: no new body (RigidBodiesManager.cpp:19-38)
    const int32_t b = (int32_t)h->bodies.size();
    h->bodies.emplace_back();
    h->bodies.back().ids = v;
    for (int64_t id : v)
        h->body_of[(size_t)id] = b;
    return BS_OK;
}

int bs_get_probe(bs_handle *h, float *x3, float *v3, float *f3)
{
    if (!h)
        return BS_ERR_ARG;
    if (!h->p.probe_enable)
        return fail(h, BS_ERR_STATE, "bs_get_probe: the probe particle is not enabled");
    int rc = ensure_ready(h);
    if (rc)
        return rc;
    ProbeDev pd;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    HIPCHK(h, hipMemcpy(&pd, h->probe_d, sizeof pd, hipMemcpyDeviceToHost));
    const float x[3] = {pd.x.x, pd.x.y, pd.x.z}, v[3] = {pd.v.x, pd.v.y, pd.v.z};
    const float f[3] = {(float)pd.f.x, (float)pd.f.y, (float)pd.f.z};
    for (int c = 0; c < 3; ++c)
    {
        if (x3)
            x3[c] = x[c];
        if (v3)
            v3[c] = v[c];
        if (f3)
            f3[c] = f[c];
    }
    return BS_OK;
}

int32_t bs_num_rigid_bodies(const bs_handle *h) { return h ? (int32_t)h->bodies.size() : 0; }

int bs_get_rigid_body(bs_handle *h, int32_t b, double *out16)
{
    if (!h || !out16)
        return BS_ERR_ARG;
    if (b < 0 || b >= (int32_t)h->bodies.size())
        return fail(h, BS_ERR_ARG, "bs_get_rigid_body: no body %d", b);
    int rc = ensure_ready(h);
    if (rc)
        return rc;
    RigidBody r;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    HIPCHK(h, hipMemcpy(&r, h->bodies_d + b, sizeof r, hipMemcpyDeviceToHost));
    for (int c = 0; c < 3; ++c)
    {
        out16[c] = r.X[c];
        out16[3 + c] = r.v[c];
        out16[6 + c] = r.L[c];
        out16[9 + c] = r.w[c];
    }
    for (int c = 0; c < 4; ++c)
        out16[12 + c] = r.q[c];
    return BS_OK;
}

int bs_set_grid(bs_handle *h, int32_t kind, const double *origin, const int64_t *shape, const double *cell,
                const float *scalar)
{
    if (!h)
        return BS_ERR_ARG;
    if (kind != BS_GRID_POTENTIAL && kind != BS_GRID_DENSITY)
        return fail(h, BS_ERR_ARG, "bs_set_grid: unknown grid kind %d", kind);
    if (scalar && (!origin || !shape || !cell))
        return fail(h, BS_ERR_ARG, "bs_set_grid: origin, shape and cell are required");
    if (scalar)
        for (int c = 0; c < 3; ++c)
            if (shape[c] <= 0 || shape[c] > (1 << 20) || !(cell[c] > 0.0))
                return fail(h, BS_ERR_ARG, "bs_set_grid: bad shape or cell size");
    if (scalar && shape[0] * shape[1] * shape[2] > (int64_t)1 << 31)
        return fail(h, BS_ERR_ARG, "bs_set_grid: grid too large");
    int rc = pull_and_release(h);
    if (rc)
        return rc;
    auto &G = h->grids[kind];
    G = bs_handle::GridHost{};
    if (!scalar)
        return BS_OK;
    const int64_t nx = shape[0], ny = shape[1], nz = shape[2];
    G.cells.resize((size_t)(nx * ny * nz));
    // PotentialGrid::compute_gradient (PotentialGrid.cpp:20-60): central
    // differences inside, zero on the faces, times -k_B * m->A * N->Da A fs^-2
    const double newton = 1.0E10 * (1.0 / 1.66054E-27) * 1.0E-15 * 1.0E-15;
    const double scale = 1.3806504E-23 * 1.0E10 * newton;
    auto at = [&](int64_t i, int64_t j, int64_t k) { return (double)scalar[(i * ny + j) * nz + k]; };
    for (int64_t i = 0; i < nx; ++i)
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t k = 0; k < nz; ++k)
            {
                double g[3] = {0.0, 0.0, 0.0};
                if (i > 0 && i < nx - 1)
                    g[0] = (at(i + 1, j, k) - at(i - 1, j, k)) / (2.0 * cell[0]);
                if (j > 0 && j < ny - 1)
                    g[1] = (at(i, j + 1, k) - at(i, j - 1, k)) / (2.0 * cell[1]);
                if (k > 0 && k < nz - 1)
                    g[2] = (at(i, j, k + 1) - at(i, j, k - 1)) / (2.0 * cell[2]);
                G.cells[(size_t)((i * ny + j) * nz + k)] =
                    make_float4((float)(-g[0] * scale), (float)(-g[1] * scale), (float)(-g[2] * scale),
                                (float)at(i, j, k));
            }
    for (int c = 0; c < 3; ++c)
    {
        G.origin[c] = origin[c];
        G.cell[c] = cell[c];
        G.shape[c] = shape[c];
    }
    G.set = true;
    return BS_OK;
}

int bs_set_burying(bs_handle *h, const float *burying)
{
    if (!h)
        return BS_ERR_ARG;
    if (!h->have_particles)
        return fail(h, BS_ERR_STATE, "bs_set_burying: particles must be set first");
    int rc = pull_and_release(h);
    if (rc)
        return rc;
    h->burying.assign((size_t)h->n, 1.0f);
    if (burying)
        std::memcpy(h->burying.data(), burying, sizeof(float) * h->n);
    return BS_OK;
}

int bs_get_insertion_vector(bs_handle *h, int64_t i1, int64_t i2, double *out6)
{
    if (!h || !out6)
        return BS_ERR_ARG;
    if (i1 < 0 || i2 < 0 || i1 >= h->n || i2 >= h->n)
        return fail(h, BS_ERR_ARG, "Insertion vector could not be initialized. One or both of the pair of "
                                   "particles defining the insertion vector have an invalid identifier.");
    int rc = ensure_ready(h);
    if (rc)
        return rc;
    launch_centroid(h->pos[h->cur], (int32_t)h->n, h->cent_part, h->cent_blocks, h->cent, h->stream);
    double cs[3];
    float4 q1, q2;
    HIPCHK(h, hipMemcpyAsync(cs, h->cent, sizeof cs, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(&q1, h->pos[h->cur] + h->iperm[(size_t)i1], sizeof q1, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(&q2, h->pos[h->cur] + h->iperm[(size_t)i2], sizeof q2, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    // InsertionVector::computeVector / computeAngle / computeRollAngle
    // (InsertionVector.cpp:18-78), in FP64; the projection uses the
    // term-by-term vector products the reference writes.
    const double p1[3] = {q1.x, q1.y, q1.z}, p2[3] = {q2.x, q2.y, q2.z};
    const double v[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
    auto nrm = [](const double *a) { return std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); };
    auto unit = [&](double *a) {
        const double l = nrm(a);
        for (int c = 0; c < 3; ++c)
            a[c] = l > 1e-40 ? a[c] / l : 0.0;
    };
    const double vn = nrm(v);
    out6[0] = std::acos(-v[2] / vn) * (180.0 / M_PI) - 90.0;
    double proj[3], n1[3], n3[3], ref[3], n2[3];
    for (int c = 0; c < 3; ++c)
        proj[c] = p1[c] + v[c] * (((cs[c] - p1[c]) * v[c]) / (vn * vn));
    for (int c = 0; c < 3; ++c)
        n1[c] = proj[c] - cs[c];
    unit(n1);
    n3[0] = n1[1] * v[2] - n1[2] * v[1];
    n3[1] = n1[2] * v[0] - n1[0] * v[2];
    n3[2] = n1[0] * v[1] - n1[1] * v[0];
    unit(n3);
    const double am = std::atan2(n3[2], n1[2]) - M_PI;
    const double d[3] = {cs[0] - proj[0], cs[1] - proj[1], cs[2] - proj[2]};
    const double r = nrm(d);
    for (int c = 0; c < 3; ++c)
        ref[c] = proj[c] + n1[c] * r * std::cos(am) + n3[c] * r * std::sin(am);
    for (int c = 0; c < 3; ++c)
        n2[c] = proj[c] - ref[c];
    unit(n2);
    const double dot12 = n1[0] * n2[0] + n1[1] * n2[1] + n1[2] * n2[2];
    const double arad = std::acos(dot12 / (nrm(n1) * nrm(n2)));
    const double c12[3] = {n1[1] * n2[2] - n1[2] * n2[1], n1[2] * n2[0] - n1[0] * n2[2],
                           n1[0] * n2[1] - n1[1] * n2[0]};
    const double sgn = c12[0] * (p2[0] - proj[0]) + c12[1] * (p2[1] - proj[1]) + c12[2] * (p2[2] - proj[2]);
    const double adeg = arad * 180.0 / M_PI;
    out6[1] = sgn > 0.0 ? 360.0 - adeg : adeg;
    out6[2] = cs[2];
    for (int c = 0; c < 3; ++c)
        out6[3 + c] = v[c];
    return BS_OK;
}

int bs_set_systems(bs_handle *h, int32_t nsys, const int64_t *offsets)
{
    if (!h)
        return BS_ERR_ARG;
    if (!h->have_particles)
        return fail(h, BS_ERR_STATE, "bs_set_systems: particles must be set first");
    if (nsys < 1 || !offsets || offsets[0] != 0 || offsets[nsys] != h->n)
        return fail(h, BS_ERR_ARG, "bs_set_systems: offsets must run from 0 to the particle count");
    for (int32_t s = 0; s < nsys; ++s)
        if (offsets[s + 1] <= offsets[s])
            return fail(h, BS_ERR_ARG, "bs_set_systems: empty or unordered system %d", s);
    std::vector<int32_t> sys((size_t)h->n);
    for (int32_t s = 0; s < nsys; ++s)
        for (int64_t i = offsets[s]; i < offsets[s + 1]; ++i)
            sys[(size_t)i] = s;
    const int hrc = springs_host(h);
    if (hrc)
        return hrc;
    for (size_t t = 0; t < h->s1.size(); ++t)
        if (sys[h->s1[t]] != sys[h->s2[t]])
            return fail(h, BS_ERR_ARG, "bs_set_systems: spring %zu joins two systems", t);
    for (const auto &B : h->bodies)
        for (int64_t id : B.ids)
            if (sys[(size_t)id] != sys[(size_t)B.ids[0]])
                return fail(h, BS_ERR_ARG, "bs_set_systems: a rigid body spans two systems");
    int rc = pull_and_release(h);
    if (rc)
        return rc;
    h->sys_off.assign(offsets, offsets + nsys + 1);
    return BS_OK;
}

int32_t bs_num_systems(const bs_handle *h) { return h ? (int32_t)h->sys_off.size() - 1 : 0; }

int bs_get_system_energies(bs_handle *h, double *out)
{
    if (!h || !out)
        return BS_ERR_ARG;
    const int32_t nsys = (int32_t)h->sys_off.size() - 1;
    if (!h->ready || !h->sys_E)
    {
        for (int32_t k = 0; k < 6 * nsys; ++k)
            out[k] = 0.0;
        return BS_OK;
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    HIPCHK(h, hipMemcpy(out, h->sys_E, sizeof(double) * 6 * nsys, hipMemcpyDeviceToHost));
    if (h->last_call_forces_only)
        for (int32_t s = 0; s < nsys; ++s)
            out[6 * s + BS_E_KINETIC] = 0.0;
    return BS_OK;
}

int bs_adjoint_forward(bs_handle *h, int32_t nsteps, int32_t checkpoint_every, const float *ext_force3)
{
    if (!h)
        return BS_ERR_ARG;
    if (nsteps < 0)
        return fail(h, BS_ERR_ARG, "bs_adjoint: negative step count");
    int rc = adjoint_check(h);
    if (rc)
        return rc;
    rc = ensure_ready(h);
    if (rc)
        return rc;
    rc = upload_ext(h, ext_force3);
    if (rc)
        return rc;
    return adjoint_record(h, nsteps, checkpoint_every);
}

// This handle's device's copy of the fixed scratch array.
static int noalloc_base(bs_handle *h, char **base)
{
    *base = nullptr;
    const int dev = h->device;
    if (dev < 0 || dev >= NOALLOC_MAXDEV)
        return fail(h, BS_ERR_STATE, "no fixed device scratch on device %d (limit %d devices)", dev, NOALLOC_MAXDEV);
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_done[dev])
        return fail(h, BS_ERR_STATE, "internal: device scratch not set up on device %d", dev);
    if (!g_noalloc_ptr[dev])
        return fail(h, BS_ERR_HIP, "hipGetSymbolAddress of the device scratch failed on device %d: %s", dev,
                    hipGetErrorString(g_noalloc_err[dev]));
    *base = g_noalloc_ptr[dev];
    return BS_OK;
}

// Checks shared by both forms of the backward call.
static int adjoint_backward_check(bs_handle *h, const double *dl_dx3, const double *dl_dv3, const char *who)
{
    if (!h->tape.ready)
        return fail(h, BS_ERR_STATE, "%s: no recorded forward pass at the current state", who);
    if (!dl_dx3 && !dl_dv3)
        return fail(h, BS_ERR_ARG, "%s: dL/dx or dL/dv required", who);
    const int rc = adjoint_check(h);
    if (rc)
        return rc;
    return check_device(h);
}

// The caller's seed (3n doubles each, particle order) uploaded into b.sx,
// b.sv, turned into lx, lv, w on the device, then the sweep.
static int adjoint_backward_on(bs_handle *h, AdjBufs &b, const double *dl_dx3, const double *dl_dv3, double *dk,
                        double *dl0, double *dq)
{
    const int64_t n = h->n;
    if (dl_dx3)
        HIPCHK(h, hipMemcpyAsync(b.sx, dl_dx3, sizeof(double) * 3 * n, hipMemcpyHostToDevice, h->stream));
    if (dl_dv3)
        HIPCHK(h, hipMemcpyAsync(b.sv, dl_dv3, sizeof(double) * 3 * n, hipMemcpyHostToDevice, h->stream));
    launch_adj_seed(dl_dx3 ? b.sx : nullptr, dl_dv3 ? b.sv : nullptr, h->perm_d, h->jprm, h->iprm, (int32_t)n,
                    (double)h->p.timestep, b.lx, b.lv, b.w[0], h->stream);
    return adjoint_sweep(h, b, dk, dl0, dq);
}

int bs_adjoint_backward(bs_handle *h, const double *dl_dx3, const double *dl_dv3, double *dk, double *dl0,
                        double *dq)
{
    if (!h)
        return BS_ERR_ARG;
    int rc = adjoint_backward_check(h, dl_dx3, dl_dv3, "bs_adjoint_backward");
    if (rc)
        return rc;
    const bool springs = (dk || dl0) && h->nsp > 0;
    if (springs)
    {
        rc = build_spring_owners(h);
        if (rc)
            return rc;
    }
    AdjBufs b;
    const size_t bytes = adj_carve(h, springs, true, nullptr, b);
    Scratch S;
    hipError_t e = hipSuccess;
    char *base = S.get<char>(bytes, e);
    HIPCHK(h, e);
    adj_carve(h, springs, true, base, b);
    return adjoint_backward_on(h, b, dl_dx3, dl_dv3, dk, dl0, dq);
}

int bs_adjoint_backward_noalloc(bs_handle *h, const double *dl_dx3, const double *dl_dv3, double *dk, double *dl0,
                                double *dq)
{
    if (!h)
        return BS_ERR_ARG;
    const char *who = "bs_adjoint_backward_noalloc";
    int rc = adjoint_backward_check(h, dl_dx3, dl_dv3, who);
    if (rc)
        return rc;
    if (h->prof)
        return fail(h, BS_ERR_STATE, "%s: profiling creates events; turn it off first", who);
    const bool springs = (dk || dl0) && h->nsp > 0;
    if (springs && h->nslots < 0)
        return fail(h, BS_ERR_STATE, "%s: spring gradient slots missing; run bs_adjoint_forward first", who);
    char *base = nullptr;
    rc = noalloc_base(h, &base);
    if (rc)
        return rc;
    AdjBufs b;
    const size_t bytes = adj_carve(h, springs, true, nullptr, b);
    if (bytes > NOALLOC_BYTES)
        return fail(h, BS_ERR_ARG,
                    "%s: this system needs %zu bytes of device scratch, more than the fixed %zu (%lld particles, "
                    "%lld springs); use bs_adjoint_backward",
                    who, bytes, NOALLOC_BYTES, (long long)h->n, (long long)h->nsp);
    adj_carve(h, springs, true, base, b);
    return adjoint_backward_on(h, b, dl_dx3, dl_dv3, dk, dl0, dq);
}

int bs_adjoint(bs_handle *h, int32_t nsteps, int32_t checkpoint_every, const float *target3, const float *ext_force3,
               double *loss, double *dk, double *dl0, double *dq)
{
    if (!h)
        return BS_ERR_ARG;
    if (!target3)
        return fail(h, BS_ERR_ARG, "bs_adjoint: target positions required");
    int rc = bs_adjoint_forward(h, nsteps, checkpoint_every, ext_force3);
    if (rc)
        return rc;
    // loss and seed lx = 2 (x_T - y) / n on the device
    const int64_t n = h->n;
    const bool springs = (dk || dl0) && h->nsp > 0;
    AdjBufs b;
    const size_t bytes = adj_carve(h, springs, false, nullptr, b);
    Scratch S;
    hipError_t e = hipSuccess;
    char *base = S.get<char>(bytes, e);
    HIPCHK(h, e);
    adj_carve(h, springs, false, base, b);
    float4 *tgt = S.get<float4>((size_t)n, e);
    HIPCHK(h, e);
    const int32_t lblocks = (int32_t)((n + BLOCK - 1) / BLOCK);
    double *lpart = S.get<double>((size_t)lblocks, e);
    HIPCHK(h, e);
    {
        std::vector<float4> y((size_t)n);
        for (int64_t t = 0; t < n; ++t)
        {
            const int64_t i = h->perm[(size_t)t];
            y[(size_t)t] = make_float4(target3[3 * i], target3[3 * i + 1], target3[3 * i + 2], 0.0f);
        }
        HIPCHK(h, hipMemcpyAsync(tgt, y.data(), sizeof(float4) * n, hipMemcpyHostToDevice, h->stream));
        launch_adj_init(h->pos[h->cur], tgt, h->jprm, h->iprm, (int32_t)n, (double)h->p.timestep, b.lx, b.lv,
                        b.w[0], lpart, lblocks, h->stream);
        std::vector<double> lp((size_t)lblocks);
        HIPCHK(h, hipMemcpyAsync(lp.data(), lpart, sizeof(double) * lblocks, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        double L = 0.0;
        for (double v : lp)
            L += v;
        if (loss)
            *loss = L / (double)n;
    }
    return adjoint_sweep(h, b, dk, dl0, dq);
}

static int springs_from_cutoff(bs_handle *h, double cutoff, float stiffness, const int32_t *group,
                               int64_t *count_out);

int bs_springs_from_cutoff(bs_handle *h, double cutoff, float stiffness, int64_t *count_out)
{
    return springs_from_cutoff(h, cutoff, stiffness, nullptr, count_out);
}

int bs_springs_between_groups_from_cutoff(bs_handle *h, double cutoff, float stiffness, const int32_t *group,
                                          int64_t *count_out)
{
    if (!h)
        return BS_ERR_ARG;
    if (!group)
        return fail(h, BS_ERR_ARG, "bs_springs_between_groups_from_cutoff: no groups");
    return springs_from_cutoff(h, cutoff, stiffness, group, count_out);
}

static int springs_from_cutoff(bs_handle *h, double cutoff, float stiffness, const int32_t *group,
                               int64_t *count_out)
{
    if (!h)
        return BS_ERR_ARG;
    if (!h->have_particles)
        return fail(h, BS_ERR_STATE, "bs_springs_from_cutoff: particles must be set first");
    const float rcf = (float)cutoff; // NeighborSearch keeps the cutoff as a float
    if (!(rcf >= 1e-6f) || !std::isfinite(rcf)) // nsearch.hpp:66-67 rejects cutoffs under 1e-6
        return fail(h, BS_ERR_ARG, "bs_springs_from_cutoff: the cutoff must be finite and at least 1e-6");
    const int64_t n = h->n;
    if (n > (int64_t)CL_JMASK)
        return fail(h, BS_ERR_ARG, "bs_springs_from_cutoff: more than %u particles", CL_JMASK);
    // the current positions (the device's, once the system has stepped)
    int rc = pull_and_release(h);
    if (rc)
        return rc;
    hipStream_t st = h->stream;
    hipError_t e = hipSuccess;
    Scratch S;
    std::vector<int32_t> rows((size_t)n + 1, 0);
    int64_t m = 0; // springs added
    // springs already loaded stay, with their own rest length and stiffness;
    // their pairs get no second spring (SpringCollection.hpp:138-151,226-233)
    const int64_t m0 = h->have_springs ? h->nsp : 0;
    if (m0 > 0)
    {
        rc = springs_host(h);
        if (rc)
            return rc;
    }
    if (n > 1)
    {
        std::vector<float4> ph((size_t)n);
        for (int64_t i = 0; i < n; ++i)
        {
            const float flag = __builtin_bit_cast(float, (uint32_t)((h->flags[(size_t)i] & BS_FLAG_STATIC) ? 1u : 0u));
            ph[(size_t)i] = make_float4(h->x[3 * i], h->x[3 * i + 1], h->x[3 * i + 2], flag);
        }
        const int32_t nb = (int32_t)std::min<int64_t>(BLOCK, (n + BLOCK - 1) / BLOCK);
        float4 *pos = S.get<float4>((size_t)n, e);
        HIPCHK(h, e);
        double *box_part = S.get<double>((size_t)(6 * nb), e);
        HIPCHK(h, e);
        double *box = S.get<double>(6, e);
        HIPCHK(h, e);
        HIPCHK(h, hipMemcpyAsync(pos, ph.data(), sizeof(float4) * n, hipMemcpyHostToDevice, st));
        launch_box_reduce(pos, nullptr, (int32_t)n, box_part, box, nb, nullptr, st);
        double bx[6];
        HIPCHK(h, hipMemcpyAsync(bx, box, sizeof bx, hipMemcpyDeviceToHost, st));
        HIPCHK(h, hipStreamSynchronize(st));
        // cells of edge a little above the cutoff (so that rounding in the
        // cell coordinate never puts a pair closer than the cutoff two cells
        // apart), at most 2^24 of them
        const float edge = rcf * (1.0f + 1e-5f);
        int32_t dims[3];
        int64_t ncells = 1;
        for (int q = 0; q < 3; ++q)
        {
            const double len = bx[3 + q] - bx[q];
            dims[q] = (int32_t)std::max(1.0, std::min(std::floor(len / (double)edge), 256.0));
            ncells *= dims[q];
        }
        const size_t e0 = h->prof ? prof_mark(h) : SIZE_MAX;
        int32_t *counts = S.get<int32_t>((size_t)ncells, e);
        HIPCHK(h, e);
        int32_t *cstart = S.get<int32_t>((size_t)ncells + 1, e);
        HIPCHK(h, e);
        int32_t *cursor = S.get<int32_t>((size_t)ncells, e);
        HIPCHK(h, e);
        int32_t *cell_of = S.get<int32_t>((size_t)n, e);
        HIPCHK(h, e);
        int32_t *tmp = S.get<int32_t>((size_t)n, e);
        HIPCHK(h, e);
        int32_t *cidx = S.get<int32_t>((size_t)n, e);
        HIPCHK(h, e);
        int32_t *rcnt = S.get<int32_t>((size_t)n, e);
        HIPCHK(h, e);
        launch_fill_i32(counts, 0, ncells, nullptr, st);
        launch_cell_ids(pos, nullptr, pos, 0u, (int32_t)n, box, nullptr, 0, nullptr, dims[0], dims[1], dims[2], 0.0, edge,
                        0, cell_of, counts, nullptr, st);
        launch_scan(counts, cstart, (int32_t)ncells, nullptr, st);
        launch_fill_i32(cursor, 0, ncells, nullptr, st);
        launch_scatter(cell_of, (int32_t)n, cstart, cursor, 0, tmp, nullptr, st);
        launch_cell_sort(cell_of, (int32_t)n, cstart, tmp, nullptr, cidx, nullptr, st);
        TopoArgs a{};
        a.pos = pos;
        a.box = box;
        a.cstart = cstart;
        a.cidx = cidx;
        a.counts = rcnt;
        a.n = (int32_t)n;
        for (int q = 0; q < 3; ++q)
            a.dims[q] = dims[q];
        a.edge = edge;
        a.rc = (double)rcf;
        if (group)
        {
            int32_t *gd = S.get<int32_t>((size_t)n, e);
            HIPCHK(h, e);
            HIPCHK(h, hipMemcpyAsync(gd, group, sizeof(int32_t) * n, hipMemcpyHostToDevice, st));
            a.group = gd;
        }
        // sqrt(r2) < rc  <=>  r2 < r2max, sqrt being monotonic and correctly
        // rounded (here and on the device)
        double r2max = a.rc * a.rc;
        while (r2max > 0.0 && std::sqrt(r2max) >= a.rc)
            r2max = std::nextafter(r2max, 0.0);
        while (std::sqrt(r2max) < a.rc)
            r2max = std::nextafter(r2max, INFINITY);
        a.r2max = r2max;
        if (m0 > 0)
        {
            // loaded springs as rows of their lower index, partners sorted
            std::vector<int32_t> eo((size_t)n + 1, 0), ej((size_t)m0);
            for (int64_t t = 0; t < m0; ++t)
                ++eo[(size_t)std::min(h->s1[(size_t)t], h->s2[(size_t)t]) + 1];
            for (int64_t i = 0; i < n; ++i)
                eo[(size_t)i + 1] += eo[(size_t)i];
            std::vector<int32_t> fill(eo.begin(), eo.end() - 1);
            for (int64_t t = 0; t < m0; ++t)
            {
                const uint32_t u = h->s1[(size_t)t], v = h->s2[(size_t)t];
                ej[(size_t)fill[std::min(u, v)]++] = (int32_t)std::max(u, v);
            }
            for (int64_t i = 0; i < n; ++i)
                std::sort(ej.begin() + eo[(size_t)i], ej.begin() + eo[(size_t)i + 1]);
            int32_t *eod = S.get<int32_t>((size_t)n + 1, e);
            HIPCHK(h, e);
            int32_t *ejd = S.get<int32_t>((size_t)m0, e);
            HIPCHK(h, e);
            HIPCHK(h, hipMemcpyAsync(eod, eo.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice, st));
            HIPCHK(h, hipMemcpyAsync(ejd, ej.data(), sizeof(int32_t) * m0, hipMemcpyHostToDevice, st));
            // the host arrays must outlive the copies
            HIPCHK(h, hipStreamSynchronize(st));
            a.ex_off = eod;
            a.ex_j = ejd;
        }
        launch_topo_springs(a, st);
        if (h->prof)
            prof_span(h, BS_PROF_TOPO_COUNT, e0, prof_mark(h));
        HIPCHK(h, hipMemcpyAsync(rows.data() + 1, rcnt, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st));
        HIPCHK(h, hipStreamSynchronize(st));
        int64_t total = 0;
        int32_t max_row = 0;
        for (int64_t i = 0; i < n; ++i)
        {
            max_row = std::max(max_row, rows[(size_t)i + 1]);
            total += rows[(size_t)i + 1];
            if (total > INT32_MAX)
                return fail(h, BS_ERR_ARG, "bs_springs_from_cutoff: more than 2^31 springs");
            rows[(size_t)i + 1] = (int32_t)total;
        }
        int32_t *offs = S.get<int32_t>((size_t)n + 1, e);
        HIPCHK(h, e);
        HIPCHK(h, hipMemcpyAsync(offs, rows.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice, st));
        // the new springs go straight into the device table, after the
        // loaded ones; the host copy is filled when something asks for it
        rc = springs_device_reserve(h, m0 + std::max<int64_t>(total, 1), m0);
        if (rc)
            return rc;
        a.offs = offs;
        a.max_row = max_row;
        a.jout = reinterpret_cast<int32_t *>(h->dsb + m0);
        a.l0out = h->dsl + m0;
        const size_t e2 = h->prof ? prof_mark(h) : SIZE_MAX;
        launch_topo_springs(a, st);
        if (h->prof)
            prof_span(h, BS_PROF_TOPO_FILL, e2, prof_mark(h));
        launch_rows_expand(offs, (int32_t)n, h->dsa + m0, st);
        launch_fill_i32(reinterpret_cast<int32_t *>(h->dsk + m0), __builtin_bit_cast(int32_t, stiffness), total,
                        nullptr, st);
        HIPCHK(h, hipStreamSynchronize(st));
        if (h->prof)
            prof_collect(h);
        m = total;
        h->nsp = m0 + m;
        h->sp_dev = true;
        h->sp_host = false;
        if (h->sys_off.size() > 2)
        {
            // independent systems (bs_set_systems): no spring joins two of
            // them; rows are sorted by j > i, so a row keeps its j below the
            // end of i's system
            rc = springs_host(h);
            if (rc)
                return rc;
            size_t w = (size_t)m0, sy = 0;
            for (int64_t i = 0; i < n; ++i)
            {
                while (h->sys_off[sy + 1] <= i)
                    ++sy;
                const uint32_t end = (uint32_t)h->sys_off[sy + 1];
                for (int32_t t = rows[(size_t)i]; t < rows[(size_t)i + 1]; ++t)
                    if (h->s2[(size_t)(m0 + t)] < end)
                    {
                        h->s1[w] = h->s1[(size_t)(m0 + t)];
                        h->s2[w] = h->s2[(size_t)(m0 + t)];
                        h->sl0[w] = h->sl0[(size_t)(m0 + t)];
                        h->sk[w] = h->sk[(size_t)(m0 + t)];
                        ++w;
                    }
            }
            m = (int64_t)w - m0;
            h->s1.resize(w);
            h->s2.resize(w);
            h->sl0.resize(w);
            h->sk.resize(w);
            springs_host_changed(h);
        }
    }
    free_device(h);
    h->have_springs = true;
    if (count_out)
        *count_out = m;
    return BS_OK;
}

int bs_get_springs(const bs_handle *h, uint32_t *id1, uint32_t *id2, float *equilibrium, float *stiffness)
{
    if (!h)
        return BS_ERR_ARG;
    // filling the host copy changes no observable state
    const int hrc = springs_host(const_cast<bs_handle *>(h));
    if (hrc)
        return hrc;
    const size_t m = h->s1.size();
    if (id1 && m)
        std::memcpy(id1, h->s1.data(), sizeof(uint32_t) * m);
    if (id2 && m)
        std::memcpy(id2, h->s2.data(), sizeof(uint32_t) * m);
    if (equilibrium && m)
        std::memcpy(equilibrium, h->sl0.data(), sizeof(float) * m);
    if (stiffness && m)
        std::memcpy(stiffness, h->sk.data(), sizeof(float) * m);
    return BS_OK;
}

int bs_autotune(bs_handle *h, int32_t steps, int32_t *group_out, int32_t *graph_out, int32_t *layout_out)
{
    if (!h)
        return BS_ERR_ARG;
    if (steps < 1)
        return fail(h, BS_ERR_ARG, "bs_autotune: steps must be positive");
    int rc = ensure_ready(h);
    if (rc)
        return rc;
    // trials run on a copy of the current system, so h is left as it is
    std::vector<float> x((size_t)(3 * h->n)), v((size_t)(3 * h->n));
    rc = bs_get_state(h, x.data(), v.data());
    if (rc)
        return rc;
    rc = springs_host(h); // the trials copy the host spring table
    if (rc)
        return rc;
    // first the pair layout (at the current lanes and graph setting), then
    // lanes per spring row x graph replay with that layout
    struct Trial
    {
        int32_t layout, group, graph;
    };
    std::vector<Trial> trials;
    for (int32_t lay : {BS_PAIR_CLUSTERS, BS_PAIR_QUEUES})
        trials.push_back({lay, h->group, h->p.graph_steps});
    for (int32_t gs : {0, 16})
        for (int32_t g : {4, 8, 16, 32})
            trials.push_back({-1, g, gs});
    double best = INFINITY;
    int32_t bg = h->group, bgr = h->p.graph_steps, blay = h->p.pair_layout == BS_PAIR_ROWS ? BS_PAIR_CLUSTERS
                                                                                        : h->p.pair_layout;
    for (size_t ti = 0; ti < trials.size(); ++ti)
    {
        if (ti == 2)
            best = INFINITY; // second phase: compare among itself
        const int32_t lay = trials[ti].layout >= 0 ? trials[ti].layout : blay;
        const int32_t g = trials[ti].group, gs = trials[ti].graph;
        {
            bs_handle *t = nullptr;
            bs_params p = h->p;
            p.graph_steps = gs;
            p.pair_layout = lay;
            if (bs_create(&p, &t) != BS_OK)
                return fail(h, BS_ERR_HIP, "bs_autotune: could not create a trial handle");
            t->n = h->n;
            t->x = x;
            t->v = v;
            t->mass = h->mass;
            t->charge = h->charge;
            t->radius = h->radius;
            t->eps = h->eps;
            t->sasa = h->sasa;
            t->transfer = h->transfer;
            t->hydro = h->hydro;
            t->flags = h->flags;
            t->s1 = h->s1;
            t->s2 = h->s2;
            t->sl0 = h->sl0;
            t->sk = h->sk;
            springs_host_changed(t);
            t->bodies = h->bodies;
            t->body_of = h->body_of;
            t->burying = h->burying;
            for (int k = 0; k < 2; ++k)
                t->grids[k] = h->grids[k];
            t->sys_off = h->sys_off;
            t->have_particles = t->have_springs = true;
            t->group = g;
            hipEvent_t e0 = nullptr, e1 = nullptr;
            int trc = do_setup(t);
            if (trc == BS_OK)
                trc = bs_step(t, 2, nullptr);
            if (trc == BS_OK && hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess)
            {
                (void)hipEventRecord(e0, t->stream);
                trc = bs_step(t, steps, nullptr);
                (void)hipEventRecord(e1, t->stream);
                float ms = 0.0f;
                if (trc == BS_OK && hipEventSynchronize(e1) == hipSuccess &&
                    hipEventElapsedTime(&ms, e0, e1) == hipSuccess && ms < best)
                {
                    best = ms;
                    if (ti < 2)
                        blay = lay;
                    else
                    {
                        bg = g;
                        bgr = gs;
                    }
                }
            }
            if (e0)
                (void)hipEventDestroy(e0);
            if (e1)
                (void)hipEventDestroy(e1);
            bs_destroy(t);
            if (trc != BS_OK)
                return fail(h, trc, "bs_autotune: trial failed");
        }
    }
    if (h->group != bg || h->p.pair_layout != blay)
        h->tape.ready = false; // recomputation would not repeat the recorded pass bit for bit
    h->group = bg;
    h->spring_blocks = (int32_t)((h->n * h->group + BLOCK - 1) / BLOCK);
    if (h->p.graph_steps != bgr || h->p.pair_layout != blay)
        drop_graph(h);
    h->p.graph_steps = bgr;
    if (h->p.pair_layout != BS_PAIR_ROWS && h->p.pair_layout != blay)
    { // row buffers are sized at setup; the other two read lists in their own order
        h->p.pair_layout = blay;
        launch_fill_i32(h->status + 7, 1, 1, nullptr, h->stream);
    }
    if (group_out)
        *group_out = bg;
    if (graph_out)
        *graph_out = bgr;
    if (layout_out)
        *layout_out = h->p.pair_layout;
    return BS_OK;
}

// ---- trajectory output ------------------------------------------------------

int bs_set_particle_labels(bs_handle *h, const char *const *name, const char *const *resname,
                           const char *const *chain, const int32_t *resid, const float *occupancy,
                           const float *tempfactor, const char *const *element, const int32_t *electron_charge)
{
    if (!h)
        return BS_ERR_ARG;
    if (!h->have_particles)
        return fail(h, BS_ERR_STATE, "bs_set_particle_labels: particles must be set first");
    const size_t n = (size_t)h->n;
    auto strs = [n](std::vector<std::string> &dst, const char *const *src) {
        dst.clear();
        if (src)
            for (size_t i = 0; i < n; ++i)
                dst.emplace_back(src[i] ? src[i] : "");
    };
    bs::TrajLabels L;
    strs(L.name, name);
    strs(L.resname, resname);
    strs(L.chain, chain);
    strs(L.element, element);
    if (resid)
        L.resid.assign(resid, resid + n);
    if (occupancy)
        L.occupancy.assign(occupancy, occupancy + n);
    if (tempfactor)
        L.tempfactor.assign(tempfactor, tempfactor + n);
    if (electron_charge)
        L.echarge.assign(electron_charge, electron_charge + n);
    h->labels = std::move(L);
    return BS_OK;
}

int bs_set_insertion_vector(bs_handle *h, int64_t i1, int64_t i2)
{
    if (!h)
        return BS_ERR_ARG;
    if (i1 < 0 && i2 < 0)
    {
        h->ivec1 = h->ivec2 = -1;
        return BS_OK;
    }
    if (i1 < 0 || i2 < 0 || i1 >= h->n || i2 >= h->n)
        return fail(h, BS_ERR_ARG, "Insertion vector could not be initialized. One or both of the pair of "
                                   "particles defining the insertion vector have an invalid identifier.");
    h->ivec1 = i1;
    h->ivec2 = i2;
    return BS_OK;
}

int bs_add_trajectory(bs_handle *h, int32_t kind, const char *path, int64_t frequency)
{
    if (!h)
        return BS_ERR_ARG;
    if (!h->have_particles)
        return fail(h, BS_ERR_STATE, "bs_add_trajectory: particles must be set first");
    if (kind != BS_TRAJ_PDB && kind != BS_TRAJ_XTC && kind != BS_TRAJ_CSV)
        return fail(h, BS_ERR_ARG, "bs_add_trajectory: unknown kind %d", kind);
    if (!path || !*path)
        return fail(h, BS_ERR_ARG, "bs_add_trajectory: empty path");
    if (frequency < 1)
        return fail(h, BS_ERR_ARG, "bs_add_trajectory: the frequency must be at least 1");
    if (kind == BS_TRAJ_XTC && h->n > INT32_MAX / 3)
        return fail(h, BS_ERR_ARG, "bs_add_trajectory: too many particles for XTC");
    if (!h->traj)
    {
        h->traj.reset(new bs::TrajectorySet(TRAJ_SLOTS));
        for (int k = 0; k < TRAJ_SLOTS; ++k)
        {
            float *x = nullptr;
            double *e = nullptr;
            hipEvent_t ev = nullptr;
            hipError_t r = hipHostMalloc((void **)&x, sizeof(float) * 3 * h->n, hipHostMallocDefault);
            if (r == hipSuccess)
            {
                h->traj_x.push_back(x);
                // the six energies, then the call's status words (8 int32)
                r = hipHostMalloc((void **)&e, sizeof(double) * (BS_NUM_ENERGIES + 4), hipHostMallocDefault);
            }
            if (r == hipSuccess)
            {
                h->traj_e.push_back(e);
                r = hipEventCreate(&ev);
            }
            if (r == hipSuccess)
                h->traj_ev.push_back(ev);
            if (r != hipSuccess)
            {
                traj_release(h);
                HIPCHK(h, r);
            }
        }
    }
    auto ctx = std::make_shared<bs::TrajContext>();
    ctx->n = h->n;
    ctx->labels = h->labels;
    ctx->echarge_default.resize((size_t)h->n);
    for (size_t i = 0; i < (size_t)h->n; ++i)
        ctx->echarge_default[i] = (int32_t)h->charge[i]; // ParticleProperty::setCharge
    const int hrc = springs_host(h);
    if (hrc)
        return hrc;
    ctx->conect.reserve(h->s1.size());
    for (size_t t = 0; t < h->s1.size(); ++t)
        ctx->conect.emplace_back(h->s1[t], h->s2[t]);
    ctx->spring = h->p.spring_enable != 0;
    ctx->steric = h->p.steric_enable != 0;
    ctx->elec = h->p.electrostatic_enable != 0;
    ctx->imp = h->p.imp_enable != 0;
    ctx->hydro = h->p.hydrophobicity_enable != 0;
    ctx->ivec1 = h->ivec1;
    ctx->ivec2 = h->ivec2;
    std::string err;
    auto w = bs::make_traj_writer(kind, path, frequency, ctx, err);
    if (!w)
        return fail(h, BS_ERR_ARG, "bs_add_trajectory: %s", err.c_str());
    h->traj->add(std::move(w));
    return BS_OK;
}

int bs_write_frame_now(bs_handle *h)
{
    if (!h)
        return BS_ERR_ARG;
    if (!h->traj || h->traj->empty())
        return fail(h, BS_ERR_STATE, "bs_write_frame_now: no trajectory writer");
    int rc = ensure_ready(h);
    if (rc)
        return rc;
    return traj_capture(h, h->steps_done, false, true);
}

int bs_flush_trajectories(bs_handle *h)
{
    if (!h)
        return BS_ERR_ARG;
    if (h->traj && !h->traj->drain())
        return fail(h, BS_ERR_ARG, "trajectory output: %s", h->traj->error().c_str());
    return BS_OK;
}

int bs_close_trajectories(bs_handle *h)
{
    if (!h)
        return BS_ERR_ARG;
    std::string err;
    if (h->traj && !h->traj->drain())
        err = h->traj->error();
    traj_release(h);
    if (!err.empty())
        return fail(h, BS_ERR_ARG, "trajectory output: %s", err.c_str());
    return BS_OK;
}

// ---- topology files -----------------------------------------------------------

int bs_load_netcdf(bs_handle *h, const char *path, char *warnings, int64_t warnings_len)
{
    if (!h)
        return BS_ERR_ARG;
    if (!path || !*path)
        return fail(h, BS_ERR_ARG, "bs_load_netcdf: empty path");
    bs::NcTopology t;
    std::string err;
    bool ok = false;
    try
    {
        ok = bs::read_netcdf_topology(path, t, err);
    }
    catch (const std::exception &x) // memory for a damaged or huge file
    {
        err = std::string("cannot read ") + path + ": " + x.what();
    }
    if (!ok)
        return fail(h, BS_ERR_ARG, "%s", err.c_str());
    std::vector<uint8_t> flags((size_t)t.n);
    for (size_t i = 0; i < flags.size(); ++i)
        flags[i] = t.is_static[i] ? BS_FLAG_STATIC : 0;
    int rc = bs_set_particles(h, t.n, t.x.data(), nullptr, t.mass.data(), t.charge.data(), t.radius.data(),
                              t.epsilon.data(), t.sasa.data(), t.transfer.data(), t.hydro.data(), flags.data());
    if (rc)
        return rc;
    // to_spring_network copies the names, residue and chain (Topology.hpp:250-262)
    h->labels.name = std::move(t.name);
    h->labels.resname = std::move(t.resname);
    h->labels.chain = std::move(t.chain);
    h->labels.resid = std::move(t.resid);
    rc = bs_set_springs(h, (int64_t)t.s1.size(), t.s1.data(), t.s2.data(), t.l0.data(), t.k.data());
    if (rc)
        return rc;
    if (warnings && warnings_len > 0)
    {
        std::string all;
        for (const auto &w : t.warnings)
            all += w + "\n";
        std::snprintf(warnings, (size_t)warnings_len, "%s", all.c_str());
    }
    return BS_OK;
}

int bs_save_netcdf(bs_handle *h, const char *path)
{
    if (!h)
        return BS_ERR_ARG;
    if (!path || !*path)
        return fail(h, BS_ERR_ARG, "bs_save_netcdf: empty path");
    if (!h->have_particles)
        return fail(h, BS_ERR_STATE, "bs_save_netcdf: no particles");
    bs::NcTopology t;
    t.n = h->n;
    t.x.resize((size_t)(3 * h->n));
    std::vector<float> v((size_t)(3 * h->n));
    int rc = bs_get_state(h, t.x.data(), v.data()); // the current positions
    if (rc)
        return rc;
    t.mass = h->mass;
    t.charge = h->charge;
    t.radius = h->radius;
    t.epsilon = h->eps;
    t.sasa = h->sasa;
    t.transfer = h->transfer;
    t.hydro = h->hydro;
    t.is_static.resize((size_t)h->n);
    for (size_t i = 0; i < t.is_static.size(); ++i)
        t.is_static[i] = (h->flags[i] & BS_FLAG_STATIC) ? 1 : 0;
    t.name = h->labels.name;
    t.resname = h->labels.resname;
    t.chain = h->labels.chain;
    t.resid = h->labels.resid;
    const int hrc = springs_host(h);
    if (hrc)
        return hrc;
    t.s1 = h->s1;
    t.s2 = h->s2;
    t.l0 = h->sl0;
    t.k = h->sk;
    std::string err;
    bool ok = false;
    try
    {
        ok = bs::write_netcdf_topology(path, t, err);
    }
    catch (const std::exception &x)
    {
        err = x.what();
    }
    if (!ok)
        return fail(h, BS_ERR_ARG, "bs_save_netcdf: %s", err.c_str());
    return BS_OK;
}

int bs_get_particle_properties(const bs_handle *h, float *mass, float *charge, float *radius, float *epsilon,
                               float *sasa, float *transfer, float *hydrophobicity, uint8_t *flags)
{
    if (!h)
        return BS_ERR_ARG;
    const size_t n = (size_t)h->n;
    auto out = [n](float *dst, const std::vector<float> &src) {
        if (dst && n)
            std::memcpy(dst, src.data(), sizeof(float) * n);
    };
    out(mass, h->mass);
    out(charge, h->charge);
    out(radius, h->radius);
    out(epsilon, h->eps);
    out(sasa, h->sasa);
    out(transfer, h->transfer);
    out(hydrophobicity, h->hydro);
    if (flags && n)
        std::memcpy(flags, h->flags.data(), n);
    return BS_OK;
}

// ---- interactive updates (the interactor bridge) --------------------------------

int bs_set_sasa(bs_handle *h, const float *sasa)
{
    if (!h || !sasa)
        return BS_ERR_ARG;
    if (!h->have_particles)
        return fail(h, BS_ERR_STATE, "bs_set_sasa: no particles");
    std::memcpy(h->sasa.data(), sasa, sizeof(float) * h->n);
    h->tape.ready = false;
    if (!h->ready)
        return BS_OK;
    // the per-particle IMPALA parameters in internal order; stream-ordered,
    // so the steps already queued keep the old surfaces
    std::vector<float4> ip((size_t)h->n);
    for (int64_t t = 0; t < h->n; ++t)
    {
        const int64_t i = h->perm[(size_t)t];
        ip[(size_t)t] = make_float4(h->mass[(size_t)i], h->sasa[(size_t)i], h->transfer[(size_t)i],
                                    h->burying.empty() ? 1.0f : h->burying[(size_t)i]);
    }
    HIPCHK(h, hipMemcpyAsync(h->iprm, ip.data(), sizeof(float4) * h->n, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return BS_OK;
}

int bs_set_membrane(bs_handle *h, float upper_offset, float lower_offset, float upper_curvature,
                    float lower_curvature)
{
    if (!h)
        return BS_ERR_ARG;
    if (!h->p.imp_enable)
        return BS_OK; // CustomData applies these only with IMPALA on
    float *dst[4] = {&h->p.imp_upper_offset, &h->p.imp_lower_offset, &h->p.imp_upper_curvature,
                     &h->p.imp_lower_curvature};
    const float val[4] = {upper_offset, lower_offset, upper_curvature, lower_curvature};
    bool changed = false;
    for (int k = 0; k < 4; ++k)
        if (!std::isnan(val[k]) && *dst[k] != val[k])
        {
            *dst[k] = val[k];
            changed = true;
        }
    if (changed)
    {
        drop_graph(h); // a captured graph holds the launch arguments
        h->tape.ready = false;
    }
    return BS_OK;
}

} // extern "C"
