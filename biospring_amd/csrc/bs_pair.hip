// bs_pair.hip -- non-bonded pair lists and pair forces on CDNA4 (gfx950).
//
// Particles are grouped into i-clusters of 8 consecutive internal indices
// (the internal order is spatial).  A neighbour list holds, per i-cluster,
// entries j | m << 24 where the 8-bit m marks which of the cluster's particles
// have j within the list radius and not spring-bonded.  Kernels run one wave64
// per i-cluster with lanes laid out as 8 i x 8 entries, so that every load of
// a j's data serves 8 lanes, and accumulate forces on i only (each pair is
// visited from both sides): no atomics, bitwise reproducible.
#include "bs_device.hpp"

#include <cmath>

namespace bs
{

namespace
{

constexpr int WAVES = BLOCK / 64;

__device__ inline uint32_t bonded_mask(const int32_t *ex_ptr, const uint32_t *ex_w, int32_t ci, int32_t j)
{
    int32_t lo = ex_ptr[ci];
    const int32_t end = ex_ptr[ci + 1];
    int32_t hi = end;
    while (lo < hi)
    {
        const int32_t mid = (lo + hi) >> 1;
        if ((int32_t)(ex_w[mid] >> 8) < j)
            lo = mid + 1;
        else
            hi = mid;
    }
    return (lo < end && (int32_t)(ex_w[lo] >> 8) == j) ? (ex_w[lo] & 0xFFu) : 0u;
}

// The cluster's spring exclusions as an open-addressing hash set in LDS
// (one table per wave; key word j << 8 | mask, never 0, so 0 marks a free
// slot).  A lookup costs about one LDS read where a binary search over the
// exclusion list in global memory costs a chain of about ten dependent loads.
constexpr int32_t EX_SLOTS = 2048;           // per wave, 8 KiB
constexpr int32_t EX_HASH_MAX = EX_SLOTS / 2; // load factor <= 1/2

__device__ inline uint32_t ex_hash(int32_t j) { return ((uint32_t)j * 0x9E3779B1u) >> 21; } // 11 bits

__device__ inline uint32_t bonded_mask_lds(const uint32_t *tab, int32_t j)
{
    uint32_t h = ex_hash(j);
    for (;;)
    {
        const uint32_t v = tab[h];
        if (v == 0u)
            return 0u;
        if ((int32_t)(v >> 8) == j)
            return v & 0xFFu;
        h = (h + 1u) & (EX_SLOTS - 1);
    }
}

// lane k's value, in a scalar register (k wave-uniform)
__device__ inline float lane_f(float x, int k) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), k)); }

// ---------------------------------------------------------------------------
// List build: one wave per i-cluster, one candidate j per lane.  Candidates
// come from the cells covering the cluster's bounding box grown by the list
// radius, so the search is complete for any cluster shape.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(BLOCK) k_cluster_list(ListArgs a)
{
    if (gated_off(a.gate))
        return;
    __shared__ uint32_t ex_tab[WAVES][EX_SLOTS];
    const int lane = threadIdx.x & 63;
    const int32_t ci = blockIdx.x * WAVES + (threadIdx.x >> 6);
    if (ci >= a.ncl)
        return; // whole wave
    // exclusions of this cluster into the wave's hash table (small sets);
    // larger sets are searched in global memory
    uint32_t *tab = ex_tab[threadIdx.x >> 6];
    bool ex_lds = false;
    if (a.exclude_bonded)
    {
        const int32_t eb = a.ex_ptr[ci], ee = a.ex_ptr[ci + 1];
        ex_lds = ee - eb <= EX_HASH_MAX;
        if (ex_lds)
        {
            for (int32_t t = lane; t < EX_SLOTS; t += 64)
                tab[t] = 0u;
            wave_lds_sync();
            for (int32_t t = eb + lane; t < ee; t += 64)
            {
                const uint32_t wv = a.ex_w[t];
                uint32_t h = ex_hash((int32_t)(wv >> 8));
                while (atomicCAS(&tab[h], 0u, wv) != 0u)
                    h = (h + 1u) & (EX_SLOTS - 1);
            }
            wave_lds_sync();
        }
    }
    const int32_t il = lane & 7;
    const int32_t iself = ci * CLUSTER + il;
    float4 pi = make_float4(0.f, 0.f, 0.f, 0.f);
    uint32_t fi = PF_STATIC, tag = 0u;
    if (iself < a.n)
    {
        pi = a.pos[iself];
        fi = fbits(a.jprm[iself].w);
        if (a.shift)
        { // independent systems: binning in shifted coordinates
            const float4 o = a.shift[iself];
            pi.x += o.x;
            pi.y += o.y;
            pi.z += o.z;
            tag = fbits(o.w);
        }
    }
    const bool vi = lane < CLUSTER && iself < a.n && !(fi & PF_STATIC) && (a.i_mask == 0u || (fi & a.i_mask));
    uint32_t todo = (uint32_t)(__ballot(vi) & 0xFFull);
    int32_t count = 0;
    // A cluster searches once per group of its particles that share a system
    // and lie within the list radius (per axis) of the group's first
    // particle.  Compact clusters are one group; a loose one (sparse, clumpy
    // systems, where 8 Morton-consecutive beads can be 30 A apart) would
    // otherwise scan every cell of its whole grown box.  Groups' entries are
    // appended in turn; a j seen by two groups appears twice with disjoint
    // masks.
    while (todo != 0u)
    {
        const int f = __ffs(todo) - 1;
        const uint32_t tag0 = (uint32_t)__builtin_amdgcn_readlane((int32_t)tag, f);
        const float fx = lane_f(pi.x, f), fy = lane_f(pi.y, f), fz = lane_f(pi.z, f);
        // the first particle always belongs to its group (also for a
        // non-finite position, so that the loop ends)
        const bool near = lane == f ||
                          (fabsf(pi.x - fx) <= a.radius && fabsf(pi.y - fy) <= a.radius && fabsf(pi.z - fz) <= a.radius);
        const uint32_t vmask =
            (uint32_t)(__ballot(lane < CLUSTER && ((todo >> lane) & 1u) && tag == tag0 && near) & 0xFFull);
        todo &= ~vmask;
        float xi[CLUSTER], yi[CLUSTER], zi[CLUSTER];
        float bmin[3] = {INFINITY, INFINITY, INFINITY}, bmax[3] = {-INFINITY, -INFINITY, -INFINITY};
#pragma unroll
        for (int k = 0; k < CLUSTER; ++k)
        {
            xi[k] = lane_f(pi.x, k); // wave-uniform: scalar registers
            yi[k] = lane_f(pi.y, k);
            zi[k] = lane_f(pi.z, k);
            if ((vmask >> k) & 1u)
            {
                bmin[0] = fminf(bmin[0], xi[k]);
                bmin[1] = fminf(bmin[1], yi[k]);
                bmin[2] = fminf(bmin[2], zi[k]);
                bmax[0] = fmaxf(bmax[0], xi[k]);
                bmax[1] = fmaxf(bmax[1], yi[k]);
                bmax[2] = fmaxf(bmax[2], zi[k]);
            }
        }
        int32_t clo[3], chi[3];
        for (int q = 0; q < 3; ++q)
        {
            const double mn = a.box[q], len = a.box[3 + q] - a.box[q];
            clo[q] = internal_coord(bmin[q] - a.radius, mn, len, a.dims[q], a.cell_edge);
            chi[q] = internal_coord(bmax[q] + a.radius, mn, len, a.dims[q], a.cell_edge);
        }
        // Rows of cells (fixed cy, cz) farther than the list radius from the
        // cluster's box are skipped and each row's x range is cut to the
        // radius left after the y and z gaps: about a third fewer candidates
        // than the whole grown box.  The first and last cell of an axis are
        // open-ended (internal_coord clamps), and the radius is padded beyond
        // the candidate test's own r2list, so no candidate is lost.
        double cmn[3], ce[3];
        for (int q = 0; q < 3; ++q)
        {
            cmn[q] = a.box[q];
            ce[q] = fmax((double)a.cell_edge, (a.box[3 + q] - a.box[q]) / (double)a.dims[q]);
        }
        const double rad = (double)a.radius * (1.0 + 1e-5) + 1e-6;
        auto gap = [&](int q, int32_t c) {
            const double lo = c == 0 ? -INFINITY : cmn[q] + c * ce[q];
            const double hi = c == a.dims[q] - 1 ? INFINITY : cmn[q] + (c + 1) * ce[q];
            return fmax(0.0, fmax(lo - (double)bmax[q], (double)bmin[q] - hi));
        };
        // The candidates form a stream, generated in row order (a row's
        // cells cx0..cx1 are contiguous in the cell array), cut into batches
        // of 64 that run across row ends, so that short rows do not leave
        // lanes idle.  A three-stage pipeline runs over the batches: while
        // batch b is tested, batch b+1's positions and batch b+2's particle
        // indices are in flight, so a batch never waits for its own loads.
        // Row setup is lane-parallel: lane l takes row rbase + l of the
        // (cy fastest, then cz) order, computes its x range and member run,
        // and a wave prefix sum places the runs in the stream, 64 rows per
        // setup; a batch never spans two setups.
        const int32_t ny = chi[1] - clo[1] + 1, nrows = ny * (chi[2] - clo[2] + 1);
        int32_t rbase = 0, my_rt = 0, my_pre = 0, my_end = 0; // lane's row: first cell entry, stream range
        int32_t total = 0, spos = 0, row = 0;                  // wave-uniform: the setup's stream length, position, row
        auto setup_rows = [&]() {
            const int32_t r = rbase + lane;
            int32_t re = 0;
            my_rt = 0;
            if (r < nrows)
            {
                const int32_t cy = clo[1] + r % ny, cz = clo[2] + r / ny;
                const double gzv = gap(2, cz), gyv = gap(1, cy);
                const double r2x = rad * rad - gzv * gzv - gyv * gyv;
                if (gzv <= rad && r2x >= 0.0)
                {
                    const double rx = sqrt(r2x);
                    const double mn0 = a.box[0], len0 = a.box[3] - a.box[0];
                    const int32_t cx0 =
                        internal_coord((float)((double)bmin[0] - rx), mn0, len0, a.dims[0], a.cell_edge);
                    const int32_t cx1 =
                        internal_coord((float)((double)bmax[0] + rx), mn0, len0, a.dims[0], a.cell_edge);
                    const int32_t row0 = cy * a.dims[0] + cz * a.dims[0] * a.dims[1];
                    my_rt = a.cell_start[row0 + cx0];
                    re = a.cell_start[row0 + cx1 + 1];
                }
            }
            const int32_t len = re > my_rt ? re - my_rt : 0;
            int32_t inc = len;
            for (int d = 1; d < 64; d <<= 1)
            {
                const int32_t o = __shfl_up(inc, d, 64);
                inc += lane >= d ? o : 0;
            }
            my_end = inc;
            my_pre = inc - len;
            total = __builtin_amdgcn_readlane(inc, 63);
            spos = 0;
            row = 0;
            rbase += 64;
        };
        // The next batch: the lane's cell-array entry, or -1 past the
        // batch's end; false once the stream is exhausted.
        auto next_batch = [&](int32_t &tl) -> bool {
            tl = -1;
            while (spos >= total)
            {
                if (rbase >= nrows)
                    return false;
                setup_rows();
            }
            const int32_t p = spos + lane, lim = min(spos + 64, total);
            for (;;) // the rows overlapping [spos, lim), in order
            {
                const int32_t pre = __builtin_amdgcn_readlane(my_pre, row);
                const int32_t end = __builtin_amdgcn_readlane(my_end, row);
                const int32_t rt0 = __builtin_amdgcn_readlane(my_rt, row);
                if (p >= pre && p < end)
                    tl = rt0 + (p - pre);
                if (end >= lim)
                {
                    if (end == lim && row < 63)
                        ++row; // the row ends with the batch
                    break;
                }
                ++row;
            }
            spos = lim;
            return true;
        };
        // Index loads are clamped into the cell array (n entries); a lane
        // past its batch's end gathers particle 0 and tests nothing.
        int32_t t0, t1, t2;
        bool h0 = next_batch(t0);
        bool h1 = next_batch(t1);
        const int32_t last = a.n > 0 ? a.n - 1 : 0;
        int32_t j0 = a.cell_idx[min(max(t0, 0), last)];
        int32_t j1 = a.cell_idx[min(max(t1, 0), last)];
        j0 = t0 >= 0 ? j0 : 0;
        float4 p0 = a.pos[j0];
        float4 o0 = a.shift ? a.shift[j0] : make_float4(0.f, 0.f, 0.f, 0.f);
        uint32_t f0 = fbits(a.jprm[j0].w);
        while (h0) // wave-uniform
        {
            const bool h2 = next_batch(t2);
            // batch b+1's positions and b+2's indices in flight
            j1 = t1 >= 0 ? j1 : 0;
            const float4 p1 = a.pos[j1];
            const float4 o1 = a.shift ? a.shift[j1] : make_float4(0.f, 0.f, 0.f, 0.f);
            const uint32_t f1 = fbits(a.jprm[j1].w);
            const int32_t j2 = a.cell_idx[min(max(t2, 0), last)];
            // batch b
            uint32_t m = 0u;
            const int32_t j = j0;
            if (t0 >= 0)
            {
                float4 pj = p0;
                bool same_sys = true;
                if (a.shift)
                { // independent systems never pair, however far they move from their setup places
                    pj.x += o0.x;
                    pj.y += o0.y;
                    pj.z += o0.z;
                    same_sys = fbits(o0.w) == tag0;
                }
                if (same_sys && (a.j_mask == 0u || (f0 & a.j_mask)))
                {
#pragma unroll
                    for (int k = 0; k < CLUSTER; ++k)
                    {
                        const float dx = pj.x - xi[k], dy = pj.y - yi[k], dz = pj.z - zi[k];
                        const float r2 = dx * dx + dy * dy + dz * dz;
                        if (r2 < a.r2list && j != ci * CLUSTER + k)
                            m |= 1u << k;
                    }
                    m &= vmask;
                    if (m != 0u && a.exclude_bonded)
                        m &= ~(ex_lds ? bonded_mask_lds(tab, j) : bonded_mask(a.ex_ptr, a.ex_w, ci, j));
                }
            }
            const uint64_t bal = __ballot(m != 0u);
            if (bal != 0ull)
            {
                const int32_t off = __popcll(bal & ((1ull << lane) - 1ull));
                if (m != 0u)
                {
                    const int32_t slot = count + off;
                    if (slot < a.cap)
                        a.cl[(size_t)ci * a.cap + slot] = (uint32_t)j | (m << 24);
                }
                count += __popcll(bal);
            }
            t0 = t1;
            h0 = h1;
            t1 = t2;
            h1 = h2;
            j0 = j1;
            j1 = j2;
            p0 = p1;
            o0 = o1;
            f0 = f1;
        }
    }
    // zero words up to the next multiple of 16, read unconditionally by the
    // pair kernel (the capacity is a multiple of 16)
    if (count < a.cap)
    {
        const int32_t pad = ((count + 15) & ~15) < a.cap ? ((count + 15) & ~15) : a.cap;
        for (int32_t slot = count + lane; slot < pad; slot += 64)
            a.cl[(size_t)ci * a.cap + slot] = 0u;
    }
    if (lane == 0)
    {
        a.cln[ci] = count < a.cap ? count : a.cap;
        if (count > a.cap)
            atomicOr(&a.status[2], 1);
        if (count > 0)
            atomicMax(&a.status[3], count);
    }
}

// ---------------------------------------------------------------------------
// One pair from i's side.  Energies are weighted 1/2 for a dynamic partner
// and 1 for a static one (Particle.cpp:231-249).  Formulas: energy/steric.hpp,
// electrostatic.hpp, hydrophobic.hpp; per-bead sqrt(r), sqrt(eps) for the
// Lorentz-Berthelot / Good-Hope modes (CombinationRules.hpp:18-31).  The
// force is accumulated in FP64 (same rate as FP32 on CDNA4), which keeps
// small net forces accurate when many pair terms cancel.
//
// Branch-free: every lane evaluates the terms with a safe distance and the
// results are selected by the pair's predicates.  Across a wave the pairs are
// half inside the cutoff at best, so branching would not skip work; it only
// costs exec-mask bookkeeping and keeps loads from being issued early.  The
// exact two-stage test runs for pairs in the guard band around rc^2 only.
//
// The kernel is bound by VALU issue, so the per-pair instruction count is
// what is minimised: the force prefactors carry C_spring, the term scales
// and i's own factor (PairI, formed once per lane); the pair predicates gate
// FP32 values (one select each) instead of FP64 results; a charge that is
// zero, or a pair gated off, zeroes the Coulomb term through q_j.
// ---------------------------------------------------------------------------
struct PairI
{
    float ks;  // steric force: C_spring x scale x {12 eps_i', 24 eps_i', eps_i} (Amber, Lewitt, Zacharias), C_spring x scale (linear)
    float kh;  // hydrophobic force: scale x C_spring x h_i
    double qe; // Coulomb energy: q_i ke
    double qf; // Coulomb force: q_i kf, doubled for eps(r) = eps r
};

// An FP32 value kept as computed: a select of it stays one FP32 select
// instead of being sunk past a later FP64 conversion (two selects).
__device__ inline float keep(float x)
{
    asm("" : "+v"(x));
    return x;
}

template <int SMODE, int EMODE>
__device__ inline PairI pair_i(float4 pi, float4 ri, const PairArgs &a)
{
    PairI c;
    const float cs = (float)a.k.c_spring;
    const float m = SMODE == 1 ? 12.0f : SMODE == 2 ? 24.0f : 1.0f;
    c.ks = SMODE == 0 ? cs * a.steric_scale : m * cs * a.steric_scale * ri.y;
    c.kh = a.hydro_scale * cs * ri.z;
    c.qe = (double)pi.w * a.ke;
    c.qf = (double)pi.w * a.kf * (EMODE == 2 ? 2.0 : 1.0);
    return c;
}

template <int SMODE, int EMODE, int HYDRO, bool ENERGY>
__device__ inline void interact(bool masked, float4 pi, float4 ri, float4 pj, float4 rj, const PairArgs &a,
                                const PairI &c, double &fx, double &fy, double &fz, double &es, double &ee,
                                double &eh)
{
    const float dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
    const float r2 = dx * dx + dy * dy + dz * dz;
    const bool lo = r2 < a.rc2lo;
    bool on = masked && lo;
    if (masked && !lo && r2 < a.rc2hi) // rare: decide exactly
        on = exact_two_stage(pi, pj, a.rc);
    const bool pos2 = r2 > 0.0f;
    const float r2s = pos2 ? r2 : 1.0f;
    // the distance itself is needed by the vacuum Coulomb term, the linear
    // steric wall and the hydrophobic term; the others need 1/r^2 only
    constexpr bool NEED_D = EMODE == 1 || SMODE == 0 || HYDRO != 0;
    float inv = 0.0f, inv2;
    double invd = 0.0, i2d = 0.0;
    if (EMODE > 0)
    {
        // 1/d or 1/r^2 to ~1e-13: the FP32 seed (its input kept at or above
        // the smallest normal float, so it is finite) and one FP64 Newton
        // step on the exact r^2 of the float components.  The FP32 values of
        // the other terms are its rounding.  r2 >= 0, so the larger bit
        // pattern is the larger value (one integer max).
        const double xd = dx, yd = dy, zd = dz;
        const double r2d = fma(xd, xd, fma(yd, yd, zd * zd));
        const float seed = __int_as_float(max(__float_as_int(r2), 0x00800000));
        if (NEED_D)
        {
            const double y = (double)__builtin_amdgcn_rsqf(seed);
            invd = y * fma(-0.5 * r2d, y * y, 1.5);
            i2d = invd * invd;
            inv = (float)invd;
            inv2 = inv * inv;
        }
        else
        {
            const double y = (double)__builtin_amdgcn_rcpf(seed);
            i2d = y * fma(-r2d, y, 2.0);
            inv2 = (float)i2d;
        }
    }
    else
    {
        inv = __builtin_amdgcn_rsqf(r2s);
        inv = inv * (1.5f - 0.5f * r2s * inv * inv);
        inv2 = inv * inv;
    }
    const uint32_t flj = fbits(rj.w);
    const float wgt = (flj & PF_STATIC) ? 1.0f : 0.5f;
    float fs = 0.0f; // FP32 terms of the force scalar: force on i = (dx, dy, dz) * (fs + fsd)
    if (SMODE >= 0)
    {
        float en = 0.0f, f = 0.0f;
        if (SMODE == 0)
        { // linear: ri + rj from raw radii; a coincident pair has energy, no force
            const float sr = ri.x + rj.x;
            const float d = r2s * inv;
            const float dv = pos2 ? d - sr : -sr;
            const bool ov = !(dv > 0.0f);
            en = ov ? 0.5f * dv * dv : 0.0f;
            f = (on && ov && pos2) ? dv * c.ks * inv : 0.0f; // -|dv| * C / d
        }
        else
        {
            const float R = ri.x * rj.x;
            const float s2 = R * R * inv2, s6 = s2 * s2 * s2;
            float g, e0 = 0.0f;
            if (SMODE == 1)
            { // Amber 12-6: F = 12 eps (s6 - s12) / r^2, E = eps (s12 - 2 s6)
                g = fmaf(-s6, s6, s6);
                if (ENERGY)
                    e0 = ri.y * rj.y * (s6 * s6 - 2.0f * s6);
            }
            else if (SMODE == 2)
            { // Lewitt 8-6: F = 24 eps (s6 - s8) / r^2, E = eps (3 s8 - 4 s6)
                const float s8 = s6 * s2;
                g = s6 - s8;
                if (ENERGY)
                    e0 = ri.y * rj.y * (3.0f * s8 - 4.0f * s6);
            }
            else
            { // Zacharias 8-6: raw r, eps, product rules
                const float s8 = s6 * s2;
                g = 6.0f * s6 - 8.0f * s8;
                if (ENERGY)
                    e0 = ri.y * rj.y * (s8 - s6);
            }
            const bool far = r2 >= 0.01f; // 0.1 A floor (steric.hpp:62)
            en = far ? e0 : 0.0f;
            f = (on && far) ? c.ks * rj.y * g * inv2 : 0.0f;
        }
        if (ENERGY)
            es += on ? (double)(wgt * a.steric_scale * en) : 0.0;
        fs += f;
    }
    double fsd = 0.0;
    if (EMODE > 0)
    {
        // Coulomb dominates the non-bonded force: its scalar is formed in FP64
        const double qj = (double)keep((on && r2 >= 1.0e-6f) ? pj.w : 0.0f);
        if (EMODE == 1)
        {
            if (ENERGY)
                ee += (double)wgt * c.qe * qj * invd;
            fsd = -(c.qf * qj) * (i2d * invd);
        }
        else
        { // eps(r) = eps * r
            if (ENERGY)
                ee += (double)wgt * c.qe * qj * i2d;
            fsd = -(c.qf * qj) * (i2d * i2d);
        }
    }
    if (HYDRO)
    {
        // h_i h_j vanishes unless both are hydrophobic
        const float e = __expf(-(r2s * inv));
        if (ENERGY)
        {
            const float hh = pos2 ? ri.z * rj.z * e : 0.0f;
            eh += on ? (double)(wgt * a.hydro_scale) * (-(double)hh * 6.02214169E23 * 1.0E-3) : 0.0;
        }
        fs += (on && pos2) ? c.kh * rj.z * e * inv : 0.0f;
    }
    const double s = EMODE > 0 ? fsd + (double)keep(fs) : (double)fs;
    fx = fma((double)dx, s, fx);
    fy = fma((double)dy, s, fy);
    fz = fma((double)dz, s, fz);
}

// Pair forces of one i-cluster per wave.  A block is 16 list words, two per
// lane.  Blocks alternate between two register sets: while block t is
// evaluated, the j data of block t+1 and the words of block t+2 are in
// flight, and no register moves rotate the pipeline (zero-padded lists: a
// zero word is a masked-out entry pointing at particle 0).
template <int SMODE, int EMODE, int HYDRO, bool ENERGY>
__global__ void __launch_bounds__(BLOCK) k_pair(PairArgs a)
{
    if (a.status && halted(a.status))
        return;
    const int lane = threadIdx.x & 63;
    // wave-uniform, made visible to the compiler so that the list walk runs
    // on scalar registers
    const int32_t ci = __builtin_amdgcn_readfirstlane(blockIdx.x * WAVES + (threadIdx.x >> 6));
    const int32_t il = lane & 7, jl = lane >> 3;
    const int32_t i = ci * CLUSTER + il;
    double es = 0.0, ee = 0.0, eh = 0.0;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    if (ci < a.ncl)
    {
        const bool iv = i < a.n;
        const float4 pi = iv ? a.pos[i] : make_float4(0.f, 0.f, 0.f, 0.f);
        const float4 ri = iv ? a.jprm[i] : make_float4(0.f, 0.f, 0.f, 0.f);
        const PairI c = pair_i<SMODE, EMODE>(pi, ri, a);
        const uint32_t ibit = 1u << (24 + il);
        const int32_t cnt16 = (a.cln[ci] + 15) & ~15;
        const uint32_t *L = a.cl + (size_t)ci * a.cap;
        // Look-ahead reads past the end are clamped to the list's last 16
        // words (valid entries, never computed with), so every load is
        // unconditional and can be issued ahead of its use.
        const int32_t last = cnt16 > 16 ? cnt16 - 16 : 0;
        uint32_t wa0 = 0u, wa1 = 0u, wb0 = 0u, wb1 = 0u; // words of the even (a) and odd (b) blocks
        if (cnt16 > 0)
        {
            const int32_t o = min(16, last);
            wa0 = L[jl];
            wa1 = L[8 + jl];
            wb0 = L[o + jl];
            wb1 = L[o + 8 + jl];
        }
        float4 pa0 = a.pos[wa0 & CL_JMASK], ra0 = a.jprm[wa0 & CL_JMASK];
        float4 pa1 = a.pos[wa1 & CL_JMASK], ra1 = a.jprm[wa1 & CL_JMASK];
        for (int32_t t = 0; t < cnt16; t += 32)
        {
            // block t (set a); in flight: words of block t+2, j data of block t+1
            const bool ma0 = (wa0 & ibit) != 0u, ma1 = (wa1 & ibit) != 0u;
            {
                const uint32_t *Lw = L + min(t + 32, last); // wave-uniform base, lane offset jl
                wa0 = Lw[jl];
                wa1 = Lw[8 + jl];
            }
            const float4 pb0 = a.pos[wb0 & CL_JMASK], rb0 = a.jprm[wb0 & CL_JMASK];
            const float4 pb1 = a.pos[wb1 & CL_JMASK], rb1 = a.jprm[wb1 & CL_JMASK];
            interact<SMODE, EMODE, HYDRO, ENERGY>(ma0, pi, ri, pa0, ra0, a, c, fx, fy, fz, es, ee, eh);
            interact<SMODE, EMODE, HYDRO, ENERGY>(ma1, pi, ri, pa1, ra1, a, c, fx, fy, fz, es, ee, eh);
            if (t + 16 >= cnt16)
                break; // wave-uniform
            // block t+1 (set b)
            const bool mb0 = (wb0 & ibit) != 0u, mb1 = (wb1 & ibit) != 0u;
            {
                const uint32_t *Lw = L + min(t + 48, last);
                wb0 = Lw[jl];
                wb1 = Lw[8 + jl];
            }
            pa0 = a.pos[wa0 & CL_JMASK];
            ra0 = a.jprm[wa0 & CL_JMASK];
            pa1 = a.pos[wa1 & CL_JMASK];
            ra1 = a.jprm[wa1 & CL_JMASK];
            interact<SMODE, EMODE, HYDRO, ENERGY>(mb0, pi, ri, pb0, rb0, a, c, fx, fy, fz, es, ee, eh);
            interact<SMODE, EMODE, HYDRO, ENERGY>(mb1, pi, ri, pb1, rb1, a, c, fx, fy, fz, es, ee, eh);
        }
    }
    // sum over the 8 entry lanes of each i (lanes il, il+8, ..., il+56)
    for (int off = 8; off < 64; off <<= 1)
    {
        fx += __shfl_xor(fx, off, 64);
        fy += __shfl_xor(fy, off, 64);
        fz += __shfl_xor(fz, off, 64);
        if (ENERGY)
        {
            es += __shfl_xor(es, off, 64);
            ee += __shfl_xor(ee, off, 64);
            eh += __shfl_xor(eh, off, 64);
        }
    }
    if (ENERGY && ci < a.ncl && jl == 0 && i < a.n)
    {
        double4 e = make_double4(es, ee, eh, 0.0);
        if (a.accumulate)
        {
            const double4 o = a.epp[i];
            e.x += o.x;
            e.y += o.y;
            e.z += o.z;
        }
        a.epp[i] = e;
    }
    if (ci < a.ncl && jl == 0 && i < a.n)
    {
        double4 f = make_double4(fx, fy, fz, 0.0);
        if (a.accumulate)
        {
            const double4 o = a.fnb[i];
            f.x += o.x;
            f.y += o.y;
            f.z += o.z;
        }
        a.fnb[i] = f;
    }
}

// ---------------------------------------------------------------------------
// Adjoint of the pair terms (no reference counterpart).  For F_i = phi(r) d,
// d = x_j - x_i, the Jacobian applied to w is
//   H w |_i = phi dw + (phi'/r) (d . dw) d,  dw = w_j - w_i,
// with phi = fm / r and phi' = (fm' - phi) / r from the force module fm of
// each term (energy/steric.hpp, electrostatic.hpp) and its r-derivative fm'.
// dL/dq_i gathers q_j (kf / r^3) (d . dw) (constant eps; 2 kf / r^4 for
// eps(r) = eps r); a static partner's share goes to it by an FP64 atomic.
// Lanes, lists, the exact cutoff test and the load pipeline are those of
// k_pair.  The terms are formed in FP32 (the weights w are FP32; a gradient
// needs about 1e-5, not the 1e-5-after-cancellation of a force) and summed
// in FP64.
// ---------------------------------------------------------------------------
template <int SMODE, int EMODE>
__device__ inline void adj_pair_term(bool masked, float4 pi, float4 ri, uint32_t fli, float4 wi, float4 pj,
                                     float4 rj, float4 wj, int32_t j, const AdjPairArgs &a, float cs, float kf,
                                     double &hx, double &hy, double &hz, double &dqi)
{
    const float dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
    const float r2f = dx * dx + dy * dy + dz * dz;
    bool on = masked && r2f < a.rc2lo;
    if (masked && r2f >= a.rc2lo && r2f < a.rc2hi)
        on = exact_two_stage(pi, pj, a.rc);
    const uint32_t flj = fbits(rj.w);
    on = on && r2f > 0.0f;
    const float r2s = on ? r2f : 1.0f;
    float ir = __builtin_amdgcn_rsqf(r2s);
    ir = ir * (1.5f - 0.5f * r2s * ir * ir);
    const float r = r2s * ir, ir2 = ir * ir;
    float fm = 0.0f, dfm = 0.0f;
    if (SMODE == 0)
    {
        const float sr = ri.x + rj.x;
        const bool ov = r <= sr;
        fm = ov ? (r - sr) * cs : 0.0f;
        dfm = ov ? cs : 0.0f;
    }
    else if (SMODE > 0)
    {
        const float eij = ri.y * rj.y, R = ri.x * rj.x;
        const float s2 = R * R * ir2, s6 = s2 * s2 * s2, s8 = s6 * s2;
        float f0, d0;
        if (SMODE == 1)
        {
            const float s12 = s6 * s6;
            f0 = 12.0f * eij * (s6 - s12) * ir;
            d0 = 12.0f * eij * (-7.0f * s6 + 13.0f * s12) * ir2;
        }
        else if (SMODE == 2)
        {
            f0 = 24.0f * eij * (s6 - s8) * ir;
            d0 = 24.0f * eij * (-7.0f * s6 + 9.0f * s8) * ir2;
        }
        else
        {
            f0 = eij * (6.0f * s6 - 8.0f * s8) * ir;
            d0 = eij * (-42.0f * s6 + 72.0f * s8) * ir2;
        }
        const bool far = r2f >= 0.01f;
        fm = far ? f0 * cs : 0.0f;
        dfm = far ? d0 * cs : 0.0f;
    }
    fm *= a.steric_scale;
    dfm *= a.steric_scale;
    const float wx = wj.x - wi.x, wy = wj.y - wi.y, wz = wj.z - wi.z;
    const float dd = dx * wx + dy * wy + dz * wz;
    if (EMODE > 0)
    {
        const bool qon = (fli & PF_CHARGED) && (flj & PF_CHARGED) && r2f >= 1.0e-6f;
        const float qq = qon ? pi.w * pj.w : 0.0f;
        const float ir3 = ir2 * ir;
        // per unit q_i q_j, the Coulomb part of phi is -c (c = kf/r^3 or 2 kf/r^4)
        const float c = EMODE == 1 ? kf * ir3 : 2.0f * kf * ir3 * ir;
        fm -= c * qq * r;                                   // phi = -c qq
        dfm += EMODE == 1 ? 2.0f * c * qq : 3.0f * c * qq;  // fm' = 2 kf qq / r^3 or 6 kf qq / r^4
        if (on && qon)
        {
            dqi += (double)(c * dd) * (double)pj.w;
            if (flj & PF_STATIC)
                atomicAdd(&a.dq[j], (double)(c * dd) * (double)pi.w);
        }
    }
    const float phi = fm * ir, dphi = (dfm - phi) * ir;
    const float g = on ? dphi * ir * dd : 0.0f;
    const float ph = on ? phi : 0.0f;
    hx += (double)fmaf(ph, wx, g * dx);
    hy += (double)fmaf(ph, wy, g * dy);
    hz += (double)fmaf(ph, wz, g * dz);
}

template <int SMODE, int EMODE>
__global__ void __launch_bounds__(BLOCK) k_adj_pair(AdjPairArgs a)
{
    const int lane = threadIdx.x & 63;
    const int32_t ci = __builtin_amdgcn_readfirstlane(blockIdx.x * WAVES + (threadIdx.x >> 6)); // wave-uniform
    const int32_t il = lane & 7, jl = lane >> 3;
    const int32_t i = ci * CLUSTER + il;
    double hx = 0.0, hy = 0.0, hz = 0.0, dqi = 0.0;
    const float cs = (float)a.k.c_spring;
    const float kf = (float)a.kf;
    if (ci < a.ncl)
    {
        const bool iv = i < a.n;
        const float4 pi = iv ? a.pos[i] : make_float4(0.f, 0.f, 0.f, 0.f);
        const float4 ri = iv ? a.jprm[i] : make_float4(0.f, 0.f, 0.f, 0.f);
        const float4 wi = iv ? a.w[i] : make_float4(0.f, 0.f, 0.f, 0.f);
        const uint32_t fli = fbits(ri.w);
        const uint32_t ibit = 1u << (24 + il);
        const int32_t cnt16 = (a.cln[ci] + 15) & ~15;
        const uint32_t *L = a.cl + (size_t)ci * a.cap;
        const int32_t last = cnt16 > 16 ? cnt16 - 16 : 0;
        // blocks of 16 words alternate between two register sets, as in k_pair
        uint32_t wa0 = 0u, wa1 = 0u, wb0 = 0u, wb1 = 0u;
        if (cnt16 > 0)
        {
            const int32_t o = min(16, last);
            wa0 = L[jl];
            wa1 = L[8 + jl];
            wb0 = L[o + jl];
            wb1 = L[o + 8 + jl];
        }
        int32_t ja0 = (int32_t)(wa0 & CL_JMASK), ja1 = (int32_t)(wa1 & CL_JMASK);
        float4 pa0 = a.pos[ja0], ra0 = a.jprm[ja0], va0 = a.w[ja0];
        float4 pa1 = a.pos[ja1], ra1 = a.jprm[ja1], va1 = a.w[ja1];
        for (int32_t t = 0; t < cnt16; t += 32)
        {
            const bool ma0 = (wa0 & ibit) != 0u, ma1 = (wa1 & ibit) != 0u;
            {
                const uint32_t *Lw = L + min(t + 32, last);
                wa0 = Lw[jl];
                wa1 = Lw[8 + jl];
            }
            const int32_t jb0 = (int32_t)(wb0 & CL_JMASK), jb1 = (int32_t)(wb1 & CL_JMASK);
            const float4 pb0 = a.pos[jb0], rb0 = a.jprm[jb0], vb0 = a.w[jb0];
            const float4 pb1 = a.pos[jb1], rb1 = a.jprm[jb1], vb1 = a.w[jb1];
            adj_pair_term<SMODE, EMODE>(ma0, pi, ri, fli, wi, pa0, ra0, va0, ja0, a, cs, kf, hx, hy, hz, dqi);
            adj_pair_term<SMODE, EMODE>(ma1, pi, ri, fli, wi, pa1, ra1, va1, ja1, a, cs, kf, hx, hy, hz, dqi);
            if (t + 16 >= cnt16)
                break; // wave-uniform
            const bool mb0 = (wb0 & ibit) != 0u, mb1 = (wb1 & ibit) != 0u;
            {
                const uint32_t *Lw = L + min(t + 48, last);
                wb0 = Lw[jl];
                wb1 = Lw[8 + jl];
            }
            ja0 = (int32_t)(wa0 & CL_JMASK);
            ja1 = (int32_t)(wa1 & CL_JMASK);
            pa0 = a.pos[ja0];
            ra0 = a.jprm[ja0];
            va0 = a.w[ja0];
            pa1 = a.pos[ja1];
            ra1 = a.jprm[ja1];
            va1 = a.w[ja1];
            adj_pair_term<SMODE, EMODE>(mb0, pi, ri, fli, wi, pb0, rb0, vb0, jb0, a, cs, kf, hx, hy, hz, dqi);
            adj_pair_term<SMODE, EMODE>(mb1, pi, ri, fli, wi, pb1, rb1, vb1, jb1, a, cs, kf, hx, hy, hz, dqi);
        }
    }
    for (int off = 8; off < 64; off <<= 1)
    {
        hx += __shfl_xor(hx, off, 64);
        hy += __shfl_xor(hy, off, 64);
        hz += __shfl_xor(hz, off, 64);
        dqi += __shfl_xor(dqi, off, 64);
    }
    if (ci < a.ncl && jl == 0 && i < a.n)
    {
        double4 h = make_double4(hx, hy, hz, 0.0);
        if (a.accumulate)
        {
            const double4 o = a.hw[i];
            h.x += o.x;
            h.y += o.y;
            h.z += o.z;
        }
        a.hw[i] = h;
        // a static particle's dq is only ever added to atomically (by its
        // dynamic partners); a plain read-modify-write here would race
        const bool dyn_i = !(fbits(a.jprm[i].w) & PF_STATIC);
        if (EMODE > 0 && dyn_i)
            a.dq[i] += dqi;
    }
}

// Processed pairs of one search kind, as (min, max) original ids
// (Particle.cpp:205-224: i dynamic, j static or later in id order).  One
// counter update per wave keeps the atomic traffic low.
__global__ void __launch_bounds__(BLOCK) k_pairs_extract(const float4 *pos, const float4 *jprm, const uint32_t *cl,
                                                         const int32_t *cln, const int32_t *perm, int32_t n,
                                                         int32_t ncl, int32_t lcap, float rc, float rc2lo,
                                                         float rc2hi, uint32_t i_mask, uint32_t j_mask,
                                                         int32_t nonzero, uint32_t *out, int32_t *count, int64_t cap)
{
    const int lane = threadIdx.x & 63;
    const int32_t ci = blockIdx.x * WAVES + (threadIdx.x >> 6);
    if (ci >= ncl)
        return; // whole wave
    const int32_t il = lane & 7, jl = lane >> 3;
    const int32_t i = ci * CLUSTER + il;
    bool iok = i < n;
    const uint32_t fi = iok ? fbits(jprm[i].w) : 0u;
    iok = iok && (i_mask == 0u || (fi & i_mask));
    const float4 pi = iok ? pos[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    const int32_t pi_id = iok ? perm[i] : 0;
    const int32_t cnt = cln[ci];
    for (int32_t t0 = 0; t0 < cnt; t0 += 8)
    {
        const int32_t t = t0 + jl;
        bool keep = false;
        uint32_t u = 0u, v = 0u;
        if (iok && t < cnt)
        {
            const uint32_t w = cl[(size_t)ci * lcap + t];
            if ((w >> (24 + il)) & 1u)
            {
                const int32_t j = (int32_t)(w & CL_JMASK);
                const uint32_t fj = fbits(jprm[j].w);
                const int32_t pj_id = perm[j];
                if ((j_mask == 0u || (fj & j_mask)) && ((fj & PF_STATIC) || pi_id < pj_id))
                {
                    const float4 pj = pos[j];
                    const float dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
                    const float r2 = dx * dx + dy * dy + dz * dz;
                    keep = in_cutoff(pi, pj, rc, rc2lo, rc2hi, r2) &&
                           !(nonzero && !(dx != 0.0f || dy != 0.0f || dz != 0.0f));
                    u = (uint32_t)(pi_id < pj_id ? pi_id : pj_id);
                    v = (uint32_t)(pi_id < pj_id ? pj_id : pi_id);
                }
            }
        }
        const uint64_t bal = __ballot(keep);
        if (bal == 0ull)
            continue;
        int32_t base = 0;
        if (lane == 0)
            base = atomicAdd(count, (int32_t)__popcll(bal));
        base = __shfl(base, 0, 64);
        if (keep)
        {
            const int64_t slot = (int64_t)base + __popcll(bal & ((1ull << lane) - 1ull));
            if (slot < cap)
            {
                out[2 * slot] = u;
                out[2 * slot + 1] = v;
            }
        }
    }
}

inline int32_t cluster_blocks(int32_t ncl) { return (ncl + WAVES - 1) / WAVES; }

} // namespace

void launch_list_build(const ListArgs &a, hipStream_t s)
{
    if (a.ncl > 0)
        k_cluster_list<<<cluster_blocks(a.ncl), BLOCK, 0, s>>>(a);
}

// ---------------------------------------------------------------------------
// Queue layout (BS_PAIR_QUEUES).  A wave takes its cluster's list one chunk
// (at most QCHUNK = 128 words) at a time.  Each word's j data is loaded
// once (one lane per word) into LDS, and each of the 8 cluster particles gets
// an LDS queue of the chunk's words it sees (one ballot per particle and 64
// words).  The 8 lanes of a particle then drain its queue, 8 entries per
// round, so the lanes evaluate listed pairs instead of idling on entries
// their particle does not see (an entry is seen by about half of a
// cluster's particles when springs empty the inner shell).  A chunk takes as
// many rounds as its fullest queue.  Each
// lane sums a fixed subset of its particle's entries in a fixed order and
// the 8 partial sums are combined as in k_pair: deterministic.  The next
// chunk's words and j data are loaded while the current one is evaluated.
// ---------------------------------------------------------------------------
__device__ inline int32_t lanes_below(uint64_t b)
{
    return (int32_t)__builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0u));
}

template <int SMODE, int EMODE, int HYDRO, bool ENERGY>
__global__ void __launch_bounds__(BLOCK) k_pair_q(PairArgs a)
{
    if (a.status && halted(a.status))
        return;
    __shared__ float4 qpos[WAVES][QCHUNK], qprm[WAVES][QCHUNK];
    __shared__ uint8_t qidx[WAVES][CLUSTER][QCHUNK];
    const int wv = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int32_t ci = __builtin_amdgcn_readfirstlane(blockIdx.x * WAVES + wv);
    if (ci >= a.ncl)
        return; // whole wave
    const int32_t il = lane & 7, sl = lane >> 3;
    const int32_t i = ci * CLUSTER + il;
    const bool iv = i < a.n;
    const float4 pi = iv ? a.pos[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 ri = iv ? a.jprm[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    const PairI c = pair_i<SMODE, EMODE>(pi, ri, a);
    double es = 0.0, ee = 0.0, eh = 0.0;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    const Chunks d = chunks_of(__builtin_amdgcn_readfirstlane(a.cln[ci]));
    const uint32_t *L = a.cl + (size_t)ci * a.cap;
    // chunk 0 in flight (a chunk holds at least one word)
    int32_t cs = 0, cn = d.q + (d.r > 0 ? 1 : 0);
    uint32_t w0 = 0u, w1 = 0u;
    if (d.nch > 0)
    {
        w0 = lane < cn ? L[lane] : 0u;
        w1 = 64 + lane < cn ? L[64 + lane] : 0u;
    }
    float4 p0 = a.pos[w0 & CL_JMASK], r0 = a.jprm[w0 & CL_JMASK];
    float4 p1 = a.pos[w1 & CL_JMASK], r1 = a.jprm[w1 & CL_JMASK];
    for (int32_t ch = 0; ch < d.nch; ++ch)
    {
        qpos[wv][lane] = p0;
        qprm[wv][lane] = r0;
        qpos[wv][64 + lane] = p1;
        qprm[wv][64 + lane] = r1;
        int32_t cnt[CLUSTER];
#pragma unroll
        for (int k = 0; k < CLUSTER; ++k)
        {
            const bool s0 = (w0 >> (24 + k)) & 1u, s1 = (w1 >> (24 + k)) & 1u;
            const uint64_t b0 = __ballot(s0), b1 = __ballot(s1);
            const int32_t n0 = __popcll(b0);
            if (s0)
                qidx[wv][k][lanes_below(b0)] = (uint8_t)lane;
            if (s1)
                qidx[wv][k][n0 + lanes_below(b1)] = (uint8_t)(64 + lane);
            cnt[k] = n0 + __popcll(b1);
        }
        // the next chunk in flight while this one is evaluated
        if (ch + 1 < d.nch)
        {
            cs += cn;
            cn = d.q + (ch + 1 < d.r ? 1 : 0);
            w0 = lane < cn ? L[cs + lane] : 0u;
            w1 = 64 + lane < cn ? L[cs + 64 + lane] : 0u;
            p0 = a.pos[w0 & CL_JMASK];
            r0 = a.jprm[w0 & CL_JMASK];
            p1 = a.pos[w1 & CL_JMASK];
            r1 = a.jprm[w1 & CL_JMASK];
        }
        wave_lds_sync();
        int32_t mine = cnt[0], most = cnt[0];
#pragma unroll
        for (int k = 1; k < CLUSTER; ++k)
        {
            mine = il == k ? cnt[k] : mine;
            most = max(most, cnt[k]);
        }
        for (int32_t e0 = 0; e0 < most; e0 += 8) // wave-uniform rounds
        {
            const int32_t e = e0 + sl;
            const bool valid = e < mine;
            const int32_t src = valid ? (int32_t)qidx[wv][il][e] : 0;
            interact<SMODE, EMODE, HYDRO, ENERGY>(valid, pi, ri, qpos[wv][src], qprm[wv][src], a, c, fx, fy, fz, es,
                                                  ee, eh);
        }
        wave_lds_sync();
    }
    for (int off = 8; off < 64; off <<= 1)
    {
        fx += __shfl_xor(fx, off, 64);
        fy += __shfl_xor(fy, off, 64);
        fz += __shfl_xor(fz, off, 64);
        if (ENERGY)
        {
            es += __shfl_xor(es, off, 64);
            ee += __shfl_xor(ee, off, 64);
            eh += __shfl_xor(eh, off, 64);
        }
    }
    if (sl != 0 || !iv)
        return;
    if (ENERGY)
    {
        double4 en = make_double4(es, ee, eh, 0.0);
        if (a.accumulate)
        {
            const double4 o = a.epp[i];
            en.x += o.x;
            en.y += o.y;
            en.z += o.z;
        }
        a.epp[i] = en;
    }
    double4 f = make_double4(fx, fy, fz, 0.0);
    if (a.accumulate)
    {
        const double4 o = a.fnb[i];
        f.x += o.x;
        f.y += o.y;
        f.z += o.z;
    }
    a.fnb[i] = f;
}

// ---------------------------------------------------------------------------
// Row layout.  With springs excluded from the pair terms, the pairs a
// particle sees lie in a shell (workload 1: 8 A to the 13 A list radius)
// thinner than twice a cluster's size, so an entry of a cluster list is seen
// by about 60% of the cluster's particles and the cluster kernel leaves the
// other lanes idle.  The rows hold each particle's own entries, so every lane
// evaluates a listed pair; the price is one gather of j's data per pair
// instead of one per eight.
//
// k_rows: one wave per cluster; 64 list words at a time, one ballot per
// cluster particle compacts the words it sees into its row, in the cluster
// list's order.  The row kernel sums a particle's terms in that order; the
// cluster kernel sums eight lane partials and then combines them, so the two
// layouts agree to FP64 rounding, not bit for bit.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(BLOCK) k_rows(const uint32_t *cl, const int32_t *cln, int32_t ncl, int32_t cap,
                                                uint32_t *rows, int32_t *rlen, int32_t rcap, const int32_t *gate)
{
    if (gated_off(gate))
        return;
    const int lane = threadIdx.x & 63;
    const int32_t ci = blockIdx.x * WAVES + (threadIdx.x >> 6);
    if (ci >= ncl)
        return; // whole wave
    const int32_t cnt = cln[ci];
    const uint32_t *L = cl + (size_t)ci * cap;
    const uint64_t below = (1ull << lane) - 1ull;
    int32_t c[CLUSTER];
#pragma unroll
    for (int k = 0; k < CLUSTER; ++k)
        c[k] = 0;
    for (int32_t t0 = 0; t0 < cnt; t0 += 64)
    {
        const uint32_t w = t0 + lane < cnt ? L[t0 + lane] : 0u;
        const uint32_t j = w & CL_JMASK;
#pragma unroll
        for (int k = 0; k < CLUSTER; ++k)
        {
            const bool on = (w >> (24 + k)) & 1u;
            const uint64_t bal = __ballot(on);
            if (on)
            {
                const int32_t i = ci * CLUSTER + k;
                const int32_t t = c[k] + __popcll(bal & below);
                rows[((size_t)(i >> 6) * rcap + t) * 64 + (i & 63)] = j;
            }
            c[k] += __popcll(bal);
        }
    }
#pragma unroll
    for (int k = 0; k < CLUSTER; ++k)
        if (lane == k)
            rlen[ci * CLUSTER + k] = c[k];
}

// Pair forces of 64 consecutive particles per wave, one particle per lane,
// two row entries per lane per iteration; row words two iterations ahead and
// j data one iteration ahead are in flight.  Entries at or past a row's end
// (never written) are replaced by the particle itself and masked off.
template <int SMODE, int EMODE, int HYDRO, bool ENERGY>
__global__ void __launch_bounds__(BLOCK) k_pair_rows(PairArgs a)
{
    if (a.status && halted(a.status))
        return;
    const int lane = threadIdx.x & 63;
    const int32_t b = blockIdx.x * WAVES + (threadIdx.x >> 6);
    const int32_t nrb = (a.n + 63) >> 6;
    if (b >= nrb)
        return; // whole wave
    const int32_t i = b * 64 + lane;
    const bool iv = i < a.n;
    const int32_t self = iv ? i : 0;
    const float4 pi = a.pos[self];
    const float4 ri = a.jprm[self];
    const PairI c = pair_i<SMODE, EMODE>(pi, ri, a);
    const int32_t len = iv ? a.rlen[i] : 0;
    int32_t mx = len;
    for (int off = 32; off > 0; off >>= 1)
        mx = max(mx, __shfl_xor(mx, off, 64));
    double es = 0.0, ee = 0.0, eh = 0.0;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    if (mx > 0)
    {
        const uint32_t *R = a.rows + (size_t)b * a.rcap * 64 + lane;
        const int32_t last = mx - 1;
        auto word = [&](int32_t e) { return R[(size_t)min(e, last) * 64]; };
        auto jof = [&](uint32_t w, int32_t e) { return e < len ? (int32_t)w : self; };
        uint32_t w0 = word(0), w1 = word(1), n0 = word(2), n1 = word(3);
        int32_t k0 = jof(w0, 0), k1 = jof(w1, 1);
        float4 pj0 = a.pos[k0], rj0 = a.jprm[k0], pj1 = a.pos[k1], rj1 = a.jprm[k1];
        for (int32_t e = 0; e < mx; e += 2)
        {
            const uint32_t m0 = word(e + 4), m1 = word(e + 5);
            const int32_t q0 = jof(n0, e + 2), q1 = jof(n1, e + 3);
            const float4 qj0 = a.pos[q0], sj0 = a.jprm[q0];
            const float4 qj1 = a.pos[q1], sj1 = a.jprm[q1];
            interact<SMODE, EMODE, HYDRO, ENERGY>(e < len, pi, ri, pj0, rj0, a, c, fx, fy, fz, es, ee, eh);
            interact<SMODE, EMODE, HYDRO, ENERGY>(e + 1 < len, pi, ri, pj1, rj1, a, c, fx, fy, fz, es, ee, eh);
            n0 = m0;
            n1 = m1;
            pj0 = qj0;
            rj0 = sj0;
            pj1 = qj1;
            rj1 = sj1;
        }
        (void)w0;
        (void)w1;
    }
    if (!iv)
        return;
    if (ENERGY)
    {
        double4 en = make_double4(es, ee, eh, 0.0);
        if (a.accumulate)
        {
            const double4 o = a.epp[i];
            en.x += o.x;
            en.y += o.y;
            en.z += o.z;
        }
        a.epp[i] = en;
    }
    double4 f = make_double4(fx, fy, fz, 0.0);
    if (a.accumulate)
    {
        const double4 o = a.fnb[i];
        f.x += o.x;
        f.y += o.y;
        f.z += o.z;
    }
    a.fnb[i] = f;
}

template <int S, int E, int H> static void pair3(const PairArgs &a, int32_t nb, hipStream_t s)
{
    if (a.queues)
    {
        if (a.energy)
            k_pair_q<S, E, H, true><<<nb, BLOCK, 0, s>>>(a);
        else
            k_pair_q<S, E, H, false><<<nb, BLOCK, 0, s>>>(a);
        return;
    }
    if (a.rows)
    {
        const int32_t nrb = (a.n + 63) / 64, blocks = (nrb + WAVES - 1) / WAVES;
        if (a.energy)
            k_pair_rows<S, E, H, true><<<blocks, BLOCK, 0, s>>>(a);
        else
            k_pair_rows<S, E, H, false><<<blocks, BLOCK, 0, s>>>(a);
        return;
    }
    if (a.energy)
        k_pair<S, E, H, true><<<nb, BLOCK, 0, s>>>(a);
    else
        k_pair<S, E, H, false><<<nb, BLOCK, 0, s>>>(a);
}

void launch_rows(const uint32_t *cl, const int32_t *cln, int32_t ncl, int32_t cap, uint32_t *rows, int32_t *rlen,
                 int32_t rcap, const int32_t *gate, hipStream_t s)
{
    if (ncl > 0)
        k_rows<<<(ncl + WAVES - 1) / WAVES, BLOCK, 0, s>>>(cl, cln, ncl, cap, rows, rlen, rcap, gate);
}

template <int S> static void pair2(const PairArgs &a, int32_t em, int32_t h, int32_t nb, hipStream_t s)
{
    if (em == 0)
        h ? pair3<S, 0, 1>(a, nb, s) : pair3<S, 0, 0>(a, nb, s);
    else if (em == 1)
        h ? pair3<S, 1, 1>(a, nb, s) : pair3<S, 1, 0>(a, nb, s);
    else
        h ? pair3<S, 2, 1>(a, nb, s) : pair3<S, 2, 0>(a, nb, s);
}

void launch_pair(const PairArgs &a, int32_t elec_mode, int32_t hydro, int32_t nb, hipStream_t s)
{
    switch (a.steric_mode)
    {
    case -1: pair2<-1>(a, elec_mode, hydro, nb, s); break;
    case 0: pair2<0>(a, elec_mode, hydro, nb, s); break;
    case 1: pair2<1>(a, elec_mode, hydro, nb, s); break;
    case 2: pair2<2>(a, elec_mode, hydro, nb, s); break;
    default: pair2<3>(a, elec_mode, hydro, nb, s); break;
    }
}

// ---------------------------------------------------------------------------
// The adjoint's pair terms in the queue layout (as k_pair_q):
// each chunk's j data (position, parameters, adjoint weight; the index j in
// the weight's unused fourth word, for a static partner's dq) is staged in
// LDS once, queued per cluster particle, and drained 8 entries per round.
// The terms and their FP64 sums are k_adj_pair's.
// ---------------------------------------------------------------------------
constexpr int32_t ACHUNK = 64; // entries per chunk: three float4 per entry in LDS

template <int SMODE, int EMODE>
__global__ void __launch_bounds__(BLOCK) k_adj_pair_q(AdjPairArgs a)
{
    __shared__ float4 qpos[WAVES][ACHUNK], qprm[WAVES][ACHUNK], qw[WAVES][ACHUNK];
    __shared__ uint8_t qidx[WAVES][CLUSTER][ACHUNK];
    const int wv = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int32_t ci = __builtin_amdgcn_readfirstlane(blockIdx.x * WAVES + wv);
    if (ci >= a.ncl)
        return; // whole wave
    const int32_t il = lane & 7, sl = lane >> 3;
    const int32_t i = ci * CLUSTER + il;
    const bool iv = i < a.n;
    const float4 pi = iv ? a.pos[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 ri = iv ? a.jprm[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 wi = iv ? a.w[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    const uint32_t fli = fbits(ri.w);
    const float cs = (float)a.k.c_spring;
    const float kf = (float)a.kf;
    double hx = 0.0, hy = 0.0, hz = 0.0, dqi = 0.0;
    const int32_t cnt = __builtin_amdgcn_readfirstlane(a.cln[ci]);
    const uint32_t *L = a.cl + (size_t)ci * a.cap;
    // 64 entries at a time in list order; the next stretch in flight
    uint32_t w = lane < cnt ? L[lane] : 0u;
    uint32_t j = w & CL_JMASK;
    float4 pj = a.pos[j], rj = a.jprm[j], vj = a.w[j];
    for (int32_t t0 = 0; t0 < cnt; t0 += ACHUNK)
    {
        qpos[wv][lane] = pj;
        qprm[wv][lane] = rj;
        qw[wv][lane] = make_float4(vj.x, vj.y, vj.z, __uint_as_float(j));
        int32_t nq[CLUSTER];
#pragma unroll
        for (int k = 0; k < CLUSTER; ++k)
        {
            const bool on = (w >> (24 + k)) & 1u;
            const uint64_t b = __ballot(on);
            if (on)
                qidx[wv][k][lanes_below(b)] = (uint8_t)lane;
            nq[k] = __popcll(b);
        }
        const int32_t tn = t0 + ACHUNK + lane;
        w = tn < cnt ? L[tn] : 0u;
        j = w & CL_JMASK;
        pj = a.pos[j];
        rj = a.jprm[j];
        vj = a.w[j];
        wave_lds_sync();
        int32_t mine = nq[0], most = nq[0];
#pragma unroll
        for (int k = 1; k < CLUSTER; ++k)
        {
            mine = il == k ? nq[k] : mine;
            most = max(most, nq[k]);
        }
        for (int32_t e0 = 0; e0 < most; e0 += 8) // wave-uniform rounds
        {
            const int32_t e = e0 + sl;
            const bool valid = e < mine;
            const int32_t src = valid ? (int32_t)qidx[wv][il][e] : 0;
            const float4 sw = qw[wv][src];
            adj_pair_term<SMODE, EMODE>(valid, pi, ri, fli, wi, qpos[wv][src], qprm[wv][src],
                                        make_float4(sw.x, sw.y, sw.z, 0.0f), (int32_t)__float_as_uint(sw.w), a, cs,
                                        kf, hx, hy, hz, dqi);
        }
        wave_lds_sync();
    }
    for (int off = 8; off < 64; off <<= 1)
    {
        hx += __shfl_xor(hx, off, 64);
        hy += __shfl_xor(hy, off, 64);
        hz += __shfl_xor(hz, off, 64);
        dqi += __shfl_xor(dqi, off, 64);
    }
    if (sl == 0 && iv)
    {
        double4 h = make_double4(hx, hy, hz, 0.0);
        if (a.accumulate)
        {
            const double4 o = a.hw[i];
            h.x += o.x;
            h.y += o.y;
            h.z += o.z;
        }
        a.hw[i] = h;
        // a static particle's dq is only ever added to atomically (by its
        // dynamic partners); a plain read-modify-write here would race
        if (EMODE > 0 && !(fli & PF_STATIC))
            a.dq[i] += dqi;
    }
}

template <int S> static void adj2(const AdjPairArgs &a, int32_t em, int32_t nb, hipStream_t s)
{
    if (a.queues)
    {
        if (em == 0)
            k_adj_pair_q<S, 0><<<nb, BLOCK, 0, s>>>(a);
        else if (em == 1)
            k_adj_pair_q<S, 1><<<nb, BLOCK, 0, s>>>(a);
        else
            k_adj_pair_q<S, 2><<<nb, BLOCK, 0, s>>>(a);
        return;
    }
    if (em == 0)
        k_adj_pair<S, 0><<<nb, BLOCK, 0, s>>>(a);
    else if (em == 1)
        k_adj_pair<S, 1><<<nb, BLOCK, 0, s>>>(a);
    else
        k_adj_pair<S, 2><<<nb, BLOCK, 0, s>>>(a);
}

void launch_adj_pair(const AdjPairArgs &a, int32_t steric_mode, int32_t elec_mode, int32_t nb, hipStream_t s)
{
    switch (steric_mode)
    {
    case -1: adj2<-1>(a, elec_mode, nb, s); break;
    case 0: adj2<0>(a, elec_mode, nb, s); break;
    case 1: adj2<1>(a, elec_mode, nb, s); break;
    case 2: adj2<2>(a, elec_mode, nb, s); break;
    default: adj2<3>(a, elec_mode, nb, s); break;
    }
}

void launch_pairs_extract(const float4 *pos, const float4 *jprm, const uint32_t *cl, const int32_t *cln,
                          const int32_t *perm, int32_t n, int32_t ncl, int32_t lcap, float rc, float rc2lo,
                          float rc2hi, uint32_t i_mask, uint32_t j_mask, int32_t nonzero, uint32_t *out,
                          int32_t *count, int64_t cap, hipStream_t s)
{
    if (ncl > 0)
        k_pairs_extract<<<cluster_blocks(ncl), BLOCK, 0, s>>>(pos, jprm, cl, cln, perm, n, ncl, lcap, rc, rc2lo,
                                                               rc2hi, i_mask, j_mask, nonzero, out, count, cap);
}

} // namespace bs
