// bs_kernels.hip -- CDNA4 (gfx950) kernels of the step: cell binning, the
// spring gather fused with the integrator, energy reductions, step control.
//
// Every kernel is wave64-shaped; none uses floating-point atomics, so the
// forces and energies of a step are bitwise reproducible run to run.  Integer
// atomics are used only for histogram counts and flags.
#include "bs_device.hpp"

#include <cmath>

namespace bs
{

namespace
{

// ---------------------------------------------------------------------------
// Bounding box, FP64 of float coordinates (box.hpp:37-69); min/max are exact.
// ---------------------------------------------------------------------------
__device__ void box_partial_block(const float4 *pos, const float4 *shift, int32_t n, double *partial)
{
    __shared__ float sh[6][BLOCK / 64];
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        float4 p = pos[i];
        if (shift)
        {
            const float4 o = shift[i];
            p.x += o.x;
            p.y += o.y;
            p.z += o.z;
        }
        mn[0] = fminf(mn[0], p.x);
        mn[1] = fminf(mn[1], p.y);
        mn[2] = fminf(mn[2], p.z);
        mx[0] = fmaxf(mx[0], p.x);
        mx[1] = fmaxf(mx[1], p.y);
        mx[2] = fmaxf(mx[2], p.z);
    }
    for (int off = 32; off > 0; off >>= 1)
        for (int c = 0; c < 3; ++c)
        {
            mn[c] = fminf(mn[c], __shfl_xor(mn[c], off, 64));
            mx[c] = fmaxf(mx[c], __shfl_xor(mx[c], off, 64));
        }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
        for (int c = 0; c < 3; ++c)
        {
            sh[c][wave] = mn[c];
            sh[3 + c][wave] = mx[c];
        }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        for (int c = 0; c < 3; ++c)
        {
            float a = sh[c][0], b = sh[3 + c][0];
            for (int w = 1; w < BLOCK / 64; ++w)
            {
                a = fminf(a, sh[c][w]);
                b = fmaxf(b, sh[3 + c][w]);
            }
            partial[6 * blockIdx.x + c] = a;
            partial[6 * blockIdx.x + 3 + c] = b;
        }
    }
}

__global__ void __launch_bounds__(BLOCK) k_box_partial(const float4 *pos, const float4 *shift, int32_t n,
                                                       double *partial, const int32_t *gate)
{
    if (gated_off(gate))
        return;
    box_partial_block(pos, shift, n, partial);
}

// Combines up to BLOCK block partials (one per thread) with wave shuffles
// into out[6]; every thread of the (BLOCK-wide) block takes part.
__device__ void block_box_final(const double *partial, int32_t nb, double *out)
{
    __shared__ double sh[6][BLOCK / 64];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    double v[6];
    for (int c = 0; c < 3; ++c)
    {
        v[c] = t < nb ? partial[6 * t + c] : INFINITY;
        v[3 + c] = t < nb ? partial[6 * t + 3 + c] : -INFINITY;
    }
    for (int off = 32; off > 0; off >>= 1)
        for (int c = 0; c < 3; ++c)
        {
            v[c] = fmin(v[c], __shfl_xor(v[c], off, 64));
            v[3 + c] = fmax(v[3 + c], __shfl_xor(v[3 + c], off, 64));
        }
    if (lane == 0)
        for (int c = 0; c < 6; ++c)
            sh[c][wave] = v[c];
    __syncthreads();
    if (t < 6)
    {
        double r = sh[t][0];
        for (int w = 1; w < BLOCK / 64; ++w)
            r = t < 3 ? fmin(r, sh[t][w]) : fmax(r, sh[t][w]);
        out[t] = r;
    }
}

__global__ void __launch_bounds__(BLOCK) k_box_final(const double *partial, int32_t nb, double *box,
                                                     const int32_t *gate)
{
    if (gated_off(gate))
        return;
    block_box_final(partial, nb, box);
}

// ---------------------------------------------------------------------------
// Cell binning: ids, counting sort, deterministic in-cell order.
// exact_mode 1: nsearch.hpp:277-293 (size_t((x - min)/edge), clamped), FP64.
// exact_mode 0: the engine's own grid, edge >= min_edge on every axis.
// ---------------------------------------------------------------------------
//
// box_part (optional): box block partials (k_prestep, k_box_partial); each workgroup
// then reduces them itself (the first one also stores the box for later
// kernels), which saves the launch of k_box_final.  xref (optional): the
// positions are also copied there (the list's reference positions).
__global__ void __launch_bounds__(BLOCK) k_cell_ids(const float4 *pos, const float4 *shift, const float4 *jprm,
                                                    uint32_t mask, int32_t n, double *box, const double *box_part,
                                                    int32_t box_nb, float4 *xref, int32_t d0, int32_t d1, int32_t d2,
                                                    double edge_exact, float min_edge, int32_t exact_mode,
                                                    int32_t *cell_of, int32_t *counts, const int32_t *gate)
{
    if (gated_off(gate))
        return;
    __shared__ double sbox[6];
    if (box_part)
    {
        block_box_final(box_part, box_nb, sbox);
        __syncthreads();
        if (blockIdx.x == 0 && threadIdx.x < 6)
            box[threadIdx.x] = sbox[threadIdx.x];
    }
    else
    {
        if (threadIdx.x < 6)
            sbox[threadIdx.x] = box[threadIdx.x];
        __syncthreads();
    }
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    if (xref)
        xref[i] = pos[i];
    const uint32_t fl = fbits(jprm[i].w);
    if (mask != 0u && (fl & mask) == 0u)
    {
        cell_of[i] = -1;
        return;
    }
    float4 p = pos[i];
    if (shift)
    {
        const float4 o = shift[i];
        p.x += o.x;
        p.y += o.y;
        p.z += o.z;
    }
    int32_t c[3];
    const float xs[3] = {p.x, p.y, p.z};
    const int32_t dims[3] = {d0, d1, d2};
    for (int a = 0; a < 3; ++a)
    {
        if (exact_mode)
        {
            const double q = ((double)xs[a] - sbox[a]) / edge_exact;
            const uint64_t u = q > 0.0 ? (uint64_t)q : 0u;
            c[a] = (int32_t)(u < (uint64_t)(dims[a] - 1) ? u : (uint64_t)(dims[a] - 1));
        }
        else
        {
            c[a] = internal_coord(xs[a], sbox[a], sbox[3 + a] - sbox[a], dims[a], min_edge);
        }
    }
    const int32_t id = c[0] + c[1] * d0 + c[2] * d0 * d1;
    cell_of[i] = id;
    atomicAdd(&counts[id], 1);
}

// Exclusive scan of counts[0..n) into start[0..n]; one workgroup of 1024.
__global__ void __launch_bounds__(1024) k_scan(const int32_t *counts, int32_t *start, int32_t n, const int32_t *gate)
{
    if (gated_off(gate))
        return;
    __shared__ int32_t wsum[16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int32_t carry = 0;
    for (int32_t base = 0; base < n; base += 4096)
    {
        const int32_t idx = base + 4 * threadIdx.x;
        int32_t a[4];
        for (int q = 0; q < 4; ++q)
            a[q] = idx + q < n ? counts[idx + q] : 0;
        const int32_t tsum = a[0] + a[1] + a[2] + a[3];
        int32_t x = tsum;
        for (int off = 1; off < 64; off <<= 1)
        {
            const int32_t y = __shfl_up(x, off, 64);
            if (lane >= off)
                x += y;
        }
        if (lane == 63)
            wsum[wave] = x;
        __syncthreads();
        if (wave == 0)
        {
            int32_t w = lane < 16 ? wsum[lane] : 0;
            for (int off = 1; off < 16; off <<= 1)
            {
                const int32_t y = __shfl_up(w, off, 64);
                if (lane >= off)
                    w += y;
            }
            if (lane < 16)
                wsum[lane] = w;
        }
        __syncthreads();
        int32_t run = x - tsum + (wave > 0 ? wsum[wave - 1] : 0) + carry;
        for (int q = 0; q < 4; ++q)
        {
            if (idx + q < n)
                start[idx + q] = run;
            run += a[q];
        }
        carry += wsum[15];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        start[n] = carry;
}

// drain: cursor holds the cell counts, which are counted back down to zero
// (so that the next binning starts from zeroed counts without a fill);
// otherwise cursor starts at zero and counts up.  The order inside a cell
// is arbitrary either way (k_cell_sort fixes it).
__global__ void k_scatter(const int32_t *cell_of, int32_t n, const int32_t *start, int32_t *cursor, int32_t drain,
                          int32_t *tmp, const int32_t *gate)
{
    if (gated_off(gate))
        return;
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int32_t c = cell_of[i];
    if (c < 0)
        return;
    const int32_t slot = drain ? atomicSub(&cursor[c], 1) - 1 : atomicAdd(&cursor[c], 1);
    tmp[start[c] + slot] = i;
}

// Places each member at its rank by key within its cell (keys are distinct).
__global__ void k_cell_sort(const int32_t *cell_of, int32_t n, const int32_t *start, const int32_t *tmp,
                            const int32_t *key, int32_t *out, const int32_t *gate)
{
    if (gated_off(gate))
        return;
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int32_t c = cell_of[i];
    if (c < 0)
        return;
    const int32_t ki = key ? key[i] : i;
    const int32_t b = start[c], e = start[c + 1];
    int32_t rank = 0;
    for (int32_t t = b; t < e; ++t)
    {
        const int32_t m = tmp[t];
        rank += (key ? key[m] : m) < ki;
    }
    out[b + rank] = i;
}

// ---------------------------------------------------------------------------
// IMPALA (imp.hpp:29-159, ForceField.cpp:62-86) in FP64: returns the energy
// and adds the force (the negated gradient vector times the scale) to f.
// The curved geometry z' = sqrt(x^2 + (z - c)^2) - R cancels two numbers of
// the size of the tube radius, so it is formed in double.
// ---------------------------------------------------------------------------
__device__ inline double imp_eval(float xf, float zf, float sasa, float tr, float uo, float lo, float uc, float lc,
                                  float scale, double c_spring, double f[3])
{
    const double alpha = (double)1.99f, z0 = 15.75, alip = (double)-0.018f;
    const bool flat = uo == 0.0f && lo == 0.0f && uc == 0.0f && lc == 0.0f;
    const double x = xf, z = zf;
    const double ru = uc == 0.0f ? 1000000.0 : fabs(1.0 / (double)uc);
    const double rl = lc == 0.0f ? 1000000.0 : fabs(1.0 / (double)lc);
    const int su = (uc > 0.0f) - (uc < 0.0f), sl = (lc > 0.0f) - (lc < 0.0f);
    const double cu = (double)uo - su * ru, cl = -(double)lo - sl * rl;
    const double vu0 = x, vu2 = z - cu, vl0 = x, vl2 = z - cl;
    const double nu = sqrt(vu0 * vu0 + vu2 * vu2);
    const double nl = sqrt(vl0 * vl0 + vl2 * vl2);
    const double zu = su == 0 ? z : (z > cu ? su * nu + uo - ru : -su * nu + uo - ru);
    const double zl = sl == 0 ? z : (z > cl ? sl * nl - lo - rl : -sl * nl - lo - rl);
    const double wu = zu - uo, wl = zl + lo;
    const double xu = exp(alpha * (fabs(wu) - z0)), xl = exp(alpha * (fabs(wl) - z0));
    const double pre = -(double)sasa * tr + alip * sasa;
    const double eu = pre * (0.5 - 1.0 / (1.0 + xu));
    const double el = pre * (0.5 - 1.0 / (1.0 + xl));
    double dcu = alpha * wu * xu / ((xu + 1.0) * (xu + 1.0) * fabs(wu));
    double dcl = alpha * wl * xl / ((xl + 1.0) * (xl + 1.0) * fabs(wl));
    if (!isfinite(dcu))
        dcu = 0.0;
    if (!isfinite(dcl))
        dcl = 0.0;
    double du[3] = {0.0, 0.0, 1.0}, dl[3] = {0.0, 0.0, 1.0};
    if (su != 0)
    {
        const double inv = nu > 1e-40 ? 1.0 / nu : 0.0;
        du[0] = vu0 * inv;
        du[2] = vu2 * inv;
    }
    if (sl != 0)
    {
        const double inv = nl > 1e-40 ? 1.0 / nl : 0.0;
        dl[0] = vl0 * inv;
        dl[2] = vl2 * inv;
    }
    const double mu = pre * dcu * c_spring, ml = pre * dcl * c_spring;
    for (int c = 0; c < 3; ++c)
        f[c] -= (double)scale * (flat ? du[c] * mu : du[c] * mu + dl[c] * ml);
    return (double)scale * (flat ? eu : eu + el);
}

// One spring edge seen from row i: force on i and the energy share.  The
// extension len - L0 is formed from an FP64 length (rsqrt seed plus one
// Newton step) so that it keeps full accuracy near rest.
template <bool COMP>
__device__ inline void spring_edge(float4 pi, float4 li, float4 pj, const float4 *plo, int32_t j, float k, float l0,
                                   bool jstatic, double cs, float scale, bool energy, double &fx, double &fy,
                                   double &fz, double &e)
{
    double dx = (double)pj.x - (double)pi.x;
    double dy = (double)pj.y - (double)pi.y;
    double dz = (double)pj.z - (double)pi.z;
    if (COMP)
    {
        const float4 lj = plo[j];
        dx += (double)lj.x - (double)li.x;
        dy += (double)lj.y - (double)li.y;
        dz += (double)lj.z - (double)li.z;
    }
    const double r2 = dx * dx + dy * dy + dz * dz;
    double len = 0.0;
    float inv = 0.0f;
    if (r2 > 0.0)
    {
        inv = rsqrtf((float)r2);
        const double s = r2 * (double)inv;
        len = fma(fma(-s, s, r2), 0.5 * (double)inv, s);
    }
    const double ext = len - (double)l0;
    const double y = inv; // 1/len refined in FP64 for the unit vector
    const double fm = (double)k * ext * cs * (y * fma(-0.5 * r2, y * y, 1.5));
    fx = fma(dx, fm, fx);
    fy = fma(dy, fm, fy);
    fz = fma(dz, fm, fz);
    if (energy)
        e += (jstatic ? 1.0 : 0.5) * 0.5 * (double)scale * (double)k * ext * ext;
}

// Cell of a field grid holding a position, -1 off the grid
// (GridCoordinatesSystem.hpp:85-93,206-218: off when x < min or x > max - 1e-6).
__device__ inline int64_t grid_cell(const GridView &g, float x, float y, float z)
{
    const double p[3] = {x, y, z};
    int64_t q[3];
    for (int c = 0; c < 3; ++c)
    {
        if (!(p[c] >= g.origin[c]) || p[c] > g.bmax[c] - 1e-6)
            return -1;
        q[c] = (int64_t)(int32_t)((p[c] - g.origin[c]) / g.cell[c]);
        if (q[c] >= g.shape[c])
            return -1;
    }
    return (q[0] * g.shape[1] + q[1]) * g.shape[2] + q[2];
}

// One particle's interactions with the probe (Particle::addStericProbeForce /
// addElectrostaticProbeForce, Particle.cpp:434-465), in FP64: the force on
// the particle goes to f; the pair energies (in full) to es, ee.  The
// formulas are the pair kernel's (energy/steric.hpp, electrostatic.hpp) with
// the probe as the first particle; no cutoff and no spring exclusion.
__device__ inline void probe_pair(const ProbeArgs &p, const float4 q, float4 pi, float4 ri, double cs, double f[3],
                                  double &es, double &ee)
{
    const double dx = (double)q.x - (double)pi.x, dy = (double)q.y - (double)pi.y, dz = (double)q.z - (double)pi.z;
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    const double inv = r > 1e-40 ? 1.0 / r : 0.0; // a zero vector stays zero (Vector3f::normalize)
    double fm = 0.0;
    if (p.steric)
    {
        double e = 0.0, m = 0.0;
        if (p.steric_mode == 0)
        { // linear: overlap of the raw radii
            const double dv = r - ((double)p.radius + (double)ri.x);
            if (!(dv > 0.0))
            {
                e = 0.5 * dv * dv;
                m = -fabs(dv) * cs;
            }
        }
        else if (r >= 0.1)
        {
            const bool prod = p.steric_mode == 3; // Zacharias: raw products; else square-root factors
            const double eij = (double)p.eps * (double)ri.y;
            const double R = (double)p.radius * (double)ri.x;
            const double sr = R * inv, s2 = sr * sr, s6 = s2 * s2 * s2;
            if (p.steric_mode == 1)
            { // Amber 12-6
                e = eij * s6 * s6 - 2.0 * eij * s6;
                m = (-12.0 * eij * s6 * s6 + 12.0 * eij * s6) * inv * cs;
            }
            else if (p.steric_mode == 2)
            { // Lewitt 8-6
                e = 3.0 * eij * s6 * s2 - 4.0 * eij * s6;
                m = (-24.0 * eij * s6 * s2 + 24.0 * eij * s6) * inv * cs;
            }
            else if (prod)
            { // Zacharias 8-6
                e = eij * s6 * s2 - eij * s6;
                m = (-8.0 * eij * s6 * s2 + 6.0 * eij * s6) * inv * cs;
            }
        }
        es += (double)p.steric_scale * e;
        fm += (double)p.steric_scale * m;
    }
    if (p.elec && r >= 0.001)
    {
        const double qq = (double)q.w * (double)pi.w;
        if (p.ddd)
        { // eps(r) = eps r
            ee += p.ke * qq * inv * inv;
            fm += -2.0 * p.kf * qq * inv * inv * inv;
        }
        else
        {
            ee += p.ke * qq * inv;
            fm += -p.kf * qq * inv * inv;
        }
    }
    f[0] += dx * inv * fm;
    f[1] += dy * inv * fm;
    f[2] += dz * inv * fm;
}

// ---------------------------------------------------------------------------
// Springs (CSR gather, G lanes per row, 8 edges in flight per lane) fused with
// viscosity, external and IMPALA forces and the viscous semi-implicit Euler
// step (SpringNetwork.cpp:63-90,202-241; Spring.cpp:20-34; Particle.cpp:80-128).
// ---------------------------------------------------------------------------
template <int G, bool INTEGRATE, bool COMP>
__global__ void __launch_bounds__(BLOCK) k_spring(SpringArgs a)
{
    const int32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int32_t row = tid / G;
    const int32_t lane = tid % G;
    if (halted(a.status))
    { // a stopped call: the state is carried over to the other position buffer
        if (INTEGRATE && row < a.n && lane == 0)
        {
            a.pos_out[row] = a.pos[row];
            if (COMP)
                a.plo_out[row] = a.plo[row];
        }
        return;
    }
    if (INTEGRATE && tid == 0 && a.steps_done)
        atomicAdd(a.steps_done, 1);
    double espr = 0.0, eke = 0.0, eimp = 0.0, efld = 0.0;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    double pf[3] = {0.0, 0.0, 0.0}, pes = 0.0, pee = 0.0; // force on the probe, probe pair energies
    const bool active = row < a.n;
    float4 pi = make_float4(0.f, 0.f, 0.f, 0.f), li = make_float4(0.f, 0.f, 0.f, 0.f);
    uint32_t fli = PF_STATIC;
    if (active)
    {
        pi = a.pos[row];
        fli = fbits(a.jprm[row].w);
        if (COMP)
            li = a.plo[row];
    }
    const bool dyn = active && !(fli & PF_STATIC);
    if (dyn && a.spring_enable)
    {
        const int32_t beg = a.sp_ptr[row], end = a.sp_ptr[row + 1];
        const double cs = a.k.c_spring * (double)a.spring_scale;
        constexpr int U = 8;
        // the edge words of the next iteration load while this one gathers
        // and computes (two dependent loads per edge: the word, then x_j)
        int32_t jr[U];
        float kk[U], ll[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
        {
            const int32_t e = beg + lane + u * G;
            const bool ok = e < end;
            jr[u] = ok ? a.sp_j[e] : 0;
            kk[u] = ok ? (a.sp_k ? a.sp_k[e] : a.k_uni) : 0.0f;
            ll[u] = ok ? a.sp_l0[e] : 0.0f;
        }
        for (int32_t e0 = beg + lane; e0 < end; e0 += U * G)
        {
            float4 pj[U];
#pragma unroll
            for (int u = 0; u < U; ++u)
                pj[u] = a.pos[jr[u] & 0x7fffffff];
            int32_t jn[U];
            float kn[U], ln[U];
#pragma unroll
            for (int u = 0; u < U; ++u)
            {
                const int32_t e = e0 + (U + u) * G;
                const bool ok = e < end;
                jn[u] = ok ? a.sp_j[e] : 0;
                kn[u] = ok ? (a.sp_k ? a.sp_k[e] : a.k_uni) : 0.0f;
                ln[u] = ok ? a.sp_l0[e] : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (e0 + u * G < end)
                    spring_edge<COMP>(pi, li, pj[u], a.plo, jr[u] & 0x7fffffff, kk[u], ll[u], jr[u] < 0, cs,
                                      a.spring_scale, a.energy != 0, fx, fy, fz, espr);
#pragma unroll
            for (int u = 0; u < U; ++u)
            {
                jr[u] = jn[u];
                kk[u] = kn[u];
                ll[u] = ln[u];
            }
        }
    }
    // reduce the force (and the spring energy) over the G lanes of the row
    for (int off = G / 2; off > 0; off >>= 1)
    {
        fx += __shfl_xor(fx, off, G);
        fy += __shfl_xor(fy, off, G);
        fz += __shfl_xor(fz, off, G);
        if (a.energy)
            espr += __shfl_xor(espr, off, G);
    }
    if (dyn && lane == 0)
    {
        float4 v = a.vel[row];
        double sx = fx, sy = fy, sz = fz;
        if (a.fnb)
        {
            const double4 fn = a.fnb[row];
            sx += fn.x;
            sy += fn.y;
            sz += fn.z;
        }
        if (a.fext)
        {
            const float4 fe = a.fext[row];
            sx += fe.x;
            sy += fe.y;
            sz += fe.z;
        }
        if (a.probe.state)
        { // the probe: the particle's share here, the opposite on the probe
            double f[3] = {0.0, 0.0, 0.0};
            probe_pair(a.probe, a.probe.state->x, pi, a.jprm[row], a.k.c_spring, f, pes, pee);
            sx += f[0];
            sy += f[1];
            sz += f[2];
            pf[0] = -f[0];
            pf[1] = -f[1];
            pf[2] = -f[2];
        }
        const float4 ip = a.iprm[row];
        // field forces: nearest-lower-cell lookup, no interpolation
        // (Particle.cpp:136-189); off-grid particles get nothing
        if (a.gpot.cells)
        {
            const int64_t c = grid_cell(a.gpot, pi.x, pi.y, pi.z);
            if (c >= 0)
            {
                const float4 g = a.gpot.cells[c];
                const double qs = (double)pi.w * (double)a.gpot.scale;
                sx += (double)g.x * qs;
                sy += (double)g.y * qs;
                sz += (double)g.z * qs;
                efld += (double)g.w * qs * a.e_field;
            }
        }
        if (a.gden.cells)
        {
            const int64_t c = grid_cell(a.gden, pi.x, pi.y, pi.z);
            if (c >= 0)
            {
                const float4 g = a.gden.cells[c];
                const double bs = (double)ip.w * (double)a.gden.scale;
                sx += (double)g.x * bs;
                sy += (double)g.y * bs;
                sz += (double)g.z * bs;
            }
        }
        if (a.imp_enable)
        {
            double f[3] = {sx, sy, sz};
            eimp += imp_eval(pi.x, pi.z, ip.y, ip.z, a.imp_uo, a.imp_lo, a.imp_uc, a.imp_lc, a.imp_scale,
                             a.k.c_spring, f);
            sx = f[0];
            sy = f[1];
            sz = f[2];
        }
        // the position-dependent sum is rounded once; the viscous term
        // -gamma v follows in FP32 (Particle.cpp:80-84)
        float fxf = (float)sx, fyf = (float)sy, fzf = (float)sz;
        if (a.visc_enable)
        {
            fxf -= v.x * a.visc;
            fyf -= v.y * a.visc;
            fzf -= v.z * a.visc;
        }
        const float fx = fxf, fy = fyf, fz = fzf;
        if (!INTEGRATE)
        {
            a.fout[row] = make_float4(fx, fy, fz, 0.0f);
        }
        else if (fli & PF_RIGID)
        {
            // placed by its body after the body solve (k_rigid_place); a rigid
            // particle's kinetic energy is reset before it is summed
            // (RigidBody.cpp:629, SpringNetwork.cpp:232)
            a.frig[row] = make_double4(fx, fy, fz, 0.0);
        }
        else
        {
            const float m = ip.x;
            const float dt = a.dt;
            if (m > 0.0f)
            {
                v.x = v.x + (fx / m) * dt;
                v.y = v.y + (fy / m) * dt;
                v.z = v.z + (fz / m) * dt;
            }
            eke += 0.5 * (double)m * ((double)v.x * v.x + (double)v.y * v.y + (double)v.z * v.z) * a.k.c_kinetic;
            float4 xn = pi;
            if (COMP)
            {
                // two-sum of hi + v*dt, carried into the low word
                float4 ln = li;
                const float u[3] = {v.x * dt, v.y * dt, v.z * dt};
                float *hp[3] = {&xn.x, &xn.y, &xn.z};
                float *lp[3] = {&ln.x, &ln.y, &ln.z};
                for (int c = 0; c < 3; ++c)
                {
                    const float h = *hp[c];
                    const float s = h + u[c];
                    const float bb = s - h;
                    const float err = (h - (s - bb)) + (u[c] - bb);
                    const float lo = *lp[c] + err;
                    const float h2 = s + lo;
                    *lp[c] = lo - (h2 - s);
                    *hp[c] = h2;
                }
                a.plo_out[row] = ln;
            }
            else
            {
                xn.x = pi.x + v.x * dt;
                xn.y = pi.y + v.y * dt;
                xn.z = pi.z + v.z * dt;
            }
            if (!isfinite(xn.x) || !isfinite(xn.y) || !isfinite(xn.z))
                atomicMin(&a.status[1], a.perm[row]);
            const float4 xr = a.xref[row];
            const float ex = xn.x - xr.x, ey = xn.y - xr.y, ez = xn.z - xr.z;
            if (!(ex * ex + ey * ey + ez * ez <= a.half_skin2))
                atomicOr(&a.status[7], 1);
            a.vel[row] = v;
            a.pos_out[row] = xn;
        }
    }
    if (a.energy && active && lane == 0)
        a.eps[row] = make_double4(espr, eke, eimp, efld);
    if (a.probe.state)
    { // block sums of the probe's force and pair energies, in a fixed order
        __shared__ double psh[BLOCK / 64];
        const double v[5] = {pf[0], pf[1], pf[2], pes, pee};
        for (int k = 0; k < 5; ++k)
        {
            const double t = block_sum(v[k], psh);
            if (threadIdx.x == 0)
                a.probe.part[8 * blockIdx.x + k] = t;
        }
    }
}

// The probe's step (SpringNetwork.cpp:156-182): its force is the sum of the
// spring kernel's block sums (in block order), minus gamma v when viscosity
// is on, rounded to float like a particle's; then the semi-implicit Euler
// step and its kinetic energy (Particle.cpp:109-128).  One workgroup.
__global__ void __launch_bounds__(BLOCK) k_probe(ProbeArgs p, int32_t nparts, float dt, int32_t visc_enable,
                                                 float visc, double ck, int32_t integrate)
{
    if (halted(p.status))
        return;
    __shared__ double sh[BLOCK / 64];
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int32_t b = threadIdx.x; b < nparts; b += blockDim.x)
        for (int k = 0; k < 5; ++k)
            acc[k] += p.part[8 * b + k];
    for (int k = 0; k < 5; ++k)
        acc[k] = block_sum(acc[k], sh);
    if (threadIdx.x != 0)
        return;
    ProbeDev &st = *p.state;
    float4 v = st.v;
    float fx = (float)acc[0], fy = (float)acc[1], fz = (float)acc[2];
    if (visc_enable)
    {
        fx -= v.x * visc;
        fy -= v.y * visc;
        fz -= v.z * visc;
    }
    st.f = make_double4(fx, fy, fz, 0.0);
    st.e[0] = acc[3];
    st.e[1] = acc[4];
    st.e[2] = 0.0;
    if (!integrate)
        return;
    const float m = p.mass;
    if (m > 0.0f)
    {
        v.x = v.x + (fx / m) * dt;
        v.y = v.y + (fy / m) * dt;
        v.z = v.z + (fz / m) * dt;
    }
    st.e[2] = 0.5 * (double)m * ((double)v.x * v.x + (double)v.y * v.y + (double)v.z * v.z) * ck;
    float4 x = st.x;
    x.x = x.x + v.x * dt;
    x.y = x.y + v.y * dt;
    x.z = x.z + v.z * dt;
    st.x = x;
    st.v = v;
}

// ---------------------------------------------------------------------------
// Rigid bodies.  k_rigid_solve: one workgroup of 1024 per body sums F and the
// torque (x_i - X) x F_i over its dynamic particles (RigidBody.cpp:567-580)
// in a fixed order (per lane, then wave shuffles, then the 16 waves in
// order), then one lane advances the body (RigidBody.cpp:430-484: linear
// Euler with the a dt^2/2 term, L += tau dt, omega = I^-1 L, rotation by the
// quaternion of axis omega -- not normalised -- and angle |omega| dt).
// k_rigid_place: x_i = X + q (p0_i, 0) q^-1 and v_i = (x_i - x_prev) / dt
// (RigidBody.cpp:600-635).  Everything in FP64; positions stored in float.
// ---------------------------------------------------------------------------
__device__ inline void qmul(const double a[4], const double b[4], double o[4])
{
    o[0] = b[0] * a[3] + a[0] * b[3] + (a[1] * b[2] - a[2] * b[1]);
    o[1] = b[1] * a[3] + a[1] * b[3] + (a[2] * b[0] - a[0] * b[2]);
    o[2] = b[2] * a[3] + a[2] * b[3] + (a[0] * b[1] - a[1] * b[0]);
    o[3] = a[3] * b[3] - (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]);
}

__device__ inline void qnormalize(double q[4])
{
    const double m = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    if (m > 0.005) // Quaternion::normalize (Quaternion.cpp:11-21)
        for (int c = 0; c < 4; ++c)
            q[c] /= m;
}

constexpr int RIGID_BLOCK = 1024;

__global__ void __launch_bounds__(RIGID_BLOCK) k_rigid_solve(RigidArgs a)
{
    if (halted(a.status))
        return;
    __shared__ double sh[RIGID_BLOCK / 64][6];
    const int32_t b = blockIdx.x;
    RigidBody &B = a.bodies[b];
    const double X0 = B.X[0], X1 = B.X[1], X2 = B.X[2];
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int32_t s = a.bstart[b] + threadIdx.x; s < a.bstart[b + 1]; s += blockDim.x)
    {
        const double4 f = a.frig[a.slot_p[s]];
        const double4 x = a.rx[s];
        const double r0 = x.x - X0, r1 = x.y - X1, r2 = x.z - X2;
        acc[0] += f.x;
        acc[1] += f.y;
        acc[2] += f.z;
        acc[3] += r1 * f.z - r2 * f.y;
        acc[4] += r2 * f.x - r0 * f.z;
        acc[5] += r0 * f.y - r1 * f.x;
    }
    for (int k = 0; k < 6; ++k)
        acc[k] = wave_sum(acc[k]);
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 6; ++k)
            sh[threadIdx.x >> 6][k] = acc[k];
    __syncthreads();
    if (threadIdx.x != 0)
        return;
    double tot[6] = {0, 0, 0, 0, 0, 0};
    for (int w = 0; w < RIGID_BLOCK / 64; ++w)
        for (int k = 0; k < 6; ++k)
            tot[k] += sh[w][k];
    const double dt = a.dt;
    double acc3[3];
    for (int c = 0; c < 3; ++c)
    {
        acc3[c] = tot[c] / B.M;
        B.v[c] += acc3[c] * dt;
    }
    for (int c = 0; c < 3; ++c)
        B.X[c] += B.v[c] * dt + acc3[c] * 0.5 * dt * dt;
    for (int c = 0; c < 3; ++c)
        B.L[c] += tot[3 + c] * dt;
    for (int r = 0; r < 3; ++r)
        B.w[r] = B.iinv[3 * r] * B.L[0] + B.iinv[3 * r + 1] * B.L[1] + B.iinv[3 * r + 2] * B.L[2];
    const double wn = sqrt(B.w[0] * B.w[0] + B.w[1] * B.w[1] + B.w[2] * B.w[2]);
    const double half = 0.5 * wn * dt;
    const double sn = sin(half), cn = cos(half);
    double dq[4] = {B.w[0] * sn, B.w[1] * sn, B.w[2] * sn, cn};
    qnormalize(dq);
    double nq[4];
    qmul(dq, B.q, nq);
    qnormalize(nq);
    for (int c = 0; c < 4; ++c)
        B.q[c] = nq[c];
}

__global__ void k_rigid_place(RigidArgs a)
{
    const int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.nr || halted(a.status))
        return;
    const int32_t p = a.slot_p[s];
    const RigidBody &B = a.bodies[a.slot_b[s]];
    const double4 p0 = a.p0[s];
    const double q[4] = {B.q[0], B.q[1], B.q[2], B.q[3]};
    const double P[4] = {p0.x, p0.y, p0.z, 0.0};
    const double n2 = q[3] * q[3] + q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
    const double qi[4] = {-q[0] / n2, -q[1] / n2, -q[2] / n2, q[3] / n2};
    double t[4], r[4];
    qmul(q, P, t);
    qmul(t, qi, r);
    const double4 xp = a.rx[s];
    const double nx = B.X[0] + r[0], ny = B.X[1] + r[1], nz = B.X[2] + r[2];
    a.rx[s] = make_double4(nx, ny, nz, 0.0);
    const float4 xn = make_float4((float)nx, (float)ny, (float)nz, a.pos[p].w);
    float4 v = a.vel[p];
    v.x = (float)((nx - xp.x) / a.dt);
    v.y = (float)((ny - xp.y) / a.dt);
    v.z = (float)((nz - xp.z) / a.dt);
    a.vel[p] = v;
    a.pos_out[p] = xn;
    if (!isfinite(xn.x) || !isfinite(xn.y) || !isfinite(xn.z))
        atomicMin(&a.status[1], a.perm[p]);
    const float4 xr = a.xref[p];
    const float ex = xn.x - xr.x, ey = xn.y - xr.y, ez = xn.z - xr.z;
    if (!(ex * ex + ey * ey + ez * ez <= a.half_skin2))
        atomicOr(&a.status[7], 1);
}

// Sum of positions in FP64, block partials then one fixed-order final sum.
__global__ void __launch_bounds__(BLOCK) k_centroid_partial(const float4 *pos, int32_t n, double *partial)
{
    __shared__ double sh[BLOCK / 64];
    double acc[3] = {0, 0, 0};
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        const float4 p = pos[i];
        acc[0] += p.x;
        acc[1] += p.y;
        acc[2] += p.z;
    }
    for (int c = 0; c < 3; ++c)
    {
        const double t = block_sum(acc[c], sh);
        if (threadIdx.x == 0)
            partial[3 * blockIdx.x + c] = t;
    }
}

__global__ void __launch_bounds__(BLOCK) k_centroid_final(const double *partial, int32_t nb, int32_t n, double *out)
{
    __shared__ double sh[BLOCK / 64];
    double acc[3] = {0, 0, 0};
    for (int32_t b = threadIdx.x; b < nb; b += blockDim.x)
        for (int c = 0; c < 3; ++c)
            acc[c] += partial[3 * b + c];
    for (int c = 0; c < 3; ++c)
    {
        const double t = block_sum(acc[c], sh);
        if (threadIdx.x == 0)
            out[c] = n > 0 ? t / (double)n : 0.0;
    }
}

// Energies per segment of contiguous particles (one workgroup each), then per
// system over its segments in order, then the total over systems in order.
// epp: steric, elec, hydro; eps: spring, kinetic, imp, field (elec).
// Slots: spring, elec, steric, kinetic, imp, hydro.
__global__ void __launch_bounds__(BLOCK) k_seg_energies(const double4 *epp, const double4 *eps,
                                                        const int32_t *seg_start, double *seg_part)
{
    __shared__ double sh[BLOCK / 64];
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int32_t i = seg_start[blockIdx.x] + threadIdx.x; i < seg_start[blockIdx.x + 1]; i += blockDim.x)
    {
        if (epp)
        {
            const double4 e = epp[i];
            acc[2] += e.x;
            acc[1] += e.y;
            acc[5] += e.z;
        }
        const double4 f = eps[i];
        acc[0] += f.x;
        acc[3] += f.y;
        acc[4] += f.z;
        acc[1] += f.w;
    }
    for (int k = 0; k < 6; ++k)
    {
        const double t = block_sum(acc[k], sh);
        if (threadIdx.x == 0)
            seg_part[6 * blockIdx.x + k] = t;
    }
}

constexpr int32_t SYS_LDS_SEGS = 1024; // segment partials staged in LDS (48 KiB)

__global__ void __launch_bounds__(BLOCK) k_sys_energies(const double *seg_part, const int32_t *sys_seg,
                                                        int32_t nsys, double *out_sys, double *out_total,
                                                        const ProbeDev *probe)
{
    // the partials are loaded by every lane at once, so the per-system sums
    // (in segment order) wait on LDS rather than on a chain of global loads
    __shared__ double sp[6 * SYS_LDS_SEGS];
    const int32_t nseg = sys_seg[nsys];
    const bool staged = nseg <= SYS_LDS_SEGS;
    if (staged)
    {
        for (int32_t t = threadIdx.x; t < 6 * nseg; t += blockDim.x)
            sp[t] = seg_part[t];
        __syncthreads();
    }
    const double *part = staged ? sp : seg_part;
    // one lane per (system, energy): each sum in segment order
    for (int32_t sk = threadIdx.x; sk < 6 * nsys; sk += blockDim.x)
    {
        const int32_t s = sk / 6, k = sk % 6;
        double t = 0.0;
        for (int32_t g = sys_seg[s]; g < sys_seg[s + 1]; ++g)
            t += part[6 * g + k];
        // the probe (a single network only): its pair energies and its
        // kinetic energy (SpringNetwork.cpp:163-166,180)
        if (probe && s == 0)
            t += k == 2 ? probe->e[0] : k == 1 ? probe->e[1] : k == 3 ? probe->e[2] : 0.0;
        out_sys[6 * s + k] = t;
    }
    __syncthreads();
    if (threadIdx.x < 6)
    {
        double t = 0.0;
        for (int32_t s = 0; s < nsys; ++s)
            t += out_sys[6 * s + threadIdx.x];
        out_total[threadIdx.x] = t;
    }
}

// ---------------------------------------------------------------------------
// Adjoint of the step (no reference counterpart).  Reverse sweep of
// v' = v + dt/m (F(x) - gamma v), x' = x + dt v' with the adjoint state
// (lx, lv) of (x, v):
//   a = lv + dt lx,  w = dt/m a,  lx += H(x) w,  lv = (1 - gamma dt/m) a,
// H = dF/dx (symmetric).  One spring edge (i, j), d = x_j - x_i, r = |d|,
// F_i = k (r - L0) C d / r = phi d:
//   H w |_i = phi dw + (phi'/r) (d . dw) d,  dw = w_j - w_i,  phi' = (k C - phi) / r
//   dL/dk += -(d . dw) (r - L0) C / r,  dL/dL0 += (d . dw) k C / r
// accumulated on the edge of the row that owns the spring (a static partner,
// or else the larger index is the partner), so each spring is read-modify-
// written once per step.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(BLOCK) k_adj_init(const float4 *pos, const float4 *target, const float4 *jprm,
                                                     const float4 *iprm, int32_t n, double dt, double4 *lx,
                                                     double4 *lv, float4 *w, double *loss_part)
{
    __shared__ double sh[BLOCK / 64];
    const int32_t i = blockIdx.x * BLOCK + threadIdx.x;
    double e2 = 0.0;
    if (i < n)
    {
        const float4 x = pos[i], y = target[i];
        const double ex = (double)x.x - (double)y.x, ey = (double)x.y - (double)y.y, ez = (double)x.z - (double)y.z;
        e2 = ex * ex + ey * ey + ez * ez;
        const bool dyn = !(fbits(jprm[i].w) & PF_STATIC);
        const double m = iprm[i].x;
        const double c = dyn ? 2.0 / (double)n : 0.0;
        const double4 l = make_double4(c * ex, c * ey, c * ez, 0.0);
        const double s = (dyn && m > 0.0) ? dt / m * dt : 0.0;
        lx[i] = l;
        lv[i] = make_double4(0.0, 0.0, 0.0, 0.0);
        w[i] = make_float4((float)(s * l.x), (float)(s * l.y), (float)(s * l.z), 0.0f);
    }
    const double t = block_sum(e2, sh);
    if (threadIdx.x == 0)
        loss_part[blockIdx.x] = t;
}

// A caller's seed dL/dx_T, dL/dv_T (3n doubles each in particle order, either
// null) into the sweep's state in internal order: lx, lv, and the first
// weights w = dt/m (lv + dt lx); static particles carry none.  Products and
// sums are rounded one by one (no contraction), as on the host.
__global__ void __launch_bounds__(BLOCK) k_adj_seed(const double *dx3, const double *dv3, const int32_t *perm,
                                                     const float4 *jprm, const float4 *iprm, int32_t n, double dt,
                                                     double4 *lx, double4 *lv, float4 *w)
{
#pragma clang fp contract(off)
    const int32_t t = blockIdx.x * BLOCK + threadIdx.x;
    if (t >= n)
        return;
    const int64_t i = perm[t];
    const bool dyn = !(fbits(jprm[t].w) & PF_STATIC);
    double x[3] = {0.0, 0.0, 0.0}, v[3] = {0.0, 0.0, 0.0};
    for (int c = 0; c < 3; ++c)
    {
        if (dyn && dx3)
            x[c] = dx3[3 * i + c];
        if (dyn && dv3)
            v[c] = dv3[3 * i + c];
    }
    lx[t] = make_double4(x[0], x[1], x[2], 0.0);
    lv[t] = make_double4(v[0], v[1], v[2], 0.0);
    const double m = (double)iprm[t].x, s = (dyn && m > 0.0) ? dt / m : 0.0;
    w[t] = make_float4((float)(s * (v[0] + dt * x[0])), (float)(s * (v[1] + dt * x[1])),
                       (float)(s * (v[2] + dt * x[2])), 0.0f);
}

// out[perm[t]] = in[t]: internal order back to particle order.
__global__ void k_unperm_f64(const double *in, const int32_t *perm, int32_t n, double *out)
{
    const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n)
        out[perm[t]] = in[t];
}

// Flush: the FP32 per-slot sums of the last steps are added to the FP64
// totals and cleared.
__global__ void k_spring_grads(float *dke, float *dle, int64_t ns, double *dk, double *dl0)
{
    const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (r >= ns)
        return;
    dk[r] += (double)dke[r];
    dl0[r] += (double)dle[r];
    dke[r] = 0.0f;
    dle[r] = 0.0f;
}

__global__ void k_spring_out(const int32_t *slot_of, int64_t m, const double *dk, const double *dl0, double *outk,
                             double *outl)
{
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= m)
        return;
    const int32_t r = slot_of[s];
    outk[s] = r >= 0 ? dk[r] : 0.0;
    outl[s] = r >= 0 ? dl0[r] : 0.0;
}

__global__ void k_drift_check(const float4 *pos, const float4 *xref, int32_t n, float half_skin2, int32_t *status)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool far = false;
    if (i < n)
    {
        const float4 p = pos[i], r = xref[i];
        const float ex = p.x - r.x, ey = p.y - r.y, ez = p.z - r.z;
        far = !(ex * ex + ey * ey + ez * ez <= half_skin2);
    }
    if (__ballot(far) != 0ull && (threadIdx.x & 63) == 0)
        atomicOr(&status[7], 1);
}

__global__ void k_differs_f32(const float *v, int64_t m, int32_t *flag)
{
    const uint32_t v0 = fbits(v[0]);
    bool d = false;
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < m; s += (int64_t)gridDim.x * blockDim.x)
        d |= fbits(v[s]) != v0;
    if (__ballot(d) != 0ull && (threadIdx.x & 63) == 0)
        atomicOr(flag, 1);
}

// Any of nstates stored states (states[k * n + i]) more than half the list
// skin from the lists' reference positions: flag = 1.
__global__ void k_drift_states(const float4 *states, int32_t nstates, int32_t n, const float4 *xref,
                               float half_skin2, int32_t *flag)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool far = false;
    if (i < n)
    {
        const float4 r = xref[i];
        for (int32_t k = blockIdx.y; k < nstates; k += gridDim.y)
        {
            const float4 p = states[(size_t)k * n + i];
            const float ex = p.x - r.x, ey = p.y - r.y, ez = p.z - r.z;
            far |= !(ex * ex + ey * ey + ez * ez <= half_skin2);
        }
    }
    if (__ballot(far) != 0ull && (threadIdx.x & 63) == 0)
        atomicOr(flag, 1);
}

template <int G>
__global__ void __launch_bounds__(BLOCK) k_adj_spring(AdjSpringArgs a)
{
    const int32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int32_t row = tid / G;
    const int32_t lane = tid % G;
    const bool active = row < a.n;
    double hx = 0.0, hy = 0.0, hz = 0.0;
    float4 pi = make_float4(0.f, 0.f, 0.f, 0.f);
    uint32_t fli = PF_STATIC;
    double4 wi = make_double4(0.0, 0.0, 0.0, 0.0); // own weights, widened
    if (active)
    {
        pi = a.pos[row];
        fli = fbits(a.jprm[row].w);
        const float4 wf = a.w[row];
        wi = make_double4(wf.x, wf.y, wf.z, 0.0);
    }
    const bool dyn = active && !(fli & PF_STATIC);
    if (dyn && a.spring_enable)
    {
        const double C = a.k.c_spring * (double)a.spring_scale;
        const int32_t beg = a.sp_ptr[row], end = a.sp_ptr[row + 1];
        // the row's next slot: owned edges take consecutive slots in edge
        // order, counted over the row's G lanes with a ballot per batch
        int32_t next_slot = a.dk ? a.slot_ptr[row] : 0;
        const int32_t gbase = (int32_t)(threadIdx.x & 63u) - lane; // first lane of the row in the wave
        const uint64_t gmask = ((1ull << G) - 1ull) << gbase; // G <= 32
        constexpr int U = 4; // edges in flight per lane
        // the edge words of the next iteration load while this one gathers
        // and computes
        int32_t jr[U];
        float kk[U], ll[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
        {
            const int32_t e = beg + lane + u * G;
            const bool ok = e < end;
            jr[u] = ok ? a.sp_j[e] : row;
            kk[u] = ok ? (a.sp_k ? a.sp_k[e] : a.k_uni) : 0.0f;
            ll[u] = ok ? a.sp_l0[e] : 0.0f;
        }
        for (int32_t e0 = beg + lane; e0 < end; e0 += U * G)
        {
            int32_t jn[U];
            float kn[U], ln[U];
            float4 pj[U], wj[U];
#pragma unroll
            for (int u = 0; u < U; ++u)
            {
                pj[u] = a.pos[jr[u] & 0x7fffffff];
                wj[u] = a.w[jr[u] & 0x7fffffff];
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
            {
                const int32_t e = e0 + (U + u) * G;
                const bool ok = e < end;
                jn[u] = ok ? a.sp_j[e] : row;
                kn[u] = ok ? (a.sp_k ? a.sp_k[e] : a.k_uni) : 0.0f;
                ln[u] = ok ? a.sp_l0[e] : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
            {
                const int32_t e = e0 + u * G;
                const int32_t j = jr[u] & 0x7fffffff;
                const bool owned = a.dk && e < end && (jr[u] < 0 || j > row);
                const uint64_t ob = __ballot(owned) & gmask;
                const int32_t slot = next_slot + (int32_t)__popcll(ob & ((1ull << (threadIdx.x & 63u)) - 1ull));
                next_slot += (int32_t)__popcll(ob);
                const double dx = (double)pj[u].x - (double)pi.x, dy = (double)pj[u].y - (double)pi.y,
                             dz = (double)pj[u].z - (double)pi.z;
                const double r2 = dx * dx + dy * dy + dz * dz;
                if (!(e < end && r2 > 0.0))
                    continue;
                const double k = kk[u], l0 = ll[u];
                const double ir = rsqrt_d(r2), r = r2 * ir;
                const double phi = k * (r - l0) * C * ir, dphi = (k * C - phi) * ir;
                const double wx = (double)wj[u].x - wi.x, wy = (double)wj[u].y - wi.y, wz = (double)wj[u].z - wi.z;
                const double dd = dx * wx + dy * wy + dz * wz;
                const double g = dphi * ir * dd;
                hx = fma(phi, wx, fma(g, dx, hx));
                hy = fma(phi, wy, fma(g, dy, hy));
                hz = fma(phi, wz, fma(g, dz, hz));
                if (owned)
                {
                    const double wu = dd * ir * C;
                    a.dk[slot] -= (float)(wu * (r - l0));
                    a.dl0[slot] += (float)(wu * k);
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
            {
                jr[u] = jn[u];
                kk[u] = kn[u];
                ll[u] = ln[u];
            }
        }
    }
    for (int off = G / 2; off > 0; off >>= 1)
    {
        hx += __shfl_xor(hx, off, G);
        hy += __shfl_xor(hy, off, G);
        hz += __shfl_xor(hz, off, G);
    }
    if (!active || lane != 0)
        return;
    if (!dyn)
    {
        a.w_next[row] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    if (a.hwp)
    {
        const double4 hp = a.hwp[row];
        hx += hp.x;
        hy += hp.y;
        hz += hp.z;
    }
    const double4 lx = a.lx[row], lv = a.lv[row];
    const double m = a.iprm[row].x;
    const double ax = fma(a.dt, lx.x, lv.x), ay = fma(a.dt, lx.y, lv.y), az = fma(a.dt, lx.z, lv.z);
    const double4 nx = make_double4(lx.x + hx, lx.y + hy, lx.z + hz, 0.0);
    const double damp = m > 0.0 ? 1.0 - a.gamma * a.dt / m : 1.0;
    const double4 nv = make_double4(ax * damp, ay * damp, az * damp, 0.0);
    const double s = m > 0.0 ? a.dt / m : 0.0;
    a.lx[row] = nx;
    a.lv[row] = nv;
    a.w_next[row] = make_float4((float)(s * fma(a.dt, nx.x, nv.x)), (float)(s * fma(a.dt, nx.y, nv.y)),
                                (float)(s * fma(a.dt, nx.z, nv.z)), 0.0f);
}

// ---------------------------------------------------------------------------
// Springs from a cutoff (Topology::add_springs_from_cutoff,
// src/topology/Topology.hpp:105-138): one spring for each pair i < j closer
// than the cutoff, unless both particles are static, with rest length the
// FP64 distance of the float positions (Spring.hpp:47-56) rounded to float.
// The distance and its test are the reference's (measure::distance, unfused,
// against the float cutoff: nsearch.hpp:230).  One wave per particle i
// scans the 27 cells around it on a grid of edge >= cutoff; the count pass
// sizes the rows, the fill pass writes each row sorted by j, so that the
// springs come out ordered by (i, j).  pos.w carries the static flag.
// Pairs linked by a spring loaded before are skipped, as the reference's
// add_spring throws for them (SpringCollection.hpp:139-160).  With a group
// per particle, only pairs of different groups are linked
// (Topology::add_springs_between_topologies_from_cutoff, Topology.hpp:140-180).
// ---------------------------------------------------------------------------
// Sorts a row (second indices j, all distinct, with their rest lengths) by
// j: insertion sort for short rows, heapsort beyond, so that a crowded row
// (many particles within the cutoff of one another) stays O(r log r).
__device__ void topo_sort_row(int32_t *j, float *l, int32_t r)
{
    if (r <= 32)
    {
        for (int32_t u = 1; u < r; ++u)
        {
            const int32_t jv = j[u];
            const float lv = l[u];
            int32_t v = u - 1;
            while (v >= 0 && j[v] > jv)
            {
                j[v + 1] = j[v];
                l[v + 1] = l[v];
                --v;
            }
            j[v + 1] = jv;
            l[v + 1] = lv;
        }
        return;
    }
    auto sift = [&](int32_t root, int32_t end) {
        const int32_t jv = j[root];
        const float lv = l[root];
        for (;;)
        {
            int32_t c = 2 * root + 1;
            if (c >= end)
                break;
            if (c + 1 < end && j[c + 1] > j[c])
                ++c;
            if (jv >= j[c])
                break;
            j[root] = j[c];
            l[root] = l[c];
            root = c;
        }
        j[root] = jv;
        l[root] = lv;
    };
    for (int32_t t = r / 2 - 1; t >= 0; --t)
        sift(t, r);
    for (int32_t e = r - 1; e > 0; --e)
    {
        const int32_t jt = j[e];
        const float lt = l[e];
        j[e] = j[0];
        l[e] = l[0];
        j[0] = jt;
        l[0] = lt;
        sift(0, e);
    }
}

// Whether i < j are already linked by a loaded spring (binary search in the
// sorted partners of row i).
__device__ inline bool topo_linked(const int32_t *off, const int32_t *js, int32_t i, int32_t j)
{
    int32_t lo = off[i], hi = off[i + 1];
    while (lo < hi)
    {
        const int32_t mid = (lo + hi) >> 1;
        if (js[mid] < j)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo < off[i + 1] && js[lo] == j;
}

// One wave per particle i.  Its candidates are the members of the 3 x 3
// rows of cells around i (each row of 3 cells is one contiguous run of the
// sorted members); the runs are laid end to end so that each pass of the
// wave fills all 64 lanes.  The pair test is on the squared distance: r2 <
// r2max, the smallest double whose correctly rounded square root reaches the
// cutoff, is the same test as sqrt(r2) < cutoff, and the square root is
// taken only for hits (their rest length).  A ballot compacts the hits; the
// fill pass ranks the row by j in LDS (rank = number of hits with a smaller
// j; the js of a row are distinct) and writes it sorted.  Rows longer than
// TOPO_CAP are written as found and sorted by k_topo_sort_long.
constexpr int32_t TOPO_WAVES = 4;
constexpr int32_t TOPO_CAP = 512;

__device__ inline double topo_r2(float4 pi, float4 pj)
{
#pragma clang fp contract(off)
    const double ax = (double)pj.x - (double)pi.x;
    const double ay = (double)pj.y - (double)pi.y;
    const double az = (double)pj.z - (double)pi.z;
    return ax * ax + ay * ay + az * az;
}

template <bool FILL>
__global__ __launch_bounds__(64 * TOPO_WAVES) void k_topo_rows(TopoArgs a)
{
    __shared__ int32_t sj[FILL ? TOPO_WAVES : 1][FILL ? TOPO_CAP : 1];
    __shared__ float sl[FILL ? TOPO_WAVES : 1][FILL ? TOPO_CAP : 1];
    const int32_t lane = (int32_t)(threadIdx.x & 63u), w = (int32_t)(threadIdx.x >> 6);
    const int32_t i = (int32_t)blockIdx.x * TOPO_WAVES + w;
    if (i >= a.n)
        return; // the whole wave: no block-wide barrier below
    const float4 pi = a.pos[i];
    const bool si = fbits(pi.w) != 0u;
    const int32_t gi = a.group ? a.group[i] : 0;
    int32_t c[3];
    const float xs[3] = {pi.x, pi.y, pi.z};
    for (int q = 0; q < 3; ++q)
        c[q] = internal_coord(xs[q], a.box[q], a.box[3 + q] - a.box[q], a.dims[q], a.edge);
    int32_t base = 0, len = 0;
    if (FILL)
    {
        base = a.offs[i];
        len = a.offs[i + 1] - base;
    }
    const bool in_lds = len <= TOPO_CAP;
    const uint64_t below = (1ull << lane) - 1ull;
    // the 9 runs: start in the member list and end in the concatenation
    const int32_t x0 = max(c[0] - 1, 0), x1 = min(c[0] + 1, a.dims[0] - 1);
    int32_t rs[9], re[9];
    int32_t tot = 0;
#pragma unroll
    for (int q = 0; q < 9; ++q)
    {
        const int32_t y = c[1] - 1 + q % 3, z = c[2] - 1 + q / 3;
        int32_t b = 0, e = 0;
        if (y >= 0 && y < a.dims[1] && z >= 0 && z < a.dims[2])
        {
            const int32_t row = y * a.dims[0] + z * a.dims[0] * a.dims[1];
            b = a.cstart[row + x0];
            e = a.cstart[row + x1 + 1];
        }
        rs[q] = b - tot; // member index = rs[q] + position in the concatenation
        tot += e - b;
        re[q] = tot;
    }
    int32_t cnt = 0;
    for (int32_t t0 = 0; t0 < tot; t0 += 64)
    {
        const int32_t t = t0 + lane;
        bool hit = false;
        int32_t j = 0;
        double r2 = 0.0;
        if (t < tot)
        {
            int32_t m = rs[8] + t;
#pragma unroll
            for (int q = 7; q >= 0; --q)
                m = t < re[q] ? rs[q] + t : m;
            j = a.cidx[m];
            if (j > i && !(a.group && a.group[j] == gi)) // same topology: skipped when merging
            {
                const float4 pj = a.pos[j];
                if (!(si && fbits(pj.w) != 0u))
                {
                    r2 = topo_r2(pi, pj);
                    // SpringAlreadyExistsException: the loaded spring stays
                    hit = r2 < a.r2max && !(a.ex_off && topo_linked(a.ex_off, a.ex_j, i, j));
                }
            }
        }
        const uint64_t bal = __ballot(hit);
        if (FILL && hit)
        {
            const int32_t slot = cnt + __popcll(bal & below);
            const float l0 = (float)sqrt(r2);
            if (slot < len)
            {
                if (in_lds)
                {
                    sj[w][slot] = j;
                    sl[w][slot] = l0;
                }
                else
                {
                    a.jout[base + slot] = j;
                    a.l0out[base + slot] = l0;
                }
            }
        }
        cnt += __popcll(bal);
    }
    if (!FILL)
    {
        if (lane == 0)
            a.counts[i] = cnt;
        return;
    }
    if (!in_lds)
        return;
    wave_lds_sync();
    const int32_t m = min(cnt, len);
    for (int32_t e = lane; e < m; e += 64)
    {
        const int32_t jv = sj[w][e];
        int32_t rank = 0;
        for (int32_t k = 0; k < m; ++k)
            rank += sj[w][k] < jv ? 1 : 0;
        a.jout[base + rank] = jv;
        a.l0out[base + rank] = sl[w][e];
    }
}

// Trajectory frames: positions back in particle order as float triples
// (x[perm[t]] = pos[t]); with compensated positions, hi + lo rounded to float.
__global__ void k_pack_positions(const float4 *pos, const float4 *plo, const int32_t *perm, int32_t n, float *out3)
{
    const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n)
        return;
    float4 p = pos[t];
    if (plo)
    {
        const float4 l = plo[t];
        p.x = (float)((double)p.x + (double)l.x);
        p.y = (float)((double)p.y + (double)l.y);
        p.z = (float)((double)p.z + (double)l.z);
    }
    const int64_t i = perm[t];
    out3[3 * i] = p.x;
    out3[3 * i + 1] = p.y;
    out3[3 * i + 2] = p.z;
}

__global__ void k_topo_sort_long(const int32_t *offs, int32_t n, int32_t *jout, float *l0out)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int32_t b = offs[i], len = offs[i + 1] - b;
    if (len > TOPO_CAP)
        topo_sort_row(jout + b, l0out + b, len);
}

// The rebuild gate of the kernels that follow: status[0] = a drift flag
// (status[7], raised by the integrator or by the host) or, with advance, a
// rebuild interval that falls due; the drift flag is consumed and the
// rebuild counted (counters[1]).  advance: the start of a step, which also
// stops the call after a non-finite position and counts the step
// (counters[0]).
__device__ void prestep_gate(int32_t *status, int32_t *counters, int32_t interval, int32_t advance, int32_t *log)
{
    if (halted(status))
        return;
    if (advance && status[1] != 0x7fffffff)
    { // the step before produced a non-finite position: stop here
        status[5] = 1;
        return;
    }
    const int32_t step = counters[0];
    const bool due = advance && interval > 0 && step > 0 && (step % interval) == 0;
    const int32_t r = (status[7] != 0 || due) ? 1 : 0;
    status[7] = 0;
    status[0] = r;
    counters[1] += r;
    if (advance)
        counters[0] = step + 1;
    if (log)
        *log = r;
}

// The gate (first lane) and, with partial, the box partials of the current
// positions, whether or not a rebuild follows: reading the positions costs
// about what a second launch would.
__global__ void __launch_bounds__(BLOCK) k_prestep(int32_t *status, int32_t *counters, int32_t interval,
                                                   int32_t advance, const float4 *pos, const float4 *shift,
                                                   int32_t n, double *partial, int32_t *log)
{
    if (blockIdx.x == 0 && threadIdx.x == 0)
        prestep_gate(status, counters, interval, advance, log);
    if (partial)
        box_partial_block(pos, shift, n, partial);
}

__global__ void k_fill_i32(int32_t *p, int32_t v, int64_t n, const int32_t *gate)
{
    if (gated_off(gate))
        return;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// ---------------------------------------------------------------------------
// Spring CSR built on the device from the spring table (external ids): both
// directions, each row sorted by (partner, spring, side); the exclusions of
// each i-cluster merged from its rows; springs between two rigid members
// left out of the force rows.
// ---------------------------------------------------------------------------
__global__ void k_edge_count(const uint32_t *sa, const uint32_t *sb, int64_t m, const int32_t *iperm, int32_t *cnt)
{
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= m)
        return;
    atomicAdd(&cnt[iperm[sa[s]]], 1);
    atomicAdd(&cnt[iperm[sb[s]]], 1);
}

// key = partner << 32 | (spring << 1 | side), placed at an arbitrary slot of
// the row (k_edge_sort orders it)
__global__ void k_edge_fill(const uint32_t *sa, const uint32_t *sb, int64_t m, const int32_t *iperm,
                            const int32_t *ptr, int32_t *cursor, uint64_t *key)
{
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= m)
        return;
    const int32_t a = iperm[sa[s]], b = iperm[sb[s]];
    const uint32_t w = (uint32_t)s << 1;
    key[ptr[a] + atomicAdd(&cursor[a], 1)] = (uint64_t)(uint32_t)b << 32 | w;
    key[ptr[b] + atomicAdd(&cursor[b], 1)] = (uint64_t)(uint32_t)a << 32 | (w | 1u);
}

constexpr int32_t EDGE_WAVES = BLOCK / 64;
constexpr int32_t EDGE_LDS = 1024; // keys per wave held in LDS; longer rows rank from global memory

// One wave per row: each key's rank is the number of smaller keys in the row
// (keys are distinct).  Writes the partner (bit 31: static partner), its
// stiffness and rest length at the rank.
__global__ void __launch_bounds__(BLOCK) k_edge_sort(const int32_t *ptr, int32_t n, const uint64_t *key,
                                                     const float4 *jprm, const float *sl, const float *sk, int32_t *ej,
                                                     float *ek, float *el)
{
    __shared__ uint64_t lk[EDGE_WAVES][EDGE_LDS];
    const int32_t lane = (int32_t)(threadIdx.x & 63u), w = (int32_t)(threadIdx.x >> 6);
    const int32_t row = (int32_t)blockIdx.x * EDGE_WAVES + w;
    if (row >= n)
        return; // whole wave; no block barrier below
    const int32_t b = ptr[row], len = ptr[row + 1] - b;
    const uint64_t *src = key + b;
    if (len <= EDGE_LDS)
    {
        for (int32_t t = lane; t < len; t += 64)
            lk[w][t] = key[b + t];
        wave_lds_sync();
        src = lk[w];
    }
    for (int32_t e = lane; e < len; e += 64)
    {
        const uint64_t kv = src[e];
        int32_t rank = 0;
        for (int32_t t = 0; t < len; ++t)
            rank += src[t] < kv ? 1 : 0;
        const int32_t j = (int32_t)(kv >> 32);
        const uint32_t s = (uint32_t)kv >> 1;
        ej[b + rank] = (fbits(jprm[j].w) & PF_STATIC) ? (int32_t)((uint32_t)j | 0x80000000u) : j;
        ek[b + rank] = sk[s];
        el[b + rank] = sl[s];
    }
}

__device__ inline int32_t row_lower_bound(const int32_t *ej, int32_t lo, int32_t hi, int32_t j, bool upper)
{
    while (lo < hi)
    {
        const int32_t mid = (lo + hi) >> 1;
        const int32_t v = ej[mid] & 0x7fffffff;
        if (v < j || (upper && v == j))
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// One wave per i-cluster: the words j << 8 | 1 << k of its rows k = 0..7
// merged in j order (stable in k) into merged[ptr[8c] ..), then the number
// of distinct j.
__global__ void k_ex_merge(const int32_t *ptr, const int32_t *ej, int32_t n, int32_t ncl, uint32_t *merged,
                           int32_t *ucount)
{
    const int32_t lane = (int32_t)(threadIdx.x & 63u);
    const int32_t c = (int32_t)((blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6);
    if (c >= ncl)
        return;
    int32_t rb[CLUSTER + 1];
#pragma unroll
    for (int k = 0; k <= CLUSTER; ++k)
        rb[k] = ptr[min(c * CLUSTER + k, n)];
    const int32_t base = rb[0], end = rb[CLUSTER];
    for (int32_t e = base + lane; e < end; e += 64)
    {
        int32_t k = 0;
#pragma unroll
        for (int q = 1; q < CLUSTER; ++q)
            k = e >= rb[q] ? q : k;
        const int32_t j = ej[e] & 0x7fffffff;
        // rank: its place in its own row, plus the smaller words of the
        // other rows (for rows before its own, the equal ones too)
        int32_t r = e - rb[k];
#pragma unroll
        for (int q = 0; q < CLUSTER; ++q)
            if (q != k)
                r += row_lower_bound(ej, rb[q], rb[q + 1], j, q < k) - rb[q];
        merged[base + r] = (uint32_t)j << 8 | 1u << k;
    }
    // the wave reads back what its own lanes stored (one CU)
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    int32_t cnt = 0;
    for (int32_t t0 = base; t0 < end; t0 += 64)
    {
        const int32_t t = t0 + lane;
        const bool first = t < end && (t == base || (merged[t - 1] >> 8) != (merged[t] >> 8));
        cnt += (int32_t)__popcll(__ballot(first));
    }
    if (lane == 0)
        ucount[c] = cnt;
}

// The distinct words of each cluster, masks of equal j or-ed, at ex_ptr[c].
__global__ void k_ex_fill(const int32_t *ptr, int32_t n, int32_t ncl, const uint32_t *merged, const int32_t *ex_ptr,
                          uint32_t *ex_w)
{
    const int32_t lane = (int32_t)(threadIdx.x & 63u);
    const int32_t c = (int32_t)((blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6);
    if (c >= ncl)
        return;
    const int32_t base = ptr[min(c * CLUSTER, n)], end = ptr[min(c * CLUSTER + CLUSTER, n)];
    int32_t out = ex_ptr[c];
    for (int32_t t0 = base; t0 < end; t0 += 64)
    {
        const int32_t t = t0 + lane;
        const bool first = t < end && (t == base || (merged[t - 1] >> 8) != (merged[t] >> 8));
        const uint64_t bal = __ballot(first);
        if (first)
        {
            uint32_t wv = merged[t];
            for (int32_t u = t + 1; u < end && (merged[u] >> 8) == (wv >> 8); ++u)
                wv |= merged[u] & 0xFFu;
            ex_w[out + (int32_t)__popcll(bal & ((1ull << lane) - 1ull))] = wv;
        }
        out += (int32_t)__popcll(bal);
    }
}

// Force rows: a row's edges but those joining two rigid-body members.
__global__ void k_edge_keep(const int32_t *ptr, int32_t n, const int32_t *ej, const float *ek, const float *el,
                            const uint8_t *rigid, const int32_t *fptr, int32_t *cnt, int32_t *fej, float *fek,
                            float *fel)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const bool ri = rigid[i] != 0;
    int32_t o = fptr ? fptr[i] : 0, c = 0;
    for (int32_t t = ptr[i]; t < ptr[i + 1]; ++t)
    {
        if (ri && rigid[ej[t] & 0x7fffffff])
            continue;
        if (fptr)
        {
            fej[o] = ej[t];
            fek[o] = ek[t];
            fel[o] = el[t];
            ++o;
        }
        ++c;
    }
    if (!fptr)
        cnt[i] = c;
}

// Row i of the constructed springs at offs[i]: id1 = i.
__global__ void k_rows_expand(const int32_t *offs, int32_t n, uint32_t *ida)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    for (int32_t t = offs[i]; t < offs[i + 1]; ++t)
        ida[t] = (uint32_t)i;
}

inline int32_t nblk(int64_t n, int32_t b = BLOCK) { return (int32_t)((n + b - 1) / b); }

} // namespace

// ---- launchers ----

void launch_box_reduce(const float4 *pos, const float4 *shift, int32_t n, double *partial, double *box,
                       int32_t nblocks, const int32_t *gate, hipStream_t s)
{
    k_box_partial<<<nblocks, BLOCK, 0, s>>>(pos, shift, n, partial, gate);
    k_box_final<<<1, BLOCK, 0, s>>>(partial, nblocks, box, gate);
}

void launch_cell_ids(const float4 *pos, const float4 *shift, const float4 *jprm, uint32_t mask, int32_t n,
                     double *box, const double *box_part, int32_t box_nb, float4 *xref, int32_t d0, int32_t d1,
                     int32_t d2, double edge_exact, float min_edge, int32_t exact_mode, int32_t *cell_of,
                     int32_t *counts, const int32_t *gate, hipStream_t s)
{
    if (n > 0)
        k_cell_ids<<<nblk(n), BLOCK, 0, s>>>(pos, shift, jprm, mask, n, box, box_part, box_nb, xref, d0, d1, d2,
                                              edge_exact, min_edge, exact_mode, cell_of, counts, gate);
}

void launch_scan(const int32_t *counts, int32_t *start, int32_t ncells, const int32_t *gate, hipStream_t s)
{
    k_scan<<<1, 1024, 0, s>>>(counts, start, ncells, gate);
}

void launch_scatter(const int32_t *cell_of, int32_t n, const int32_t *start, int32_t *cursor, int32_t drain,
                    int32_t *tmp, const int32_t *gate, hipStream_t s)
{
    if (n > 0)
        k_scatter<<<nblk(n), BLOCK, 0, s>>>(cell_of, n, start, cursor, drain, tmp, gate);
}

void launch_cell_sort(const int32_t *cell_of, int32_t n, const int32_t *start, const int32_t *tmp,
                      const int32_t *key, int32_t *out, const int32_t *gate, hipStream_t s)
{
    if (n > 0)
        k_cell_sort<<<nblk(n), BLOCK, 0, s>>>(cell_of, n, start, tmp, key, out, gate);
}

template <int G> static void spring_g(const SpringArgs &a, int32_t nb, hipStream_t s)
{
    const bool comp = a.plo != nullptr;
    if (a.integrate)
        comp ? k_spring<G, true, true><<<nb, BLOCK, 0, s>>>(a) : k_spring<G, true, false><<<nb, BLOCK, 0, s>>>(a);
    else
        comp ? k_spring<G, false, true><<<nb, BLOCK, 0, s>>>(a) : k_spring<G, false, false><<<nb, BLOCK, 0, s>>>(a);
}

void launch_spring(const SpringArgs &a, int32_t group, int32_t nb, hipStream_t s)
{
    switch (group)
    {
    case 4: spring_g<4>(a, nb, s); break;
    case 8: spring_g<8>(a, nb, s); break;
    case 32: spring_g<32>(a, nb, s); break;
    default: spring_g<16>(a, nb, s); break;
    }
}

template <int G> static void adj_spring_g(const AdjSpringArgs &a, int32_t nb, hipStream_t s)
{
    k_adj_spring<G><<<nb, BLOCK, 0, s>>>(a);
}

void launch_adj_spring(const AdjSpringArgs &a, int32_t group, int32_t nb, hipStream_t s)
{
    switch (group)
    {
    case 4: adj_spring_g<4>(a, nb, s); break;
    case 8: adj_spring_g<8>(a, nb, s); break;
    case 32: adj_spring_g<32>(a, nb, s); break;
    default: adj_spring_g<16>(a, nb, s); break;
    }
}

void launch_adj_init(const float4 *pos, const float4 *target, const float4 *jprm, const float4 *iprm, int32_t n,
                     double dt, double4 *lx, double4 *lv, float4 *w, double *loss_part, int32_t nblocks,
                     hipStream_t s)
{
    k_adj_init<<<nblocks, BLOCK, 0, s>>>(pos, target, jprm, iprm, n, dt, lx, lv, w, loss_part);
}

void launch_adj_seed(const double *dx3, const double *dv3, const int32_t *perm, const float4 *jprm,
                     const float4 *iprm, int32_t n, double dt, double4 *lx, double4 *lv, float4 *w, hipStream_t s)
{
    if (n > 0)
        k_adj_seed<<<nblk(n), BLOCK, 0, s>>>(dx3, dv3, perm, jprm, iprm, n, dt, lx, lv, w);
}

void launch_unperm_f64(const double *in, const int32_t *perm, int32_t n, double *out, hipStream_t s)
{
    if (n > 0)
        k_unperm_f64<<<nblk(n), BLOCK, 0, s>>>(in, perm, n, out);
}

void launch_spring_grads(float *dke, float *dle, int64_t ns, double *dk, double *dl0, hipStream_t s)
{
    if (ns > 0)
        k_spring_grads<<<nblk(ns), BLOCK, 0, s>>>(dke, dle, ns, dk, dl0);
}

void launch_spring_out(const int32_t *slot_of, int64_t m, const double *dk, const double *dl0, double *outk,
                       double *outl, hipStream_t s)
{
    if (m > 0)
        k_spring_out<<<nblk(m), BLOCK, 0, s>>>(slot_of, m, dk, dl0, outk, outl);
}

void launch_topo_springs(const TopoArgs &a, hipStream_t s)
{
    if (a.n <= 0)
        return;
    const int32_t blocks = (a.n + TOPO_WAVES - 1) / TOPO_WAVES;
    if (!a.offs)
    {
        k_topo_rows<false><<<blocks, 64 * TOPO_WAVES, 0, s>>>(a);
        return;
    }
    k_topo_rows<true><<<blocks, 64 * TOPO_WAVES, 0, s>>>(a);
    if (a.max_row > TOPO_CAP)
        k_topo_sort_long<<<nblk(a.n), BLOCK, 0, s>>>(a.offs, a.n, a.jout, a.l0out);
}

void launch_pack_positions(const float4 *pos, const float4 *plo, const int32_t *perm, int32_t n, float *out3,
                           hipStream_t s)
{
    if (n > 0)
        k_pack_positions<<<nblk(n), BLOCK, 0, s>>>(pos, plo, perm, n, out3);
}

void launch_differs_f32(const float *v, int64_t m, int32_t *flag, hipStream_t s)
{
    if (m > 0)
    {
        const int64_t b = (m + BLOCK - 1) / BLOCK;
        k_differs_f32<<<(int32_t)(b < 2048 ? b : 2048), BLOCK, 0, s>>>(v, m, flag);
    }
}

void launch_drift_states(const float4 *states, int32_t nstates, int32_t n, const float4 *xref, float half_skin2,
                         int32_t *flag, hipStream_t s)
{
    if (n > 0 && nstates > 0)
        k_drift_states<<<dim3(nblk(n), nstates < 64 ? nstates : 64), BLOCK, 0, s>>>(states, nstates, n, xref,
                                                                                     half_skin2, flag);
}

void launch_drift_check(const float4 *pos, const float4 *xref, int32_t n, float half_skin2, int32_t *status,
                        hipStream_t s)
{
    k_drift_check<<<nblk(n), BLOCK, 0, s>>>(pos, xref, n, half_skin2, status);
}

void launch_energies(const double4 *epp, const double4 *eps, const int32_t *seg_start, int32_t nseg,
                     const int32_t *sys_seg, int32_t nsys, double *seg_part, double *out_sys, double *out_total,
                     const ProbeDev *probe, hipStream_t s)
{
    if (nseg > 0)
        k_seg_energies<<<nseg, BLOCK, 0, s>>>(epp, eps, seg_start, seg_part);
    k_sys_energies<<<1, BLOCK, 0, s>>>(seg_part, sys_seg, nsys, out_sys, out_total, probe);
}

void launch_probe(const ProbeArgs &p, int32_t nparts, float dt, int32_t visc_enable, float visc, double c_kinetic,
                  int32_t integrate, hipStream_t s)
{
    k_probe<<<1, BLOCK, 0, s>>>(p, nparts, dt, visc_enable, visc, c_kinetic, integrate);
}

void launch_prestep(int32_t *status, int32_t *counters, int32_t interval, int32_t advance, const float4 *pos,
                    const float4 *shift, int32_t n, double *partial, int32_t nblocks, int32_t *log, hipStream_t s)
{
    if (partial && n > 0)
        k_prestep<<<nblocks, BLOCK, 0, s>>>(status, counters, interval, advance, pos, shift, n, partial, log);
    else
        k_prestep<<<1, BLOCK, 0, s>>>(status, counters, interval, advance, nullptr, nullptr, 0, nullptr, log);
}

void launch_rigid(const RigidArgs &a, hipStream_t s)
{
    if (a.nb <= 0)
        return;
    if (a.solve)
        k_rigid_solve<<<a.nb, RIGID_BLOCK, 0, s>>>(a);
    if (a.nr > 0)
        k_rigid_place<<<nblk(a.nr), BLOCK, 0, s>>>(a);
}

void launch_centroid(const float4 *pos, int32_t n, double *partial, int32_t nblocks, double *out3, hipStream_t s)
{
    k_centroid_partial<<<nblocks, BLOCK, 0, s>>>(pos, n, partial);
    k_centroid_final<<<1, BLOCK, 0, s>>>(partial, nblocks, n, out3);
}

void launch_fill_i32(int32_t *p, int32_t v, int64_t n, const int32_t *gate, hipStream_t s)
{
    if (n > 0)
    {
        const int64_t b = (n + BLOCK - 1) / BLOCK;
        k_fill_i32<<<(int32_t)(b < 4096 ? b : 4096), BLOCK, 0, s>>>(p, v, n, gate);
    }
}

void launch_edge_count(const uint32_t *sa, const uint32_t *sb, int64_t m, const int32_t *iperm, int32_t *cnt,
                       hipStream_t s)
{
    if (m > 0)
        k_edge_count<<<(int32_t)((m + BLOCK - 1) / BLOCK), BLOCK, 0, s>>>(sa, sb, m, iperm, cnt);
}

void launch_edge_fill(const uint32_t *sa, const uint32_t *sb, int64_t m, const int32_t *iperm, const int32_t *ptr,
                      int32_t *cursor, uint64_t *key, hipStream_t s)
{
    if (m > 0)
        k_edge_fill<<<(int32_t)((m + BLOCK - 1) / BLOCK), BLOCK, 0, s>>>(sa, sb, m, iperm, ptr, cursor, key);
}

void launch_edge_sort(const int32_t *ptr, int32_t n, const uint64_t *key, const float4 *jprm, const float *sl,
                      const float *sk, int32_t *ej, float *ek, float *el, hipStream_t s)
{
    if (n > 0)
        k_edge_sort<<<(n + EDGE_WAVES - 1) / EDGE_WAVES, BLOCK, 0, s>>>(ptr, n, key, jprm, sl, sk, ej, ek, el);
}

void launch_ex_merge(const int32_t *ptr, const int32_t *ej, int32_t n, int32_t ncl, uint32_t *merged,
                     int32_t *ucount, hipStream_t s)
{
    if (ncl > 0)
        k_ex_merge<<<(ncl + EDGE_WAVES - 1) / EDGE_WAVES, BLOCK, 0, s>>>(ptr, ej, n, ncl, merged, ucount);
}

void launch_ex_fill(const int32_t *ptr, int32_t n, int32_t ncl, const uint32_t *merged, const int32_t *ex_ptr,
                    uint32_t *ex_w, hipStream_t s)
{
    if (ncl > 0)
        k_ex_fill<<<(ncl + EDGE_WAVES - 1) / EDGE_WAVES, BLOCK, 0, s>>>(ptr, n, ncl, merged, ex_ptr, ex_w);
}

void launch_edge_keep(const int32_t *ptr, int32_t n, const int32_t *ej, const float *ek, const float *el,
                      const uint8_t *rigid, const int32_t *fptr, int32_t *cnt, int32_t *fej, float *fek, float *fel,
                      hipStream_t s)
{
    if (n > 0)
        k_edge_keep<<<nblk(n), BLOCK, 0, s>>>(ptr, n, ej, ek, el, rigid, fptr, cnt, fej, fek, fel);
}

void launch_rows_expand(const int32_t *offs, int32_t n, uint32_t *ida, hipStream_t s)
{
    if (n > 0)
        k_rows_expand<<<nblk(n), BLOCK, 0, s>>>(offs, n, ida);
}

} // namespace bs

