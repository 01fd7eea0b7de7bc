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
__global__ void k_box_partial(const float4 *pos, int32_t n, double *partial, const int32_t *gate)
{
    if (gated_off(gate))
        return;
    __shared__ float sh[6][BLOCK / 64];
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        const float4 p = pos[i];
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

// Combines up to BLOCK block partials (one per thread) with wave shuffles.
__global__ void __launch_bounds__(BLOCK) k_box_final(const double *partial, int32_t nb, double *box,
                                                     const int32_t *gate)
{
    if (gated_off(gate))
        return;
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
        box[t] = r;
    }
}

// ---------------------------------------------------------------------------
// Cell binning: ids, counting sort, deterministic in-cell order.
// exact_mode 1: nsearch.hpp:277-293 (size_t((x - min)/edge), clamped), FP64.
// exact_mode 0: the engine's own grid, edge >= min_edge on every axis.
// ---------------------------------------------------------------------------
__global__ void k_cell_ids(const float4 *pos, const float4 *jprm, uint32_t mask, int32_t n, const double *box,
                           int32_t d0, int32_t d1, int32_t d2, double edge_exact, float min_edge, int32_t exact_mode,
                           int32_t *cell_of, int32_t *counts, const int32_t *gate)
{
    if (gated_off(gate))
        return;
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t fl = fbits(jprm[i].w);
    if (mask != 0u && (fl & mask) == 0u)
    {
        cell_of[i] = -1;
        return;
    }
    const float4 p = pos[i];
    int32_t c[3];
    const float xs[3] = {p.x, p.y, p.z};
    const int32_t dims[3] = {d0, d1, d2};
    for (int a = 0; a < 3; ++a)
    {
        if (exact_mode)
        {
            const double q = ((double)xs[a] - box[a]) / edge_exact;
            const uint64_t u = q > 0.0 ? (uint64_t)q : 0u;
            c[a] = (int32_t)(u < (uint64_t)(dims[a] - 1) ? u : (uint64_t)(dims[a] - 1));
        }
        else
        {
            c[a] = internal_coord(xs[a], box[a], box[3 + a] - box[a], dims[a], min_edge);
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

__global__ void k_scatter(const int32_t *cell_of, int32_t n, const int32_t *start, int32_t *cursor, int32_t *tmp,
                          const int32_t *gate)
{
    if (gated_off(gate))
        return;
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int32_t c = cell_of[i];
    if (c < 0)
        return;
    tmp[start[c] + atomicAdd(&cursor[c], 1)] = i;
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
// IMPALA (imp.hpp:29-159), FP32 geometry with FP64 energy/derivative, as the
// reference; returns the energy and adds the force (ForceField.cpp:80-86).
// ---------------------------------------------------------------------------
__device__ inline double imp_eval(float x, float z, float sasa, float tr, float uo, float lo, float uc, float lc,
                                  float scale, double c_spring, float f[3])
{
    const float alpha = 1.99f, z0 = 15.75f, alip = -0.018f;
    const bool flat = uo == 0.0f && lo == 0.0f && uc == 0.0f && lc == 0.0f;
    const float ru = uc == 0.0f ? 1000000.0f : fabsf(1.0f / uc);
    const float rl = lc == 0.0f ? 1000000.0f : fabsf(1.0f / lc);
    const int su = (uc > 0.0f) - (uc < 0.0f), sl = (lc > 0.0f) - (lc < 0.0f);
    const float cu = uo - su * ru, cl = -lo - sl * rl;
    const float vu[3] = {x, 0.0f, z - cu}, vl[3] = {x, 0.0f, z - cl};
    const float nu = sqrtf(vu[0] * vu[0] + vu[2] * vu[2]);
    const float nl = sqrtf(vl[0] * vl[0] + vl[2] * vl[2]);
    const float zu = su == 0 ? z : (z > cu ? su * nu + uo - ru : -su * nu + uo - ru);
    const float zl = sl == 0 ? z : (z > cl ? sl * nl - lo - rl : -sl * nl - lo - rl);
    const double wu = (double)zu - uo, wl = (double)zl + lo;
    const double xu = exp(alpha * (fabs(wu) - z0)), xl = exp(alpha * (fabs(wl) - z0));
    const double pre = (double)(-sasa * tr) + (double)(alip * sasa);
    const double eu = pre * (0.5 - 1.0 / (1.0 + xu));
    const double el = pre * (0.5 - 1.0 / (1.0 + xl));
    double dcu = alpha * wu * xu / ((xu + 1.0) * (xu + 1.0) * fabs(wu));
    double dcl = alpha * wl * xl / ((xl + 1.0) * (xl + 1.0) * fabs(wl));
    if (!isfinite(dcu))
        dcu = 0.0;
    if (!isfinite(dcl))
        dcl = 0.0;
    float du[3] = {0.0f, 0.0f, 1.0f}, dl[3] = {0.0f, 0.0f, 1.0f};
    if (su != 0)
    {
        const float inv = nu > 0.0f ? 1.0f / nu : 0.0f;
        du[0] = vu[0] * inv;
        du[1] = 0.0f;
        du[2] = vu[2] * inv;
    }
    if (sl != 0)
    {
        const float inv = nl > 0.0f ? 1.0f / nl : 0.0f;
        dl[0] = vl[0] * inv;
        dl[1] = 0.0f;
        dl[2] = vl[2] * inv;
    }
    const float mu = (float)(pre * dcu * c_spring), ml = (float)(pre * dcl * c_spring);
    for (int c = 0; c < 3; ++c)
        f[c] -= scale * (flat ? du[c] * mu : du[c] * mu + dl[c] * ml);
    return (double)scale * (flat ? eu : eu + el);
}

// One spring edge seen from row i: force on i and the energy share.  The
// extension len - L0 is formed from an FP64 length (rsqrt seed plus one
// Newton step) so that it keeps full accuracy near rest.
template <bool COMP>
__device__ inline void spring_edge(float4 pi, float4 li, float4 pj, const float4 *plo, int32_t j, float k, float l0,
                                   bool jstatic, double cs, float scale, double &fx, double &fy, double &fz, double &e)
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
    e += (jstatic ? 1.0 : 0.5) * 0.5 * (double)scale * (double)k * ext * ext;
}

// ---------------------------------------------------------------------------
// Springs (CSR gather, G lanes per row, 4 edges in flight per lane) fused with
// viscosity, external and IMPALA forces and the viscous semi-implicit Euler
// step (SpringNetwork.cpp:63-90,202-241; Spring.cpp:20-34; Particle.cpp:80-128).
// ---------------------------------------------------------------------------
template <int G, bool INTEGRATE, bool COMP>
__global__ void __launch_bounds__(BLOCK) k_spring(SpringArgs a)
{
    __shared__ double sh[BLOCK / 64];
    const int32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int32_t row = tid / G;
    const int32_t lane = tid % G;
    double espr = 0.0, eke = 0.0, eimp = 0.0;
    double fx = 0.0, fy = 0.0, fz = 0.0;
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
        constexpr int U = 4;
        for (int32_t e0 = beg + lane; e0 < end; e0 += U * G)
        {
            int32_t jr[U];
            float kk[U], ll[U];
            float4 pj[U];
#pragma unroll
            for (int u = 0; u < U; ++u)
            {
                const int32_t e = e0 + u * G;
                const bool ok = e < end;
                jr[u] = ok ? a.sp_j[e] : 0;
                kk[u] = ok ? a.sp_k[e] : 0.0f;
                ll[u] = ok ? a.sp_l0[e] : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
                pj[u] = a.pos[jr[u] & 0x7fffffff];
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (e0 + u * G < end)
                    spring_edge<COMP>(pi, li, pj[u], a.plo, jr[u] & 0x7fffffff, kk[u], ll[u], jr[u] < 0, cs,
                                      a.spring_scale, fx, fy, fz, espr);
        }
    }
    // reduce the force over the G lanes of the row
    for (int off = G / 2; off > 0; off >>= 1)
    {
        fx += __shfl_xor(fx, off, G);
        fy += __shfl_xor(fy, off, G);
        fz += __shfl_xor(fz, off, G);
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
        // the conservative sum is rounded once; the viscous and IMPALA terms
        // follow in FP32 as in the reference (Particle.cpp:80-84, 259-267)
        float fxf = (float)sx, fyf = (float)sy, fzf = (float)sz;
        if (a.visc_enable)
        {
            fxf -= v.x * a.visc;
            fyf -= v.y * a.visc;
            fzf -= v.z * a.visc;
        }
        const float4 ip = a.iprm[row];
        if (a.imp_enable)
        {
            float f[3] = {fxf, fyf, fzf};
            eimp += imp_eval(pi.x, pi.z, ip.y, ip.z, a.imp_uo, a.imp_lo, a.imp_uc, a.imp_lc, a.imp_scale,
                             a.k.c_spring, f);
            fxf = f[0];
            fyf = f[1];
            fzf = f[2];
        }
        const float fx = fxf, fy = fyf, fz = fzf;
        if (!INTEGRATE)
        {
            a.fout[row] = make_float4(fx, fy, fz, 0.0f);
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
                atomicOr(&a.status[0], 1);
            a.vel[row] = v;
            a.pos_out[row] = xn;
        }
    }
    const double b0 = block_sum(espr, sh);
    const double b1 = block_sum(eke, sh);
    const double b2 = block_sum(eimp, sh);
    if (threadIdx.x == 0)
    {
        a.epart[3 * blockIdx.x + 0] = b0;
        a.epart[3 * blockIdx.x + 1] = b1;
        a.epart[3 * blockIdx.x + 2] = b2;
    }
}

// Sums the block partials in a fixed order.  energies: spring, elec, steric,
// kinetic, imp, hydro.  pair partials are [list][block][steric, elec, hydro].
__global__ void __launch_bounds__(BLOCK) k_reduce_energies(const double *pp, int32_t npb, int32_t nlists,
                                                           const double *spp, int32_t nsb, double *energies)
{
    __shared__ double sh[BLOCK / 64];
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int32_t l = 0; l < nlists; ++l)
        for (int32_t b = threadIdx.x; b < npb; b += blockDim.x)
        {
            acc[2] += pp[(size_t)(l * npb + b) * 3 + 0];
            acc[1] += pp[(size_t)(l * npb + b) * 3 + 1];
            acc[5] += pp[(size_t)(l * npb + b) * 3 + 2];
        }
    if (spp)
        for (int32_t b = threadIdx.x; b < nsb; b += blockDim.x)
        {
            acc[0] += spp[3 * b + 0];
            acc[3] += spp[3 * b + 1];
            acc[4] += spp[3 * b + 2];
        }
    for (int k = 0; k < 6; ++k)
    {
        const double s = block_sum(acc[k], sh);
        if (threadIdx.x == 0)
            energies[k] = s;
    }
}

__global__ void k_prestep(int32_t *status, int32_t *step, int32_t interval)
{
    if (interval > 0 && *step > 0 && (*step % interval) == 0)
        status[0] = 1;
    *step += 1;
}

__global__ void k_copy_xref(const float4 *pos, float4 *xref, int32_t n, const int32_t *status)
{
    if (status[0] == 0)
        return;
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        xref[i] = pos[i];
}

__global__ void k_clear_rebuild(int32_t *status, int32_t *counters)
{
    if (status[0] != 0)
    {
        status[0] = 0;
        counters[1] += 1;
    }
}

__global__ void k_fill_i32(int32_t *p, int32_t v, int64_t n, const int32_t *gate)
{
    if (gated_off(gate))
        return;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

inline int32_t nblk(int64_t n, int32_t b = BLOCK) { return (int32_t)((n + b - 1) / b); }

} // namespace

// ---- launchers ----

void launch_box_reduce(const float4 *pos, int32_t n, double *partial, double *box, int32_t nblocks,
                       const int32_t *gate, hipStream_t s)
{
    k_box_partial<<<nblocks, BLOCK, 0, s>>>(pos, n, partial, gate);
    k_box_final<<<1, BLOCK, 0, s>>>(partial, nblocks, box, gate);
}

void launch_cell_ids(const float4 *pos, const float4 *jprm, uint32_t mask, int32_t n, const double *box,
                     int32_t d0, int32_t d1, int32_t d2, double edge_exact, float min_edge, int32_t exact_mode,
                     int32_t *cell_of, int32_t *counts, const int32_t *gate, hipStream_t s)
{
    if (n > 0)
        k_cell_ids<<<nblk(n), BLOCK, 0, s>>>(pos, jprm, mask, n, box, d0, d1, d2, edge_exact, min_edge, exact_mode,
                                              cell_of, counts, gate);
}

void launch_scan(const int32_t *counts, int32_t *start, int32_t ncells, const int32_t *gate, hipStream_t s)
{
    k_scan<<<1, 1024, 0, s>>>(counts, start, ncells, gate);
}

void launch_scatter(const int32_t *cell_of, int32_t n, const int32_t *start, int32_t *cursor, int32_t *tmp,
                    const int32_t *gate, hipStream_t s)
{
    if (n > 0)
        k_scatter<<<nblk(n), BLOCK, 0, s>>>(cell_of, n, start, cursor, tmp, gate);
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

void launch_reduce_energies(const double *pp, int32_t npb, int32_t nlists, const double *spp, int32_t nsb,
                            double *energies, hipStream_t s)
{
    k_reduce_energies<<<1, BLOCK, 0, s>>>(pp, npb, nlists, spp, nsb, energies);
}

void launch_prestep(int32_t *status, int32_t *step, int32_t interval, hipStream_t s)
{
    k_prestep<<<1, 1, 0, s>>>(status, step, interval);
}

void launch_finish_rebuild(const float4 *pos, float4 *xref, int32_t n, int32_t *status, int32_t *counters,
                           hipStream_t s)
{
    if (n > 0)
        k_copy_xref<<<nblk(n), BLOCK, 0, s>>>(pos, xref, n, status);
    (void)counters;
}

void launch_clear_rebuild(int32_t *status, int32_t *counters, hipStream_t s)
{
    k_clear_rebuild<<<1, 1, 0, s>>>(status, counters);
}

void launch_fill_i32(int32_t *p, int32_t v, int64_t n, const int32_t *gate, hipStream_t s)
{
    if (n > 0)
    {
        const int64_t b = (n + BLOCK - 1) / BLOCK;
        k_fill_i32<<<(int32_t)(b < 4096 ? b : 4096), BLOCK, 0, s>>>(p, v, n, gate);
    }
}

} // namespace bs
