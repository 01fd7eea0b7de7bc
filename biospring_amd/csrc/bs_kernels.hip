// bs_kernels.hip -- CDNA4 (gfx950) kernels of the force-and-integrate step.
//
// Every kernel is wave64-shaped; none uses floating-point atomics, so the
// forces and energies of a step are bitwise reproducible run to run.  Integer
// atomics are used only for histogram counts, flags and the parity-hook pair
// extraction (whose output is sorted on the host).
#include "bs_engine.hpp"

#include <cmath>

namespace bs
{

namespace
{

__device__ inline uint32_t fbits(float f) { return __float_as_uint(f); }

__device__ inline bool gated_off(const int32_t *gate) { return gate != nullptr && *gate == 0; }

__device__ inline double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off, 64);
    return v;
}

// Block-wide sum, valid in thread 0.  Fixed order: deterministic.
__device__ inline double block_sum(double v, double *sh)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
        sh[wave] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w)
            r += sh[w];
    __syncthreads();
    return r;
}

// The reference's two-stage cutoff test, exactly (nsearch.hpp:230 then
// Particle.cpp:400-402 through Vector3f::norm, measure.hpp:163-196): an FP64
// distance of exact float differences, then an FP64 norm of float-rounded
// differences rounded to float; both must be below the float cutoff.
// Products and sums are kept unfused, as in the reference's x86-64 build.
__device__ __noinline__ bool exact_two_stage(float4 pi, float4 pj, float rc)
{
#pragma clang fp contract(off)
    const double ax = (double)pi.x - (double)pj.x;
    const double ay = (double)pi.y - (double)pj.y;
    const double az = (double)pi.z - (double)pj.z;
    const double s1 = ax * ax + ay * ay + az * az;
    if (!(sqrt(s1) < (double)rc))
        return false;
    const double bx = (double)(pj.x - pi.x);
    const double by = (double)(pj.y - pi.y);
    const double bz = (double)(pj.z - pi.z);
    float d = 0.0f;
    if (!(fabs(bx) < 1e-40 && fabs(by) < 1e-40 && fabs(bz) < 1e-40))
        d = (float)sqrt(bx * bx + by * by + bz * bz);
    return d < rc;
}

__device__ inline bool in_cutoff(float4 pi, float4 pj, float rc, float rc2lo, float rc2hi, float r2)
{
    if (r2 < rc2lo)
        return true;
    if (r2 >= rc2hi)
        return false;
    return exact_two_stage(pi, pj, rc);
}

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

__global__ void k_box_final(const double *partial, int32_t nb, double *box, const int32_t *gate)
{
    if (gated_off(gate))
        return;
    if (threadIdx.x < 6)
    {
        const int c = threadIdx.x;
        double v = partial[c];
        for (int b = 1; b < nb; ++b)
            v = c < 3 ? fmin(v, partial[6 * b + c]) : fmax(v, partial[6 * b + c]);
        box[c] = v;
    }
}

// ---------------------------------------------------------------------------
// Cell binning: ids, counting sort, deterministic in-cell order.
// ---------------------------------------------------------------------------
__device__ inline int32_t internal_coord(float x, double mn, double len, int32_t dim, float min_edge)
{
    const double e = fmax((double)min_edge, len / (double)dim);
    int32_t c = (int32_t)((x - mn) / e);
    c = c < 0 ? 0 : c;
    return c < dim - 1 ? c : dim - 1;
}

// exact_mode 1: nsearch.hpp:277-293 (size_t((x - min)/edge), clamped), FP64.
// exact_mode 0: the engine's own grid, edge >= min_edge on every axis.
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
            const uint64_t u = (uint64_t)q;
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
// Verlet list build: full list (both directions), spring-bonded pairs
// excluded, radius = largest cutoff + list skin.
// ---------------------------------------------------------------------------
__device__ inline bool bonded(const int32_t *sp_ptr, const int32_t *sp_j, int32_t i, int32_t j)
{
    int32_t lo = sp_ptr[i], hi = sp_ptr[i + 1];
    while (lo < hi)
    {
        const int32_t mid = (lo + hi) >> 1;
        if ((sp_j[mid] & 0x7fffffff) < j)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo < sp_ptr[i + 1] && (sp_j[lo] & 0x7fffffff) == j;
}

__global__ void __launch_bounds__(BLOCK) k_list_build(ListArgs a)
{
    if (gated_off(a.gate))
        return;
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n)
        return;
    const uint32_t fi = fbits(a.jprm[i].w);
    int32_t cnt = 0;
    if (!(fi & PF_STATIC) && (a.i_mask == 0u || (fi & a.i_mask)))
    {
        const float4 pi = a.pos[i];
        const float xs[3] = {pi.x, pi.y, pi.z};
        int32_t c[3];
        for (int q = 0; q < 3; ++q)
            c[q] = internal_coord(xs[q], a.box[q], a.box[3 + q] - a.box[q], a.dims[q], a.cell_edge);
        for (int dz = -1; dz <= 1; ++dz)
        {
            const int32_t cz = c[2] + dz;
            if (cz < 0 || cz >= a.dims[2])
                continue;
            for (int dy = -1; dy <= 1; ++dy)
            {
                const int32_t cy = c[1] + dy;
                if (cy < 0 || cy >= a.dims[1])
                    continue;
                for (int dx = -1; dx <= 1; ++dx)
                {
                    const int32_t cx = c[0] + dx;
                    if (cx < 0 || cx >= a.dims[0])
                        continue;
                    const int32_t cell = cx + cy * a.dims[0] + cz * a.dims[0] * a.dims[1];
                    for (int32_t t = a.cell_start[cell]; t < a.cell_start[cell + 1]; ++t)
                    {
                        const int32_t j = a.cell_idx[t];
                        if (j == i)
                            continue;
                        const float4 pj = a.pos[j];
                        const float ddx = pj.x - pi.x, ddy = pj.y - pi.y, ddz = pj.z - pi.z;
                        const float r2 = ddx * ddx + ddy * ddy + ddz * ddz;
                        if (!(r2 < a.r2list))
                            continue;
                        if (a.j_mask != 0u && !(fbits(a.jprm[j].w) & a.j_mask))
                            continue;
                        if (a.exclude_bonded && bonded(a.sp_ptr, a.sp_j, i, j))
                            continue;
                        if (cnt < a.cap)
                            a.nbl[(size_t)cnt * a.npad + i] = j;
                        ++cnt;
                    }
                }
            }
        }
    }
    a.nbn[i] = cnt < a.cap ? cnt : a.cap;
    if (cnt > a.cap)
        atomicOr(&a.status[2], 1);
    if (cnt > 0)
        atomicMax(&a.status[3], cnt);
}

// ---------------------------------------------------------------------------
// Pair forces over the list.  One lane per particle i; forces on i only
// (each dynamic-dynamic pair is visited from both sides), energies weighted
// 1/2 for a dynamic partner and 1 for a static one (Particle.cpp:231-249).
// Formulas: energy/steric.hpp, electrostatic.hpp, hydrophobic.hpp.
// ---------------------------------------------------------------------------
template <int SMODE, int EMODE, int HYDRO>
__global__ void __launch_bounds__(BLOCK) k_pair(PairArgs a)
{
    __shared__ double sh[BLOCK / 64];
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double es = 0.0, ee = 0.0, eh = 0.0;
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    const float cs = (float)a.k.c_spring;
    // Coulomb prefactors: energy kJ/mol for q1 q2 / (eps d), force module
    // q1 q2 / (4 pi eps d^2) * C_elec (electrostatic.hpp:16-49).
    const float ke = (float)(a.k.e_coul / a.diel) * a.elec_scale;
    const float kf = (float)(a.k.c_elec / (4.0 * a.k.pi * a.diel)) * a.elec_scale;
    if (i < a.n)
    {
        const int32_t cnt = a.nbn[i];
        const float4 pi = a.pos[i];
        const float4 ri = a.jprm[i];
        const uint32_t fli = fbits(ri.w);
        for (int32_t t = 0; t < cnt; ++t)
        {
            const int32_t j = a.nbl[(size_t)t * a.npad + i];
            const float4 pj = a.pos[j];
            const float dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
            const float r2 = dx * dx + dy * dy + dz * dz;
            if (!in_cutoff(pi, pj, a.rc, a.rc2lo, a.rc2hi, r2))
                continue;
            const float4 rj = a.jprm[j];
            const uint32_t flj = fbits(rj.w);
            const float w = (flj & PF_STATIC) ? 1.0f : 0.5f;
            float fscal = 0.0f; // force = (dx,dy,dz) * fscal
            if (r2 > 0.0f)
            {
                float inv = rsqrtf(r2);
                inv = inv * (1.5f - 0.5f * r2 * inv * inv);
                const float inv2 = inv * inv;
                if (SMODE >= 0)
                {
                    float e = 0.0f, f = 0.0f;
                    if (SMODE == 0)
                    { // linear: ri + rj from raw radii
                        const float d = r2 * inv;
                        const float dv = d - (ri.x + rj.x);
                        if (!(dv > 0.0f))
                        {
                            e = 0.5f * dv * dv;
                            f = dv * cs * inv; // -|dv| * C / d
                        }
                    }
                    else if (r2 >= 0.01f)
                    {
                        if (SMODE == 1)
                        { // Amber 12-6: sqrt(r), sqrt(eps) per bead
                            const float eij = ri.y * rj.y;
                            const float R = ri.x * rj.x;
                            const float s2 = R * R * inv2, s6 = s2 * s2 * s2, s12 = s6 * s6;
                            e = eij * (s12 - 2.0f * s6);
                            f = 12.0f * eij * (s6 - s12) * inv2 * cs;
                        }
                        else if (SMODE == 2)
                        { // Lewitt 8-6
                            const float eij = ri.y * rj.y;
                            const float R = ri.x * rj.x;
                            const float s2 = R * R * inv2, s6 = s2 * s2 * s2, s8 = s6 * s2;
                            e = eij * (3.0f * s8 - 4.0f * s6);
                            f = 24.0f * eij * (s6 - s8) * inv2 * cs;
                        }
                        else
                        { // Zacharias 8-6: raw r, eps, product rules
                            const float eij = ri.y * rj.y;
                            const float R = ri.x * rj.x;
                            const float s2 = R * R * inv2, s6 = s2 * s2 * s2, s8 = s6 * s2;
                            e = eij * (s8 - s6);
                            f = eij * (6.0f * s6 - 8.0f * s8) * inv2 * cs;
                        }
                    }
                    es += (double)(w * a.steric_scale * e);
                    fscal += a.steric_scale * f;
                }
                if (EMODE > 0 && (fli & PF_CHARGED) && (flj & PF_CHARGED) && r2 >= 1.0e-6f)
                {
                    const float qq = pi.w * pj.w;
                    if (EMODE == 1)
                    {
                        ee += (double)(w * ke * qq * inv);
                        fscal -= kf * qq * inv * inv2;
                    }
                    else
                    { // eps(r) = eps * r
                        ee += (double)(w * ke * qq * inv2);
                        fscal -= 2.0f * kf * qq * inv2 * inv2;
                    }
                }
                if (HYDRO && (fli & PF_HYDRO) && (flj & PF_HYDRO))
                {
                    const float d = r2 * inv;
                    const float hh = ri.z * rj.z * __expf(-d);
                    eh += (double)(w * a.hydro_scale) * (-(double)hh * 6.02214169E23 * 1.0E-3);
                    fscal += a.hydro_scale * hh * cs * inv;
                }
            }
            else if (SMODE == 0)
            {
                const float s = ri.x + rj.x; // coincident pair: linear energy, zero force
                es += (double)(w * a.steric_scale * 0.5f * s * s);
            }
            fx += dx * fscal;
            fy += dy * fscal;
            fz += dz * fscal;
        }
    }
    const double bs = block_sum(es, sh);
    const double be = block_sum(ee, sh);
    const double bh = block_sum(eh, sh);
    if (threadIdx.x == 0)
    {
        a.epart[3 * blockIdx.x + 0] = bs;
        a.epart[3 * blockIdx.x + 1] = be;
        a.epart[3 * blockIdx.x + 2] = bh;
    }
    if (i < a.n)
    {
        float4 f = make_float4(fx, fy, fz, 0.0f);
        if (a.accumulate)
        {
            const float4 o = a.fnb[i];
            f.x += o.x;
            f.y += o.y;
            f.z += o.z;
        }
        a.fnb[i] = f;
    }
}

// ---------------------------------------------------------------------------
// IMPALA (imp.hpp:29-159), FP32 geometry with FP64 energy/derivative, as the
// reference; returns the energy and adds the force (ForceField.cpp:80-86).
// ---------------------------------------------------------------------------
__device__ inline double imp_eval(float x, float y, float z, float sasa, float tr, float uo, float lo, float uc,
                                  float lc, float scale, double c_spring, float f[3])
{
    (void)y;
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

// ---------------------------------------------------------------------------
// Springs (CSR gather, G lanes per row) fused with viscosity, external and
// IMPALA forces and the viscous semi-implicit Euler step
// (SpringNetwork.cpp:63-90,202-241; Spring.cpp:20-34; Particle.cpp:80-128).
// The extension len - L0 is formed from an FP64 length so that it keeps full
// accuracy when the spring is near rest.
// ---------------------------------------------------------------------------
template <int G, bool INTEGRATE, bool COMP>
__global__ void __launch_bounds__(BLOCK) k_spring(SpringArgs a)
{
    __shared__ double sh[BLOCK / 64];
    const int32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int32_t row = tid / G;
    const int32_t lane = tid % G;
    double espr = 0.0, eke = 0.0, eimp = 0.0;
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
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
        for (int32_t e = beg + lane; e < end; e += G)
        {
            const int32_t jr = a.sp_j[e];
            const int32_t j = jr & 0x7fffffff;
            const float4 pj = a.pos[j];
            double dx = (double)pj.x - (double)pi.x;
            double dy = (double)pj.y - (double)pi.y;
            double dz = (double)pj.z - (double)pi.z;
            if (COMP)
            {
                const float4 lj = a.plo[j];
                dx += (double)lj.x - (double)li.x;
                dy += (double)lj.y - (double)li.y;
                dz += (double)lj.z - (double)li.z;
            }
            const double r2 = dx * dx + dy * dy + dz * dz;
            const double k = a.sp_k[e];
            double len = 0.0;
            float inv = 0.0f;
            if (r2 > 0.0)
            {
                inv = rsqrtf((float)r2);
                const double s = r2 * (double)inv;
                len = fma(fma(-s, s, r2), 0.5 * (double)inv, s);
            }
            const double ext = len - (double)a.sp_l0[e];
            const float fm = (float)(k * ext * cs) * inv;
            fx += (float)dx * fm;
            fy += (float)dy * fm;
            fz += (float)dz * fm;
            espr += (jr < 0 ? 1.0 : 0.5) * 0.5 * (double)a.spring_scale * k * ext * ext;
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
        if (a.fnb)
        {
            const float4 fn = a.fnb[row];
            fx += fn.x;
            fy += fn.y;
            fz += fn.z;
        }
        if (a.fext)
        {
            const float4 fe = a.fext[row];
            fx += fe.x;
            fy += fe.y;
            fz += fe.z;
        }
        if (a.visc_enable)
        {
            fx -= v.x * a.visc;
            fy -= v.y * a.visc;
            fz -= v.z * a.visc;
        }
        const float4 ip = a.iprm[row];
        if (a.imp_enable)
        {
            float f[3] = {fx, fy, fz};
            eimp += imp_eval(pi.x, pi.y, pi.z, ip.y, ip.z, a.imp_uo, a.imp_lo, a.imp_uc, a.imp_lc, a.imp_scale,
                             a.k.c_spring, f);
            fx = f[0];
            fy = f[1];
            fz = f[2];
        }
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

// Processed pairs of one search kind, as (min, max) original ids.
__global__ void k_pairs_extract(const float4 *pos, const float4 *jprm, const int32_t *nbl, const int32_t *nbn,
                                const int32_t *perm, int32_t n, int32_t npad, float rc, float rc2lo, float rc2hi,
                                uint32_t i_mask, uint32_t j_mask, int32_t nonzero, uint32_t *out, int32_t *count,
                                int64_t cap)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t fi = fbits(jprm[i].w);
    if (i_mask != 0u && !(fi & i_mask))
        return;
    const float4 pi = pos[i];
    const int32_t cnt = nbn[i];
    for (int32_t t = 0; t < cnt; ++t)
    {
        const int32_t j = nbl[(size_t)t * npad + i];
        const uint32_t fj = fbits(jprm[j].w);
        if (j_mask != 0u && !(fj & j_mask))
            continue;
        if (!(fj & PF_STATIC) && perm[j] < perm[i])
            continue;
        const float4 pj = pos[j];
        const float dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
        const float r2 = dx * dx + dy * dy + dz * dz;
        if (!in_cutoff(pi, pj, rc, rc2lo, rc2hi, r2))
            continue;
        if (nonzero && !(dx != 0.0f || dy != 0.0f || dz != 0.0f))
            continue;
        const int32_t slot = atomicAdd(count, 1);
        if (slot < cap)
        {
            const uint32_t a = (uint32_t)perm[i], b = (uint32_t)perm[j];
            out[2 * (size_t)slot] = a < b ? a : b;
            out[2 * (size_t)slot + 1] = a < b ? b : a;
        }
    }
}

inline int32_t nblk(int64_t n, int32_t b = BLOCK) { return (int32_t)((n + b - 1) / b); }

} // namespace

// ---- launchers ----

void launch_box_reduce(const float4 *pos, int32_t n, double *partial, double *box, int32_t nblocks,
                       const int32_t *gate, hipStream_t s)
{
    k_box_partial<<<nblocks, BLOCK, 0, s>>>(pos, n, partial, gate);
    k_box_final<<<1, 64, 0, s>>>(partial, nblocks, box, gate);
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

void launch_list_build(const ListArgs &a, hipStream_t s)
{
    if (a.n > 0)
        k_list_build<<<nblk(a.n), BLOCK, 0, s>>>(a);
}

template <int S, int E, int H> static void pair3(const PairArgs &a, int32_t nb, hipStream_t s)
{
    k_pair<S, E, H><<<nb, BLOCK, 0, s>>>(a);
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

void launch_pairs_extract(const float4 *pos, const float4 *jprm, const int32_t *nbl, const int32_t *nbn,
                          const int32_t *perm, int32_t n, int32_t npad, float rc, float rc2lo, float rc2hi,
                          uint32_t i_mask, uint32_t j_mask, int32_t nonzero, uint32_t *out, int32_t *count,
                          int64_t cap, hipStream_t s)
{
    if (n > 0)
        k_pairs_extract<<<nblk(n), BLOCK, 0, s>>>(pos, jprm, nbl, nbn, perm, n, npad, rc, rc2lo, rc2hi, i_mask,
                                                   j_mask, nonzero, out, count, cap);
}

} // namespace bs
