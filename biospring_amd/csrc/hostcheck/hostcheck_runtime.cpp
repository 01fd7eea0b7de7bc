// hostcheck_runtime.cpp -- CPU-side checker of the engine's host logic.
//
// Links the engine's real host code (bs_api.cpp) against a recording fake of
// the HIP runtime and against launchers that, instead of launching, verify
// that every buffer a kernel would touch is a live allocation of the extent
// the kernel assumes.  The binning and list-build launchers also perform
// their work on the CPU so that the host's read-backs (box, list lengths) see
// consistent data.  Any violation aborts the process with a message.  Test
// infrastructure only; never loaded by the engine.
#include "hip/hip_runtime.h"

#include "../bs_engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <vector>

namespace
{
std::map<uintptr_t, size_t> &allocs()
{
    static std::map<uintptr_t, size_t> m;
    return m;
}
std::mutex &mu()
{
    static std::mutex m;
    return m;
}
// file-scope device arrays, registered by HIP_SYMBOL
std::map<uintptr_t, size_t> &symbols()
{
    static std::map<uintptr_t, size_t> m;
    return m;
}
int64_t g_malloc_calls = 0;

bool inside(const std::map<uintptr_t, size_t> &m, uintptr_t a, size_t bytes)
{
    auto it = m.upper_bound(a);
    if (it == m.begin())
        return false;
    --it;
    return a >= it->first && a + bytes <= it->first + it->second;
}
} // namespace

bool hc_valid(const void *p, size_t bytes)
{
    std::lock_guard<std::mutex> lk(mu());
    const uintptr_t a = (uintptr_t)p;
    return inside(allocs(), a, bytes) || inside(symbols(), a, bytes);
}

const void *hc_symbol(const void *p, size_t bytes)
{
    std::lock_guard<std::mutex> lk(mu());
    symbols()[(uintptr_t)p] = bytes;
    return p;
}

hipError_t hipGetSymbolAddress(void **p, const void *symbol)
{
    *p = const_cast<void *>(symbol);
    return hipSuccess;
}

void hc_fail(const char *what)
{
    std::fprintf(stderr, "HOSTCHECK VIOLATION: %s\n", what);
    std::fflush(stderr);
    std::abort();
}

static void need(const void *p, size_t bytes, const char *what)
{
    if (bytes == 0)
        return;
    if (!hc_valid(p, bytes))
    {
        char buf[256];
        std::snprintf(buf, sizeof buf, "%s: %p (+%zu bytes) is not inside a live device allocation", what, p, bytes);
        hc_fail(buf);
    }
}

hipError_t hipMalloc(void **p, size_t bytes)
{
    *p = std::calloc(bytes ? bytes : 1, 1);
    std::lock_guard<std::mutex> lk(mu());
    ++g_malloc_calls;
    allocs()[(uintptr_t)*p] = bytes;
    return *p ? hipSuccess : hipErrorOutOfMemory;
}

hipError_t hipFree(void *p)
{
    if (!p)
        return hipSuccess;
    {
        std::lock_guard<std::mutex> lk(mu());
        auto it = allocs().find((uintptr_t)p);
        if (it == allocs().end())
            hc_fail("hipFree of a pointer that is not a live allocation");
        allocs().erase(it);
    }
    std::free(p);
    return hipSuccess;
}

std::map<uintptr_t, size_t> &host_allocs()
{
    static std::map<uintptr_t, size_t> m;
    return m;
}

hipError_t hipHostMalloc(void **p, size_t bytes, unsigned)
{
    *p = std::calloc(bytes ? bytes : 1, 1);
    std::lock_guard<std::mutex> lk(mu());
    host_allocs()[(uintptr_t)*p] = bytes;
    return *p ? hipSuccess : hipErrorOutOfMemory;
}

hipError_t hipHostFree(void *p)
{
    if (!p)
        return hipSuccess;
    {
        std::lock_guard<std::mutex> lk(mu());
        auto it = host_allocs().find((uintptr_t)p);
        if (it == host_allocs().end())
            hc_fail("hipHostFree of a pointer that is not a live pinned allocation");
        host_allocs().erase(it);
    }
    std::free(p);
    return hipSuccess;
}

hipError_t hipMemGetInfo(size_t *free, size_t *total)
{
    *free = *total = (size_t)1 << 36; // 64 GiB
    return hipSuccess;
}

hipError_t hipMemcpy(void *dst, const void *src, size_t bytes, hipMemcpyKind kind)
{
    if (kind == hipMemcpyHostToDevice)
        need(dst, bytes, "hipMemcpy H2D destination");
    if (kind == hipMemcpyDeviceToHost)
        need(src, bytes, "hipMemcpy D2H source");
    std::memcpy(dst, src, bytes);
    return hipSuccess;
}

hipError_t hipMemcpyAsync(void *dst, const void *src, size_t bytes, hipMemcpyKind kind, hipStream_t)
{
    return hipMemcpy(dst, src, bytes, kind);
}

hipError_t hipMemsetAsync(void *dst, int v, size_t bytes, hipStream_t)
{
    need(dst, bytes, "hipMemsetAsync");
    std::memset(dst, v, bytes);
    return hipSuccess;
}

hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned)
{
    *s = reinterpret_cast<hipStream_t>(new int(1));
    return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s)
{
    delete reinterpret_cast<int *>(s);
    return hipSuccess;
}
hipError_t hipGetDeviceCount(int *n)
{
    *n = 1;
    return hipSuccess;
}
hipError_t hipGetDevice(int *d)
{
    *d = 0;
    return hipSuccess;
}
hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int)
{
    std::strcpy(p->gcnArchName, "gfx950:sramecc+:xnack-");
    return hipSuccess;
}
hipError_t hipEventCreate(hipEvent_t *e)
{
    *e = reinterpret_cast<hipEvent_t>(new int(0));
    return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float *ms, hipEvent_t, hipEvent_t)
{
    *ms = 0.0f;
    return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t e)
{
    delete reinterpret_cast<int *>(e);
    return hipSuccess;
}
const char *hipGetErrorString(hipError_t) { return "hostcheck error"; }

extern "C" int64_t hc_live_allocations()
{
    std::lock_guard<std::mutex> lk(mu());
    return (int64_t)(allocs().size() + host_allocs().size());
}

// hipMalloc calls so far (a call that allocates nothing leaves it unchanged).
extern "C" int64_t hc_malloc_calls()
{
    std::lock_guard<std::mutex> lk(mu());
    return g_malloc_calls;
}

// ---------------------------------------------------------------------------
// Launchers: extent checks + CPU work where the host reads results back.
// ---------------------------------------------------------------------------
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { return hipErrorNotSupported; }
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t *g)
{
    *g = nullptr;
    return hipErrorNotSupported;
}
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t, hipGraphNode_t *, char *, size_t)
{
    *e = nullptr;
    return hipErrorNotSupported;
}
hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipErrorNotSupported; }
hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }
hipError_t hipGraphDestroy(hipGraph_t) { return hipSuccess; }

namespace bs
{

// The kernels' stop rule (bs_device.hpp): a list overflowed, or a step
// before this one produced a non-finite position.
static bool halted(const int32_t *status) { return status[2] != 0 || status[5] != 0; }

static bool gated_off(const int32_t *gate)
{
    if (!gate)
        return false;
    need(gate, 32, "gate");
    return *gate == 0 || halted(gate);
}

static uint32_t fbits(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

static float4 shifted(const float4 *pos, const float4 *shift, int32_t i)
{
    float4 p = pos[i];
    if (shift)
    {
        p.x += shift[i].x;
        p.y += shift[i].y;
        p.z += shift[i].z;
    }
    return p;
}

void launch_box_reduce(const float4 *pos, const float4 *shift, int32_t n, double *partial, double *box,
                       int32_t nblocks, const int32_t *gate, hipStream_t)
{
    need(pos, sizeof(float4) * n, "box: pos");
    if (shift)
        need(shift, sizeof(float4) * n, "box: shift");
    need(partial, sizeof(double) * 6 * nblocks, "box: partial");
    need(box, sizeof(double) * 6, "box: box");
    if (nblocks < 1 || nblocks > 4096)
        hc_fail("box: bad block count");
    if (gated_off(gate))
        return;
    for (int c = 0; c < 3; ++c)
    {
        box[c] = INFINITY;
        box[3 + c] = -INFINITY;
    }
    for (int32_t i = 0; i < n; ++i)
    {
        const float4 p = shifted(pos, shift, i);
        const float xs[3] = {p.x, p.y, p.z};
        for (int c = 0; c < 3; ++c)
        {
            box[c] = std::min(box[c], (double)xs[c]);
            box[3 + c] = std::max(box[3 + c], (double)xs[c]);
        }
    }
}

// One partial per block, as the device writes them: the whole box in
// partial 0, empty boxes in the others (their min/max change nothing).
static void launch_box_partial(const float4 *pos, const float4 *shift, int32_t n, double *partial, int32_t nblocks,
                               const int32_t *gate, hipStream_t s)
{
    need(partial, sizeof(double) * 6 * nblocks, "box_partial: partial");
    if (nblocks < 1 || nblocks > 256)
        hc_fail("box_partial: block count outside 1..BLOCK");
    if (gated_off(gate))
        return;
    launch_box_reduce(pos, shift, n, partial, partial, 1, nullptr, s);
    for (int32_t b = 1; b < nblocks; ++b)
        for (int c = 0; c < 3; ++c)
        {
            partial[6 * b + c] = INFINITY;
            partial[6 * b + 3 + c] = -INFINITY;
        }
}

static int32_t internal_coord(float x, double mn, double len, int32_t dim, float min_edge)
{
    const double e = std::max((double)min_edge, len / (double)dim);
    int32_t c = (int32_t)((x - mn) / e);
    c = c < 0 ? 0 : c;
    return c < dim - 1 ? c : dim - 1;
}

void launch_cell_ids(const float4 *pos, const float4 *shift, const float4 *jprm, uint32_t mask, int32_t n,
                     double *box, const double *box_part, int32_t box_nb, float4 *xref, int32_t d0, int32_t d1,
                     int32_t d2, double edge_exact, float min_edge, int32_t exact_mode, int32_t *cell_of,
                     int32_t *counts, const int32_t *gate, hipStream_t)
{
    const int64_t nc = (int64_t)d0 * d1 * d2;
    need(pos, sizeof(float4) * n, "cell_ids: pos");
    if (shift)
        need(shift, sizeof(float4) * n, "cell_ids: shift");
    need(jprm, sizeof(float4) * n, "cell_ids: jprm");
    need(box, sizeof(double) * 6, "cell_ids: box");
    need(cell_of, sizeof(int32_t) * n, "cell_ids: cell_of");
    need(counts, sizeof(int32_t) * nc, "cell_ids: counts");
    if (box_part)
    {
        need(box_part, sizeof(double) * 6 * box_nb, "cell_ids: box partials");
        if (box_nb < 1 || box_nb > 256)
            hc_fail("cell_ids: partial count outside 1..BLOCK");
    }
    if (xref)
        need(xref, sizeof(float4) * n, "cell_ids: xref");
    if (gated_off(gate))
        return;
    if (box_part)
        for (int c = 0; c < 3; ++c)
        {
            box[c] = INFINITY;
            box[3 + c] = -INFINITY;
            for (int32_t b = 0; b < box_nb; ++b)
            {
                box[c] = std::min(box[c], box_part[6 * b + c]);
                box[3 + c] = std::max(box[3 + c], box_part[6 * b + 3 + c]);
            }
        }
    if (xref && n > 0)
        std::memcpy(xref, pos, sizeof(float4) * n);
    const int32_t dims[3] = {d0, d1, d2};
    for (int32_t i = 0; i < n; ++i)
    {
        if (mask && !(fbits(jprm[i].w) & mask))
        {
            cell_of[i] = -1;
            continue;
        }
        const float4 p = shifted(pos, shift, i);
        const float xs[3] = {p.x, p.y, p.z};
        int32_t c[3];
        for (int a = 0; a < 3; ++a)
        {
            if (exact_mode)
            {
                const uint64_t u = (uint64_t)(((double)xs[a] - box[a]) / edge_exact);
                c[a] = (int32_t)(u < (uint64_t)(dims[a] - 1) ? u : (uint64_t)(dims[a] - 1));
            }
            else
                c[a] = internal_coord(xs[a], box[a], box[3 + a] - box[a], dims[a], min_edge);
        }
        const int64_t id = c[0] + (int64_t)c[1] * d0 + (int64_t)c[2] * d0 * d1;
        if (id < 0 || id >= nc)
            hc_fail("cell_ids: cell id out of range");
        cell_of[i] = (int32_t)id;
        counts[id]++;
    }
}

void launch_scan(const int32_t *counts, int32_t *start, int32_t ncells, const int32_t *gate, hipStream_t)
{
    need(counts, sizeof(int32_t) * ncells, "scan: counts");
    need(start, sizeof(int32_t) * (ncells + 1), "scan: start");
    if (gated_off(gate))
        return;
    int32_t run = 0;
    for (int32_t c = 0; c < ncells; ++c)
    {
        start[c] = run;
        run += counts[c];
    }
    start[ncells] = run;
}

void launch_scatter(const int32_t *cell_of, int32_t n, const int32_t *start, int32_t *cursor, int32_t drain,
                    int32_t *tmp, const int32_t *gate, hipStream_t)
{
    need(cell_of, sizeof(int32_t) * n, "scatter: cell_of");
    need(tmp, sizeof(int32_t) * n, "scatter: tmp");
    if (gated_off(gate))
        return;
    for (int32_t i = 0; i < n; ++i)
    {
        const int32_t c = cell_of[i];
        if (c < 0)
            continue;
        need(start + c, 8, "scatter: start[c..c+1]");
        need(cursor + c, 4, "scatter: cursor[c]");
        const int32_t slot = start[c] + (drain ? --cursor[c] : cursor[c]++);
        if (slot < 0 || slot >= n || (drain && cursor[c] < 0))
            hc_fail("scatter: slot out of range");
        tmp[slot] = i;
    }
}

void launch_cell_sort(const int32_t *cell_of, int32_t n, const int32_t *start, const int32_t *tmp, const int32_t *key,
                      int32_t *out, const int32_t *gate, hipStream_t)
{
    need(cell_of, sizeof(int32_t) * n, "cell_sort: cell_of");
    need(tmp, sizeof(int32_t) * n, "cell_sort: tmp");
    need(out, sizeof(int32_t) * n, "cell_sort: out");
    if (key)
        need(key, sizeof(int32_t) * n, "cell_sort: key");
    if (gated_off(gate))
        return;
    for (int32_t i = 0; i < n; ++i)
    {
        const int32_t c = cell_of[i];
        if (c < 0)
            continue;
        need(start + c, 8, "cell_sort: start");
        const int32_t ki = key ? key[i] : i;
        int32_t rank = 0;
        for (int32_t t = start[c]; t < start[c + 1]; ++t)
            rank += (key ? key[tmp[t]] : tmp[t]) < ki;
        out[start[c] + rank] = i;
    }
}

// Cluster list build on the CPU with the kernel's semantics (search range =
// cells covering the cluster bounding box grown by the radius).
void launch_list_build(const ListArgs &a, hipStream_t)
{
    const int64_t nc = (int64_t)a.dims[0] * a.dims[1] * a.dims[2];
    need(a.pos, sizeof(float4) * a.n, "list: pos");
    need(a.jprm, sizeof(float4) * a.n, "list: jprm");
    need(a.cell_start, sizeof(int32_t) * (nc + 1), "list: cell_start");
    need(a.cell_idx, sizeof(int32_t) * a.n, "list: cell_idx");
    need(a.box, sizeof(double) * 6, "list: box");
    need(a.cln, sizeof(int32_t) * a.ncl, "list: cln");
    need(a.status, sizeof(int32_t) * 4, "list: status");
    if (a.cap > 0)
    {
        need(a.cl, sizeof(uint32_t) * (size_t)a.ncl * a.cap, "list: cl");
    }
    if ((int64_t)a.ncl * CLUSTER < a.n || (int64_t)(a.ncl - 1) * CLUSTER >= a.n)
        hc_fail("list: bad cluster count");
    if (a.exclude_bonded)
        need(a.ex_ptr, sizeof(int32_t) * (a.ncl + 1), "list: ex_ptr");
    if (gated_off(a.gate))
        return;
    if (a.exclude_bonded)
        need(a.ex_w, sizeof(uint32_t) * std::max(a.ex_ptr[a.ncl], 1), "list: ex_w");
    if (a.shift)
        need(a.shift, sizeof(float4) * a.n, "list: shift");
    for (int32_t ci = 0; ci < a.ncl; ++ci)
    {
        uint32_t valid = 0, tag[CLUSTER] = {};
        float xi[CLUSTER], yi[CLUSTER], zi[CLUSTER];
        for (int k = 0; k < CLUSTER; ++k)
        {
            const int32_t i = ci * CLUSTER + k;
            xi[k] = yi[k] = zi[k] = 0.0f;
            if (i >= a.n)
                continue;
            const uint32_t fi = fbits(a.jprm[i].w);
            const float4 p = shifted(a.pos, a.shift, i);
            xi[k] = p.x;
            yi[k] = p.y;
            zi[k] = p.z;
            tag[k] = a.shift ? fbits(a.shift[i].w) : 0u;
            if (!(fi & PF_STATIC) && (a.i_mask == 0u || (fi & a.i_mask)))
                valid |= 1u << k;
        }
        int32_t count = 0;
        uint32_t todo = valid;
        while (todo)
        { // one search per system among the cluster's particles
            const uint32_t t0 = tag[__builtin_ctz(todo)];
            uint32_t vmask = 0;
            double bmin[3] = {INFINITY, INFINITY, INFINITY}, bmax[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (int k = 0; k < CLUSTER; ++k)
                if (((todo >> k) & 1u) && tag[k] == t0)
                {
                    vmask |= 1u << k;
                    const float v[3] = {xi[k], yi[k], zi[k]};
                    for (int q = 0; q < 3; ++q)
                    {
                        bmin[q] = std::min(bmin[q], (double)v[q]);
                        bmax[q] = std::max(bmax[q], (double)v[q]);
                    }
                }
            todo &= ~vmask;
            int32_t clo[3], chi[3];
            for (int q = 0; q < 3; ++q)
            {
                const double len = a.box[3 + q] - a.box[q];
                clo[q] = internal_coord((float)bmin[q] - a.radius, a.box[q], len, a.dims[q], a.cell_edge);
                chi[q] = internal_coord((float)bmax[q] + a.radius, a.box[q], len, a.dims[q], a.cell_edge);
            }
            for (int32_t cz = clo[2]; cz <= chi[2]; ++cz)
                for (int32_t cy = clo[1]; cy <= chi[1]; ++cy)
                {
                    const int32_t row = cy * a.dims[0] + cz * a.dims[0] * a.dims[1];
                    if (row + chi[0] + 1 > nc)
                        hc_fail("list: cell row out of range");
                    for (int32_t t = a.cell_start[row + clo[0]]; t < a.cell_start[row + chi[0] + 1]; ++t)
                    {
                        const int32_t j = a.cell_idx[t];
                        if (j < 0 || j >= a.n)
                            hc_fail("list: member out of range");
                        if (a.j_mask && !(fbits(a.jprm[j].w) & a.j_mask))
                            continue;
                        const float4 pj = shifted(a.pos, a.shift, j);
                        uint32_t m = 0;
                        for (int k = 0; k < CLUSTER; ++k)
                        {
                            const float dx = pj.x - xi[k], dy = pj.y - yi[k], dz = pj.z - zi[k];
                            if (dx * dx + dy * dy + dz * dz < a.r2list && j != ci * CLUSTER + k)
                                m |= 1u << k;
                        }
                        m &= vmask;
                        if (m && a.exclude_bonded)
                            for (int32_t x = a.ex_ptr[ci]; x < a.ex_ptr[ci + 1]; ++x)
                                if ((int32_t)(a.ex_w[x] >> 8) == j)
                                    m &= ~(a.ex_w[x] & 0xFFu);
                        if (!m)
                            continue;
                        if (count < a.cap)
                            a.cl[(size_t)ci * a.cap + count] = (uint32_t)j | (m << 24);
                        ++count;
                    }
                }
        }
        if (a.cap % 16)
            hc_fail("list: capacity not a multiple of 16");
        const int32_t kept = std::min(count, a.cap);
        for (int32_t slot = count; slot < std::min((count + 15) & ~15, a.cap); ++slot)
            a.cl[(size_t)ci * a.cap + slot] = 0u;
        a.cln[ci] = std::min(count, a.cap);
        if (count > a.cap)
            a.status[2] |= 1;
        a.status[3] = std::max(a.status[3], count);
    }
}

void launch_rows(const uint32_t *cl, const int32_t *cln, int32_t ncl, int32_t cap, uint32_t *rows, int32_t *rlen,
                 int32_t rcap, const int32_t *gate, hipStream_t)
{
    if (gated_off(gate))
        return;
    need(cln, sizeof(int32_t) * ncl, "rows: cln");
    need(cl, sizeof(uint32_t) * (size_t)ncl * cap, "rows: cl");
    const int64_t nrb = ((int64_t)ncl * CLUSTER + 63) / 64;
    need(rlen, sizeof(int32_t) * nrb * 64, "rows: rlen");
    need(rows, sizeof(uint32_t) * (size_t)nrb * rcap * 64, "rows: rows");
    if (rcap < cap)
        hc_fail("rows: row capacity below the cluster capacity");
    for (int32_t ci = 0; ci < ncl; ++ci)
        for (int k = 0; k < CLUSTER; ++k)
        {
            const int32_t i = ci * CLUSTER + k;
            int32_t c = 0;
            for (int32_t t = 0; t < cln[ci]; ++t)
            {
                const uint32_t w = cl[(size_t)ci * cap + t];
                if ((w >> (24 + k)) & 1u)
                    rows[((size_t)(i >> 6) * rcap + c++) * 64 + (i & 63)] = w & CL_JMASK;
            }
            rlen[i] = c;
        }
}

void launch_pair(const PairArgs &a, int32_t, int32_t, int32_t nb, hipStream_t)
{
    if (a.rows)
    {
        const int64_t nrb = ((int64_t)a.n + 63) / 64;
        need(a.rlen, sizeof(int32_t) * nrb * 64, "pair rows: rlen");
        need(a.rows, sizeof(uint32_t) * (size_t)nrb * a.rcap * 64, "pair rows: rows");
        for (int32_t i = 0; i < a.n; ++i)
        {
            if (a.rlen[i] < 0 || a.rlen[i] > a.rcap)
                hc_fail("pair rows: row length out of range");
            for (int32_t t = 0; t < a.rlen[i]; ++t)
                if ((int32_t)a.rows[((size_t)(i >> 6) * a.rcap + t) * 64 + (i & 63)] >= a.n)
                    hc_fail("pair rows: neighbour index out of range");
        }
    }
    need(a.pos, sizeof(float4) * a.n, "pair: pos");
    need(a.jprm, sizeof(float4) * a.n, "pair: jprm");
    need(a.cln, sizeof(int32_t) * a.ncl, "pair: cln");
    need(a.cl, sizeof(uint32_t) * (size_t)a.ncl * a.cap, "pair: cl");
    need(a.fnb, sizeof(double4) * a.n, "pair: fnb");
    if (a.energy)
        need(a.epp, sizeof(double4) * a.n, "pair: epp");
    if ((int64_t)nb * (BLOCK / 64) < a.ncl)
        hc_fail("pair: grid too small");
    for (int32_t ci = 0; ci < a.ncl; ++ci)
    {
        if (a.cln[ci] < 0 || a.cln[ci] > a.cap)
            hc_fail("pair: list length out of range");
        const int32_t cnt16 = (a.cln[ci] + 15) & ~15;
        if (cnt16 > a.cap)
            hc_fail("pair: padded list length exceeds the capacity");
        for (int32_t t = a.cln[ci]; t < cnt16; ++t)
            if (a.cl[(size_t)ci * a.cap + t] != 0u)
                hc_fail("pair: list padding is not zero");
        for (int32_t t = 0; t < a.cln[ci]; ++t)
        {
            const uint32_t w = a.cl[(size_t)ci * a.cap + t];
            if ((int32_t)(w & CL_JMASK) >= a.n)
                hc_fail("pair: neighbour index out of range");
            for (int k = 0; k < CLUSTER; ++k)
                if (((w >> (24 + k)) & 1u) && ci * CLUSTER + k >= a.n)
                    hc_fail("pair: mask bit on a padding lane");
        }
    }
    for (int32_t i = 0; i < a.n; ++i)
        if (!a.accumulate)
        {
            a.fnb[i] = make_double4(0, 0, 0, 0);
            if (a.energy)
                a.epp[i] = make_double4(0, 0, 0, 0);
        }
}

void launch_spring(const SpringArgs &a, int32_t group, int32_t nb, hipStream_t)
{
    const int32_t n = a.n;
    need(a.pos, sizeof(float4) * n, "spring: pos");
    need(a.vel, sizeof(float4) * n, "spring: vel");
    need(a.iprm, sizeof(float4) * n, "spring: iprm");
    need(a.jprm, sizeof(float4) * n, "spring: jprm");
    need(a.xref, sizeof(float4) * n, "spring: xref");
    need(a.sp_ptr, sizeof(int32_t) * (n + 1), "spring: sp_ptr");
    need(a.perm, sizeof(int32_t) * n, "spring: perm");
    need(a.status, sizeof(int32_t) * 8, "spring: status");
    if (a.energy)
        need(a.eps, sizeof(double4) * n, "spring: eps");
    for (const GridView *g : {&a.gpot, &a.gden})
        if (g->cells)
            need(g->cells, sizeof(float4) * (size_t)g->shape[0] * g->shape[1] * g->shape[2], "spring: grid");
    bool any_rigid = false;
    for (int32_t i = 0; i < n; ++i)
        any_rigid |= (fbits(a.jprm[i].w) & PF_RIGID) != 0;
    if (any_rigid && a.integrate)
        need(a.frig, sizeof(double4) * n, "spring: frig");
    if (a.fnb)
        need(a.fnb, sizeof(double4) * n, "spring: fnb");
    if (a.fext)
        need(a.fext, sizeof(float4) * n, "spring: fext");
    if (a.plo)
        need(a.plo, sizeof(float4) * n, "spring: plo");
    if (a.integrate)
    {
        need(a.pos_out, sizeof(float4) * n, "spring: pos_out");
        if (a.plo)
            need(a.plo_out, sizeof(float4) * n, "spring: plo_out");
    }
    else
        need(a.fout, sizeof(float4) * n, "spring: fout");
    if ((int64_t)nb * BLOCK < (int64_t)n * group)
        hc_fail("spring: grid too small");
    if (a.probe.state)
    {
        need(a.probe.state, sizeof(ProbeDev), "spring: probe state");
        need(a.probe.part, sizeof(double) * 8 * nb, "spring: probe block sums");
    }
    const int32_t ne = a.sp_ptr[n];
    if (ne > 0)
    {
        need(a.sp_j, sizeof(int32_t) * ne, "spring: sp_j");
        if (a.sp_k)
            need(a.sp_k, sizeof(float) * ne, "spring: sp_k");
        need(a.sp_l0, sizeof(float) * ne, "spring: sp_l0");
    }
    for (int32_t e = 0; e < ne; ++e)
        if ((a.sp_j[e] & 0x7fffffff) >= n)
            hc_fail("spring: edge index out of range");
    need(a.status, 32, "spring: status");
    if (a.integrate && halted(a.status))
    { // a stopped call carries every position over
        for (int32_t i = 0; i < n; ++i)
            const_cast<float4 *>(a.pos_out)[i] = a.pos[i];
        return;
    }
    if (a.integrate && a.steps_done)
        *a.steps_done += 1;
    if (a.integrate)
        for (int32_t i = 0; i < n; ++i)
            if (!(fbits(a.jprm[i].w) & (PF_STATIC | PF_RIGID)))
                const_cast<float4 *>(a.pos_out)[i] = a.pos[i];
    if (a.energy)
        for (int32_t i = 0; i < n; ++i)
            a.eps[i] = make_double4(0, 0, 0, 0);
}

// Adjoint launchers: extents and list invariants as the kernels assume them.
// The stand-ins compute the loss partials exactly, leave H w at zero and
// count one unit of dk per reverse step on each owning edge (dynamic row,
// static or larger partner), so that the host's edge-to-spring mapping can
// be checked.
void launch_adj_pair(const AdjPairArgs &a, int32_t, int32_t, int32_t nb, hipStream_t)
{
    need(a.pos, sizeof(float4) * a.n, "adj_pair: pos");
    need(a.jprm, sizeof(float4) * a.n, "adj_pair: jprm");
    need(a.cln, sizeof(int32_t) * a.ncl, "adj_pair: cln");
    need(a.cl, sizeof(uint32_t) * (size_t)a.ncl * a.cap, "adj_pair: cl");
    need(a.w, sizeof(float4) * a.n, "adj_pair: w");
    need(a.hw, sizeof(double4) * a.n, "adj_pair: hw");
    need(a.dq, sizeof(double) * a.n, "adj_pair: dq");
    if ((int64_t)nb * (BLOCK / 64) < a.ncl)
        hc_fail("adj_pair: grid too small");
    for (int32_t ci = 0; ci < a.ncl; ++ci)
    {
        if (a.cln[ci] < 0 || a.cln[ci] > a.cap)
            hc_fail("adj_pair: list length out of range");
        const int32_t cnt16 = (a.cln[ci] + 15) & ~15;
        if (cnt16 > a.cap)
            hc_fail("adj_pair: padded list length exceeds the capacity");
        for (int32_t t = 0; t < cnt16; ++t)
            if ((int32_t)(a.cl[(size_t)ci * a.cap + t] & CL_JMASK) >= a.n)
                hc_fail("adj_pair: neighbour index out of range");
    }
    if (!a.accumulate)
        for (int32_t i = 0; i < a.n; ++i)
            a.hw[i] = make_double4(0, 0, 0, 0);
}

void launch_adj_spring(const AdjSpringArgs &a, int32_t group, int32_t nb, hipStream_t)
{
    const int32_t n = a.n;
    need(a.pos, sizeof(float4) * n, "adj_spring: pos");
    need(a.jprm, sizeof(float4) * n, "adj_spring: jprm");
    need(a.iprm, sizeof(float4) * n, "adj_spring: iprm");
    need(a.sp_ptr, sizeof(int32_t) * (n + 1), "adj_spring: sp_ptr");
    need(a.w, sizeof(float4) * n, "adj_spring: w");
    need(a.lx, sizeof(double4) * n, "adj_spring: lx");
    need(a.lv, sizeof(double4) * n, "adj_spring: lv");
    need(a.w_next, sizeof(float4) * n, "adj_spring: w_next");
    if ((const float4 *)a.w_next == a.w)
        hc_fail("adj_spring: w_next aliases w");
    if (a.hwp)
        need(a.hwp, sizeof(double4) * n, "adj_spring: hwp");
    if ((int64_t)nb * BLOCK < (int64_t)n * group)
        hc_fail("adj_spring: grid too small");
    const int32_t ne = a.sp_ptr[n];
    if (ne > 0)
    {
        need(a.sp_j, sizeof(int32_t) * ne, "adj_spring: sp_j");
        if (a.sp_k)
            need(a.sp_k, sizeof(float) * ne, "adj_spring: sp_k");
        need(a.sp_l0, sizeof(float) * ne, "adj_spring: sp_l0");
    }
    if (a.dk)
    {
        need(a.slot_ptr, sizeof(int32_t) * (n + 1), "adj_spring: slot_ptr");
        need(a.dk, sizeof(float) * a.slot_ptr[n], "adj_spring: dk");
        need(a.dl0, sizeof(float) * a.slot_ptr[n], "adj_spring: dl0");
    }
    for (int32_t i = 0; i < n; ++i)
    {
        const bool dyn = !(fbits(a.jprm[i].w) & PF_STATIC);
        int32_t slot = a.dk ? a.slot_ptr[i] : 0;
        if (dyn && a.spring_enable)
            for (int32_t e = a.sp_ptr[i]; e < a.sp_ptr[i + 1]; ++e)
            {
                if ((a.sp_j[e] & 0x7fffffff) >= n)
                    hc_fail("adj_spring: edge index out of range");
                if (a.dk && (a.sp_j[e] < 0 || a.sp_j[e] > i))
                {
                    if (slot >= a.slot_ptr[i + 1])
                        hc_fail("adj_spring: more owned edges than slots in a row");
                    a.dk[slot++] += 1.0f;
                }
            }
        if (a.dk && dyn && a.spring_enable && slot != a.slot_ptr[i + 1])
            hc_fail("adj_spring: fewer owned edges than slots in a row");
        a.w_next[i] = make_float4(0, 0, 0, 0);
    }
}

void launch_spring_grads(float *dke, float *dle, int64_t ns, double *dk, double *dl0, hipStream_t)
{
    need(dke, sizeof(float) * ns, "spring_grads: dke");
    need(dle, sizeof(float) * ns, "spring_grads: dle");
    need(dk, sizeof(double) * ns, "spring_grads: dk");
    need(dl0, sizeof(double) * ns, "spring_grads: dl0");
    for (int64_t r = 0; r < ns; ++r)
    {
        dk[r] += dke[r];
        dl0[r] += dle[r];
        dke[r] = dle[r] = 0.0f;
    }
}

void launch_spring_out(const int32_t *slot_of, int64_t m, const double *dk, const double *dl0, double *outk,
                       double *outl, hipStream_t)
{
    need(slot_of, sizeof(int32_t) * m, "spring_out: slot_of");
    need(outk, sizeof(double) * m, "spring_out: outk");
    need(outl, sizeof(double) * m, "spring_out: outl");
    for (int64_t s = 0; s < m; ++s)
    {
        const int32_t r = slot_of[s];
        if (r >= 0)
        {
            need(dk + r, sizeof(double), "spring_out: dk");
            need(dl0 + r, sizeof(double), "spring_out: dl0");
        }
        outk[s] = r >= 0 ? dk[r] : 0.0;
        outl[s] = r >= 0 ? dl0[r] : 0.0;
    }
}

void launch_adj_init(const float4 *pos, const float4 *target, const float4 *jprm, const float4 *iprm, int32_t n,
                     double, double4 *lx, double4 *lv, float4 *w, double *loss_part, int32_t nblocks, hipStream_t)
{
    need(pos, sizeof(float4) * n, "adj_init: pos");
    need(target, sizeof(float4) * n, "adj_init: target");
    need(jprm, sizeof(float4) * n, "adj_init: jprm");
    need(iprm, sizeof(float4) * n, "adj_init: iprm");
    need(lx, sizeof(double4) * n, "adj_init: lx");
    need(lv, sizeof(double4) * n, "adj_init: lv");
    need(w, sizeof(float4) * n, "adj_init: w");
    need(loss_part, sizeof(double) * nblocks, "adj_init: loss_part");
    if ((int64_t)nblocks * BLOCK < n)
        hc_fail("adj_init: grid too small");
    for (int32_t b = 0; b < nblocks; ++b)
    {
        double t = 0.0;
        for (int32_t i = b * BLOCK; i < std::min(n, (b + 1) * BLOCK); ++i)
        {
            const double ex = (double)pos[i].x - target[i].x, ey = (double)pos[i].y - target[i].y,
                         ez = (double)pos[i].z - target[i].z;
            t += ex * ex + ey * ey + ez * ez;
        }
        loss_part[b] = t;
    }
    for (int32_t i = 0; i < n; ++i)
    {
        lx[i] = lv[i] = make_double4(0, 0, 0, 0);
        w[i] = make_float4(0, 0, 0, 0);
    }
}

void launch_adj_seed(const double *dx3, const double *dv3, const int32_t *perm, const float4 *jprm,
                     const float4 *iprm, int32_t n, double, double4 *lx, double4 *lv, float4 *w, hipStream_t)
{
    if (n <= 0)
        return;
    if (dx3)
        need(dx3, sizeof(double) * 3 * n, "adj_seed: dx3");
    if (dv3)
        need(dv3, sizeof(double) * 3 * n, "adj_seed: dv3");
    need(perm, sizeof(int32_t) * n, "adj_seed: perm");
    need(jprm, sizeof(float4) * n, "adj_seed: jprm");
    need(iprm, sizeof(float4) * n, "adj_seed: iprm");
    need(lx, sizeof(double4) * n, "adj_seed: lx");
    need(lv, sizeof(double4) * n, "adj_seed: lv");
    need(w, sizeof(float4) * n, "adj_seed: w");
    for (int32_t t = 0; t < n; ++t)
    {
        if (perm[t] < 0 || perm[t] >= n)
            hc_fail("adj_seed: perm out of range");
        lx[t] = lv[t] = make_double4(0, 0, 0, 0);
        w[t] = make_float4(0, 0, 0, 0);
    }
}

void launch_unperm_f64(const double *in, const int32_t *perm, int32_t n, double *out, hipStream_t)
{
    if (n <= 0)
        return;
    need(in, sizeof(double) * n, "unperm: in");
    need(perm, sizeof(int32_t) * n, "unperm: perm");
    need(out, sizeof(double) * n, "unperm: out");
    for (int32_t t = 0; t < n; ++t)
    {
        if (perm[t] < 0 || perm[t] >= n)
            hc_fail("unperm: perm out of range");
        out[perm[t]] = in[t];
    }
}

// Springs from a cutoff, computed like the kernel (same grid walk and test).
// Spring CSR, exclusions and construction expansion: the device kernels'
// results computed serially, with every access checked.
void launch_edge_count(const uint32_t *sa, const uint32_t *sb, int64_t m, const int32_t *iperm, int32_t *cnt,
                       hipStream_t)
{
    need(sa, sizeof(uint32_t) * m, "edge_count: sa");
    need(sb, sizeof(uint32_t) * m, "edge_count: sb");
    for (int64_t s = 0; s < m; ++s)
        for (uint32_t id : {sa[s], sb[s]})
        {
            need(iperm + id, 4, "edge_count: iperm");
            need(cnt + iperm[id], 4, "edge_count: cnt");
            ++cnt[iperm[id]];
        }
}

void launch_edge_fill(const uint32_t *sa, const uint32_t *sb, int64_t m, const int32_t *iperm, const int32_t *ptr,
                      int32_t *cursor, uint64_t *key, hipStream_t)
{
    need(sa, sizeof(uint32_t) * m, "edge_fill: sa");
    need(sb, sizeof(uint32_t) * m, "edge_fill: sb");
    for (int64_t s = 0; s < m; ++s)
    {
        const int32_t a = iperm[sa[s]], b = iperm[sb[s]];
        for (int side = 0; side < 2; ++side)
        {
            const int32_t r = side ? b : a, o = side ? a : b;
            need(ptr + r, 8, "edge_fill: ptr");
            need(cursor + r, 4, "edge_fill: cursor");
            const int32_t slot = ptr[r] + cursor[r]++;
            if (slot >= ptr[r + 1])
                hc_fail("edge_fill: row overflow");
            need(key + slot, 8, "edge_fill: key");
            key[slot] = (uint64_t)(uint32_t)o << 32 | ((uint32_t)s << 1 | (uint32_t)side);
        }
    }
}

void launch_edge_sort(const int32_t *ptr, int32_t n, const uint64_t *key, const float4 *jprm, const float *sl,
                      const float *sk, int32_t *ej, float *ek, float *el, hipStream_t)
{
    need(ptr, sizeof(int32_t) * (n + 1), "edge_sort: ptr");
    const int32_t E = n > 0 ? ptr[n] : 0;
    need(key, sizeof(uint64_t) * E, "edge_sort: key");
    need(ej, sizeof(int32_t) * E, "edge_sort: ej");
    need(ek, sizeof(float) * E, "edge_sort: ek");
    need(el, sizeof(float) * E, "edge_sort: el");
    need(jprm, sizeof(float4) * n, "edge_sort: jprm");
    for (int32_t i = 0; i < n; ++i)
    {
        std::vector<uint64_t> row(key + ptr[i], key + ptr[i + 1]);
        std::sort(row.begin(), row.end());
        for (size_t t = 0; t < row.size(); ++t)
        {
            if (t > 0 && row[t] == row[t - 1])
                hc_fail("edge_sort: repeated key");
            const int32_t j = (int32_t)(row[t] >> 32);
            const uint32_t s = (uint32_t)row[t] >> 1;
            if (j < 0 || j >= n)
                hc_fail("edge_sort: partner out of range");
            need(sl + s, 4, "edge_sort: rest length");
            need(sk + s, 4, "edge_sort: stiffness");
            const int32_t o = ptr[i] + (int32_t)t;
            ej[o] = (fbits(jprm[j].w) & PF_STATIC) ? (int32_t)((uint32_t)j | 0x80000000u) : j;
            ek[o] = sk[s];
            el[o] = sl[s];
        }
    }
}

static std::vector<uint32_t> hc_cluster_words(const int32_t *ptr, const int32_t *ej, int32_t n, int32_t c)
{
    std::vector<uint32_t> w;
    for (int32_t k = 0; k < CLUSTER && c * CLUSTER + k < n; ++k)
    {
        const int32_t i = c * CLUSTER + k;
        for (int32_t t = ptr[i]; t < ptr[i + 1]; ++t)
            w.push_back((uint32_t)(ej[t] & 0x7fffffff) << 8 | (1u << k));
    }
    std::sort(w.begin(), w.end());
    return w;
}

void launch_ex_merge(const int32_t *ptr, const int32_t *ej, int32_t n, int32_t ncl, uint32_t *merged,
                     int32_t *ucount, hipStream_t)
{
    need(ptr, sizeof(int32_t) * (n + 1), "ex_merge: ptr");
    need(ej, sizeof(int32_t) * (n > 0 ? ptr[n] : 0), "ex_merge: ej");
    need(merged, sizeof(uint32_t) * (n > 0 ? ptr[n] : 0), "ex_merge: merged");
    need(ucount, sizeof(int32_t) * ncl, "ex_merge: ucount");
    if (ncl != (n + CLUSTER - 1) / CLUSTER)
        hc_fail("ex_merge: cluster count");
    for (int32_t c = 0; c < ncl; ++c)
    {
        const std::vector<uint32_t> w = hc_cluster_words(ptr, ej, n, c);
        const int32_t base = ptr[c * CLUSTER];
        int32_t u = 0;
        for (size_t t = 0; t < w.size(); ++t)
        {
            merged[base + (int32_t)t] = w[t];
            u += (t == 0 || (w[t] >> 8) != (w[t - 1] >> 8)) ? 1 : 0;
        }
        ucount[c] = u;
    }
}

void launch_ex_fill(const int32_t *ptr, int32_t n, int32_t ncl, const uint32_t *merged, const int32_t *ex_ptr,
                    uint32_t *ex_w, hipStream_t)
{
    need(ex_ptr, sizeof(int32_t) * (ncl + 1), "ex_fill: ex_ptr");
    need(ex_w, sizeof(uint32_t) * ex_ptr[ncl], "ex_fill: ex_w");
    for (int32_t c = 0; c < ncl; ++c)
    {
        const int32_t base = ptr[std::min(c * CLUSTER, n)], end = ptr[std::min(c * CLUSTER + CLUSTER, n)];
        int32_t o = ex_ptr[c];
        for (int32_t t = base; t < end; ++t)
        {
            if (t > base && (merged[t] >> 8) == (merged[t - 1] >> 8))
            {
                ex_w[o - 1] |= merged[t] & 0xFFu;
                continue;
            }
            if (o >= ex_ptr[c + 1])
                hc_fail("ex_fill: more words than counted");
            ex_w[o++] = merged[t];
        }
        if (o != ex_ptr[c + 1])
            hc_fail("ex_fill: fewer words than counted");
    }
}

void launch_edge_keep(const int32_t *ptr, int32_t n, const int32_t *ej, const float *ek, const float *el,
                      const uint8_t *rigid, const int32_t *fptr, int32_t *cnt, int32_t *fej, float *fek, float *fel,
                      hipStream_t)
{
    need(ptr, sizeof(int32_t) * (n + 1), "edge_keep: ptr");
    need(rigid, (size_t)n, "edge_keep: rigid");
    if (fptr)
    {
        need(fptr, sizeof(int32_t) * (n + 1), "edge_keep: fptr");
        need(fej, sizeof(int32_t) * fptr[n], "edge_keep: fej");
        need(fek, sizeof(float) * fptr[n], "edge_keep: fek");
        need(fel, sizeof(float) * fptr[n], "edge_keep: fel");
    }
    else
        need(cnt, sizeof(int32_t) * n, "edge_keep: cnt");
    for (int32_t i = 0; i < n; ++i)
    {
        int32_t o = fptr ? fptr[i] : 0, c = 0;
        for (int32_t t = ptr[i]; t < ptr[i + 1]; ++t)
        {
            if (rigid[i] && rigid[ej[t] & 0x7fffffff])
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
        if (fptr && o != fptr[i + 1])
            hc_fail("edge_keep: row length differs from its count");
        if (!fptr)
            cnt[i] = c;
    }
}

void launch_rows_expand(const int32_t *offs, int32_t n, uint32_t *ida, hipStream_t)
{
    need(offs, sizeof(int32_t) * (n + 1), "rows_expand: offs");
    need(ida, sizeof(uint32_t) * (n > 0 ? offs[n] : 0), "rows_expand: ida");
    for (int32_t i = 0; i < n; ++i)
        for (int32_t t = offs[i]; t < offs[i + 1]; ++t)
            ida[t] = (uint32_t)i;
}

void launch_topo_springs(const TopoArgs &a, hipStream_t)
{
    const int32_t n = a.n;
    need(a.pos, sizeof(float4) * n, "topo: pos");
    need(a.box, sizeof(double) * 6, "topo: box");
    const int64_t ncells = (int64_t)a.dims[0] * a.dims[1] * a.dims[2];
    need(a.cstart, sizeof(int32_t) * (ncells + 1), "topo: cstart");
    need(a.cidx, sizeof(int32_t) * n, "topo: cidx");
    if (a.group)
        need(a.group, sizeof(int32_t) * n, "topo: group");
    if (!(std::sqrt(a.r2max) >= a.rc && std::sqrt(std::nextafter(a.r2max, 0.0)) < a.rc))
        hc_fail("topo: r2max is not the squared-distance threshold of the cutoff");
    if (!a.offs)
        need(a.counts, sizeof(int32_t) * n, "topo: counts");
    else
        need(a.offs, sizeof(int32_t) * (n + 1), "topo: offs");
    for (int32_t i = 0; i < n; ++i)
    {
        const float4 pi = a.pos[i];
        const bool si = fbits(pi.w) != 0u;
        int32_t c[3];
        const float xs[3] = {pi.x, pi.y, pi.z};
        for (int q = 0; q < 3; ++q)
            c[q] = internal_coord(xs[q], a.box[q], a.box[3 + q] - a.box[q], a.dims[q], a.edge);
        std::vector<std::pair<int32_t, float>> row;
        for (int32_t z = std::max(c[2] - 1, 0); z <= std::min(c[2] + 1, a.dims[2] - 1); ++z)
            for (int32_t y = std::max(c[1] - 1, 0); y <= std::min(c[1] + 1, a.dims[1] - 1); ++y)
                for (int32_t x = std::max(c[0] - 1, 0); x <= std::min(c[0] + 1, a.dims[0] - 1); ++x)
                {
                    const int32_t cell = x + y * a.dims[0] + z * a.dims[0] * a.dims[1];
                    for (int32_t t = a.cstart[cell]; t < a.cstart[cell + 1]; ++t)
                    {
                        const int32_t j = a.cidx[t];
                        if (j < 0 || j >= n)
                            hc_fail("topo: member index out of range");
                        if (j <= i)
                            continue;
                        if (a.group && a.group[j] == a.group[i])
                            continue;
                        const float4 pj = a.pos[j];
                        if (si && fbits(pj.w) != 0u)
                            continue;
                        const double ax = (double)pj.x - (double)pi.x, ay = (double)pj.y - (double)pi.y,
                                     az = (double)pj.z - (double)pi.z;
                        volatile double s2 = ax * ax;
                        s2 = s2 + ay * ay;
                        s2 = s2 + az * az;
                        const double d = std::sqrt((double)s2);
                        if (!(d < a.rc))
                            continue;
                        if (a.ex_off)
                        {
                            need(a.ex_off + i, 8, "topo: ex_off[i..i+1]");
                            const int32_t b = a.ex_off[i], e = a.ex_off[i + 1];
                            need(a.ex_j + b, sizeof(int32_t) * (e - b), "topo: ex_j row");
                            if (!std::is_sorted(a.ex_j + b, a.ex_j + e))
                                hc_fail("topo: existing partners not sorted");
                            if (std::binary_search(a.ex_j + b, a.ex_j + e, j))
                                continue;
                        }
                        row.push_back({j, (float)d});
                    }
                }
        if (!a.offs)
        {
            a.counts[i] = (int32_t)row.size();
            continue;
        }
        if ((int64_t)row.size() != (int64_t)a.offs[i + 1] - a.offs[i])
            hc_fail("topo: fill pass row length differs from the count pass");
        if ((int32_t)row.size() > a.max_row)
            hc_fail("topo: row longer than max_row");
        std::sort(row.begin(), row.end(), [](auto &u, auto &v) { return u.first < v.first; });
        need(a.jout + a.offs[i], sizeof(int32_t) * row.size(), "topo: jout");
        need(a.l0out + a.offs[i], sizeof(float) * row.size(), "topo: l0out");
        for (size_t t = 0; t < row.size(); ++t)
        {
            a.jout[a.offs[i] + t] = row[t].first;
            a.l0out[a.offs[i] + t] = row[t].second;
        }
    }
}

void launch_pack_positions(const float4 *pos, const float4 *plo, const int32_t *perm, int32_t n, float *out3,
                           hipStream_t)
{
    need(pos, sizeof(float4) * n, "pack: pos");
    if (plo)
        need(plo, sizeof(float4) * n, "pack: plo");
    need(perm, sizeof(int32_t) * n, "pack: perm");
    need(out3, sizeof(float) * 3 * n, "pack: out3");
    for (int32_t t = 0; t < n; ++t)
    {
        float4 p = pos[t];
        if (plo)
        {
            p.x = (float)((double)p.x + (double)plo[t].x);
            p.y = (float)((double)p.y + (double)plo[t].y);
            p.z = (float)((double)p.z + (double)plo[t].z);
        }
        const int32_t i = perm[t];
        if (i < 0 || i >= n)
            hc_fail("pack: permutation entry out of range");
        out3[3 * i] = p.x;
        out3[3 * i + 1] = p.y;
        out3[3 * i + 2] = p.z;
    }
}

void launch_differs_f32(const float *v, int64_t m, int32_t *flag, hipStream_t)
{
    need(v, sizeof(float) * m, "differs: v");
    need(flag, sizeof(int32_t), "differs: flag");
    for (int64_t s = 1; s < m; ++s)
        if (std::memcmp(&v[s], &v[0], 4) != 0)
            *flag |= 1;
}

void launch_drift_states(const float4 *states, int32_t nstates, int32_t n, const float4 *xref, float half_skin2,
                         int32_t *flag, hipStream_t)
{
    need(states, sizeof(float4) * (size_t)n * (size_t)std::max(nstates, 0), "drift_states: states");
    need(xref, sizeof(float4) * n, "drift_states: xref");
    need(flag, sizeof(int32_t), "drift_states: flag");
    for (int32_t k = 0; k < nstates; ++k)
        for (int32_t i = 0; i < n; ++i)
        {
            const float4 p = states[(size_t)k * n + i], r = xref[i];
            const float ex = p.x - r.x, ey = p.y - r.y, ez = p.z - r.z;
            if (!(ex * ex + ey * ey + ez * ez <= half_skin2))
                *flag |= 1;
        }
}

void launch_drift_check(const float4 *pos, const float4 *xref, int32_t n, float, int32_t *status, hipStream_t)
{
    need(pos, sizeof(float4) * n, "drift_check: pos");
    need(xref, sizeof(float4) * n, "drift_check: xref");
    need(status, sizeof(int32_t) * 8, "drift_check: status");
}

void launch_energies(const double4 *epp, const double4 *eps, const int32_t *seg_start, int32_t nseg,
                     const int32_t *sys_seg, int32_t nsys, double *seg_part, double *out_sys, double *out_total,
                     const ProbeDev *probe, hipStream_t)
{
    if (probe)
    {
        need(probe, sizeof(ProbeDev), "energies: probe");
        if (nsys != 1)
            hc_fail("energies: the probe with a batch of systems");
    }
    need(seg_start, sizeof(int32_t) * (nseg + 1), "energies: seg_start");
    need(sys_seg, sizeof(int32_t) * (nsys + 1), "energies: sys_seg");
    need(seg_part, sizeof(double) * 6 * std::max(nseg, 1), "energies: seg_part");
    need(out_sys, sizeof(double) * 6 * nsys, "energies: out_sys");
    need(out_total, sizeof(double) * 6, "energies: out_total");
    const int32_t n = seg_start[nseg];
    need(eps, sizeof(double4) * n, "energies: eps");
    if (epp)
        need(epp, sizeof(double4) * n, "energies: epp");
    if (sys_seg[0] != 0 || sys_seg[nsys] != nseg)
        hc_fail("energies: system segment ranges do not cover the segments");
    for (int32_t g = 0; g < nseg; ++g)
        if (seg_start[g + 1] <= seg_start[g])
            hc_fail("energies: empty segment");
    for (int k = 0; k < 6 * nsys; ++k)
        out_sys[k] = 0.0;
    for (int k = 0; k < 6; ++k)
        out_total[k] = 0.0;
}

void launch_probe(const ProbeArgs &p, int32_t nparts, float, int32_t, float, double, int32_t, hipStream_t)
{
    need(p.state, sizeof(ProbeDev), "probe: state");
    need(p.part, sizeof(double) * 8 * std::max(nparts, 1), "probe: block sums");
}

void launch_rigid(const RigidArgs &a, hipStream_t)
{
    if (a.nb <= 0)
        return;
    need(a.bodies, sizeof(RigidBody) * a.nb, "rigid: bodies");
    need(a.bstart, sizeof(int32_t) * (a.nb + 1), "rigid: bstart");
    if (a.bstart[0] != 0 || a.bstart[a.nb] != a.nr)
        hc_fail("rigid: slot ranges do not cover the slots");
    for (int32_t b = 0; b < a.nb; ++b)
        if (a.bstart[b + 1] < a.bstart[b])
            hc_fail("rigid: slot ranges not ascending");
    if (a.nr == 0)
        return;
    need(a.slot_p, sizeof(int32_t) * a.nr, "rigid: slot_p");
    need(a.slot_b, sizeof(int32_t) * a.nr, "rigid: slot_b");
    need(a.p0, sizeof(double4) * a.nr, "rigid: p0");
    need(a.rx, sizeof(double4) * a.nr, "rigid: rx");
    need(a.status, sizeof(int32_t) * 8, "rigid: status");
    int32_t n = 0;
    for (int32_t t = 0; t < a.nr; ++t)
    {
        if (a.slot_b[t] < 0 || a.slot_b[t] >= a.nb || t < a.bstart[a.slot_b[t]] || t >= a.bstart[a.slot_b[t] + 1])
            hc_fail("rigid: slot outside its body's range");
        n = std::max(n, a.slot_p[t] + 1);
        if (a.slot_p[t] < 0)
            hc_fail("rigid: negative particle index");
    }
    need(a.pos, sizeof(float4) * n, "rigid: pos");
    need(a.pos_out, sizeof(float4) * n, "rigid: pos_out");
    need(a.vel, sizeof(float4) * n, "rigid: vel");
    need(a.frig, sizeof(double4) * n, "rigid: frig");
    need(a.xref, sizeof(float4) * n, "rigid: xref");
    need(a.perm, sizeof(int32_t) * n, "rigid: perm");
    for (int32_t t = 0; t < a.nr; ++t)
        const_cast<float4 *>(a.pos_out)[a.slot_p[t]] = a.pos[a.slot_p[t]];
}

void launch_centroid(const float4 *pos, int32_t n, double *partial, int32_t nblocks, double *out3, hipStream_t)
{
    need(pos, sizeof(float4) * n, "centroid: pos");
    need(partial, sizeof(double) * 3 * nblocks, "centroid: partial");
    need(out3, sizeof(double) * 3, "centroid: out");
    double c[3] = {0, 0, 0};
    for (int32_t i = 0; i < n; ++i)
    {
        c[0] += pos[i].x;
        c[1] += pos[i].y;
        c[2] += pos[i].z;
    }
    for (int k = 0; k < 3; ++k)
        out3[k] = n > 0 ? c[k] / n : 0.0;
}

void launch_prestep(int32_t *status, int32_t *counters, int32_t interval, int32_t advance, const float4 *pos,
                    const float4 *shift, int32_t n, double *partial, int32_t nblocks, int32_t *log, hipStream_t st)
{
    if (log)
        need(log, 4, "prestep: log");
    need(status, 32, "prestep: status");
    need(counters, 8, "prestep: counters");
    if (partial && n > 0)
        launch_box_partial(pos, shift, n, partial, nblocks, nullptr, st);
    if (halted(status))
        return;
    if (advance && status[1] != 0x7fffffff)
    {
        status[5] = 1;
        return;
    }
    const bool due = advance && interval > 0 && counters[0] > 0 && counters[0] % interval == 0;
    const int32_t r = (status[7] != 0 || due) ? 1 : 0;
    status[7] = 0;
    status[0] = r;
    counters[1] += r;
    if (advance)
        counters[0] += 1;
    if (log)
        *log = r;
}

void launch_fill_i32(int32_t *p, int32_t v, int64_t n, const int32_t *gate, hipStream_t)
{
    need(p, sizeof(int32_t) * n, "fill");
    if (gated_off(gate))
        return;
    for (int64_t i = 0; i < n; ++i)
        p[i] = v;
}

void launch_pairs_extract(const float4 *pos, const float4 *jprm, const uint32_t *cl, const int32_t *cln,
                          const int32_t *perm, int32_t n, int32_t ncl, int32_t lcap, float rc, float, float,
                          uint32_t i_mask, uint32_t j_mask, int32_t, uint32_t *out, int32_t *count, int64_t cap,
                          hipStream_t)
{
    need(pos, sizeof(float4) * n, "extract: pos");
    need(jprm, sizeof(float4) * n, "extract: jprm");
    need(cln, sizeof(int32_t) * ncl, "extract: cln");
    need(cl, sizeof(uint32_t) * (size_t)ncl * lcap, "extract: cl");
    need(perm, sizeof(int32_t) * n, "extract: perm");
    need(count, 4, "extract: count");
    if (cap > 0)
        need(out, sizeof(uint32_t) * 2 * cap, "extract: out");
    for (int32_t i = 0; i < n; ++i)
    {
        if (i_mask && !(fbits(jprm[i].w) & i_mask))
            continue;
        const int32_t ci = i / CLUSTER, il = i % CLUSTER;
        for (int32_t t = 0; t < cln[ci]; ++t)
        {
            const uint32_t w = cl[(size_t)ci * lcap + t];
            if (!((w >> (24 + il)) & 1u))
                continue;
            const int32_t j = (int32_t)(w & CL_JMASK);
            const uint32_t fj = fbits(jprm[j].w);
            if (j_mask && !(fj & j_mask))
                continue;
            if (!(fj & PF_STATIC) && perm[j] < perm[i])
                continue;
            // the reference's two-stage test (nsearch.hpp:230, Particle.cpp:400-402)
            const double ax = (double)pos[i].x - (double)pos[j].x, ay = (double)pos[i].y - (double)pos[j].y,
                         az = (double)pos[i].z - (double)pos[j].z;
            if (!(std::sqrt(ax * ax + ay * ay + az * az) < (double)rc))
                continue;
            const float dx = pos[j].x - pos[i].x, dy = pos[j].y - pos[i].y, dz = pos[j].z - pos[i].z;
            const double bx = dx, by = dy, bz = dz;
            const float d = (std::fabs(bx) < 1e-40 && std::fabs(by) < 1e-40 && std::fabs(bz) < 1e-40)
                                ? 0.0f
                                : (float)std::sqrt(bx * bx + by * by + bz * bz);
            if (!(d < rc))
                continue;
            const int32_t slot = (*count)++;
            if (slot < cap)
            {
                out[2 * slot] = std::min(perm[i], perm[j]);
                out[2 * slot + 1] = std::max(perm[i], perm[j]);
            }
        }
    }
}

} // namespace bs
