// bs_gen.cpp -- seeded synthetic spring networks for the benchmark and tests.
//
// Project-owned PRNG (splitmix64) with explicit transforms, so the same seed
// gives the same system on every machine and compiler.  Springs follow the
// reference's cutoff rule (Topology::add_springs_from_cutoff,
// src/topology/Topology.hpp:105-138): one spring per pair closer than the
// cutoff (FP64 distance), rest length = that FP64 distance stored as float
// (src/topology/Spring.hpp:47-56 then SpringNetwork::addSpring's float).
//
// Workloads (lengths in A, k in kJ/mol/A^2; the nm figures of the workload
// text converted with 1 nm = 10 A and 1 kJ/mol/nm^2 = 0.01 kJ/mol/A^2):
//   0  CA elastic network: 2,000 beads, self-avoiding chain, 3.8 A bonds,
//      inside a 30 A sphere; springs < 12 A, k = 0.1; linear steric 4 A;
//      20% charged +-1; Coulomb 16 A; viscosity 1 Da/fs; dt = 1 fs.
//   1  250,000 beads uniform in a 140 A cube, 1.2 A minimum separation;
//      springs < 8 A, k lognormal (mean 0.5, sigma_log 0.5); Amber 12-6 with
//      radius U[3,4] and epsilon U[0.1,1]; Coulomb with eps(r) = r, q ~ N(0,0.4),
//      12 A cutoffs; viscosity 1 Da/fs; dt = 1 fs.
//   2  IMPALA: CG beads in 4-bead residues, sasa U[10,100] A^2, transfer
//      U[-1,1]; springs < 10 A; membrane set by the caller.
//   3  one system of a batch: workload 0's CA network with the spring cutoff
//      drawn from U[9,14] A by the system's seed (the caller draws the sizes).
//   4  fitting network: workload 1's beads and force field at 100,000 beads
//      (same density), springs < 10 A.
// Any workload accepts an explicit bead count (n > 0) at the same density.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "bs_params_defaults.h"

namespace
{

struct Rng
{
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull) {}
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    // uniform in [0, 1)
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    // standard normal, Box-Muller (first output only)
    double normal()
    {
        double u1 = uniform();
        while (u1 <= 0.0)
            u1 = uniform();
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

struct System
{
    std::vector<float> x, v, mass, charge, radius, eps, sasa, transfer, hydro;
    std::vector<uint8_t> flags;
    std::vector<uint32_t> id1, id2;
    std::vector<float> l0, k;
    int64_t n() const { return (int64_t)mass.size(); }
    void resize(int64_t n)
    {
        x.assign(3 * n, 0.0f);
        v.assign(3 * n, 0.0f);
        mass.assign(n, 1.0f);
        charge.assign(n, 0.0f);
        radius.assign(n, 1.0f);
        eps.assign(n, 0.0f);
        sasa.assign(n, 0.0f);
        transfer.assign(n, 0.0f);
        hydro.assign(n, 0.0f);
        flags.assign(n, 0);
    }
};

// Uniform hash grid for proximity queries during generation.
struct Grid
{
    double lo[3], h;
    int64_t d[3];
    std::vector<std::vector<int32_t>> cells;
    Grid(const double lo_[3], const double hi_[3], double h_) : h(h_)
    {
        for (int c = 0; c < 3; ++c)
        {
            lo[c] = lo_[c];
            d[c] = std::max<int64_t>(1, (int64_t)std::ceil((hi_[c] - lo_[c]) / h) + 1);
        }
        cells.resize((size_t)(d[0] * d[1] * d[2]));
    }
    int64_t coord(double v, int c) const
    {
        int64_t q = (int64_t)std::floor((v - lo[c]) / h);
        return std::min(std::max<int64_t>(q, 0), d[c] - 1);
    }
    int64_t id(int64_t a, int64_t b, int64_t c) const { return a + d[0] * (b + d[1] * c); }
    void add(const float * p, int32_t i) { cells[(size_t)id(coord(p[0], 0), coord(p[1], 1), coord(p[2], 2))].push_back(i); }
    template <typename F> void near(const double * p, F && f) const
    {
        const int64_t a = coord(p[0], 0), b = coord(p[1], 1), c = coord(p[2], 2);
        for (int64_t z = std::max<int64_t>(c - 1, 0); z <= std::min(c + 1, d[2] - 1); ++z)
            for (int64_t y = std::max<int64_t>(b - 1, 0); y <= std::min(b + 1, d[1] - 1); ++y)
                for (int64_t xx = std::max<int64_t>(a - 1, 0); xx <= std::min(a + 1, d[0] - 1); ++xx)
                    for (int32_t j : cells[(size_t)id(xx, y, z)])
                        if (!f(j))
                            return;
    }
};

double dist2(const float * a, const double * b)
{
    const double dx = (double)a[0] - b[0], dy = (double)a[1] - b[1], dz = (double)a[2] - b[2];
    return dx * dx + dy * dy + dz * dz;
}

// Springs for every pair closer than cutoff (FP64 distance of the float
// positions), i < j, ordered by i then j.
void cutoff_springs(System & s, double cutoff, Rng & rng, int kmode, double kval)
{
    const int64_t n = s.n();
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int64_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c)
        {
            lo[c] = std::min(lo[c], (double)s.x[3 * i + c]);
            hi[c] = std::max(hi[c], (double)s.x[3 * i + c]);
        }
    Grid g(lo, hi, cutoff);
    for (int64_t i = 0; i < n; ++i)
        g.add(&s.x[3 * i], (int32_t)i);
    std::vector<std::vector<std::pair<uint32_t, float>>> rows((size_t)n);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i)
    {
        const double p[3] = {s.x[3 * i], s.x[3 * i + 1], s.x[3 * i + 2]};
        auto & r = rows[(size_t)i];
        g.near(p, [&](int32_t j) {
            if (j > i)
            {
                const double d2 = dist2(&s.x[3 * (size_t)j], p);
                const double d = std::sqrt(d2);
                if (d < cutoff && !(s.flags[i] & BS_FLAG_STATIC && s.flags[j] & BS_FLAG_STATIC))
                    r.emplace_back((uint32_t)j, (float)d);
            }
            return true;
        });
        std::sort(r.begin(), r.end());
    }
    size_t m = 0;
    for (auto & r : rows)
        m += r.size();
    s.id1.resize(m);
    s.id2.resize(m);
    s.l0.resize(m);
    s.k.resize(m);
    size_t w = 0;
    // lognormal with arithmetic mean kval and sigma_log 0.5
    const double sig = 0.5, mu = std::log(kval) - 0.5 * sig * sig;
    for (int64_t i = 0; i < n; ++i)
        for (auto & e : rows[(size_t)i])
        {
            s.id1[w] = (uint32_t)i;
            s.id2[w] = e.first;
            s.l0[w] = e.second;
            s.k[w] = kmode == 0 ? (float)kval : (float)std::exp(mu + sig * rng.normal());
            ++w;
        }
}

// Uniform random beads in [0, L)^3 with a hard-core minimum separation.
void uniform_beads(System & s, int64_t n, double L, double dmin, Rng & rng)
{
    s.resize(n);
    const double lo[3] = {0, 0, 0}, hi[3] = {L, L, L};
    Grid g(lo, hi, std::max(dmin, 1e-3));
    const double d2min = dmin * dmin;
    int64_t placed = 0;
    while (placed < n)
    {
        const double p[3] = {rng.uniform(0, L), rng.uniform(0, L), rng.uniform(0, L)};
        const float pf[3] = {(float)p[0], (float)p[1], (float)p[2]};
        const double pd[3] = {pf[0], pf[1], pf[2]};
        bool ok = true;
        g.near(pd, [&](int32_t j) {
            if (dist2(&s.x[3 * (size_t)j], pd) < d2min)
                ok = false;
            return ok;
        });
        if (!ok)
            continue;
        std::memcpy(&s.x[3 * placed], pf, sizeof pf);
        g.add(pf, (int32_t)placed);
        ++placed;
    }
}

// Self-avoiding chain of n beads with bond length b inside a sphere of radius R
// (min non-bonded separation dmin); backtracks when a bead cannot be placed.
void chain_in_sphere(System & s, int64_t n, double b, double R, double dmin, Rng & rng)
{
    s.resize(n);
    std::vector<std::array<float, 3>> pos;
    pos.reserve(n);
    const double d2min = dmin * dmin;
    // hashed grid of edge dmin over the placed beads: the self-avoidance test
    // looks at the 27 cells around a trial point
    std::unordered_map<uint64_t, std::vector<int64_t>> grid;
    auto cell = [&](double v) { return (int64_t)std::floor(v / dmin) + (1 << 20); };
    auto key = [](int64_t a, int64_t b, int64_t c) { return (uint64_t)a << 42 | (uint64_t)b << 21 | (uint64_t)c; };
    auto fits = [&](const double * p, int64_t upto) {
        const int64_t c0 = cell(p[0]), c1 = cell(p[1]), c2 = cell(p[2]);
        for (int64_t a = c0 - 1; a <= c0 + 1; ++a)
            for (int64_t bb = c1 - 1; bb <= c1 + 1; ++bb)
                for (int64_t c = c2 - 1; c <= c2 + 1; ++c)
                {
                    const auto it = grid.find(key(a, bb, c));
                    if (it == grid.end())
                        continue;
                    for (int64_t j : it->second)
                    {
                        if (j + 1 >= upto) // the chain predecessor is not tested
                            continue;
                        const double dx = pos[j][0] - p[0], dy = pos[j][1] - p[1], dz = pos[j][2] - p[2];
                        if (dx * dx + dy * dy + dz * dz < d2min)
                            return false;
                    }
                }
        return true;
    };
    auto push = [&](const std::array<float, 3> & q) {
        grid[key(cell(q[0]), cell(q[1]), cell(q[2]))].push_back((int64_t)pos.size());
        pos.push_back(q);
    };
    auto pop = [&]() {
        const auto & q = pos.back();
        grid[key(cell(q[0]), cell(q[1]), cell(q[2]))].pop_back(); // the largest index is last in its cell
        pos.pop_back();
    };
    push({0.0f, 0.0f, 0.0f});
    int fails = 0;
    while ((int64_t)pos.size() < n)
    {
        const auto prev = pos.back();
        bool placed = false;
        for (int t = 0; t < 400 && !placed; ++t)
        {
            double u[3] = {rng.normal(), rng.normal(), rng.normal()};
            const double nrm = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
            if (nrm == 0.0)
                continue;
            const double p[3] = {prev[0] + b * u[0] / nrm, prev[1] + b * u[1] / nrm, prev[2] + b * u[2] / nrm};
            if (p[0] * p[0] + p[1] * p[1] + p[2] * p[2] > R * R)
                continue;
            const float pf[3] = {(float)p[0], (float)p[1], (float)p[2]};
            const double pd[3] = {pf[0], pf[1], pf[2]};
            if (!fits(pd, (int64_t)pos.size()))
                continue;
            push({pf[0], pf[1], pf[2]});
            placed = true;
        }
        if (!placed)
        {
            ++fails;
            const int64_t back = std::min<int64_t>((int64_t)pos.size() - 1, 1 + fails % 8);
            for (int64_t t = 0; t < back; ++t)
                pop();
        }
    }
    for (int64_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c)
            s.x[3 * i + c] = pos[i][c];
}

} // namespace

extern "C" {

struct bsg_system;

// Parameters of workload `config` (see the header comment).
void bsg_params(int config, bs_params * p)
{
    bs_params_fill_defaults(p);
    p->timestep = 1.0f;
    p->viscosity_enable = 1;
    p->viscosity_value = 1.0f;
    p->spring_enable = 1;
    switch (config)
    {
    case 0:
    case 3:
        p->steric_enable = 1;
        p->steric_mode = BS_STERIC_LINEAR;
        p->steric_cutoff = 4.0f;
        p->electrostatic_enable = 1;
        p->electrostatic_cutoff = 16.0f;
        p->list_skin = 1.0f;
        break;
    case 1:
    case 4:
        p->steric_enable = 1;
        p->steric_mode = BS_STERIC_AMBER_12_6;
        p->steric_cutoff = 12.0f;
        p->electrostatic_enable = 1;
        p->electrostatic_cutoff = 12.0f;
        p->dielectric_model = BS_DIELECTRIC_DISTANCE;
        p->list_skin = 0.25f; // beads move < 0.125 A in 10 steps: no rebuilds beyond the interval
        p->rebuild_interval = 10;
        break;
    case 2:
        p->steric_enable = 1;
        p->steric_mode = BS_STERIC_LINEAR;
        p->steric_cutoff = 4.0f;
        p->imp_enable = 1;
        p->list_skin = 1.0f;
        p->graph_steps = 16; // a short, launch-bound step: replayed from a graph
        break;
    default:
        break;
    }
}

bsg_system * bsg_generate(int config, uint64_t seed, int64_t n)
{
    auto * s = new System();
    Rng rng(seed);
    switch (config)
    {
    case 0:
    case 3:
    {
        if (n <= 0)
            n = 2000;
        const double cut = config == 3 ? rng.uniform(9.0, 14.0) : 12.0;
        const double R = 30.0 * std::cbrt((double)n / 2000.0);
        chain_in_sphere(*s, n, 3.8, R, 3.0, rng);
        for (int64_t i = 0; i < n; ++i)
            if (rng.uniform() < 0.2)
                s->charge[i] = rng.uniform() < 0.5 ? -1.0f : 1.0f;
        cutoff_springs(*s, cut, rng, 0, 0.1);
        break;
    }
    case 1:
    case 4:
    {
        if (n <= 0)
            n = config == 4 ? 100000 : 250000;
        const double L = 140.0 * std::cbrt((double)n / 250000.0);
        uniform_beads(*s, n, L, 1.2, rng);
        for (int64_t i = 0; i < n; ++i)
        {
            s->radius[i] = (float)rng.uniform(3.0, 4.0);
            s->eps[i] = (float)rng.uniform(0.1, 1.0);
            s->charge[i] = (float)(0.4 * rng.normal());
        }
        cutoff_springs(*s, config == 4 ? 10.0 : 8.0, rng, 1, 0.5);
        break;
    }
    case 2:
    {
        if (n <= 0)
            n = 50000;
        // CG beads: residue centres uniform at CG density, 4 beads per residue
        // within 2.5 A of the centre.
        const int64_t nres = (n + 3) / 4;
        const double L = 140.0 * std::cbrt((double)nres * 4 / 250000.0 * 6.0);
        System c;
        uniform_beads(c, nres, L, 4.0, rng);
        s->resize(n);
        for (int64_t i = 0; i < n; ++i)
        {
            const int64_t r = i / 4;
            for (int ax = 0; ax < 3; ++ax)
                s->x[3 * i + ax] = (float)(c.x[3 * r + ax] - 0.5 * L + rng.uniform(-1.25, 1.25));
            s->sasa[i] = (float)rng.uniform(10.0, 100.0);
            s->transfer[i] = (float)rng.uniform(-1.0, 1.0);
        }
        cutoff_springs(*s, 10.0, rng, 0, 0.5);
        break;
    }
    default:
        delete s;
        return nullptr;
    }
    return reinterpret_cast<bsg_system *>(s);
}

int64_t bsg_num_particles(const bsg_system * h) { return reinterpret_cast<const System *>(h)->n(); }
int64_t bsg_num_springs(const bsg_system * h) { return (int64_t)reinterpret_cast<const System *>(h)->id1.size(); }

// Copies the system out; every pointer is nullable.
void bsg_copy(const bsg_system * h, float * x3, float * v3, float * mass, float * charge, float * radius,
              float * eps, float * sasa, float * transfer, float * hydro, uint8_t * flags, uint32_t * id1,
              uint32_t * id2, float * l0, float * k)
{
    const System & s = *reinterpret_cast<const System *>(h);
    auto cp = [](void * dst, const void * src, size_t bytes) {
        if (dst && bytes)
            std::memcpy(dst, src, bytes);
    };
    const size_t n = (size_t)s.n(), m = s.id1.size();
    cp(x3, s.x.data(), 12 * n);
    cp(v3, s.v.data(), 12 * n);
    cp(mass, s.mass.data(), 4 * n);
    cp(charge, s.charge.data(), 4 * n);
    cp(radius, s.radius.data(), 4 * n);
    cp(eps, s.eps.data(), 4 * n);
    cp(sasa, s.sasa.data(), 4 * n);
    cp(transfer, s.transfer.data(), 4 * n);
    cp(hydro, s.hydro.data(), 4 * n);
    cp(flags, s.flags.data(), n);
    cp(id1, s.id1.data(), 4 * m);
    cp(id2, s.id2.data(), 4 * m);
    cp(l0, s.l0.data(), 4 * m);
    cp(k, s.k.data(), 4 * m);
}

void bsg_free(bsg_system * h) { delete reinterpret_cast<System *>(h); }

} // extern "C"
