// NetCDF topology reader (see bs_netcdf.hpp).
//
// The classic file layout (header: magic, record count, dimension list,
// global attributes, variable list with each variable's type, size and
// offset; data big-endian at those offsets, record variables interleaved per
// record) is parsed here directly.  What is read, the checks and their
// messages follow NetCDFReader::readParticles / readSprings /
// addParticlesToSpn / addSpringsToSpn (src/IO/NetCDFReader.cpp:8-280).
#include "bs_netcdf.hpp"

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <set>
#include <utility>

namespace bs
{
namespace
{

const size_t CHAIN_LEN = 4, RESNAME_LEN = 4, NAME_LEN = 5; // SpnBuffer.h:10-12

enum NcType
{
    NC_BYTE = 1,
    NC_CHAR = 2,
    NC_SHORT = 3,
    NC_INT = 4,
    NC_FLOAT = 5,
    NC_DOUBLE = 6,
    NC_UBYTE = 7,
    NC_USHORT = 8,
    NC_UINT = 9,
    NC_INT64 = 10,
    NC_UINT64 = 11,
};

int type_size(int t)
{
    switch (t)
    {
    case NC_BYTE: case NC_CHAR: case NC_UBYTE: return 1;
    case NC_SHORT: case NC_USHORT: return 2;
    case NC_INT: case NC_FLOAT: case NC_UINT: return 4;
    case NC_DOUBLE: case NC_INT64: case NC_UINT64: return 8;
    default: return 0;
    }
}

std::string format(const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    return buf;
}

struct Var
{
    std::string name;
    std::vector<uint64_t> dimids;
    int type = 0;
    uint64_t vsize = 0, begin = 0;
};

class NcFile
{
  public:
    std::vector<uint8_t> buf;
    int version = 0;
    uint64_t numrecs = 0, recsize = 0;
    std::vector<std::pair<std::string, uint64_t>> dims;
    int64_t recdim = -1;
    std::vector<Var> vars;
    std::string err;

    bool parse()
    {
        if (buf.size() >= 8 && std::memcmp(buf.data(), "\x89HDF\r\n\x1a\n", 8) == 0)
            return fail("NetCDF-4 (HDF5) files are not read here; convert it to the classic format "
                        "(nccopy -k classic)");
        if (buf.size() < 4 || buf[0] != 'C' || buf[1] != 'D' || buf[2] != 'F' ||
            (buf[3] != 1 && buf[3] != 2 && buf[3] != 5))
            return fail("not a NetCDF classic file");
        version = buf[3];
        pos_ = 4;
        numrecs = count();
        if ((version != 5 && numrecs == 0xffffffffull) || (version == 5 && numrecs == ~0ull))
            return fail("NetCDF file with an unknown record count (streaming)");
        // dimensions
        uint32_t tag = u32();
        uint64_t nd = count();
        if (tag == 0x0A)
            for (uint64_t d = 0; d < nd && ok_; ++d)
            {
                std::string nm = name();
                uint64_t len = count();
                if (len == 0)
                {
                    recdim = (int64_t)d;
                    len = numrecs;
                }
                dims.emplace_back(nm, len);
            }
        else if (tag != 0 || nd != 0)
            return fail("malformed NetCDF header (dimension list)");
        if (!attributes())
            return false;
        tag = u32();
        uint64_t nv = count();
        if (tag == 0x0B)
            for (uint64_t v = 0; v < nv && ok_; ++v)
            {
                Var var;
                var.name = name();
                const uint64_t ndims = count();
                for (uint64_t k = 0; k < ndims && ok_; ++k)
                {
                    const uint64_t id = count();
                    if (id >= dims.size())
                        return fail("malformed NetCDF header (dimension id)");
                    var.dimids.push_back(id);
                }
                if (!attributes())
                    return false;
                var.type = (int)u32();
                var.vsize = count();
                var.begin = version == 1 ? u32() : u64();
                if (type_size(var.type) == 0)
                    return fail(format("variable \"%s\": unknown NetCDF type %d", var.name.c_str(), var.type));
                vars.push_back(var);
            }
        else if (tag != 0 || nv != 0)
            return fail("malformed NetCDF header (variable list)");
        if (!ok_)
            return fail("truncated NetCDF header");
        // record size: the record variables' padded sizes, except a single
        // record variable of 1- or 2-byte type, stored unpadded
        int nrec = 0;
        const Var *only = nullptr;
        for (const Var &v : vars)
            if (is_record(v))
            {
                ++nrec;
                only = &v;
                if (__builtin_add_overflow(recsize, v.vsize, &recsize))
                    return fail("malformed NetCDF header (record size)");
            }
        if (nrec == 1 && type_size(only->type) < 4)
        {
            recsize = type_size(only->type);
            for (size_t k = 1; k < only->dimids.size(); ++k)
                if (__builtin_mul_overflow(recsize, dims[only->dimids[k]].second, &recsize))
                    return fail("malformed NetCDF header (record size)");
        }
        return true;
    }

    const Var *var(const char *nm) const
    {
        for (const Var &v : vars)
            if (v.name == nm)
                return &v;
        return nullptr;
    }
    bool has_dim(const char *nm, uint64_t &len) const
    {
        for (const auto &d : dims)
            if (d.first == nm)
            {
                len = d.second;
                return true;
            }
        return false;
    }
    uint64_t dim_len(const Var &v, size_t k) const { return dims[v.dimids[k]].second; }

    // All elements of v, converted to double, in row-major order.
    bool values(const Var &v, std::vector<double> &out)
    {
        const int ts = type_size(v.type);
        // every element takes at least one byte of the file: larger shapes
        // are a damaged header, not something to allocate
        const std::string big = format("variable \"%s\": shape larger than the file", v.name.c_str());
        uint64_t per = 1; // elements per record (or in all, for a fixed variable)
        for (size_t k = is_record(v) ? 1 : 0; k < v.dimids.size(); ++k)
        {
            const uint64_t d = dim_len(v, k);
            if (d != 0 && per > buf.size() / d)
                return fail(big);
            per *= d;
        }
        const uint64_t nrec = is_record(v) ? numrecs : 1;
        if (nrec != 0 && per > buf.size() / nrec)
            return fail(big);
        // per * ts <= 8 * size (both factors bounded above) cannot wrap; the
        // record offsets come from the file, so every product and sum is
        // checked before it is used
        const uint64_t size = buf.size(), bytes = per * (uint64_t)ts;
        const std::string beyond = format("variable \"%s\": data beyond the end of the file", v.name.c_str());
        if (bytes > size)
            return fail(beyond);
        out.resize((size_t)(per * nrec));
        for (uint64_t r = 0; r < nrec; ++r)
        {
            uint64_t off = 0, at = 0;
            if (__builtin_mul_overflow(r, is_record(v) ? recsize : 0, &off) ||
                __builtin_add_overflow(v.begin, off, &at) || at > size || bytes > size - at)
                return fail(beyond);
            const uint8_t *p = buf.data() + at;
            for (uint64_t e = 0; e < per; ++e, p += ts)
                out[(size_t)(r * per + e)] = decode(v.type, p);
        }
        return true;
    }

    bool chars(const Var &v, std::vector<char> &out)
    {
        if (v.type != NC_CHAR)
            return fail(format("variable \"%s\" is not a character array", v.name.c_str()));
        std::vector<double> tmp;
        if (!values(v, tmp))
            return false;
        out.resize(tmp.size());
        for (size_t i = 0; i < tmp.size(); ++i)
            out[i] = (char)(int)tmp[i];
        return true;
    }

    bool fail(const std::string &m)
    {
        if (err.empty())
            err = m;
        return false;
    }

  private:
    size_t pos_ = 0;
    bool ok_ = true;

    bool is_record(const Var &v) const { return !v.dimids.empty() && (int64_t)v.dimids[0] == recdim; }
    uint32_t u32()
    {
        if (pos_ + 4 > buf.size())
        {
            ok_ = false;
            return 0;
        }
        const uint8_t *p = buf.data() + pos_;
        pos_ += 4;
        return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3];
    }
    uint64_t u64()
    {
        const uint64_t hi = u32();
        return hi << 32 | u32();
    }
    uint64_t count() { return version == 5 ? u64() : u32(); }
    std::string name()
    {
        const uint64_t len = count();
        if (!ok_ || len > buf.size() - pos_)
        {
            ok_ = false;
            return {};
        }
        std::string s((const char *)buf.data() + pos_, (size_t)len);
        pos_ += (size_t)((len + 3) / 4 * 4);
        return s;
    }
    bool attributes()
    {
        const uint32_t tag = u32();
        const uint64_t na = count();
        if (tag == 0 && na == 0)
            return true;
        if (tag != 0x0C)
            return fail("malformed NetCDF header (attribute list)");
        for (uint64_t a = 0; a < na && ok_; ++a)
        {
            name();
            const int t = (int)u32();
            const uint64_t ne = count();
            uint64_t bytes = 0;
            if (!ok_ || type_size(t) == 0 || __builtin_mul_overflow(ne, (uint64_t)type_size(t), &bytes) ||
                bytes > buf.size() - pos_)
                return fail("malformed NetCDF header (attribute)");
            pos_ += (size_t)((bytes + 3) / 4 * 4);
        }
        return ok_ || fail("truncated NetCDF header");
    }
    static double decode(int t, const uint8_t *p)
    {
        uint64_t u = 0;
        for (int b = 0; b < type_size(t); ++b)
            u = u << 8 | p[b];
        switch (t)
        {
        case NC_BYTE: return (double)(int8_t)u;
        case NC_CHAR: return (double)(uint8_t)u;
        case NC_SHORT: return (double)(int16_t)u;
        case NC_INT: return (double)(int32_t)u;
        case NC_FLOAT:
        {
            const uint32_t w = (uint32_t)u;
            float f;
            std::memcpy(&f, &w, 4);
            return f;
        }
        case NC_DOUBLE:
        {
            double d;
            std::memcpy(&d, &u, 8);
            return d;
        }
        case NC_UBYTE: return (double)(uint8_t)u;
        case NC_USHORT: return (double)(uint16_t)u;
        case NC_UINT: return (double)(uint32_t)u;
        case NC_INT64: return (double)(int64_t)u;
        default: return (double)u;
        }
    }
};

// NetCDFReader::getNcVar + checkNDims + checkDim (NetCDFReader.cpp:248-280).
const Var *checked(NcFile &f, const char *nm, bool mandatory, std::vector<uint64_t> shape, std::string &warn)
{
    const Var *v = f.var(nm);
    if (!v)
    {
        if (mandatory)
            f.fail(format("Array \"%s\" not found in topology file.", nm));
        else
            warn = format("Array \"%s\" not found in topology file.", nm);
        return nullptr;
    }
    if (v->dimids.size() != shape.size())
    {
        f.fail(format("\"%s\" array has %d dimension(s) whereas it should be %d.", nm, (int)v->dimids.size(),
                      (int)shape.size()));
        return nullptr;
    }
    for (size_t k = 0; k < shape.size(); ++k)
        if (f.dim_len(*v, k) != shape[k])
        {
            f.fail(format("Dimension %d of array \"%s\" has size %d whereas it should be %d.", (int)k, nm,
                          (int)f.dim_len(*v, k), (int)shape[k]));
            return nullptr;
        }
    return v;
}

template <class T> bool numbers(NcFile &f, const Var *v, std::vector<T> &out)
{
    std::vector<double> d;
    if (!f.values(*v, d))
        return false;
    out.resize(d.size());
    for (size_t i = 0; i < d.size(); ++i)
        out[i] = (T)d[i];
    return true;
}

// strncpy into a zeroed buffer of len characters (NetCDFReader.cpp:40-50).
std::vector<std::string> names(const std::vector<char> &c, size_t n, size_t len)
{
    std::vector<std::string> out(n);
    for (size_t i = 0; i < n; ++i)
    {
        const char *s = c.data() + i * len;
        out[i] = std::string(s, strnlen(s, len));
    }
    return out;
}

// ---------------------------------------------------------------------------
// Writing (CDF-2: the classic layout with 64-bit variable offsets).
// ---------------------------------------------------------------------------
struct OutVar
{
    std::string name;
    std::vector<int> dims;
    int type;
    std::vector<std::pair<std::string, std::string>> atts; // text attributes
    std::vector<uint8_t> data;                              // big-endian, unpadded
};

void be32(std::vector<uint8_t> &o, uint32_t v)
{
    for (int s = 24; s >= 0; s -= 8)
        o.push_back((uint8_t)(v >> s));
}
void be64(std::vector<uint8_t> &o, uint64_t v)
{
    be32(o, (uint32_t)(v >> 32));
    be32(o, (uint32_t)v);
}
void pad4(std::vector<uint8_t> &o)
{
    while (o.size() % 4)
        o.push_back(0);
}
void nc_name(std::vector<uint8_t> &o, const std::string &s)
{
    be32(o, (uint32_t)s.size());
    o.insert(o.end(), s.begin(), s.end());
    pad4(o);
}
template <class T> std::vector<uint8_t> be_values(const T *v, size_t n)
{
    std::vector<uint8_t> o;
    o.reserve(n * sizeof(T));
    for (size_t i = 0; i < n; ++i)
    {
        uint64_t u = 0;
        std::memcpy(&u, &v[i], sizeof(T));
        for (int b = (int)sizeof(T) - 1; b >= 0; --b)
            o.push_back((uint8_t)(u >> (8 * b)));
    }
    return o;
}
std::vector<uint8_t> fixed_chars(const std::vector<std::string> &names, size_t n, size_t len)
{
    std::vector<uint8_t> o(n * len, 0);
    for (size_t i = 0; i < n && i < names.size(); ++i)
        std::memcpy(o.data() + i * len, names[i].data(), std::min(len, names[i].size())); // strncpy
    return o;
}

} // namespace

bool write_netcdf_topology(const std::string &path, const NcTopology &t, std::string &err)
{
    const size_t n = (size_t)t.n, m = t.s1.size();
    std::vector<std::pair<std::string, uint64_t>> dims = {{"spatialdim", 3},
                                                          {"particle_number", n},
                                                          {"particlename_length", NAME_LEN},
                                                          {"chainname_length", CHAIN_LEN},
                                                          {"resname_length", RESNAME_LEN}};
    enum
    {
        D_SPATIAL,
        D_PNB,
        D_PNAME,
        D_CHAIN,
        D_RESNAME,
        D_SPRINGDIM,
        D_SNB
    };
    std::vector<OutVar> vars;
    auto add = [&](const char *nm, std::vector<int> d, int type, std::vector<std::pair<std::string, std::string>> atts,
                   std::vector<uint8_t> data) { vars.push_back(OutVar{nm, std::move(d), type, std::move(atts), std::move(data)}); };
    std::vector<int32_t> ids(n);
    std::vector<int8_t> dyn8(n);
    for (size_t i = 0; i < n; ++i)
    {
        ids[i] = (int32_t)i;
        dyn8[i] = t.is_static[i] ? 0 : 1;
    }
    std::vector<int32_t> resid(t.resid);
    resid.resize(n, 0);
    add("coordinates", {D_PNB, D_SPATIAL}, NC_FLOAT, {{"units", "angstrom"}, {"long_name", "Particle coordinates"}},
        be_values(t.x.data(), 3 * n));
    add("particleids", {D_PNB}, NC_INT, {{"long_name", "Particle ids in source database"}}, be_values(ids.data(), n));
    add("charges", {D_PNB}, NC_FLOAT, {{"units", "electron"}, {"long_name", "Particle charge id"}},
        be_values(t.charge.data(), n));
    add("radii", {D_PNB}, NC_FLOAT, {{"units", "A"}, {"long_name", "Particle radius"}}, be_values(t.radius.data(), n));
    add("epsilon", {D_PNB}, NC_FLOAT, {{"units", "kJ.mol-1"}, {"long_name", "Particle epsilon for Lennard-Jones"}},
        be_values(t.epsilon.data(), n));
    add("mass", {D_PNB}, NC_FLOAT, {{"units", "Da"}, {"long_name", "Particle mass"}}, be_values(t.mass.data(), n));
    add("surfaceaccessibility", {D_PNB}, NC_FLOAT,
        {{"units", "A2 or percent"}, {"long_name", "Particle surface accessibility"}}, be_values(t.sasa.data(), n));
    add("hydrophobicityscale", {D_PNB}, NC_FLOAT,
        {{"units", "kJ.mol-1"}, {"long_name", "Particle hydrophobicity scale (transfer energy)"}},
        be_values(t.transfer.data(), n));
    add("hydrophobicity", {D_PNB}, NC_FLOAT, {{"long_name", "Particle hydrophobicity (pairwise hydrophobic force)"}},
        be_values(t.hydro.data(), n));
    add("resids", {D_PNB}, NC_INT, {{"long_name", "Particle residue id"}}, be_values(resid.data(), n));
    add("dynamicstate", {D_PNB}, NC_BYTE, {{"long_name", "Particle dynamic state (static 0, dynamic 1)"}},
        be_values(dyn8.data(), n));
    add("particlenames", {D_PNB, D_PNAME}, NC_CHAR, {{"long_name", "Particle name"}}, fixed_chars(t.name, n, NAME_LEN));
    add("resnames", {D_PNB, D_RESNAME}, NC_CHAR, {{"long_name", "Particle residue name"}},
        fixed_chars(t.resname, n, RESNAME_LEN));
    add("chainnames", {D_PNB, D_CHAIN}, NC_CHAR, {{"long_name", "Particle chain name"}},
        fixed_chars(t.chain, n, CHAIN_LEN));
    if (m > 0)
    {
        dims.emplace_back("springdim", 2);
        dims.emplace_back("spring_number", m);
        std::vector<int32_t> ij(2 * m), per(n, 0);
        for (size_t s = 0; s < m; ++s)
        {
            ij[2 * s] = (int32_t)t.s1[s];
            ij[2 * s + 1] = (int32_t)t.s2[s];
            ++per[t.s1[s]];
            ++per[t.s2[s]];
        }
        add("springs", {D_SNB, D_SPRINGDIM}, NC_INT,
            {{"long_name", "Spring between particles referenced by 2 particle ids"}}, be_values(ij.data(), 2 * m));
        add("springsstiffness", {D_SNB}, NC_FLOAT, {{"long_name", "Spring stiffness"}}, be_values(t.k.data(), m));
        add("springsequilibrium", {D_SNB}, NC_FLOAT, {{"long_name", "Spring distance equilibrium"}},
            be_values(t.l0.data(), m));
        add("nbspringsperparticle", {D_PNB}, NC_INT, {{"long_name", "Number of springs per particle"}},
            be_values(per.data(), n));
    }
    // header with placeholder offsets, then the data
    auto header = [&](const std::vector<uint64_t> &begins) {
        std::vector<uint8_t> h = {'C', 'D', 'F', 2};
        be32(h, 0); // numrecs: no record dimension
        be32(h, 0x0A);
        be32(h, (uint32_t)dims.size());
        for (const auto &d : dims)
        {
            nc_name(h, d.first);
            be32(h, (uint32_t)d.second);
        }
        be32(h, 0); // no global attributes
        be32(h, 0);
        be32(h, 0x0B);
        be32(h, (uint32_t)vars.size());
        for (size_t v = 0; v < vars.size(); ++v)
        {
            const OutVar &x = vars[v];
            nc_name(h, x.name);
            be32(h, (uint32_t)x.dims.size());
            for (int d : x.dims)
                be32(h, (uint32_t)d);
            be32(h, 0x0C);
            be32(h, (uint32_t)x.atts.size());
            for (const auto &a : x.atts)
            {
                nc_name(h, a.first);
                be32(h, NC_CHAR);
                be32(h, (uint32_t)a.second.size());
                h.insert(h.end(), a.second.begin(), a.second.end());
                pad4(h);
            }
            be32(h, (uint32_t)x.type);
            be32(h, (uint32_t)std::min<uint64_t>((x.data.size() + 3) / 4 * 4, 0xffffffffu)); // vsize
            be64(h, begins[v]);
        }
        return h;
    };
    std::vector<uint64_t> begins(vars.size(), 0);
    uint64_t at = header(begins).size();
    for (size_t v = 0; v < vars.size(); ++v)
    {
        begins[v] = at;
        at += (vars[v].data.size() + 3) / 4 * 4;
    }
    const std::vector<uint8_t> h = header(begins);
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f)
    {
        err = "cannot open output file '" + path + "': " + std::strerror(errno);
        return false;
    }
    bool ok = std::fwrite(h.data(), 1, h.size(), f) == h.size();
    static const uint8_t zeros[4] = {0, 0, 0, 0};
    for (const OutVar &x : vars)
    {
        ok = ok && std::fwrite(x.data.data(), 1, x.data.size(), f) == x.data.size();
        const size_t padn = (4 - x.data.size() % 4) % 4;
        ok = ok && std::fwrite(zeros, 1, padn, f) == padn;
    }
    ok = (std::fclose(f) == 0) && ok;
    if (!ok)
        err = "cannot write " + path;
    return ok;
}

bool read_netcdf_topology(const std::string &path, NcTopology &t, std::string &err)
{
    NcFile f;
    {
        std::ifstream in(path, std::ios::binary);
        if (!in)
        {
            err = path + " not found.";
            return false;
        }
        f.buf.assign(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
    }
    if (!f.parse())
    {
        err = f.err;
        return false;
    }
    // readNumberOfParticles (NetCDFReader.cpp:212-228)
    uint64_t n = 0;
    if (!f.has_dim("particle_number", n))
    {
        err = "Dimension \"particle_number\" not present in nc file.";
        return false;
    }
    if (n == 0)
    {
        err = "No particle found in topology";
        return false;
    }
    t = NcTopology{};
    t.n = (int64_t)n;
    std::string w;
    std::vector<char> cbuf;
    auto req = [&](const char *nm, std::vector<uint64_t> shape) { return checked(f, nm, true, shape, w); };
    // readParticles (NetCDFReader.cpp:89-176), in its order
    const Var *v;
    if (!(v = req("particleids", {n})) || !numbers(f, v, t.atom_id))
        return err = f.err, false;
    if (!(v = req("resids", {n})) || !numbers(f, v, t.resid))
        return err = f.err, false;
    if (!(v = req("resnames", {n, RESNAME_LEN})) || !f.chars(*v, cbuf))
        return err = f.err, false;
    t.resname = names(cbuf, n, RESNAME_LEN);
    if (!(v = req("particlenames", {n, NAME_LEN})) || !f.chars(*v, cbuf))
        return err = f.err, false;
    t.name = names(cbuf, n, NAME_LEN);
    if (!(v = req("chainnames", {n, CHAIN_LEN})) || !f.chars(*v, cbuf))
        return err = f.err, false;
    t.chain = names(cbuf, n, CHAIN_LEN);
    std::vector<double> dyn;
    if (!(v = req("dynamicstate", {n})) || !f.values(*v, dyn))
        return err = f.err, false;
    t.is_static.resize(n);
    for (uint64_t i = 0; i < n; ++i)
        t.is_static[i] = (uint8_t)dyn[i] == 0 ? 1 : 0; // read as bool / unsigned char
    if (!(v = req("coordinates", {n, 3})) || !numbers(f, v, t.x))
        return err = f.err, false;
    if (!(v = req("radii", {n})) || !numbers(f, v, t.radius))
        return err = f.err, false;
    if (!(v = req("mass", {n})) || !numbers(f, v, t.mass))
        return err = f.err, false;
    if (!(v = req("epsilon", {n})) || !numbers(f, v, t.epsilon))
        return err = f.err, false;
    if (!(v = req("charges", {n})) || !numbers(f, v, t.charge))
        return err = f.err, false;
    struct Opt
    {
        const char *nm;
        std::vector<float> *dst;
    } opts[] = {{"surfaceaccessibility", &t.sasa}, {"hydrophobicityscale", &t.transfer}, {"hydrophobicity", &t.hydro}};
    for (const Opt &o : opts)
    {
        w.clear();
        v = checked(f, o.nm, false, {n}, w);
        if (!v && !f.err.empty())
            return err = f.err, false;
        if (v && !numbers(f, v, *o.dst))
            return err = f.err, false;
        if (!v)
        {
            o.dst->assign(n, 0.0f);
            t.warnings.push_back(w);
        }
    }
    // addParticlesToSpn: a zero mass becomes 1 (NetCDFReader.cpp:34-38)
    for (uint64_t i = 0; i < n; ++i)
        if (std::fabs(t.mass[i]) < 1e-6f)
        {
            t.warnings.push_back(format("Mass of particle %d is 0.0. Changing it to 1.0", (int)i));
            t.mass[i] = 1.0f;
        }
    // readSprings (NetCDFReader.cpp:178-210, 230-246)
    uint64_t m = 0;
    if (!f.has_dim("spring_number", m))
        t.warnings.push_back("Dimension \"spring_number\" not present in nc file.");
    else if (m == 0)
        t.warnings.push_back("No spring found in topology");
    if (m > 0)
    {
        std::vector<int64_t> ij;
        if (!(v = req("springs", {m, 2})) || !numbers(f, v, ij))
            return err = f.err, false;
        if (!(v = req("springsstiffness", {m})) || !numbers(f, v, t.k))
            return err = f.err, false;
        if (!(v = req("springsequilibrium", {m})) || !numbers(f, v, t.l0))
            return err = f.err, false;
        // addSpringsToSpn -> Topology::add_spring -> SpringCollection::add_spring
        // (SpringCollection.hpp:138-160): unknown particles, self springs and
        // repeated pairs are errors there
        std::set<std::pair<uint32_t, uint32_t>> seen;
        for (uint64_t s = 0; s < m; ++s)
        {
            const int64_t a = ij[2 * s], b = ij[2 * s + 1];
            if (a < 0 || b < 0 || (uint64_t)a >= n || (uint64_t)b >= n)
            {
                err = format("spring %d: particle index out of range", (int)s);
                return false;
            }
            if (a == b)
            {
                err = "Cannot add a spring between a particle and itself.";
                return false;
            }
            const auto key = std::make_pair((uint32_t)std::min(a, b), (uint32_t)std::max(a, b));
            if (!seen.insert(key).second)
            {
                err = "Cannot add a spring that already exists.";
                return false;
            }
            t.s1.push_back((uint32_t)a);
            t.s2.push_back((uint32_t)b);
        }
    }
    return true;
}

} // namespace bs
