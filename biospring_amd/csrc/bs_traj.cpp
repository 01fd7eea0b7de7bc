// Trajectory writers and their frame queue (see bs_traj.hpp).
//
// Formats follow the reference's writers field for field:
//   PDB  MODEL/ATOM/CONECT/ENDMDL records (PDBTrajectoryWriter.cpp:10-33,
//        pdb_format.hpp:17-58); CONECT records in the first model only.
//   XTC  positions / 10 (A to nm), step = time = the writer's frame count, a
//        unit box and precision 1000 (XTCTrajectoryWriter.cpp:13-31), in the
//        xdrfile frame layout and coordinate compression (xdrfile_xtc.c:22-72,
//        xdrfile.c:470-700,965-1256).
//   CSV  the tab-separated energy table (CSVTrajectoryWriter.cpp:11-60), the
//        values printed as the reference's float getters through an ostream
//        (6 significant digits).
#include "bs_traj.hpp"

#include <algorithm>
#include <cerrno>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <exception>

namespace bs
{
namespace
{

// ---------------------------------------------------------------------------
// XDR primitives: big-endian 4-byte ints and floats, opaque data padded to 4.
// ---------------------------------------------------------------------------
void put_u32(std::vector<uint8_t> &o, uint32_t v)
{
    o.push_back((uint8_t)(v >> 24));
    o.push_back((uint8_t)(v >> 16));
    o.push_back((uint8_t)(v >> 8));
    o.push_back((uint8_t)v);
}
void put_i32(std::vector<uint8_t> &o, int32_t v) { put_u32(o, (uint32_t)v); }
void put_f32(std::vector<uint8_t> &o, float v)
{
    uint32_t u;
    std::memcpy(&u, &v, 4);
    put_u32(o, u);
}

// The size table of the small-integer code (xdrfile.c:744-752): entry k is
// about 2^(k/3), so three integers below entry k fit in k bits.
const int kMagic[] = {0,       0,       0,       0,       0,       0,       0,       0,       0,       8,
                      10,      12,      16,      20,      25,      32,      40,      50,      64,      80,
                      101,     128,     161,     203,     256,     322,     406,     512,     645,     812,
                      1024,    1290,    1625,    2048,    2580,    3250,    4096,    5060,    6501,    8192,
                      10321,   13003,   16384,   20642,   26007,   32768,   41285,   52015,   65536,   82570,
                      104031,  131072,  165140,  208063,  262144,  330280,  416127,  524287,  660561,  832255,
                      1048576, 1321122, 1664510, 2097152, 2642245, 3329021, 4194304, 5284491, 6658042, 8388607,
                      10568983, 13316085, 16777216};
constexpr int kFirstIdx = 9;
constexpr int kLastIdx = (int)(sizeof(kMagic) / sizeof(kMagic[0]));

int magic(int k) { return kMagic[std::min(std::max(k, 0), kLastIdx - 1)]; }

// Bits written most significant first into a byte stream.
struct BitSink
{
    std::vector<uint8_t> bytes;
    uint64_t acc = 0;
    int nacc = 0; // bits pending in acc (< 8 between calls)

    // nbits <= 32 bits of v, most significant first
    void put32(int nbits, uint32_t v)
    {
        if (nbits <= 0)
            return;
        const uint64_t m = nbits == 32 ? 0xffffffffull : ((1ull << nbits) - 1ull);
        acc = (acc << nbits) | ((uint64_t)v & m);
        nacc += nbits;
        while (nacc >= 8)
        {
            nacc -= 8;
            bytes.push_back((uint8_t)(acc >> nacc));
        }
        acc &= (1ull << nacc) - 1ull;
    }
    // any number of bits; beyond 32 the leading bits are zero (v is 32 bits)
    void put(int nbits, uint32_t v)
    {
        while (nbits > 32)
        {
            put32(32, 0u);
            nbits -= 32;
        }
        put32(nbits, v);
    }
    void finish()
    {
        if (nacc > 0)
            bytes.push_back((uint8_t)(acc << (8 - nacc)));
        acc = 0;
        nacc = 0;
    }
};

using u128 = unsigned __int128;

int bitlen(u128 v)
{
    const uint64_t hi = (uint64_t)(v >> 64), lo = (uint64_t)v;
    return hi ? 128 - __builtin_clzll(hi) : lo ? 64 - __builtin_clzll(lo) : 0;
}

// Bits of the integer that sizeofint reports: the smallest b with 2^b > size.
int bits_for(uint32_t size)
{
    int b = 0;
    uint64_t num = 1;
    while ((uint64_t)size >= num && b < 32)
    {
        ++b;
        num <<= 1;
    }
    return b;
}

// Bits of the product of three sizes (sizeofints, xdrfile.c:497-530).
int bits_for3(const uint32_t s[3]) { return bitlen((u128)s[0] * s[1] * s[2]); }

// Three integers as one mixed-radix number ((n0 s1 + n1) s2 + n2), written
// as nbits bits in little-endian byte order (encodeints, xdrfile.c:592-640).
void put_ints3(BitSink &b, int nbits, const uint32_t s[3], const uint32_t v[3])
{
    const u128 x = ((u128)v[0] * s[1] + v[1]) * s[2] + v[2];
    int nbytes = std::max(1, (bitlen(x) + 7) / 8);
    if (nbits >= nbytes * 8)
    {
        for (int k = 0; k < nbytes; ++k)
            b.put(8, (uint32_t)(x >> (8 * k)) & 0xffu);
        b.put(nbits - nbytes * 8, 0u);
    }
    else
    {
        for (int k = 0; k < nbytes - 1; ++k)
            b.put(8, (uint32_t)(x >> (8 * k)) & 0xffu);
        b.put(nbits - (nbytes - 1) * 8, (uint32_t)(x >> (8 * (nbytes - 1))) & 0xffu);
    }
}

// One coordinate to its integer at the given precision, with the float
// rounding of the reference: product in float, +-0.5 in double, back to float,
// truncated (xdrfile.c:1020-1028).
int32_t quantize(float x, float prec)
{
    const float p = x * prec;
    const float lf = (float)(x >= 0.0f ? (double)p + 0.5 : (double)p - 0.5);
    return (int32_t)lf;
}

void xtc_coords(std::vector<uint8_t> &o, int32_t n, const float *x, float prec)
{
    put_i32(o, n);
    if (n <= 9)
    { // uncompressed
        for (int64_t k = 0; k < 3 * (int64_t)n; ++k)
            put_f32(o, x[k]);
        return;
    }
    if (prec <= 0.0f)
        prec = 1000.0f;
    put_f32(o, prec);
    std::vector<int32_t> q((size_t)3 * n);
    int32_t mn[3] = {INT_MAX, INT_MAX, INT_MAX}, mx[3] = {INT_MIN, INT_MIN, INT_MIN};
    int32_t mindiff = INT_MAX;
    for (int32_t i = 0; i < n; ++i)
    {
        for (int c = 0; c < 3; ++c)
        {
            const int32_t v = quantize(x[3 * i + c], prec);
            q[3 * (size_t)i + c] = v;
            mn[c] = std::min(mn[c], v);
            mx[c] = std::max(mx[c], v);
        }
        if (i > 0)
        {
            // the difference to the previous atom (the first against 0 is not counted)
            const int32_t d = std::abs(q[3 * (size_t)i] - q[3 * (size_t)i - 3]) +
                              std::abs(q[3 * (size_t)i + 1] - q[3 * (size_t)i - 2]) +
                              std::abs(q[3 * (size_t)i + 2] - q[3 * (size_t)i - 1]);
            mindiff = std::min(mindiff, d);
        }
    }
    for (int c = 0; c < 3; ++c)
        put_i32(o, mn[c]);
    for (int c = 0; c < 3; ++c)
        put_i32(o, mx[c]);
    uint32_t sizeint[3], bitsizeint[3] = {0, 0, 0};
    for (int c = 0; c < 3; ++c)
        sizeint[c] = (uint32_t)(mx[c] - mn[c]) + 1u;
    int bitsize = 0;
    if ((sizeint[0] | sizeint[1] | sizeint[2]) > 0xffffffu)
        for (int c = 0; c < 3; ++c)
            bitsizeint[c] = (uint32_t)bits_for(sizeint[c]); // bitsize 0: each coordinate on its own
    else
        bitsize = bits_for3(sizeint);
    int smallidx = kFirstIdx;
    while (smallidx < kLastIdx && kMagic[smallidx] < mindiff)
        ++smallidx;
    put_i32(o, smallidx);
    const int maxidx = std::min(kLastIdx, smallidx + 8);
    const int minidx = maxidx - 8;
    int smaller = magic(std::max(kFirstIdx, smallidx - 1)) / 2;
    int smallnum = magic(smallidx) / 2;
    uint32_t sizesmall[3];
    sizesmall[0] = sizesmall[1] = sizesmall[2] = (uint32_t)magic(smallidx);
    const int larger = magic(maxidx) / 2;

    BitSink b;
    b.bytes.reserve((size_t)n * 4);
    int32_t prev[3] = {0, 0, 0};
    int prevrun = -1;
    int32_t i = 0;
    uint32_t small[24];
    while (i < n)
    {
        int32_t *cur = &q[3 * (size_t)i];
        int is_smaller;
        if (smallidx < maxidx && i >= 1 && std::abs(cur[0] - prev[0]) < larger &&
            std::abs(cur[1] - prev[1]) < larger && std::abs(cur[2] - prev[2]) < larger)
            is_smaller = 1;
        else if (smallidx > minidx)
            is_smaller = -1;
        else
            is_smaller = 0;
        bool is_small = false;
        if (i + 1 < n && std::abs(cur[0] - cur[3]) < smallnum && std::abs(cur[1] - cur[4]) < smallnum &&
            std::abs(cur[2] - cur[5]) < smallnum)
        {
            // the first of a close pair is sent second (water molecules)
            for (int c = 0; c < 3; ++c)
                std::swap(cur[c], cur[3 + c]);
            is_small = true;
        }
        const uint32_t full[3] = {(uint32_t)(cur[0] - mn[0]), (uint32_t)(cur[1] - mn[1]), (uint32_t)(cur[2] - mn[2])};
        if (bitsize == 0)
            for (int c = 0; c < 3; ++c)
                b.put((int)bitsizeint[c], full[c]);
        else
            put_ints3(b, bitsize, sizeint, full);
        for (int c = 0; c < 3; ++c)
            prev[c] = cur[c];
        ++i;
        cur += 3;
        int run = 0;
        if (!is_small && is_smaller == -1)
            is_smaller = 0;
        while (is_small && run < 24)
        {
            int32_t sq = 0;
            for (int c = 0; c < 3; ++c)
            {
                const int32_t d = cur[c] - prev[c];
                sq += d * d;
            }
            if (is_smaller == -1 && sq >= smaller * smaller)
                is_smaller = 0;
            for (int c = 0; c < 3; ++c)
                small[run++] = (uint32_t)(cur[c] - prev[c] + smallnum);
            for (int c = 0; c < 3; ++c)
                prev[c] = cur[c];
            ++i;
            cur += 3;
            is_small = i < n && std::abs(cur[0] - prev[0]) < smallnum && std::abs(cur[1] - prev[1]) < smallnum &&
                       std::abs(cur[2] - prev[2]) < smallnum;
        }
        if (run != prevrun || is_smaller != 0)
        {
            prevrun = run;
            b.put(1, 1u);
            b.put(5, (uint32_t)(run + is_smaller + 1));
        }
        else
        {
            b.put(1, 0u);
        }
        for (int k = 0; k < run; k += 3)
            put_ints3(b, smallidx, sizesmall, &small[k]);
        if (is_smaller != 0)
        {
            smallidx += is_smaller;
            if (is_smaller < 0)
            {
                smallnum = smaller;
                smaller = magic(smallidx - 1) / 2;
            }
            else
            {
                smaller = smallnum;
                smallnum = magic(smallidx) / 2;
            }
            sizesmall[0] = sizesmall[1] = sizesmall[2] = (uint32_t)magic(smallidx);
        }
    }
    b.finish();
    put_i32(o, (int32_t)b.bytes.size());
    o.insert(o.end(), b.bytes.begin(), b.bytes.end());
    while (o.size() % 4 != 0)
        o.push_back(0);
}

// ---------------------------------------------------------------------------
// Writers.
// ---------------------------------------------------------------------------
class FileWriter : public TrajWriter
{
  public:
    FILE *f = nullptr;
    std::shared_ptr<const TrajContext> ctx;
    ~FileWriter() override
    {
        if (f)
            std::fclose(f);
    }
    bool emit(const std::string &s) override
    {
        if (!s.empty() && std::fwrite(s.data(), 1, s.size(), f) != s.size())
            return false;
        return std::fflush(f) == 0;
    }
};

void appendf(std::string &s, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void appendf(std::string &s, const char *fmt, ...)
{
    char buf[256];
    va_list ap;
    va_start(ap, fmt);
    const int k = std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (k > 0)
        s.append(buf, (size_t)std::min(k, (int)sizeof buf - 1));
}

template <class T> const T &at_or(const std::vector<T> &v, int64_t i, const T &dflt)
{
    return (size_t)i < v.size() ? v[(size_t)i] : dflt;
}

class PdbWriter : public FileWriter
{
  public:
    bool encode(const TrajFrame &fr, int64_t index, std::string &s) override
    {
        const TrajContext &c = *ctx;
        const TrajLabels &L = c.labels;
        static const std::string empty;
        s.clear();
        s.reserve((size_t)(c.n + 2) * 81);
        s += "MODEL    " + std::to_string(index) + "\n";
        for (int64_t i = 0; i < c.n; ++i)
        {
            const std::string &nm = at_or(L.name, i, empty);
            char name[8];
            if (nm.size() < 4)
                std::snprintf(name, sizeof name, " %-3.3s", nm.c_str());
            else
                std::snprintf(name, sizeof name, "%-4.4s", nm.c_str());
            const int32_t q = at_or(L.echarge, i, at_or(c.echarge_default, i, 0));
            char charge[16] = "";
            if (q > 0)
                std::snprintf(charge, sizeof charge, "%+d", q);
            else if (q < 0)
                std::snprintf(charge, sizeof charge, "%d", q);
            appendf(s, "%-6s%5d %4s%1s%3s %1s%4d%1s   %8.3f%8.3f%8.3f%6.2f%6.2f          %2s%2s\n", "ATOM",
                    (int)(i + 1), name, " ", at_or(L.resname, i, empty).c_str(), at_or(L.chain, i, empty).c_str(),
                    (int)at_or(L.resid, i, 0), " ", (double)fr.x3[3 * i], (double)fr.x3[3 * i + 1],
                    (double)fr.x3[3 * i + 2], (double)at_or(L.occupancy, i, 0.0f),
                    (double)at_or(L.tempfactor, i, 0.0f), at_or(L.element, i, empty).c_str(), charge);
        }
        if (index == 0)
            for (const auto &sp : c.conect)
                appendf(s, "CONECT%5d%5d\n", (int)sp.first + 1, (int)sp.second + 1);
        s += "ENDMDL\n";
        return true;
    }
};

class XtcWriter : public FileWriter
{
  public:
    bool encode(const TrajFrame &fr, int64_t index, std::string &s) override
    {
        const int64_t n = ctx->n;
        std::vector<float> nm((size_t)(3 * n));
        for (int64_t k = 0; k < 3 * n; ++k)
            nm[(size_t)k] = (float)((double)fr.x3[k] / 10.0);
        const float box[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        std::vector<uint8_t> out;
        xtc_encode_frame(out, (int32_t)n, (int32_t)index, (float)index, box, nm.data(), 1000.0f);
        s.assign(reinterpret_cast<const char *>(out.data()), out.size());
        return true;
    }
};

void put_g(std::string &s, float v) { appendf(s, "%g", (double)v); }

class CsvWriter : public FileWriter
{
  public:
    bool header()
    {
        const TrajContext &c = *ctx;
        std::string s = "# Step\tFrameRate(Hz)\tKinetic energy (kJ.mol-1)";
        if (c.spring)
            s += "\tSpring energy (kJ.mol-1)";
        if (c.steric)
            s += "\tSteric energy (kJ.mol-1)";
        if (c.elec)
            s += "\tElectrostatic energy (kJ.mol-1)";
        if (c.imp)
            s += "\tIMP energy (kJ.mol-1)";
        if (c.hydro)
            s += "\tHydrophobic energy (kJ.mol-1)";
        if (c.ivec1 >= 0)
            s += "\tInsertion Angle (degrees) \tInsertion Depth (A)";
        s += "\n";
        return emit(s);
    }
    // a row holds the frame rate, known on the ordered path only
    bool ordered() const override { return true; }
    bool encode(const TrajFrame &fr, int64_t, std::string &s) override
    {
        const TrajContext &c = *ctx;
        s = std::to_string(fr.nbiter) + "\t";
        put_g(s, (float)fr.framerate);
        s += "\t";
        put_g(s, (float)fr.energies[3]);
        // columns in the reference's order; energies in BS_E_* order
        // (spring, electrostatic, steric, kinetic, IMP, hydrophobic)
        const bool on[5] = {c.spring, c.steric, c.elec, c.imp, c.hydro};
        const int idx[5] = {0, 2, 1, 4, 5};
        for (int t = 0; t < 5; ++t)
            if (on[t])
            {
                s += "\t";
                put_g(s, (float)fr.energies[idx[t]]);
            }
        if (c.ivec1 >= 0)
        {
            // InsertionVector::computeAngle (InsertionVector.cpp:18-40) and the
            // depth, the centroid's z (measure.hpp:121-140), in FP64
            double cz = 0.0;
            for (int64_t i = 0; i < c.n; ++i)
                cz += fr.x3[3 * i + 2];
            cz /= (double)c.n;
            const float *p1 = fr.x3 + 3 * c.ivec1, *p2 = fr.x3 + 3 * c.ivec2;
            const double v[3] = {(double)p2[0] - p1[0], (double)p2[1] - p1[1], (double)p2[2] - p1[2]};
            const double vn = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            const double angle = std::acos(-v[2] / vn) * (180.0 / M_PI) - 90.0;
            s += "\t";
            put_g(s, (float)angle);
            s += "\t";
            put_g(s, (float)cz);
        }
        s += "\n";
        return true;
    }
};

} // namespace

void xtc_encode_frame(std::vector<uint8_t> &out, int32_t natoms, int32_t step, float time, const float box[9],
                      const float *x3nm, float precision)
{
    put_i32(out, 1995); // magic
    put_i32(out, natoms);
    put_i32(out, step);
    put_f32(out, time);
    for (int k = 0; k < 9; ++k)
        put_f32(out, box[k]);
    xtc_coords(out, natoms, x3nm, precision);
}

std::unique_ptr<TrajWriter> make_traj_writer(int kind, const std::string &path, int64_t frequency,
                                             std::shared_ptr<const TrajContext> ctx, std::string &err)
{
    std::unique_ptr<FileWriter> w;
    if (kind == TRAJ_PDB)
        w.reset(new PdbWriter());
    else if (kind == TRAJ_XTC)
        w.reset(new XtcWriter());
    else if (kind == TRAJ_CSV)
        w.reset(new CsvWriter());
    else
    {
        err = "unknown trajectory kind";
        return nullptr;
    }
    w->f = std::fopen(path.c_str(), kind == TRAJ_XTC ? "wb" : "w");
    if (!w->f)
    {
        err = "cannot open " + path + ": " + std::strerror(errno);
        return nullptr;
    }
    w->path = path;
    w->frequency = frequency;
    w->ctx = std::move(ctx);
    if (kind == TRAJ_CSV && !static_cast<CsvWriter *>(w.get())->header())
    {
        err = "cannot write " + path;
        return nullptr;
    }
    return w;
}

// ---------------------------------------------------------------------------
// Frame queue.
// ---------------------------------------------------------------------------
void TrajectorySet::add(std::unique_ptr<TrajWriter> w)
{
    drain();
    std::lock_guard<std::mutex> lk(mu_);
    writers_.push_back(std::move(w));
}
void TrajectorySet::release(int s)
{
    {
        std::lock_guard<std::mutex> lk(mu_);
        free_[(size_t)s] = true;
    }
    cv_.notify_all();
}
bool TrajectorySet::due(int64_t nb) const
{
    for (const auto &w : writers_)
        if (nb % w->frequency == 0)
            return true;
    return false;
}
int64_t TrajectorySet::next_due(int64_t nb) const
{
    int64_t best = INT64_MAX;
    for (const auto &w : writers_)
        best = std::min(best, (nb / w->frequency + 1) * w->frequency);
    return best;
}
int TrajectorySet::acquire()
{
    std::unique_lock<std::mutex> lk(mu_);
    if (!th_.joinable())
    {
        th_ = std::thread([this] { run(); });
        const int hw = (int)std::thread::hardware_concurrency();
        const int ne = std::max(1, std::min((int)free_.size(), hw / 2));
        for (int k = 0; k < ne; ++k)
            enc_.emplace_back([this] { encode_loop(); });
    }
    cv_.wait(lk, [&] { return std::find(free_.begin(), free_.end(), true) != free_.end(); });
    const int s = (int)(std::find(free_.begin(), free_.end(), true) - free_.begin());
    free_[(size_t)s] = false;
    return s;
}
void TrajectorySet::submit(int s, const TrajFrame &f, std::function<std::string()> wait, bool all)
{
    auto job = std::make_shared<Job>();
    job->slot = s;
    job->f = f;
    job->wait = std::move(wait);
    {
        std::lock_guard<std::mutex> lk(mu_);
        // each due writer's frame number, in submission order
        for (size_t k = 0; k < writers_.size(); ++k)
            if (all || f.nbiter % writers_[k]->frequency == 0)
            {
                job->writer.push_back((int)k);
                job->index.push_back(writers_[k]->frame++);
            }
        job->bytes.resize(job->writer.size());
        q_.push_back(job);
        todo_.push_back(job);
    }
    cv_.notify_all();
}

// Encoder threads: wait for a frame's copy, then format it for every due
// writer that does not need the ordered path, frames in parallel.
void TrajectorySet::encode_loop()
{
    for (;;)
    {
        std::shared_ptr<Job> job;
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [&] { return stop_ || !todo_.empty(); });
            if (todo_.empty())
                return;
            job = todo_.front();
            todo_.pop_front();
        }
        std::string e;
        try
        {
            e = job->wait();
        }
        catch (const std::exception &x)
        {
            e = std::string("frame copy failed: ") + x.what();
        }
        const int32_t *st = job->f.status;
        const bool stopped = st && (st[2] != 0 || st[5] != 0 || st[1] != INT32_MAX);
        if (e.empty() && !stopped)
            for (size_t j = 0; j < job->writer.size(); ++j)
            {
                TrajWriter &w = *writers_[(size_t)job->writer[j]];
                if (w.ordered())
                    continue;
                try
                {
                    if (!w.encode(job->f, job->index[j], job->bytes[j]))
                        e = "write failed: " + w.path;
                }
                catch (const std::exception &x)
                {
                    e = std::string("write failed: ") + x.what();
                }
                if (!e.empty())
                    break;
            }
        {
            std::lock_guard<std::mutex> lk(mu_);
            job->err = e;
            job->skip = !e.empty() || stopped;
            job->encoded = true;
        }
        cv_.notify_all();
    }
}

// The ordered path: frames in submission order, each once its encoding is
// done; the ordered writers' rows, then every payload appended to its file.
void TrajectorySet::run()
{
    for (;;)
    {
        std::shared_ptr<Job> job;
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [&] { return (stop_ && q_.empty()) || (!q_.empty() && q_.front()->encoded); });
            if (q_.empty())
                return;
            job = q_.front();
            q_.pop_front();
            busy_ = true;
        }
        std::string e = job->err;
        if (!job->skip)
        {
            const double t =
                std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
            // steps per second between the completions of consecutive frames
            job->f.framerate = (last_nb_ >= 0 && t > last_t_ && job->f.nbiter > last_nb_)
                                   ? (double)(job->f.nbiter - last_nb_) / (t - last_t_)
                                   : 0.0;
            last_nb_ = job->f.nbiter;
            last_t_ = t;
            for (size_t j = 0; j < job->writer.size(); ++j)
            {
                TrajWriter &w = *writers_[(size_t)job->writer[j]];
                bool ok = false;
                try
                {
                    ok = (!w.ordered() || w.encode(job->f, job->index[j], job->bytes[j])) && w.emit(job->bytes[j]);
                }
                catch (const std::exception &x)
                {
                    ok = false;
                    e = e.empty() ? std::string("write failed: ") + x.what() : e;
                }
                if (!ok && e.empty())
                    e = "write failed: " + w.path;
            }
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            if (!e.empty() && err_.empty())
                err_ = e;
            free_[(size_t)job->slot] = true;
            busy_ = false;
        }
        cv_.notify_all();
    }
}
bool TrajectorySet::drain()
{
    std::unique_lock<std::mutex> lk(mu_);
    cv_.wait(lk, [&] { return q_.empty() && !busy_; });
    return err_.empty();
}
void TrajectorySet::close()
{
    if (th_.joinable())
    {
        drain();
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        th_.join();
        for (auto &t : enc_)
            t.join();
        enc_.clear();
    }
    writers_.clear();
}

} // namespace bs
