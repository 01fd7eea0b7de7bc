// Trajectory and energy output: the writers of io::modern::TrajectoryManager
// (src/IO/modern/TrajectoryManager.cpp:9-26) -- PDB models
// (PDBTrajectoryWriter.cpp:10-33, pdb_format.hpp:17-58), XTC frames
// (XTCTrajectoryWriter.cpp:13-31 over the xdrfile XTC format) and the CSV
// energy table (CSVTrajectoryWriter.cpp:11-60) -- fed from frames that the
// engine copies out of HBM into pinned host slots at step boundaries.
// Encoder threads format frames in parallel once their copies have landed
// (an XTC frame of 250k beads takes milliseconds to compress); one writer
// thread appends them to the files in frame order, so the stream keeps
// stepping while files are written and the files hold the same bytes as a
// sequential writer's.
//
// Host code only: no HIP types here.  The engine hands each frame over with a
// `wait` callback (an event synchronisation) and gets the slot back through
// the slot pool below.
#pragma once

#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace bs
{

enum TrajKind
{
    TRAJ_PDB = 0,
    TRAJ_XTC = 1,
    TRAJ_CSV = 2,
};

// Per-particle PDB fields (spn::Particle, Particle.h:36-41, and
// ParticleProperty.h:20-21 defaults when a field is left empty).
struct TrajLabels
{
    std::vector<std::string> name, resname, chain, element;
    std::vector<int32_t> resid, echarge;
    std::vector<float> occupancy, tempfactor;
};

// What a writer needs besides the frames, fixed when it is added.
struct TrajContext
{
    int64_t n = 0;
    TrajLabels labels;                                  // empty vectors: defaults
    std::vector<int32_t> echarge_default;               // (int)charge per particle
    std::vector<std::pair<uint32_t, uint32_t>> conect;  // springs, for the first PDB model
    bool spring = false, steric = false, elec = false, imp = false, hydro = false;
    int64_t ivec1 = -1, ivec2 = -1;                     // insertion-vector columns when >= 0
};

struct TrajFrame
{
    int64_t nbiter = 0;
    const double *energies = nullptr; // [6]: spring, electrostatic, steric, kinetic, IMP, hydrophobic
    const float *x3 = nullptr;        // n x 3, particle order, A
    double framerate = 0.0;     // steps per second before this frame (0 for the first)
    // nullable: the engine's status words when the frame was taken; a frame
    // taken after its call stopped (an overflowed list, or a non-finite
    // position in an earlier step: [2], [5], [1] != INT32_MAX) is not written
    const int32_t *status = nullptr;
};

class TrajWriter
{
  public:
    virtual ~TrajWriter() = default;
    // The bytes of frame f as this writer's frame number `index`; false on a
    // formatting failure.  Called from any thread unless ordered().
    virtual bool encode(const TrajFrame &f, int64_t index, std::string &out) = 0;
    // Appends bytes to the file; false after an I/O error (the file stays as
    // written so far).  Called in frame order.
    virtual bool emit(const std::string &bytes) = 0;
    // encode needs the frame order (the CSV rows' frame rate).
    virtual bool ordered() const { return false; }
    int64_t frequency = 1;
    int64_t frame = 0; // frames given to this writer
    std::string path;
};

// nullptr and err set when the file cannot be opened.
std::unique_ptr<TrajWriter> make_traj_writer(int kind, const std::string &path, int64_t frequency,
                                             std::shared_ptr<const TrajContext> ctx, std::string &err);

// One XTC frame in the xdrfile layout (xdrfile_xtc.c:22-72, xdrfile.c:965-1256):
// header, box, then the coordinates compressed at `precision`.  Appended to out.
void xtc_encode_frame(std::vector<uint8_t> &out, int32_t natoms, int32_t step, float time, const float box[9],
                      const float *x3nm, float precision);

// The writers and the frame queue of one handle.
class TrajectorySet
{
  public:
    explicit TrajectorySet(int nslots) : free_(nslots, true) {}
    ~TrajectorySet() { close(); }

    // Adds a writer once the queued frames are written.
    void add(std::unique_ptr<TrajWriter> w);
    bool empty() const { return writers_.empty(); }
    int nslots() const { return (int)free_.size(); }
    // Whether a writer takes a frame at iteration nb (TrajectoryManager.cpp:9-16).
    bool due(int64_t nb) const;
    // Smallest iteration > nb at which a writer takes a frame.
    int64_t next_due(int64_t nb) const;

    // A free slot (waits for the writer thread to hand one back).
    int acquire();
    // Queues frame f held in slot s; `wait` returns once the slot's copy has
    // landed.  all = true writes to every writer (writeNextStepNow).
    // The callback returns an error message, empty on success; a frame whose
    // copy failed is not written and the error is kept (error()).
    void submit(int s, const TrajFrame &f, std::function<std::string()> wait, bool all);
    // Returns a slot without writing it (a frame whose copy failed).
    void release(int s);
    // Waits until every queued frame is written; false if a write failed.
    bool drain();
    // Drains, stops the thread and closes the files.
    void close();
    const std::string &error() const { return err_; }

  private:
    struct Job
    {
        int slot = -1;
        TrajFrame f;
        std::function<std::string()> wait;
        std::vector<int> writer;        // the due writers
        std::vector<int64_t> index;     // their frame numbers
        std::vector<std::string> bytes; // their payloads
        std::string err;
        bool skip = false, encoded = false;
    };
    void run();
    void encode_loop();
    std::vector<std::unique_ptr<TrajWriter>> writers_;
    std::vector<bool> free_;
    std::deque<std::shared_ptr<Job>> q_;    // submission order, until written
    std::deque<std::shared_ptr<Job>> todo_; // waiting for an encoder
    std::mutex mu_;
    std::condition_variable cv_;
    std::thread th_;
    std::vector<std::thread> enc_;
    bool stop_ = false, busy_ = false;
    std::string err_;
    int64_t last_nb_ = -1;
    double last_t_ = 0.0;
};

} // namespace bs
