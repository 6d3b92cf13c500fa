// engine.hpp -- internal host runtime: a set of foams sharing one stream,
// one device array of foam headers, concatenated block tables, a device
// memory arena and a captured HIP graph.  A single simulation is an engine
// with one foam; a batch holds many foams of uneven size; the serving loop
// adds and retires foams between runs.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "cubble_amd.h"
#include "foam.hpp"
#include "host_init.hpp"
#include "launch.hpp"

namespace cubble_amd {

// First-fit device memory arena over one or more hipMalloc chunks, 256-byte
// granules, coalescing frees.  A fixed arena (growable = false) is the HBM
// budget of the serving loop: allocation fails instead of growing.
class DeviceArena {
  public:
    ~DeviceArena();
    int init(size_t first_chunk, bool growable, std::string &err);
    void *alloc(size_t bytes);
    void release(void *p);
    size_t capacity() const { return capacity_; }
    size_t in_use() const { return in_use_; }
    size_t largest_free() const;

  private:
    struct Chunk {
        char *base;
        size_t size;
        std::map<size_t, size_t> free;  // offset -> length
    };
    std::vector<Chunk> chunks_;
    std::map<void *, size_t> live_;     // pointer -> length
    bool growable_ = true;
    size_t capacity_ = 0, in_use_ = 0;
    bool add_chunk(size_t bytes);
};

struct Tuning {
    int ipt = 0;       // bubbles per thread (0 = from the size)
    int stile = 256;   // bubbles per neighbour-search tile
    int groups = 16;   // attempt groups per captured graph replay
};

// Smallest neighbour-search tile, in bubbles.
constexpr int kMinSearchTile = 32;

// Most attempt groups between two host checks of the foams' flags in
// run_steps (halts, list-capacity growth).
constexpr int kRoundGroups = 128;

// Device copies of every foam's buffers and host copies of their headers,
// taken by Engine::snapshot and put back by Engine::restore.
struct Snapshot {
    std::vector<Foam> h;
    std::vector<void *> blk, nbr;
    std::vector<size_t> blk_bytes, nbr_bytes;
};

// Host-side record of one foam.
struct FoamRec {
    Foam h{};                 // host mirror of the device header
    void *block = nullptr;    // all per-foam buffers except the lists
    void *nbr_block = nullptr;
    int32_t num_end = 0;
    int64_t max_steps = 0;
    int failed = 0;           // status code once the foam was failed alone
};

// Bytes of device memory one foam needs (the admission footprint).
size_t foam_footprint(int dim, int precision, int cap, int nbr_cap, int cell_cap,
                      const Tuning &t);

class Engine;
// Copies of foam k's state between the device and caller-owned host views
// (state_io.hip).
int get_scalars(Engine &e, int k, cubble_scalars *out);
int get_state(Engine &e, int k, cubble_state_view *view);
int get_attempt(Engine &e, int k, cubble_attempt_view *view);
int set_state(Engine &e, int k, const cubble_state_view *view, const cubble_scalars *sc);

class Engine {
  public:
    ~Engine();
    int init(size_t arena_bytes, bool growable);
    // Draw the synthetic foam, allocate and upload it (no search yet).
    int add_foam(const cubble_config &cfg, const Tuning &t, int *index);
    // Allocate and upload a foam drawn earlier by generate_foam.
    int add_generated(const cubble_config &cfg, const Tuning &t, const std::vector<double> &x4,
                      const InitOut &io, int *index);
    // Free a foam's memory and drop it (layout must be committed again).
    int drop_foam(int index);
    // Rebuild block tables, upload every header, drop the captured graph.
    int commit_layout();
    // First neighbour search of the foams whose need_search flag is set.
    int search_pending();

    Grid grid() const;
    void enqueue_attempt(uint32_t flags);
    void enqueue_group(uint32_t flags);
    int capture(uint32_t flags, int groups);

    int read_dyn_all(std::vector<Dyn> &out);
    int write_dyn(int k, const Dyn &d);
    int read_header(int k, Foam &f);
    int sync();
    // Read every foam's scalars; grow and rebuild the neighbour lists of
    // foams whose search overflowed the list capacity.
    int settle(std::vector<Dyn> &d);

    // Host-driven: one attempt group at a time until each foam accepted one
    // step (integrate()).
    int integrate(uint32_t flags, std::vector<Dyn> &d);
    // Device-driven: `steps` accepted steps for every foam (or per_foam[k]
    // for foam k) through graph replays; foams that reach their limit or
    // their dissolution stop halt on the device.
    int run_steps(uint32_t flags, int64_t steps, std::vector<Dyn> &d,
                  const std::vector<int64_t> *per_foam = nullptr);
    // Step limits for the next run (0 = none); clears halt and overflow.
    int set_limits(int64_t steps, const std::vector<int64_t> *per_foam);

    // New launch shapes for every foam (within the block partials carved for
    // the smallest shapes); the state is untouched.
    int retune(const Tuning &t);
    const Tuning &tuning() const { return tuning_; }
    // Copy the whole device state aside / put it back (launch shapes stay).
    int snapshot(Snapshot &s);
    int restore(const Snapshot &s);
    void free_snapshot(Snapshot &s);

    int fail(int code, const std::string &m) {
        err = m;
        return code;
    }

    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<FoamRec> foams;
    int precision = CUBBLE_PREC_FP64;
    std::string err;
    DeviceArena arena;
    // With a fixed arena (the serving loop), a foam whose neighbour lists
    // cannot grow is failed alone (FoamRec::failed, halted on the device)
    // instead of failing the whole run.
    bool isolate_failures = false;

    // Mark foam k failed with `code` and halt it for good (isolate_failures).
    int fail_foam(int k, int code);

  private:
    int grow_lists(int k, int need, bool rebuild, bool keep = false);
    int pull_headers();
    Tuning tuning_;
    Foam *dfoams_ = nullptr;
    int dfoams_cap_ = 0;
    int2 *bt_bub_ = nullptr, *bt_cell_ = nullptr, *bt_scell_ = nullptr;
    size_t bt_bub_cap_ = 0, bt_cell_cap_ = 0, bt_scell_cap_ = 0;
    int nb_bub_ = 0, nb_cell_ = 0, nb_scell_ = 0;
    Dyn *pinned_ = nullptr;
    int pinned_cap_ = 0;
    std::vector<int> dev_index_;  // position of each foam in the device array
    hipGraphExec_t graph_ = nullptr;
    uint32_t graph_flags_ = 0;
    int graph_groups_ = 0;
    bool graph_walls_ = false;
    int committed_nf_ = 0;  // foams in the device array at the last commit

  public:
    Foam *dfoams() const { return dfoams_; }
    // Raise foam k's list capacity without rebuilding (before uploading lists).
    int grow(int k, int need) { return grow_lists(k, need, false); }
};

}  // namespace cubble_amd
