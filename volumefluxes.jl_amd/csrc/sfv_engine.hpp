// sfv_engine.hpp — the batch engine behind sfv_batch_* and sfv_server_*: a
// device member table, one step-control block per member slot, and a
// captured graph of graph_steps RK2 steps of the batched kernels.
//
// Members are sfv_handle objects; the engine borrows their buffers.  A member
// lives in a stable control slot (its StepCtl), while its row in the member
// table may move when membership changes, so members can join and leave
// between graph replays without disturbing the others.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "sfv_handle.hpp"

namespace sfv {

// Segment height of batch members when the batch fills the device.
constexpr int SFV_BATCH_DEFAULT_ROWS = 12;

// How an engine's persistent grid shares the device with another engine's.
//   ALONE: the grid fills every resident block slot.
//   HOST:  one block slot per CU is left free for a GUEST engine, and the
//          waves' rows-left issue priority stays below the guest's.
//   GUEST: one block per CU, and the waves raise their issue priority; for
//          small, latency-bound grids stepping beside a HOST engine.
enum class Share { ALONE, HOST, GUEST };

struct RunParams {
    double t0 = 0.0;
    double endtime = 0.0;         // NaN / unused when until_endtime = 0
    int until_endtime = 0;
    int match_endtime = 0;
    double maximum_timestep = 0;  // <= 0: none
    long long max_steps = 0;      // <= 0: unlimited
};

class BatchEngine {
public:
    ~BatchEngine();
    // Set up for `precision` members on the current device.  capacity is the
    // initial number of member slots (grows on demand).
    int init(int precision, int capacity, std::string& why);

    // Add a member (RK2, fusable, same precision) and prepare its step
    // control for `rp`, from the member's current buffers and parity.
    // Returns the member's slot in *slot_out.  Takes effect at the next replay.
    int add(sfv_handle* h, const RunParams& rp, int* slot_out, std::string& why);
    // Remove the member in `slot` (finished or cancelled), accounting its
    // steps into the handle (buffer parity, last dt).
    void remove(int slot);

    // Replay the graph `k` times (k * graph_steps launched steps), then copy
    // every slot's step control to the host and wait.
    int replay(int k, std::string& why);
    // The same in two halves, so the host can work while the device steps:
    // replay_begin queues the replays and the step-control copy; replay_end
    // waits for them.  Between the two, add / remove / finished / ctl must
    // not be called (they use the table and the step-control mirror).
    int replay_begin(int k, std::string& why);
    int replay_end(std::string& why);
    // Without waiting: has the work queued by replay_begin finished?  (Also
    // true after a device error, which replay_end then reports.)
    bool replay_ready() const;
    // Online segment-height tuning (off by default).  While on, the replays
    // after a membership change time the height rule's choice and its
    // neighbours in the list {1, 2, 3, 4, 6, 8, 12, 16, 24} (one below, two
    // above), one replay each, with the rule's choice first and last, and
    // then keep the fastest: the rule's choice unless another is more than
    // 3% faster than the rule's time interpolated to the trial's place
    // (heights within 1.5% of the fastest tie, the nearest to the rule's wins).
    // The choice is kept for later runs with the same member sizes.  Every
    // height gives bitwise the same results, so the timed replays are
    // ordinary steps of the run.
    void set_autotune(bool on) { autotune_ = on; }
    // A timed trial is pending, or a membership change not yet replayed may
    // start one: replay one graph at a time so that each trial is timed alone.
    bool tuning() const { return (autotune_ && members_changed_) || trial_ < (int)trials_.size(); }
    // Segment height of the members without a forced one (0: no members yet).
    int rows_in_use() const { return rows_in_use_; }

    // The grid's share of the device; takes effect at the next replay_begin
    // (the graph is recaptured when it changes).  Results do not depend on it.
    // Not to be called between replay_begin and replay_end.
    void set_share(Share s);

    int members() const { return (int)rows_.size(); }
    // After replay: has the member in `slot` finished its run?
    bool finished(int slot) const;
    const StepCtl& ctl(int slot) const { return h_ctl_[slot]; }
    sfv_handle* handle(int slot) const { return slot_handle_[slot]; }
    hipStream_t stream() const { return stream_; }
    int graph_steps() const { return graph_steps_; }

private:
    int grow(int capacity, std::string& why);
    int upload_table(std::string& why);
    void choose_rows();
    int rule_rows() const;
    void plan_tuning();
    std::vector<long long> signature() const;
    int capture(std::string& why);
    void size_grid();

    int precision_ = 0;
    int device_ = 0;
    int capacity_ = 0;
    int graph_steps_ = 8;         // even: launch parity is 0 at every replay start
    int blocks_stage_ = 0;
    int per_cu_ = 1;              // resident stage-kernel blocks per CU
    Share share_ = Share::ALONE;
    hipStream_t stream_ = nullptr;
    BatchMember* d_mem_ = nullptr;
    BatchHeader* d_hdr_ = nullptr;
    BatchMember* h_mem_ = nullptr;
    BatchHeader* h_hdr_ = nullptr;
    StepCtl* d_ctl_ = nullptr;
    StepCtl* h_ctl_ = nullptr;
    hipGraphExec_t graph_ = nullptr;
    bool table_dirty_ = true;
    bool members_changed_ = false;
    bool autotune_ = false;
    int rows_override_ = 0;                  // > 0: every member without a forced height takes it
    int rows_in_use_ = 0;
    std::vector<int> trials_;                // heights to time, in order
    std::vector<float> trial_ms_;            // ms per replay of each trial
    int trial_ = 0;                          // next trial; == trials_.size(): none pending
    int timed_k_ = 0;                        // replays between the events (0: none timed)
    std::vector<long long> tuned_sig_;       // member sizes the kept choice was made for
    int tuned_rows_ = 0;
    hipEvent_t ev_[2] = {nullptr, nullptr};
    bool fast_ = false;           // the captured graph uses the specialised march kernels
    int cus_ = 256;
    std::vector<int> rows_;                  // slot of each table row
    std::vector<sfv_handle*> slot_handle_;   // nullptr: free slot
    std::vector<BatchMember> slot_row_;      // the member's row (host copy)
    std::vector<char> ran_;                  // slot took part in the last replay
};

}  // namespace sfv
