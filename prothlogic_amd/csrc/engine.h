// engine.h -- device-resident modular squaring of Proth residues on one
// gfx950 GPU.  A Group holds `capacity` slots that share one transform
// length L (and its tables) but each carry their own candidate
// N = k*2^n + 1, residue, top scalar and schedule.  Every kernel launch
// covers all slots of a group (blockIdx.y = slot), so a batch of candidates
// of one length advances with one launch per kernel kind.  See engine.hip
// for the algorithm.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <map>
#include <string>
#include <utility>
#include <vector>

namespace pleng {

// Transform shape for one length: L = R*C complex points (= digits of a
// residue), R = strided (column) DFT length, C = contiguous (row) length.
struct Plan {
    uint32_t n = 0;       // candidate the plan was chosen for (single use)
    uint64_t k = 0;
    uint32_t L = 0, logL = 0, R = 0, C = 0;
    uint32_t s = 0;       // log2 of the low half of the two-level tables
    uint32_t nt = 0;      // split outer pass with runs of nt digits (8 or 16), 0 = unfused
    uint32_t run = 16;    // digits per thread in the flat carry kernel (8 or 16)
    // largest radix of the steady-state row pass and of the steady-state
    // outer (column) pass, 8 or 16; the entry, exit and product passes use 8
    uint32_t row_radix = 8, col_radix = 8;
};

// Shapes a plan's length allows: log2 R for L = R*C with R in [32, 1024],
// C in [32, 4096] (the kernels' instantiated lengths), and R in [64, 512]
// where the split outer pass is wanted.
std::vector<uint32_t> shape_options(uint32_t logL, bool split);
// Sets R = 2^logR, C = L/R (and the two-level table split) on a plan.
void set_shape(Plan& p, uint32_t logR);
// log2 R of the preferred shape whose steady state runs as one resident
// launch (k_loop) at length L on this device; 0 when no such shape fits.
uint32_t resident_logR(int device, uint32_t L);

// Smallest power of two L in [2^10, 2^22] with n/L <= max_bits_per_digit(L, k)
// and n/L >= 7.  force_L, when nonzero, is checked and used instead.
bool choose_plan(uint32_t n, uint64_t k, uint32_t force_L, Plan& out, std::string& why);
double max_bits_per_digit(uint32_t logL, uint64_t k);
// Fills nt and run for slots whose bits per digit are at least bmin, whose
// k is below kmax and whose per-squaring multiplier is at most mult_max.
// Returns false when no carry layout can absorb those carries.
bool choose_layout(Plan& p, double bmin, uint64_t kmax, uint32_t mult_max, bool allow_split, std::string& why);

struct DeviceTables;
struct Params;

// Per-slot constants on the device.
struct SlotDesc {
    unsigned long long n;
    long long k;
    unsigned a;       // base of the seed powering (0: no schedule multipliers)
    unsigned topbit;  // index of the top set bit of k
    unsigned long long total;  // squarings of the whole test (topbit + n - 1)
};

// Per-slot progress on the device (scheduled stepping only).
struct SlotState {
    unsigned long long carried;  // squarings completed (carried) so far
    int active;                  // 1 while steps still apply to the slot
    int pad;
};

class Group {
  public:
    Group(const Plan& p, uint32_t capacity, int device);
    // A group in memory the caller provides (at least layout_bytes(p,
    // capacity) bytes of device memory on `device`) and on the caller's
    // stream: nothing is allocated on the device and nothing is freed at
    // destruction.  Such a group takes the split outer pass only.
    Group(const Plan& p, uint32_t capacity, int device, void* mem, size_t mem_bytes, hipStream_t stream);
    ~Group();
    static size_t layout_bytes(const Plan& p, uint32_t capacity);
    Group(const Group&) = delete;
    Group& operator=(const Group&) = delete;

    const Plan& plan() const { return plan_; }
    uint32_t capacity() const { return cap_; }
    hipStream_t stream() const { return stream_; }
    size_t device_bytes() const { return bytes_; }
    static size_t slot_bytes(const Plan& p);  // HBM per slot (residue + carry staging)

    // Slot set-up.  a = 0 disables the schedule multipliers; total = the
    // squarings of the slot's test (used by scheduled stepping).
    void set_slot(uint32_t slot, uint32_t n, uint64_t k, uint32_t a, uint64_t total);
    void set_state(uint32_t slot, uint64_t carried, bool active);

    // Digits (variable base, balanced) + top scalar in/out of a slot.
    void upload(uint32_t slot, const std::vector<int32_t>& digits, int64_t xhi);
    void download(uint32_t slot, std::vector<int32_t>& digits, int64_t& xhi);

    // Plain squaring: every slot squared `iters` times (no multipliers),
    // from and back to the digit state.  err_trace (iters entries, nullable)
    // gets slot 0's per-iteration max round-off.  Returns the max over all.
    double square(uint64_t iters, double* err_trace);
    // The same for slots [slot0, slot0 + slots) only (err_trace: slot0's).
    double square_slots(uint32_t slot0, uint32_t slots, uint64_t iters, double* err_trace);

    // Slot arithmetic in the digit state.  copy_slot: dst <- src.
    // mul_slot: dst <- dst*src mod N (returns its round-off); src is left
    // transformed, so copy it first when it is still needed.
    void copy_slot(uint32_t dst, uint32_t src);
    double mul_slot(uint32_t dst, uint32_t src);

    // Scheduled stepping.  enter: digit state -> first squaring's FFT done;
    // step: carry squaring i (x <- x^2 * a^bit) and start squaring i+1 for
    // every active slot, `steps` times; leave: carry the slot's last
    // squaring and return to the digit state.  A slot goes inactive after
    // total-1 steps.  step() returns the max round-off of the chunk and
    // fills per-slot maxima (capacity entries, nullable).
    void enter(uint32_t slot);
    double step(uint32_t steps, double* slot_err);
    // step() in two halves, so that several groups' chunks run at once on
    // their own streams: step_begin enqueues one chunk of g <= kStepChunk
    // steps; step_end waits for it and reads its round-off like step().
    // slots > 0: only slots [0, slots) can be active, so the chunk's launches
    // cover the smallest power of two >= max(slots, 8) of them (capped at
    // the capacity) instead of every slot; step_end then reports round-off
    // for those slots only (0 for the rest).
    void step_begin(uint32_t g, uint32_t slots = 0);
    double step_end(double* slot_err);
    // Whether the chunk step_begin enqueued has finished (step_end then
    // returns without waiting); true when none is in flight.
    bool step_ready();
    double leave(uint32_t slot);
    std::vector<SlotState> states();

    int status();
    void clear_status();
    // kernels per steady-state squaring: 2 (one-launch outer pass + row
    // pass), 3 (split outer pass in two launches + row pass) or 5 (unfused)
    // (0: the steady state is one resident launch per chunk, k_loop)
    int launches_per_squaring() const
    {
        return resident_ && turn() && uniform_ && plan_.row_radix == 8 && plan_.col_radix == 8 ? 0
             : turn()                                                                       ? 2
             : plan_.nt                                                                     ? 3
                                                                                            : 5;
    }
    // the one-launch outer pass can be used (runs of 8 digits; with 16 its
    // registers would spill)
    bool turn() const { return turn_ && plan_.nt == 8; }
    // Plain chunks of this group's slot run their steady state as one
    // resident launch (k_loop: every workgroup on its CU for the whole chunk,
    // outer and row passes handing points over without launch boundaries).
    // Needs the one-launch pass, a shape k_loop is built for whose workgroups
    // all fit on the device at once, one slot per chunk and every slot
    // holding the plan's candidate, and the plan's radices at 8 (the loop
    // variants fix their own; plans pinned at radix 16 keep two kernels).
    // PL_RESIDENT=0 turns it off.
    bool resident_on() const
    {
        return resident_ && turn() && sqn_ == 1 && uniform_ && plan_.row_radix == 8 && plan_.col_radix == 8;
    }
    void no_resident() { resident_ = false; }
    // Use the split outer pass even where the one-launch pass would run.  The
    // one-launch pass's block 0 waits on its slot's last block, which starts
    // after it: a launch only finishes when all its blocks can eventually be
    // resident.  Two such launches in flight at once can each hold the CUs
    // the other's last blocks need, so callers that overlap groups on several
    // streams (the pool) take the split pass.  Call before the first step.
    void no_turn() { turn_ = false; }
    // Chunks of plain squarings run the one-launch pass only while they hold
    // their device's turn token (at most one holder per device in the
    // process; others take the split pass for that chunk), and a chunk whose
    // carry hand-off gave up is recomputed from its saved start on the split
    // pass.  Number of such recomputations so far:
    uint32_t handoff_fallbacks() const { return fallbacks_; }
    // Device time of the resident launches of completed plain chunks (HIP
    // events around each launch, on this group's stream), how many launches
    // and how many squarings they ran; reset clears the sums after reading.
    void loop_time(bool reset, double* ms, uint64_t* launches, uint64_t* squarings);

    // Average duration (ms) of each kernel over all slots, timed with HIP
    // events between consecutive kernels of `reps` real plain squarings on
    // this group's stream (kernels in their steady-state order, each on its
    // predecessor's output), less the events' own cost (measured on the same
    // loop timed end to end): [0] forward column pass, [1] row pass, [2]
    // inverse column pass, [3] flat carry, [4] block fix-up, [5] inverse
    // column DFT + carry, [6] carry-in + forward column DFT ([5], [6] are 0
    // without the split outer pass; with it, [0], [2]-[4] come from the one
    // entry/exit pass around the loop).  The kernels cover slots
    // [0, nslots); slot 0's residue is saved and restored, the others are
    // squared `reps` times.
    void kernel_times(int reps, double* ms, uint32_t nslots);
    static constexpr int kNumKernels = 8;
    // squarings per plain chunk (one graph replay, one round-off read-back)
    // and columns of the round-off arrays; a chunk enters and leaves the
    // steady state once, so longer chunks spread that cost thinner
    static constexpr int kChunk = 256;
    // steps per pool chunk (the pool retires and admits between chunks)
    static constexpr int kStepChunk = 64;

  private:
    void init(void* ext_mem, size_t ext_bytes, hipStream_t ext_stream);
    Params params(uint32_t slot0, bool check_active) const;
    void launch(int kind, const Params& p, uint32_t slots);
    // parity >= 0: a pipelined resident chunk (its backup and events)
    void enqueue_plain_chunk(int g, bool use_turn, int parity = -1);
    uint64_t square_pipelined(uint64_t iters, double* err_trace, double& emax);
    void ensure_pipeline();
    void enqueue_exit(const Params& p, uint32_t slots);
    void enqueue_step_chunk(int g);
    void build_graph(hipGraphExec_t& exec, hipGraph_t& graph, bool stepping, int g, bool use_turn);
    void run_plain_chunk(int g, bool use_turn);
    void ensure_backup();
    void ensure_loop_events();
    void restore_start(uint32_t slot0, uint32_t slots);
    void skip_epochs(uint32_t slot);
    void sync_status(const char* what);
    double read_err(int g, double* trace, double* slot_err, uint32_t trace_slot = 0, uint32_t nslots = 0);
    void enqueue_err_clear(uint32_t slots, unsigned col0);
    void enqueue_set_col(uint32_t slot, unsigned col);
    void enqueue_err_reduce(uint32_t slot0, uint32_t slots);

    Plan plan_;
    uint32_t cap_;
    int device_;
    hipStream_t stream_ = nullptr;
    size_t bytes_ = 0;
    void* mem_ = nullptr;
    double2* z_ = nullptr;
    long long* carries_ = nullptr;
    long long* runc_ = nullptr;   // split pass: plain carry-out triples of every run
    long long* runt_ = nullptr;   // one-launch pass: tagged carry words (written by nothing else,
                                  // so every word carries the last launch's tag)
    double2* zbak_ = nullptr;     // start of the running chunk (one-launch pass fallback)
    long long* xbak_ = nullptr;
    unsigned* epoch_ = nullptr;   // cap x 2: next and last carry hand-off epoch
    int* rdig_ = nullptr;
    long long* xhi_ = nullptr;
    unsigned long long* xw_ = nullptr;  // per slot: X_hi as a tagged word (k_loop's row passes poll it)
    unsigned long long* err_ = nullptr;   // cap x kChunk (bit patterns of non-negative doubles)
    unsigned long long* errp_ = nullptr;  // cap x eb_ x kChunk per-block partials
    uint32_t eb_ = 1;                     // blocks that report round-off per squaring
    int* status_ = nullptr;
    int32_t* digits_dev_ = nullptr;
    SlotDesc* desc_ = nullptr;
    SlotState* state_ = nullptr;
    DeviceTables* tabs_ = nullptr;
    uint32_t carry_blocks_ = 0, carry_tpb_ = 64;
    std::vector<SlotDesc> desc_host_;
    // scheduled-step chunks, one graph per launched slot count
    std::map<uint32_t, std::pair<hipGraphExec_t, hipGraph_t>> step_graphs_;
    uint32_t step_n_ = 0;  // slots the step chunk being enqueued launches
    // plain squaring chunks, one graph per slot range and outer-pass form
    // (slot0 << 33 | slots << 1 | one-launch)
    std::map<uint64_t, std::pair<hipGraphExec_t, hipGraph_t>> plain_graphs_;
    uint32_t sq0_ = 0, sqn_ = 1;
    uint32_t pending_steps_ = 0;  // steps enqueued by step_begin, not yet read back
    bool uniform_ = true;  // every slot holds the plan's n and k
    bool turn_ = true;     // split outer pass as one launch (k_carry_turn); PL_CARRY_TURN=0: two
    bool resident_ = false;  // k_loop exists for the shape and fits the device (resident_on())
    int loop_var_ = -1;      // its variant (engine.hip kLoopShapes)
    unsigned poll_sleep_ = 1;  // k_loop's pause between reloads of points not yet ready
    bool give_up_ = false; // PL_HANDOFF_GIVE_UP=1: every carry hand-off wait gives up (tests)
    // PL_NO_GRAPHS=1: whole chunks run as direct launches like partial ones
    // (the same kernels in the same order; for tools that trace dispatches
    // and must not see graph replays)
    bool no_graphs_ = false;
    uint32_t fallbacks_ = 0;
    hipEvent_t lev_[2][2] = { { nullptr, nullptr }, { nullptr, nullptr } };  // around a chunk's resident launch,
                                                                             // per pipeline parity
    // pipelined resident chunks (square_pipelined): per parity a one-slot
    // backup of the chunk's start, pinned round-off columns + status word,
    // and the chunk's end event
    void* pipe_mem_ = nullptr;
    double2* pipe_z_[2] = { nullptr, nullptr };
    long long* pipe_x_[2] = { nullptr, nullptr };
    unsigned long long* pipe_err_[2] = { nullptr, nullptr };
    int* pipe_st_[2] = { nullptr, nullptr };
    hipEvent_t pipe_done_[2] = { nullptr, nullptr };
    double loop_ms_ = 0;
    uint64_t loop_launches_ = 0, loop_squarings_ = 0;
    bool owns_ = true;     // the memory and the stream are the group's own
    // the status word as read at the last wait on the stream, and whether
    // kernels were launched since
    int seen_status_ = 0;
    bool launched_ = true;
    int status_word_ = 0;
    int* status_host_ = &status_word_;
    std::vector<unsigned long long> err_host_;
};

void check(hipError_t e, const char* what);

// Average ms of a plain device-to-device copy of `bytes` (reads and writes
// that many bytes) over `reps` launches after one warm-up launch.
double copy_probe(int device, size_t bytes, int reps);

}  // namespace pleng
