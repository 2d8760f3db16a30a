// ensemble.h -- ragged ensembles: flows of different grids advanced together, one launch per
// stage and kernel class for all of them (the tile shapes the autotuner picks decide which
// groups share a class), the step captured in one HIP graph.
//
// Members of equal shape form a group (a batched dns_ctx with a number of slots); a slot is
// active or idle.  Per stage, the tiles of every active slot of every group that share a kernel
// class (plane-size class and precision for the transforms, row blocks per wave for the
// wall-normal kernels) go into one tile table and one launch: block b works on tile map[b] of
// group map[b].g with that group's argument block.  Tiles are independent and each is computed
// by the same arithmetic as in a context of one, so a member's result does not depend on the
// other members.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "context.h"
#include "kernels.h"

namespace dns {

// the flow controller runs inside K_HELM and mode (0,0) inside K_POIS
enum StageKind { K_ZVORT = 0, K_C2R, K_R2C, K_HELM, K_POIS, K_COUNT };

// most streams a step is split over (GPU_MAX_HW_QUEUES is 4 by default)
constexpr int ENS_LANES = 4;

// Tunable launch shapes (chosen by Ensemble::autotune or set by hand).
struct EnsTuning {
    int promote[2][FFT_CLASSES] = {};   // [f32][size class]: levels up in capacity order (more threads)
    int lanes = 1;                                       // streams the step is split over (1 to ENS_LANES)
    int interleave = 0;                                  // 1: z-kernel tiles ordered column-block-major across members
    // tile shapes of the wall-normal kernels (vorticity z-derivative, Helmholtz, Poisson): the
    // least waves per column block (z_variant_widen: 0 = one per row-block pair, 2, 4); groups
    // of different nz at one width share a launch
    int zwaves[3] = {0, 0, 0};
};

class Ensemble {
public:
    struct Group {
        dns_ctx* c = nullptr;        // null: removed
        std::vector<char> active;    // per slot
        int lane = 0;
    };
    std::string err;
    std::vector<Group> groups;
    EnsTuning tune;

    Ensemble() = default;
    ~Ensemble();
    int init();
    // a group of n slots of one shape (params[s] for slot s); returns its index, or -1 (err set)
    int add_group(const dns_params* params, int n, bool active);
    int remove_group(int g);
    int set_active(int g, int slot, bool on);
    int live_slots() const;
    size_t bytes() const;
    // tile tables and argument blocks for the current membership and tuning (synchronises)
    int rebuild();
    // one step of every active slot on stream s0 (forked over the lanes' streams when split), eager
    int enqueue_step();
    // n steps, eager or replayed from the captured graph; per-step device time kept
    int step(int n, bool graph);
    // the tables (and with graph the captured step) for the current membership, ahead of step
    int prepare(bool graph);
    // launches of one step per stage kind (K_ZVORT .. K_POIS) for the current tables
    void launch_counts(int* counts) const;
    int synchronize();
    int timing(double* out4);
    // choose tuning by timing graph replays; restores every field and controller state
    int autotune(int steps, std::string& report);
    hipStream_t stream() const { return s0; }
    // the next graph step rebuilds the tables and re-captures (membership or tuning changed)
    bool stale() const { return dirty || !gexec; }
    long long steps_done = 0;

private:
    struct Launch {
        int kind, cls, f32, off, count, lane, nf;
        size_t lds;
    };
    hipStream_t s0 = nullptr;                 // lane 0, the stream of the step
    hipStream_t sl[ENS_LANES] = {};           // sl[0] = s0, the other lanes' streams
    hipEvent_t fork = nullptr;
    hipEvent_t join[ENS_LANES] = {};
    FFTArgs *d_c2r = nullptr, *d_r2c = nullptr;
    ZArgs* d_z = nullptr;
    HelmArgs* d_h = nullptr;
    PoisArgs* d_p = nullptr;
    Tile* d_tiles = nullptr;
    size_t cap_groups = 0, cap_tiles = 0;
    std::vector<Launch> launches;
    bool dirty = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    std::vector<hipEvent_t> ev;
    std::vector<double> step_ms;
    int pending = -1;
    int launch(const Launch& L, hipStream_t s);
    int drop_graph();
    int ensure_graph();
    int save_state(std::vector<void*>& bufs);
    int restore_state(std::vector<void*>& bufs);
    int time_steps(int n, double& ms_per_step);
};

// HBM bytes dns_create_batch allocates for B members of shape p (what dns_device_bytes reports)
bool ctx_bytes_for(const dns_params& p, int B, size_t& bytes, std::string& err);

}  // namespace dns
