// slot_exec.hpp -- host-side executor of multi-cell slots: a fixed sequence of tile kernels whose contents
// (cells and tiles) change from one slot (or batch) to the next. Used by the slot plans
// (pe_slot_plan_*, fixed cells) and by the serving loop (pe_serve_run, a new batch of jobs every launch).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "pusch_eq.h"

namespace pe {

// Device view of one cell-job.
struct CellDev {
    const float2 *H;
    const float *sigma2;
    float2 *C;
    float *noise_inv;
    const float2 *Y;
    float *llr;
    uint8_t *flags;  // stage-1 per-subcarrier flags (executor-owned)
    int *cexp;       // binary16 coefficient cells (PE_PREC_F16_C16): exponents after the C16 block, else null
    int nSC, L, S, A, npb, prec;
    int mmse_p;      // stage-1 lanes per subcarrier (the shape's choice, as in the single-cell path)
    int rec;         // stage 1 keeps solve records in C, stage 2 forms the coefficients on chip
};

struct Tile {
    int cell, sc0, n, pad;
};

// Per-launch tile range, read by the kernels from device memory so that one captured graph serves any batch.
struct LaunchRange {
    int tile0, ntiles;
};

class SlotExec {
public:
    static constexpr int kApplyTile = 4;  // subcarriers per stage-2 tile (one per wave of a 4-wave workgroup)

    struct Launch {
        int kind;  // 0 mmse, 2 apply, 3 apply forming coefficients from solve records
        int cls;   // apply: largest antenna class KS of the launch (4 or 8)
        int prec;
        int grid, block;  // workgroups (each takes tiles g, g + grid, ...) and threads per workgroup
        int cap;          // largest tile count of the launch
    };

    SlotExec() = default;
    ~SlotExec();
    SlotExec(const SlotExec &) = delete;
    SlotExec &operator=(const SlotExec &) = delete;

    // Sizes device state for batches of up to max_cells cell-jobs drawn from `population` (shapes and
    // precisions; pointers ignored). fixed: every batch is the population itself (slot plans), so launch grids
    // are exact; else they cover max_cells jobs of any population shape. flags: PE_MMSE_UNBIASED.
    pe_status init(const pe_cell_desc *population, int n_population, int max_cells, unsigned flags, bool fixed,
                   std::string &err);
    // Builds the tiles of a batch into the pinned staging area (cells must match population classes).
    pe_status stage(const pe_cell_desc *cells, int n, std::string &err);
    // Copies the staged batch to the device on `st` (one asynchronous copy).
    pe_status upload(hipStream_t st, std::string &err);
    // Launches the kernel sequence on `st` (kernels with no tiles exit at once). exact: grids sized to the
    // staged batch (eager launches of a small batch) instead of the capacity grids a captured graph uses.
    pe_status launch(hipStream_t st, std::string &err, bool exact = false);
    // Captures launch() into a graph on first use, then replays it.
    pe_status graph_launch(hipStream_t st, std::string &err);

    int tiles_staged() const { return staged_tiles_; }
    int max_grid() const;
    const std::vector<Launch> &launches() const { return launches_; }

private:
    int launch_index(int kind, int prec) const;
    // the cell's coefficients go through solve records (PE_SLOT_SOLVE_RECORDS and a shape that allows it)
    bool uses_records(const pe_cell_desc &c) const;
    hipError_t launch_one(int i, hipStream_t st, bool exact) const;

    unsigned flags_ = 0;
    bool fixed_ = false;  // every batch is the population (grids = tile counts, one tile per workgroup)
    bool c16_ = false;    // the cells' coefficients are binary16 (PE_PREC_F16_C16): stage 1 writes that format
    int max_cells_ = 0, max_tiles_ = 0;
    int mmse_tile_ = 16, apply_tile_ = kApplyTile;  // subcarriers per tile
    int mmse_waves_ = 1;                                     // waves per stage-1 workgroup
    int mmse_keep_ = 0;                                      // stage 1 keeps antenna rows in registers
    int mmse_lds_ = 0;                                       // LDS bytes of a stage-1 wave
    int max_layers_ = 1;                                     // largest layer count the LDS is sized for
    size_t flags_per_cell_ = 0;
    std::vector<Launch> launches_;
    // one contiguous block, same layout on host (pinned) and device: [LaunchRange x nL][CellDev x max_cells]
    // [Tile x max_tiles]
    size_t off_cells_ = 0, off_tiles_ = 0, block_bytes_ = 0, used_bytes_ = 0;
    unsigned char *h_block_ = nullptr, *d_block_ = nullptr;
    uint8_t *d_flags_ = nullptr;
    int staged_tiles_ = 0;
    std::vector<int> staged_ntiles_;  // per launch, of the staged batch
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
};

int ks_class(int A);
// Validates one cell descriptor for the slot executors (shape limits, MFMA precision, pointers).
pe_status check_cell_desc(const pe_cell_desc &c, int i, std::string &err);

}  // namespace pe
