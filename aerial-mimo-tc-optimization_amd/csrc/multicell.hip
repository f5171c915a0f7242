// multicell.hip -- many uneven cells per launch as persistent work-queue kernels (SlotExec), and the slot-plan
// C ABI (pe_slot_plan_*) on top of it.
//
// The reference runs one launch per cell on its own stream (multicell_contention.cu:194-283,
// multicell_evidence.cu:497-566). Here a batch of cell-jobs is cut into subcarrier tiles on the host and
// executed by a short fixed sequence of persistent kernels that pull tiles from device work queues:
//   stage 1 (cells with H): MMSE FP32 pass, then the FP64 pass over flagged subcarriers (one kernel each; the
//                           L x L solve is unrolled per L and selected by the tile's cell);
//   stage 2 (all cells):    fused coefficient apply + demap, one kernel per (antenna class KS, precision); the
//                           QAM order is read per tile.
// Tiles of a kernel are ordered by decreasing per-subcarrier cost so the queue drains with a short tail. The
// kernels read their tile range from device memory, so one captured graph serves every batch; a one-thread
// kernel at the head of the sequence zeroes the queue counters, so graph replays need no memset node.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "apply_device.hpp"
#include "mmse_device.hpp"
#include "pe_internal.hpp"
#include "pusch_eq.h"
#include "slot_exec.hpp"

namespace pe {

namespace {

// Work distribution: wave w of a persistent grid of W waves takes tile w first (no atomic), then tiles
// W + atomicAdd(next, 1) until none are left. Waves without a first tile exit at once, so a small batch on a
// large grid costs no atomics. The counters are zeroed by queue_reset_kernel at the head of every slot (a
// kernel node, so a captured graph replays with no memset node).
__device__ __forceinline__ unsigned next_tile(Queue *q, int lane, unsigned nwaves) {
    unsigned t = 0;
    if (lane == 0) t = atomicAdd(&q->next, 1u);
    return __shfl(t, 0, 64) + nwaves;
}

__global__ void queue_reset_kernel(Queue *q, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        q[i].next = 0u;
        q[i].flagged = 0u;
    }
}

struct BlockView {
    const LaunchRange *ranges;
    const CellDev *cells;
    const Tile *tiles;
};

__device__ __forceinline__ BlockView view(const unsigned char *block, int off_cells, int off_tiles) {
    return BlockView{reinterpret_cast<const LaunchRange *>(block), reinterpret_cast<const CellDev *>(block + off_cells),
                     reinterpret_cast<const Tile *>(block + off_tiles)};
}

template <int KS, int PREC>
__global__ __launch_bounds__(64) void slot_apply_kernel(const unsigned char *__restrict__ block, int li,
                                                        int off_cells, int off_tiles, Queue *q, unsigned nwaves) {
    using namespace mfma;
    __shared__ __attribute__((aligned(16))) unsigned char lds[LdsLayout<KS>::bytes(PREC)];
    const int lane = threadIdx.x;
    const BlockView v = view(block, off_cells, off_tiles);
    const LaunchRange rg = v.ranges[li];
    for (unsigned t = blockIdx.x; t < (unsigned)rg.ntiles; t = next_tile(q, lane, nwaves)) {
        const Tile tl = v.tiles[rg.tile0 + t];
        const CellDev cd = v.cells[tl.cell];
        if (cd.A < 16 * KS) zero_operands<KS, PREC>(lds, lane);
        const long first = tl.sc0, last = (long)tl.sc0 + tl.n;
        apply_range<KS, PREC, -1, false>(cd.C, cd.Y, nullptr, cd.noise_inv, cd.llr, first, last, cd.L, cd.S, cd.A,
                                         lds, lane, cd.npb);
        lds_order_wave();  // LDS of this tile consumed before the next tile may zero it
    }
}

// Stage 1: a tile is up to mmse_tile subcarriers of one cell. FP32 pass: one workgroup of W waves solves the
// tile (tile solve). FP64 pass: one-wave workgroups revisit the tile's flagged subcarriers, 64 at a time.
// Lanes per subcarrier are the cell shape's choice, the same as in the single-cell kernels, so the
// coefficients are bit-identical to the cell solved alone.
template <bool FP64, int KR>
__device__ __forceinline__ void mmse_tile(const CellDev &cd, const Tile &tl, float *lds, int tid, int W,
                                          int unbiased, float thr, Queue *q) {
    using namespace mmse;
#define PE_MMSE_L(LL)                                                                                            \
    case LL:                                                                                                     \
        if constexpr (FP64) {                                                                                    \
            for (int c = 0; c < tl.n; c += 64)                                                                   \
                tile_fp64<LL>(cd.H, cd.sigma2, cd.C, cd.noise_inv, cd.flags, tl.sc0 + c, min(64, tl.n - c),      \
                              cd.A, cd.mmse_p, unbiased != 0, tid);                                              \
        } else {                                                                                                 \
            tile_fp32<LL, KR>(cd.H, cd.sigma2, cd.C, cd.noise_inv, cd.flags, tl.sc0, tl.n, cd.A, cd.mmse_p,      \
                              unbiased != 0, thr, lds, tid, W, &q->flagged);                                     \
        }                                                                                                        \
        break;
    switch (cd.L) {
        PE_MMSE_L(1) PE_MMSE_L(2) PE_MMSE_L(3) PE_MMSE_L(4) PE_MMSE_L(5) PE_MMSE_L(6) PE_MMSE_L(7)
    default:
        PE_MMSE_L(8)
    }
#undef PE_MMSE_L
}

// All layer counts in one persistent kernel per pass (the solve is unrolled per L; the tile's cell selects it).
// Workgroups pull whole tiles: the first from blockIdx.x, then from the queue (one atomic per workgroup,
// broadcast through LDS). The FP32 pass uses the workgroup's dynamic LDS for its tile slots.
template <bool FP64, int KR>
__device__ __forceinline__ void slot_mmse(const unsigned char *__restrict__ block, int li, int off_cells,
                                          int off_tiles, Queue *q, unsigned nblocks, int unbiased, float thr,
                                          float *lds, unsigned *next) {
    const int tid = threadIdx.x, W = blockDim.x >> 6;
    const BlockView v = view(block, off_cells, off_tiles);
    const LaunchRange rg = v.ranges[li];
    // the FP64 pass has nothing to do unless the FP32 pass (same queue slot, li - 1) flagged a subcarrier
    if (FP64 && __atomic_load_n(&(q - 1)->flagged, __ATOMIC_RELAXED) == 0u) return;
    unsigned t = blockIdx.x;
    while (t < (unsigned)rg.ntiles) {
        const Tile tl = v.tiles[rg.tile0 + t];
        const CellDev cd = v.cells[tl.cell];
        mmse_tile<FP64, KR>(cd, tl, lds, tid, W, unbiased, thr, q);
        if (tid == 0) *next = atomicAdd(&q->next, 1u) + nblocks;
        __syncthreads();
        t = *next;
        __syncthreads();  // every thread has read `next` before thread 0 rewrites it
    }
}

// KR = 4: antenna rows kept in registers where a tile allows it (see tile_fp32)
template <int KR>
__global__ __launch_bounds__(512) void slot_mmse_fp32_kernel(const unsigned char *__restrict__ block, int li,
                                                             int off_cells, int off_tiles, Queue *q,
                                                             unsigned nblocks, int unbiased, float thr) {
    extern __shared__ __attribute__((aligned(16))) float lds[];
    __shared__ unsigned next;
    slot_mmse<false, KR>(block, li, off_cells, off_tiles, q, nblocks, unbiased, thr, lds, &next);
}

__global__ __launch_bounds__(64) void slot_mmse_fp64_kernel(const unsigned char *__restrict__ block, int li,
                                                            int off_cells, int off_tiles, Queue *q,
                                                            unsigned nblocks, int unbiased) {
    __shared__ unsigned next;
    slot_mmse<true, 0>(block, li, off_cells, off_tiles, q, nblocks, unbiased, 0.f, nullptr, &next);
}

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

int apply_cost(const pe_cell_desc &c) { return (c.nLayer + c.nSym) * c.nAnt + c.nLayer * c.nSym * c.nPamBits; }

}  // namespace

int ks_class(int A) { return A <= 16 ? 1 : A <= 32 ? 2 : A <= 64 ? 4 : 8; }

SlotExec::~SlotExec() {
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    if (d_block_) (void)hipFree(d_block_);
    if (d_queues_) (void)hipFree(d_queues_);
    if (d_flags_) (void)hipFree(d_flags_);
    if (h_block_) (void)hipHostFree(h_block_);
}

int SlotExec::max_grid() const {
    int g = 0;
    for (const Launch &l : launches_) g = std::max(g, l.grid);
    return g;
}

int SlotExec::launch_index(int kind, int cls, int prec) const {
    for (size_t i = 0; i < launches_.size(); ++i)
        if (launches_[i].kind == kind && launches_[i].cls == cls && (kind != 2 || launches_[i].prec == prec))
            return (int)i;
    return -1;
}

pe_status SlotExec::init(const pe_cell_desc *pop, int npop, int max_cells, unsigned flags, std::string &err) {
    flags_ = flags;
    max_cells_ = max_cells;
    const int cus = device_cus();
    int max_sc = 1;
    bool any_mmse = false;
    std::vector<std::pair<int, int>> apply_cls;
    for (int i = 0; i < npop; ++i) {
        const pe_cell_desc &c = pop[i];
        max_sc = std::max(max_sc, c.nSC);
        any_mmse = any_mmse || c.H != nullptr;
        const std::pair<int, int> k(ks_class(c.nAnt), c.prec);
        if (std::find(apply_cls.begin(), apply_cls.end(), k) == apply_cls.end()) apply_cls.push_back(k);
    }
    std::sort(apply_cls.begin(), apply_cls.end(), std::greater<std::pair<int, int>>());
    const SlotTile tt = slot_tile();
    apply_tile_ = tt.apply_tile > 0 ? tt.apply_tile : kApplyTileDefault;
    if (tt.mmse_tile > 0) {
        mmse_tile_ = tt.mmse_tile;
        mmse_waves_ = tt.mmse_waves;
        mmse_keep_ = tt.mmse_keep;
    } else {  // the single-cell choice for a batch of the population's size at its largest shape
        long total = 0;
        int mL = 1, mA = 4;
        for (int i = 0; i < npop; ++i) {
            total += pop[i].nSC;
            if (pop[i].nLayer * pop[i].nAnt > mL * mA) {
                mL = pop[i].nLayer;
                mA = pop[i].nAnt;
            }
        }
        total = total * max_cells / std::max(1, npop);
        const MmseTile mt = mmse_tile_for(mL, mA, (int)std::min<long>(total, 1L << 30));
        mmse_tile_ = mt.tile;
        mmse_waves_ = mt.waves;
        mmse_keep_ = mt.keep;
    }
    max_layers_ = 1;
    for (int i = 0; i < npop; ++i) max_layers_ = std::max(max_layers_, pop[i].nLayer);
    mmse_lds_ = mmse::tile_lds_bytes(max_layers_, mmse_tile_);
    const int mmse_tiles_per_cell = (max_sc + mmse_tile_ - 1) / mmse_tile_;
    const int apply_tiles_per_cell = (max_sc + apply_tile_ - 1) / apply_tile_;
    max_tiles_ = max_cells * (mmse_tiles_per_cell + apply_tiles_per_cell);
    flags_per_cell_ = (size_t)max_sc;
    if (any_mmse) {
        const int grid = std::max(1, std::min(max_cells * mmse_tiles_per_cell, cus * 16 / mmse_waves_));
        launches_.push_back({0, 0, 0, grid, 64 * mmse_waves_, (unsigned)grid});
        const int grid64 = std::max(1, std::min(max_cells * mmse_tiles_per_cell, cus * 8));
        launches_.push_back({1, 0, 0, grid64, 64, (unsigned)grid64});
    }
    for (const auto &k : apply_cls) {
        const int grid = std::max(1, std::min(max_cells * apply_tiles_per_cell, cus * 24));
        launches_.push_back({2, k.first, k.second, grid, 64, (unsigned)grid});
    }
    off_cells_ = align_up(sizeof(LaunchRange) * launches_.size(), 64);
    off_tiles_ = align_up(off_cells_ + sizeof(CellDev) * (size_t)max_cells, 64);
    block_bytes_ = off_tiles_ + sizeof(Tile) * (size_t)std::max(1, max_tiles_);
    hipError_t e = hipMalloc(&d_block_, block_bytes_);
    if (e == hipSuccess) e = hipHostMalloc(&h_block_, block_bytes_, hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc(&d_queues_, sizeof(Queue) * std::max<size_t>(1, launches_.size()));
    if (e == hipSuccess) e = hipMalloc(&d_flags_, flags_per_cell_ * (size_t)max_cells);
    if (e == hipSuccess) e = hipMemset(d_queues_, 0, sizeof(Queue) * std::max<size_t>(1, launches_.size()));
    if (e == hipSuccess) e = hipMemset(d_block_, 0, block_bytes_);
    if (e == hipSuccess) e = hipMemset(d_flags_, 0, flags_per_cell_ * (size_t)max_cells);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        err = std::string("slot executor allocation: ") + hipGetErrorString(e);
        return PE_ERR_NOMEM;
    }
    memset(h_block_, 0, block_bytes_);
    return PE_OK;
}

pe_status SlotExec::stage(const pe_cell_desc *cells, int n, std::string &err) {
    if (n > max_cells_) {
        err = "batch of " + std::to_string(n) + " cells exceeds the capacity " + std::to_string(max_cells_);
        return PE_ERR_CAPACITY;
    }
    LaunchRange *ranges = reinterpret_cast<LaunchRange *>(h_block_);
    CellDev *hc = reinterpret_cast<CellDev *>(h_block_ + off_cells_);
    Tile *ht = reinterpret_cast<Tile *>(h_block_ + off_tiles_);
    for (int i = 0; i < n; ++i) {
        const pe_cell_desc &c = cells[i];
        if ((size_t)c.nSC > flags_per_cell_) {
            err = "cell " + std::to_string(i) + " has more subcarriers than the executor was sized for";
            return PE_ERR_CAPACITY;
        }
        if ((c.H && launch_index(0, 0, 0) < 0) || launch_index(2, ks_class(c.nAnt), c.prec) < 0 ||
            c.nLayer > max_layers_) {
            err = "cell " + std::to_string(i) + " has a shape class the executor was not built for";
            return PE_ERR_INVALID;
        }
        CellDev &d = hc[i];
        d.H = reinterpret_cast<const float2 *>(c.H);
        d.sigma2 = c.sigma2;
        d.C = reinterpret_cast<float2 *>(c.C);
        d.noise_inv = c.noise_inv;
        d.Y = reinterpret_cast<const float2 *>(c.Y_rx);
        d.llr = c.llr;
        d.flags = d_flags_ + flags_per_cell_ * (size_t)i;
        d.nSC = c.nSC;
        d.L = c.nLayer;
        d.S = c.nSym;
        d.A = c.nAnt;
        d.npb = c.nPamBits;
        d.prec = c.prec;
        d.mmse_p = mmse_lanes_for(c.nLayer, c.nAnt);
    }
    int nt = 0;
    std::vector<int> order;
    for (size_t li = 0; li < launches_.size(); ++li) {
        const Launch &l = launches_[li];
        if (l.kind == 1) {  // FP64 pass: same tiles as the FP32 pass just before it
            ranges[li] = ranges[li - 1];
            continue;
        }
        order.clear();
        for (int i = 0; i < n; ++i) {
            const pe_cell_desc &c = cells[i];
            if (l.kind == 0 ? (c.H != nullptr) : (ks_class(c.nAnt) == l.cls && c.prec == l.prec))
                order.push_back(i);
        }
        if (l.kind == 0)
            std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
                return cells[a].nAnt * cells[a].nLayer > cells[b].nAnt * cells[b].nLayer;
            });
        else
            std::stable_sort(order.begin(), order.end(),
                             [&](int a, int b) { return apply_cost(cells[a]) > apply_cost(cells[b]); });
        const int tile = l.kind == 0 ? mmse_tile_ : apply_tile_;
        ranges[li].tile0 = nt;
        for (int i : order)
            for (int s = 0; s < cells[i].nSC; s += tile) {
                if (nt >= max_tiles_) {
                    err = "tile capacity exceeded";
                    return PE_ERR_CAPACITY;
                }
                ht[nt++] = Tile{i, s, std::min(tile, cells[i].nSC - s), 0};
            }
        ranges[li].ntiles = nt - ranges[li].tile0;
    }
    staged_tiles_ = nt;
    used_bytes_ = off_tiles_ + sizeof(Tile) * (size_t)nt;
    return PE_OK;
}

pe_status SlotExec::upload(hipStream_t st, std::string &err) {
    hipError_t e = hipMemcpyAsync(d_block_, h_block_, used_bytes_, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) {
        err = std::string("slot upload: ") + hipGetErrorString(e);
        return PE_ERR_DEVICE;
    }
    return PE_OK;
}

hipError_t SlotExec::launch_one(int i, hipStream_t st) const {
    const Launch &l = launches_[i];
    const int unb = (flags_ & PE_MMSE_UNBIASED) ? 1 : 0;
    const float thr = 1e4f;
    const int oc = (int)off_cells_, ot = (int)off_tiles_;
    Queue *q = d_queues_ + i;
    dim3 g(l.grid), b(l.block);
    if (l.kind == 0) {
        const size_t lds = (size_t)mmse_lds_;
        if (mmse_keep_)
            hipLaunchKernelGGL(slot_mmse_fp32_kernel<4>, g, b, lds, st, d_block_, i, oc, ot, q, l.nwaves, unb, thr);
        else
            hipLaunchKernelGGL(slot_mmse_fp32_kernel<0>, g, b, lds, st, d_block_, i, oc, ot, q, l.nwaves, unb, thr);
        return hipGetLastError();
    }
    if (l.kind == 1) {
        hipLaunchKernelGGL(slot_mmse_fp64_kernel, g, b, 0, st, d_block_, i, oc, ot, q, l.nwaves, unb);
        return hipGetLastError();
    }
#define PE_APPLY_CASE(KK, PP)                                                                                    \
    if (l.cls == KK && l.prec == PP) {                                                                           \
        hipLaunchKernelGGL((slot_apply_kernel<KK, PP>), g, b, 0, st, d_block_, i, oc, ot, q, l.nwaves);          \
        return hipGetLastError();                                                                                \
    }
    PE_APPLY_CASE(1, 1) PE_APPLY_CASE(1, 2) PE_APPLY_CASE(1, 3) PE_APPLY_CASE(1, 4)
    PE_APPLY_CASE(2, 1) PE_APPLY_CASE(2, 2) PE_APPLY_CASE(2, 3) PE_APPLY_CASE(2, 4)
    PE_APPLY_CASE(4, 1) PE_APPLY_CASE(4, 2) PE_APPLY_CASE(4, 3) PE_APPLY_CASE(4, 4)
    PE_APPLY_CASE(8, 1) PE_APPLY_CASE(8, 2) PE_APPLY_CASE(8, 3) PE_APPLY_CASE(8, 4)
#undef PE_APPLY_CASE
    return hipErrorInvalidValue;
}

pe_status SlotExec::launch(hipStream_t st, std::string &err) {
    hipLaunchKernelGGL(queue_reset_kernel, dim3(1), dim3(64), 0, st, d_queues_, (int)launches_.size());
    if (hipGetLastError() != hipSuccess) {
        err = "queue reset launch failed";
        return PE_ERR_DEVICE;
    }
    for (size_t i = 0; i < launches_.size(); ++i) {
        hipError_t e = launch_one((int)i, st);
        if (e != hipSuccess) {
            err = std::string("slot launch: ") + hipGetErrorString(e);
            return PE_ERR_DEVICE;
        }
    }
    return PE_OK;
}

pe_status SlotExec::graph_launch(hipStream_t st, std::string &err) {
    hipError_t e;
    if (!exec_) {
        hipStream_t cap;
        e = hipStreamCreateWithFlags(&cap, hipStreamNonBlocking);
        if (e != hipSuccess) {
            err = hipGetErrorString(e);
            return PE_ERR_DEVICE;
        }
        e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
        pe_status s = PE_OK;
        if (e == hipSuccess) s = launch(cap, err);
        hipGraph_t g = nullptr;
        hipError_t e2 = hipStreamEndCapture(cap, &g);
        (void)hipStreamDestroy(cap);
        if (e != hipSuccess || s != PE_OK || e2 != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            err = "graph capture failed";
            return PE_ERR_DEVICE;
        }
        e = hipGraphInstantiate(&exec_, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            (void)hipGraphDestroy(g);
            err = std::string("graph instantiate: ") + hipGetErrorString(e);
            return PE_ERR_DEVICE;
        }
        graph_ = g;
    }
    e = hipGraphLaunch(exec_, st);
    if (e != hipSuccess) {
        err = std::string("graph launch: ") + hipGetErrorString(e);
        return PE_ERR_DEVICE;
    }
    return PE_OK;
}

pe_status check_cell_desc(const pe_cell_desc &c, int i, std::string &err) {
    char b[256];
    if (c.nSC < 1 || c.nLayer < 1 || c.nLayer > 8 || c.nSym < 1 || c.nSym > 16 || c.nAnt < 4 || c.nAnt % 4 ||
        c.nAnt > 128 || c.nPamBits < 1 || c.nPamBits > 4) {
        snprintf(b, sizeof(b), "cell %d: shape nSC=%d L=%d S=%d A=%d q=%d outside the plan limits", i, c.nSC,
                 c.nLayer, c.nSym, c.nAnt, c.nPamBits);
        err = b;
        return PE_ERR_INVALID;
    }
    if (c.prec < PE_PREC_F16 || c.prec > PE_PREC_BF16_SPLIT) {
        snprintf(b, sizeof(b), "cell %d: precision %d not available in a slot plan (MFMA precisions only)", i, c.prec);
        err = b;
        return PE_ERR_UNSUPPORTED;
    }
    if (!c.C || !c.noise_inv || !c.Y_rx || !c.llr || (c.H && !c.sigma2)) {
        snprintf(b, sizeof(b), "cell %d: null pointer", i);
        err = b;
        return PE_ERR_INVALID;
    }
    return PE_OK;
}

}  // namespace pe

// ---------------------------------------------------------------------------------------------------------
// slot plans: a fixed set of cells on one SlotExec
// ---------------------------------------------------------------------------------------------------------
struct pe_slot_plan {
    std::vector<pe_cell_desc> cells;
    pe::SlotExec exec;
};

namespace {

pe_status plan_fail(pe_status s, const std::string &m) {
    pe::set_error(m.c_str());
    return s;
}

}  // namespace

extern "C" {

pe_status pe_slot_plan_create(const pe_cell_desc *cells, int nCells, unsigned flags, pe_slot_plan **out) {
    if (!out) return plan_fail(PE_ERR_INVALID, "plan output pointer is null");
    *out = nullptr;
    if (!cells || nCells < 1) return plan_fail(PE_ERR_INVALID, "no cells");
    if (flags & ~PE_MMSE_UNBIASED) return plan_fail(PE_ERR_INVALID, "unknown plan flags");
    std::string err;
    for (int i = 0; i < nCells; ++i) {
        pe_status s = pe::check_cell_desc(cells[i], i, err);
        if (s != PE_OK) return plan_fail(s, err);
    }
    pe_slot_plan *p = new (std::nothrow) pe_slot_plan();
    if (!p) return plan_fail(PE_ERR_NOMEM, "host allocation");
    p->cells.assign(cells, cells + nCells);
    pe_status s = p->exec.init(cells, nCells, nCells, flags, err);
    if (s == PE_OK) s = p->exec.stage(cells, nCells, err);
    if (s == PE_OK) s = p->exec.upload(nullptr, err);
    if (s == PE_OK && hipDeviceSynchronize() != hipSuccess) {
        s = PE_ERR_DEVICE;
        err = "plan upload failed";
    }
    if (s != PE_OK) {
        delete p;
        return plan_fail(s, err);
    }
    *out = p;
    return PE_OK;
}

pe_status pe_slot_plan_update(pe_slot_plan *p, const pe_cell_desc *cells, int nCells, void *stream) {
    if (!p || !cells) return plan_fail(PE_ERR_INVALID, "null plan or cells");
    if (nCells != (int)p->cells.size()) return plan_fail(PE_ERR_INVALID, "cell count differs from the plan");
    std::string err;
    for (int i = 0; i < nCells; ++i) {
        const pe_cell_desc &a = cells[i], &b = p->cells[i];
        if (a.nSC != b.nSC || a.nLayer != b.nLayer || a.nSym != b.nSym || a.nAnt != b.nAnt ||
            a.nPamBits != b.nPamBits || a.prec != b.prec || (!a.H) != (!b.H))
            return plan_fail(PE_ERR_INVALID, "cell " + std::to_string(i) + ": shape differs from the plan");
        pe_status s = pe::check_cell_desc(a, i, err);
        if (s != PE_OK) return plan_fail(s, err);
    }
    hipStream_t st = (hipStream_t)stream;
    // the pinned staging block may still be read by an earlier asynchronous upload on this stream
    hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) return plan_fail(PE_ERR_DEVICE, hipGetErrorString(e));
    p->cells.assign(cells, cells + nCells);
    pe_status s = p->exec.stage(cells, nCells, err);
    if (s == PE_OK) s = p->exec.upload(st, err);
    return s == PE_OK ? PE_OK : plan_fail(s, err);
}

pe_status pe_slot_plan_launch(pe_slot_plan *p, void *stream) {
    if (!p) return plan_fail(PE_ERR_INVALID, "null plan");
    std::string err;
    pe_status s = p->exec.launch((hipStream_t)stream, err);
    return s == PE_OK ? PE_OK : plan_fail(s, err);
}

pe_status pe_slot_plan_graph_launch(pe_slot_plan *p, void *stream) {
    if (!p) return plan_fail(PE_ERR_INVALID, "null plan");
    std::string err;
    pe_status s = p->exec.graph_launch((hipStream_t)stream, err);
    return s == PE_OK ? PE_OK : plan_fail(s, err);
}

void pe_slot_plan_destroy(pe_slot_plan *p) {
    if (p) (void)hipDeviceSynchronize();
    delete p;
}

pe_status pe_slot_plan_info(const pe_slot_plan *p, int *nTiles, int *nWorkgroups) {
    if (!p) return plan_fail(PE_ERR_INVALID, "null plan");
    if (nTiles) *nTiles = p->exec.tiles_staged();
    if (nWorkgroups) *nWorkgroups = p->exec.max_grid();
    return PE_OK;
}

}  // extern "C"
