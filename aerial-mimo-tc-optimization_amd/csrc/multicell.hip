// multicell.hip -- many uneven cells per launch (SlotExec: tile kernels over a batch of cell-jobs), and the
// slot-plan C ABI (pe_slot_plan_*) on top of it.
//
// The reference runs one launch per cell on its own stream (multicell_contention.cu:194-283,
// multicell_evidence.cu:497-566). Here a batch of cell-jobs is cut into subcarrier tiles on the host and
// executed by a fixed sequence of kernels, whatever the number and shapes of the cells:
//   the MMSE kernel over the cells with H (the L x L solve is unrolled per L and selected by the tile's cell;
//   subcarriers the FP32 algebra flags are solved in FP64 by the same workgroup),
//   one fused coefficient apply + demap kernel per precision (the antenna class and the QAM order are read
//   per tile).
// Tiles of a kernel are ordered by decreasing per-subcarrier cost. The kernels read their tile ranges from
// device memory, so one captured graph serves every batch, and no memset node is needed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "pe_internal.hpp"
#include "pusch_eq.h"
#include "slot_exec.hpp"
#include "solve_record.hpp"

namespace pe {

namespace {

// Work distribution: workgroup g of a kernel takes tiles g, g + grid, g + 2 grid, ... of its class. For a slot
// plan the grid is the class's exact tile count (one tile per workgroup); for the serving loop, whose batches
// vary, it is bounded by the device's size. Tiles of one class cost about the same, so this static assignment
// balances without atomics. Tile counts are read from device memory, so one captured graph serves every
// batch.
struct BlockView {
    const LaunchRange *ranges;
    const CellDev *cells;
    const Tile *tiles;
};

__device__ __forceinline__ BlockView view(const unsigned char *block, int off_cells, int off_tiles) {
    return BlockView{reinterpret_cast<const LaunchRange *>(block), reinterpret_cast<const CellDev *>(block + off_cells),
                     reinterpret_cast<const Tile *>(block + off_tiles)};
}

// Fused apply + demap of one tile by a workgroup of 4 waves: wave w takes subcarriers sc0 + w, sc0 + w + 4, ...
// of the tile (one subcarrier per wave at tile = 4), each in its own LDS operand region. One kernel per
// precision serves every antenna class up to KSMAX: the tile's cell selects the class (the same per-class
// code as the single-cell kernel, so results are bit-identical). LOOP = false (slot plans, grid = tile count):
// one tile per workgroup, written without loops so the subcarrier code keeps the single-cell kernel's
// registers; LOOP = true (serving, bounded grid): workgroups stride over the tiles.
constexpr int kApplyWaves = 4;

__host__ __device__ __forceinline__ int ks_class_dev(int A) { return A <= 16 ? 1 : A <= 32 ? 2 : A <= 64 ? 4 : 8; }

// One subcarrier of a cell through the noinline per-subcarrier function (used inside loops, where inlining
// would make the compiler hoist per-lane address arithmetic and double the registers).
// PREC 5 (PE_PREC_F16_C16): binary16 product on binary16 coefficients (C holds the C16 block, cexp the
// exponents), the code of the single-cell pe_coef_apply_demap_c16.
// Cells with solve records (rec): stage 2 forms the coefficients from H and the record, except on subcarriers
// solved in FP64 (their coefficient block is in C).
template <int KS, int PREC>
__device__ __forceinline__ void apply_one_sc_cell(const CellDev &cd, long sc, unsigned char *lds, int lane) {
    constexpr int P = PREC == 5 ? 1 : PREC;
    if (cd.A < 16 * KS) mfma::zero_operands<KS, P>(lds, lane);
    if constexpr (PREC == 5)
        mfma::apply_one_sc_c16<KS, -1>(reinterpret_cast<const uint32_t *>(cd.C), cd.cexp, cd.Y, nullptr,
                                       cd.noise_inv, cd.llr, sc, cd.L, cd.S, cd.A, lds, lane, cd.npb);
    else
        mfma::apply_one_sc<KS, PREC, -1>(cd.C, cd.Y, nullptr, cd.noise_inv, cd.llr, sc, cd.L, cd.S, cd.A, lds, lane,
                                         cd.npb);
}

template <int KS, int PREC>
__device__ __forceinline__ void apply_sc_class(const CellDev &cd, long sc, unsigned char *lds, int lane) {
    using namespace mfma;
    constexpr int P = PREC == 5 ? 1 : PREC;
    if (cd.A < 16 * KS) zero_operands<KS, P>(lds, lane);
    const int nY = cd.S * (cd.A >> 1), ry = (nY + 63) >> 6;
    RawSC<KS> ra;
    if constexpr (PREC == 5) {
        const int nC = cd.L * (cd.A >> 2), rc = (nC + 63) >> 6;
        load_raw_c16<KS>(ra, reinterpret_cast<const uint32_t *>(cd.C), cd.Y, (size_t)sc, cd.L, cd.S, cd.A, lane, nC,
                         nY, rc, ry);
        process_sc<KS, 1, -1, true>(ra, (size_t)sc, lds, nullptr, cd.noise_inv, cd.llr, cd.L, cd.S, cd.A, lane, nC,
                                    nY, rc, ry, cd.npb, 0, cd.cexp[sc]);
    } else {
        const int nC = cd.L * (cd.A >> 1), rc = (nC + 63) >> 6;
        load_raw<KS>(ra, cd.C, cd.Y, (size_t)sc, cd.L, cd.S, cd.A, lane, nC, nY, rc, ry);
        process_sc<KS, PREC, -1>(ra, (size_t)sc, lds, nullptr, cd.noise_inv, cd.llr, cd.L, cd.S, cd.A, lane, nC, nY,
                                 rc, ry, cd.npb);
    }
}

template <int KSMAX, int PREC, bool LOOP>
__global__ __launch_bounds__(64 * kApplyWaves) void slot_apply_kernel(const unsigned char *__restrict__ block,
                                                                      int li, int off_cells, int off_tiles) {
    using namespace mfma;
    __shared__ __attribute__((aligned(16))) unsigned char lds[kApplyWaves][LdsLayout<KSMAX>::bytes(PREC)];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const BlockView v = view(block, off_cells, off_tiles);
    const LaunchRange rg = v.ranges[li];
    if constexpr (!LOOP) {
        if ((int)blockIdx.x >= rg.ntiles) return;
        const Tile tl = v.tiles[rg.tile0 + blockIdx.x];
        const long sc = tl.sc0 + w;
        if (sc >= (long)tl.sc0 + tl.n) return;  // uniform per wave; no workgroup barrier follows
        const CellDev cd = v.cells[tl.cell];
        const int ks = ks_class_dev(cd.A);
        if (ks == 1) apply_sc_class<1, PREC>(cd, sc, lds[w], lane);
        else if (ks == 2) apply_sc_class<2, PREC>(cd, sc, lds[w], lane);
        else if (KSMAX >= 4 && ks == 4) apply_sc_class<(KSMAX >= 4 ? 4 : 1), PREC>(cd, sc, lds[w], lane);
        else if (KSMAX >= 8) apply_sc_class<(KSMAX >= 8 ? 8 : 1), PREC>(cd, sc, lds[w], lane);
    } else {
        for (int t = blockIdx.x; t < rg.ntiles; t += gridDim.x) {
            const Tile tl = v.tiles[rg.tile0 + t];
            const CellDev cd = v.cells[tl.cell];
            const int ks = ks_class_dev(cd.A);
            for (long sc = tl.sc0 + w; sc < (long)tl.sc0 + tl.n; sc += kApplyWaves) {
                if (ks == 1) apply_one_sc_cell<1, PREC>(cd, sc, lds[w], lane);
                else if (ks == 2) apply_one_sc_cell<2, PREC>(cd, sc, lds[w], lane);
                else if (KSMAX >= 4 && ks == 4) apply_one_sc_cell<(KSMAX >= 4 ? 4 : 1), PREC>(cd, sc, lds[w], lane);
                else if (KSMAX >= 8) apply_one_sc_cell<(KSMAX >= 8 ? 8 : 1), PREC>(cd, sc, lds[w], lane);
            }
            lds_order_wave();  // LDS of this tile consumed before the next tile may zero it
        }
    }
}

// Cells with solve records (SlotExec::uses_records): their own apply launch per precision, so that the other
// cells' kernel keeps its registers. Per subcarrier, the coefficients are formed on chip from H and the record,
// or read from C where the solve ran in FP64; the code is inlined (a called function is scheduled without the
// kernel's occupancy target and takes twice the registers).
template <int KS, int PREC>
__device__ __forceinline__ void apply_record_class(const CellDev &cd, long sc, unsigned char *lds, float2 *gfull,
                                                   int lane) {
    using namespace mfma;
    if (cd.A < 16 * KS) zero_operands<KS, PREC>(lds, lane);
    if (cd.flags[sc] & PE_SCFLAG_FP64) {  // solved in FP64: the coefficient block is in C
        const int nC = cd.L * (cd.A >> 1), rc = (nC + 63) >> 6, nY = cd.S * (cd.A >> 1), ry = (nY + 63) >> 6;
        RawSC<KS> ra;
        load_raw<KS>(ra, cd.C, cd.Y, (size_t)sc, cd.L, cd.S, cd.A, lane, nC, nY, rc, ry);
        process_sc<KS, PREC, -1>(ra, (size_t)sc, lds, nullptr, cd.noise_inv, cd.llr, cd.L, cd.S, cd.A, lane, nC, nY,
                                 rc, ry, cd.npb);
        return;
    }
    rec::apply_from_record<0, KS, PREC>(cd.L, cd.H, cd.Y, cd.C, cd.noise_inv, cd.llr, sc, cd.S, cd.A, cd.npb, 0, 0,
                                        lds, gfull, lane);
}

// One tile (4 subcarriers, one per wave) per workgroup: plans and eager batches (graph replays of varying batches
// keep the coefficient path). Five waves per SIMD asked for: without the hint the compiler schedules the inlined
// layer and antenna variants for 128-148 registers; with it, 93 for the binary16 and bfloat16 products, with no
// spills.
template <int PREC>
__global__ __launch_bounds__(64 * kApplyWaves) __attribute__((amdgpu_waves_per_eu(5, 8))) void slot_apply_rec_kernel(
    const unsigned char *__restrict__ block, int li, int off_cells, int off_tiles) {
    using namespace mfma;
    __shared__ __attribute__((aligned(16))) unsigned char lds[kApplyWaves][LdsLayout<4>::bytes(PREC)];
    __shared__ __attribute__((aligned(16))) float2 gfull_all[kApplyWaves][rec::kGinvBytes / sizeof(float2)];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const BlockView v = view(block, off_cells, off_tiles);
    const LaunchRange rg = v.ranges[li];
    auto one = [&](const CellDev &cd, long sc) {  // 32 <= A <= 64 (record_pays)
        if (cd.A <= 32) apply_record_class<2, PREC>(cd, sc, lds[w], gfull_all[w], lane);
        else apply_record_class<4, PREC>(cd, sc, lds[w], gfull_all[w], lane);
    };
    if ((int)blockIdx.x >= rg.ntiles) return;
    const Tile tl = v.tiles[rg.tile0 + blockIdx.x];
    const long sc = tl.sc0 + w;
    if (sc >= (long)tl.sc0 + tl.n) return;  // uniform per wave; no workgroup barrier follows
    one(v.cells[tl.cell], sc);
}

// Stage 1: a tile is up to mmse_tile subcarriers of one cell, solved by one workgroup of W waves (tile solve),
// which then solves the subcarriers its FP32 algebra flagged in FP64 (tile_fp64_inline). Lanes per subcarrier are the cell shape's choice, the same as in the single-cell kernels, so the
// coefficients are bit-identical to the cell solved alone.
// C16: coefficients written in the binary16 operand format (cells of precision PE_PREC_F16_C16).
template <int KR, bool C16 = false>
__device__ __forceinline__ void mmse_tile(const CellDev &cd, const Tile &tl, float *lds, int tid, int W,
                                          int unbiased, float thr) {
    using namespace mmse;
    uint32_t *c16 = reinterpret_cast<uint32_t *>(cd.C);
#define PE_MMSE_L(LL)                                                                                            \
    case LL:                                                                                                     \
        if (!C16 && cd.rec) {                                                                                    \
            float kr[1][LL], ki[1][LL];                                                                          \
            tile_solve<LL, 1, true>(cd.H, cd.sigma2, cd.noise_inv, cd.flags, tl.sc0, tl.n, cd.A, cd.mmse_p,      \
                                    unbiased != 0, thr, lds, tid, W, false, kr, ki, cd.C); /* FP32 records */    \
            tile_fp64_inline<LL, false, true>(cd.H, cd.sigma2, cd.C, cd.noise_inv, cd.flags, tl.sc0, tl.n, cd.A, \
                                              unbiased != 0, lds, tid, W);                                       \
            tile_barrier(W); /* slots read before the workgroup's next tile rewrites them */                     \
        } else {                                                                                                 \
            tile_fp32<LL, C16 ? 0 : KR, C16, 2, true>(cd.H, cd.sigma2, cd.C, cd.noise_inv, cd.flags, tl.sc0,     \
                                                      tl.n, cd.A, cd.mmse_p, unbiased != 0, thr, lds, tid, W,    \
                                                      c16, cd.cexp);                                             \
        }                                                                                                        \
        break;
    switch (cd.L) {
        PE_MMSE_L(1) PE_MMSE_L(2) PE_MMSE_L(3) PE_MMSE_L(4) PE_MMSE_L(5) PE_MMSE_L(6) PE_MMSE_L(7)
    default:
        PE_MMSE_L(8)
    }
#undef PE_MMSE_L
}

// All layer counts in one kernel (the tile's cell selects the unrolled solve). KR = 4: antenna rows kept in
// registers where a tile allows it (see tile_fp32). LOOP as for the apply kernel: one tile per workgroup, or
// workgroups striding over the tiles.
template <int KR, bool LOOP, bool C16 = false>
__global__ __launch_bounds__(512) void slot_mmse_fp32_kernel(const unsigned char *__restrict__ block, int li,
                                                             int off_cells, int off_tiles, int unbiased,
                                                             float thr) {
    extern __shared__ __attribute__((aligned(16))) float lds[];
    const BlockView v = view(block, off_cells, off_tiles);
    const LaunchRange rg = v.ranges[li];
    if constexpr (!LOOP) {
        if ((int)blockIdx.x >= rg.ntiles) return;  // uniform per workgroup (tile_fp32 has barriers)
        const Tile tl = v.tiles[rg.tile0 + blockIdx.x];
        const CellDev cd = v.cells[tl.cell];
        mmse_tile<KR, C16>(cd, tl, lds, threadIdx.x, blockDim.x >> 6, unbiased, thr);
    } else {
        for (int t = blockIdx.x; t < rg.ntiles; t += gridDim.x) {  // uniform per workgroup
            const Tile tl = v.tiles[rg.tile0 + t];
            const CellDev cd = v.cells[tl.cell];
            mmse_tile<KR, C16>(cd, tl, lds, threadIdx.x, blockDim.x >> 6, unbiased, thr);
        }
    }
}

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

int apply_cost(const pe_cell_desc &c) { return (c.nLayer + c.nSym) * c.nAnt + c.nLayer * c.nSym * c.nPamBits; }

}  // namespace

int ks_class(int A) { return A <= 16 ? 1 : A <= 32 ? 2 : A <= 64 ? 4 : 8; }

SlotExec::~SlotExec() {
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    if (d_block_) (void)hipFree(d_block_);
    if (d_flags_) (void)hipFree(d_flags_);
    if (h_block_) (void)hipHostFree(h_block_);
}

int SlotExec::max_grid() const {
    int g = 0;
    for (const Launch &l : launches_) g = std::max(g, l.grid);
    return g;
}

bool SlotExec::uses_records(const pe_cell_desc &c) const {
    return (flags_ & PE_SLOT_SOLVE_RECORDS) && c.H && c.prec != PE_PREC_F16_C16 && rec::record_pays(c.nLayer, c.nAnt);
}

int SlotExec::launch_index(int kind, int prec) const {
    for (size_t i = 0; i < launches_.size(); ++i)
        if (launches_[i].kind == kind && (kind < 2 || launches_[i].prec == prec)) return (int)i;
    return -1;
}

pe_status SlotExec::init(const pe_cell_desc *pop, int npop, int max_cells, unsigned flags, bool fixed,
                         std::string &err) {
    flags_ = flags;
    max_cells_ = max_cells;
    fixed_ = fixed;
    c16_ = false;
    for (int i = 0; i < npop; ++i) c16_ = c16_ || pop[i].prec == PE_PREC_F16_C16;
    for (int i = 0; i < npop; ++i)
        if (c16_ && pop[i].H && pop[i].prec != PE_PREC_F16_C16) {
            err = "cell " + std::to_string(i) + ": a plan whose cells compute coefficients uses one coefficient "
                  "format; mix PE_PREC_F16_C16 only with cells that bring their coefficients (H == NULL)";
            return PE_ERR_UNSUPPORTED;
        }
    int max_sc = 1;
    struct ApplyClass {
        int kind, prec, ks;  // launch kind 2 (coefficients in C) or 3 (solve records), largest antenna class
    };
    std::vector<ApplyClass> apply_cls;
    for (int i = 0; i < npop; ++i) {
        const pe_cell_desc &c = pop[i];
        max_sc = std::max(max_sc, c.nSC);
        const int kind = uses_records(c) ? 3 : 2;
        auto it = std::find_if(apply_cls.begin(), apply_cls.end(),
                               [&](const ApplyClass &k) { return k.kind == kind && k.prec == c.prec; });
        const int ks = ks_class(c.nAnt) <= 4 ? 4 : 8;
        if (it == apply_cls.end())
            apply_cls.push_back({kind, c.prec, ks});
        else
            it->ks = std::max(it->ks, ks);
    }
    std::sort(apply_cls.begin(), apply_cls.end(), [](const ApplyClass &a, const ApplyClass &b) {
        return a.kind != b.kind ? a.kind > b.kind : a.prec != b.prec ? a.prec > b.prec : a.ks > b.ks;
    });
    const SlotTile tt = slot_tile();
    // one subcarrier per wave of a 4-wave workgroup: the loop-free kernels (fixed plans, eager launches) take one
    // tile per workgroup, the striding kernels of graph replays one tile per iteration
    static_assert(kApplyTile == kApplyWaves, "one subcarrier per wave");
    apply_tile_ = kApplyTile;
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
    // tile capacity of a class: exact for a fixed batch, else max_cells jobs of its largest cell; the grid is
    // the capacity for a fixed batch, else bounded by `waves_per_cu` waves per compute unit
    const int cus = device_cus();
    auto class_tiles = [&](auto in_class, int tile) {
        long sum = 0, mx = 0;
        for (int i = 0; i < npop; ++i)
            if (in_class(pop[i])) {
                const long t = (pop[i].nSC + tile - 1) / tile;
                sum += t;
                mx = std::max(mx, t);
            }
        return (int)std::max<long>(1, fixed ? sum : mx * max_cells);
    };
    auto grid_for = [&](int cap, int waves_per_wg, int waves_per_cu) {
        return fixed ? cap : std::max(1, std::min(cap, cus * waves_per_cu / waves_per_wg));
    };
    // the MMSE kernel over every cell with H, then one fused apply/demap launch per precision
    max_tiles_ = 0;
    flags_per_cell_ = (size_t)max_sc;
    bool mm = false;
    for (int i = 0; i < npop; ++i) mm = mm || pop[i].H != nullptr;
    if (mm) {
        const int cap = class_tiles([](const pe_cell_desc &c) { return c.H != nullptr; }, mmse_tile_);
        launches_.push_back({0, 0, 0, grid_for(cap, mmse_waves_, 16), 64 * mmse_waves_, cap});
        max_tiles_ += cap;
    }
    for (const auto &k : apply_cls) {  // one fused apply/demap launch per (kind, precision), every antenna class
        const int cap = class_tiles(
            [&](const pe_cell_desc &c) { return c.prec == k.prec && (uses_records(c) ? 3 : 2) == k.kind; },
            apply_tile_);
        launches_.push_back({k.kind, k.ks, k.prec, grid_for(cap, kApplyWaves, 32), 64 * kApplyWaves, cap});
        max_tiles_ += cap;
    }
    off_cells_ = align_up(sizeof(LaunchRange) * launches_.size(), 64);
    off_tiles_ = align_up(off_cells_ + sizeof(CellDev) * (size_t)max_cells, 64);
    block_bytes_ = off_tiles_ + sizeof(Tile) * (size_t)std::max(1, max_tiles_);
    hipError_t e = hipMalloc(&d_block_, block_bytes_);
    if (e == hipSuccess) e = hipHostMalloc(&h_block_, block_bytes_, hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc(&d_flags_, flags_per_cell_ * (size_t)max_cells);
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
        const int ai = launch_index(uses_records(c) ? 3 : 2, c.prec);
        if ((c.H && launch_index(0, 0) < 0) || ai < 0 || (ks_class(c.nAnt) == 8 && launches_[ai].cls < 8) ||
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
        d.cexp = c.prec == PE_PREC_F16_C16
                     ? reinterpret_cast<int *>(reinterpret_cast<unsigned char *>(c.C) +
                                               (size_t)c.nSC * c.nLayer * c.nAnt * 4)
                     : nullptr;
        d.nSC = c.nSC;
        d.L = c.nLayer;
        d.S = c.nSym;
        d.A = c.nAnt;
        d.npb = c.nPamBits;
        d.prec = c.prec;
        d.mmse_p = mmse_lanes_for(c.nLayer, c.nAnt);
        d.rec = uses_records(c) ? 1 : 0;
    }
    int nt = 0;
    std::vector<int> order;
    staged_ntiles_.assign(launches_.size(), 0);
    for (size_t li = 0; li < launches_.size(); ++li) {
        const Launch &l = launches_[li];
        order.clear();
        for (int i = 0; i < n; ++i) {
            const pe_cell_desc &c = cells[i];
            if (l.kind == 0 ? c.H != nullptr : c.prec == l.prec && (uses_records(c) ? 3 : 2) == l.kind)
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
                if (nt >= max_tiles_ || nt - ranges[li].tile0 >= l.cap) {
                    err = "tile capacity exceeded";
                    return PE_ERR_CAPACITY;
                }
                ht[nt++] = Tile{i, s, std::min(tile, cells[i].nSC - s), 0};
            }
        ranges[li].ntiles = nt - ranges[li].tile0;
        staged_ntiles_[li] = ranges[li].ntiles;
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

hipError_t SlotExec::launch_one(int i, hipStream_t st, bool exact) const {
    const Launch &l = launches_[i];
    const int unb = (flags_ & PE_MMSE_UNBIASED) ? 1 : 0;
    const float thr = 1e4f;
    const int oc = (int)off_cells_, ot = (int)off_tiles_;
    // loop-free kernels (one tile per workgroup, fewer registers) for fixed plans and for eager launches, whose
    // grid is the staged tile count; the striding kernels for graph replays of varying batches
    const bool lf = fixed_ || exact;
    const int st_tiles = std::max(1, staged_ntiles_[i]);
    dim3 g(exact ? st_tiles : l.grid), b(l.block);
    if (l.kind == 0) {
        const size_t lds = (size_t)mmse_lds_;
        if (c16_ && lf)
            hipLaunchKernelGGL((slot_mmse_fp32_kernel<0, false, true>), g, b, lds, st, d_block_, i, oc, ot, unb, thr);
        else if (c16_)
            hipLaunchKernelGGL((slot_mmse_fp32_kernel<0, true, true>), g, b, lds, st, d_block_, i, oc, ot, unb, thr);
        else if (mmse_keep_ && lf)
            hipLaunchKernelGGL((slot_mmse_fp32_kernel<4, false>), g, b, lds, st, d_block_, i, oc, ot, unb, thr);
        else if (mmse_keep_)
            hipLaunchKernelGGL((slot_mmse_fp32_kernel<4, true>), g, b, lds, st, d_block_, i, oc, ot, unb, thr);
        else if (lf)
            hipLaunchKernelGGL((slot_mmse_fp32_kernel<0, false>), g, b, lds, st, d_block_, i, oc, ot, unb, thr);
        else
            hipLaunchKernelGGL((slot_mmse_fp32_kernel<0, true>), g, b, lds, st, d_block_, i, oc, ot, unb, thr);
        return hipGetLastError();
    }
    if (l.kind == 3) {  // cells with solve records: plans and eager batches only (one tile per workgroup)
        if (!lf) return hipErrorInvalidValue;
#define PE_REC_CASE(PP)                                                                                          \
    if (l.prec == PP) {                                                                                          \
        hipLaunchKernelGGL((slot_apply_rec_kernel<PP>), g, b, 0, st, d_block_, i, oc, ot);                   \
        return hipGetLastError();                                                                                \
    }
        PE_REC_CASE(1) PE_REC_CASE(2) PE_REC_CASE(3) PE_REC_CASE(4)
#undef PE_REC_CASE
        return hipErrorInvalidValue;
    }
#define PE_APPLY_CASE(KK, PP)                                                                                    \
    if (l.cls == KK && l.prec == PP) {                                                                           \
        if (lf)                                                                                                  \
            hipLaunchKernelGGL((slot_apply_kernel<KK, PP, false>), g, b, 0, st, d_block_, i, oc, ot);        \
        else                                                                                                     \
            hipLaunchKernelGGL((slot_apply_kernel<KK, PP, true>), g, b, 0, st, d_block_, i, oc, ot);         \
        return hipGetLastError();                                                                                \
    }
    PE_APPLY_CASE(4, 1) PE_APPLY_CASE(4, 2) PE_APPLY_CASE(4, 3) PE_APPLY_CASE(4, 4)
    PE_APPLY_CASE(8, 1) PE_APPLY_CASE(8, 2) PE_APPLY_CASE(8, 3) PE_APPLY_CASE(8, 4)
    PE_APPLY_CASE(4, 5) PE_APPLY_CASE(8, 5)
#undef PE_APPLY_CASE
    return hipErrorInvalidValue;
}

// The whole sequence on `st`: the MMSE kernel, one fused apply/demap launch per precision.
pe_status SlotExec::launch(hipStream_t st, std::string &err, bool exact) {
    hipError_t e = hipSuccess;
    for (size_t i = 0; e == hipSuccess && i < launches_.size(); ++i) e = launch_one((int)i, st, exact);
    if (e != hipSuccess) {
        err = std::string("slot launch: ") + hipGetErrorString(e);
        return PE_ERR_DEVICE;
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
    if (c.prec < PE_PREC_F16 || c.prec > PE_PREC_F16_C16) {
        snprintf(b, sizeof(b), "cell %d: precision %d not available in a slot plan (MFMA precisions only)", i, c.prec);
        err = b;
        return PE_ERR_UNSUPPORTED;
    }
    if (!c.C || !c.noise_inv || !c.Y_rx || !c.llr || (c.H && !c.sigma2)) {
        snprintf(b, sizeof(b), "cell %d: null pointer", i);
        err = b;
        return PE_ERR_INVALID;
    }
    if (c.prec == PE_PREC_F16_C16 && !is_aligned(c.C, 16)) {
        snprintf(b, sizeof(b), "cell %d: C must be 16-byte aligned for binary16 coefficients", i);
        err = b;
        return PE_ERR_INVALID;
    }
    // the widest accesses of the tile kernels: 16-byte loads of Y_rx, C and H rows (even layer counts),
    // LLR stores of 2 nPamBits floats
    const struct {
        const void *p;
        size_t a;
        const char *name;
    } req[] = {{c.Y_rx, 16, "Y_rx"}, {c.C, 16, "C"}, {c.H, h_align(c.nLayer), "H"}, {c.sigma2, 4, "sigma2"},
               {c.noise_inv, 4, "noise_inv"}, {c.llr, llr_align(c.nPamBits, false), "llr"}};
    for (const auto &r : req)
        if (r.p && !is_aligned(r.p, r.a)) {
            snprintf(b, sizeof(b), "cell %d: %s must be %zu-byte aligned", i, r.name, r.a);
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
    if (flags & ~(PE_MMSE_UNBIASED | PE_SLOT_SOLVE_RECORDS)) return plan_fail(PE_ERR_INVALID, "unknown plan flags");
    std::string err;
    for (int i = 0; i < nCells; ++i) {
        pe_status s = pe::check_cell_desc(cells[i], i, err);
        if (s != PE_OK) return plan_fail(s, err);
    }
    pe_slot_plan *p = new (std::nothrow) pe_slot_plan();
    if (!p) return plan_fail(PE_ERR_NOMEM, "host allocation");
    p->cells.assign(cells, cells + nCells);
    pe_status s = p->exec.init(cells, nCells, nCells, flags, true, err);
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
