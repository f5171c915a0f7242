// multicell.hip -- one slot of many uneven cells as persistent work-queue kernels (pe_slot_plan_*).
//
// The reference runs one launch per cell on its own stream (multicell_contention.cu:194-283,
// multicell_evidence.cu:497-566). Here a plan cuts every cell into subcarrier tiles once, on the host, and a
// slot is a short fixed sequence of persistent kernels that pull tiles from a device work queue:
//   stage 1 (cells with H): MMSE FP32 pass, then the FP64 pass over flagged subcarriers, one kernel pair per
//                           layer count L (the solve is unrolled for a fixed L);
//   stage 2 (all cells):    fused coefficient apply + demap, one kernel per (antenna class KS, precision); the
//                           QAM order is read per tile.
// Tiles of a kernel are ordered by decreasing cost (largest per-subcarrier work first) so the queue drains
// with a short tail. Each queue counter resets itself: the last wave to leave a kernel zeroes it, so a
// captured graph replays without memset nodes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "apply_device.hpp"
#include "mmse_device.hpp"
#include "pe_internal.hpp"
#include "pusch_eq.h"

namespace pe {

struct CellDev {
    const float2 *H;
    const float *sigma2;
    float2 *C;
    float *noise_inv;
    const float2 *Y;
    float *llr;
    uint8_t *flags;  // stage-1 per-subcarrier flags (plan-owned)
    int nSC, L, S, A, npb, prec;
};

struct Tile {
    int cell, sc0, n, pad;
};

struct Queue {
    unsigned next, done;
};

namespace {

// Pulls the next tile index for the calling wave (one atomic per tile, broadcast from lane 0).
__device__ __forceinline__ unsigned next_tile(Queue *q, int lane) {
    unsigned t = 0;
    if (lane == 0) t = atomicAdd(&q->next, 1u);
    return __shfl(t, 0, 64);
}

// Every wave of the grid calls this once after its last (failed) pull; the last one resets the queue.
__device__ __forceinline__ void leave_queue(Queue *q, int lane, unsigned nwaves) {
    if (lane == 0) {
        __threadfence();
        const unsigned d = atomicAdd(&q->done, 1u);
        if (d + 1 == nwaves) {
            atomicExch(&q->next, 0u);
            atomicExch(&q->done, 0u);
        }
    }
}

template <int KS, int PREC>
__global__ __launch_bounds__(64) void slot_apply_kernel(const CellDev *__restrict__ cells,
                                                        const Tile *__restrict__ tiles, int nTiles, Queue *q,
                                                        unsigned nwaves) {
    using namespace mfma;
    __shared__ __attribute__((aligned(16))) unsigned char lds[LdsLayout<KS>::bytes(PREC)];
    const int lane = threadIdx.x;
    for (;;) {
        const unsigned t = next_tile(q, lane);
        if (t >= (unsigned)nTiles) break;
        const Tile tl = tiles[t];
        const CellDev cd = cells[tl.cell];
        if (cd.A < 16 * KS) zero_operands<KS, PREC>(lds, lane);
        const long first = tl.sc0, last = (long)tl.sc0 + tl.n;
        switch (cd.npb) {
        case 1: apply_range<KS, PREC, 1>(cd.C, cd.Y, nullptr, cd.noise_inv, cd.llr, first, last, cd.L, cd.S, cd.A, lds, lane); break;
        case 2: apply_range<KS, PREC, 2>(cd.C, cd.Y, nullptr, cd.noise_inv, cd.llr, first, last, cd.L, cd.S, cd.A, lds, lane); break;
        case 3: apply_range<KS, PREC, 3>(cd.C, cd.Y, nullptr, cd.noise_inv, cd.llr, first, last, cd.L, cd.S, cd.A, lds, lane); break;
        default: apply_range<KS, PREC, 4>(cd.C, cd.Y, nullptr, cd.noise_inv, cd.llr, first, last, cd.L, cd.S, cd.A, lds, lane); break;
        }
        lds_order_wave();  // LDS of this tile consumed before the next tile may zero it
    }
    leave_queue(q, lane, nwaves);
}

// Stage 1: a wave takes a tile of up to 64 subcarriers; its 8 lane groups of 8 walk the tile with stride 8.
template <int L, bool FP64>
__global__ __launch_bounds__(256) void slot_mmse_kernel(const CellDev *__restrict__ cells,
                                                        const Tile *__restrict__ tiles, int nTiles, Queue *q,
                                                        unsigned nwaves, int unbiased, float thr) {
    using namespace mmse;
    constexpr int P = kMmseP;
    const int lane = threadIdx.x & 63;
    const int p = lane & (P - 1);
    const int grp = lane / P;
    for (;;) {
        const unsigned t = next_tile(q, lane);
        if (t >= (unsigned)nTiles) break;
        const Tile tl = tiles[t];
        const CellDev cd = cells[tl.cell];
        for (int i = grp; i < tl.n; i += 64 / P) {
            const size_t sc = (size_t)tl.sc0 + i;
            float2 *c = cd.C + sc * L * cd.A;
            float *ni = cd.noise_inv + sc * L;
            const float2 *h = cd.H + sc * cd.A * L;
            if constexpr (!FP64) {
                const bool ok = mmse_one<float, L, P>(h, cd.sigma2[sc], cd.A, p, unbiased != 0, thr, true, c, ni);
                if (p == 0) cd.flags[sc] = ok ? 0 : kPendingFp64;
            } else {
                if (!(cd.flags[sc] & kPendingFp64)) continue;  // uniform per group
                const bool ok = mmse_one<double, L, P>(h, (double)cd.sigma2[sc], cd.A, p, unbiased != 0, 0.0, false,
                                                       c, ni);
                if (!ok) {
                    for (int a = p; a < cd.A; a += P)
                        for (int l = 0; l < L; ++l) c[(size_t)l * cd.A + a] = make_float2(0.f, 0.f);
                    for (int l = 0; l < L; ++l)
                        if (p == l % P) ni[l] = 0.f;
                }
                if (p == 0) cd.flags[sc] = ok ? PE_SCFLAG_FP64 : (PE_SCFLAG_FP64 | PE_SCFLAG_SINGULAR);
            }
        }
    }
    leave_queue(q, lane, nwaves);
}

}  // namespace

}  // namespace pe

// ---------------------------------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------------------------------
struct pe_slot_plan {
    struct Launch {
        int kind;  // 0 mmse fp32, 1 mmse fp64, 2 apply
        int cls;   // L (mmse) or KS (apply)
        int prec;
        int tile0, ntiles;
        int queue;
        int grid, block;
        unsigned nwaves;
    };
    int device = 0;
    unsigned flags = 0;
    std::vector<pe_cell_desc> cells;
    std::vector<pe::CellDev> host_cells;
    pe::CellDev *d_cells = nullptr;
    pe::CellDev *h_pinned = nullptr;  // staging for pe_slot_plan_update
    pe::Tile *d_tiles = nullptr;
    pe::Queue *d_queues = nullptr;
    uint8_t *d_flags = nullptr;
    std::vector<Launch> launches;
    int ntiles = 0, nwg = 0;
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
};

namespace {

int ks_class(int A) { return A <= 16 ? 1 : A <= 32 ? 2 : A <= 64 ? 4 : 8; }

hipError_t launch_one(const pe_slot_plan *p, const pe_slot_plan::Launch &l, hipStream_t st) {
    using namespace pe;
    const Tile *tiles = p->d_tiles + l.tile0;
    Queue *q = p->d_queues + l.queue;
    const int unb = (p->flags & PE_MMSE_UNBIASED) ? 1 : 0;
    const float thr = 1e4f;
    dim3 g(l.grid), b(l.block);
#define PE_MMSE_CASE(LL)                                                                                        \
    case LL:                                                                                                    \
        if (l.kind == 0)                                                                                        \
            hipLaunchKernelGGL((slot_mmse_kernel<LL, false>), g, b, 0, st, p->d_cells, tiles, l.ntiles, q,       \
                               l.nwaves, unb, thr);                                                             \
        else                                                                                                    \
            hipLaunchKernelGGL((slot_mmse_kernel<LL, true>), g, b, 0, st, p->d_cells, tiles, l.ntiles, q,        \
                               l.nwaves, unb, thr);                                                             \
        break;
    if (l.kind <= 1) {
        switch (l.cls) {
            PE_MMSE_CASE(1) PE_MMSE_CASE(2) PE_MMSE_CASE(3) PE_MMSE_CASE(4)
            PE_MMSE_CASE(5) PE_MMSE_CASE(6) PE_MMSE_CASE(7) PE_MMSE_CASE(8)
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
#undef PE_MMSE_CASE
#define PE_APPLY_CASE(KK, PP)                                                                                   \
    if (l.cls == KK && l.prec == PP) {                                                                          \
        hipLaunchKernelGGL((slot_apply_kernel<KK, PP>), g, b, 0, st, p->d_cells, tiles, l.ntiles, q, l.nwaves); \
        return hipGetLastError();                                                                               \
    }
    PE_APPLY_CASE(1, 1) PE_APPLY_CASE(1, 2) PE_APPLY_CASE(1, 3) PE_APPLY_CASE(1, 4)
    PE_APPLY_CASE(2, 1) PE_APPLY_CASE(2, 2) PE_APPLY_CASE(2, 3) PE_APPLY_CASE(2, 4)
    PE_APPLY_CASE(4, 1) PE_APPLY_CASE(4, 2) PE_APPLY_CASE(4, 3) PE_APPLY_CASE(4, 4)
    PE_APPLY_CASE(8, 1) PE_APPLY_CASE(8, 2) PE_APPLY_CASE(8, 3) PE_APPLY_CASE(8, 4)
#undef PE_APPLY_CASE
    return hipErrorInvalidValue;
}

pe_status plan_fail(pe_status s, const std::string &m) {
    pe::set_error(m.c_str());
    return s;
}

void plan_free(pe_slot_plan *p) {
    if (!p) return;
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->d_cells) (void)hipFree(p->d_cells);
    if (p->d_tiles) (void)hipFree(p->d_tiles);
    if (p->d_queues) (void)hipFree(p->d_queues);
    if (p->d_flags) (void)hipFree(p->d_flags);
    if (p->h_pinned) (void)hipHostFree(p->h_pinned);
    delete p;
}

// Device view of the caller's cells (flags point into the plan-owned buffer).
void fill_host_cells(pe_slot_plan *p) {
    p->host_cells.resize(p->cells.size());
    size_t off = 0;
    for (size_t i = 0; i < p->cells.size(); ++i) {
        const pe_cell_desc &c = p->cells[i];
        pe::CellDev &d = p->host_cells[i];
        d.H = reinterpret_cast<const float2 *>(c.H);
        d.sigma2 = c.sigma2;
        d.C = reinterpret_cast<float2 *>(c.C);
        d.noise_inv = c.noise_inv;
        d.Y = reinterpret_cast<const float2 *>(c.Y_rx);
        d.llr = c.llr;
        d.flags = p->d_flags + off;
        d.nSC = c.nSC;
        d.L = c.nLayer;
        d.S = c.nSym;
        d.A = c.nAnt;
        d.npb = c.nPamBits;
        d.prec = c.prec;
        off += (size_t)c.nSC;
    }
}

pe_status check_cell(const pe_cell_desc &c, int i) {
    char b[256];
    if (c.nSC < 1 || c.nLayer < 1 || c.nLayer > 8 || c.nSym < 1 || c.nSym > 16 || c.nAnt < 4 || c.nAnt % 4 ||
        c.nAnt > 128 || c.nPamBits < 1 || c.nPamBits > 4) {
        snprintf(b, sizeof(b), "cell %d: shape nSC=%d L=%d S=%d A=%d q=%d outside the plan limits", i, c.nSC,
                 c.nLayer, c.nSym, c.nAnt, c.nPamBits);
        return plan_fail(PE_ERR_INVALID, b);
    }
    if (c.prec < PE_PREC_F16 || c.prec > PE_PREC_BF16_SPLIT) {
        snprintf(b, sizeof(b), "cell %d: precision %d not available in a slot plan (MFMA precisions only)", i, c.prec);
        return plan_fail(PE_ERR_UNSUPPORTED, b);
    }
    if (!c.C || !c.noise_inv || !c.Y_rx || !c.llr || (c.H && !c.sigma2)) {
        snprintf(b, sizeof(b), "cell %d: null pointer", i);
        return plan_fail(PE_ERR_INVALID, b);
    }
    return PE_OK;
}

}  // namespace

extern "C" {

pe_status pe_slot_plan_create(const pe_cell_desc *cells, int nCells, unsigned flags, pe_slot_plan **out) {
    using namespace pe;
    if (!out) return plan_fail(PE_ERR_INVALID, "plan output pointer is null");
    *out = nullptr;
    if (!cells || nCells < 1) return plan_fail(PE_ERR_INVALID, "no cells");
    if (flags & ~PE_MMSE_UNBIASED) return plan_fail(PE_ERR_INVALID, "unknown plan flags");
    for (int i = 0; i < nCells; ++i) {
        pe_status s = check_cell(cells[i], i);
        if (s != PE_OK) return s;
    }
    pe_slot_plan *p = new (std::nothrow) pe_slot_plan();
    if (!p) return plan_fail(PE_ERR_NOMEM, "host allocation");
    p->flags = flags;
    p->cells.assign(cells, cells + nCells);
    (void)hipGetDevice(&p->device);
    const int cus = device_cus();

    // ---- tiles --------------------------------------------------------------------------------------
    std::vector<Tile> tiles;
    std::vector<int> order(nCells);
    for (int i = 0; i < nCells; ++i) order[i] = i;
    const int kApplyTile = 8;   // subcarriers per stage-2 tile (one wave)
    const int kMmseTile = 64;   // subcarriers per stage-1 tile (one wave of 8 groups)
    int nq = 0;
    // stage 1: per L class, cells with H, larger per-subcarrier cost first
    for (int L = 1; L <= 8; ++L) {
        std::vector<int> cs;
        for (int i : order)
            if (cells[i].H && cells[i].nLayer == L) cs.push_back(i);
        if (cs.empty()) continue;
        std::stable_sort(cs.begin(), cs.end(), [&](int a, int b) { return cells[a].nAnt > cells[b].nAnt; });
        const int t0 = (int)tiles.size();
        for (int i : cs)
            for (int s = 0; s < cells[i].nSC; s += kMmseTile)
                tiles.push_back(Tile{i, s, std::min(kMmseTile, cells[i].nSC - s), 0});
        const int nt = (int)tiles.size() - t0;
        const int grid = std::max(1, std::min((nt + 3) / 4, cus * 2));
        p->launches.push_back({0, L, 0, t0, nt, nq++, grid, 256, (unsigned)grid * 4});
        p->launches.push_back({1, L, 0, t0, nt, nq++, grid, 256, (unsigned)grid * 4});
    }
    // stage 2: per (KS class, precision)
    for (int ks : {8, 4, 2, 1})
        for (int prec = 1; prec <= 4; ++prec) {
            std::vector<int> cs;
            for (int i : order)
                if (ks_class(cells[i].nAnt) == ks && cells[i].prec == prec) cs.push_back(i);
            if (cs.empty()) continue;
            auto cost = [&](int i) {
                const pe_cell_desc &c = cells[i];
                return (c.nLayer + c.nSym) * c.nAnt + c.nLayer * c.nSym * c.nPamBits;
            };
            std::stable_sort(cs.begin(), cs.end(), [&](int a, int b) { return cost(a) > cost(b); });
            const int t0 = (int)tiles.size();
            for (int i : cs)
                for (int s = 0; s < cells[i].nSC; s += kApplyTile)
                    tiles.push_back(Tile{i, s, std::min(kApplyTile, cells[i].nSC - s), 0});
            const int nt = (int)tiles.size() - t0;
            const int grid = std::max(1, std::min(nt, cus * 8));
            p->launches.push_back({2, ks, prec, t0, nt, nq++, grid, 64, (unsigned)grid});
        }
    p->ntiles = (int)tiles.size();
    for (auto &l : p->launches) p->nwg = std::max(p->nwg, l.grid);

    // ---- device state -----------------------------------------------------------------------------
    size_t total_sc = 0;
    for (int i = 0; i < nCells; ++i) total_sc += (size_t)cells[i].nSC;
    hipError_t e = hipMalloc(&p->d_flags, total_sc);
    if (e == hipSuccess) e = hipMalloc(&p->d_cells, sizeof(CellDev) * nCells);
    if (e == hipSuccess) e = hipMalloc(&p->d_tiles, sizeof(Tile) * std::max<size_t>(1, tiles.size()));
    if (e == hipSuccess) e = hipMalloc(&p->d_queues, sizeof(Queue) * std::max(1, nq));
    if (e == hipSuccess) e = hipHostMalloc(&p->h_pinned, sizeof(CellDev) * nCells, hipHostMallocDefault);
    if (e != hipSuccess) {
        plan_free(p);
        return plan_fail(PE_ERR_NOMEM, std::string("device allocation: ") + hipGetErrorString(e));
    }
    fill_host_cells(p);
    e = hipMemcpy(p->d_cells, p->host_cells.data(), sizeof(CellDev) * nCells, hipMemcpyHostToDevice);
    if (e == hipSuccess && !tiles.empty())
        e = hipMemcpy(p->d_tiles, tiles.data(), sizeof(Tile) * tiles.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(p->d_queues, 0, sizeof(Queue) * std::max(1, nq));
    if (e == hipSuccess) e = hipMemset(p->d_flags, 0, total_sc);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        plan_free(p);
        return plan_fail(PE_ERR_DEVICE, std::string("plan upload: ") + hipGetErrorString(e));
    }
    *out = p;
    return PE_OK;
}

pe_status pe_slot_plan_update(pe_slot_plan *p, const pe_cell_desc *cells, int nCells, void *stream) {
    if (!p || !cells) return plan_fail(PE_ERR_INVALID, "null plan or cells");
    if (nCells != (int)p->cells.size()) return plan_fail(PE_ERR_INVALID, "cell count differs from the plan");
    for (int i = 0; i < nCells; ++i) {
        const pe_cell_desc &a = cells[i], &b = p->cells[i];
        if (a.nSC != b.nSC || a.nLayer != b.nLayer || a.nSym != b.nSym || a.nAnt != b.nAnt ||
            a.nPamBits != b.nPamBits || a.prec != b.prec || (!a.H) != (!b.H))
            return plan_fail(PE_ERR_INVALID, "cell " + std::to_string(i) + ": shape differs from the plan");
        pe_status s = check_cell(a, i);
        if (s != PE_OK) return s;
    }
    hipStream_t st = (hipStream_t)stream;
    // the staging buffer may still be read by an earlier asynchronous copy on this stream
    hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) return plan_fail(PE_ERR_DEVICE, hipGetErrorString(e));
    p->cells.assign(cells, cells + nCells);
    fill_host_cells(p);
    memcpy(p->h_pinned, p->host_cells.data(), sizeof(pe::CellDev) * nCells);
    e = hipMemcpyAsync(p->d_cells, p->h_pinned, sizeof(pe::CellDev) * nCells, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return plan_fail(PE_ERR_DEVICE, hipGetErrorString(e));
    return PE_OK;
}

pe_status pe_slot_plan_launch(pe_slot_plan *p, void *stream) {
    if (!p) return plan_fail(PE_ERR_INVALID, "null plan");
    for (const auto &l : p->launches) {
        if (l.ntiles == 0) continue;
        hipError_t e = launch_one(p, l, (hipStream_t)stream);
        if (e != hipSuccess) return plan_fail(PE_ERR_DEVICE, std::string("slot launch: ") + hipGetErrorString(e));
    }
    return PE_OK;
}

pe_status pe_slot_plan_graph_launch(pe_slot_plan *p, void *stream) {
    if (!p) return plan_fail(PE_ERR_INVALID, "null plan");
    hipStream_t st = (hipStream_t)stream;
    hipError_t e;
    if (!p->exec) {
        hipStream_t cap;
        e = hipStreamCreateWithFlags(&cap, hipStreamNonBlocking);
        if (e != hipSuccess) return plan_fail(PE_ERR_DEVICE, hipGetErrorString(e));
        e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
        pe_status s = PE_OK;
        if (e == hipSuccess) s = pe_slot_plan_launch(p, cap);
        hipGraph_t g = nullptr;
        hipError_t e2 = hipStreamEndCapture(cap, &g);
        (void)hipStreamDestroy(cap);
        if (e != hipSuccess || s != PE_OK || e2 != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return plan_fail(PE_ERR_DEVICE, "graph capture failed");
        }
        e = hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            (void)hipGraphDestroy(g);
            return plan_fail(PE_ERR_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(e));
        }
        p->graph = g;
    }
    e = hipGraphLaunch(p->exec, st);
    if (e != hipSuccess) return plan_fail(PE_ERR_DEVICE, std::string("graph launch: ") + hipGetErrorString(e));
    return PE_OK;
}

void pe_slot_plan_destroy(pe_slot_plan *p) {
    if (p) (void)hipDeviceSynchronize();
    plan_free(p);
}

pe_status pe_slot_plan_info(const pe_slot_plan *p, int *nTiles, int *nWorkgroups) {
    if (!p) return plan_fail(PE_ERR_INVALID, "null plan");
    if (nTiles) *nTiles = p->ntiles;
    if (nWorkgroups) *nWorkgroups = p->nwg;
    return PE_OK;
}

}  // extern "C"
