// Building blocks of the batched and served forms.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "pm_kernels.h"
#include "pm_plan.h"

namespace pm {

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Device tables for a set of points over the shared prime set: one twiddle
// table for all of them, RNS/CRT tables per point.
struct Catalog {
    SharedPrimes S;
    std::vector<HostPlan> plans;
    std::vector<DevPlan> dev;  // pointers into the device tables
    void* d_shared = nullptr;
    void* d_tables = nullptr;
    size_t bytes = 0;

    void build(uint32_t count, const uint32_t* N, const uint32_t* W, const uint32_t* q_concat);
    size_t residue_words(size_t i) const { return 2 * (size_t)plans[i].K * plans[i].N; }
    size_t io_words(size_t i) const { return (size_t)plans[i].N * plans[i].L; }
    ~Catalog();
};

struct BatchConfig {
    int col_tc = 32;        // columns per column tile (capped by the block and ncols)
    int ceil_split = 0;     // odd logN: 1 puts the larger half on the column pass
    int blocks_per_cu = 0;  // 0: one block per tile; k > 0: k persistent blocks per CU
    // Products run in groups whose residues (2 K N words each) stay within
    // group_mb MiB, one group's five phases after another's, so that a
    // group's residues stay in the 256 MB last-level cache (MALL) from phase
    // to phase; 0: one group.
    int group_mb = 96;
    int single_kernels = 1;  // a group of one product runs the single-product kernels
    // Graph branches: groups alternate over this many independent chains, so
    // one group's latency-bound kernels can overlap the next group's.
    int branches = 2;
};

// Tile lists and the five phase launches for a list of products.
struct Runner {
    std::vector<BatchProduct> prods;
    std::vector<BatchTask> tasks;  // group-major, then phase-major
    std::vector<uint32_t> off;     // tiles of group g, phase k: [off[5g + k], off[5g + k + 1])
    // A group of one product runs the single-product kernels instead of the
    // tile lists (kernel arguments instead of a tile lookup per block):
    // single[g] is that product, or -1.
    std::vector<int> single;
    std::vector<LaunchCfg> single_cfg;
    size_t lds[PH_COUNT] = {};
    BatchProduct* d_prods = nullptr;
    size_t prod_cap = 0;
    BatchTask* d_tasks = nullptr;
    size_t task_cap = 0;
    uint32_t* d_ctr = nullptr;  // one tile counter per group and phase (persistent mode)
    size_t ctr_cap = 0;
    BatchProduct* h_prods = nullptr;  // pinned mirrors
    size_t h_prod_cap = 0;
    BatchTask* h_tasks = nullptr;
    size_t h_task_cap = 0;
    BatchConfig cfg;
    int cus = 0, max_lds = 0;

    void init();
    // Sets every product's split and tile shapes, groups the products and
    // builds the tile lists and the LDS sizes. Throws if a product cannot fit
    // the LDS.
    void plan(const BatchConfig& c);
    // Copies products and tiles to the device (stream-ordered). With
    // `pinned`, through pinned mirrors, so the call does not wait for the
    // copy; the mirrors are rewritten by the next upload, so the caller must
    // know the previous one has completed.
    void upload(hipStream_t s, bool pinned = false);
    // Sizes the device buffers and the pinned mirrors (touched, so a later
    // upload takes no page faults) for up to nprods products and ntasks tiles.
    void reserve(size_t nprods, size_t ntasks);
    size_t groups() const { return off.empty() ? 0 : (off.size() - 1) / PH_COUNT; }
    // Enqueues every group's phases (and the counter reset of the persistent
    // mode) on one stream.
    hipError_t launch(hipStream_t s) const;
    hipError_t launch_reset(hipStream_t s) const;   // persistent mode only
    hipError_t launch_group(size_t g, hipStream_t s) const;
    ~Runner();
};

}  // namespace pm
