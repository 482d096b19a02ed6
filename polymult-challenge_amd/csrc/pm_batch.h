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
    int col_tc = 16;        // columns per column tile (capped by the block and ncols)
    int ceil_split = 0;     // odd logN: 1 puts the larger half on the column pass
    int blocks_per_cu = 0;  // 0: one block per tile; k > 0: k persistent blocks per CU
};

// Tile lists and the five phase launches for a list of products.
struct Runner {
    std::vector<BatchProduct> prods;
    std::vector<BatchTask> tasks;  // phase-major
    uint32_t off[PH_COUNT + 1] = {};  // tiles of phase k: [off[k], off[k+1])
    size_t lds[PH_COUNT] = {};
    BatchProduct* d_prods = nullptr;
    size_t prod_cap = 0;
    BatchTask* d_tasks = nullptr;
    size_t task_cap = 0;
    uint32_t* d_ctr = nullptr;  // [PH_COUNT] tile counters of the persistent mode
    BatchProduct* h_prods = nullptr;  // pinned mirrors
    size_t h_prod_cap = 0;
    BatchTask* h_tasks = nullptr;
    size_t h_task_cap = 0;
    BatchConfig cfg;
    int cus = 0, max_lds = 0;

    void init();
    // Sets every product's split and tile shapes and builds the tile lists and
    // the LDS sizes. Throws if a product cannot fit the LDS.
    void plan(const BatchConfig& c);
    // Copies products and tiles to the device (stream-ordered). With
    // `pinned`, through pinned mirrors, so the call does not wait for the
    // copy; the mirrors are rewritten by the next upload, so the caller must
    // know the previous one has completed.
    void upload(hipStream_t s, bool pinned = false);
    // Sizes the device buffers and the pinned mirrors (touched, so a later
    // upload takes no page faults) for up to nprods products and ntasks tiles.
    void reserve(size_t nprods, size_t ntasks);
    // Enqueues the phases (and the counter reset of the persistent mode).
    hipError_t launch(hipStream_t s) const;
    ~Runner();
};

}  // namespace pm
