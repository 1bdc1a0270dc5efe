// Launches of one kernel over the blocks of several problems.
//
// Every kernel of the solve path is a struct K with
//   using Args = ...;                     its argument block (one problem, one launch),
//   static constexpr int NT;              threads per block (the launch bound),
//   static __device__ void run(const Args&, Blk);   the work of one block,
// and runs either as solo<K> (one problem: grid and Args as given) or as grouped<K> (up to
// kGroupMax problems in one launch: the flat block index is mapped to a problem and to that
// problem's own block through the offset table of GroupArgs, the Args blocks are read from a
// device table). A block does the same work either way, so a grouped launch gives bit for bit
// the results of the solo launches it replaces.
//
// On the host, launch_k<K> launches solo<K>, or, while a LaunchSink is installed on the calling
// thread, appends the launch to the sink instead (the batch engine of serve.cpp records the
// solves of its problems this way and replays them as grouped launches).
#pragma once
#include <hip/hip_runtime.h>

#include <cstring>
#include <type_traits>
#include <vector>

#include "core.hpp"

namespace flups {

// the block of one problem's grid that a workgroup runs
struct Blk {
    int x, y, nx;   // block index (x, y) and the x extent of the problem's grid
};

constexpr int kGroupMax = 32;
struct GroupArgs {
    const char* tab;             // n Args blocks, `stride` bytes apart (device memory)
    int stride, n;
    int start[kGroupMax + 1];    // first flat block of each entry; start[n] = blocks of the launch
    int gx[kGroupMax];           // x extent of each entry's grid (its y extent: its blocks / gx)
};

template <class K>
__device__ __forceinline__ void grouped_body(const GroupArgs& g) {
    const int b = blockIdx.x;
    int lo = 0, hi = g.n - 1;   // the last entry whose first block is <= b
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (g.start[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    const int l = b - g.start[lo], gx = g.gx[lo];
    K::run(*reinterpret_cast<const typename K::Args*>(g.tab + (size_t)lo * g.stride), Blk{l % gx, l / gx, gx});
}

template <class K>
__global__ void __launch_bounds__(K::NT) solo(typename K::Args a) {
    K::run(a, Blk{(int)blockIdx.x, (int)blockIdx.y, (int)gridDim.x});
}
template <class K>
__global__ void __launch_bounds__(K::NT) grouped(GroupArgs g) {
    grouped_body<K>(g);
}
// kernels whose K sets kWaves = W > 0: at least W waves per SIMD (at most 512 / W registers per lane)
template <class K, int W>
__global__ void __launch_bounds__(K::NT) __attribute__((amdgpu_waves_per_eu(W))) solo_w(typename K::Args a) {
    K::run(a, Blk{(int)blockIdx.x, (int)blockIdx.y, (int)gridDim.x});
}
template <class K, int W>
__global__ void __launch_bounds__(K::NT) __attribute__((amdgpu_waves_per_eu(W))) grouped_w(GroupArgs g) {
    grouped_body<K>(g);
}
template <class K, class = void>
struct waves_of {
    static constexpr int value = 0;
};
template <class K>
struct waves_of<K, std::void_t<decltype(K::kWaves)>> {
    static constexpr int value = K::kWaves;
};

// one recorded launch
struct LaunchRec {
    const void* solo = nullptr;      // solo<K>
    const void* grouped = nullptr;   // grouped<K>
    int threads = 0, gx = 1, gy = 1;
    size_t smem = 0;
    int stage = 0;                   // launches of one stage of a problem are independent of each other
    std::vector<char> args;
};

struct LaunchSink {
    std::vector<LaunchRec> recs;
    int stage = 0;
    int same = 0;   // > 0: the following launches share the current stage (SameStage)
    void add(const void* s, const void* g, dim3 grid, dim3 block, size_t smem, const void* a, size_t n) {
        FLUPS_REQUIRE(grid.z == 1 && block.y == 1 && block.z == 1, "recorded launches are 2-D grids of 1-D blocks");
        LaunchRec r;
        r.solo = s;
        r.grouped = g;
        r.threads = (int)block.x;
        r.gx = (int)grid.x;
        r.gy = (int)grid.y;
        r.smem = smem;
        r.stage = stage;
        r.args.resize(n);
        std::memcpy(r.args.data(), a, n);
        recs.push_back(std::move(r));
        if (!same) ++stage;
    }
};

// the sink of the calling thread (nullptr: launches run)
LaunchSink*& launch_sink();

// launches issued while one SameStage is alive go to the same stage (the components of a pass)
struct SameStage {
    LaunchSink* k;
    SameStage() : k(launch_sink()) {
        if (k) ++k->same;
    }
    ~SameStage() {
        if (k && --k->same == 0) ++k->stage;
    }
    SameStage(const SameStage&) = delete;
    SameStage& operator=(const SameStage&) = delete;
};

template <class K>
void launch_k(dim3 grid, dim3 block, size_t smem, hipStream_t s, const typename K::Args& a) {
    constexpr int W = waves_of<K>::value;
    if (LaunchSink* k = launch_sink()) {
        if constexpr (W > 0) k->add((const void*)solo_w<K, W>, (const void*)grouped_w<K, W>, grid, block, smem, &a, sizeof(a));
        else k->add((const void*)solo<K>, (const void*)grouped<K>, grid, block, smem, &a, sizeof(a));
        return;
    }
    if constexpr (W > 0) hipLaunchKernelGGL((solo_w<K, W>), grid, block, smem, s, a);
    else hipLaunchKernelGGL((solo<K>), grid, block, smem, s, a);
}

}  // namespace flups
