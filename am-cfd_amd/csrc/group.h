// Grouped launches: one launch of a kernel family over many solver
// instances.
//
// Every kernel of the mixed mode's SIMPLE iteration is written as a struct K
// with a device body K::run(vb, args...) that takes its workgroup's
// coordinates from a virtual block vb instead of blockIdx / gridDim.  The
// named __global__ kernel runs it with the hardware block (vb_hw()); the
// grouped kernel k_group<K, NT, A...> runs it for many instances in one
// launch: instance i owns the blocks [start_i, start_i + blocks_i) of a 1-D
// grid and reads its arguments from a GroupEntry in device memory.  Physical
// blocks are renumbered so that each XCD takes one contiguous run of the
// instances' blocks (workgroups go to the 8 XCDs round-robin by id).
//
// Launchers call dispatch<K, NT>(k_named, ...).  Normally that launches
// k_named; while a LaunchRecorder is installed on the calling thread it
// records the launch (the grouped kernel, the grid and the argument entry)
// instead, so that a grouped runner can replay one iteration of many solvers
// as one launch per kernel position (group.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstring>
#include <utility>
#include <vector>

namespace amcfd {

// the coordinates a kernel body sees: its block and the grid, and whether
// they are the hardware's (bodies that renumber blocks for XCD locality do so
// only then; grouped launches renumber before the instance lookup)
struct VB {
    int x, y, z, gx, gy, gz;
    int hw;
    __device__ __forceinline__ int lin() const { return x + gx * (y + gy * z); }
    __device__ __forceinline__ int count() const { return gx * gy * gz; }
};
__device__ __forceinline__ VB vb_hw() {
    return VB{(int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z, (int)gridDim.x, (int)gridDim.y, (int)gridDim.z, 1};
}

// workgroup id -> a numbering in which each of the 8 XCDs holds one
// contiguous run (the hardware deals ids to XCDs round-robin)
__host__ __device__ __forceinline__ int xcd_contiguous(int b, int nb) {
    const int q8 = nb / 8, r8 = nb % 8, x8 = b % 8, i8 = b / 8;
    return (x8 < r8) ? x8 * (q8 + 1) + i8 : r8 * (q8 + 1) + (x8 - r8) * q8 + i8;
}

// a plain tuple of kernel arguments (trivially copyable, same layout on the
// host and the device)
template <class... A>
struct Tup {};
template <class H, class... T>
struct Tup<H, T...> {
    H h;
    Tup<T...> t;
};
__host__ __device__ inline Tup<> make_tup() { return Tup<>{}; }
template <class H, class... T>
__host__ __device__ inline Tup<H, T...> make_tup(H h, T... t) {
    Tup<H, T...> r;
    r.h = h;
    r.t = make_tup(t...);
    return r;
}
template <size_t I, class H, class... T>
__host__ __device__ __forceinline__ const auto &tup_get(const Tup<H, T...> &t) {
    if constexpr (I == 0) return t.h;
    else return tup_get<I - 1>(t.t);
}

template <class... A>
struct GroupEntry {
    int start;        // first block of this instance in the grouped grid
    int gx, gy, gz;   // the instance's own grid
    Tup<A...> args;
};

template <class K, class... A, size_t... I>
__device__ __forceinline__ void group_apply(const VB &vb, const Tup<A...> &t, std::index_sequence<I...>) {
    K::run(vb, tup_get<I>(t)...);
}

template <class K, int NT, class... A>
__global__ __launch_bounds__(NT) void k_group(const GroupEntry<A...> *__restrict__ e, int n) {
    const int b = xcd_contiguous((int)blockIdx.x, (int)gridDim.x);
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (e[mid].start <= b) lo = mid;
        else hi = mid - 1;
    }
    const GroupEntry<A...> &g = e[lo];
    const int local = b - g.start;
    VB vb;
    vb.gx = g.gx;
    vb.gy = g.gy;
    vb.gz = g.gz;
    vb.x = local % g.gx;
    vb.y = (local / g.gx) % g.gy;
    vb.z = local / (g.gx * g.gy);
    vb.hw = 0;
    group_apply<K>(vb, g.args, std::index_sequence_for<A...>{});
}

// one recorded launch: the grouped kernel of its family, its shape and the
// bytes of its GroupEntry (start filled in when the launches are merged)
struct LaunchRecord {
    const void *group_kernel = nullptr;
    int family = 0, rank = 0;   // rank: the size class; the largest of a position serves all
    dim3 grid, block;
    size_t lds = 0;
    std::vector<unsigned char> entry;
};
struct LaunchRecorder {
    std::vector<LaunchRecord> recs;
    bool unsupported = false;   // a launch that has no grouped form was reached
};
// the calling thread's recorder (null: launch normally)
LaunchRecorder *&tls_recorder();

template <class T>
struct NoDeduce {
    using type = T;
};

// launch k_named, or record it as a grouped launch of K
template <class K, int NT, class... A>
void dispatch(void (*k_named)(A...), dim3 grid, dim3 block, size_t lds, hipStream_t st,
              typename NoDeduce<A>::type... args) {
    if (LaunchRecorder *r = tls_recorder()) {
        LaunchRecord rec;
        rec.group_kernel = reinterpret_cast<const void *>(&k_group<K, NT, A...>);
        rec.family = K::family;
        rec.rank = K::rank;
        rec.grid = grid;
        rec.block = block;
        rec.lds = lds;
        GroupEntry<A...> e;
        std::memset(&e, 0, sizeof(e));
        e.start = 0;
        e.gx = (int)grid.x;
        e.gy = (int)grid.y;
        e.gz = (int)grid.z;
        e.args = make_tup<A...>(args...);
        rec.entry.resize(sizeof(e));
        std::memcpy(rec.entry.data(), &e, sizeof(e));
        r->recs.push_back(std::move(rec));
        return;
    }
    hipLaunchKernelGGL(k_named, grid, block, lds, st, args...);
}

}  // namespace amcfd
