// Several cases' launches of one kernel merged into one grid.
//
// A batch of cases steps together by running each kernel of the step once
// for all of them: the launches a case's step would make are recorded (the
// kernel's merged-grid launcher, its argument record and its block count),
// the records of the same launch of every case are placed side by side in
// device memory, and one grid covers all their blocks.  Block b of that grid
// reads map[b] = job << MULTI_LB_BITS | block-within-job, and runs the
// kernel's body with that job's arguments as that block.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstring>
#include <vector>

namespace pic {

constexpr int MULTI_LB_BITS = 22;   // blocks per job < 4 M, jobs < 1024

__device__ inline int multi_job(const uint32_t *map, int &lb) {
    const uint32_t m = map[blockIdx.x];   // one load per block; uniform
    lb = (int)(m & ((1u << MULTI_LB_BITS) - 1));
    return (int)(m >> MULTI_LB_BITS);
}

// launches a merged grid of `total` blocks over the job records at `jobs`
typedef void (*MultiLaunchFn)(const void *jobs, const uint32_t *map, int total, hipStream_t s);

// the MultiLaunchFn of merged-grid kernel K (job record type J, LB threads per block)
template <typename J, void (*K)(const J *, const uint32_t *), int LB>
void launch_m(const void *jobs, const uint32_t *map, int total, hipStream_t s) {
    hipLaunchKernelGGL(K, dim3(total), dim3(LB), 0, s, (const J *)jobs, map);
}

struct RecordedLaunch {
    MultiLaunchFn fn;
    size_t job_size, job_align;
    int blocks;
    std::vector<unsigned char> job;
};

struct Recorder {
    std::vector<RecordedLaunch> launches;
    template <typename J>
    void add(MultiLaunchFn fn, const J &job, int blocks) {
        RecordedLaunch r{fn, sizeof(J), alignof(J), blocks, {}};
        r.job.resize(sizeof(J));
        memcpy(r.job.data(), &job, sizeof(J));
        launches.push_back(std::move(r));
    }
};

// The launch recorder of the calling host thread: when set, the launchers of
// the batched step's kernels record their launch instead of making it.
Recorder *&launch_recorder();

}  // namespace pic
