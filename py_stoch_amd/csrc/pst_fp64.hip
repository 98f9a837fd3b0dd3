// pst_fp64.hip -- table-driven kernels of the size classes compiled into the library; the engine is in pst_engine_f64.hpp.
#include "pst_engine_f64.hpp"

namespace pst {

template <int NP, int MP>
__global__ void __launch_bounds__(256) sim_fp64_kernel(const JobDev<NP, MP>* __restrict__ jobs, int n_jobs,
                                                       unsigned int* __restrict__ queue,
                                                       unsigned long long* __restrict__ stats, Range range) {
  run_driver<EngineF64<JobRef<NP, MP>>, JobRef<NP, MP>>(jobs, n_jobs, queue, stats, range.f_end, range.resume);
}

template <int NP, int MP>
static int launch_cls(const void* jobs, int n_jobs, unsigned int* queue, unsigned long long* stats, int waves, Range range,
                      hipStream_t stream) {
  const int block = 256;
  const int grid = (waves + 3) / 4;
  if (grid <= 0) return 0;
  hipLaunchKernelGGL((sim_fp64_kernel<NP, MP>), dim3(grid), dim3(block), 0, stream,
                     (const JobDev<NP, MP>*)jobs, n_jobs, queue, stats, range);
  return (int)hipGetLastError();
}

template <int NP, int MP>
static int occupancy_cls() {
  int blocks = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, sim_fp64_kernel<NP, MP>, 256, 0) != hipSuccess) return 1;
  return blocks > 0 ? blocks : 1;
}

int launch_fp64(int cls, const void* jobs, int n_jobs, unsigned int* queue, unsigned long long* stats, int waves,
                 Range range, hipStream_t stream) {
  switch (cls) {
    case 0: return launch_cls<4, 4>(jobs, n_jobs, queue, stats, waves, range, stream);
    case 1: return launch_cls<8, 8>(jobs, n_jobs, queue, stats, waves, range, stream);
    case 2: return launch_cls<16, 16>(jobs, n_jobs, queue, stats, waves, range, stream);
    default: return -1;
  }
}

int blocks_per_cu_fp64(int cls) {
  switch (cls) {
    case 0: return occupancy_cls<4, 4>();
    case 1: return occupancy_cls<8, 8>();
    case 2: return occupancy_cls<16, 16>();
    default: return 1;
  }
}

}  // namespace pst
