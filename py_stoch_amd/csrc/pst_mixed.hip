// pst_mixed.hip -- table-driven kernels of the size classes compiled into the library; the engine is in pst_engine_f32.hpp.
#include "pst_engine_f32.hpp"

namespace pst {

template <int NP, int MP>
__global__ void __launch_bounds__(256) sim_mixed_kernel(const JobDev<NP, MP>* __restrict__ jobs, int n_jobs,
                                                        unsigned int* __restrict__ queue,
                                                        unsigned long long* __restrict__ stats, Range range) {
  run_driver<EngineF32<JobRef<NP, MP>>, JobRef<NP, MP>>(jobs, n_jobs, queue, stats, range.f_end, range.resume);
}

template <int NP, int MP>
static int launch_cls_m(const void* jobs, int n_jobs, unsigned int* queue, unsigned long long* stats, int waves, Range range,
                        hipStream_t stream) {
  const int block = 256;
  const int grid = (waves + 3) / 4;
  if (grid <= 0) return 0;
  hipLaunchKernelGGL((sim_mixed_kernel<NP, MP>), dim3(grid), dim3(block), 0, stream,
                     (const JobDev<NP, MP>*)jobs, n_jobs, queue, stats, range);
  return (int)hipGetLastError();
}

template <int NP, int MP>
static int occupancy_cls_m() {
  int blocks = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, sim_mixed_kernel<NP, MP>, 256, 0) != hipSuccess) return 1;
  return blocks > 0 ? blocks : 1;
}

int launch_mixed(int cls, const void* jobs, int n_jobs, unsigned int* queue, unsigned long long* stats, int waves,
                 Range range, hipStream_t stream) {
  switch (cls) {
    case 0: return launch_cls_m<4, 4>(jobs, n_jobs, queue, stats, waves, range, stream);
    case 1: return launch_cls_m<8, 8>(jobs, n_jobs, queue, stats, waves, range, stream);
    case 2: return launch_cls_m<16, 16>(jobs, n_jobs, queue, stats, waves, range, stream);
    default: return -1;
  }
}

// The mixed-precision Poisson sampler on given means and Philox words, one
// draw per thread, with the code of the kernels' leap: search inversion below
// POIS_ASYM_MIN, the asymptotic inversion (and its exact FP64 decisions) up to
// POIS_ASYM_MAX, PTRS above (on the stream of simulation i, block 0).  path:
// 0 mean <= 0, 1 search, 2 asymptotic, 3 asymptotic decided in FP64, 4 PTRS.
__global__ void __launch_bounds__(256) poisson_probe_kernel(const float* __restrict__ lam,
                                                            const uint32_t* __restrict__ words, int64_t n,
                                                            int32_t* __restrict__ k_out, int32_t* __restrict__ path) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float l = lam[i];
  const float u = u24(words[i]);
  int32_t k = 0, p = 0;
  if (l > 0.0f && l < POIS_ASYM_MIN) {
    float pk = fast_exp(-l), Fc = pk;
    while (u > Fc && k < INV_MAX_K) {
      k += 1;
      pk = pk * l * fast_rcp((float)k);
      const float Fn = Fc + pk;
      if (!(Fn > Fc)) break;
      Fc = Fn;
    }
    p = 1;
  } else if (l >= POIS_ASYM_MIN && l < POIS_ASYM_MAX) {
    bool ex = false;
    int32_t m = -1;
    k = poisson_asym_f32(l, u, ex, m);
    p = 2;
    if (ex) {
      k = poisson_exact_f64((double)l, (double)u, m);
      p = 3;
    }
  } else if (l >= POIS_ASYM_MAX) {
    Key key;
    key.init(0, (uint64_t)i);
    const Draw d = ptrs_wide(key.k0, key.k1, key.s_lo, key.s_hi, 0, 0, l);
    k = (int32_t)d.k;
    p = 4;
  }
  k_out[i] = k;
  path[i] = p;
}

int launch_poisson_probe(const float* lam, const uint32_t* words, int64_t n, int32_t* k, int32_t* path,
                         hipStream_t stream) {
  if (n <= 0) return 0;
  const int64_t grid = (n + 255) / 256;
  hipLaunchKernelGGL(poisson_probe_kernel, dim3((unsigned)grid), dim3(256), 0, stream, lam, words, n, k, path);
  return (int)hipGetLastError();
}

int blocks_per_cu_mixed(int cls) {
  switch (cls) {
    case 0: return occupancy_cls_m<4, 4>();
    case 1: return occupancy_cls_m<8, 8>();
    case 2: return occupancy_cls_m<16, 16>();
    default: return 1;
  }
}

}  // namespace pst
