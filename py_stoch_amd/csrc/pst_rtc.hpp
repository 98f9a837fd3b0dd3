// pst_rtc.hpp -- network-specific kernels compiled at run time (host side).
//
// The reference turns each model's kinetic laws into C and compiles its two
// kernels with nvcc for every run (mod/parser/parser.py:32-55 spliced in at
// P1_P2.py:586, compiled at cuTauLeaping.py:34-35).  Here the same engine
// source that the library's table-driven kernels use is compiled by hiprtc
// with the network (stoichiometry, propensity orders, HOR tables, recorded
// species) and the network-level settings (kappa, n_c, eta, method, burst)
// as compile-time constants.  Compiled modules are cached for the life of
// the process (per device and source text) and, if a cache directory is
// usable, on disk.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "pst_args.hpp"

namespace pst {

// largest network a specialised kernel is generated for (state and
// propensities live in registers)
constexpr int RTC_MAX_SPECIES = 16;
constexpr int RTC_MAX_REACTIONS = 32;

struct RtcKernel {
  hipFunction_t fn = nullptr;
  int block = 256;        // lanes per workgroup: 256, or 512 for the paired driver
  int blocks_per_cu = 1;  // resident workgroups per CU (each with four simulation wavefronts)
};

// Source text of the specialised kernel for a network and the constant
// settings in `c` (N, M, kappa, n_c, eta, method, ssa_burst, n_rates are
// read).  `hazards` empty: mass-action propensities from the table; else one
// C expression per reaction (X(i), C(k), R(v), see pst_network.hazards).
// waves_per_eu > 0 asks the compiler for register use that fits that many
// wavefronts per SIMD (the kernel's arithmetic, and so its results, do not
// change).  paired: the kernel runs run_driver_paired (workgroups of 512
// lanes: four decide and four burst wavefronts), named pst_net_*_paired.
std::string rtc_source(const HostNet& net, const JobParams& c, int precision,
                       const std::vector<std::string>& hazards, int waves_per_eu = 0, bool paired = false);

// Compile (or fetch from the cache) the kernel for `source` on the current
// device (`paired` as given to rtc_source).  Returns 0, or a PST_* status
// with a message in err.
int rtc_kernel(const std::string& source, int precision, RtcKernel* out, std::string* err, bool paired = false);

// Launch a specialised kernel: `jobs` is a device array of n_jobs JobParams;
// `waves` simulation wavefronts, four per workgroup.
hipError_t rtc_launch(const RtcKernel& k, const JobParams* jobs, int n_jobs, unsigned int* queue,
                      unsigned long long* stats, int waves, Range range, hipStream_t stream);

}  // namespace pst
