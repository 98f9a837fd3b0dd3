// pst_rtc.hip -- hiprtc compilation, caching and launch of network-specific
// kernels (see pst_rtc.hpp).
#include "pst_rtc.hpp"

#include <hip/hiprtc.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "../../include/pst.h"
#include "pst_rtc_sources.inc"

namespace pst {
namespace {

std::mutex g_mu;
std::map<std::pair<int, std::string>, RtcKernel> g_cache;  // (device, source) -> kernel

const char* const kOptsMixed[] = {"--offload-arch=gfx950", "-O3", "-std=c++17",
                                  "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
const char* const kOptsFp64[] = {"--offload-arch=gfx950", "-O3", "-std=c++17",
                                 "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                                 "-ffp-contract=off"};

void arr(std::string& s, const char* name, const int8_t* v, int n) {
  s += "  static constexpr int8_t ";
  s += name;
  s += "[" + std::to_string(n > 0 ? n : 1) + "] = {";
  for (int i = 0; i < n; ++i) {
    if (i) s += ", ";
    s += std::to_string((int)v[i]);
  }
  if (n == 0) s += "0";
  s += "};\n";
}

uint64_t fnv1a(const std::string& s, uint64_t h = 1469598103934665603ull) {
  for (unsigned char c : s) {
    h ^= c;
    h *= 1099511628211ull;
  }
  return h;
}

std::string cache_dir() {
  const char* d = std::getenv("PST_KERNEL_CACHE");
  if (d && !*d) return "";  // set and empty: no disk cache
  std::string dir;
  if (d) {
    dir = d;
  } else if (const char* x = std::getenv("XDG_CACHE_HOME")) {
    dir = std::string(x) + "/py_stoch_amd";
  } else if (const char* h = std::getenv("HOME")) {
    dir = std::string(h) + "/.cache/py_stoch_amd";
  } else {
    return "";
  }
  // create the directory (and one parent level) if needed
  std::string parent = dir.substr(0, dir.find_last_of('/'));
  if (!parent.empty()) (void)mkdir(parent.c_str(), 0755);
  (void)mkdir(dir.c_str(), 0755);
  struct stat st;
  if (stat(dir.c_str(), &st) != 0 || !S_ISDIR(st.st_mode) || access(dir.c_str(), W_OK) != 0) return "";
  return dir;
}

bool read_file(const std::string& path, std::vector<char>* out) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  std::vector<char> buf;
  char tmp[65536];
  size_t n;
  while ((n = std::fread(tmp, 1, sizeof(tmp), f)) > 0) buf.insert(buf.end(), tmp, tmp + n);
  std::fclose(f);
  *out = std::move(buf);
  return true;
}

void write_file(const std::string& path, const char* data, size_t n) {
  const std::string tmp = path + ".tmp." + std::to_string((long)getpid());
  FILE* f = std::fopen(tmp.c_str(), "wb");
  if (!f) return;
  const bool ok = std::fwrite(data, 1, n, f) == n;
  std::fclose(f);
  if (ok) (void)std::rename(tmp.c_str(), path.c_str());
  else (void)std::remove(tmp.c_str());
}

int compile(const std::string& source, int precision, std::vector<char>* code, std::string* err) {
  hiprtcProgram prog;
  hiprtcResult r = hiprtcCreateProgram(&prog, source.c_str(), "pst_fixed.hip", pst_rtc_src::kCount,
                                       pst_rtc_src::kTexts, pst_rtc_src::kNames);
  if (r != HIPRTC_SUCCESS) {
    *err = std::string("hiprtcCreateProgram: ") + hiprtcGetErrorString(r);
    return PST_EHIP;
  }
  const bool fp64 = precision == PST_PRECISION_FP64;
  r = fp64 ? hiprtcCompileProgram(prog, (int)(sizeof(kOptsFp64) / sizeof(kOptsFp64[0])), kOptsFp64)
           : hiprtcCompileProgram(prog, (int)(sizeof(kOptsMixed) / sizeof(kOptsMixed[0])), kOptsMixed);
  if (r != HIPRTC_SUCCESS) {
    size_t n = 0;
    std::string log;
    if (hiprtcGetProgramLogSize(prog, &n) == HIPRTC_SUCCESS && n > 1) {
      log.resize(n);
      (void)hiprtcGetProgramLog(prog, &log[0]);
    }
    (void)hiprtcDestroyProgram(&prog);
    *err = std::string("hiprtc compile of the network kernel failed: ") + hiprtcGetErrorString(r) + "\n" +
           log.substr(0, 400);
    return PST_EHIP;
  }
  size_t n = 0;
  r = hiprtcGetCodeSize(prog, &n);
  if (r == HIPRTC_SUCCESS) {
    code->resize(n);
    r = hiprtcGetCode(prog, code->data());
  }
  (void)hiprtcDestroyProgram(&prog);
  if (r != HIPRTC_SUCCESS) {
    *err = std::string("hiprtcGetCode: ") + hiprtcGetErrorString(r);
    return PST_EHIP;
  }
  return PST_OK;
}

}  // namespace

std::string rtc_source(const HostNet& net, const JobParams& c, int precision,
                       const std::vector<std::string>& hazards, int waves_per_eu, bool paired) {
  const int N = net.N, M = net.M;
  std::vector<int8_t> v(N * M), ord(N * M), stp(N * M), E(N, 0);
  for (int i = 0; i < N * M; ++i) {
    v[i] = net.v[i];
    ord[i] = net.ord[i];
    stp[i] = net.stp[i];
  }
  for (int q = 0; q < net.kappa; ++q) E[q] = net.E[q];
  char eta[64];
  std::snprintf(eta, sizeof(eta), "%a", c.eta);  // exact bits of the double
  std::string s;
  s += precision == PST_PRECISION_FP64 ? "#include \"pst_engine_f64.hpp\"\n" : "#include \"pst_engine_f32.hpp\"\n";
  s += "namespace pst {\nstruct NetGen {\n";
  s += "  static constexpr int N = " + std::to_string(N) + ", M = " + std::to_string(M) +
       ", kappa = " + std::to_string(c.kappa) + ", n_c = " + std::to_string(c.n_c) +
       ", method = " + std::to_string(c.method) + ", ssa_burst = " + std::to_string(c.ssa_burst) + ";\n";
  s += std::string("  static constexpr double eta = ") + eta + ";\n";
  arr(s, "v", v.data(), N * M);
  arr(s, "ord", ord.data(), N * M);
  arr(s, "stp", stp.data(), N * M);
  arr(s, "H", net.H, N);
  arr(s, "Ht", net.Ht, N);
  arr(s, "irs", net.irs, N);
  arr(s, "E", E.data(), N);
  const bool custom = !hazards.empty();
  const int n_rates = custom ? c.n_rates : M;
  s += std::string("  static constexpr bool custom = ") + (custom ? "true" : "false") + ";\n";
  s += "  static constexpr int n_rates = " + std::to_string(n_rates > 0 ? n_rates : 1) + ";\n";
  s += "  template <class R>\n  __device__ __forceinline__ static void props(const int32_t (&xs)[N], const R (&cs)[n_rates],"
       " R (&a)[M]) {\n";
  if (custom) {
    s += "#define X(i) ((R)xs[i])\n#define C(k) (cs[k])\n";
    for (int j = 0; j < M; ++j) s += "    a[" + std::to_string(j) + "] = " + hazards[j] + ";\n";
    s += "#undef X\n#undef C\n";
  }
  s += "  }\n};\n}  // namespace pst\n";
  s += precision == PST_PRECISION_FP64 ? "#define PST_ENGINE EngineF64\n" : "#define PST_ENGINE EngineF32\n";
  s += std::string("#define PST_KERNEL_NAME ") + (precision == PST_PRECISION_FP64 ? "pst_net_fp64" : "pst_net_mixed") +
       (paired ? "_paired" : "") + "\n";
  s += paired ? "extern \"C\" __global__ void __launch_bounds__(512)" : "extern \"C\" __global__ void __launch_bounds__(256)";
  if (waves_per_eu > 0) s += " __attribute__((amdgpu_waves_per_eu(" + std::to_string(waves_per_eu) + ")))";
  s += "\n"
       "PST_KERNEL_NAME(const pst::JobParams* __restrict__ jobs, int n_jobs, unsigned int* __restrict__ queue,\n"
       "                 unsigned long long* __restrict__ stats, pst::Range range) {\n";
  s += std::string("  pst::") + (paired ? "run_driver_paired" : "run_driver") +
       "<pst::PST_ENGINE<pst::FixedRef<pst::NetGen>>, pst::FixedRef<pst::NetGen>>(jobs, n_jobs, queue, stats, "
       "range.f_end, range.resume);\n}\n";
  return s;
}

int rtc_kernel(const std::string& source, int precision, RtcKernel* out, std::string* err, bool paired) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    *err = std::string("hipGetDevice: ") + hipGetErrorString(e);
    return PST_EHIP;
  }
  std::lock_guard<std::mutex> lock(g_mu);
  const auto key = std::make_pair(dev, source);
  auto it = g_cache.find(key);
  if (it != g_cache.end()) {
    *out = it->second;
    return PST_OK;
  }
  // disk cache: code object keyed by the source, the options and the hiprtc version
  int vmaj = 0, vmin = 0;
  (void)hiprtcVersion(&vmaj, &vmin);
  std::string id = source + "|" + std::to_string(precision) + "|" + std::to_string(vmaj) + "." +
                   std::to_string(vmin);
  for (const char* o : (precision == PST_PRECISION_FP64 ? std::vector<const char*>(std::begin(kOptsFp64), std::end(kOptsFp64))
                                                         : std::vector<const char*>(std::begin(kOptsMixed), std::end(kOptsMixed))))
    id += std::string("|") + o;
  for (int i = 0; i < pst_rtc_src::kCount; ++i) id += pst_rtc_src::kTexts[i];
  const std::string dir = cache_dir();
  char hname[32];
  std::snprintf(hname, sizeof(hname), "%016llx", (unsigned long long)fnv1a(id));
  std::vector<char> code;
  bool have = false;
  if (!dir.empty()) {
    std::vector<char> key_text;
    if (read_file(dir + "/" + hname + ".key", &key_text) && std::string(key_text.begin(), key_text.end()) == id &&
        read_file(dir + "/" + hname + ".co", &code) && !code.empty())
      have = true;
  }
  if (!have) {
    int rc = compile(source, precision, &code, err);
    if (rc != PST_OK) return rc;
    if (!dir.empty()) {
      write_file(dir + "/" + hname + ".co", code.data(), code.size());
      write_file(dir + "/" + hname + ".key", id.data(), id.size());
    }
  }
  hipModule_t mod;
  if ((e = hipModuleLoadData(&mod, code.data())) != hipSuccess) {
    *err = std::string("hipModuleLoadData: ") + hipGetErrorString(e);
    return PST_EHIP;
  }
  RtcKernel k;
  const std::string name = std::string(precision == PST_PRECISION_FP64 ? "pst_net_fp64" : "pst_net_mixed") +
                           (paired ? "_paired" : "");
  if ((e = hipModuleGetFunction(&k.fn, mod, name.c_str())) != hipSuccess) {
    *err = std::string("hipModuleGetFunction: ") + hipGetErrorString(e);
    return PST_EHIP;
  }
  k.block = paired ? 512 : 256;
  int blocks = 0;
  if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k.fn, k.block, 0) == hipSuccess && blocks > 0)
    k.blocks_per_cu = blocks;
  g_cache.emplace(key, k);  // the module stays loaded for the life of the process
  *out = k;
  return PST_OK;
}

hipError_t rtc_launch(const RtcKernel& k, const JobParams* jobs, int n_jobs, unsigned int* queue,
                      unsigned long long* stats, int waves, Range range, hipStream_t stream) {
  const int grid = (waves + 3) / 4;  // four simulation wavefronts per workgroup
  if (grid <= 0) return hipSuccess;
  void* args[] = {(void*)&jobs, (void*)&n_jobs, (void*)&queue, (void*)&stats, (void*)&range};
  return hipModuleLaunchKernel(k.fn, grid, 1, 1, k.block, 1, 1, 0, stream, args, nullptr);
}

}  // namespace pst
