// kt_noalloc.hip — kt_price_noalloc: kt_price for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).  Every device buffer the call needs (the
// program tables, the block partials, the error word and the statistics)
// lives in one file-scope __device__ array, `scratch`, of kNoAllocBytes per
// GPU, held by each GPU the library runs on until the process ends.
//
// The array's name is a host-side stand-in in host code: its device address
// is taken once per device with hipGetSymbolAddress (at kt_engine_create,
// with the device current) and only that address is handed to kernels and
// copies.  Every call on one GPU shares that GPU's array and leaves the last
// call's contents in it; the error word is cleared by a kernel at the start
// of each call.  Calls that overlap on one GPU (two engines, two host
// threads) race on the array: callers serialise them.
#include <mutex>

#include "kt_engine_internal.h"

namespace kt {

// Layout of the array: [header | tables | partials].
constexpr size_t kNoAllocBytes = (size_t)8 << 20;          // 8 MiB per GPU
constexpr size_t kNoAllocHeader = 4096;                    // error word, statistics, error copy
constexpr size_t kNoAllocPartials = (size_t)6 << 20;       // 262,144 KStats: up to 2^26 paths
constexpr size_t kNoAllocTables = kNoAllocBytes - kNoAllocHeader - kNoAllocPartials;
constexpr uint64_t kNoAllocMaxPaths = (uint64_t)(kNoAllocPartials / sizeof(KStats)) * kBlock;
constexpr int kNoAllocMaxDevices = 64;

}  // namespace kt

static __device__ __align__(256) unsigned char scratch[kt::kNoAllocBytes];

namespace kt {

namespace {

std::mutex g_scratch_mu;
void* g_scratch_ptr[kNoAllocMaxDevices] = {};

__global__ void noalloc_clear_kernel(unsigned* w) {
  if (threadIdx.x == 0) *w = 0u;
}

void* scratch_for(int device) {
  if (device < 0 || device >= kNoAllocMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_ptr[device];
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

hipError_t noalloc_scratch_init(int device) {
  if (device < 0 || device >= kNoAllocMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return hipSuccess;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return e;
  }
  g_scratch_ptr[device] = p;
  return hipSuccess;
}

}  // namespace kt

using namespace kt;

extern "C" kt_status kt_price_noalloc(kt_engine* E, const char* json, const kt_model* model,
                                      const kt_mc_config* cfg, kt_price_result* out) {
  KT_GUARD_BEGIN
  if (!E || !out || !cfg) raise(KT_ERR_INVALID_ARGUMENT, "null engine, config or output");
  if (cfg->antithetic || cfg->control_variate)
    raise(KT_ERR_UNSUPPORTED, "kt_price_noalloc: variance reduction runs extra launches; use kt_price");
  check_hip(hipSetDevice(E->device), "hipSetDevice");
  char* base = (char*)scratch_for(E->device);
  if (!base) raise(KT_ERR_DEVICE, "kt_price_noalloc: the scratch array has no device address on this GPU");
  ModelDesc md = model_from_abi(model);
  PreparedContract pc = prepare(json, md, cfg, false);
  *out = kt_price_result{};
  if (cfg->n_paths == 0) return KT_OK;
  if (cfg->n_paths > kNoAllocMaxPaths)
    raise(KT_ERR_INVALID_ARGUMENT, "kt_price_noalloc: at most " + std::to_string(kNoAllocMaxPaths) +
                                       " paths fit its scratch array (use kt_price)");
  Program prog = lower({pc.con.get()}, pc.grid, md.names);
  LaunchTables T = launch_tables(prog, {md}, {cfg->rate}, pc.grid);
  const int kind = model_kernel_kind(md);
  const int K = (int)md.names.size();
  int ns = 0, nout = 0;
  if (!path_variant_exists(kind, K, prog.n_slots, prog.n_outputs, false, 1, &ns, &nout))
    raise(KT_ERR_UNSUPPORTED, "no path kernel variant for this program shape");
  if (prog.n_slots > ns) raise(KT_ERR_UNSUPPORTED, "program needs more registers than the kernel variant holds");
  const int nstat = nout, n_used = prog.n_outputs;
  if (n_used < 1 || n_used > 8) raise(KT_ERR_UNSUPPORTED, "kt_price_noalloc: one to eight outputs");
  const uint64_t n_blocks = (cfg->n_paths + kBlock - 1) / kBlock;
  if (n_blocks * (uint64_t)nstat * sizeof(KStats) > kNoAllocPartials)
    raise(KT_ERR_INVALID_ARGUMENT, "kt_price_noalloc: the program's block partials exceed its scratch array");

  // Tables packed on the host, then one copy into the array's table region.
  std::vector<KOp> ops;
  const auto blobs = program_blobs(prog, T, &ops);
  std::vector<size_t> offs;
  size_t total = 0;
  for (const auto& b : blobs) {
    offs.push_back(total);
    total += align256(b.second);
  }
  if (total > kNoAllocTables)
    raise(KT_ERR_INVALID_ARGUMENT, "kt_price_noalloc: the program's tables (" + std::to_string(total) +
                                       " bytes) exceed its scratch array (" + std::to_string(kNoAllocTables) +
                                       " bytes; use kt_price)");
  std::vector<char> host(std::max<size_t>(total, 256), 0);
  for (size_t i = 0; i < blobs.size(); ++i)
    if (blobs[i].second) std::memcpy(host.data() + offs[i], blobs[i].first, blobs[i].second);

  unsigned* err = (unsigned*)base;
  KStats* stats = (KStats*)(base + 256);
  unsigned* err_copy = (unsigned*)(stats + n_used);
  char* tables = base + kNoAllocHeader;
  KStats* partials = (KStats*)(base + kNoAllocHeader + kNoAllocTables);
  std::vector<void*> dev;
  for (size_t o : offs) dev.push_back(tables + o);

  check_hip(hipMemcpyAsync(tables, host.data(), total, hipMemcpyHostToDevice, E->stream), "upload tables");
  hipLaunchKernelGGL(noalloc_clear_kernel, dim3(1), dim3(64), 0, E->stream, err);
  check_hip(hipGetLastError(), "clear launch");
  PathArgs A = program_path_args(prog, {md}, T, pc.grid, cfg->seed, dev);
  A.err = err;
  A.path_base = 0;
  A.n_paths = cfg->n_paths;
  A.partials = partials;
  E->prof_begin();
  check_hip(launch_path(kind, K, ns, nout, false, 1, A, (int)n_blocks, E->stream), "path kernel launch");
  E->prof_end((double)cfg->n_paths * (double)(pc.grid.size() > 0 ? pc.grid.size() - 1 : 0));
  check_hip(launch_finalize(partials, (int)n_blocks, nstat, n_used, stats, E->stream, err, err_copy),
            "finalize launch");
  struct {
    KStats st[8 + 1];
  } rb;
  check_hip(hipMemcpyAsync(&rb, stats, (size_t)(n_used + 1) * sizeof(KStats), hipMemcpyDeviceToHost, E->stream),
            "copy stats");
  check_hip(hipStreamSynchronize(E->stream), "stream sync");
  unsigned err_bits = 0;
  std::memcpy(&err_bits, &rb.st[n_used], sizeof(unsigned));
  if (err_bits) {
    const int code = lowest_error_code(err_bits);
    raise(code, error_text(code, prog));
  }
  *out = result_from_stats(rb.st[0]);
  E->err.clear();
  return KT_OK;
  KT_GUARD_END(E)
}
