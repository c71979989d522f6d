// Single-source BFS that makes no device allocation (ll_gpu_bfs_noalloc).
//
// Callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc) get the same search as
// ll_gpu_bfs_ex with its working memory in `scratch`, a fixed array of this
// library on every GPU it runs on (64 MiB), held until the process ends.  In host code
// the array's name is not a device address: ll_gpu_create asks HIP for the
// array's address on its device (hipGetSymbolAddress, once per device, with
// that device current) and every use goes through that address.  All calls on
// one GPU share the array, so calls that overlap on one GPU (two contexts or
// two threads) race; every region a search reads is cleared or written by the
// search itself first, so a call never sees the previous call's contents.
#include <mutex>

#include "ll_gpu_internal.h"

namespace {

static __device__ __align__(256) unsigned char scratch[llgpu::kScratchBytes];

std::mutex g_scratch_mu;
std::map<int, void*> g_scratch_addr;  // device -> address of `scratch` there

}  // namespace

namespace llgpu {

void* scratch_setup(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_addr.find(device);
  if (it != g_scratch_addr.end()) return it->second;
  int cur = -1;
  LLG_HIP(hipGetDevice(&cur));
  LLG_CHECK(cur == device, LL_GPU_ESTATE, "scratch set-up needs its device current");
  void* p = nullptr;
  LLG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  LLG_CHECK(p != nullptr, LL_GPU_EHIP, "no device address for the scratch array");
  g_scratch_addr[device] = p;
  return p;
}

}  // namespace llgpu

extern "C" {

int ll_gpu_bfs_noalloc(ll_gpu_graph* g, int level, int64_t root, int32_t* levels_dev,
                       uint64_t* reached, uint64_t* edges_examined) {
  if (!g) return LL_GPU_EINVAL;
  try {
    ll_gpu_ctx* ctx = g->ctx;
    LLG_HIP(hipSetDevice(ctx->device));
    LLG_CHECK(levels_dev != nullptr, LL_GPU_EINVAL, "levels_dev (device int32[max_nodes]) is required");
    LLG_CHECK(root >= 0 && uint64_t(root) < g->V, LL_GPU_EINVAL, "root outside [0, max_nodes)");
    LLG_CHECK(ctx->scratch != nullptr, LL_GPU_ESTATE, "context has no scratch address");
    const size_t need = llgpu::bfs_work_bytes(g->V);
    LLG_CHECK(need <= llgpu::kScratchBytes, LL_GPU_ERANGE,
              "ll_gpu_bfs_noalloc: " + std::to_string(g->V) + " vertices need " +
                  std::to_string(need) + " bytes of working memory; the fixed array holds " +
                  std::to_string(llgpu::kScratchBytes));
    llgpu::BfsPrepared p;
    LLG_CHECK(llgpu::bfs_peek(g, level, &p), LL_GPU_ESTATE,
              "ll_gpu_bfs_noalloc: the snapshot's BFS structures are not built; call ll_gpu_bfs "
              "on this level once first");
    const llgpu::BfsResult r =
        llgpu::bfs_single_ws(g, p, uint32_t(root), ctx->scratch, levels_dev, false);
    if (reached) *reached = r.reached;
    if (edges_examined) *edges_examined = r.examined;
    return LL_GPU_OK;
  } catch (const llgpu::Error& e) {
    g->ctx->err = e.msg;
    return e.code;
  } catch (...) {
    return llgpu::api_exception(g->ctx);
  }
}

}  // extern "C"
