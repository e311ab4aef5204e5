// solhip_render_noalloc: the plain GPU renderer (megakernel) into a device image without any
// allocation, for callers that cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc). Every scratch buffer the megakernel needs - its
// work counter, ray and sample counters, the stack-overflow flag and the traversal stack's spill
// area - lives in one file-scope device array per GPU, `scratch`, whose device address is looked up
// once per device when a context is created (solhip_create calls noalloc_setup with the device
// current). Calls on one GPU share that array: calls that overlap in time on the same GPU race, so
// the caller serialises them (as every call on one context already is).
#include <algorithm>
#include <mutex>
#include <string>

#include "solhip.h"
#include "launch.h"
#include "state.h"

namespace solhip {

namespace {

// Layout of the scratch array: eight 64-bit words (work counter, extension rays, shadow rays,
// samples, overflow flag, unused), padded to 256 bytes, then the spill area of the traversal stack:
// kNoAllocSpill entries for each of up to kNoAllocLanes lanes (entry e of lane l at
// spill[e * lanes + l]). 196,608 lanes are three waves per SIMD on 256 compute units, the
// megakernel's occupancy; a larger device runs with fewer workgroups than it could hold.
constexpr int kNoAllocLanes = 196608;
constexpr int kNoAllocSpill = 64;
constexpr size_t kNoAllocHeader = 256;
constexpr size_t kNoAllocBytes = kNoAllocHeader + sizeof(int) * size_t(kNoAllocLanes) * kNoAllocSpill;  // 48 MiB + 256 B
constexpr int kMaxDevices = 64;

std::mutex g_mutex;
void* g_scratch[kMaxDevices] = {};   // device address of `scratch` on each device, once looked up

} // namespace

// The scratch array (one copy per GPU). Host code never passes its name to a kernel: the name is
// a host-side stand-in, the device address comes from hipGetSymbolAddress in noalloc_setup.
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

// Looks up the device address of `scratch` on `device`, which must be current. Called once per
// device by solhip_create; later calls return at once.
int noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return set_error(SOLHIP_E_NO_DEVICE, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_mutex);
    if (g_scratch[device]) return SOLHIP_OK;
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        (void)hipGetLastError();
        return set_error(SOLHIP_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    }
    g_scratch[device] = ptr;
    return SOLHIP_OK;
}

} // namespace solhip

using namespace solhip;

extern "C" {

int solhip_render_noalloc(solhip_scene* s, const sol_pt_config* cfg, uint32_t width, uint32_t height,
                          uint64_t sample_index, uint32_t sample_count, float* accum) {
    if (!s || !cfg || !accum) return set_error(SOLHIP_E_INVALID, "null argument");
    if (width == 0 || height == 0) return set_error(SOLHIP_E_INVALID, "empty image");
    if (cfg->min_survival_prob < 0 || cfg->max_survival_prob > 1 || cfg->min_survival_prob > cfg->max_survival_prob)
        return set_error(SOLHIP_E_INVALID, "survival probabilities must satisfy 0 <= min <= max <= 1");
    if (!s->num_lights) return set_error(SOLHIP_E_NO_LIGHTS, "scene without light cannot be rendered");
    if (sample_count == 0) return SOLHIP_OK;
    Context* ctx = s->ctx;
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_mutex);
        base = ctx->device < kMaxDevices ? g_scratch[ctx->device] : nullptr;
    }
    if (!base) return set_error(SOLHIP_E_HIP, "scratch array of this device not set up");
    LaunchShape shape = ctx->shape(ctx->tuning.block_size, ctx->tuning.lds_stack);
    shape.grid = std::min(shape.grid, kNoAllocLanes / shape.block);
    // the stack holds LDS + kNoAllocSpill entries: deeper trees are refused, not truncated
    const bool wide = shape.wide && s->view.nodes4 != nullptr;
    const uint32_t room = static_cast<uint32_t>(shape.lds_stack + kNoAllocSpill);
    if (wide ? 3 * (s->bvh4_depth > 0 ? s->bvh4_depth - 1 : 0) > room : s->bvh_depth > room + 1)
        return set_error(SOLHIP_E_TOO_DEEP, "BVH depth " + std::to_string(wide ? s->bvh4_depth : s->bvh_depth) +
                                                " exceeds the stack of solhip_render_noalloc (" +
                                                std::to_string(room) + " entries)");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return set_error(SOLHIP_E_HIP, hipGetErrorString(e));
    unsigned long long* words = static_cast<unsigned long long*>(base);
    TraversalScratch sc;
    sc.spill = reinterpret_cast<int*>(static_cast<unsigned char*>(base) + kNoAllocHeader);
    sc.spill_entries = kNoAllocSpill;
    sc.max_threads = shape.grid * shape.block;
    sc.overflow = reinterpret_cast<int*>(words + 4);
    // counters and the overflow flag start at zero (a kernel, not a memset: capturable in graphs)
    if ((e = launch_zero_u64(words, 8, ctx->stream)) != hipSuccess ||
        (e = launch_megakernel(s->view, *cfg, width, height, sample_index, sample_count, accum, words, words + 1, sc,
                               shape, ctx->stream)) != hipSuccess)
        return set_error(SOLHIP_E_HIP, std::string("megakernel launch: ") + hipGetErrorString(e));
    unsigned long long host[5];
    if ((e = hipMemcpyAsync(host, words, sizeof(host), hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess ||
        (e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
        return set_error(SOLHIP_E_HIP, std::string("solhip_render_noalloc: ") + hipGetErrorString(e));
    ctx->stats.extension_rays += host[1];
    ctx->stats.shadow_rays += host[2];
    ctx->stats.samples += host[3];
    ctx->stats.launches += 2;
    if (static_cast<int>(host[4] & 0xFFFFFFFFu) != 0) return set_error(SOLHIP_E_TOO_DEEP, "traversal stack overflow");
    return SOLHIP_OK;
}

} // extern "C"
