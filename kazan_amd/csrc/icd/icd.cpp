// icd.cpp — Vulkan entry points of the Kazan draw path on MI355X.
//
// The exported ICD interface and the name -> entry point table follow the
// reference (vulkan-driver/src/lib.rs:37-130, api_impl.rs:432-860). Objects
// live in host memory; buffers and images live in HBM or pinned host memory;
// command buffers record a list of engine Ops that vkQueueSubmit replays on
// the device's HIP stream.
#include <hip/hip_runtime_api.h>

#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <thread>
#include <unordered_map>

#include "kazan/kazan_icd.h"
#include "kernels.h"
#include "objects.h"

namespace kz {
bool create_graphics_pipeline(Device *dev, const VkGraphicsPipelineCreateInfo *ci, PipelineCache *cache,
                              Pipeline **out, std::string &err);
bool create_compute_pipeline(Device *dev, const VkComputePipelineCreateInfo *ci, Pipeline **out, std::string &err);
void destroy_pipeline(Pipeline *p);
}  // namespace kz

using namespace kz;

namespace {

static_assert(sizeof(VkPhysicalDeviceLimits) == 504, "VkPhysicalDeviceLimits layout");
static_assert(sizeof(VkPhysicalDeviceProperties) == 824, "VkPhysicalDeviceProperties layout");
static_assert(sizeof(VkPhysicalDeviceFeatures) == 220, "VkPhysicalDeviceFeatures layout");
static_assert(sizeof(VkGraphicsPipelineCreateInfo) == 144, "VkGraphicsPipelineCreateInfo layout");
static_assert(sizeof(VkComputePipelineCreateInfo) == 96, "VkComputePipelineCreateInfo layout");
static_assert(sizeof(VkPhysicalDeviceMemoryProperties) == 520, "VkPhysicalDeviceMemoryProperties layout");

constexpr VkDeviceSize kBufferAlignment = 256;
constexpr VkDeviceSize kImageAlignment = 256;
constexpr VkDeviceSize kMapAlignment = 128;  // reference MIN_MEMORY_MAP_ALIGNMENT, lib.rs:29

template <class T, class H>
T *from(H h) {
    return reinterpret_cast<T *>(h);
}
template <class H, class T>
H to(T *p) {
    return reinterpret_cast<H>(p);
}

uint32_t encode_unorm8(float v) {
    v = std::fmin(std::fmax(v, 0.0f), 1.0f);
    return (uint32_t)(v * 255.0f + 0.5f);
}

bool set_device(Device *d) { return hipSetDevice(d->physical->hip_device) == hipSuccess; }

// --- physical device description --------------------------------------

void fill_limits(VkPhysicalDeviceLimits &l) {
    // The values Kazan advertises (vulkan-driver/src/api_impl.rs:1501-1612):
    // they fix the raster precision contract of this driver too.
    std::memset(&l, 0, sizeof l);
    const uint32_t big = ~0u;
    l.maxImageDimension1D = l.maxImageDimension2D = l.maxImageDimension3D = l.maxImageDimensionCube = 16384;
    l.maxImageArrayLayers = 1;
    l.maxTexelBufferElements = big;
    l.maxUniformBufferRange = big;
    l.maxStorageBufferRange = big;
    l.maxPushConstantsSize = KZ_PUSH_WORDS * 4;
    l.maxMemoryAllocationCount = big;
    l.maxSamplerAllocationCount = big;
    l.bufferImageGranularity = 1;
    l.maxBoundDescriptorSets = 8;
    l.maxPerStageDescriptorSamplers = KZ_MAX_TEXTURES;
    l.maxPerStageDescriptorUniformBuffers = KZ_MAX_UBOS;
    l.maxPerStageDescriptorSampledImages = KZ_MAX_TEXTURES;
    l.maxPerStageResources = KZ_MAX_TEXTURES + KZ_MAX_UBOS;
    l.maxDescriptorSetSamplers = KZ_MAX_TEXTURES;
    l.maxDescriptorSetUniformBuffers = KZ_MAX_UBOS;
    l.maxDescriptorSetSampledImages = KZ_MAX_TEXTURES;
    l.maxVertexInputAttributes = 16;
    l.maxVertexInputBindings = KZ_MAX_BINDINGS;
    l.maxVertexInputAttributeOffset = 2047;
    l.maxVertexInputBindingStride = 2048;
    l.maxVertexOutputComponents = 64;
    l.maxFragmentInputComponents = KZ_MAX_VARYINGS;
    l.maxFragmentOutputAttachments = 1;
    l.maxFragmentCombinedOutputResources = 1;
    l.subPixelPrecisionBits = 4;     // api_impl.rs:1559
    l.subTexelPrecisionBits = 4;     // api_impl.rs:1560
    l.mipmapPrecisionBits = 4;       // api_impl.rs:1561
    l.maxDrawIndexedIndexValue = big;
    l.maxDrawIndirectCount = 1;
    l.maxSamplerLodBias = 2.0f;
    l.maxSamplerAnisotropy = 1.0f;
    l.maxViewports = 1;
    l.maxViewportDimensions[0] = l.maxViewportDimensions[1] = 4096;  // api_impl.rs:1567
    l.viewportBoundsRange[0] = -8192.0f;                              // api_impl.rs:1568
    l.viewportBoundsRange[1] = 8191.0f;
    l.viewportSubPixelBits = 0;
    l.minMemoryMapAlignment = kMapAlignment;
    l.minTexelBufferOffsetAlignment = 64;
    l.minUniformBufferOffsetAlignment = 64;
    l.minStorageBufferOffsetAlignment = 64;
    l.minTexelOffset = -8;
    l.maxTexelOffset = 7;
    l.maxFramebufferWidth = 4096;  // api_impl.rs:1581-1582
    l.maxFramebufferHeight = 4096;
    l.maxFramebufferLayers = 1;
    const uint32_t s14 = VK_SAMPLE_COUNT_1_BIT | VK_SAMPLE_COUNT_4_BIT;  // api_impl.rs:1584
    l.framebufferColorSampleCounts = l.framebufferDepthSampleCounts = s14;
    l.framebufferStencilSampleCounts = l.framebufferNoAttachmentsSampleCounts = s14;
    l.maxColorAttachments = 1;
    l.sampledImageColorSampleCounts = VK_SAMPLE_COUNT_1_BIT;
    l.sampledImageIntegerSampleCounts = VK_SAMPLE_COUNT_1_BIT;
    l.sampledImageDepthSampleCounts = VK_SAMPLE_COUNT_1_BIT;
    l.sampledImageStencilSampleCounts = VK_SAMPLE_COUNT_1_BIT;
    l.storageImageSampleCounts = VK_SAMPLE_COUNT_1_BIT;
    l.maxSampleMaskWords = 1;
    l.timestampComputeAndGraphics = VK_FALSE;
    l.discreteQueuePriorities = 2;
    l.pointSizeRange[0] = l.pointSizeRange[1] = 1.0f;
    l.lineWidthRange[0] = l.lineWidthRange[1] = 1.0f;
    l.standardSampleLocations = VK_TRUE;  // api_impl.rs:1609
    l.optimalBufferCopyOffsetAlignment = 16;
    l.optimalBufferCopyRowPitchAlignment = 16;
    l.nonCoherentAtomSize = 1;
}

void init_physical_device(PhysicalDevice &pd, int hip_device) {
    pd.hip_device = hip_device;
    hipDeviceProp_t hp{};
    hipGetDeviceProperties(&hp, hip_device);
    VkPhysicalDeviceProperties &p = pd.props;
    std::memset(&p, 0, sizeof p);
    p.apiVersion = VK_API_VERSION_1_1;
    p.driverVersion = VK_MAKE_VERSION(0, 1, 0);
    p.vendorID = 0x1002;
    p.deviceID = (uint32_t)hp.pciDeviceID;
    p.deviceType = VK_PHYSICAL_DEVICE_TYPE_DISCRETE_GPU;
    std::snprintf(p.deviceName, sizeof p.deviceName, "Kazan on %s (%s)", hp.name, hp.gcnArchName);
    // pipelineCacheUUID: bump with any change to the code-object format
    const char uuid[16] = {'k', 'a', 'z', 'a', 'n', '-', 'g', 'f', 'x', '9', '5', '0', '-', 'v', '0', '1'};
    std::memcpy(p.pipelineCacheUUID, uuid, 16);
    fill_limits(p.limits);
    VkPhysicalDeviceMemoryProperties &m = pd.memory;
    std::memset(&m, 0, sizeof m);
    m.memoryTypeCount = 2;
    m.memoryTypes[0].propertyFlags = VK_MEMORY_PROPERTY_DEVICE_LOCAL_BIT;
    m.memoryTypes[0].heapIndex = 0;
    m.memoryTypes[1].propertyFlags =
        VK_MEMORY_PROPERTY_HOST_VISIBLE_BIT | VK_MEMORY_PROPERTY_HOST_COHERENT_BIT | VK_MEMORY_PROPERTY_HOST_CACHED_BIT;
    m.memoryTypes[1].heapIndex = 1;
    m.memoryHeapCount = 2;
    m.memoryHeaps[0].size = hp.totalGlobalMem;
    m.memoryHeaps[0].flags = VK_MEMORY_HEAP_DEVICE_LOCAL_BIT;
    m.memoryHeaps[1].size = 64ull << 30;
    std::memset(&pd.features, 0, sizeof pd.features);
    pd.features.fullDrawIndexUint32 = VK_TRUE;
}

// --- entry point table --------------------------------------------------

struct Entry {
    PFN_vkVoidFunction fn;
    bool physical;  // physical-device level
    bool device;    // device level (vkGetDeviceProcAddr)
};

std::unordered_map<std::string, Entry> &entries();

PFN_vkVoidFunction lookup(const char *name, bool device_only) {
    if (!name) return nullptr;
    auto &t = entries();
    auto it = t.find(name);
    if (it == t.end()) return nullptr;
    if (device_only && !it->second.device) return nullptr;
    return it->second.fn;
}

}  // namespace

// ===========================================================================
// instance / physical device

extern "C" {

VKAPI_ATTR VkResult VKAPI_CALL vkEnumerateInstanceExtensionProperties(const char *pLayerName, uint32_t *pCount,
                                                                      VkExtensionProperties *pProps) {
    if (pLayerName) return VK_ERROR_LAYER_NOT_PRESENT;
    (void)pProps;
    *pCount = 0;
    return VK_SUCCESS;
}

VKAPI_ATTR VkResult VKAPI_CALL vkEnumerateInstanceLayerProperties(uint32_t *pCount, VkLayerProperties *) {
    *pCount = 0;
    return VK_SUCCESS;
}

VKAPI_ATTR VkResult VKAPI_CALL vkEnumerateInstanceVersion(uint32_t *pApiVersion) {
    *pApiVersion = VK_API_VERSION_1_1;
    return VK_SUCCESS;
}

VKAPI_ATTR VkResult VKAPI_CALL vkCreateInstance(const VkInstanceCreateInfo *ci, const VkAllocationCallbacks *,
                                                VkInstance *pInstance) {
    if (ci->enabledLayerCount) return VK_ERROR_LAYER_NOT_PRESENT;          // api_impl.rs:2950
    if (ci->enabledExtensionCount) return VK_ERROR_EXTENSION_NOT_PRESENT;  // no instance extensions
    if (ci->pApplicationInfo && ci->pApplicationInfo->apiVersion &&
        ci->pApplicationInfo->apiVersion > VK_MAKE_VERSION(1, 1, 0xFFF) &&
        (ci->pApplicationInfo->apiVersion >> 22) > 1)
        return VK_ERROR_INCOMPATIBLE_DRIVER;
    auto *inst = new (std::nothrow) Instance();
    if (!inst) return VK_ERROR_OUT_OF_HOST_MEMORY;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    for (int i = 0; i < n; ++i) {
        auto pd = std::make_unique<PhysicalDevice>();
        pd->instance = inst;
        init_physical_device(*pd, i);
        inst->physical_devices.push_back(std::move(pd));
    }
    *pInstance = to<VkInstance>(inst);
    return VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkDestroyInstance(VkInstance instance, const VkAllocationCallbacks *) {
    delete from<Instance>(instance);
}

VKAPI_ATTR VkResult VKAPI_CALL vkEnumeratePhysicalDevices(VkInstance instance, uint32_t *pCount,
                                                          VkPhysicalDevice *pDevices) {
    Instance *inst = from<Instance>(instance);
    const uint32_t n = (uint32_t)inst->physical_devices.size();
    if (!pDevices) {
        *pCount = n;
        return VK_SUCCESS;
    }
    const uint32_t k = std::min(*pCount, n);
    for (uint32_t i = 0; i < k; ++i) pDevices[i] = to<VkPhysicalDevice>(inst->physical_devices[i].get());
    *pCount = k;
    return k < n ? VK_INCOMPLETE : VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkGetPhysicalDeviceProperties(VkPhysicalDevice pd, VkPhysicalDeviceProperties *p) {
    *p = from<PhysicalDevice>(pd)->props;
}

VKAPI_ATTR void VKAPI_CALL vkGetPhysicalDeviceFeatures(VkPhysicalDevice pd, VkPhysicalDeviceFeatures *f) {
    *f = from<PhysicalDevice>(pd)->features;
}

VKAPI_ATTR void VKAPI_CALL vkGetPhysicalDeviceMemoryProperties(VkPhysicalDevice pd,
                                                               VkPhysicalDeviceMemoryProperties *m) {
    *m = from<PhysicalDevice>(pd)->memory;
}

VKAPI_ATTR void VKAPI_CALL vkGetPhysicalDeviceQueueFamilyProperties(VkPhysicalDevice, uint32_t *pCount,
                                                                    VkQueueFamilyProperties *p) {
    // one family with one queue (reference QUEUE_COUNTS = [1], lib.rs:30-32)
    if (!p) {
        *pCount = 1;
        return;
    }
    if (*pCount < 1) return;
    std::memset(p, 0, sizeof *p);
    p->queueFlags = VK_QUEUE_GRAPHICS_BIT | VK_QUEUE_COMPUTE_BIT | VK_QUEUE_TRANSFER_BIT;
    p->queueCount = 1;
    p->timestampValidBits = 0;
    p->minImageTransferGranularity = {1, 1, 1};
    *pCount = 1;
}

VKAPI_ATTR void VKAPI_CALL vkGetPhysicalDeviceFormatProperties(VkPhysicalDevice, VkFormat format,
                                                               VkFormatProperties *p) {
    std::memset(p, 0, sizeof *p);
    const uint32_t SAMPLED = 0x1, VERTEX = 0x40, COLOR = 0x80, BLEND = 0x100, DEPTH = 0x200, FILTER = 0x1000,
                   XFER = 0x4000 | 0x8000;
    switch (format) {
    case VK_FORMAT_R8G8B8A8_UNORM:
        p->linearTilingFeatures = p->optimalTilingFeatures = SAMPLED | COLOR | BLEND | FILTER | XFER;
        p->bufferFeatures = VERTEX;
        break;
    case VK_FORMAT_D32_SFLOAT: p->linearTilingFeatures = p->optimalTilingFeatures = DEPTH | XFER; break;
    case VK_FORMAT_R32_SFLOAT:
    case VK_FORMAT_R32G32_SFLOAT:
    case VK_FORMAT_R32G32B32_SFLOAT:
    case VK_FORMAT_R32G32B32A32_SFLOAT:
    case VK_FORMAT_R32_UINT: p->bufferFeatures = VERTEX; break;
    default: break;
    }
}

VKAPI_ATTR VkResult VKAPI_CALL vkEnumerateDeviceExtensionProperties(VkPhysicalDevice, const char *pLayerName,
                                                                    uint32_t *pCount, VkExtensionProperties *) {
    if (pLayerName) return VK_ERROR_LAYER_NOT_PRESENT;
    *pCount = 0;
    return VK_SUCCESS;
}

VKAPI_ATTR VkResult VKAPI_CALL vkEnumerateDeviceLayerProperties(VkPhysicalDevice, uint32_t *pCount,
                                                                VkLayerProperties *) {
    *pCount = 0;
    return VK_SUCCESS;
}

// ===========================================================================
// device / queue

VKAPI_ATTR VkResult VKAPI_CALL vkCreateDevice(VkPhysicalDevice physicalDevice, const VkDeviceCreateInfo *ci,
                                              const VkAllocationCallbacks *, VkDevice *pDevice) {
    PhysicalDevice *pd = from<PhysicalDevice>(physicalDevice);
    if (ci->enabledLayerCount) return VK_ERROR_LAYER_NOT_PRESENT;          // api_impl.rs:1340
    if (ci->enabledExtensionCount) return VK_ERROR_EXTENSION_NOT_PRESENT;  // api_impl.rs:1349-1351
    if (ci->pEnabledFeatures) {
        const VkBool32 *want = (const VkBool32 *)ci->pEnabledFeatures;
        const VkBool32 *have = (const VkBool32 *)&pd->features;
        for (size_t i = 0; i < sizeof(VkPhysicalDeviceFeatures) / sizeof(VkBool32); ++i)
            if (want[i] && !have[i]) return VK_ERROR_FEATURE_NOT_PRESENT;  // api_impl.rs:1418
    }
    for (uint32_t i = 0; i < ci->queueCreateInfoCount; ++i)
        if (ci->pQueueCreateInfos[i].queueFamilyIndex != 0 || ci->pQueueCreateInfos[i].queueCount > 1)
            return VK_ERROR_INITIALIZATION_FAILED;
    auto *dev = new (std::nothrow) Device();
    if (!dev) return VK_ERROR_OUT_OF_HOST_MEMORY;
    dev->physical = pd;
    dev->queue = std::make_unique<Queue>();
    dev->queue->device = dev;
    dev->engine = std::make_unique<Engine>();
    std::string err;
    if (!dev->engine->init(pd->hip_device, err)) {
        std::fprintf(stderr, "kazan: device initialisation failed: %s\n", err.c_str());
        delete dev;
        return VK_ERROR_INITIALIZATION_FAILED;
    }
    *pDevice = to<VkDevice>(dev);
    return VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkDestroyDevice(VkDevice device, const VkAllocationCallbacks *) {
    Device *d = from<Device>(device);
    if (!d) return;
    set_device(d);
    hipDeviceSynchronize();
    delete d;
}

VKAPI_ATTR void VKAPI_CALL vkGetDeviceQueue(VkDevice device, uint32_t family, uint32_t index, VkQueue *pQueue) {
    Device *d = from<Device>(device);
    *pQueue = (family == 0 && index == 0) ? to<VkQueue>(d->queue.get()) : nullptr;
}

VKAPI_ATTR VkResult VKAPI_CALL vkDeviceWaitIdle(VkDevice device) {
    Device *d = from<Device>(device);
    set_device(d);
    if (hipStreamSynchronize(d->engine->stream()) != hipSuccess) return VK_ERROR_DEVICE_LOST;
    d->engine->collect_stats();
    return VK_SUCCESS;
}

VKAPI_ATTR VkResult VKAPI_CALL vkQueueWaitIdle(VkQueue queue) {
    return vkDeviceWaitIdle(to<VkDevice>(from<Queue>(queue)->device));
}

VKAPI_ATTR VkResult VKAPI_CALL vkGetDrawStatsKAZAN(VkDevice device, uint64_t *pCounters, float *pGpuMs,
                                                   uint32_t *pBatches, uint32_t *pLaunches) {
    Device *d = from<Device>(device);
    VkResult r = vkDeviceWaitIdle(device);
    if (r != VK_SUCCESS) return r;
    const SubmitStats &s = d->engine->last_stats();
    if (pCounters)
        for (int i = 0; i < KZ_STAT_COUNT; ++i) pCounters[i] = s.counters[i];
    if (pGpuMs) *pGpuMs = s.gpu_ms;
    if (pBatches) *pBatches = s.batches;
    if (pLaunches) *pLaunches = s.launches;
    return VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkSetKernelTimingKAZAN(VkDevice device, VkBool32 enable) {
    from<Device>(device)->engine->set_kernel_timing(enable != 0);
}

VKAPI_ATTR VkResult VKAPI_CALL vkGetKernelTimesKAZAN(VkDevice device, float *pMs, uint32_t *pLaunches) {
    Device *d = from<Device>(device);
    VkResult r = vkDeviceWaitIdle(device);
    if (r != VK_SUCCESS) return r;
    const SubmitStats &s = d->engine->last_stats();
    for (int i = 0; i < KS_COUNT; ++i) {
        if (pMs) pMs[i] = s.kernel_ms[i];
        if (pLaunches) pLaunches[i] = s.kernel_launches[i];
    }
    return VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkSetTileProfileKAZAN(VkDevice device, VkBool32 enable) {
    from<Device>(device)->engine->set_tile_profile(enable != 0);
}

VKAPI_ATTR VkResult VKAPI_CALL vkGetTileProfileKAZAN(VkDevice device, uint32_t *pTileCount, float *pFeatures,
                                                     float *pFragments, float *pCycles, float *pPredicted) {
    Device *d = from<Device>(device);
    VkResult r = vkDeviceWaitIdle(device);
    if (r != VK_SUCCESS) return r;
    return d->engine->read_tile_profile(pTileCount, pFeatures, pFragments, pCycles, pPredicted) == hipSuccess
               ? VK_SUCCESS
               : VK_ERROR_DEVICE_LOST;
}

VKAPI_ATTR VkResult VKAPI_CALL vkSetCostModelKAZAN(VkDevice device, const void *pBlob, size_t size) {
    return from<Device>(device)->engine->set_cost_model(pBlob, size) ? VK_SUCCESS
                                                                       : VK_ERROR_INITIALIZATION_FAILED;
}

VKAPI_ATTR VkBool32 VKAPI_CALL vkHasCostModelKAZAN(VkDevice device) {
    return from<Device>(device)->engine->has_cost_model() ? VK_TRUE : VK_FALSE;
}

VKAPI_ATTR void VKAPI_CALL vkGetGraphStatsKAZAN(VkDevice device, uint32_t *pCaptures, uint32_t *pLaunches) {
    const Engine &e = *from<Device>(device)->engine;
    if (pCaptures) *pCaptures = e.graph_captures();
    if (pLaunches) *pLaunches = e.graph_launches();
}

VKAPI_ATTR void VKAPI_CALL vkGetTuningKAZAN(VkDevice device, uint32_t *pTileSize, uint32_t *pGridDivisor,
                                            uint32_t *pCostModel, uint32_t *pPersistentRaster,
                                            uint32_t *pMaxSplitLevel, uint32_t *pFragWeightScale) {
    const kz::BatchConfig &c = from<Device>(device)->engine->last_tuned();
    if (pTileSize) *pTileSize = c.shift ? 1u << c.shift : 0u;
    if (pGridDivisor) *pGridDivisor = c.grid_div;
    if (pCostModel) *pCostModel = c.model ? 1u : 0u;
    if (pPersistentRaster) *pPersistentRaster = c.persistent ? 1u : 0u;
    if (pMaxSplitLevel) *pMaxSplitLevel = c.max_level;
    if (pFragWeightScale) *pFragWeightScale = c.frag_scale;
}

VKAPI_ATTR VkResult VKAPI_CALL vkFillClusterSceneKAZAN(VkDevice device, VkBuffer buffer, VkDeviceSize offset,
                                                       const KazanClusterScene *p) {
    Device *d = from<Device>(device);
    Buffer *b = from<Buffer>(buffer);
    if (!p || !b || !b->address() || p->width == 0 || p->height == 0) return VK_ERROR_INITIALIZATION_FAILED;
    const VkDeviceSize bytes = (VkDeviceSize)p->triangles * 3 * 8 * sizeof(float);
    if (offset % 4 != 0 || offset > b->size || bytes > b->size - offset) return VK_ERROR_INITIALIZATION_FAILED;
    const uint32_t nc = (uint32_t)std::llround((double)p->triangles * std::min(std::max(p->cluster_frac, 0.0), 1.0));
    hipSetDevice(d->engine->device());
    hipStream_t st = d->engine->stream();
    if (kz_launch_gen_cluster(st, (float *)(b->address() + offset), p->triangles, nc, p->width, p->height,
                              p->sigma_frac, p->radius, p->seed) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return VK_ERROR_DEVICE_LOST;
    return VK_SUCCESS;
}

// ===========================================================================
// memory, buffers, images

VKAPI_ATTR VkResult VKAPI_CALL vkAllocateMemory(VkDevice device, const VkMemoryAllocateInfo *ai,
                                                const VkAllocationCallbacks *, VkDeviceMemory *pMemory) {
    Device *d = from<Device>(device);
    if (ai->allocationSize > (VkDeviceSize)INT64_MAX) return VK_ERROR_OUT_OF_DEVICE_MEMORY;  // api_impl.rs:3460
    if (ai->memoryTypeIndex >= d->physical->memory.memoryTypeCount) return VK_ERROR_OUT_OF_DEVICE_MEMORY;
    set_device(d);
    auto *m = new (std::nothrow) DeviceMemory();
    if (!m) return VK_ERROR_OUT_OF_HOST_MEMORY;
    m->size = ai->allocationSize;
    m->type = ai->memoryTypeIndex;
    m->host = ai->memoryTypeIndex == 1;
    const size_t bytes = std::max<size_t>(ai->allocationSize, 1);
    const bool host = m->host;
    hipError_t e = host ? hipHostMalloc((void **)&m->ptr, bytes, hipHostMallocDefault)
                        : hipMalloc((void **)&m->ptr, bytes);
    if (e != hipSuccess) {
        delete m;
        return host ? VK_ERROR_OUT_OF_HOST_MEMORY : VK_ERROR_OUT_OF_DEVICE_MEMORY;
    }
    *pMemory = to<VkDeviceMemory>(m);
    return VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkFreeMemory(VkDevice device, VkDeviceMemory memory, const VkAllocationCallbacks *) {
    DeviceMemory *m = from<DeviceMemory>(memory);
    if (!m) return;
    set_device(from<Device>(device));
    if (m->host) hipHostFree(m->ptr);
    else hipFree(m->ptr);
    delete m;
}

VKAPI_ATTR VkResult VKAPI_CALL vkMapMemory(VkDevice, VkDeviceMemory memory, VkDeviceSize offset, VkDeviceSize,
                                           VkMemoryMapFlags, void **ppData) {
    DeviceMemory *m = from<DeviceMemory>(memory);
    if (!m->host) return VK_ERROR_MEMORY_MAP_FAILED;
    *ppData = m->ptr + offset;
    return VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkUnmapMemory(VkDevice, VkDeviceMemory) {}

VKAPI_ATTR VkResult VKAPI_CALL vkFlushMappedMemoryRanges(VkDevice, uint32_t, const VkMappedMemoryRange *) {
    return VK_SUCCESS;
}

VKAPI_ATTR VkResult VKAPI_CALL vkInvalidateMappedMemoryRanges(VkDevice, uint32_t, const VkMappedMemoryRange *) {
    return VK_SUCCESS;
}

VKAPI_ATTR VkResult VKAPI_CALL vkCreateBuffer(VkDevice, const VkBufferCreateInfo *ci, const VkAllocationCallbacks *,
                                              VkBuffer *pBuffer) {
    if (ci->size > (VkDeviceSize)INT64_MAX) return VK_ERROR_OUT_OF_DEVICE_MEMORY;  // api_impl.rs:3812
    auto *b = new (std::nothrow) Buffer();
    if (!b) return VK_ERROR_OUT_OF_HOST_MEMORY;
    b->size = ci->size;
    *pBuffer = to<VkBuffer>(b);
    return VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkDestroyBuffer(VkDevice, VkBuffer buffer, const VkAllocationCallbacks *) {
    delete from<Buffer>(buffer);
}

VKAPI_ATTR void VKAPI_CALL vkGetBufferMemoryRequirements(VkDevice, VkBuffer buffer, VkMemoryRequirements *r) {
    Buffer *b = from<Buffer>(buffer);
    r->size = (b->size + kBufferAlignment - 1) / kBufferAlignment * kBufferAlignment;
    r->alignment = kBufferAlignment;
    r->memoryTypeBits = 0x3;
}

VKAPI_ATTR VkResult VKAPI_CALL vkBindBufferMemory(VkDevice, VkBuffer buffer, VkDeviceMemory memory,
                                                  VkDeviceSize offset) {
    Buffer *b = from<Buffer>(buffer);
    b->memory = from<DeviceMemory>(memory);
    b->offset = offset;
    return VK_SUCCESS;
}

VKAPI_ATTR VkResult VKAPI_CALL vkCreateImage(VkDevice, const VkImageCreateInfo *ci, const VkAllocationCallbacks *,
                                             VkImage *pImage) {
    if (ci->imageType != VK_IMAGE_TYPE_2D || ci->mipLevels != 1 || ci->arrayLayers != 1 || ci->extent.depth != 1)
        return VK_ERROR_FORMAT_NOT_SUPPORTED;
    if (ci->format != VK_FORMAT_R8G8B8A8_UNORM && ci->format != VK_FORMAT_D32_SFLOAT &&
        ci->format != VK_FORMAT_B8G8R8A8_UNORM)
        return VK_ERROR_FORMAT_NOT_SUPPORTED;
    if (ci->samples != VK_SAMPLE_COUNT_1_BIT && ci->samples != VK_SAMPLE_COUNT_4_BIT)
        return VK_ERROR_FORMAT_NOT_SUPPORTED;
    auto *im = new (std::nothrow) Image();
    if (!im) return VK_ERROR_OUT_OF_HOST_MEMORY;
    im->format = ci->format;
    im->width = ci->extent.width;
    im->height = ci->extent.height;
    im->samples = (uint32_t)ci->samples;
    // linear layout, samples interleaved per pixel (image.rs:58-88 lays out
    // RGBA8 as 4*w*h; D32 and 4x samples extend it the same way)
    im->size = (VkDeviceSize)im->width * im->height * im->samples * 4;
    *pImage = to<VkImage>(im);
    return VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkDestroyImage(VkDevice, VkImage image, const VkAllocationCallbacks *) {
    delete from<Image>(image);
}

VKAPI_ATTR void VKAPI_CALL vkGetImageMemoryRequirements(VkDevice, VkImage image, VkMemoryRequirements *r) {
    Image *im = from<Image>(image);
    r->size = (im->size + kImageAlignment - 1) / kImageAlignment * kImageAlignment;
    r->alignment = kImageAlignment;
    r->memoryTypeBits = 0x3;
}

VKAPI_ATTR VkResult VKAPI_CALL vkBindImageMemory(VkDevice, VkImage image, VkDeviceMemory memory,
                                                 VkDeviceSize offset) {
    Image *im = from<Image>(image);
    im->memory = from<DeviceMemory>(memory);
    im->offset = offset;
    return VK_SUCCESS;
}

VKAPI_ATTR VkResult VKAPI_CALL vkCreateImageView(VkDevice, const VkImageViewCreateInfo *ci,
                                                 const VkAllocationCallbacks *, VkImageView *pView) {
    auto *v = new (std::nothrow) ImageView();
    if (!v) return VK_ERROR_OUT_OF_HOST_MEMORY;
    v->image = from<Image>(ci->image);
    v->format = ci->format;
    *pView = to<VkImageView>(v);
    return VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkDestroyImageView(VkDevice, VkImageView view, const VkAllocationCallbacks *) {
    delete from<ImageView>(view);
}

VKAPI_ATTR VkResult VKAPI_CALL vkCreateSampler(VkDevice, const VkSamplerCreateInfo *ci,
                                               const VkAllocationCallbacks *, VkSampler *pSampler) {
    auto *s = new (std::nothrow) Sampler();
    if (!s) return VK_ERROR_OUT_OF_HOST_MEMORY;
    s->mag_filter = ci->magFilter;
    s->min_filter = ci->minFilter;
    s->addr_u = ci->addressModeU == VK_SAMPLER_ADDRESS_MODE_REPEAT ? 0u : 2u;
    s->addr_v = ci->addressModeV == VK_SAMPLER_ADDRESS_MODE_REPEAT ? 0u : 2u;
    *pSampler = to<VkSampler>(s);
    return VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkDestroySampler(VkDevice, VkSampler sampler, const VkAllocationCallbacks *) {
    delete from<Sampler>(sampler);
}

// ===========================================================================
// shaders, descriptors, layouts, render passes

VKAPI_ATTR VkResult VKAPI_CALL vkCreateShaderModule(VkDevice, const VkShaderModuleCreateInfo *ci,
                                                    const VkAllocationCallbacks *, VkShaderModule *pModule) {
    // shader_module.rs:3-5 / api_impl.rs:3946-3967: codeSize % 4 == 0, non-zero
    if (ci->codeSize == 0 || ci->codeSize % 4 != 0) return VK_ERROR_INVALID_SHADER_NV;
    auto *m = new (std::nothrow) ShaderModule();
    if (!m) return VK_ERROR_OUT_OF_HOST_MEMORY;
    m->code.assign(ci->pCode, ci->pCode + ci->codeSize / 4);
    *pModule = to<VkShaderModule>(m);
    return VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkDestroyShaderModule(VkDevice, VkShaderModule module, const VkAllocationCallbacks *) {
    delete from<ShaderModule>(module);
}

// Pipeline caches (the reference leaves vkCreatePipelineCache a stub,
// vulkan-driver/src/api_impl.rs:3979). Code objects live in the on-disk
// cache keyed by content hash; a VkPipelineCache carries them between
// processes: its data holds the code objects of the pipelines created with
// it, and creating a cache from such data (same vendor, device and
// pipelineCacheUUID) puts them into the on-disk cache, where pipeline
// creation finds them instead of compiling.
namespace {
constexpr size_t kCacheHeader = 32, kCacheName = 32;

void cache_add(PipelineCache *c, const std::string &name, std::vector<char> code) {
    for (const auto &e : c->entries)
        if (e.first == name) return;
    c->entries.emplace_back(name, std::move(code));
}

bool read_file(const std::string &path, std::vector<char> &out) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    out.clear();
    char buf[65536];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) out.insert(out.end(), buf, buf + n);
    std::fclose(f);
    return !out.empty();
}


void cache_header(const PhysicalDevice &pd, uint8_t out[kCacheHeader]) {
    const uint32_t w[4] = {(uint32_t)kCacheHeader, 1u /* VK_PIPELINE_CACHE_HEADER_VERSION_ONE */, pd.props.vendorID,
                           pd.props.deviceID};
    std::memcpy(out, w, 16);
    std::memcpy(out + 16, pd.props.pipelineCacheUUID, 16);
}

// Entries of data written by vkGetPipelineCacheData for this device; other
// data (another device, driver version or format) is ignored, as Vulkan
// allows, and so is an entry whose checksum does not match its bytes or
// whose bytes are not a gfx950 code object. Entries stay in the cache
// object: pipelines created with it load them (nothing is written to the
// shared on-disk cache), and one that does not load is compiled instead.
void cache_load(const PhysicalDevice &pd, PipelineCache *c, const uint8_t *data, size_t size) {
    uint8_t want[kCacheHeader];
    cache_header(pd, want);
    if (size < kCacheHeader + 4 || std::memcmp(data, want, kCacheHeader) != 0) return;
    uint32_t count;
    std::memcpy(&count, data + kCacheHeader, 4);
    size_t at = kCacheHeader + 4;
    for (uint32_t i = 0; i < count; ++i) {
        if (size - at < kCacheName + 16) return;
        char name[kCacheName + 1] = {};
        std::memcpy(name, data + at, kCacheName);
        uint64_t n, sum;
        std::memcpy(&n, data + at + kCacheName, 8);
        std::memcpy(&sum, data + at + kCacheName + 8, 8);
        at += kCacheName + 16;
        if (n == 0 || n > size - at) return;
        const std::string nm(name);
        const char *bytes = (const char *)data + at;
        at += n;
        if (nm.empty() || nm.find('/') != std::string::npos || nm.find("..") != std::string::npos) continue;
        if (kz::fnv1a64(bytes, n) != sum || !kz::is_amdgpu_code_object(bytes, n)) continue;
        cache_add(c, nm, std::vector<char>(bytes, bytes + n));
    }
}
}  // namespace

VKAPI_ATTR VkResult VKAPI_CALL vkCreatePipelineCache(VkDevice device, const VkPipelineCacheCreateInfo *ci,
                                                     const VkAllocationCallbacks *, VkPipelineCache *pCache) {
    auto *c = new (std::nothrow) PipelineCache();
    if (!c) return VK_ERROR_OUT_OF_HOST_MEMORY;
    if (ci && ci->initialDataSize && ci->pInitialData)
        cache_load(*from<Device>(device)->physical, c, (const uint8_t *)ci->pInitialData, ci->initialDataSize);
    *pCache = to<VkPipelineCache>(c);
    return VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkDestroyPipelineCache(VkDevice, VkPipelineCache cache, const VkAllocationCallbacks *) {
    delete from<PipelineCache>(cache);
}

// Vulkan's two-call form; with a short buffer, the header and the entries
// that fit (still valid initial data) and VK_INCOMPLETE.
VKAPI_ATTR VkResult VKAPI_CALL vkGetPipelineCacheData(VkDevice device, VkPipelineCache cache, size_t *pDataSize,
                                                      void *pData) {
    PipelineCache *c = from<PipelineCache>(cache);
    std::lock_guard<std::mutex> lock(c->mu);
    if (!c->pending.empty()) {  // variants compiled since their pipeline was created (tile-size tuning)
        const std::string dir = kz::pipeline_code_dir();
        std::vector<std::string> still;
        for (const std::string &name : c->pending) {
            std::vector<char> code;
            if (read_file(dir + "/" + name, code) && kz::is_amdgpu_code_object(code.data(), code.size()))
                cache_add(c, name, std::move(code));
            else
                still.push_back(name);
        }
        c->pending.swap(still);
    }
    size_t total = kCacheHeader + 4;
    for (const auto &e : c->entries) total += kCacheName + 16 + e.second.size();
    if (!pData) {
        *pDataSize = total;
        return VK_SUCCESS;
    }
    if (*pDataSize < kCacheHeader + 4) {
        *pDataSize = 0;
        return VK_INCOMPLETE;
    }
    uint8_t *out = (uint8_t *)pData;
    cache_header(*from<Device>(device)->physical, out);
    size_t at = kCacheHeader + 4;
    uint32_t count = 0;
    for (const auto &e : c->entries) {
        const size_t need = kCacheName + 16 + e.second.size();
        if (*pDataSize - at < need) break;
        std::memset(out + at, 0, kCacheName);
        std::memcpy(out + at, e.first.data(), std::min(e.first.size(), kCacheName));
        const uint64_t n = e.second.size(), sum = kz::fnv1a64(e.second.data(), n);
        std::memcpy(out + at + kCacheName, &n, 8);
        std::memcpy(out + at + kCacheName + 8, &sum, 8);
        std::memcpy(out + at + kCacheName + 16, e.second.data(), n);
        at += need;
        ++count;
    }
    std::memcpy(out + kCacheHeader, &count, 4);
    const bool all = count == c->entries.size();
    *pDataSize = at;
    return all ? VK_SUCCESS : VK_INCOMPLETE;
}

VKAPI_ATTR VkResult VKAPI_CALL vkMergePipelineCaches(VkDevice, VkPipelineCache dst, uint32_t n,
                                                     const VkPipelineCache *srcs) {
    PipelineCache *d = from<PipelineCache>(dst);
    for (uint32_t i = 0; i < n; ++i) {
        PipelineCache *s = from<PipelineCache>(srcs[i]);
        if (s == d) continue;
        // a snapshot of the source under its lock, then the destination under its own
        std::vector<std::pair<std::string, std::vector<char>>> entries;
        std::vector<std::string> pending;
        {
            std::lock_guard<std::mutex> lock(s->mu);
            entries = s->entries;
            pending = s->pending;
        }
        std::lock_guard<std::mutex> lock(d->mu);
        for (auto &e : entries) cache_add(d, e.first, std::move(e.second));
        for (auto &name : pending)
            if (std::find(d->pending.begin(), d->pending.end(), name) == d->pending.end()) d->pending.push_back(name);
    }
    return VK_SUCCESS;
}

VKAPI_ATTR VkResult VKAPI_CALL vkCreateDescriptorSetLayout(VkDevice, const VkDescriptorSetLayoutCreateInfo *ci,
                                                           const VkAllocationCallbacks *,
                                                           VkDescriptorSetLayout *pLayout) {
    auto *l = new (std::nothrow) DescriptorSetLayout();
    if (!l) return VK_ERROR_OUT_OF_HOST_MEMORY;
    for (uint32_t i = 0; i < ci->bindingCount; ++i) {
        DescriptorBinding b;
        b.binding = ci->pBindings[i].binding;
        b.type = ci->pBindings[i].descriptorType;
        b.count = ci->pBindings[i].descriptorCount;
        if (ci->pBindings[i].pImmutableSamplers)
            for (uint32_t k = 0; k < b.count; ++k)
                b.immutable_samplers.push_back(from<Sampler>(ci->pBindings[i].pImmutableSamplers[k]));
        l->bindings.push_back(b);
    }
    std::sort(l->bindings.begin(), l->bindings.end(),
              [](const DescriptorBinding &a, const DescriptorBinding &b) { return a.binding < b.binding; });
    *pLayout = to<VkDescriptorSetLayout>(l);
    return VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkDestroyDescriptorSetLayout(VkDevice, VkDescriptorSetLayout l,
                                                        const VkAllocationCallbacks *) {
    delete from<DescriptorSetLayout>(l);
}

VKAPI_ATTR VkResult VKAPI_CALL vkCreateDescriptorPool(VkDevice, const VkDescriptorPoolCreateInfo *,
                                                      const VkAllocationCallbacks *, VkDescriptorPool *pPool) {
    auto *p = new (std::nothrow) DescriptorPool();
    if (!p) return VK_ERROR_OUT_OF_HOST_MEMORY;
    *pPool = to<VkDescriptorPool>(p);
    return VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkDestroyDescriptorPool(VkDevice, VkDescriptorPool pool, const VkAllocationCallbacks *) {
    delete from<DescriptorPool>(pool);
}

VKAPI_ATTR VkResult VKAPI_CALL vkResetDescriptorPool(VkDevice, VkDescriptorPool pool, VkDescriptorPoolResetFlags) {
    from<DescriptorPool>(pool)->sets.clear();
    return VK_SUCCESS;
}

VKAPI_ATTR VkResult VKAPI_CALL vkAllocateDescriptorSets(VkDevice, const VkDescriptorSetAllocateInfo *ai,
                                                        VkDescriptorSet *pSets) {
    DescriptorPool *pool = from<DescriptorPool>(ai->descriptorPool);
    for (uint32_t i = 0; i < ai->descriptorSetCount; ++i) {
        auto s = std::make_unique<DescriptorSet>();
        s->layout = from<DescriptorSetLayout>(ai->pSetLayouts[i]);
        for (const auto &b : s->layout->bindings) {
            DescriptorEntry e;
            e.type = b.type;
            if (!b.immutable_samplers.empty()) e.sampler = b.immutable_samplers[0];
            s->entries[b.binding] = e;
        }
        pSets[i] = to<VkDescriptorSet>(s.get());
        pool->sets.push_back(std::move(s));
    }
    return VK_SUCCESS;
}

VKAPI_ATTR VkResult VKAPI_CALL vkFreeDescriptorSets(VkDevice, VkDescriptorPool, uint32_t, const VkDescriptorSet *) {
    return VK_SUCCESS;  // sets are owned by their pool
}

VKAPI_ATTR void VKAPI_CALL vkUpdateDescriptorSets(VkDevice, uint32_t writeCount, const VkWriteDescriptorSet *w,
                                                  uint32_t copyCount, const VkCopyDescriptorSet *c) {
    for (uint32_t i = 0; i < writeCount; ++i) {
        DescriptorSet *s = from<DescriptorSet>(w[i].dstSet);
        for (uint32_t k = 0; k < w[i].descriptorCount; ++k) {
            DescriptorEntry &e = s->entries[w[i].dstBinding + w[i].dstArrayElement + k];
            e.type = w[i].descriptorType;
            switch (w[i].descriptorType) {
            case VK_DESCRIPTOR_TYPE_UNIFORM_BUFFER:
            case VK_DESCRIPTOR_TYPE_STORAGE_BUFFER:
                e.buffer = from<Buffer>(w[i].pBufferInfo[k].buffer);
                e.offset = w[i].pBufferInfo[k].offset;
                e.range = w[i].pBufferInfo[k].range;
                break;
            case VK_DESCRIPTOR_TYPE_COMBINED_IMAGE_SAMPLER:
            case VK_DESCRIPTOR_TYPE_SAMPLED_IMAGE:
            case VK_DESCRIPTOR_TYPE_SAMPLER:
                if (w[i].pImageInfo[k].imageView) e.view = from<ImageView>(w[i].pImageInfo[k].imageView);
                if (w[i].pImageInfo[k].sampler) e.sampler = from<Sampler>(w[i].pImageInfo[k].sampler);
                break;
            default: break;
            }
        }
    }
    for (uint32_t i = 0; i < copyCount; ++i) {
        DescriptorSet *src = from<DescriptorSet>(c[i].srcSet), *dst = from<DescriptorSet>(c[i].dstSet);
        for (uint32_t k = 0; k < c[i].descriptorCount; ++k)
            dst->entries[c[i].dstBinding + c[i].dstArrayElement + k] =
                src->entries[c[i].srcBinding + c[i].srcArrayElement + k];
    }
}

VKAPI_ATTR VkResult VKAPI_CALL vkCreatePipelineLayout(VkDevice, const VkPipelineLayoutCreateInfo *ci,
                                                      const VkAllocationCallbacks *, VkPipelineLayout *pLayout) {
    auto *l = new (std::nothrow) PipelineLayout();
    if (!l) return VK_ERROR_OUT_OF_HOST_MEMORY;
    for (uint32_t i = 0; i < ci->setLayoutCount; ++i) l->set_layouts.push_back(from<DescriptorSetLayout>(ci->pSetLayouts[i]));
    // push-constant size = max(offset + size) (api_impl.rs:4061-4153)
    for (uint32_t i = 0; i < ci->pushConstantRangeCount; ++i)
        l->push_constant_size =
            std::max(l->push_constant_size, ci->pPushConstantRanges[i].offset + ci->pPushConstantRanges[i].size);
    if (l->push_constant_size > KZ_PUSH_WORDS * 4) {
        delete l;
        return VK_ERROR_OUT_OF_HOST_MEMORY;
    }
    *pLayout = to<VkPipelineLayout>(l);
    return VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkDestroyPipelineLayout(VkDevice, VkPipelineLayout l, const VkAllocationCallbacks *) {
    delete from<PipelineLayout>(l);
}

VKAPI_ATTR VkResult VKAPI_CALL vkCreateRenderPass(VkDevice, const VkRenderPassCreateInfo *ci,
                                                  const VkAllocationCallbacks *, VkRenderPass *pRenderPass) {
    auto *rp = new (std::nothrow) RenderPass();
    if (!rp) return VK_ERROR_OUT_OF_HOST_MEMORY;
    rp->attachments.assign(ci->pAttachments, ci->pAttachments + ci->attachmentCount);
    for (uint32_t i = 0; i < ci->subpassCount; ++i) {
        const VkSubpassDescription &s = ci->pSubpasses[i];
        SubpassInfo si;
        for (uint32_t k = 0; k < s.colorAttachmentCount; ++k) {
            si.color.push_back(s.pColorAttachments[k].attachment);
            si.resolve.push_back(s.pResolveAttachments ? s.pResolveAttachments[k].attachment : VK_ATTACHMENT_UNUSED);
        }
        if (s.pDepthStencilAttachment) si.depth = s.pDepthStencilAttachment->attachment;
        if (si.color.size() > 1) {
            delete rp;
            return VK_ERROR_FEATURE_NOT_PRESENT;
        }
        rp->subpasses.push_back(si);
    }
    *pRenderPass = to<VkRenderPass>(rp);
    return VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkDestroyRenderPass(VkDevice, VkRenderPass rp, const VkAllocationCallbacks *) {
    delete from<RenderPass>(rp);
}

VKAPI_ATTR VkResult VKAPI_CALL vkCreateFramebuffer(VkDevice, const VkFramebufferCreateInfo *ci,
                                                   const VkAllocationCallbacks *, VkFramebuffer *pFb) {
    auto *fb = new (std::nothrow) Framebuffer();
    if (!fb) return VK_ERROR_OUT_OF_HOST_MEMORY;
    fb->render_pass = from<RenderPass>(ci->renderPass);
    for (uint32_t i = 0; i < ci->attachmentCount; ++i) fb->views.push_back(from<ImageView>(ci->pAttachments[i]));
    fb->width = ci->width;
    fb->height = ci->height;
    *pFb = to<VkFramebuffer>(fb);
    return VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkDestroyFramebuffer(VkDevice, VkFramebuffer fb, const VkAllocationCallbacks *) {
    delete from<Framebuffer>(fb);
}

// ===========================================================================
// pipelines

VKAPI_ATTR VkResult VKAPI_CALL vkCreateGraphicsPipelines(VkDevice device, VkPipelineCache cache, uint32_t count,
                                                         const VkGraphicsPipelineCreateInfo *ci,
                                                         const VkAllocationCallbacks *, VkPipeline *pPipelines) {
    Device *d = from<Device>(device);
    set_device(d);
    VkResult result = VK_SUCCESS;
    for (uint32_t i = 0; i < count; ++i) {
        Pipeline *p = nullptr;
        std::string err;
        if (!create_graphics_pipeline(d, &ci[i], cache ? from<PipelineCache>(cache) : nullptr, &p, err)) {
            std::fprintf(stderr, "kazan: vkCreateGraphicsPipelines: %s\n", err.c_str());
            pPipelines[i] = nullptr;
            result = VK_ERROR_INVALID_SHADER_NV;
            continue;
        }
        pPipelines[i] = to<VkPipeline>(p);
        if (cache) {  // the pipeline's code objects (each tile size's variant; those compiled later at data time)
            PipelineCache *c = from<PipelineCache>(cache);
            const std::string dir = kz::pipeline_code_dir();
            std::lock_guard<std::mutex> lock(c->mu);
            for (uint32_t s = kz::kMinTileShift; s <= kz::kMaxTileShift; ++s) {
                if (!p->variants.fits(s)) continue;  // a tile size this pipeline can never compile
                FixedState fs = p->fixed;
                fs.tile_shift = s;
                const std::string name = kz::pipeline_code_name(fs, p->shader_source);
                bool have = false;
                for (const auto &e : c->entries) have = have || e.first == name;
                if (have) continue;
                std::vector<char> code;
                if (read_file(dir + "/" + name, code) && kz::is_amdgpu_code_object(code.data(), code.size()))
                    cache_add(c, name, std::move(code));
                else if (std::find(c->pending.begin(), c->pending.end(), name) == c->pending.end())
                    c->pending.push_back(name);
            }
        }
    }
    return result;
}

// Reference: vkCreateComputePipelines (api_impl.rs) -> ComputePipeline::create
// (pipeline.rs:122-154).
VKAPI_ATTR VkResult VKAPI_CALL vkCreateComputePipelines(VkDevice device, VkPipelineCache, uint32_t count,
                                                        const VkComputePipelineCreateInfo *ci,
                                                        const VkAllocationCallbacks *, VkPipeline *pPipelines) {
    Device *d = from<Device>(device);
    set_device(d);
    VkResult result = VK_SUCCESS;
    for (uint32_t i = 0; i < count; ++i) {
        Pipeline *p = nullptr;
        std::string err;
        if (!create_compute_pipeline(d, &ci[i], &p, err)) {
            std::fprintf(stderr, "kazan: vkCreateComputePipelines: %s\n", err.c_str());
            pPipelines[i] = nullptr;
            result = VK_ERROR_INVALID_SHADER_NV;
            continue;
        }
        pPipelines[i] = to<VkPipeline>(p);
    }
    return result;
}

VKAPI_ATTR void VKAPI_CALL vkDestroyPipeline(VkDevice device, VkPipeline pipeline, const VkAllocationCallbacks *) {
    if (!pipeline) return;
    set_device(from<Device>(device));
    destroy_pipeline(from<Pipeline>(pipeline));
}

// ===========================================================================
// command pools and buffers

VKAPI_ATTR VkResult VKAPI_CALL vkCreateCommandPool(VkDevice, const VkCommandPoolCreateInfo *ci,
                                                   const VkAllocationCallbacks *, VkCommandPool *pPool) {
    if (ci->queueFamilyIndex != 0) return VK_ERROR_INITIALIZATION_FAILED;
    auto *p = new (std::nothrow) CommandPool();
    if (!p) return VK_ERROR_OUT_OF_HOST_MEMORY;
    *pPool = to<VkCommandPool>(p);
    return VK_SUCCESS;
}

static void reset_cb(CommandBuffer *cb) {
    if (!cb->device_allocs.empty() || cb->graph.exec) {
        set_device(cb->device);
        hipStreamSynchronize(cb->device->engine->stream());
        for (void *p : cb->device_allocs) hipFree(p);
    }
    cb->graph.reset();
    cb->graph.eager_runs = 0;
    cb->device_allocs.clear();
    cb->ops.clear();
    cb->st = RecordState();
    cb->recording = false;
}

VKAPI_ATTR void VKAPI_CALL vkDestroyCommandPool(VkDevice, VkCommandPool pool, const VkAllocationCallbacks *) {
    CommandPool *p = from<CommandPool>(pool);
    if (!p) return;
    for (CommandBuffer *cb : p->buffers) {
        reset_cb(cb);
        delete cb;
    }
    delete p;
}

VKAPI_ATTR VkResult VKAPI_CALL vkResetCommandPool(VkDevice, VkCommandPool pool, VkCommandPoolResetFlags) {
    for (CommandBuffer *cb : from<CommandPool>(pool)->buffers) reset_cb(cb);
    return VK_SUCCESS;
}

VKAPI_ATTR VkResult VKAPI_CALL vkAllocateCommandBuffers(VkDevice device, const VkCommandBufferAllocateInfo *ai,
                                                        VkCommandBuffer *pCbs) {
    CommandPool *p = from<CommandPool>(ai->commandPool);
    if (ai->level != VK_COMMAND_BUFFER_LEVEL_PRIMARY) return VK_ERROR_FEATURE_NOT_PRESENT;
    for (uint32_t i = 0; i < ai->commandBufferCount; ++i) {
        auto *cb = new (std::nothrow) CommandBuffer();
        if (!cb) return VK_ERROR_OUT_OF_HOST_MEMORY;
        cb->device = from<Device>(device);
        cb->pool = p;
        p->buffers.push_back(cb);
        pCbs[i] = to<VkCommandBuffer>(cb);
    }
    return VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkFreeCommandBuffers(VkDevice, VkCommandPool pool, uint32_t n,
                                                const VkCommandBuffer *pCbs) {
    CommandPool *p = from<CommandPool>(pool);
    for (uint32_t i = 0; i < n; ++i) {
        CommandBuffer *cb = from<CommandBuffer>(pCbs[i]);
        if (!cb) continue;
        reset_cb(cb);
        p->buffers.erase(std::remove(p->buffers.begin(), p->buffers.end(), cb), p->buffers.end());
        delete cb;
    }
}

VKAPI_ATTR VkResult VKAPI_CALL vkBeginCommandBuffer(VkCommandBuffer commandBuffer, const VkCommandBufferBeginInfo *) {
    CommandBuffer *cb = from<CommandBuffer>(commandBuffer);
    reset_cb(cb);
    cb->recording = true;
    return VK_SUCCESS;
}

VKAPI_ATTR VkResult VKAPI_CALL vkEndCommandBuffer(VkCommandBuffer commandBuffer) {
    CommandBuffer *cb = from<CommandBuffer>(commandBuffer);
    if (!cb->recording) return VK_ERROR_INITIALIZATION_FAILED;
    set_device(cb->device);
    // upload each batch's draw records once; every submission reuses them
    for (Op &op : cb->ops) {
        if (op.kind != Op::Batch || op.draws.empty()) continue;
        // the records, then their first triangles as a dense array, then
        // the draw of the first triangle of every KZ_BIN_BLOCK-triangle
        // block (and the last draw), which narrows the device's search
        const size_t n = op.draws.size(), rec_bytes = n * sizeof(KzDraw);
        const size_t nblk = ((size_t)op.frame.total_tris + KZ_BIN_BLOCK - 1) / KZ_BIN_BLOCK + 1;
        std::vector<char> host(rec_bytes + (n + nblk) * sizeof(uint32_t));
        std::memcpy(host.data(), op.draws.data(), rec_bytes);
        uint32_t *begin = (uint32_t *)(host.data() + rec_bytes);
        for (size_t i = 0; i < n; ++i) begin[i] = op.draws[i].tri_begin;
        uint32_t *block_draw = begin + n;
        for (size_t b = 0, d = 0; b < nblk; ++b) {
            const uint64_t g = std::min<uint64_t>((uint64_t)b * KZ_BIN_BLOCK, op.frame.total_tris ? op.frame.total_tris - 1 : 0);
            while (d + 1 < n && op.draws[d + 1].tri_begin <= g) ++d;
            block_draw[b] = (uint32_t)d;
        }
        void *p = nullptr;
        if (hipMalloc(&p, host.size()) != hipSuccess) return VK_ERROR_OUT_OF_DEVICE_MEMORY;
        cb->device_allocs.push_back(p);
        if (hipMemcpy(p, host.data(), host.size(), hipMemcpyHostToDevice) != hipSuccess) return VK_ERROR_DEVICE_LOST;
        op.d_draws = (KzDraw *)p;
        op.d_draw_begin = (uint32_t *)((char *)p + rec_bytes);
        op.d_block_draw = op.d_draw_begin + n;
        op.frame.ndraws = (uint32_t)n;
    }
    cb->recording = false;
    return VK_SUCCESS;
}

VKAPI_ATTR VkResult VKAPI_CALL vkResetCommandBuffer(VkCommandBuffer commandBuffer, VkCommandBufferResetFlags) {
    reset_cb(from<CommandBuffer>(commandBuffer));
    return VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkCmdBindPipeline(VkCommandBuffer commandBuffer, VkPipelineBindPoint point,
                                             VkPipeline pipeline) {
    CommandBuffer *cb = from<CommandBuffer>(commandBuffer);
    Pipeline *p = from<Pipeline>(pipeline);
    if (point == VK_PIPELINE_BIND_POINT_COMPUTE) {
        cb->st.compute_pipeline = p;
        return;
    }
    cb->st.pipeline = p;
    if (!p->dynamic_viewport) cb->st.viewport = p->viewport;
    if (!p->dynamic_scissor) cb->st.scissor = p->scissor;
    cb->st.batch_open = false;
}

VKAPI_ATTR void VKAPI_CALL vkCmdSetViewport(VkCommandBuffer commandBuffer, uint32_t first, uint32_t n,
                                            const VkViewport *v) {
    CommandBuffer *cb = from<CommandBuffer>(commandBuffer);
    if (first == 0 && n >= 1) cb->st.viewport = v[0];
    cb->st.batch_open = false;
}

VKAPI_ATTR void VKAPI_CALL vkCmdSetScissor(VkCommandBuffer commandBuffer, uint32_t first, uint32_t n,
                                           const VkRect2D *s) {
    CommandBuffer *cb = from<CommandBuffer>(commandBuffer);
    if (first == 0 && n >= 1) cb->st.scissor = s[0];
    cb->st.batch_open = false;
}

VKAPI_ATTR void VKAPI_CALL vkCmdBindVertexBuffers(VkCommandBuffer commandBuffer, uint32_t first, uint32_t n,
                                                  const VkBuffer *buffers, const VkDeviceSize *offsets) {
    CommandBuffer *cb = from<CommandBuffer>(commandBuffer);
    for (uint32_t i = 0; i < n && first + i < KZ_MAX_BINDINGS; ++i) {
        cb->st.vb[first + i] = from<Buffer>(buffers[i]);
        cb->st.vb_offset[first + i] = offsets[i];
    }
}

VKAPI_ATTR void VKAPI_CALL vkCmdBindIndexBuffer(VkCommandBuffer commandBuffer, VkBuffer buffer, VkDeviceSize offset,
                                                VkIndexType type) {
    CommandBuffer *cb = from<CommandBuffer>(commandBuffer);
    cb->st.ib = from<Buffer>(buffer);
    cb->st.ib_offset = offset;
    cb->st.ib_type = type;
}

VKAPI_ATTR void VKAPI_CALL vkCmdBindDescriptorSets(VkCommandBuffer commandBuffer, VkPipelineBindPoint point,
                                                   VkPipelineLayout, uint32_t firstSet, uint32_t n,
                                                   const VkDescriptorSet *sets, uint32_t, const uint32_t *) {
    CommandBuffer *cb = from<CommandBuffer>(commandBuffer);
    DescriptorSet **slots = point == VK_PIPELINE_BIND_POINT_COMPUTE ? cb->st.compute_sets : cb->st.sets;
    for (uint32_t i = 0; i < n && firstSet + i < 8; ++i) slots[firstSet + i] = from<DescriptorSet>(sets[i]);
    if (point != VK_PIPELINE_BIND_POINT_COMPUTE) cb->st.batch_open = false;
}

// vkCmdDispatch / vkCmdDispatchBase (reference: api_impl.rs:4677, a stub):
// the bound compute pipeline's kernel over x * y * z workgroups, with the
// descriptors and push constants bound now, in order with the other
// commands of the buffer (one stream executes them).
static void record_dispatch(CommandBuffer *cb, uint32_t bx, uint32_t by, uint32_t bz, uint32_t x, uint32_t y,
                            uint32_t z) {
    RecordState &st = cb->st;
    Pipeline *p = st.compute_pipeline;
    if (!p || !p->compute || !p->cs_kernel || x == 0 || y == 0 || z == 0) return;
    Op op;
    op.kind = Op::Dispatch;
    op.cs_kernel = p->cs_kernel;
    op.groups[0] = x;
    op.groups[1] = y;
    op.groups[2] = z;
    for (int k = 0; k < 3; ++k) op.block[k] = p->local_size[k];
    KzComputeFrame &f = op.cframe;
    f.base[0] = bx;
    f.base[1] = by;
    f.base[2] = bz;
    std::memcpy(f.push, st.push, sizeof f.push);
    auto entry_at = [&](uint32_t set, uint32_t binding) -> const DescriptorEntry * {
        if (set >= 8 || !st.compute_sets[set]) return nullptr;
        auto it = st.compute_sets[set]->entries.find(binding);
        if (it == st.compute_sets[set]->entries.end() || !it->second.buffer) return nullptr;
        return &it->second;
    };
    auto buffer_at = [&](uint32_t set, uint32_t binding) -> uint8_t * {
        const DescriptorEntry *e = entry_at(set, binding);
        return e ? e->buffer->address() + e->offset : nullptr;
    };
    for (const auto &kv : p->ubo_slot) f.ubo[kv.second] = buffer_at(kv.first.first, kv.first.second);
    for (const auto &kv : p->ssbo_slot) {
        f.ssbo[kv.second] = buffer_at(kv.first.first, kv.first.second);
        // the bound range (VK_WHOLE_SIZE: to the end of the buffer), for OpArrayLength
        const DescriptorEntry *e = entry_at(kv.first.first, kv.first.second);
        uint64_t bytes = 0;
        if (e) {
            const uint64_t size = e->buffer->size, off = e->offset;
            bytes = off >= size ? 0 : (e->range == VK_WHOLE_SIZE ? size - off : std::min<uint64_t>(e->range, size - off));
        }
        f.ssbo_bytes[kv.second] = (uint32_t)std::min<uint64_t>(bytes, 0x7FFFFFFFu);
    }
    cb->ops.push_back(op);
    st.batch_open = false;
}

VKAPI_ATTR void VKAPI_CALL vkCmdDispatch(VkCommandBuffer commandBuffer, uint32_t x, uint32_t y, uint32_t z) {
    record_dispatch(from<CommandBuffer>(commandBuffer), 0, 0, 0, x, y, z);
}

VKAPI_ATTR void VKAPI_CALL vkCmdDispatchBase(VkCommandBuffer commandBuffer, uint32_t bx, uint32_t by, uint32_t bz,
                                             uint32_t x, uint32_t y, uint32_t z) {
    record_dispatch(from<CommandBuffer>(commandBuffer), bx, by, bz, x, y, z);
}

VKAPI_ATTR void VKAPI_CALL vkCmdPushConstants(VkCommandBuffer commandBuffer, VkPipelineLayout, VkShaderStageFlags,
                                              uint32_t offset, uint32_t size, const void *values) {
    CommandBuffer *cb = from<CommandBuffer>(commandBuffer);
    if (offset + size <= KZ_PUSH_WORDS * 4) std::memcpy((uint8_t *)cb->st.push + offset, values, size);
}

VKAPI_ATTR void VKAPI_CALL vkCmdBeginRenderPass(VkCommandBuffer commandBuffer, const VkRenderPassBeginInfo *bi,
                                                VkSubpassContents) {
    CommandBuffer *cb = from<CommandBuffer>(commandBuffer);
    RecordState &st = cb->st;
    st.render_pass = from<RenderPass>(bi->renderPass);
    st.framebuffer = from<Framebuffer>(bi->framebuffer);
    st.render_area = bi->renderArea;
    st.subpass = 0;
    st.batch_open = false;
    const VkRect2D &a = bi->renderArea;
    for (uint32_t i = 0; i < st.render_pass->attachments.size(); ++i) {
        const VkAttachmentDescription &ad = st.render_pass->attachments[i];
        if (ad.loadOp != VK_ATTACHMENT_LOAD_OP_CLEAR || i >= bi->clearValueCount) continue;
        Image *im = st.framebuffer->views[i]->image;
        Op op;
        op.kind = Op::Fill;
        op.fill_ptr = (uint32_t *)im->address();
        op.fill_width = im->width;
        op.fill_spp = im->samples;
        op.x0 = a.offset.x;
        op.y0 = a.offset.y;
        op.w = std::min<uint32_t>(a.extent.width, im->width - (uint32_t)a.offset.x);
        op.h = std::min<uint32_t>(a.extent.height, im->height - (uint32_t)a.offset.y);
        const VkClearValue &cv = bi->pClearValues[i];
        if (ad.format == VK_FORMAT_D32_SFLOAT) {
            std::memcpy(&op.fill_value, &cv.depthStencil.depth, 4);
        } else {
            uint32_t c[4];
            for (int k = 0; k < 4; ++k) c[k] = encode_unorm8(cv.color.float32[k]);
            if (ad.format == VK_FORMAT_B8G8R8A8_UNORM) std::swap(c[0], c[2]);
            op.fill_value = c[0] | (c[1] << 8) | (c[2] << 16) | (c[3] << 24);
        }
        cb->ops.push_back(op);
    }
}

VKAPI_ATTR void VKAPI_CALL vkCmdEndRenderPass(VkCommandBuffer commandBuffer) {
    CommandBuffer *cb = from<CommandBuffer>(commandBuffer);
    RecordState &st = cb->st;
    if (!st.render_pass) return;
    const SubpassInfo &sp = st.render_pass->subpasses[st.subpass];
    for (size_t k = 0; k < sp.color.size(); ++k) {
        if (sp.resolve[k] == VK_ATTACHMENT_UNUSED) continue;
        Image *src = st.framebuffer->views[sp.color[k]]->image;
        Image *dst = st.framebuffer->views[sp.resolve[k]]->image;
        Op op;
        op.kind = Op::Resolve;
        op.rsrc = (const uint32_t *)src->address();
        op.rdst = (uint32_t *)dst->address();
        op.rwidth = dst->width;
        op.x0 = st.render_area.offset.x;
        op.y0 = st.render_area.offset.y;
        op.w = std::min<uint32_t>(st.render_area.extent.width, dst->width - (uint32_t)op.x0);
        op.h = std::min<uint32_t>(st.render_area.extent.height, dst->height - (uint32_t)op.y0);
        cb->ops.push_back(op);
    }
    st.render_pass = nullptr;
    st.framebuffer = nullptr;
    st.batch_open = false;
}

static bool same_textures(const KzFrame &a, const KzFrame &b) {
    return std::memcmp(a.tex, b.tex, sizeof a.tex) == 0 && std::memcmp(a.ubo, b.ubo, sizeof a.ubo) == 0;
}

static void record_draw(CommandBuffer *cb, uint32_t count, uint32_t instances, uint32_t first, int32_t vertex_offset,
                        uint32_t first_instance, bool indexed) {
    RecordState &st = cb->st;
    Pipeline *p = st.pipeline;
    if (!p || !st.render_pass || count < 3 || instances == 0) return;
    const SubpassInfo &sp = st.render_pass->subpasses[st.subpass];
    if (sp.color.empty() || sp.depth == VK_ATTACHMENT_UNUSED) return;  // colour + depth attachments required
    Image *color = st.framebuffer->views[sp.color[0]]->image;
    Image *depth = st.framebuffer->views[sp.depth]->image;

    // the batch's frame description
    KzFrame f{};
    f.color = color->address();
    f.depth = (float *)depth->address();
    f.width = st.framebuffer->width;
    f.height = st.framebuffer->height;
    const uint32_t ts = 1u << p->fixed.tile_shift;
    f.tiles_x = (f.width + ts - 1) / ts;
    f.tiles_y = (f.height + ts - 1) / ts;
    const FixedState &fx = p->fixed;
    f.pipe_info = KZ_PIPE_PACK(fx.samples == 4 ? 2u : 0u, fx.blend ? 1u : 0u, fx.depth_test ? 1u : 0u,
                               std::min<uint32_t>(fx.num_varyings, 31u),
                               std::min<uint32_t>((uint32_t)p->tex_slot.size(), 7u),
                               std::min<uint32_t>((uint32_t)p->ubo_slot.size(), 7u));
    const VkViewport &v = st.viewport;
    f.vp_hw = v.width * 0.5f;
    f.vp_hh = v.height * 0.5f;
    f.vp_ox = v.x + f.vp_hw;
    f.vp_oy = v.y + f.vp_hh;
    f.vp_dz = v.maxDepth - v.minDepth;
    f.vp_zmin = v.minDepth;
    f.z_lo = std::fmin(v.minDepth, v.maxDepth);
    f.z_hi = std::fmax(v.minDepth, v.maxDepth);
    const VkRect2D &s = st.scissor, &a = st.render_area;
    f.sc_x0 = std::max({s.offset.x, a.offset.x, 0});
    f.sc_y0 = std::max({s.offset.y, a.offset.y, 0});
    f.sc_x1 = (int32_t)std::min<int64_t>({(int64_t)s.offset.x + s.extent.width - 1,
                                          (int64_t)a.offset.x + a.extent.width - 1, (int64_t)f.width - 1});
    f.sc_y1 = (int32_t)std::min<int64_t>({(int64_t)s.offset.y + s.extent.height - 1,
                                          (int64_t)a.offset.y + a.extent.height - 1, (int64_t)f.height - 1});
    for (const auto &kv : p->ubo_slot) {
        DescriptorSet *ds = kv.first.first < 8 ? st.sets[kv.first.first] : nullptr;
        if (!ds) continue;
        auto it = ds->entries.find(kv.first.second);
        if (it != ds->entries.end() && it->second.buffer)
            f.ubo[kv.second] = it->second.buffer->address() + it->second.offset;
    }
    for (const auto &kv : p->tex_slot) {
        DescriptorSet *ds = kv.first.first < 8 ? st.sets[kv.first.first] : nullptr;
        if (!ds) continue;
        auto it = ds->entries.find(kv.first.second);
        if (it == ds->entries.end() || !it->second.view) continue;
        Image *im = it->second.view->image;
        KzTexture &t = f.tex[kv.second];
        t.data = im->address();
        t.w = im->width;
        t.h = im->height;
        t.filter = it->second.sampler ? it->second.sampler->mag_filter : 0;
        t.addr_u = it->second.sampler ? it->second.sampler->addr_u : 2;
        t.addr_v = it->second.sampler ? it->second.sampler->addr_v : 2;
    }

    Op *batch = nullptr;
    if (st.batch_open && cb->device->engine->merge_draws() && !cb->ops.empty() && cb->ops.back().kind == Op::Batch) {
        Op &last = cb->ops.back();
        if (last.variants == &p->variants && std::memcmp(&last.frame.vp_hw, &f.vp_hw, sizeof(float) * 8) == 0 &&
            last.frame.sc_x0 == f.sc_x0 && last.frame.sc_y0 == f.sc_y0 && last.frame.sc_x1 == f.sc_x1 &&
            last.frame.sc_y1 == f.sc_y1 && last.frame.color == f.color && last.frame.depth == f.depth &&
            same_textures(last.frame, f))
            batch = &last;
    }
    if (!batch) {
        // The render pass's clears of this batch's colour and depth, recorded
        // just before it over the whole framebuffer, fold into the batch: its
        // raster starts every tile from the clear values instead of a fill
        // kernel and a load.
        while (!cb->ops.empty() && cb->ops.back().kind == Op::Fill) {
            const Op &o = cb->ops.back();
            const bool whole = o.x0 == 0 && o.y0 == 0 && o.w == f.width && o.h == f.height &&
                               o.fill_width == f.width && o.fill_spp == fx.samples;
            const bool col = whole && o.fill_ptr == (const uint32_t *)f.color && !(f.clear_mask & 1u);
            const bool dep = whole && o.fill_ptr == (const uint32_t *)f.depth && !(f.clear_mask & 2u);
            if (col) {
                f.clear_mask |= 1u;
                f.clear_color = o.fill_value;
            } else if (dep) {
                f.clear_mask |= 2u;
                std::memcpy(&f.clear_depth, &o.fill_value, 4);
            } else {
                break;
            }
            cb->ops.pop_back();
        }
        Op op;
        op.kind = Op::Batch;
        op.variants = &p->variants;
        op.tile_shift = p->fixed.tile_shift;
        op.frame = f;
        cb->ops.push_back(std::move(op));
        batch = &cb->ops.back();
        st.batch_open = true;
    }
    const uint32_t tris = count / 3;
    for (uint32_t inst = 0; inst < instances; ++inst) {
        KzDraw d{};
        for (int b = 0; b < KZ_MAX_BINDINGS; ++b)
            d.vb[b] = st.vb[b] ? st.vb[b]->address() + st.vb_offset[b] : nullptr;
        if (indexed) {
            d.ib = st.ib ? st.ib->address() + st.ib_offset : nullptr;
            d.index_type = st.ib_type == VK_INDEX_TYPE_UINT16 ? 0u : 1u;
        } else {
            d.index_type = 2u;
        }
        d.first = first;
        d.vertex_offset = vertex_offset;
        d.instance = first_instance + inst;
        d.tri_count = tris;
        d.tri_begin = batch->frame.total_tris;
        std::memcpy(d.push, st.push, sizeof d.push);
        batch->draws.push_back(d);
        batch->frame.total_tris += tris;
    }
}

VKAPI_ATTR void VKAPI_CALL vkCmdDraw(VkCommandBuffer commandBuffer, uint32_t vertexCount, uint32_t instanceCount,
                                     uint32_t firstVertex, uint32_t firstInstance) {
    record_draw(from<CommandBuffer>(commandBuffer), vertexCount, instanceCount, firstVertex, 0, firstInstance, false);
}

VKAPI_ATTR void VKAPI_CALL vkCmdDrawIndexed(VkCommandBuffer commandBuffer, uint32_t indexCount,
                                            uint32_t instanceCount, uint32_t firstIndex, int32_t vertexOffset,
                                            uint32_t firstInstance) {
    record_draw(from<CommandBuffer>(commandBuffer), indexCount, instanceCount, firstIndex, vertexOffset,
                firstInstance, true);
}

VKAPI_ATTR void VKAPI_CALL vkCmdCopyBuffer(VkCommandBuffer commandBuffer, VkBuffer src, VkBuffer dst, uint32_t n,
                                           const VkBufferCopy *regions) {
    CommandBuffer *cb = from<CommandBuffer>(commandBuffer);
    for (uint32_t i = 0; i < n; ++i) {
        Op op;
        op.kind = Op::Copy;
        op.src = from<Buffer>(src)->address() + regions[i].srcOffset;
        op.dst = from<Buffer>(dst)->address() + regions[i].dstOffset;
        op.bytes = regions[i].size;
        cb->ops.push_back(op);
    }
    cb->st.batch_open = false;
}

static void image_buffer_copy(CommandBuffer *cb, Image *im, Buffer *buf, const VkBufferImageCopy &r, bool to_image) {
    const uint32_t tb = im->texel_bytes() * im->samples;
    const uint32_t row_len = r.bufferRowLength ? r.bufferRowLength : r.imageExtent.width;
    Op op;
    op.kind = Op::Copy2D;
    uint8_t *ip = im->address() + ((size_t)r.imageOffset.y * im->width + r.imageOffset.x) * tb;
    uint8_t *bp = buf->address() + r.bufferOffset;
    op.row_bytes = (size_t)r.imageExtent.width * tb;
    op.rows = r.imageExtent.height;
    if (to_image) {
        op.src = bp;
        op.src_pitch = (size_t)row_len * tb;
        op.dst = ip;
        op.dst_pitch = (size_t)im->width * tb;
    } else {
        op.src = ip;
        op.src_pitch = (size_t)im->width * tb;
        op.dst = bp;
        op.dst_pitch = (size_t)row_len * tb;
    }
    cb->ops.push_back(op);
    cb->st.batch_open = false;
}

VKAPI_ATTR void VKAPI_CALL vkCmdCopyBufferToImage(VkCommandBuffer commandBuffer, VkBuffer src, VkImage dst,
                                                  VkImageLayout, uint32_t n, const VkBufferImageCopy *regions) {
    for (uint32_t i = 0; i < n; ++i)
        image_buffer_copy(from<CommandBuffer>(commandBuffer), from<Image>(dst), from<Buffer>(src), regions[i], true);
}

VKAPI_ATTR void VKAPI_CALL vkCmdCopyImageToBuffer(VkCommandBuffer commandBuffer, VkImage src, VkImageLayout,
                                                  VkBuffer dst, uint32_t n, const VkBufferImageCopy *regions) {
    for (uint32_t i = 0; i < n; ++i)
        image_buffer_copy(from<CommandBuffer>(commandBuffer), from<Image>(src), from<Buffer>(dst), regions[i], false);
}

VKAPI_ATTR void VKAPI_CALL vkCmdPipelineBarrier(VkCommandBuffer, VkPipelineStageFlags, VkPipelineStageFlags,
                                                VkDependencyFlags, uint32_t, const VkMemoryBarrier *, uint32_t,
                                                const VkBufferMemoryBarrier *, uint32_t, const VkImageMemoryBarrier *) {
    // the queue is one in-order HIP stream: every command already waits for
    // the previous one
}

// ===========================================================================
// submission and synchronisation

VKAPI_ATTR VkResult VKAPI_CALL vkCreateFence(VkDevice device, const VkFenceCreateInfo *ci,
                                             const VkAllocationCallbacks *, VkFence *pFence) {
    Device *d = from<Device>(device);
    set_device(d);
    auto *f = new (std::nothrow) Fence();
    if (!f) return VK_ERROR_OUT_OF_HOST_MEMORY;
    f->device = d;
    if (hipEventCreateWithFlags(&f->event, hipEventDisableTiming) != hipSuccess) {
        delete f;
        return VK_ERROR_OUT_OF_DEVICE_MEMORY;
    }
    f->signaled = (ci->flags & VK_FENCE_CREATE_SIGNALED_BIT) != 0;
    *pFence = to<VkFence>(f);
    return VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkDestroyFence(VkDevice, VkFence fence, const VkAllocationCallbacks *) {
    Fence *f = from<Fence>(fence);
    if (!f) return;
    set_device(f->device);
    if (f->submitted) hipEventSynchronize(f->event);
    hipEventDestroy(f->event);
    delete f;
}

VKAPI_ATTR VkResult VKAPI_CALL vkResetFences(VkDevice, uint32_t n, const VkFence *fences) {
    for (uint32_t i = 0; i < n; ++i) {
        Fence *f = from<Fence>(fences[i]);
        f->signaled = false;
        f->submitted = false;
    }
    return VK_SUCCESS;
}

VKAPI_ATTR VkResult VKAPI_CALL vkGetFenceStatus(VkDevice, VkFence fence) {
    Fence *f = from<Fence>(fence);
    if (f->signaled) return VK_SUCCESS;
    if (!f->submitted) return VK_NOT_READY;
    hipError_t e = hipEventQuery(f->event);
    if (e == hipSuccess) {
        f->signaled = true;
        return VK_SUCCESS;
    }
    return e == hipErrorNotReady ? VK_NOT_READY : VK_ERROR_DEVICE_LOST;
}

VKAPI_ATTR VkResult VKAPI_CALL vkWaitForFences(VkDevice device, uint32_t n, const VkFence *fences, VkBool32 waitAll,
                                               uint64_t timeout) {
    Device *d = from<Device>(device);
    set_device(d);
    // 1 signaled, 0 not yet, -1 device error
    auto poll = [](Fence *f) {
        if (f->signaled) return 1;
        if (!f->submitted) return 0;
        const hipError_t e = hipEventQuery(f->event);
        if (e == hipSuccess) {
            f->signaled = true;
            return 1;
        }
        return e == hipErrorNotReady ? 0 : -1;
    };
    const auto t0 = std::chrono::steady_clock::now();
    // timeouts beyond ~292 years are infinite (and would overflow the clock)
    const bool infinite = timeout >= (uint64_t)INT64_MAX;
    for (;;) {
        uint32_t done = 0, unsubmitted = 0;
        for (uint32_t i = 0; i < n; ++i) {
            Fence *f = from<Fence>(fences[i]);
            const int r = poll(f);
            if (r < 0) return VK_ERROR_DEVICE_LOST;
            done += (uint32_t)r;
            unsubmitted += (r == 0 && !f->submitted);
        }
        if (waitAll ? done == n : (done > 0 || n == 0)) break;
        if (timeout == 0) return VK_TIMEOUT;
        const uint64_t elapsed = (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
                                     std::chrono::steady_clock::now() - t0).count();
        if (!infinite && elapsed >= timeout) return VK_TIMEOUT;
        if (waitAll && unsubmitted == 0 && infinite) {
            // every fence is on its way: block on each instead of polling
            for (uint32_t i = 0; i < n; ++i) {
                Fence *f = from<Fence>(fences[i]);
                if (f->signaled) continue;
                if (hipEventSynchronize(f->event) != hipSuccess) return VK_ERROR_DEVICE_LOST;
                f->signaled = true;
            }
            continue;
        }
        // a fence another thread has yet to submit, a partial wait or a
        // finite timeout: poll
        std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
    // the statistics fold waits for the whole queue: only when that costs
    // nothing more than the wait itself
    if (waitAll || hipStreamQuery(d->engine->stream()) == hipSuccess) d->engine->collect_stats();
    return VK_SUCCESS;
}

VKAPI_ATTR VkResult VKAPI_CALL vkCreateSemaphore(VkDevice, const VkSemaphoreCreateInfo *,
                                                 const VkAllocationCallbacks *, VkSemaphore *pSem) {
    *pSem = to<VkSemaphore>(new (std::nothrow) Semaphore());
    return VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkDestroySemaphore(VkDevice, VkSemaphore sem, const VkAllocationCallbacks *) {
    delete from<Semaphore>(sem);
}

VKAPI_ATTR VkResult VKAPI_CALL vkQueueSubmit(VkQueue queue, uint32_t submitCount, const VkSubmitInfo *submits,
                                             VkFence fence) {
    Device *d = from<Queue>(queue)->device;
    set_device(d);
    Engine &eng = *d->engine;
    if (eng.begin_submit() != hipSuccess) return VK_ERROR_DEVICE_LOST;
    for (uint32_t s = 0; s < submitCount; ++s)
        for (uint32_t i = 0; i < submits[s].commandBufferCount; ++i) {
            CommandBuffer *cb = from<CommandBuffer>(submits[s].pCommandBuffers[i]);
            hipError_t e = eng.execute(cb->ops, &cb->graph);
            if (e != hipSuccess) {
                std::fprintf(stderr, "kazan: vkQueueSubmit: %s\n", hipGetErrorString(e));
                return VK_ERROR_DEVICE_LOST;
            }
        }
    Fence *f = from<Fence>(fence);
    if (eng.end_submit(f ? f->event : nullptr) != hipSuccess) return VK_ERROR_DEVICE_LOST;
    if (f) {
        f->submitted = true;
        f->signaled = false;
    }
    return VK_SUCCESS;
}

// ===========================================================================
// proc addresses

VKAPI_ATTR PFN_vkVoidFunction VKAPI_CALL vkGetDeviceProcAddr(VkDevice, const char *pName) {
    return lookup(pName, true);
}

VKAPI_ATTR PFN_vkVoidFunction VKAPI_CALL vkGetInstanceProcAddr(VkInstance, const char *pName) {
    return lookup(pName, false);
}

VKAPI_ATTR PFN_vkVoidFunction VKAPI_CALL vk_icdGetInstanceProcAddr(VkInstance instance, const char *pName) {
    return vkGetInstanceProcAddr(instance, pName);
}

VKAPI_ATTR VkResult VKAPI_CALL vk_icdNegotiateLoaderICDInterfaceVersion(uint32_t *pSupportedVersion) {
    if (*pSupportedVersion > KAZAN_ICD_VERSION) *pSupportedVersion = KAZAN_ICD_VERSION;
    return VK_SUCCESS;
}

VKAPI_ATTR PFN_vkVoidFunction VKAPI_CALL vk_icdGetPhysicalDeviceProcAddr(VkInstance, const char *pName) {
    if (!pName) return nullptr;
    auto &t = entries();
    auto it = t.find(pName);
    if (it == t.end() || !it->second.physical) return nullptr;
    return it->second.fn;
}

}  // extern "C"

namespace {

std::unordered_map<std::string, Entry> &entries() {
    static std::unordered_map<std::string, Entry> t = [] {
        std::unordered_map<std::string, Entry> m;
#define KZ_I(name) m[#name] = Entry{(PFN_vkVoidFunction)&name, false, false}
#define KZ_P(name) m[#name] = Entry{(PFN_vkVoidFunction)&name, true, false}
#define KZ_D(name) m[#name] = Entry{(PFN_vkVoidFunction)&name, false, true}
        KZ_I(vkCreateInstance);
        KZ_I(vkDestroyInstance);
        KZ_I(vkEnumerateInstanceExtensionProperties);
        KZ_I(vkEnumerateInstanceLayerProperties);
        KZ_I(vkEnumerateInstanceVersion);
        KZ_I(vkEnumeratePhysicalDevices);
        KZ_I(vkGetInstanceProcAddr);
        KZ_P(vkGetPhysicalDeviceProperties);
        KZ_P(vkGetPhysicalDeviceFeatures);
        KZ_P(vkGetPhysicalDeviceMemoryProperties);
        KZ_P(vkGetPhysicalDeviceQueueFamilyProperties);
        KZ_P(vkGetPhysicalDeviceFormatProperties);
        KZ_P(vkEnumerateDeviceExtensionProperties);
        KZ_P(vkEnumerateDeviceLayerProperties);
        KZ_P(vkCreateDevice);
        KZ_D(vkGetDeviceProcAddr);
        KZ_D(vkDestroyDevice);
        KZ_D(vkGetDeviceQueue);
        KZ_D(vkDeviceWaitIdle);
        KZ_D(vkQueueWaitIdle);
        KZ_D(vkQueueSubmit);
        KZ_D(vkAllocateMemory);
        KZ_D(vkFreeMemory);
        KZ_D(vkMapMemory);
        KZ_D(vkUnmapMemory);
        KZ_D(vkFlushMappedMemoryRanges);
        KZ_D(vkInvalidateMappedMemoryRanges);
        KZ_D(vkCreateBuffer);
        KZ_D(vkDestroyBuffer);
        KZ_D(vkGetBufferMemoryRequirements);
        KZ_D(vkBindBufferMemory);
        KZ_D(vkCreateImage);
        KZ_D(vkDestroyImage);
        KZ_D(vkGetImageMemoryRequirements);
        KZ_D(vkBindImageMemory);
        KZ_D(vkCreateImageView);
        KZ_D(vkDestroyImageView);
        KZ_D(vkCreateSampler);
        KZ_D(vkDestroySampler);
        KZ_D(vkCreateShaderModule);
        KZ_D(vkDestroyShaderModule);
        KZ_D(vkCreatePipelineCache);
        KZ_D(vkDestroyPipelineCache);
        KZ_D(vkGetPipelineCacheData);
        KZ_D(vkMergePipelineCaches);
        KZ_D(vkCreateDescriptorSetLayout);
        KZ_D(vkDestroyDescriptorSetLayout);
        KZ_D(vkCreateDescriptorPool);
        KZ_D(vkDestroyDescriptorPool);
        KZ_D(vkResetDescriptorPool);
        KZ_D(vkAllocateDescriptorSets);
        KZ_D(vkFreeDescriptorSets);
        KZ_D(vkUpdateDescriptorSets);
        KZ_D(vkCreatePipelineLayout);
        KZ_D(vkDestroyPipelineLayout);
        KZ_D(vkCreateRenderPass);
        KZ_D(vkDestroyRenderPass);
        KZ_D(vkCreateFramebuffer);
        KZ_D(vkDestroyFramebuffer);
        KZ_D(vkCreateGraphicsPipelines);
        KZ_D(vkCreateComputePipelines);
        KZ_D(vkDestroyPipeline);
        KZ_D(vkCreateCommandPool);
        KZ_D(vkDestroyCommandPool);
        KZ_D(vkResetCommandPool);
        KZ_D(vkAllocateCommandBuffers);
        KZ_D(vkFreeCommandBuffers);
        KZ_D(vkBeginCommandBuffer);
        KZ_D(vkEndCommandBuffer);
        KZ_D(vkResetCommandBuffer);
        KZ_D(vkCmdBindPipeline);
        KZ_D(vkCmdSetViewport);
        KZ_D(vkCmdSetScissor);
        KZ_D(vkCmdBindVertexBuffers);
        KZ_D(vkCmdBindIndexBuffer);
        KZ_D(vkCmdBindDescriptorSets);
        KZ_D(vkCmdPushConstants);
        KZ_D(vkCmdBeginRenderPass);
        KZ_D(vkCmdEndRenderPass);
        KZ_D(vkCmdDraw);
        KZ_D(vkCmdDrawIndexed);
        KZ_D(vkCmdDispatch);
        KZ_D(vkCmdDispatchBase);
        KZ_D(vkCmdCopyBuffer);
        KZ_D(vkCmdCopyBufferToImage);
        KZ_D(vkCmdCopyImageToBuffer);
        KZ_D(vkCmdPipelineBarrier);
        KZ_D(vkCreateFence);
        KZ_D(vkDestroyFence);
        KZ_D(vkResetFences);
        KZ_D(vkGetFenceStatus);
        KZ_D(vkWaitForFences);
        KZ_D(vkCreateSemaphore);
        KZ_D(vkDestroySemaphore);
        KZ_D(vkGetDrawStatsKAZAN);
        KZ_D(vkSetKernelTimingKAZAN);
        KZ_D(vkGetKernelTimesKAZAN);
        KZ_D(vkSetTileProfileKAZAN);
        KZ_D(vkGetTileProfileKAZAN);
        KZ_D(vkSetCostModelKAZAN);
        KZ_D(vkHasCostModelKAZAN);
        KZ_D(vkFillClusterSceneKAZAN);
        KZ_D(vkGetGraphStatsKAZAN);
        KZ_D(vkGetTuningKAZAN);
#undef KZ_I
#undef KZ_P
#undef KZ_D
        return m;
    }();
    return t;
}

}  // namespace
