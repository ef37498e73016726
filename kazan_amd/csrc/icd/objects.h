// objects.h — the driver's Vulkan objects.
//
// Counterparts of the reference's object model (vulkan-driver/src/handle.rs,
// device_memory.rs, buffer.rs, image.rs, pipeline.rs, descriptor_set.rs,
// sampler.rs, render_pass.rs) with device memory in HBM instead of host heap:
// memory type 0 is DEVICE_LOCAL (hipMalloc), type 1 is
// HOST_VISIBLE | HOST_COHERENT | HOST_CACHED pinned host memory
// (hipHostMalloc) that the GPU reads directly.
#pragma once

#include <hip/hip_runtime_api.h>

#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../engine/engine.h"
#include "kazan/vk_abi.h"

namespace kz {

struct Dispatchable {
    uintptr_t loader_magic = ICD_LOADER_MAGIC;
};

struct PhysicalDevice;
struct Device;
struct Queue;

struct Instance : Dispatchable {
    std::vector<std::unique_ptr<PhysicalDevice>> physical_devices;
};

struct PhysicalDevice : Dispatchable {
    Instance *instance = nullptr;
    int hip_device = 0;
    VkPhysicalDeviceProperties props{};
    VkPhysicalDeviceMemoryProperties memory{};
    VkPhysicalDeviceFeatures features{};
};

struct Queue : Dispatchable {
    Device *device = nullptr;
};

struct Device : Dispatchable {
    PhysicalDevice *physical = nullptr;
    std::unique_ptr<Queue> queue;
    std::unique_ptr<Engine> engine;
};

struct DeviceMemory {
    uint8_t *ptr = nullptr;
    VkDeviceSize size = 0;
    uint32_t type = 0;
    bool host = false;
};

struct Buffer {
    VkDeviceSize size = 0;
    DeviceMemory *memory = nullptr;
    VkDeviceSize offset = 0;
    uint8_t *address() const { return memory ? memory->ptr + offset : nullptr; }
};

struct Image {
    VkFormat format = VK_FORMAT_UNDEFINED;
    uint32_t width = 0, height = 0, samples = 1;
    VkDeviceSize size = 0;
    DeviceMemory *memory = nullptr;
    VkDeviceSize offset = 0;
    uint8_t *address() const { return memory ? memory->ptr + offset : nullptr; }
    uint32_t texel_bytes() const { return 4; }
};

struct ImageView {
    Image *image = nullptr;
    VkFormat format = VK_FORMAT_UNDEFINED;
};

struct Sampler {
    uint32_t mag_filter = 0, min_filter = 0, addr_u = 0, addr_v = 0;
};

struct ShaderModule {
    std::vector<uint32_t> code;
};

struct DescriptorBinding {
    uint32_t binding = 0;
    VkDescriptorType type = VK_DESCRIPTOR_TYPE_MAX_ENUM;
    uint32_t count = 1;
    std::vector<Sampler *> immutable_samplers;
};

struct DescriptorSetLayout {
    std::vector<DescriptorBinding> bindings;
};

struct DescriptorEntry {
    VkDescriptorType type = VK_DESCRIPTOR_TYPE_MAX_ENUM;
    Buffer *buffer = nullptr;
    VkDeviceSize offset = 0, range = 0;
    ImageView *view = nullptr;
    Sampler *sampler = nullptr;
};

struct DescriptorSet {
    DescriptorSetLayout *layout = nullptr;
    std::map<uint32_t, DescriptorEntry> entries;
};

struct DescriptorPool {
    std::vector<std::unique_ptr<DescriptorSet>> sets;
};

struct PipelineLayout {
    std::vector<DescriptorSetLayout *> set_layouts;
    uint32_t push_constant_size = 0;
};

struct SubpassInfo {
    std::vector<uint32_t> color, resolve;
    uint32_t depth = VK_ATTACHMENT_UNUSED;
};

struct RenderPass {
    std::vector<VkAttachmentDescription> attachments;
    std::vector<SubpassInfo> subpasses;
};

struct Framebuffer {
    RenderPass *render_pass = nullptr;
    std::vector<ImageView *> views;
    uint32_t width = 0, height = 0;
};

struct Pipeline {
    // a compute pipeline (vkCreateComputePipelines): its kernel and its
    // workgroup size; the graphics fields below are unused
    bool compute = false;
    hipModule_t cs_module = nullptr;
    hipFunction_t cs_kernel = nullptr;
    uint32_t local_size[3] = {1, 1, 1};
    std::map<std::pair<uint32_t, uint32_t>, uint32_t> ssbo_slot;
    ~Pipeline() {
        if (cs_module) hipModuleUnload(cs_module);
    }

    FixedState fixed;
    PipelineVariants variants;  // compiled kernels per tile size
    bool dynamic_viewport = false, dynamic_scissor = false;
    VkViewport viewport{};
    VkRect2D scissor{};
    PipelineLayout *layout = nullptr;
    std::map<std::pair<uint32_t, uint32_t>, uint32_t> ubo_slot, tex_slot;
    std::string shader_source;
    uint32_t subpass = 0;
};

// A VkPipelineCache: the code objects (by cache file name) of the pipeline
// variants created with it or loaded into it. Its data is a Vulkan pipeline
// cache header (32 bytes) followed by a u32 entry count and, per entry, the
// file name (32 bytes, zero-padded), a u64 size and the code object.
struct PipelineCache {
    // Vulkan lets several threads use one cache at once (unless it was
    // created externally synchronized): every access takes the lock
    std::mutex mu;
    std::vector<std::pair<std::string, std::vector<char>>> entries;
    std::vector<std::string> pending;  // variants not compiled yet when their pipeline was created
};

struct CommandPool {
    std::vector<struct CommandBuffer *> buffers;
};

struct RecordState {
    Pipeline *pipeline = nullptr;
    // the compute bind point (vkCmdBindPipeline / vkCmdBindDescriptorSets
    // with VK_PIPELINE_BIND_POINT_COMPUTE); push constants are shared
    Pipeline *compute_pipeline = nullptr;
    DescriptorSet *compute_sets[8] = {};
    VkViewport viewport{};
    VkRect2D scissor{};
    Buffer *vb[KZ_MAX_BINDINGS] = {};
    VkDeviceSize vb_offset[KZ_MAX_BINDINGS] = {};
    Buffer *ib = nullptr;
    VkDeviceSize ib_offset = 0;
    VkIndexType ib_type = VK_INDEX_TYPE_UINT32;
    DescriptorSet *sets[8] = {};
    uint32_t push[KZ_PUSH_WORDS] = {};
    RenderPass *render_pass = nullptr;
    Framebuffer *framebuffer = nullptr;
    VkRect2D render_area{};
    uint32_t subpass = 0;
    bool batch_open = false;  // the last op is a batch that may take more draws
};

struct CommandBuffer : Dispatchable {
    Device *device = nullptr;
    CommandPool *pool = nullptr;
    std::vector<Op> ops;
    std::vector<void *> device_allocs;
    GraphCache graph;  // the ops as a hipGraph, once captured
    RecordState st;
    bool recording = false;
};

struct Fence {
    hipEvent_t event = nullptr;
    bool submitted = false;
    bool signaled = false;
    Device *device = nullptr;
};

struct Semaphore {};

}  // namespace kz
