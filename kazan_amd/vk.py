"""ctypes binding of the Vulkan 1.1 subset exported by libkazan_driver.so.

This is what a Vulkan loader does with the ICD: load the shared object, call
vk_icdNegotiateLoaderICDInterfaceVersion, then resolve every entry point by
name through vk_icdGetInstanceProcAddr / vkGetDeviceProcAddr (the
reference's exported interface, vulkan-driver/src/lib.rs:37-128). Struct
layouts mirror include/kazan/vk_abi.h; tests compare every size with the C
compiler's.
"""
import ctypes
import os

from ctypes import POINTER, Structure, Union, c_char, c_char_p, c_float, c_int32, c_size_t, c_uint8, c_uint32, \
    c_uint64, c_void_p

_PKG = os.path.dirname(os.path.abspath(__file__))
DRIVER_PATH = os.path.join(_PKG, "libkazan_driver.so")

# --- enums and constants (public Vulkan 1.1 values) -----------------------
VK_SUCCESS = 0
VK_NOT_READY = 1
VK_TIMEOUT = 2
VK_INCOMPLETE = 5
VK_ERROR_OUT_OF_HOST_MEMORY = -1
VK_ERROR_OUT_OF_DEVICE_MEMORY = -2
VK_ERROR_INITIALIZATION_FAILED = -3
VK_ERROR_DEVICE_LOST = -4
VK_ERROR_MEMORY_MAP_FAILED = -5
VK_ERROR_LAYER_NOT_PRESENT = -6
VK_ERROR_EXTENSION_NOT_PRESENT = -7
VK_ERROR_FEATURE_NOT_PRESENT = -8
VK_ERROR_FORMAT_NOT_SUPPORTED = -11
VK_ERROR_INVALID_SHADER_NV = -1000012000

ST_APPLICATION_INFO = 0
ST_INSTANCE_CREATE_INFO = 1
ST_DEVICE_QUEUE_CREATE_INFO = 2
ST_DEVICE_CREATE_INFO = 3
ST_SUBMIT_INFO = 4
ST_MEMORY_ALLOCATE_INFO = 5
ST_FENCE_CREATE_INFO = 8
ST_BUFFER_CREATE_INFO = 12
ST_IMAGE_CREATE_INFO = 14
ST_IMAGE_VIEW_CREATE_INFO = 15
ST_SHADER_MODULE_CREATE_INFO = 16
ST_PIPELINE_CACHE_CREATE_INFO = 17
ST_PIPELINE_SHADER_STAGE_CREATE_INFO = 18
ST_PIPELINE_VERTEX_INPUT_STATE_CREATE_INFO = 19
ST_PIPELINE_INPUT_ASSEMBLY_STATE_CREATE_INFO = 20
ST_PIPELINE_VIEWPORT_STATE_CREATE_INFO = 22
ST_PIPELINE_RASTERIZATION_STATE_CREATE_INFO = 23
ST_PIPELINE_MULTISAMPLE_STATE_CREATE_INFO = 24
ST_PIPELINE_DEPTH_STENCIL_STATE_CREATE_INFO = 25
ST_PIPELINE_COLOR_BLEND_STATE_CREATE_INFO = 26
ST_PIPELINE_DYNAMIC_STATE_CREATE_INFO = 27
ST_GRAPHICS_PIPELINE_CREATE_INFO = 28
ST_COMPUTE_PIPELINE_CREATE_INFO = 29
ST_PIPELINE_LAYOUT_CREATE_INFO = 30
ST_SAMPLER_CREATE_INFO = 31
ST_DESCRIPTOR_SET_LAYOUT_CREATE_INFO = 32
ST_DESCRIPTOR_POOL_CREATE_INFO = 33
ST_DESCRIPTOR_SET_ALLOCATE_INFO = 34
ST_WRITE_DESCRIPTOR_SET = 35
ST_FRAMEBUFFER_CREATE_INFO = 37
ST_RENDER_PASS_CREATE_INFO = 38
ST_COMMAND_POOL_CREATE_INFO = 39
ST_COMMAND_BUFFER_ALLOCATE_INFO = 40
ST_COMMAND_BUFFER_BEGIN_INFO = 42
ST_RENDER_PASS_BEGIN_INFO = 43

FORMAT_R8G8B8A8_UNORM = 37
FORMAT_R32_SFLOAT = 100
FORMAT_R32G32_SFLOAT = 103
FORMAT_R32G32B32_SFLOAT = 106
FORMAT_R32G32B32A32_SFLOAT = 109
FORMAT_D32_SFLOAT = 126

MEMORY_PROPERTY_DEVICE_LOCAL = 0x1
MEMORY_PROPERTY_HOST_VISIBLE = 0x2
MEMORY_PROPERTY_HOST_COHERENT = 0x4

SHADER_STAGE_VERTEX = 0x1
SHADER_STAGE_FRAGMENT = 0x10
SHADER_STAGE_COMPUTE = 0x20
DESCRIPTOR_TYPE_COMBINED_IMAGE_SAMPLER = 1
DESCRIPTOR_TYPE_UNIFORM_BUFFER = 6
DESCRIPTOR_TYPE_STORAGE_BUFFER = 7
PIPELINE_BIND_POINT_GRAPHICS = 0
PIPELINE_BIND_POINT_COMPUTE = 1
ATTACHMENT_LOAD_OP_LOAD = 0
ATTACHMENT_LOAD_OP_CLEAR = 1
ATTACHMENT_UNUSED = 0xFFFFFFFF
IMAGE_LAYOUT_GENERAL = 1
INDEX_TYPE_UINT16 = 0
INDEX_TYPE_UINT32 = 1
TOPOLOGY_TRIANGLE_LIST = 3
DYNAMIC_STATE_VIEWPORT = 0
DYNAMIC_STATE_SCISSOR = 1


def VK_MAKE_VERSION(a, b, c):
    return (a << 22) | (b << 12) | c


Handle = c_void_p


class VkApplicationInfo(Structure):
    _fields_ = [("sType", c_int32), ("pNext", c_void_p), ("pApplicationName", c_char_p),
                ("applicationVersion", c_uint32), ("pEngineName", c_char_p), ("engineVersion", c_uint32),
                ("apiVersion", c_uint32)]


class VkInstanceCreateInfo(Structure):
    _fields_ = [("sType", c_int32), ("pNext", c_void_p), ("flags", c_uint32),
                ("pApplicationInfo", POINTER(VkApplicationInfo)), ("enabledLayerCount", c_uint32),
                ("ppEnabledLayerNames", POINTER(c_char_p)), ("enabledExtensionCount", c_uint32),
                ("ppEnabledExtensionNames", POINTER(c_char_p))]


_LIMIT_U32 = ["maxImageDimension1D", "maxImageDimension2D", "maxImageDimension3D", "maxImageDimensionCube",
              "maxImageArrayLayers", "maxTexelBufferElements", "maxUniformBufferRange", "maxStorageBufferRange",
              "maxPushConstantsSize", "maxMemoryAllocationCount", "maxSamplerAllocationCount"]
_LIMIT_U32B = ["maxBoundDescriptorSets", "maxPerStageDescriptorSamplers", "maxPerStageDescriptorUniformBuffers",
               "maxPerStageDescriptorStorageBuffers", "maxPerStageDescriptorSampledImages",
               "maxPerStageDescriptorStorageImages", "maxPerStageDescriptorInputAttachments", "maxPerStageResources",
               "maxDescriptorSetSamplers", "maxDescriptorSetUniformBuffers", "maxDescriptorSetUniformBuffersDynamic",
               "maxDescriptorSetStorageBuffers", "maxDescriptorSetStorageBuffersDynamic",
               "maxDescriptorSetSampledImages", "maxDescriptorSetStorageImages", "maxDescriptorSetInputAttachments",
               "maxVertexInputAttributes", "maxVertexInputBindings", "maxVertexInputAttributeOffset",
               "maxVertexInputBindingStride", "maxVertexOutputComponents", "maxTessellationGenerationLevel",
               "maxTessellationPatchSize", "maxTessellationControlPerVertexInputComponents",
               "maxTessellationControlPerVertexOutputComponents", "maxTessellationControlPerPatchOutputComponents",
               "maxTessellationControlTotalOutputComponents", "maxTessellationEvaluationInputComponents",
               "maxTessellationEvaluationOutputComponents", "maxGeometryShaderInvocations",
               "maxGeometryInputComponents", "maxGeometryOutputComponents", "maxGeometryOutputVertices",
               "maxGeometryTotalOutputComponents", "maxFragmentInputComponents", "maxFragmentOutputAttachments",
               "maxFragmentDualSrcAttachments", "maxFragmentCombinedOutputResources", "maxComputeSharedMemorySize"]


class VkPhysicalDeviceLimits(Structure):
    _fields_ = ([(n, c_uint32) for n in _LIMIT_U32] +
                [("bufferImageGranularity", c_uint64), ("sparseAddressSpaceSize", c_uint64)] +
                [(n, c_uint32) for n in _LIMIT_U32B] +
                [("maxComputeWorkGroupCount", c_uint32 * 3), ("maxComputeWorkGroupInvocations", c_uint32),
                 ("maxComputeWorkGroupSize", c_uint32 * 3), ("subPixelPrecisionBits", c_uint32),
                 ("subTexelPrecisionBits", c_uint32), ("mipmapPrecisionBits", c_uint32),
                 ("maxDrawIndexedIndexValue", c_uint32), ("maxDrawIndirectCount", c_uint32),
                 ("maxSamplerLodBias", c_float), ("maxSamplerAnisotropy", c_float), ("maxViewports", c_uint32),
                 ("maxViewportDimensions", c_uint32 * 2), ("viewportBoundsRange", c_float * 2),
                 ("viewportSubPixelBits", c_uint32), ("minMemoryMapAlignment", c_size_t),
                 ("minTexelBufferOffsetAlignment", c_uint64), ("minUniformBufferOffsetAlignment", c_uint64),
                 ("minStorageBufferOffsetAlignment", c_uint64), ("minTexelOffset", c_int32),
                 ("maxTexelOffset", c_uint32), ("minTexelGatherOffset", c_int32), ("maxTexelGatherOffset", c_uint32),
                 ("minInterpolationOffset", c_float), ("maxInterpolationOffset", c_float),
                 ("subPixelInterpolationOffsetBits", c_uint32), ("maxFramebufferWidth", c_uint32),
                 ("maxFramebufferHeight", c_uint32), ("maxFramebufferLayers", c_uint32),
                 ("framebufferColorSampleCounts", c_uint32), ("framebufferDepthSampleCounts", c_uint32),
                 ("framebufferStencilSampleCounts", c_uint32), ("framebufferNoAttachmentsSampleCounts", c_uint32),
                 ("maxColorAttachments", c_uint32), ("sampledImageColorSampleCounts", c_uint32),
                 ("sampledImageIntegerSampleCounts", c_uint32), ("sampledImageDepthSampleCounts", c_uint32),
                 ("sampledImageStencilSampleCounts", c_uint32), ("storageImageSampleCounts", c_uint32),
                 ("maxSampleMaskWords", c_uint32), ("timestampComputeAndGraphics", c_uint32),
                 ("timestampPeriod", c_float), ("maxClipDistances", c_uint32), ("maxCullDistances", c_uint32),
                 ("maxCombinedClipAndCullDistances", c_uint32), ("discreteQueuePriorities", c_uint32),
                 ("pointSizeRange", c_float * 2), ("lineWidthRange", c_float * 2), ("pointSizeGranularity", c_float),
                 ("lineWidthGranularity", c_float), ("strictLines", c_uint32), ("standardSampleLocations", c_uint32),
                 ("optimalBufferCopyOffsetAlignment", c_uint64), ("optimalBufferCopyRowPitchAlignment", c_uint64),
                 ("nonCoherentAtomSize", c_uint64)])


class VkPhysicalDeviceProperties(Structure):
    _fields_ = [("apiVersion", c_uint32), ("driverVersion", c_uint32), ("vendorID", c_uint32),
                ("deviceID", c_uint32), ("deviceType", c_int32), ("deviceName", c_char * 256),
                ("pipelineCacheUUID", c_uint8 * 16), ("limits", VkPhysicalDeviceLimits),
                ("sparseProperties", c_uint32 * 5)]


class VkMemoryType(Structure):
    _fields_ = [("propertyFlags", c_uint32), ("heapIndex", c_uint32)]


class VkMemoryHeap(Structure):
    _fields_ = [("size", c_uint64), ("flags", c_uint32)]


class VkPhysicalDeviceMemoryProperties(Structure):
    _fields_ = [("memoryTypeCount", c_uint32), ("memoryTypes", VkMemoryType * 32), ("memoryHeapCount", c_uint32),
                ("memoryHeaps", VkMemoryHeap * 16)]


class VkDeviceQueueCreateInfo(Structure):
    _fields_ = [("sType", c_int32), ("pNext", c_void_p), ("flags", c_uint32), ("queueFamilyIndex", c_uint32),
                ("queueCount", c_uint32), ("pQueuePriorities", POINTER(c_float))]


class VkDeviceCreateInfo(Structure):
    _fields_ = [("sType", c_int32), ("pNext", c_void_p), ("flags", c_uint32), ("queueCreateInfoCount", c_uint32),
                ("pQueueCreateInfos", POINTER(VkDeviceQueueCreateInfo)), ("enabledLayerCount", c_uint32),
                ("ppEnabledLayerNames", POINTER(c_char_p)), ("enabledExtensionCount", c_uint32),
                ("ppEnabledExtensionNames", POINTER(c_char_p)), ("pEnabledFeatures", c_void_p)]


class VkMemoryAllocateInfo(Structure):
    _fields_ = [("sType", c_int32), ("pNext", c_void_p), ("allocationSize", c_uint64), ("memoryTypeIndex", c_uint32)]


class VkMemoryRequirements(Structure):
    _fields_ = [("size", c_uint64), ("alignment", c_uint64), ("memoryTypeBits", c_uint32)]


class VkBufferCreateInfo(Structure):
    _fields_ = [("sType", c_int32), ("pNext", c_void_p), ("flags", c_uint32), ("size", c_uint64), ("usage", c_uint32),
                ("sharingMode", c_int32), ("queueFamilyIndexCount", c_uint32), ("pQueueFamilyIndices", c_void_p)]


class VkExtent2D(Structure):
    _fields_ = [("width", c_uint32), ("height", c_uint32)]


class VkExtent3D(Structure):
    _fields_ = [("width", c_uint32), ("height", c_uint32), ("depth", c_uint32)]


class VkOffset2D(Structure):
    _fields_ = [("x", c_int32), ("y", c_int32)]


class VkOffset3D(Structure):
    _fields_ = [("x", c_int32), ("y", c_int32), ("z", c_int32)]


class VkRect2D(Structure):
    _fields_ = [("offset", VkOffset2D), ("extent", VkExtent2D)]


class VkImageCreateInfo(Structure):
    _fields_ = [("sType", c_int32), ("pNext", c_void_p), ("flags", c_uint32), ("imageType", c_int32),
                ("format", c_int32), ("extent", VkExtent3D), ("mipLevels", c_uint32), ("arrayLayers", c_uint32),
                ("samples", c_uint32), ("tiling", c_int32), ("usage", c_uint32), ("sharingMode", c_int32),
                ("queueFamilyIndexCount", c_uint32), ("pQueueFamilyIndices", c_void_p), ("initialLayout", c_int32)]


class VkImageSubresourceRange(Structure):
    _fields_ = [("aspectMask", c_uint32), ("baseMipLevel", c_uint32), ("levelCount", c_uint32),
                ("baseArrayLayer", c_uint32), ("layerCount", c_uint32)]


class VkImageViewCreateInfo(Structure):
    _fields_ = [("sType", c_int32), ("pNext", c_void_p), ("flags", c_uint32), ("image", Handle),
                ("viewType", c_int32), ("format", c_int32), ("components", c_int32 * 4),
                ("subresourceRange", VkImageSubresourceRange)]


class VkSamplerCreateInfo(Structure):
    _fields_ = [("sType", c_int32), ("pNext", c_void_p), ("flags", c_uint32), ("magFilter", c_int32),
                ("minFilter", c_int32), ("mipmapMode", c_int32), ("addressModeU", c_int32),
                ("addressModeV", c_int32), ("addressModeW", c_int32), ("mipLodBias", c_float),
                ("anisotropyEnable", c_uint32), ("maxAnisotropy", c_float), ("compareEnable", c_uint32),
                ("compareOp", c_int32), ("minLod", c_float), ("maxLod", c_float), ("borderColor", c_int32),
                ("unnormalizedCoordinates", c_uint32)]


class VkShaderModuleCreateInfo(Structure):
    _fields_ = [("sType", c_int32), ("pNext", c_void_p), ("flags", c_uint32), ("codeSize", c_size_t),
                ("pCode", c_void_p)]


class VkPipelineCacheCreateInfo(Structure):
    _fields_ = [("sType", c_int32), ("pNext", c_void_p), ("flags", c_uint32), ("initialDataSize", c_size_t),
                ("pInitialData", c_void_p)]


class VkDescriptorSetLayoutBinding(Structure):
    _fields_ = [("binding", c_uint32), ("descriptorType", c_int32), ("descriptorCount", c_uint32),
                ("stageFlags", c_uint32), ("pImmutableSamplers", c_void_p)]


class VkDescriptorSetLayoutCreateInfo(Structure):
    _fields_ = [("sType", c_int32), ("pNext", c_void_p), ("flags", c_uint32), ("bindingCount", c_uint32),
                ("pBindings", POINTER(VkDescriptorSetLayoutBinding))]


class VkDescriptorPoolSize(Structure):
    _fields_ = [("type", c_int32), ("descriptorCount", c_uint32)]


class VkDescriptorPoolCreateInfo(Structure):
    _fields_ = [("sType", c_int32), ("pNext", c_void_p), ("flags", c_uint32), ("maxSets", c_uint32),
                ("poolSizeCount", c_uint32), ("pPoolSizes", POINTER(VkDescriptorPoolSize))]


class VkDescriptorSetAllocateInfo(Structure):
    _fields_ = [("sType", c_int32), ("pNext", c_void_p), ("descriptorPool", Handle), ("descriptorSetCount", c_uint32),
                ("pSetLayouts", POINTER(Handle))]


class VkDescriptorImageInfo(Structure):
    _fields_ = [("sampler", Handle), ("imageView", Handle), ("imageLayout", c_int32)]


class VkDescriptorBufferInfo(Structure):
    _fields_ = [("buffer", Handle), ("offset", c_uint64), ("range", c_uint64)]


class VkWriteDescriptorSet(Structure):
    _fields_ = [("sType", c_int32), ("pNext", c_void_p), ("dstSet", Handle), ("dstBinding", c_uint32),
                ("dstArrayElement", c_uint32), ("descriptorCount", c_uint32), ("descriptorType", c_int32),
                ("pImageInfo", POINTER(VkDescriptorImageInfo)), ("pBufferInfo", POINTER(VkDescriptorBufferInfo)),
                ("pTexelBufferView", c_void_p)]


class VkPushConstantRange(Structure):
    _fields_ = [("stageFlags", c_uint32), ("offset", c_uint32), ("size", c_uint32)]


class VkPipelineLayoutCreateInfo(Structure):
    _fields_ = [("sType", c_int32), ("pNext", c_void_p), ("flags", c_uint32), ("setLayoutCount", c_uint32),
                ("pSetLayouts", POINTER(Handle)), ("pushConstantRangeCount", c_uint32),
                ("pPushConstantRanges", POINTER(VkPushConstantRange))]


class VkPipelineShaderStageCreateInfo(Structure):
    _fields_ = [("sType", c_int32), ("pNext", c_void_p), ("flags", c_uint32), ("stage", c_uint32),
                ("module", Handle), ("pName", c_char_p), ("pSpecializationInfo", c_void_p)]


class VkVertexInputBindingDescription(Structure):
    _fields_ = [("binding", c_uint32), ("stride", c_uint32), ("inputRate", c_int32)]


class VkVertexInputAttributeDescription(Structure):
    _fields_ = [("location", c_uint32), ("binding", c_uint32), ("format", c_int32), ("offset", c_uint32)]


class VkPipelineVertexInputStateCreateInfo(Structure):
    _fields_ = [("sType", c_int32), ("pNext", c_void_p), ("flags", c_uint32),
                ("vertexBindingDescriptionCount", c_uint32),
                ("pVertexBindingDescriptions", POINTER(VkVertexInputBindingDescription)),
                ("vertexAttributeDescriptionCount", c_uint32),
                ("pVertexAttributeDescriptions", POINTER(VkVertexInputAttributeDescription))]


class VkPipelineInputAssemblyStateCreateInfo(Structure):
    _fields_ = [("sType", c_int32), ("pNext", c_void_p), ("flags", c_uint32), ("topology", c_int32),
                ("primitiveRestartEnable", c_uint32)]


class VkViewport(Structure):
    _fields_ = [("x", c_float), ("y", c_float), ("width", c_float), ("height", c_float), ("minDepth", c_float),
                ("maxDepth", c_float)]


class VkPipelineViewportStateCreateInfo(Structure):
    _fields_ = [("sType", c_int32), ("pNext", c_void_p), ("flags", c_uint32), ("viewportCount", c_uint32),
                ("pViewports", POINTER(VkViewport)), ("scissorCount", c_uint32), ("pScissors", POINTER(VkRect2D))]


class VkPipelineRasterizationStateCreateInfo(Structure):
    _fields_ = [("sType", c_int32), ("pNext", c_void_p), ("flags", c_uint32), ("depthClampEnable", c_uint32),
                ("rasterizerDiscardEnable", c_uint32), ("polygonMode", c_int32), ("cullMode", c_uint32),
                ("frontFace", c_int32), ("depthBiasEnable", c_uint32), ("depthBiasConstantFactor", c_float),
                ("depthBiasClamp", c_float), ("depthBiasSlopeFactor", c_float), ("lineWidth", c_float)]


class VkPipelineMultisampleStateCreateInfo(Structure):
    _fields_ = [("sType", c_int32), ("pNext", c_void_p), ("flags", c_uint32), ("rasterizationSamples", c_uint32),
                ("sampleShadingEnable", c_uint32), ("minSampleShading", c_float), ("pSampleMask", c_void_p),
                ("alphaToCoverageEnable", c_uint32), ("alphaToOneEnable", c_uint32)]


class VkStencilOpState(Structure):
    _fields_ = [("failOp", c_int32), ("passOp", c_int32), ("depthFailOp", c_int32), ("compareOp", c_int32),
                ("compareMask", c_uint32), ("writeMask", c_uint32), ("reference", c_uint32)]


class VkPipelineDepthStencilStateCreateInfo(Structure):
    _fields_ = [("sType", c_int32), ("pNext", c_void_p), ("flags", c_uint32), ("depthTestEnable", c_uint32),
                ("depthWriteEnable", c_uint32), ("depthCompareOp", c_int32), ("depthBoundsTestEnable", c_uint32),
                ("stencilTestEnable", c_uint32), ("front", VkStencilOpState), ("back", VkStencilOpState),
                ("minDepthBounds", c_float), ("maxDepthBounds", c_float)]


class VkPipelineColorBlendAttachmentState(Structure):
    _fields_ = [("blendEnable", c_uint32), ("srcColorBlendFactor", c_int32), ("dstColorBlendFactor", c_int32),
                ("colorBlendOp", c_int32), ("srcAlphaBlendFactor", c_int32), ("dstAlphaBlendFactor", c_int32),
                ("alphaBlendOp", c_int32), ("colorWriteMask", c_uint32)]


class VkPipelineColorBlendStateCreateInfo(Structure):
    _fields_ = [("sType", c_int32), ("pNext", c_void_p), ("flags", c_uint32), ("logicOpEnable", c_uint32),
                ("logicOp", c_int32), ("attachmentCount", c_uint32),
                ("pAttachments", POINTER(VkPipelineColorBlendAttachmentState)), ("blendConstants", c_float * 4)]


class VkPipelineDynamicStateCreateInfo(Structure):
    _fields_ = [("sType", c_int32), ("pNext", c_void_p), ("flags", c_uint32), ("dynamicStateCount", c_uint32),
                ("pDynamicStates", POINTER(c_int32))]


class VkGraphicsPipelineCreateInfo(Structure):
    _fields_ = [("sType", c_int32), ("pNext", c_void_p), ("flags", c_uint32), ("stageCount", c_uint32),
                ("pStages", POINTER(VkPipelineShaderStageCreateInfo)),
                ("pVertexInputState", POINTER(VkPipelineVertexInputStateCreateInfo)),
                ("pInputAssemblyState", POINTER(VkPipelineInputAssemblyStateCreateInfo)),
                ("pTessellationState", c_void_p), ("pViewportState", POINTER(VkPipelineViewportStateCreateInfo)),
                ("pRasterizationState", POINTER(VkPipelineRasterizationStateCreateInfo)),
                ("pMultisampleState", POINTER(VkPipelineMultisampleStateCreateInfo)),
                ("pDepthStencilState", POINTER(VkPipelineDepthStencilStateCreateInfo)),
                ("pColorBlendState", POINTER(VkPipelineColorBlendStateCreateInfo)),
                ("pDynamicState", POINTER(VkPipelineDynamicStateCreateInfo)), ("layout", Handle),
                ("renderPass", Handle), ("subpass", c_uint32), ("basePipelineHandle", Handle),
                ("basePipelineIndex", c_int32)]


class VkComputePipelineCreateInfo(Structure):
    _fields_ = [("sType", c_int32), ("pNext", c_void_p), ("flags", c_uint32),
                ("stage", VkPipelineShaderStageCreateInfo), ("layout", Handle), ("basePipelineHandle", Handle),
                ("basePipelineIndex", c_int32)]


class VkAttachmentDescription(Structure):
    _fields_ = [("flags", c_uint32), ("format", c_int32), ("samples", c_uint32), ("loadOp", c_int32),
                ("storeOp", c_int32), ("stencilLoadOp", c_int32), ("stencilStoreOp", c_int32),
                ("initialLayout", c_int32), ("finalLayout", c_int32)]


class VkAttachmentReference(Structure):
    _fields_ = [("attachment", c_uint32), ("layout", c_int32)]


class VkSubpassDescription(Structure):
    _fields_ = [("flags", c_uint32), ("pipelineBindPoint", c_int32), ("inputAttachmentCount", c_uint32),
                ("pInputAttachments", POINTER(VkAttachmentReference)), ("colorAttachmentCount", c_uint32),
                ("pColorAttachments", POINTER(VkAttachmentReference)),
                ("pResolveAttachments", POINTER(VkAttachmentReference)),
                ("pDepthStencilAttachment", POINTER(VkAttachmentReference)), ("preserveAttachmentCount", c_uint32),
                ("pPreserveAttachments", c_void_p)]


class VkRenderPassCreateInfo(Structure):
    _fields_ = [("sType", c_int32), ("pNext", c_void_p), ("flags", c_uint32), ("attachmentCount", c_uint32),
                ("pAttachments", POINTER(VkAttachmentDescription)), ("subpassCount", c_uint32),
                ("pSubpasses", POINTER(VkSubpassDescription)), ("dependencyCount", c_uint32),
                ("pDependencies", c_void_p)]


class VkFramebufferCreateInfo(Structure):
    _fields_ = [("sType", c_int32), ("pNext", c_void_p), ("flags", c_uint32), ("renderPass", Handle),
                ("attachmentCount", c_uint32), ("pAttachments", POINTER(Handle)), ("width", c_uint32),
                ("height", c_uint32), ("layers", c_uint32)]


class VkCommandPoolCreateInfo(Structure):
    _fields_ = [("sType", c_int32), ("pNext", c_void_p), ("flags", c_uint32), ("queueFamilyIndex", c_uint32)]


class VkCommandBufferAllocateInfo(Structure):
    _fields_ = [("sType", c_int32), ("pNext", c_void_p), ("commandPool", Handle), ("level", c_int32),
                ("commandBufferCount", c_uint32)]


class VkCommandBufferBeginInfo(Structure):
    _fields_ = [("sType", c_int32), ("pNext", c_void_p), ("flags", c_uint32), ("pInheritanceInfo", c_void_p)]


class VkClearColorValue(Union):
    _fields_ = [("float32", c_float * 4), ("int32", c_int32 * 4), ("uint32", c_uint32 * 4)]


class VkClearDepthStencilValue(Structure):
    _fields_ = [("depth", c_float), ("stencil", c_uint32)]


class VkClearValue(Union):
    _fields_ = [("color", VkClearColorValue), ("depthStencil", VkClearDepthStencilValue)]


class VkRenderPassBeginInfo(Structure):
    _fields_ = [("sType", c_int32), ("pNext", c_void_p), ("renderPass", Handle), ("framebuffer", Handle),
                ("renderArea", VkRect2D), ("clearValueCount", c_uint32), ("pClearValues", POINTER(VkClearValue))]


class VkBufferCopy(Structure):
    _fields_ = [("srcOffset", c_uint64), ("dstOffset", c_uint64), ("size", c_uint64)]


class VkImageSubresourceLayers(Structure):
    _fields_ = [("aspectMask", c_uint32), ("mipLevel", c_uint32), ("baseArrayLayer", c_uint32),
                ("layerCount", c_uint32)]


class VkBufferImageCopy(Structure):
    _fields_ = [("bufferOffset", c_uint64), ("bufferRowLength", c_uint32), ("bufferImageHeight", c_uint32),
                ("imageSubresource", VkImageSubresourceLayers), ("imageOffset", VkOffset3D),
                ("imageExtent", VkExtent3D)]


class VkSubmitInfo(Structure):
    _fields_ = [("sType", c_int32), ("pNext", c_void_p), ("waitSemaphoreCount", c_uint32),
                ("pWaitSemaphores", c_void_p), ("pWaitDstStageMask", c_void_p), ("commandBufferCount", c_uint32),
                ("pCommandBuffers", POINTER(Handle)), ("signalSemaphoreCount", c_uint32),
                ("pSignalSemaphores", c_void_p)]


class VkFenceCreateInfo(Structure):
    _fields_ = [("sType", c_int32), ("pNext", c_void_p), ("flags", c_uint32)]


# Sizes the C compiler gives the same structs on LP64 (checked by the tests).
ABI_STRUCTS = [VkApplicationInfo, VkPipelineCacheCreateInfo, VkInstanceCreateInfo, VkPhysicalDeviceLimits, VkPhysicalDeviceProperties,
               VkPhysicalDeviceMemoryProperties, VkDeviceQueueCreateInfo, VkDeviceCreateInfo, VkMemoryAllocateInfo,
               VkMemoryRequirements, VkBufferCreateInfo, VkImageCreateInfo, VkImageViewCreateInfo,
               VkSamplerCreateInfo, VkShaderModuleCreateInfo, VkDescriptorSetLayoutBinding,
               VkDescriptorSetLayoutCreateInfo, VkDescriptorPoolCreateInfo, VkDescriptorSetAllocateInfo,
               VkDescriptorImageInfo, VkDescriptorBufferInfo, VkWriteDescriptorSet, VkPushConstantRange,
               VkPipelineLayoutCreateInfo, VkPipelineShaderStageCreateInfo, VkVertexInputBindingDescription,
               VkVertexInputAttributeDescription, VkPipelineVertexInputStateCreateInfo,
               VkPipelineInputAssemblyStateCreateInfo, VkViewport, VkPipelineViewportStateCreateInfo,
               VkPipelineRasterizationStateCreateInfo, VkPipelineMultisampleStateCreateInfo,
               VkPipelineDepthStencilStateCreateInfo, VkPipelineColorBlendAttachmentState,
               VkPipelineColorBlendStateCreateInfo, VkPipelineDynamicStateCreateInfo, VkGraphicsPipelineCreateInfo,
               VkComputePipelineCreateInfo, VkAttachmentDescription, VkAttachmentReference, VkSubpassDescription, VkRenderPassCreateInfo,
               VkFramebufferCreateInfo, VkCommandPoolCreateInfo, VkCommandBufferAllocateInfo,
               VkCommandBufferBeginInfo, VkClearValue, VkRenderPassBeginInfo, VkBufferCopy, VkBufferImageCopy,
               VkSubmitInfo, VkFenceCreateInfo]

P = POINTER
_SIGS = {
    # name: (restype, argtypes)
    "vkCreateInstance": (c_int32, [P(VkInstanceCreateInfo), c_void_p, P(Handle)]),
    "vkDestroyInstance": (None, [Handle, c_void_p]),
    "vkEnumeratePhysicalDevices": (c_int32, [Handle, P(c_uint32), P(Handle)]),
    "vkEnumerateInstanceVersion": (c_int32, [P(c_uint32)]),
    "vkGetPhysicalDeviceProperties": (None, [Handle, P(VkPhysicalDeviceProperties)]),
    "vkGetPhysicalDeviceMemoryProperties": (None, [Handle, P(VkPhysicalDeviceMemoryProperties)]),
    "vkCreateDevice": (c_int32, [Handle, P(VkDeviceCreateInfo), c_void_p, P(Handle)]),
    "vkDestroyDevice": (None, [Handle, c_void_p]),
    "vkGetDeviceQueue": (None, [Handle, c_uint32, c_uint32, P(Handle)]),
    "vkDeviceWaitIdle": (c_int32, [Handle]),
    "vkQueueWaitIdle": (c_int32, [Handle]),
    "vkAllocateMemory": (c_int32, [Handle, P(VkMemoryAllocateInfo), c_void_p, P(Handle)]),
    "vkFreeMemory": (None, [Handle, Handle, c_void_p]),
    "vkMapMemory": (c_int32, [Handle, Handle, c_uint64, c_uint64, c_uint32, P(c_void_p)]),
    "vkUnmapMemory": (None, [Handle, Handle]),
    "vkCreateBuffer": (c_int32, [Handle, P(VkBufferCreateInfo), c_void_p, P(Handle)]),
    "vkDestroyBuffer": (None, [Handle, Handle, c_void_p]),
    "vkGetBufferMemoryRequirements": (None, [Handle, Handle, P(VkMemoryRequirements)]),
    "vkBindBufferMemory": (c_int32, [Handle, Handle, Handle, c_uint64]),
    "vkCreateImage": (c_int32, [Handle, P(VkImageCreateInfo), c_void_p, P(Handle)]),
    "vkDestroyImage": (None, [Handle, Handle, c_void_p]),
    "vkGetImageMemoryRequirements": (None, [Handle, Handle, P(VkMemoryRequirements)]),
    "vkBindImageMemory": (c_int32, [Handle, Handle, Handle, c_uint64]),
    "vkCreateImageView": (c_int32, [Handle, P(VkImageViewCreateInfo), c_void_p, P(Handle)]),
    "vkDestroyImageView": (None, [Handle, Handle, c_void_p]),
    "vkCreateSampler": (c_int32, [Handle, P(VkSamplerCreateInfo), c_void_p, P(Handle)]),
    "vkDestroySampler": (None, [Handle, Handle, c_void_p]),
    "vkCreateShaderModule": (c_int32, [Handle, P(VkShaderModuleCreateInfo), c_void_p, P(Handle)]),
    "vkDestroyShaderModule": (None, [Handle, Handle, c_void_p]),
    "vkCreateDescriptorSetLayout": (c_int32, [Handle, P(VkDescriptorSetLayoutCreateInfo), c_void_p, P(Handle)]),
    "vkDestroyDescriptorSetLayout": (None, [Handle, Handle, c_void_p]),
    "vkCreateDescriptorPool": (c_int32, [Handle, P(VkDescriptorPoolCreateInfo), c_void_p, P(Handle)]),
    "vkDestroyDescriptorPool": (None, [Handle, Handle, c_void_p]),
    "vkAllocateDescriptorSets": (c_int32, [Handle, P(VkDescriptorSetAllocateInfo), P(Handle)]),
    "vkUpdateDescriptorSets": (None, [Handle, c_uint32, P(VkWriteDescriptorSet), c_uint32, c_void_p]),
    "vkCreatePipelineLayout": (c_int32, [Handle, P(VkPipelineLayoutCreateInfo), c_void_p, P(Handle)]),
    "vkCreatePipelineCache": (c_int32, [Handle, P(VkPipelineCacheCreateInfo), c_void_p, P(Handle)]),
    "vkDestroyPipelineCache": (None, [Handle, Handle, c_void_p]),
    "vkGetPipelineCacheData": (c_int32, [Handle, Handle, P(c_size_t), c_void_p]),
    "vkMergePipelineCaches": (c_int32, [Handle, Handle, c_uint32, P(Handle)]),
    "vkDestroyPipelineLayout": (None, [Handle, Handle, c_void_p]),
    "vkCreateRenderPass": (c_int32, [Handle, P(VkRenderPassCreateInfo), c_void_p, P(Handle)]),
    "vkDestroyRenderPass": (None, [Handle, Handle, c_void_p]),
    "vkCreateFramebuffer": (c_int32, [Handle, P(VkFramebufferCreateInfo), c_void_p, P(Handle)]),
    "vkDestroyFramebuffer": (None, [Handle, Handle, c_void_p]),
    "vkCreateGraphicsPipelines": (c_int32, [Handle, Handle, c_uint32, P(VkGraphicsPipelineCreateInfo), c_void_p,
                                            P(Handle)]),
    "vkCreateComputePipelines": (c_int32, [Handle, Handle, c_uint32, P(VkComputePipelineCreateInfo), c_void_p,
                                           P(Handle)]),
    "vkDestroyPipeline": (None, [Handle, Handle, c_void_p]),
    "vkCreateCommandPool": (c_int32, [Handle, P(VkCommandPoolCreateInfo), c_void_p, P(Handle)]),
    "vkDestroyCommandPool": (None, [Handle, Handle, c_void_p]),
    "vkAllocateCommandBuffers": (c_int32, [Handle, P(VkCommandBufferAllocateInfo), P(Handle)]),
    "vkFreeCommandBuffers": (None, [Handle, Handle, c_uint32, P(Handle)]),
    "vkBeginCommandBuffer": (c_int32, [Handle, P(VkCommandBufferBeginInfo)]),
    "vkEndCommandBuffer": (c_int32, [Handle]),
    "vkCmdBindPipeline": (None, [Handle, c_int32, Handle]),
    "vkCmdSetViewport": (None, [Handle, c_uint32, c_uint32, P(VkViewport)]),
    "vkCmdSetScissor": (None, [Handle, c_uint32, c_uint32, P(VkRect2D)]),
    "vkCmdBindVertexBuffers": (None, [Handle, c_uint32, c_uint32, P(Handle), P(c_uint64)]),
    "vkCmdBindIndexBuffer": (None, [Handle, Handle, c_uint64, c_int32]),
    "vkCmdBindDescriptorSets": (None, [Handle, c_int32, Handle, c_uint32, c_uint32, P(Handle), c_uint32, c_void_p]),
    "vkCmdPushConstants": (None, [Handle, Handle, c_uint32, c_uint32, c_uint32, c_void_p]),
    "vkCmdBeginRenderPass": (None, [Handle, P(VkRenderPassBeginInfo), c_int32]),
    "vkCmdEndRenderPass": (None, [Handle]),
    "vkCmdDraw": (None, [Handle, c_uint32, c_uint32, c_uint32, c_uint32]),
    "vkCmdDrawIndexed": (None, [Handle, c_uint32, c_uint32, c_uint32, c_int32, c_uint32]),
    "vkCmdDispatch": (None, [Handle, c_uint32, c_uint32, c_uint32]),
    "vkCmdDispatchBase": (None, [Handle, c_uint32, c_uint32, c_uint32, c_uint32, c_uint32, c_uint32]),
    "vkCmdCopyBuffer": (None, [Handle, Handle, Handle, c_uint32, P(VkBufferCopy)]),
    "vkCmdCopyBufferToImage": (None, [Handle, Handle, Handle, c_int32, c_uint32, P(VkBufferImageCopy)]),
    "vkCmdCopyImageToBuffer": (None, [Handle, Handle, c_int32, Handle, c_uint32, P(VkBufferImageCopy)]),
    "vkCreateFence": (c_int32, [Handle, P(VkFenceCreateInfo), c_void_p, P(Handle)]),
    "vkDestroyFence": (None, [Handle, Handle, c_void_p]),
    "vkResetFences": (c_int32, [Handle, c_uint32, P(Handle)]),
    "vkWaitForFences": (c_int32, [Handle, c_uint32, P(Handle), c_uint32, c_uint64]),
    "vkQueueSubmit": (c_int32, [Handle, c_uint32, P(VkSubmitInfo), Handle]),
    "vkGetDrawStatsKAZAN": (c_int32, [Handle, P(c_uint64), P(c_float), P(c_uint32), P(c_uint32)]),
    "vkSetKernelTimingKAZAN": (None, [Handle, c_uint32]),
    "vkGetKernelTimesKAZAN": (c_int32, [Handle, P(c_float), P(c_uint32)]),
    "vkSetTileProfileKAZAN": (None, [Handle, c_uint32]),
    "vkGetTileProfileKAZAN": (c_int32, [Handle, P(c_uint32), P(c_float), P(c_float), P(c_float), P(c_float)]),
    "vkSetCostModelKAZAN": (c_int32, [Handle, c_void_p, c_size_t]),
    "vkHasCostModelKAZAN": (c_uint32, [Handle]),
    "vkFillClusterSceneKAZAN": (c_int32, [Handle, Handle, c_uint64, c_void_p]),
    "vkGetGraphStatsKAZAN": (None, [Handle, P(c_uint32), P(c_uint32)]),
    "vkGetTuningKAZAN": (None, [Handle] + [P(c_uint32)] * 6),
}

KERNEL_NAMES = ["clear_counts", "bin_count", "bin_hist", "bin_scan", "bin_scatter", "bin_sort", "plan", "raster"]

STAT_NAMES = ["triangles_in", "triangles_setup", "triangles_clipped", "fragments", "depth_passed",
              "list_entries", "list_overflow", "split_entries"]


class VulkanError(RuntimeError):
    def __init__(self, fn, code):
        super().__init__("%s failed: VkResult %d" % (fn, code))
        self.code = code


class Driver:
    """The ICD loaded as a Vulkan loader would load it."""

    def __init__(self, path=DRIVER_PATH):
        if not os.path.exists(path):
            raise FileNotFoundError("%s is not built (run python -m kazan_amd.build)" % path)
        self.lib = ctypes.CDLL(path)
        gipa = self.lib.vk_icdGetInstanceProcAddr
        gipa.restype = c_void_p
        gipa.argtypes = [Handle, c_char_p]
        self._gipa = gipa
        neg = self.lib.vk_icdNegotiateLoaderICDInterfaceVersion
        neg.restype = c_int32
        neg.argtypes = [P(c_uint32)]
        v = c_uint32(7)
        if neg(ctypes.byref(v)) != VK_SUCCESS:
            raise RuntimeError("ICD interface negotiation failed")
        self.icd_version = v.value
        self._fns = {}

    def proc(self, name):
        if name not in self._fns:
            addr = self._gipa(None, name.encode())
            if not addr:
                raise AttributeError("driver does not export %s" % name)
            restype, argtypes = _SIGS[name]
            self._fns[name] = ctypes.CFUNCTYPE(restype, *argtypes)(addr)
        return self._fns[name]

    def __getattr__(self, name):
        if name.startswith("vk"):
            return self.proc(name)
        raise AttributeError(name)

    def check(self, name, *args):
        r = self.proc(name)(*args)
        if r != VK_SUCCESS:
            raise VulkanError(name, r)
        return r


# --- shader-compiler tools exported beside the ICD (include/kazan/kazan_tools.h)

class KazanPipelineDesc(ctypes.Structure):
    _fields_ = [("vs_words", c_void_p), ("vs_nwords", ctypes.c_size_t), ("fs_words", c_void_p),
                ("fs_nwords", ctypes.c_size_t), ("num_bindings", c_uint32), ("binding_stride", c_uint32 * 8),
                ("binding_rate", c_uint32 * 8), ("num_attributes", c_uint32), ("attr_location", c_uint32 * 16),
                ("attr_binding", c_uint32 * 16), ("attr_format", c_uint32 * 16), ("attr_offset", c_uint32 * 16),
                ("tile_shift", c_uint32), ("samples", c_uint32), ("cull_mode", c_uint32), ("front_face", c_uint32),
                ("depth_test", c_uint32), ("depth_write", c_uint32), ("depth_op", c_uint32), ("blend", c_uint32),
                ("src_color", c_uint32), ("dst_color", c_uint32), ("color_op", c_uint32), ("src_alpha", c_uint32),
                ("dst_alpha", c_uint32), ("alpha_op", c_uint32), ("write_mask", c_uint32), ("num_ubos", c_uint32),
                ("ubo_key", c_uint32 * 4), ("num_textures", c_uint32), ("texture_key", c_uint32 * 4)]


def tool_call(driver, name, *args, cap=1 << 22):
    """Calls one of the kazan_* text-returning tools; returns (rc, text)."""
    fn = getattr(driver.lib, name)
    fn.restype = ctypes.c_int
    out = ctypes.create_string_buffer(cap)
    n = ctypes.c_size_t(0)
    rc = fn(*args, out, ctypes.c_size_t(cap), ctypes.byref(n))
    return rc, out.value.decode(errors="replace")


class KazanClusterScene(ctypes.Structure):
    _fields_ = [("seed", c_uint64), ("triangles", c_uint32), ("width", c_uint32), ("height", c_uint32),
                ("pad", c_uint32), ("cluster_frac", ctypes.c_double), ("sigma_frac", ctypes.c_double),
                ("radius", ctypes.c_double)]
