/*
 * kazan_icd.h — the C ABI libkazan_driver.so exports.
 *
 * These are the three symbols the reference's cdylib exports
 * (vulkan-driver/Cargo.toml:10-12 crate-type cdylib, vulkan-driver/src/lib.rs)
 * and through which the Vulkan loader reaches every vk* entry point:
 *
 *   vk_icdGetInstanceProcAddr
 *       replaces vulkan-driver/src/lib.rs:37-43 (forwards to
 *       vkGetInstanceProcAddr, api_impl.rs:3067-3086, which resolves names
 *       through get_proc_address, api_impl.rs:432-860).
 *   vk_icdNegotiateLoaderICDInterfaceVersion
 *       replaces lib.rs:53-61: clamps *pSupportedVersion to ICD_VERSION = 5
 *       (lib.rs:51) and returns VK_SUCCESS.
 *   vk_icdGetPhysicalDeviceProcAddr
 *       replaces lib.rs:67-128: returns the entry point for physical-device
 *       level names, NULL for any other name.
 *
 * Dispatchable handles (VkInstance, VkPhysicalDevice, VkDevice, VkQueue,
 * VkCommandBuffer) point at objects whose first word holds ICD_LOADER_MAGIC
 * (handle.rs:23-36); non-dispatchable handles are object pointers
 * (handle.rs:121-159). pAllocator is accepted and ignored, as in the
 * reference (e.g. api_impl.rs:3437).
 *
 * The draw-path entry points the reference leaves unimplemented and this
 * driver implements, all reached through vk_icdGetInstanceProcAddr /
 * vkGetDeviceProcAddr:
 *   vkCmdDraw               api_impl.rs:4632-4640
 *   vkCmdDrawIndexed        api_impl.rs:4643-4652
 *   vkCreateGraphicsPipelines api_impl.rs:4018-4032 -> pipeline.rs:165-183
 *   vkQueueSubmit           api_impl.rs:3414-3421
 *   vkCmdBeginRenderPass / vkCmdEndRenderPass   :4944 / :4961
 *   vkCmdBindPipeline :4506, vkCmdSetViewport :4515, vkCmdSetScissor :4525,
 *   vkCmdBindDescriptorSets :4597, vkCmdBindIndexBuffer :4611,
 *   vkCmdBindVertexBuffers :4621, vkCmdPushConstants :4932,
 *   vkCmdCopyBuffer, vkCmdCopyBufferToImage, vkCmdCopyImageToBuffer :4746,
 *   vkCreateCommandPool :4436, vkAllocateCommandBuffers :4464,
 *   vkBegin/EndCommandBuffer :4483/:4491, vkCreateFramebuffer :4383-4390,
 *   vkCreateRenderPass :4402-4415, vkCreateFence :3653, vkWaitForFences :3689,
 *   vkQueueWaitIdle :3424.
 *
 * One device-level extension entry point, resolved by name through
 * vkGetDeviceProcAddr, has no counterpart in the reference: it reports the
 * draw statistics of the submissions completed since the last call.
 */
#ifndef KAZAN_ICD_H
#define KAZAN_ICD_H

#include "vk_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

#define KAZAN_ICD_VERSION 5u

/* The library is built with hidden visibility: these three (and the tools
 * of kazan_tools.h) are its only exported C names. */
#ifndef KAZAN_EXPORT
#define KAZAN_EXPORT __attribute__((visibility("default")))
#endif

KAZAN_EXPORT VKAPI_ATTR PFN_vkVoidFunction VKAPI_CALL vk_icdGetInstanceProcAddr(VkInstance instance,
                                                                                const char *pName);
KAZAN_EXPORT VKAPI_ATTR VkResult VKAPI_CALL vk_icdNegotiateLoaderICDInterfaceVersion(uint32_t *pSupportedVersion);
KAZAN_EXPORT VKAPI_ATTR PFN_vkVoidFunction VKAPI_CALL vk_icdGetPhysicalDeviceProcAddr(VkInstance instance,
                                                                                      const char *pName);

/* Counters of vkGetDrawStatsKAZAN, in order. */
enum KazanDrawStat {
    KAZAN_STAT_TRIANGLES_IN = 0,          /* primitives assembled */
    KAZAN_STAT_TRIANGLES_SETUP = 1,       /* survived clip-reject, cull, zero area (clipped: a piece did) */
    KAZAN_STAT_TRIANGLES_CLIPPED = 2,     /* clipped against the near/far planes or the guard band */
    KAZAN_STAT_FRAGMENTS = 3,             /* covered (triangle, sample) pairs */
    KAZAN_STAT_DEPTH_PASSED = 4,          /* samples that passed the depth test */
    KAZAN_STAT_LIST_ENTRIES = 5,          /* binned (triangle, tile) pairs */
    KAZAN_STAT_LIST_OVERFLOW = 6,         /* batches whose tile lists overflowed */
    KAZAN_STAT_SPLIT_ENTRIES = 7,         /* entries of split tiles' sub-tile lists */
    KAZAN_STAT_COUNT = 8
};

/* vkGetDrawStatsKAZAN: waits for the device's queue, then writes the
 * KAZAN_STAT_COUNT counters accumulated over the submissions since the
 * previous call, the GPU time from the first of them to the end of the last
 * (pGpuMs), and the number of batches and kernel launches. */
typedef VkResult(VKAPI_PTR *PFN_vkGetDrawStatsKAZAN)(VkDevice device, uint64_t *pCounters, float *pGpuMs,
                                                     uint32_t *pBatches, uint32_t *pLaunches);

/* Kernels of one draw batch, in launch order, for per-kernel timing. */
enum KazanKernel {
    KAZAN_KERNEL_CLEAR_COUNTS = 0, /* counter resets */
    KAZAN_KERNEL_BIN_COUNT = 1,    /* vertex shader + setup: pixel boxes */
    KAZAN_KERNEL_BIN_HIST = 2,     /* (triangle, tile) pairs per tile */
    KAZAN_KERNEL_BIN_SCAN = 3,     /* list offsets */
    KAZAN_KERNEL_BIN_SCATTER = 4,  /* tile lists */
    KAZAN_KERNEL_BIN_SORT = 5,     /* lists into primitive order */
    KAZAN_KERNEL_PLAN = 6,         /* cost model, split, sub-tile lists, queue */
    KAZAN_KERNEL_RASTER = 7,
    KAZAN_KERNEL_COUNT = 8
};

/* vkSetKernelTimingKAZAN: records a HIP event pair around every batch kernel
 * on the queue's stream from the next submission on. vkGetKernelTimesKAZAN
 * waits for the queue and writes, per KazanKernel, the summed milliseconds
 * and the number of launches of the submissions since the previous
 * statistics read. Neither has a counterpart in the reference. */
typedef void(VKAPI_PTR *PFN_vkSetKernelTimingKAZAN)(VkDevice device, VkBool32 enable);
typedef VkResult(VKAPI_PTR *PFN_vkGetKernelTimesKAZAN)(VkDevice device, float *pMs, uint32_t *pLaunches);

/* The learned tile-cost model of the balanced pipeline form (no counterpart
 * in the reference, which has no draw path to balance).
 *
 * vkSetTileProfileKAZAN: from the next submission on, every batch also
 * computes the 16 per-tile cost-model features (and, with a model, the
 * predicted fragments), and the plain form records each tile's fragment
 * count (the model's training label) and its raster workgroup's
 * shader-clock cycles. vkGetTileProfileKAZAN waits for the queue and copies
 * the last batch's tile count, features [tiles * 16], fragments, cycles and
 * predicted fragments [tiles each] (zeros where the batch produced none);
 * null arrays are skipped.
 *
 * vkSetCostModelKAZAN replaces the model with a KzCostModel blob (the
 * format of cost_model.bin: magic "KZCM", version 1, size-checked);
 * VK_ERROR_INITIALIZATION_FAILED leaves the current one in place.
 * vkHasCostModelKAZAN: whether the balanced form predicts costs (VK_FALSE:
 * it splits tiles by binned triangle count). */
typedef void(VKAPI_PTR *PFN_vkSetTileProfileKAZAN)(VkDevice device, VkBool32 enable);
typedef VkResult(VKAPI_PTR *PFN_vkGetTileProfileKAZAN)(VkDevice device, uint32_t *pTileCount, float *pFeatures,
                                                        float *pFragments, float *pCycles, float *pPredicted);
typedef VkResult(VKAPI_PTR *PFN_vkSetCostModelKAZAN)(VkDevice device, const void *pBlob, size_t size);
typedef VkBool32(VKAPI_PTR *PFN_vkHasCostModelKAZAN)(VkDevice device);

/* vkFillClusterSceneKAZAN: the cost model's training-data generator. Writes
 * the Gaussian-cluster scene of `seed` (round(triangles * cluster_frac)
 * triangles around the target's centre with sigma = sigma_frac * (width,
 * height) pixels, the rest uniform; vertices at offsets U[-radius,
 * radius]^2 px; z and RGBA uniform [0, 1)) into `buffer` at `offset` as
 * 3 * triangles vertices of vec4 position (NDC, w = 1) + vec4 colour, on the
 * device, and waits for it. It follows the host generator
 * (kazan_amd/scenes/generators.py gaussian_cluster) in double precision, so
 * a fresh frame costs no host work or upload. */
typedef struct KazanClusterScene {
    uint64_t seed;
    uint32_t triangles, width, height, pad;
    double cluster_frac, sigma_frac, radius;
} KazanClusterScene;
typedef VkResult(VKAPI_PTR *PFN_vkFillClusterSceneKAZAN)(VkDevice device, VkBuffer buffer, VkDeviceSize offset,
                                                         const KazanClusterScene *pScene);

/* Batched launch. In the balanced form a command buffer's batches are
 * autotuned over their first submissions (one timed candidate per
 * submission: tile size 8..64, then the persistent raster's grid, then the
 * plan without the cost model and the plain one-workgroup-per-tile raster;
 * KAZAN_AUTOTUNE=0 turns this off), after which the whole command buffer is
 * captured into a hipGraph on its next submission and replayed as one graph
 * launch (KAZAN_GRAPHS=0 turns this off; re-recording, or any reallocation
 * of the device's scratch buffers, makes it capture again).
 * vkGetGraphStatsKAZAN: graph captures and graph launches so far.
 * vkGetTuningKAZAN: what the most recently finished tuning chose: tile size
 * (8..64, 0 before any tuning finished), persistent-grid divisor, whether
 * the plan uses the cost model, whether the raster is the persistent one,
 * the deepest split the plan may make (2: 16 sub-tiles, 1: 4, 0: none), the
 * factor on the cost model's fragment weight (any pointer may be null). */
typedef void(VKAPI_PTR *PFN_vkGetGraphStatsKAZAN)(VkDevice device, uint32_t *pCaptures, uint32_t *pLaunches);
typedef void(VKAPI_PTR *PFN_vkGetTuningKAZAN)(VkDevice device, uint32_t *pTileSize, uint32_t *pGridDivisor,
                                              uint32_t *pCostModel, uint32_t *pPersistentRaster,
                                              uint32_t *pMaxSplitLevel, uint32_t *pFragWeightScale);

#ifdef __cplusplus
}
#endif

#endif /* KAZAN_ICD_H */
