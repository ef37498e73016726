"""A compute pipeline driven through the ICD's entry points, as an
application drives one: storage buffers in descriptor set 0 (bindings
0 .. n-1), an optional push-constant block, vkCmdDispatch /
vkCmdDispatchBase on the compute bind point (the reference's
vkCreateComputePipelines / vkCmdDispatch, vulkan-driver/src/api_impl.rs).
"""
import ctypes

import numpy as np

from kazan_amd import vk

_ref = ctypes.byref


class ComputeKernel:
    def __init__(self, ctx, code, n_buffers, push_bytes=0):
        self.ctx = ctx
        d, dev = ctx.d, ctx.device
        self.push_bytes = push_bytes
        self.n = n_buffers
        self._keep = []
        binds = (vk.VkDescriptorSetLayoutBinding * n_buffers)(
            *[vk.VkDescriptorSetLayoutBinding(i, vk.DESCRIPTOR_TYPE_STORAGE_BUFFER, 1, vk.SHADER_STAGE_COMPUTE, None)
              for i in range(n_buffers)])
        dl = vk.VkDescriptorSetLayoutCreateInfo(vk.ST_DESCRIPTOR_SET_LAYOUT_CREATE_INFO, None, 0, n_buffers, binds)
        self.set_layout = vk.Handle()
        d.check("vkCreateDescriptorSetLayout", dev, _ref(dl), None, _ref(self.set_layout))
        layouts = (vk.Handle * 1)(self.set_layout)
        pcr = vk.VkPushConstantRange(vk.SHADER_STAGE_COMPUTE, 0, push_bytes)
        pl = vk.VkPipelineLayoutCreateInfo(vk.ST_PIPELINE_LAYOUT_CREATE_INFO, None, 0, 1, layouts,
                                           1 if push_bytes else 0, ctypes.pointer(pcr) if push_bytes else None)
        self.layout = vk.Handle()
        d.check("vkCreatePipelineLayout", dev, _ref(pl), None, _ref(self.layout))
        buf = ctypes.create_string_buffer(code, len(code))
        self._keep.append(buf)
        smi = vk.VkShaderModuleCreateInfo(vk.ST_SHADER_MODULE_CREATE_INFO, None, 0, len(code), ctypes.addressof(buf))
        self.module = vk.Handle()
        d.check("vkCreateShaderModule", dev, _ref(smi), None, _ref(self.module))
        stage = vk.VkPipelineShaderStageCreateInfo(vk.ST_PIPELINE_SHADER_STAGE_CREATE_INFO, None, 0,
                                                   vk.SHADER_STAGE_COMPUTE, self.module, b"main", None)
        cci = vk.VkComputePipelineCreateInfo(vk.ST_COMPUTE_PIPELINE_CREATE_INFO, None, 0, stage, self.layout, None, -1)
        self.pipeline = vk.Handle()
        d.check("vkCreateComputePipelines", dev, ctx.pipeline_cache, 1, _ref(cci), None, _ref(self.pipeline))
        ps = vk.VkDescriptorPoolSize(vk.DESCRIPTOR_TYPE_STORAGE_BUFFER, n_buffers)
        dpi = vk.VkDescriptorPoolCreateInfo(vk.ST_DESCRIPTOR_POOL_CREATE_INFO, None, 0, 1, 1, ctypes.pointer(ps))
        self.dpool = vk.Handle()
        d.check("vkCreateDescriptorPool", dev, _ref(dpi), None, _ref(self.dpool))
        dai = vk.VkDescriptorSetAllocateInfo(vk.ST_DESCRIPTOR_SET_ALLOCATE_INFO, None, self.dpool, 1, layouts)
        self.set = vk.Handle()
        d.check("vkAllocateDescriptorSets", dev, _ref(dai), _ref(self.set))

    def bind(self, buffers):
        """buffers: one (VkBuffer, size in bytes) per binding."""
        infos = [vk.VkDescriptorBufferInfo(b, 0, size) for b, size in buffers]
        writes = (vk.VkWriteDescriptorSet * len(infos))(
            *[vk.VkWriteDescriptorSet(vk.ST_WRITE_DESCRIPTOR_SET, None, self.set, i, 0, 1,
                                      vk.DESCRIPTOR_TYPE_STORAGE_BUFFER, None, ctypes.pointer(info), None)
              for i, info in enumerate(infos)])
        self._keep.append(infos)
        self.ctx.d.vkUpdateDescriptorSets(self.ctx.device, len(infos), writes, 0, None)

    def record(self, cb, groups, push=None, base=None):
        d = self.ctx.d
        d.vkCmdBindPipeline(cb, vk.PIPELINE_BIND_POINT_COMPUTE, self.pipeline)
        d.vkCmdBindDescriptorSets(cb, vk.PIPELINE_BIND_POINT_COMPUTE, self.layout, 0, 1, (vk.Handle * 1)(self.set), 0,
                                  None)
        if push is not None:
            p = np.ascontiguousarray(push).view(np.uint8)
            assert p.nbytes <= self.push_bytes
            self._keep.append(p)
            d.vkCmdPushConstants(cb, self.layout, vk.SHADER_STAGE_COMPUTE, 0, p.nbytes, p.ctypes.data)
        gx, gy, gz = (tuple(groups) + (1, 1))[:3]
        if base is None:
            d.vkCmdDispatch(cb, gx, gy, gz)
        else:
            d.vkCmdDispatchBase(cb, base[0], base[1], base[2], gx, gy, gz)

    def run(self, groups, push=None, base=None):
        """Records one dispatch into a command buffer of its own, submits it
        and waits."""
        ctx = self.ctx
        cb = ctx.command_buffer()
        ctx.begin(cb)
        self.record(cb, groups, push, base)
        ctx.d.check("vkEndCommandBuffer", cb)
        ctx.submit(cb)
        ctx.d.vkFreeCommandBuffers(ctx.device, ctx.pool, 1, (vk.Handle * 1)(cb))

    def destroy(self):
        d, dev = self.ctx.d, self.ctx.device
        d.vkDestroyPipeline(dev, self.pipeline, None)
        d.vkDestroyShaderModule(dev, self.module, None)
        d.vkDestroyDescriptorPool(dev, self.dpool, None)
        d.vkDestroyPipelineLayout(dev, self.layout, None)
        d.vkDestroyDescriptorSetLayout(dev, self.set_layout, None)


def read_buffer(ctx, buf, nbytes, dtype):
    """Copies a device buffer back to the host (a staging buffer and
    vkCmdCopyBuffer), as a numpy array of `dtype`."""
    stage = ctx.buffer(nbytes, True)
    cb = ctx.command_buffer()
    ctx.begin(cb)
    ctx.d.vkCmdCopyBuffer(cb, buf, stage[0], 1, _ref(vk.VkBufferCopy(0, 0, nbytes)))
    ctx.d.check("vkEndCommandBuffer", cb)
    ctx.submit(cb)
    ctx.d.vkFreeCommandBuffers(ctx.device, ctx.pool, 1, (vk.Handle * 1)(cb))
    out = np.empty(nbytes, np.uint8)
    ctypes.memmove(out.ctypes.data, ctx.map(stage[1], nbytes), nbytes)
    ctx.destroy_buffer(stage)
    return out.view(dtype)
