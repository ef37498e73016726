"""A minimal Vulkan application that renders a Scene through the Kazan ICD.

It does what any Vulkan app does: create an instance and device, upload the
geometry and texture through a staging buffer into device-local memory,
create shader modules from SPIR-V, a render pass with colour (RGBA8) and
depth (D32) attachments, a graphics pipeline, record one command buffer
(clear + draws), submit it, and read the attachments back. Every call goes
through the entry points vk_icdGetInstanceProcAddr resolves.
"""
import ctypes

import numpy as np

from .. import vk
from . import generators as g
from . import shaders

_VS = {
    g.VS_POS4_COLOR4: (shaders.vs_pos4_color4, [(0, vk.FORMAT_R32G32B32A32_SFLOAT, 0), (1, vk.FORMAT_R32G32B32A32_SFLOAT, 16)]),
    g.VS_POS3_UV2: (shaders.vs_pos3_uv2, [(0, vk.FORMAT_R32G32B32_SFLOAT, 0), (1, vk.FORMAT_R32G32_SFLOAT, 12)]),
    g.VS_MVP_POS3_COLOR4: (shaders.vs_mvp_pos3_color4,
                           [(0, vk.FORMAT_R32G32B32_SFLOAT, 0), (1, vk.FORMAT_R32G32B32A32_SFLOAT, 12)]),
    g.VS_MVP_POS3_NORMAL3: (shaders.vs_mvp_pos3_normal3,
                            [(0, vk.FORMAT_R32G32B32_SFLOAT, 0), (1, vk.FORMAT_R32G32B32_SFLOAT, 12)]),
}
_FS = {g.FS_COLOR: shaders.fs_color, g.FS_TEXTURE: shaders.fs_texture, g.FS_PHONG: shaders.fs_phong,
       g.FS_ALPHA_TEST: shaders.fs_alpha_test}

BUFFER_USAGE_ALL = 0x1FF
IMAGE_USAGE_ALL = 0xFF


def _ref(x):
    return ctypes.byref(x)


class Context:
    """Instance, device, queue, command pool: one per process is enough."""

    def __init__(self, driver=None):
        self.d = driver or vk.Driver()
        d = self.d
        app = vk.VkApplicationInfo(vk.ST_APPLICATION_INFO, None, b"kazan-scene", 1, b"kazan", 1, vk.VK_MAKE_VERSION(1, 1, 0))
        ici = vk.VkInstanceCreateInfo(vk.ST_INSTANCE_CREATE_INFO, None, 0, ctypes.pointer(app), 0, None, 0, None)
        self.instance = vk.Handle()
        d.check("vkCreateInstance", _ref(ici), None, _ref(self.instance))
        n = ctypes.c_uint32(0)
        d.check("vkEnumeratePhysicalDevices", self.instance, _ref(n), None)
        if n.value == 0:
            raise RuntimeError("no Vulkan physical device (no GPU visible to HIP)")
        pds = (vk.Handle * n.value)()
        d.check("vkEnumeratePhysicalDevices", self.instance, _ref(n), pds)
        self.physical = vk.Handle(pds[0])
        self.pipeline_cache = None  # VkPipelineCache handle passed to vkCreateGraphicsPipelines, if set
        self.props = vk.VkPhysicalDeviceProperties()
        d.vkGetPhysicalDeviceProperties(self.physical, _ref(self.props))
        prio = ctypes.c_float(1.0)
        qci = vk.VkDeviceQueueCreateInfo(vk.ST_DEVICE_QUEUE_CREATE_INFO, None, 0, 0, 1, ctypes.pointer(prio))
        dci = vk.VkDeviceCreateInfo(vk.ST_DEVICE_CREATE_INFO, None, 0, 1, ctypes.pointer(qci), 0, None, 0, None, None)
        self.device = vk.Handle()
        d.check("vkCreateDevice", self.physical, _ref(dci), None, _ref(self.device))
        self.queue = vk.Handle()
        d.vkGetDeviceQueue(self.device, 0, 0, _ref(self.queue))
        pci = vk.VkCommandPoolCreateInfo(vk.ST_COMMAND_POOL_CREATE_INFO, None, 0, 0)
        self.pool = vk.Handle()
        d.check("vkCreateCommandPool", self.device, _ref(pci), None, _ref(self.pool))
        fci = vk.VkFenceCreateInfo(vk.ST_FENCE_CREATE_INFO, None, 0)
        self.fence = vk.Handle()
        d.check("vkCreateFence", self.device, _ref(fci), None, _ref(self.fence))
        self._keep = []

    @property
    def device_name(self):
        return self.props.deviceName.decode()

    # -- memory ------------------------------------------------------------
    def _alloc(self, reqs, host):
        mai = vk.VkMemoryAllocateInfo(vk.ST_MEMORY_ALLOCATE_INFO, None, reqs.size, 1 if host else 0)
        mem = vk.Handle()
        self.d.check("vkAllocateMemory", self.device, _ref(mai), None, _ref(mem))
        return mem

    def buffer(self, size, host):
        bci = vk.VkBufferCreateInfo(vk.ST_BUFFER_CREATE_INFO, None, 0, max(int(size), 4), BUFFER_USAGE_ALL, 0, 0, None)
        buf = vk.Handle()
        self.d.check("vkCreateBuffer", self.device, _ref(bci), None, _ref(buf))
        reqs = vk.VkMemoryRequirements()
        self.d.vkGetBufferMemoryRequirements(self.device, buf, _ref(reqs))
        mem = self._alloc(reqs, host)
        self.d.check("vkBindBufferMemory", self.device, buf, mem, 0)
        return buf, mem

    def destroy_buffer(self, bm):
        self.d.vkDestroyBuffer(self.device, bm[0], None)
        self.d.vkFreeMemory(self.device, bm[1], None)

    def map(self, mem, size):
        p = ctypes.c_void_p()
        self.d.check("vkMapMemory", self.device, mem, 0, size, 0, _ref(p))
        return p.value

    def image(self, w, h, fmt, samples=1):
        ici = vk.VkImageCreateInfo(vk.ST_IMAGE_CREATE_INFO, None, 0, 1, fmt, vk.VkExtent3D(w, h, 1), 1, 1, samples,
                                   1, IMAGE_USAGE_ALL, 0, 0, None, 0)
        img = vk.Handle()
        self.d.check("vkCreateImage", self.device, _ref(ici), None, _ref(img))
        reqs = vk.VkMemoryRequirements()
        self.d.vkGetImageMemoryRequirements(self.device, img, _ref(reqs))
        mem = self._alloc(reqs, False)
        self.d.check("vkBindImageMemory", self.device, img, mem, 0)
        aspect = 2 if fmt == vk.FORMAT_D32_SFLOAT else 1
        vci = vk.VkImageViewCreateInfo(vk.ST_IMAGE_VIEW_CREATE_INFO, None, 0, img, 1, fmt, (ctypes.c_int32 * 4)(),
                                       vk.VkImageSubresourceRange(aspect, 0, 1, 0, 1))
        view = vk.Handle()
        self.d.check("vkCreateImageView", self.device, _ref(vci), None, _ref(view))
        return img, mem, view

    # -- command buffers ---------------------------------------------------
    def command_buffer(self):
        cai = vk.VkCommandBufferAllocateInfo(vk.ST_COMMAND_BUFFER_ALLOCATE_INFO, None, self.pool, 0, 1)
        cb = vk.Handle()
        self.d.check("vkAllocateCommandBuffers", self.device, _ref(cai), _ref(cb))
        return cb

    def begin(self, cb):
        bi = vk.VkCommandBufferBeginInfo(vk.ST_COMMAND_BUFFER_BEGIN_INFO, None, 0, None)
        self.d.check("vkBeginCommandBuffer", cb, _ref(bi))

    def submit(self, cb, wait=True):
        cbs = (vk.Handle * 1)(cb)
        si = vk.VkSubmitInfo(vk.ST_SUBMIT_INFO, None, 0, None, None, 1, cbs, 0, None)
        fences = (vk.Handle * 1)(self.fence)
        if wait:
            self.d.check("vkResetFences", self.device, 1, fences)
            self.d.check("vkQueueSubmit", self.queue, 1, _ref(si), self.fence)
            self.d.check("vkWaitForFences", self.device, 1, fences, 1, 2**63)
        else:
            self.d.check("vkQueueSubmit", self.queue, 1, _ref(si), None)

    def wait_idle(self):
        self.d.check("vkQueueWaitIdle", self.queue)

    def stats(self):
        c = (ctypes.c_uint64 * 8)()
        ms = ctypes.c_float()
        nb, nl = ctypes.c_uint32(), ctypes.c_uint32()
        self.d.check("vkGetDrawStatsKAZAN", self.device, c, _ref(ms), _ref(nb), _ref(nl))
        out = dict(zip(vk.STAT_NAMES, list(c)))
        out.update(gpu_ms=ms.value, batches=nb.value, launches=nl.value)
        return out

    def set_kernel_timing(self, on):
        self.d.vkSetKernelTimingKAZAN(self.device, 1 if on else 0)

    def kernel_times(self):
        """{kernel: (summed ms, launches)} of the submissions since the last read."""
        ms = (ctypes.c_float * len(vk.KERNEL_NAMES))()
        n = (ctypes.c_uint32 * len(vk.KERNEL_NAMES))()
        self.d.check("vkGetKernelTimesKAZAN", self.device, ms, n)
        return {k: (ms[i], n[i]) for i, k in enumerate(vk.KERNEL_NAMES)}

    def upload(self, dst_buf, data):
        data = np.ascontiguousarray(data)
        n = data.nbytes
        stage = self.buffer(n, True)
        ctypes.memmove(self.map(stage[1], n), data.ctypes.data, n)
        cb = self.command_buffer()
        self.begin(cb)
        self.d.vkCmdCopyBuffer(cb, stage[0], dst_buf, 1, _ref(vk.VkBufferCopy(0, 0, n)))
        self.d.check("vkEndCommandBuffer", cb)
        self.submit(cb)
        self.d.vkFreeCommandBuffers(self.device, self.pool, 1, (vk.Handle * 1)(cb))
        self.destroy_buffer(stage)

    def device_buffer_with(self, data):
        data = np.ascontiguousarray(data)
        bm = self.buffer(data.nbytes, False)
        self.upload(bm[0], data)
        return bm


class SceneRenderer:
    """Everything one Scene needs; render() records nothing new.
    `before_pass(renderer, cb)`, if given, records commands into the frame's
    command buffer ahead of the render pass (a dispatch that writes the
    vertex buffer, say)."""

    def __init__(self, ctx, scene, before_pass=None):
        self.ctx, self.s = ctx, scene
        self.before_pass = before_pass
        d, dev = ctx.d, ctx.device
        s = scene
        self.keep = []
        W, H, S = s.width, s.height, s.samples
        self.vbuf = ctx.device_buffer_with(s.vertices)
        self.ibuf = ctx.device_buffer_with(s.indices) if s.indices is not None else None
        # attachments (+ a single-sample resolve target for 4x)
        self.color = ctx.image(W, H, vk.FORMAT_R8G8B8A8_UNORM, S)
        self.depth = ctx.image(W, H, vk.FORMAT_D32_SFLOAT, S)
        self.resolve = ctx.image(W, H, vk.FORMAT_R8G8B8A8_UNORM, 1) if S > 1 else None
        # descriptors
        set_layouts = []
        self.desc_set = None
        binding = None
        if s.fs == g.FS_TEXTURE:
            th, tw = s.texture.shape[:2]
            self.tex = ctx.image(tw, th, vk.FORMAT_R8G8B8A8_UNORM, 1)
            self._upload_image(self.tex[0], s.texture, tw, th)
            sci = vk.VkSamplerCreateInfo(vk.ST_SAMPLER_CREATE_INFO, None, 0, s.tex_filter, s.tex_filter, 0,
                                         s.tex_addr[0], s.tex_addr[1], 2, 0.0, 0, 1.0, 0, 0, 0.0, 0.0, 0, 0)
            self.sampler = vk.Handle()
            d.check("vkCreateSampler", dev, _ref(sci), None, _ref(self.sampler))
            binding = vk.VkDescriptorSetLayoutBinding(0, vk.DESCRIPTOR_TYPE_COMBINED_IMAGE_SAMPLER, 1,
                                                      vk.SHADER_STAGE_FRAGMENT, None)
        elif s.fs == g.FS_PHONG:
            self.ubo = ctx.device_buffer_with(s.ubo)
            binding = vk.VkDescriptorSetLayoutBinding(0, vk.DESCRIPTOR_TYPE_UNIFORM_BUFFER, 1,
                                                      vk.SHADER_STAGE_FRAGMENT, None)
        if binding is not None:
            dl = vk.VkDescriptorSetLayoutCreateInfo(vk.ST_DESCRIPTOR_SET_LAYOUT_CREATE_INFO, None, 0, 1,
                                                    ctypes.pointer(binding))
            self.set_layout = vk.Handle()
            d.check("vkCreateDescriptorSetLayout", dev, _ref(dl), None, _ref(self.set_layout))
            set_layouts.append(self.set_layout)
            ps = vk.VkDescriptorPoolSize(binding.descriptorType, 1)
            dpi = vk.VkDescriptorPoolCreateInfo(vk.ST_DESCRIPTOR_POOL_CREATE_INFO, None, 0, 1, 1, ctypes.pointer(ps))
            self.dpool = vk.Handle()
            d.check("vkCreateDescriptorPool", dev, _ref(dpi), None, _ref(self.dpool))
            layouts = (vk.Handle * 1)(self.set_layout)
            dai = vk.VkDescriptorSetAllocateInfo(vk.ST_DESCRIPTOR_SET_ALLOCATE_INFO, None, self.dpool, 1, layouts)
            self.desc_set = vk.Handle()
            d.check("vkAllocateDescriptorSets", dev, _ref(dai), _ref(self.desc_set))
            if s.fs == g.FS_TEXTURE:
                ii = vk.VkDescriptorImageInfo(self.sampler, self.tex[2], vk.IMAGE_LAYOUT_GENERAL)
                w = vk.VkWriteDescriptorSet(vk.ST_WRITE_DESCRIPTOR_SET, None, self.desc_set, 0, 0, 1,
                                            vk.DESCRIPTOR_TYPE_COMBINED_IMAGE_SAMPLER, ctypes.pointer(ii), None, None)
            else:
                bi = vk.VkDescriptorBufferInfo(self.ubo[0], 0, s.ubo.nbytes)
                w = vk.VkWriteDescriptorSet(vk.ST_WRITE_DESCRIPTOR_SET, None, self.desc_set, 0, 0, 1,
                                            vk.DESCRIPTOR_TYPE_UNIFORM_BUFFER, None, ctypes.pointer(bi), None)
            d.vkUpdateDescriptorSets(dev, 1, _ref(w), 0, None)
        push = vk.VkPushConstantRange(vk.SHADER_STAGE_VERTEX, 0, 64)
        sl = (vk.Handle * max(1, len(set_layouts)))(*set_layouts)
        plci = vk.VkPipelineLayoutCreateInfo(vk.ST_PIPELINE_LAYOUT_CREATE_INFO, None, 0, len(set_layouts), sl, 1,
                                             ctypes.pointer(push))
        self.layout = vk.Handle()
        d.check("vkCreatePipelineLayout", dev, _ref(plci), None, _ref(self.layout))
        self._render_pass()
        self._pipeline()
        self._record()

    def _upload_image(self, img, pixels, w, h):
        ctx = self.ctx
        n = pixels.nbytes
        stage = ctx.buffer(n, True)
        ctypes.memmove(ctx.map(stage[1], n), np.ascontiguousarray(pixels).ctypes.data, n)
        cb = ctx.command_buffer()
        ctx.begin(cb)
        region = vk.VkBufferImageCopy(0, 0, 0, vk.VkImageSubresourceLayers(1, 0, 0, 1), vk.VkOffset3D(0, 0, 0),
                                      vk.VkExtent3D(w, h, 1))
        ctx.d.vkCmdCopyBufferToImage(cb, stage[0], img, vk.IMAGE_LAYOUT_GENERAL, 1, _ref(region))
        ctx.d.check("vkEndCommandBuffer", cb)
        ctx.submit(cb)
        ctx.d.vkFreeCommandBuffers(ctx.device, ctx.pool, 1, (vk.Handle * 1)(cb))
        ctx.destroy_buffer(stage)

    def _render_pass(self):
        d, dev, s = self.ctx.d, self.ctx.device, self.s
        atts = [vk.VkAttachmentDescription(0, vk.FORMAT_R8G8B8A8_UNORM, s.samples, vk.ATTACHMENT_LOAD_OP_CLEAR, 0, 2, 1,
                                           0, vk.IMAGE_LAYOUT_GENERAL),
                vk.VkAttachmentDescription(0, vk.FORMAT_D32_SFLOAT, s.samples, vk.ATTACHMENT_LOAD_OP_CLEAR, 0, 2, 1, 0,
                                           vk.IMAGE_LAYOUT_GENERAL)]
        if s.samples > 1:
            atts.append(vk.VkAttachmentDescription(0, vk.FORMAT_R8G8B8A8_UNORM, 1, 2, 0, 2, 1, 0,
                                                   vk.IMAGE_LAYOUT_GENERAL))
        arr = (vk.VkAttachmentDescription * len(atts))(*atts)
        cref = vk.VkAttachmentReference(0, vk.IMAGE_LAYOUT_GENERAL)
        dref = vk.VkAttachmentReference(1, vk.IMAGE_LAYOUT_GENERAL)
        rref = vk.VkAttachmentReference(2, vk.IMAGE_LAYOUT_GENERAL)
        sub = vk.VkSubpassDescription(0, 0, 0, None, 1, ctypes.pointer(cref),
                                      ctypes.pointer(rref) if s.samples > 1 else None, ctypes.pointer(dref), 0, None)
        rpci = vk.VkRenderPassCreateInfo(vk.ST_RENDER_PASS_CREATE_INFO, None, 0, len(atts), arr, 1,
                                         ctypes.pointer(sub), 0, None)
        self.rp = vk.Handle()
        d.check("vkCreateRenderPass", dev, _ref(rpci), None, _ref(self.rp))
        views = [self.color[2], self.depth[2]] + ([self.resolve[2]] if self.resolve else [])
        varr = (vk.Handle * len(views))(*views)
        fci = vk.VkFramebufferCreateInfo(vk.ST_FRAMEBUFFER_CREATE_INFO, None, 0, self.rp, len(views), varr, s.width,
                                         s.height, 1)
        self.fb = vk.Handle()
        d.check("vkCreateFramebuffer", dev, _ref(fci), None, _ref(self.fb))

    def _shader(self, code):
        d, dev = self.ctx.d, self.ctx.device
        buf = ctypes.create_string_buffer(code, len(code))
        self.keep.append(buf)
        ci = vk.VkShaderModuleCreateInfo(vk.ST_SHADER_MODULE_CREATE_INFO, None, 0, len(code), ctypes.addressof(buf))
        m = vk.Handle()
        d.check("vkCreateShaderModule", dev, _ref(ci), None, _ref(m))
        return m

    def _pipeline(self):
        d, dev, s = self.ctx.d, self.ctx.device, self.s
        vs_fn, attrs = _VS[s.vs]
        vsm, fsm = self._shader(vs_fn()), self._shader(_FS[s.fs]())
        stages = (vk.VkPipelineShaderStageCreateInfo * 2)(
            vk.VkPipelineShaderStageCreateInfo(vk.ST_PIPELINE_SHADER_STAGE_CREATE_INFO, None, 0,
                                               vk.SHADER_STAGE_VERTEX, vsm, b"main", None),
            vk.VkPipelineShaderStageCreateInfo(vk.ST_PIPELINE_SHADER_STAGE_CREATE_INFO, None, 0,
                                               vk.SHADER_STAGE_FRAGMENT, fsm, b"main", None))
        bind = vk.VkVertexInputBindingDescription(0, s.stride * 4, 0)
        at = (vk.VkVertexInputAttributeDescription * len(attrs))(
            *[vk.VkVertexInputAttributeDescription(loc, 0, fmt, off) for loc, fmt, off in attrs])
        vis = vk.VkPipelineVertexInputStateCreateInfo(vk.ST_PIPELINE_VERTEX_INPUT_STATE_CREATE_INFO, None, 0, 1,
                                                      ctypes.pointer(bind), len(attrs), at)
        ias = vk.VkPipelineInputAssemblyStateCreateInfo(vk.ST_PIPELINE_INPUT_ASSEMBLY_STATE_CREATE_INFO, None, 0,
                                                        vk.TOPOLOGY_TRIANGLE_LIST, 0)
        vp = vk.VkViewport(*s.viewport)
        sc = vk.VkRect2D(vk.VkOffset2D(s.scissor[0], s.scissor[1]), vk.VkExtent2D(s.scissor[2], s.scissor[3]))
        vps = vk.VkPipelineViewportStateCreateInfo(vk.ST_PIPELINE_VIEWPORT_STATE_CREATE_INFO, None, 0, 1,
                                                   ctypes.pointer(vp), 1, ctypes.pointer(sc))
        rs = vk.VkPipelineRasterizationStateCreateInfo(vk.ST_PIPELINE_RASTERIZATION_STATE_CREATE_INFO, None, 0, 0, 0,
                                                       0, s.cull_mode, s.front_face, 0, 0.0, 0.0, 0.0, 1.0)
        ms = vk.VkPipelineMultisampleStateCreateInfo(vk.ST_PIPELINE_MULTISAMPLE_STATE_CREATE_INFO, None, 0, s.samples,
                                                     0, 0.0, None, 0, 0)
        ds = vk.VkPipelineDepthStencilStateCreateInfo(vk.ST_PIPELINE_DEPTH_STENCIL_STATE_CREATE_INFO, None, 0,
                                                      s.depth_test, s.depth_write, s.depth_op, 0, 0,
                                                      vk.VkStencilOpState(), vk.VkStencilOpState(), 0.0, 1.0)
        b = s.blend
        cba = vk.VkPipelineColorBlendAttachmentState(b[0], b[1], b[2], b[3], b[4], b[5], b[6], s.write_mask)
        cbs = vk.VkPipelineColorBlendStateCreateInfo(vk.ST_PIPELINE_COLOR_BLEND_STATE_CREATE_INFO, None, 0, 0, 0, 1,
                                                     ctypes.pointer(cba), (ctypes.c_float * 4)())
        gci = vk.VkGraphicsPipelineCreateInfo(vk.ST_GRAPHICS_PIPELINE_CREATE_INFO, None, 0, 2, stages,
                                              ctypes.pointer(vis), ctypes.pointer(ias), None, ctypes.pointer(vps),
                                              ctypes.pointer(rs), ctypes.pointer(ms), ctypes.pointer(ds),
                                              ctypes.pointer(cbs), None, self.layout, self.rp, 0, None, -1)
        self.pipeline = vk.Handle()
        d.check("vkCreateGraphicsPipelines", dev, self.ctx.pipeline_cache, 1, _ref(gci), None,
                _ref(self.pipeline))

    def _record(self):
        d, s = self.ctx.d, self.s
        cb = self.cb = self.ctx.command_buffer()
        self.ctx.begin(cb)
        clears = (vk.VkClearValue * 2)()
        for k in range(4):
            clears[0].color.float32[k] = s.clear_color[k]
        clears[1].depthStencil.depth = s.clear_depth
        ra = s.render_area or (0, 0, s.width, s.height)
        rbi = vk.VkRenderPassBeginInfo(vk.ST_RENDER_PASS_BEGIN_INFO, None, self.rp, self.fb,
                                       vk.VkRect2D(vk.VkOffset2D(ra[0], ra[1]), vk.VkExtent2D(ra[2], ra[3])), 2, clears)
        if self.before_pass is not None:
            self.before_pass(self, cb)
        d.vkCmdBeginRenderPass(cb, _ref(rbi), 0)
        d.vkCmdBindPipeline(cb, 0, self.pipeline)
        d.vkCmdBindVertexBuffers(cb, 0, 1, (vk.Handle * 1)(self.vbuf[0]), (ctypes.c_uint64 * 1)(0))
        if self.ibuf is not None:
            d.vkCmdBindIndexBuffer(cb, self.ibuf[0], 0, vk.INDEX_TYPE_UINT32)
        if self.desc_set is not None:
            d.vkCmdBindDescriptorSets(cb, 0, self.layout, 0, 1, (vk.Handle * 1)(self.desc_set), 0, None)
        for dr in s.draws:
            if dr.get("mvp") is not None:
                m = np.ascontiguousarray(dr["mvp"], dtype=np.float32)
                d.vkCmdPushConstants(cb, self.layout, vk.SHADER_STAGE_VERTEX, 0, 64, m.ctypes.data)
            if self.ibuf is not None:
                d.vkCmdDrawIndexed(cb, dr["count"], 1, dr["first"], dr["vertex_offset"], 0)
            else:
                d.vkCmdDraw(cb, dr["count"], 1, dr["first"], 0)
        d.vkCmdEndRenderPass(cb)
        d.check("vkEndCommandBuffer", cb)

    def render(self, frames=1):
        """Submits the recorded frame `frames` times, waits, returns stats."""
        for i in range(frames):
            self.ctx.submit(self.cb, wait=(i == frames - 1))
        return self.ctx.stats()

    def read_back(self):
        """(colour (H,W,S,4) u8, depth (H,W,S) f32, resolved (H,W,4) u8)."""
        ctx, s = self.ctx, self.s
        W, H, S = s.width, s.height, s.samples
        imgs = [(self.color[0], W * H * S * 4), (self.depth[0], W * H * S * 4)]
        if self.resolve:
            imgs.append((self.resolve[0], W * H * 4))
        outs = []
        for img, n in imgs:
            stage = ctx.buffer(n, True)
            cb = ctx.command_buffer()
            ctx.begin(cb)
            region = vk.VkBufferImageCopy(0, 0, 0, vk.VkImageSubresourceLayers(1, 0, 0, 1), vk.VkOffset3D(0, 0, 0),
                                          vk.VkExtent3D(W, H, 1))
            ctx.d.vkCmdCopyImageToBuffer(cb, img, vk.IMAGE_LAYOUT_GENERAL, stage[0], 1, _ref(region))
            ctx.d.check("vkEndCommandBuffer", cb)
            ctx.submit(cb)
            buf = (ctypes.c_uint8 * n).from_address(ctx.map(stage[1], n))
            outs.append(np.frombuffer(bytes(buf), dtype=np.uint8).copy())
            ctx.d.vkFreeCommandBuffers(ctx.device, ctx.pool, 1, (vk.Handle * 1)(cb))
            ctx.destroy_buffer(stage)
        color = outs[0].reshape(H, W, S, 4)
        depth = outs[1].view(np.float32).reshape(H, W, S)
        resolved = outs[2].reshape(H, W, 4) if self.resolve else color[:, :, 0, :].copy()
        return color, depth, resolved


def pipeline_desc(s, tile_shift=4):
    """The KazanPipelineDesc of the pipeline SceneRenderer creates for scene
    `s` (same shaders, vertex input and fixed-function state), for the
    device-less compiler entry point. Returns (desc, keep-alive buffers)."""
    vs_fn, attrs = _VS[s.vs]
    vs = np.frombuffer(vs_fn(), dtype=np.uint32).copy()
    fs = np.frombuffer(_FS[s.fs](), dtype=np.uint32).copy()
    d = vk.KazanPipelineDesc()
    d.vs_words, d.vs_nwords = vs.ctypes.data, vs.size
    d.fs_words, d.fs_nwords = fs.ctypes.data, fs.size
    d.num_bindings = 1
    d.binding_stride[0] = s.stride * 4
    d.num_attributes = len(attrs)
    for i, (loc, fmt, off) in enumerate(attrs):
        d.attr_location[i], d.attr_binding[i], d.attr_format[i], d.attr_offset[i] = loc, 0, fmt, off
    d.tile_shift, d.samples, d.cull_mode, d.front_face = tile_shift, s.samples, s.cull_mode, s.front_face
    d.depth_test, d.depth_write, d.depth_op = s.depth_test, s.depth_write, s.depth_op
    (d.blend, d.src_color, d.dst_color, d.color_op, d.src_alpha, d.dst_alpha, d.alpha_op) = s.blend
    d.write_mask = s.write_mask
    if s.fs == g.FS_TEXTURE:
        d.num_textures, d.texture_key[0] = 1, 0
    if s.fs == g.FS_PHONG:
        d.num_ubos, d.ubo_key[0] = 1, 0
    return d, (vs, fs)


def compile_offline(driver, s, what=2, tile_shift=4):
    """Runs pipeline compilation for scene `s` without a device: what = 0
    generated shader source, 1 full kernel source, 2 code object (through the
    on-disk cache; the text reports its size). Returns (rc, text)."""
    d, keep = pipeline_desc(s, tile_shift)
    return vk.tool_call(driver, "kazan_compile_pipeline_spirv", ctypes.byref(d), ctypes.c_uint32(what))
