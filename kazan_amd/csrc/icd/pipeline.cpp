// pipeline.cpp — vkCreateGraphicsPipelines / vkCreateComputePipelines:
// state capture, SPIR-V -> IR -> HIP, kernel compilation and loading.
//
// The reference collects the shader stages with get_shader_stages!
// (vulkan-driver/src/pipeline.rs:69-108: vertex required, fragment optional,
// a repeated stage is an error) and specialisations (:46-67), then stops at
// unimplemented!() (:182). This continues from there: it captures the fixed
// function state the draw path needs and compiles the stages into the
// pipeline's binning and raster kernels.
#include <cstdlib>
#include <cstring>
#include <string>

#include "../shader/backend_hip.h"
#include "../shader/translate.h"
#include "objects.h"

namespace kz {

namespace {

uint32_t env_tile_shift() {
    if (const char *e = std::getenv("KAZAN_TILE_SHIFT")) {
        const int v = std::atoi(e);
        if (v >= 3 && v <= 6) return (uint32_t)v;
    }
    return 4;
}

std::vector<SpecConstant> specs(const VkSpecializationInfo *si) {
    std::vector<SpecConstant> out;
    if (!si) return out;
    for (uint32_t i = 0; i < si->mapEntryCount; ++i) {
        const VkSpecializationMapEntry &e = si->pMapEntries[i];
        SpecConstant s;
        s.id = e.constantID;
        const uint8_t *p = (const uint8_t *)si->pData + e.offset;
        s.bytes.assign(p, p + e.size);
        out.push_back(s);
    }
    return out;
}

}  // namespace

void build_slot_maps(const PipelineLayout *layout, std::map<std::pair<uint32_t, uint32_t>, uint32_t> &ubo,
                     std::map<std::pair<uint32_t, uint32_t>, uint32_t> &tex,
                     std::map<std::pair<uint32_t, uint32_t>, uint32_t> *ssbo) {
    if (!layout) return;
    uint32_t nu = 0, nt = 0, ns = 0;
    for (uint32_t s = 0; s < layout->set_layouts.size(); ++s) {
        const DescriptorSetLayout *l = layout->set_layouts[s];
        if (!l) continue;
        for (const auto &b : l->bindings) {
            if (b.type == VK_DESCRIPTOR_TYPE_UNIFORM_BUFFER && nu < KZ_MAX_UBOS) ubo[{s, b.binding}] = nu++;
            if (b.type == VK_DESCRIPTOR_TYPE_COMBINED_IMAGE_SAMPLER && nt < KZ_MAX_TEXTURES) tex[{s, b.binding}] = nt++;
            if (ssbo && b.type == VK_DESCRIPTOR_TYPE_STORAGE_BUFFER && ns < KZ_MAX_SSBOS) (*ssbo)[{s, b.binding}] = ns++;
        }
    }
}

bool create_graphics_pipeline(Device *dev, const VkGraphicsPipelineCreateInfo *ci, PipelineCache *cache,
                              Pipeline **out, std::string &err) {
    (void)dev;
    const VkPipelineShaderStageCreateInfo *vs = nullptr, *fs = nullptr;
    for (uint32_t i = 0; i < ci->stageCount; ++i) {
        const VkPipelineShaderStageCreateInfo &s = ci->pStages[i];
        const VkPipelineShaderStageCreateInfo **slot = nullptr;
        if (s.stage == VK_SHADER_STAGE_VERTEX_BIT) slot = &vs;
        else if (s.stage == VK_SHADER_STAGE_FRAGMENT_BIT) slot = &fs;
        else {
            err = "only vertex and fragment stages are supported";
            return false;
        }
        if (*slot) {
            err = "duplicate shader stage";
            return false;
        }
        *slot = &s;
    }
    if (!vs) {
        err = "a graphics pipeline needs a vertex shader";
        return false;
    }
    if (!ci->pInputAssemblyState || ci->pInputAssemblyState->topology != VK_PRIMITIVE_TOPOLOGY_TRIANGLE_LIST) {
        err = "only VK_PRIMITIVE_TOPOLOGY_TRIANGLE_LIST is supported";
        return false;
    }
    const VkPipelineRasterizationStateCreateInfo *rs = ci->pRasterizationState;
    if (!rs || rs->polygonMode != VK_POLYGON_MODE_FILL || rs->rasterizerDiscardEnable || rs->depthClampEnable ||
        rs->depthBiasEnable) {
        err = "unsupported rasterization state (fill mode only, no discard, depth clamp or bias)";
        return false;
    }
    auto p = std::make_unique<Pipeline>();
    FixedState &fx = p->fixed;
    fx.tile_shift = env_tile_shift();
    // VK_PIPELINE_CREATE_DISABLE_OPTIMIZATION_BIT: the backend's
    // NoOptimizations mode (reference pipeline.rs:110-120)
    fx.optimize = (ci->flags & 0x1u) == 0;
    fx.cull_mode = rs->cullMode;
    fx.front_face = rs->frontFace;
    fx.samples = ci->pMultisampleState ? (uint32_t)ci->pMultisampleState->rasterizationSamples : 1u;
    if (fx.samples != 1 && fx.samples != 4) {
        err = "sample count must be 1 or 4";
        return false;
    }
    // the largest tile whose raster workgroup fits LDS at this sample count
    while (fx.tile_shift > kMinTileShift && raster_lds_bytes(fx.tile_shift, fx.samples) > 160u * 1024u)
        --fx.tile_shift;
    if (const VkPipelineDepthStencilStateCreateInfo *ds = ci->pDepthStencilState) {
        fx.depth_test = ds->depthTestEnable ? 1u : 0u;
        fx.depth_write = ds->depthWriteEnable ? 1u : 0u;
        fx.depth_op = ds->depthCompareOp;
    }
    if (const VkPipelineColorBlendStateCreateInfo *cb = ci->pColorBlendState) {
        if (cb->attachmentCount >= 1) {
            const VkPipelineColorBlendAttachmentState &a = cb->pAttachments[0];
            fx.blend = a.blendEnable ? 1u : 0u;
            fx.src_color = a.srcColorBlendFactor;
            fx.dst_color = a.dstColorBlendFactor;
            fx.color_op = a.colorBlendOp;
            fx.src_alpha = a.srcAlphaBlendFactor;
            fx.dst_alpha = a.dstAlphaBlendFactor;
            fx.alpha_op = a.alphaBlendOp;
            fx.write_mask = a.colorWriteMask;
            const uint32_t f[4] = {a.srcColorBlendFactor, a.dstColorBlendFactor, a.srcAlphaBlendFactor,
                                   a.dstAlphaBlendFactor};
            for (uint32_t k : f)
                if (fx.blend && k >= 10 && k <= 13) {
                    err = "constant blend factors are not supported";
                    return false;
                }
        }
    }
    if (const VkPipelineDynamicStateCreateInfo *dy = ci->pDynamicState)
        for (uint32_t i = 0; i < dy->dynamicStateCount; ++i) {
            if (dy->pDynamicStates[i] == VK_DYNAMIC_STATE_VIEWPORT) p->dynamic_viewport = true;
            else if (dy->pDynamicStates[i] == VK_DYNAMIC_STATE_SCISSOR) p->dynamic_scissor = true;
        }
    if (const VkPipelineViewportStateCreateInfo *vp = ci->pViewportState) {
        if (!p->dynamic_viewport && vp->pViewports && vp->viewportCount) p->viewport = vp->pViewports[0];
        if (!p->dynamic_scissor && vp->pScissors && vp->scissorCount) p->scissor = vp->pScissors[0];
    }
    p->layout = reinterpret_cast<PipelineLayout *>(ci->layout);
    p->subpass = ci->subpass;
    build_slot_maps(p->layout, p->ubo_slot, p->tex_slot, nullptr);

    ir::Module vmod, fmod;
    const ShaderModule *vsm = reinterpret_cast<const ShaderModule *>(vs->module);
    if (!translate_spirv(vsm->code.data(), vsm->code.size(), vs->pName ? vs->pName : "main", ir::Stage::Vertex,
                         specs(vs->pSpecializationInfo), vmod, err)) {
        err = "vertex shader: " + err;
        return false;
    }
    if (fs) {
        const ShaderModule *fsm = reinterpret_cast<const ShaderModule *>(fs->module);
        if (!translate_spirv(fsm->code.data(), fsm->code.size(), fs->pName ? fs->pName : "main", ir::Stage::Fragment,
                             specs(fs->pSpecializationInfo), fmod, err)) {
            err = "fragment shader: " + err;
            return false;
        }
    }
    PipelineCodegenInput cg;
    cg.vs = &vmod;
    cg.fs = fs ? &fmod : nullptr;
    if (const VkPipelineVertexInputStateCreateInfo *vi = ci->pVertexInputState) {
        for (uint32_t i = 0; i < vi->vertexBindingDescriptionCount; ++i) {
            const auto &b = vi->pVertexBindingDescriptions[i];
            if (b.binding >= KZ_MAX_BINDINGS) {
                err = "vertex binding index out of range";
                return false;
            }
            cg.bindings.push_back({b.binding, b.stride, (uint32_t)b.inputRate});
        }
        for (uint32_t i = 0; i < vi->vertexAttributeDescriptionCount; ++i) {
            const auto &a = vi->pVertexAttributeDescriptions[i];
            cg.attributes.push_back({a.location, a.binding, (uint32_t)a.format, a.offset});
        }
    }
    cg.ubo_slot = p->ubo_slot;
    cg.tex_slot = p->tex_slot;
    PipelineCodegenOutput co;
    if (!generate_pipeline_source(cg, co, err)) return false;
    if (co.num_varyings > KZ_MAX_VARYINGS) {
        err = "too many varyings";
        return false;
    }
    fx.num_varyings = co.num_varyings;
    p->shader_source = co.source;
    p->variants.fixed = fx;
    p->variants.shader_source = co.source;
    // code objects the application's pipeline cache holds for this pipeline
    if (cache) {
        std::lock_guard<std::mutex> lock(cache->mu);
        for (uint32_t s = kMinTileShift; s <= kMaxTileShift; ++s) {
            if (!p->variants.fits(s)) continue;
            FixedState fs = fx;
            fs.tile_shift = s;
            const std::string name = pipeline_code_name(fs, co.source);
            for (const auto &e : cache->entries)
                if (e.first == name) p->variants.preload(s, e.second);
        }
    }
    // the default tile size now; others when the autotuner asks for them
    if (!p->variants.get(fx.tile_shift, &err)) return false;
    *out = p.release();
    return true;
}

// vkCreateComputePipelines: reference ComputePipeline::create
// (vulkan-driver/src/pipeline.rs:122-154: the one compute stage, the
// layout, the optimisation mode), continued here to a HIP kernel.
bool create_compute_pipeline(Device *dev, const VkComputePipelineCreateInfo *ci, Pipeline **out, std::string &err) {
    (void)dev;
    const VkPipelineShaderStageCreateInfo &st = ci->stage;
    if (st.stage != VK_SHADER_STAGE_COMPUTE_BIT) {
        err = "a compute pipeline's stage must be VK_SHADER_STAGE_COMPUTE_BIT";
        return false;
    }
    auto p = std::make_unique<Pipeline>();
    p->compute = true;
    p->layout = reinterpret_cast<PipelineLayout *>(ci->layout);
    build_slot_maps(p->layout, p->ubo_slot, p->tex_slot, &p->ssbo_slot);
    ir::Module cmod;
    const ShaderModule *sm = reinterpret_cast<const ShaderModule *>(st.module);
    if (!translate_spirv(sm->code.data(), sm->code.size(), st.pName ? st.pName : "main", ir::Stage::Compute,
                         specs(st.pSpecializationInfo), cmod, err)) {
        err = "compute shader: " + err;
        return false;
    }
    PipelineCodegenInput cg;
    cg.ubo_slot = p->ubo_slot;
    cg.ssbo_slot = p->ssbo_slot;
    std::string source;
    if (!generate_compute_source(cmod, cg, source, err)) return false;
    std::vector<char> code;
    std::string log;
    if (!compile_compute_code(source, (ci->flags & 0x1u) == 0, code, log, nullptr)) {
        err = "kernel compilation failed:\n" + log;
        return false;
    }
    hipError_t e = hipModuleLoadData(&p->cs_module, code.data());
    if (e == hipSuccess) e = hipModuleGetFunction(&p->cs_kernel, p->cs_module, "kz_compute");
    if (e != hipSuccess) {
        err = std::string("loading the compute kernel: ") + hipGetErrorString(e);
        return false;
    }
    for (int k = 0; k < 3; ++k) p->local_size[k] = cmod.local_size[k];
    p->shader_source = source;
    *out = p.release();
    return true;
}

void destroy_pipeline(Pipeline *p) { delete p; }

}  // namespace kz
