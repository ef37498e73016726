// tools.cpp — the shader-compiler entry points of kazan_tools.h.
#include "kazan/kazan_tools.h"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../engine/engine.h"
#include "../engine/kernels.h"
#include "../shader/backend_hip.h"
#include "../shader/ir.h"
#include "../shader/spirv_reader.h"
#include "../shader/translate.h"

namespace {

int put(const std::string &s, char *out, size_t cap, size_t *out_len, int rc) {
    if (out_len) *out_len = s.size();
    if (out && cap) {
        const size_t n = s.size() < cap - 1 ? s.size() : cap - 1;
        std::memcpy(out, s.data(), n);
        out[n] = 0;
    }
    return rc;
}

kz::ir::Stage stage_of(uint32_t s) {
    return s == 0 ? kz::ir::Stage::Vertex : s == 4 ? kz::ir::Stage::Fragment : kz::ir::Stage::Compute;
}

}  // namespace

extern "C" int kazan_spirv_disassemble(const uint32_t *words, size_t nwords, char *out, size_t cap, size_t *out_len) {
    kz::SpirvModule m;
    std::string err;
    // as the module's bytes (either byte order), like vkCreateShaderModule
    std::vector<uint32_t> w;
    if (!kz::spirv_words_from_bytes((const uint8_t *)words, nwords * 4, w, err) ||
        !kz::spirv_parse(w.data(), w.size(), m, err))
        return put(err, out, cap, out_len, -1);
    return put(kz::spirv_disassemble(m), out, cap, out_len, 0);
}

extern "C" int kazan_translate_to_ir_text(const uint32_t *words, size_t nwords, uint32_t stage, const char *entry,
                                          char *out, size_t cap, size_t *out_len) {
    kz::ir::Module m;
    std::string err;
    if (!kz::translate_spirv(words, nwords, entry ? entry : "main", stage_of(stage), {}, m, err))
        return put(err, out, cap, out_len, -1);
    return put(kz::ir::module_text(m), out, cap, out_len, 0);
}

extern "C" int kazan_ir_round_trip(const char *text, uint32_t what, char *out, size_t cap, size_t *out_len) {
    if (!text) return put("no text", out, cap, out_len, -1);
    kz::ir::Module m;
    std::string err;
    if (what == 0) {
        if (!kz::ir::parse_module_text(text, m, err)) return put(err, out, cap, out_len, -1);
        return put(kz::ir::module_text(m), out, cap, out_len, 0);
    }
    kz::ir::Instruction ins;
    if (!kz::ir::parse_instruction_text(text, m, ins, err)) return put(err, out, cap, out_len, -1);
    return put(kz::ir::instruction_text(m, ins), out, cap, out_len, 0);
}

namespace kz {

// A KazanPipelineDesc's shaders translated and lowered to HIP source, and
// its fixed-function state (shared by the offline compile tool and the
// backend entry points of kazan_backend.h).
bool pipeline_from_desc(const KazanPipelineDesc *d, FixedState &fx, std::string &source, std::string &err) {
    ir::Module vm, fm;
    if (!translate_spirv(d->vs_words, d->vs_nwords, "main", ir::Stage::Vertex, {}, vm, err)) {
        err = "vertex shader: " + err;
        return false;
    }
    if (d->fs_words && !translate_spirv(d->fs_words, d->fs_nwords, "main", ir::Stage::Fragment, {}, fm, err)) {
        err = "fragment shader: " + err;
        return false;
    }
    PipelineCodegenInput in;
    in.vs = &vm;
    in.fs = d->fs_words ? &fm : nullptr;
    for (uint32_t i = 0; i < d->num_bindings && i < 8; ++i) in.bindings.push_back({i, d->binding_stride[i], d->binding_rate[i]});
    for (uint32_t i = 0; i < d->num_attributes && i < 16; ++i)
        in.attributes.push_back({d->attr_location[i], d->attr_binding[i], d->attr_format[i], d->attr_offset[i]});
    for (uint32_t i = 0; i < d->num_ubos && i < 4; ++i) in.ubo_slot[{d->ubo_key[i] >> 16, d->ubo_key[i] & 0xFFFF}] = i;
    for (uint32_t i = 0; i < d->num_textures && i < 4; ++i)
        in.tex_slot[{d->texture_key[i] >> 16, d->texture_key[i] & 0xFFFF}] = i;
    PipelineCodegenOutput co;
    if (!generate_pipeline_source(in, co, err)) return false;
    source = co.source;
    fx.tile_shift = d->tile_shift ? d->tile_shift : 4;
    fx.samples = d->samples ? d->samples : 1;
    fx.num_varyings = co.num_varyings;
    fx.cull_mode = d->cull_mode;
    fx.front_face = d->front_face;
    fx.depth_test = d->depth_test;
    fx.depth_write = d->depth_write;
    fx.depth_op = d->depth_op;
    fx.blend = d->blend;
    fx.src_color = d->src_color;
    fx.dst_color = d->dst_color;
    fx.color_op = d->color_op;
    fx.src_alpha = d->src_alpha;
    fx.dst_alpha = d->dst_alpha;
    fx.alpha_op = d->alpha_op;
    fx.write_mask = d->write_mask;
    return true;
}

}  // namespace kz

extern "C" int kazan_compile_pipeline_spirv(const KazanPipelineDesc *d, uint32_t what, char *out, size_t cap,
                                            size_t *out_len) {
    kz::FixedState fx;
    std::string source, err;
    if (!kz::pipeline_from_desc(d, fx, source, err)) return put(err, out, cap, out_len, -1);
    if (what == 0) return put(source, out, cap, out_len, 0);
    if (what == 1) return put(kz::pipeline_full_source(fx, source), out, cap, out_len, 0);
    std::vector<char> code;
    std::string log;
    bool hit = false;
    if (!kz::compile_pipeline_code(fx, source, code, log, &hit)) return put(log, out, cap, out_len, -1);
    std::string rep = "code_object_bytes=" + std::to_string(code.size()) + " cache_hit=" + (hit ? "1" : "0") +
                      " varyings=" + std::to_string(fx.num_varyings);
    for (const char *k : {"kz_bin_count", "kz_raster", "kz_raster_persistent"}) {
        kz::KernelResources r;
        if (!kz::kernel_resources(code, k, r)) continue;
        rep += std::string(" ") + k + "_vgprs=" + std::to_string(r.vgprs) + " " + k + "_scratch=" +
               std::to_string(r.scratch) + " " + k + "_lds=" + std::to_string(r.lds) + " " + k +
               "_waves=" + std::to_string(r.waves);
    }
    return put(rep, out, cap, out_len, 0);
}

extern "C" int kazan_compile_compute_spirv(const uint32_t *words, size_t nwords, const char *entry, uint32_t what,
                                           char *out, size_t cap, size_t *out_len) {
    kz::ir::Module m;
    std::string err;
    if (!words) return put("null module", out, cap, out_len, -1);
    if (!kz::translate_spirv(words, nwords, entry ? entry : "main", kz::ir::Stage::Compute, {}, m, err))
        return put(err, out, cap, out_len, -1);
    // storage buffers and uniform buffers take slots in (set, binding) order,
    // as a pipeline layout listing them in that order gives them
    kz::PipelineCodegenInput cg;
    std::vector<std::pair<uint32_t, uint32_t>> ssbos, ubos;
    for (const auto &r : m.resources) {
        if (r.kind == kz::ir::ResourceKind::StorageBuffer) ssbos.push_back({r.set, r.binding});
        if (r.kind == kz::ir::ResourceKind::UniformBuffer) ubos.push_back({r.set, r.binding});
    }
    std::sort(ssbos.begin(), ssbos.end());
    std::sort(ubos.begin(), ubos.end());
    for (size_t i = 0; i < ssbos.size() && i < KZ_MAX_SSBOS; ++i) cg.ssbo_slot[ssbos[i]] = (uint32_t)i;
    for (size_t i = 0; i < ubos.size() && i < KZ_MAX_UBOS; ++i) cg.ubo_slot[ubos[i]] = (uint32_t)i;
    std::string source;
    if (!kz::generate_compute_source(m, cg, source, err)) return put(err, out, cap, out_len, -1);
    if (what == 0) return put(source, out, cap, out_len, 0);
    if (what == 1) return put(kz::compute_full_source(source), out, cap, out_len, 0);
    std::vector<char> code;
    std::string log;
    bool hit = false;
    if (!kz::compile_compute_code(source, true, code, log, &hit)) return put(log, out, cap, out_len, -1);
    return put("code_object_bytes=" + std::to_string(code.size()) + " cache_hit=" + (hit ? "1" : "0") +
                   " local_size=" + std::to_string(m.local_size[0]) + "x" + std::to_string(m.local_size[1]) + "x" +
                   std::to_string(m.local_size[2]),
               out, cap, out_len, 0);
}

// --- the learned tile-cost model's kernels, outside a Vulkan device ------

static_assert(KAZAN_COST_PARAMS == KZ_PARAMS, "parameter count of kazan_tools.h");
static_assert(sizeof(KzCostModel) == 19616, "cost_model.bin layout (kazan_amd/balancer writes it)");

namespace {

struct DeviceBuffers {
    std::vector<void *> ptrs;
    ~DeviceBuffers() {
        for (void *p : ptrs) hipFree(p);
    }
    template <class T>
    T *alloc(size_t count) {
        void *p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) return nullptr;
        ptrs.push_back(p);
        return (T *)p;
    }
};

}  // namespace

extern "C" int kazan_cost_model_eval(const void *blob, size_t bytes, const float *features, uint32_t n, float *out) {
    KzCostModel m;
    if (!blob || bytes != sizeof m || (!features && n) || (!out && n)) return -1;
    std::memcpy(&m, blob, sizeof m);
    if (m.magic != KZ_COST_MODEL_MAGIC || m.version != 1 || m.bytes != sizeof m) return -2;
    if (n == 0) return 0;
    DeviceBuffers b;
    KzCostModel *dm = b.alloc<KzCostModel>(1);
    float *df = b.alloc<float>((size_t)n * KZ_FEATURES), *dp = b.alloc<float>(n);
    if (!dm || !df || !dp) return -3;
    if (hipMemcpy(dm, &m, sizeof m, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(df, features, (size_t)n * KZ_FEATURES * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        kz_launch_cost_mlp(nullptr, dm, df, n, dp) != hipSuccess ||
        hipMemcpy(out, dp, n * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
        return -4;
    return 0;
}

extern "C" int kazan_cost_model_train(const float *x, const float *y, uint32_t n, float *state, uint32_t steps,
                                      uint32_t step0, uint32_t batch, float lr, uint64_t seed, float *loss_out) {
    if (!x || !y || !state || n == 0 || batch == 0 || batch % 64 != 0) return -1;
    DeviceBuffers b;
    float *dx = b.alloc<float>((size_t)n * KZ_FEATURES), *dy = b.alloc<float>(n);
    float *ds = b.alloc<float>(3 * (size_t)KZ_PARAMS), *dl = b.alloc<float>(steps);
    if (!dx || !dy || !ds || !dl) return -3;
    if (hipMemcpy(dx, x, (size_t)n * KZ_FEATURES * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dy, y, n * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(ds, state, 3 * (size_t)KZ_PARAMS * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
        return -4;
    // runs of at most 1000 steps per launch
    for (uint32_t done = 0; done < steps;) {
        KzTrainParams p{};
        p.lr = lr;
        p.beta1 = 0.9f;
        p.beta2 = 0.999f;
        p.eps = 1e-8f;
        p.steps = std::min<uint32_t>(steps - done, 1000u);
        p.step0 = step0 + done;
        p.batch = batch;
        p.n = n;
        p.seed = seed;
        if (kz_launch_cost_train(nullptr, ds, dx, dy, p, dl + done) != hipSuccess) return -4;
        done += p.steps;
    }
    if (hipMemcpy(state, ds, 3 * (size_t)KZ_PARAMS * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) return -4;
    if (loss_out && steps && hipMemcpy(loss_out, dl, steps * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
        return -4;
    return 0;
}
