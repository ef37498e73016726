/*
 * kazan_tools.h — shader-compiler entry points exported beside the ICD for
 * tests and tooling. They need no GPU and no Vulkan objects.
 *
 *   kazan_spirv_disassemble      counterpart of the reference's dump of a
 *                                parsed module (spirv-parser/src/lib.rs:45-216
 *                                golden tests)
 *   kazan_translate_to_ir_text   counterpart of TranslatedSPIRVShader::new
 *                                (shader-compiler-translate-spirv-to-ir/src/
 *                                lib.rs:191-215) followed by the IR's text form
 *                                (assert_ir_matches_file!, shader-compiler-ir/
 *                                src/macros.rs:39-48)
 *   kazan_ir_round_trip          counterpart of the reference's IR text parser
 *                                (shader-compiler-ir/src/text.rs, used by
 *                                assert_ir_matches_file!, macros.rs:39-80, and
 *                                the instruction text tests,
 *                                instructions_impl.rs:560-622): parses IR text
 *                                and prints it back
 *   kazan_compile_pipeline_spirv runs the whole pipeline compile (SPIR-V -> IR
 *                                -> HIP source -> gfx950 code object) without
 *                                a device, as vkCreateGraphicsPipelines does
 *   kazan_compile_compute_spirv  the same for a compute shader, as
 *                                vkCreateComputePipelines does (reference
 *                                ComputePipeline::create, vulkan-driver/src/
 *                                pipeline.rs:122-154)
 *
 * All return 0 on success and a negative value on failure, with a message in
 * `out`. Output longer than `cap` is truncated; *out_len receives the full
 * length.
 */
#ifndef KAZAN_TOOLS_H
#define KAZAN_TOOLS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifndef KAZAN_EXPORT
#define KAZAN_EXPORT __attribute__((visibility("default")))
#endif

/* stage: 0 = vertex, 4 = fragment, 5 = compute (SPIR-V execution models) */
KAZAN_EXPORT int kazan_spirv_disassemble(const uint32_t *words, size_t nwords, char *out, size_t cap,
                                         size_t *out_len);
KAZAN_EXPORT int kazan_translate_to_ir_text(const uint32_t *words, size_t nwords, uint32_t stage, const char *entry,
                                            char *out, size_t cap, size_t *out_len);

typedef struct KazanPipelineDesc {
    const uint32_t *vs_words;
    size_t vs_nwords;
    const uint32_t *fs_words; /* may be NULL */
    size_t fs_nwords;
    /* vertex input: up to 8 bindings and 16 attributes */
    uint32_t num_bindings;
    uint32_t binding_stride[8];
    uint32_t binding_rate[8];
    uint32_t num_attributes;
    uint32_t attr_location[16], attr_binding[16], attr_format[16], attr_offset[16];
    /* fixed function */
    uint32_t tile_shift, samples, cull_mode, front_face;
    uint32_t depth_test, depth_write, depth_op;
    uint32_t blend, src_color, dst_color, color_op, src_alpha, dst_alpha, alpha_op, write_mask;
    /* descriptor slots: (set << 16 | binding) of up to 4 uniform buffers and
     * 4 combined image samplers, in slot order */
    uint32_t num_ubos, ubo_key[4];
    uint32_t num_textures, texture_key[4];
} KazanPipelineDesc;

/* The learned tile-cost model's kernels on the current HIP device (these
 * need a GPU; no counterpart in the reference, which has no draw path to
 * balance).
 *
 * kazan_cost_model_eval: the bf16 MFMA inference kernel of the balanced
 * pipeline over n rows of raw tile features ([n][16], the features
 * vkGetTileProfileKAZAN returns) with a KzCostModel blob (cost_model.bin's
 * bytes); writes the predicted fragments per row to out[n].
 *
 * kazan_cost_model_train: `steps` minibatch Adam steps (beta 0.9/0.999,
 * eps 1e-8) of the MFMA training kernel on the squared error between the
 * model output (log2(1 + fragments)) and y[n], from normalised features
 * x[n][16]. state = [weights | first moments | second moments], each
 * KAZAN_COST_PARAMS floats laid out w1 [64][16], w2 [64][64], w3 [64][64],
 * b1, b2, b3, w4 [64], b4; it is updated in place. step0 = steps already
 * taken (Adam bias correction, minibatch stream); batch = samples per
 * step, a multiple of 64; loss_out[steps] (may be null) = each step's mean
 * squared error.
 *
 * Both return 0, or a negative value on invalid input or a HIP error. */
#define KAZAN_COST_PARAMS 9473
KAZAN_EXPORT int kazan_cost_model_eval(const void *blob, size_t bytes, const float *features, uint32_t n,
                                       float *out);
KAZAN_EXPORT int kazan_cost_model_train(const float *x, const float *y, uint32_t n, float *state, uint32_t steps,
                                        uint32_t step0, uint32_t batch, float lr, uint64_t seed, float *loss_out);

/* kazan_cost_model_eval_noalloc: kazan_cost_model_eval with the same
 * arguments and results, for callers that cannot allocate device memory
 * (tight serving loops, code that forbids hipMalloc). It has no
 * counterpart in the reference. It calls no hipMalloc/hipFree: the model,
 * the features and the predictions are staged through one static device
 * array of KAZAN_NOALLOC_SCRATCH_BYTES per GPU, held until the process
 * ends, so n is limited to KAZAN_NOALLOC_MAX_ROWS (larger n returns -5 and
 * touches nothing). Calls on one GPU share that array: within a process
 * they are serialised by a lock; concurrent calls from several processes
 * each use their own copy. Returns the codes of kazan_cost_model_eval, or
 * -5 for n > KAZAN_NOALLOC_MAX_ROWS. */
#define KAZAN_NOALLOC_SCRATCH_BYTES (16u << 20)
#define KAZAN_NOALLOC_MAX_ROWS 245760u
KAZAN_EXPORT int kazan_cost_model_eval_noalloc(const void *blob, size_t bytes, const float *features, uint32_t n,
                                               float *out);

/* what: 0 = a module's text; 1 = value definitions "[name: type, ...]"
 * followed by one instruction, which is printed back alone */
KAZAN_EXPORT int kazan_ir_round_trip(const char *text, uint32_t what, char *out, size_t cap, size_t *out_len);

/* what: 0 = generated shader source, 1 = full kernel source, 2 = compile to a
 * code object and report its size, whether the cache held it, and per kernel
 * (kz_bin_count, kz_raster, kz_raster_persistent) <kernel>_vgprs, _scratch
 * (bytes per lane), _lds (bytes per workgroup) and _waves (resident waves per
 * SIMD) */
KAZAN_EXPORT int kazan_compile_pipeline_spirv(const KazanPipelineDesc *desc, uint32_t what, char *out, size_t cap,
                                              size_t *out_len);

/* A GLCompute entry point; storage and uniform buffers take kernel slots in
 * (set, binding) order. what: 0 = generated kernel body, 1 = full kernel
 * source, 2 = compile to a code object and report its size, whether the
 * cache held it and the workgroup size */
KAZAN_EXPORT int kazan_compile_compute_spirv(const uint32_t *words, size_t nwords, const char *entry, uint32_t what,
                                             char *out, size_t cap, size_t *out_len);

#ifdef __cplusplus
}
#endif

#endif /* KAZAN_TOOLS_H */
