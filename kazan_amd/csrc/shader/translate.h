// translate.h — SPIR-V -> shader IR.
//
// Counterpart of TranslatedSPIRVShader::new
// (shader-compiler-translate-spirv-to-ir/src/lib.rs:191-215): decode the
// module, find the entry point by name and execution model, build the
// interface blocks and translate the entry function. The reference's
// translator stops with todo!() at OpLoad/OpStore/OpAccessChain
// (parse/memory_instructions.rs:99,124,141) and at every arithmetic,
// composite and image opcode (parse/unimplemented_instructions.rs:40-513);
// this one implements the subset the draw path's shaders use and reports any
// other opcode as an error naming it.
#pragma once

#include <string>
#include <vector>

#include "ir.h"

namespace kz {

struct SpecConstant {
    uint32_t id;
    std::vector<uint8_t> bytes;
};

bool translate_spirv(const uint32_t *words, size_t nwords, const std::string &entry_name, ir::Stage stage,
                     const std::vector<SpecConstant> &spec, ir::Module &out, std::string &err);

}  // namespace kz
