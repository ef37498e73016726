// spirv_reader.h — decoding of SPIR-V binaries into an instruction list.
//
// Mirrors what spirv_parser::Parser::start does for the reference
// (spirv-parser/src/generated_parser.rs:16195): validate the 5-word header,
// then split the word stream into instructions by each first word's
// (word count << 16 | opcode). convert_bytes_to_words (:38-52) byte-swaps a
// module whose magic number reads reversed; spirv_words_from_bytes does the
// same.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace kz {

constexpr uint32_t kSpirvMagic = 0x07230203u;

struct SpirvInstruction {
    uint16_t opcode = 0;
    uint32_t word_offset = 0;        // offset of the first word in the module
    std::vector<uint32_t> operands;  // words after the first
};

struct SpirvModule {
    uint32_t version = 0, generator = 0, bound = 0, schema = 0;
    std::vector<SpirvInstruction> instructions;
};

// Returns false and sets err for a stream that is not whole words or whose
// magic number is wrong either way round.
bool spirv_words_from_bytes(const uint8_t *bytes, size_t n, std::vector<uint32_t> &words, std::string &err);

bool spirv_parse(const uint32_t *words, size_t nwords, SpirvModule &out, std::string &err);

// Decodes a literal string operand starting at operands[start]; sets `next`
// to the first word after it.
std::string spirv_string(const std::vector<uint32_t> &ops, size_t start, size_t *next = nullptr);

// A textual dump in the reference disassembler's layout (one instruction per
// line, result ids right-aligned, byte offset comment), for the opcodes and
// enumerants this path understands.
std::string spirv_disassemble(const SpirvModule &m);

}  // namespace kz
