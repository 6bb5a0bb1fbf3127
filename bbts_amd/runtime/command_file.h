// Readers for the reference's two command-file formats, so its own generated
// streams can drive this runtime unchanged.
//
// * Binary (.bbts): parsed_command_list_t::serialize / deserialize
//   (src/commands/parsed_command.h:194-383): u64 command count, then per command
//   i32 type (APPLY 0, REDUCE 1, MOVE 2, DELETE 3), ud name (u64 length + bytes),
//   is_gpu (1 byte), input and output type lists (u64 count, then strings),
//   inputs and outputs (u64 count, then the raw bytes of std::tuple<tid_t,
//   node_id_t>: libstdc++ lays a tuple out last element first, so each is
//   i32 node then i32 tid), parameters (u64 count, then 4-byte words).
// * Text (.sbbts): compile_source_file_t (src/commands/command_compiler.h:391-501,
//   977-1027): function specs "id name n_in types... n_out types...", then
//   commands "ud_id APPLY|REDUCE|DELETE n_in tids... n_out tids... n_params
//   (int|float|uint value)...", whitespace separated.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace bbts {

struct file_command_t {
  int32_t type = 0;  // 0 APPLY, 1 REDUCE, 2 MOVE, 3 DELETE
  std::string ud_name;
  bool is_gpu = false;
  bool gpu_unspecified = false;  // text files leave CPU / GPU to the runtime
  std::vector<std::string> input_types, output_types;
  std::vector<int32_t> inputs, outputs;         // tids
  std::vector<int32_t> input_nodes, output_nodes;  // node ids (binary files; 0 in text files)
  std::vector<uint32_t> params;
};

// Throw std::runtime_error with the byte offset / token of the first malformed field.
std::vector<file_command_t> parse_bbts_binary(const uint8_t *data, size_t len);
std::vector<file_command_t> parse_bbts_text(const char *data, size_t len);

}  // namespace bbts
