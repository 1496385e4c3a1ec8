// Driver inputs from a TOML file and command-line overrides
// (reference: apps/driver_config.hpp:22-48, apps/driver_config.cpp).
// Keys: run_name, dataset_path (--dataset), queries (--query, -q),
// file_reader.type, index_builder.{type,file_partition_size,grid_size,
// block_size,reduction_grid_size,reduction_block_size},
// query_executor.{grid_size,block_size}.  CLI names replace '_' with '-'.
#pragma once

#include "gpjson/engine.hpp"

#include <exception>
#include <stdexcept>
#include <string>
#include <vector>

namespace gpjson::driver {

struct DriverInputs {
  std::string run_name;
  std::string dataset_path;
  std::vector<std::string> queries;
  gpjson::EngineOptions engine_options;
};

class CliExit final : public std::exception {
public:
  explicit CliExit(int exit_code) : exit_code_(exit_code) {}
  const char *what() const noexcept override { return "CLI requested exit"; }
  int exit_code() const noexcept { return exit_code_; }

private:
  int exit_code_;
};

DriverInputs default_driver_inputs();
DriverInputs parse_driver_inputs(int argc, char **argv);
void validate_driver_inputs(const DriverInputs &inputs);
std::string engine_options_to_json(const gpjson::EngineOptions &options);
std::string driver_inputs_to_json(const DriverInputs &inputs);

template <class Enum> Enum parse_enum_value(const std::string &text);
template <> file::FileReaderType parse_enum_value<file::FileReaderType>(const std::string &text);
template <>
index::IndexBuilderType parse_enum_value<index::IndexBuilderType>(const std::string &text);

} // namespace gpjson::driver
