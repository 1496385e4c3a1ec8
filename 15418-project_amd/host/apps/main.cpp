// gpjson_driver: run queries over a dataset from a TOML config and/or the
// command line (reference: apps/main.cpp:16-50, same output markers).
#include "driver_config.hpp"

#include "gpjson/engine.hpp"

#include <iostream>

namespace {
void summarize(const gpjson::query::MaterializedBatchResult &r) {
  std::cout << "Result has " << r.queries().size() << " queries\n";
  for (const auto &q : r.queries()) {
    unsigned long matches = 0;
    for (const auto &line : q.lines()) matches += line.values().size();
    std::cout << "Query '" << q.query_text() << "' has " << q.lines().size()
              << " lines of result, and " << matches << " matched values\n";
  }
}
} // namespace

int main(int argc, char **argv) {
  try {
    const auto in = gpjson::driver::parse_driver_inputs(argc, argv);
    std::cout << "### RUN BEGIN run_name=" << in.run_name << " ###\n";
    std::cout << "GPJSON_RUN_CONFIG " << gpjson::driver::driver_inputs_to_json(in) << "\n";
    gpjson::Engine engine{};
    summarize(engine.query(in.dataset_path, in.queries, in.engine_options));
    std::cout << "### RUN END run_name=" << in.run_name << " ###\n";
    return 0;
  } catch (const gpjson::driver::CliExit &e) {
    return e.exit_code();
  } catch (const std::exception &e) {
    std::cerr << "gpjson_driver: " << e.what() << '\n';
    return 1;
  }
}
