// jsonpath.h -- compiler from the reference's JSONPath subset to its IR bytes.
#pragma once
#include <stdint.h>
#include <string>
#include <vector>

namespace gj {

enum class PathStatus { kOk = 0, kParseError = 1, kUnsupported = 2 };

struct CompiledPath {
  std::vector<uint8_t> ir;
  int max_depth = 0;    // levels the query walks (J/jsonpath/JSONPathResult.java:5)
  int num_results = 0;  // STORE_RESULT count
};

// Compiles `path` (UTF-8).  On failure returns the status and a message in *err.
PathStatus compile_path(const std::string& path, CompiledPath* out, std::string* err);

}  // namespace gj
