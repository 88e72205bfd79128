// CPU build of split_item (gpjson_amd/csrc/query_common.h), the interpreter's
// map from a file's work item to (query, line) when the file's queries take
// turns over blocks of lines.
#include <cstdint>

#include "query_common.h"

extern "C" {
// q[r], i[r] for every item r of a file with `lines` lines and nq queries.
void split_all(int64_t lines, int64_t nq, int64_t B, int64_t* q, int64_t* i) {
  for (int64_t r = 0; r < lines * nq; r++) gj::split_item(r, lines, nq, B, &q[r], &i[r]);
}
}
