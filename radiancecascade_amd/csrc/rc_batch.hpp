// rc_batch.hpp -- the work-list engine behind rc_batch (rc_batch.hip).
#pragma once
#include <string>

#include "rc_internal.hpp"

namespace rc {

struct WorkList;

// levels[i]: the (already resolved) level count of view i
hipError_t worklist_create(const rc_view *views, uint32_t n, const uint32_t *levels, Storage st, int device,
                           WorkList **out, std::string &err);
void worklist_destroy(WorkList *w);
hipError_t worklist_run(WorkList *w, hipStream_t s, std::string &err);
hipError_t worklist_autotune(WorkList *w, int iters, hipStream_t s, std::string &err);
size_t worklist_workspace_bytes(const WorkList *w);
int worklist_launches(const WorkList *w);

}  // namespace rc
