// Definitions of the opaque handle types of the C ABI (shared by the C ABI
// translation units).
#pragma once
#include "internal.hpp"
#include "batch.hpp"

struct p3g_ctx {
  p3g::Ctx c;
};
struct p3g_tree {
  p3g::Tree* t;
};
struct p3g_transcript {
  p3g::Transcript tx;
};
struct p3g_commit_plan {
  p3g::CommitPlan* p;
};
