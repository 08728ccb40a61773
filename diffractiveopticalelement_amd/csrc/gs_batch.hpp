// The batched far-field engine's host object (shared by gs_batch.hip and the
// warm-start hand-off in warmstart.hip).
#pragma once
#include <utility>
#include <vector>

#include "gs_ops.hpp"

struct doe_gs_batch {
  struct Class {
    int n = 0;
    std::vector<int> users;  // user index of each design of the class
    size_t off = 0;          // element offset of the class's first design
    doe::GsState* state = nullptr;
    double* partials = nullptr;
    size_t hist_off = 0;     // design d's histories at hist + hist_off + d*3K
    doe::Plan2 pl;
    doe::GsBuf<float> b{};
  };
  doe_ctx* ctx = nullptr;
  int count = 0;
  int algorithm = DOE_ALG_GS;
  double beta = 0.9;
  bool has_roi = false;
  std::vector<int> side;                    // per user index
  std::vector<std::pair<int, int>> where;   // user index -> (class, position)
  std::vector<Class> cls;
  size_t total = 0;
  std::vector<void*> allocs;
  float* re = nullptr;
  float* im = nullptr;
  float* illum = nullptr;
  float* tgt = nullptr;
  float* roi = nullptr;
  float* w = nullptr;
  float* phi0 = nullptr;
  float* ph[2] = {nullptr, nullptr};
  float* am[2] = {nullptr, nullptr};
  double* hist = nullptr;
  int hist_k = 0;  // history capacity per design
  int last_k = 0;  // iterations of the last run
  hipGraphExec_t graph = nullptr;
  int graph_k = -1;
  bool graph_merged = false;  // the captured loop's launch layout

  ~doe_gs_batch() {
    if (graph) hipGraphExecDestroy(graph);
    for (void* p : allocs) hipFree(p);
  }
  int alloc(void** p, size_t bytes) {
    if (hipMalloc(p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      doe::set_error("doe_gs_batch: device allocation failed");
      return DOE_ERR_OOM;
    }
    allocs.push_back(*p);
    return DOE_OK;
  }
};

