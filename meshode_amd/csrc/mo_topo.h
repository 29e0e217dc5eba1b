// Device-side adjacency builder for the batched LM solver.
#pragma once

#include "mo_common.h"

namespace mo {

struct TopoPairDesc {
  const int* idx;       // faces [nf,3] (kind 0) or edges [ne,2] (kind 1), device
  long long nedges;     // number of directed edges: 3*nf or ne
  long long vbase;      // first vertex of the pair in the batch arrays
  int nv;
  int kind;
};

size_t scan_workspace_bytes(long long n);
int exclusive_scan_launch(const unsigned* in, long long n, long long* out, long long* total_dev, void* ws,
                          hipStream_t s);

size_t topo_workspace_bytes(long long nvtot);

// raw: 2 * (number of directed edges) entries; on return vertex v's list is
// adj[v] = (offset into raw, unique neighbours), entries packed
// neighbour | multiplicity << 24; deg[v] = sum of multiplicities.
// err_dev bit 0: edge index out of range; bit 1: neighbour index >= 2^24 or
// multiplicity > 255.
// pstat[2p] = unique adjacency entries of pair p, pstat[2p+1] = largest
// multiplicity.
int topo_build_launch(const TopoPairDesc* pairs_dev, int npairs, long long max_edges, long long max_nv,
                      long long nvtot, unsigned* raw, int2* adj, float* deg, int* pstat, void* ws, int* err_dev,
                      hipStream_t s);

}  // namespace mo
