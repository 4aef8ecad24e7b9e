// Host-side setup of the sparse path: CSR sparsity from element connectivity,
// row/column submatrices for constrained dofs, and the natural-order greedy
// distance-2 coloring.  Plain C++ (no GPU needed); runs once per mesh/pattern.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "../../include/tatva_amd.h"

namespace tatva {

static thread_local std::string g_error;

void set_error(const std::string &msg) { g_error = msg; }

template <class F>
static void parallel_for(int64_t n, F &&f) {
  unsigned hw = std::thread::hardware_concurrency();
  int nt = (int)std::min<unsigned>(hw == 0 ? 1 : hw, 16);
  if (n < 4096 || nt <= 1) {
    f(0, n, 0);
    return;
  }
  std::vector<std::thread> th;
  const int64_t chunk = (n + nt - 1) / nt;
  for (int t = 0; t < nt; ++t) {
    const int64_t b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    th.emplace_back([&, b, e, t] { f(b, e, t); });
  }
  for (auto &x : th) x.join();
}

}  // namespace tatva

struct tatva_csr_host {
  int64_t n_rows = 0;
  std::vector<int32_t> indptr;
  std::vector<int32_t> indices;
};

using namespace tatva;

extern "C" const char *tatva_last_error(void) { return g_error.c_str(); }

extern "C" int tatva_version(void) { return 1000; }

extern "C" int tatva_device_count(int *out) {
  if (!out) {
    set_error("out is NULL");
    return TATVA_EINVAL;
  }
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) {
    *out = 0;
    return TATVA_OK;
  }
  *out = n;
  return TATVA_OK;
}

// Node graph: nodes sharing an element with node i (including i), ascending.
static int node_graph(const int32_t *elements, int64_t E, int nne, int64_t N,
                      std::vector<int64_t> &gptr, std::vector<int32_t> &gadj) {
  // node -> incident elements (counting sort, ascending element id)
  std::vector<int64_t> iptr((size_t)N + 1, 0);
  for (int64_t i = 0; i < E * nne; ++i) iptr[(size_t)elements[i] + 1]++;
  for (int64_t i = 0; i < N; ++i) iptr[(size_t)i + 1] += iptr[(size_t)i];
  std::vector<int32_t> ilist((size_t)(E * nne));
  {
    std::vector<int64_t> fill(iptr.begin(), iptr.end() - 1);
    for (int64_t e = 0; e < E; ++e)
      for (int a = 0; a < nne; ++a) ilist[(size_t)fill[(size_t)elements[e * nne + a]]++] = (int32_t)e;
  }
  std::vector<int32_t> deg((size_t)N, 0);
  // pass 1: degrees; pass 2: fill.  Per-thread scratch of candidate nodes.
  for (int pass = 0; pass < 2; ++pass) {
    parallel_for(N, [&](int64_t b, int64_t e, int) {
      std::vector<int32_t> cand;
      cand.reserve(256);
      for (int64_t i = b; i < e; ++i) {
        cand.clear();
        for (int64_t p = iptr[(size_t)i]; p < iptr[(size_t)i + 1]; ++p) {
          const int64_t el = ilist[(size_t)p];
          for (int a = 0; a < nne; ++a) cand.push_back(elements[el * nne + a]);
        }
        std::sort(cand.begin(), cand.end());
        cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
        if (pass == 0) {
          deg[(size_t)i] = (int32_t)cand.size();
        } else {
          std::copy(cand.begin(), cand.end(), gadj.begin() + gptr[(size_t)i]);
        }
      }
    });
    if (pass == 0) {
      gptr.assign((size_t)N + 1, 0);
      for (int64_t i = 0; i < N; ++i) gptr[(size_t)i + 1] = gptr[(size_t)i] + deg[(size_t)i];
      gadj.assign((size_t)gptr[(size_t)N], 0);
    }
  }
  return TATVA_OK;
}

extern "C" int tatva_pattern_from_element_dofs(const int32_t *element_dofs, int64_t E, int k,
                                               int64_t size, const int32_t *diag_dofs,
                                               int64_t n_diag, tatva_csr_host **out) {
  if (!out) {
    set_error("out is NULL");
    return TATVA_EINVAL;
  }
  *out = nullptr;
  if (E < 0 || k < 0 || size < 0 || n_diag < 0 || (E * k > 0 && !element_dofs) ||
      (n_diag > 0 && !diag_dofs) || size >= (int64_t(1) << 31)) {
    set_error("invalid arguments");
    return TATVA_EINVAL;
  }
  for (int64_t i = 0; i < E * k; ++i)
    if (element_dofs[i] < -1 || element_dofs[i] >= size) {
      set_error("element dofs outside [-1, size)");
      return TATVA_EINVAL;
    }
  std::vector<char> diag((size_t)size, 0);
  for (int64_t i = 0; i < n_diag; ++i) {
    if (diag_dofs[i] < 0 || diag_dofs[i] >= size) {
      set_error("diagonal dofs outside [0, size)");
      return TATVA_EINVAL;
    }
    diag[(size_t)diag_dofs[i]] = 1;
  }
  // dof -> incident elements (ascending), -1 entries skipped
  std::vector<int64_t> iptr((size_t)size + 1, 0);
  for (int64_t i = 0; i < E * k; ++i)
    if (element_dofs[i] >= 0) iptr[(size_t)element_dofs[i] + 1]++;
  for (int64_t i = 0; i < size; ++i) iptr[(size_t)i + 1] += iptr[(size_t)i];
  std::vector<int64_t> ilist((size_t)iptr[(size_t)size]);
  {
    std::vector<int64_t> fill(iptr.begin(), iptr.end() - 1);
    for (int64_t e = 0; e < E; ++e)
      for (int a = 0; a < k; ++a) {
        const int32_t d = element_dofs[e * k + a];
        if (d >= 0) ilist[(size_t)fill[(size_t)d]++] = e;
      }
  }
  auto *h = new tatva_csr_host();
  h->n_rows = size;
  std::vector<int64_t> deg((size_t)size, 0), ptr;
  for (int pass = 0; pass < 2; ++pass) {
    parallel_for(size, [&](int64_t b, int64_t e, int) {
      std::vector<int32_t> cand;
      for (int64_t i = b; i < e; ++i) {
        cand.clear();
        for (int64_t p = iptr[(size_t)i]; p < iptr[(size_t)i + 1]; ++p) {
          const int64_t el = ilist[(size_t)p];
          for (int a = 0; a < k; ++a) {
            const int32_t d = element_dofs[el * k + a];
            if (d >= 0) cand.push_back(d);
          }
        }
        if (diag[(size_t)i]) cand.push_back((int32_t)i);
        std::sort(cand.begin(), cand.end());
        cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
        if (pass == 0) deg[(size_t)i] = (int64_t)cand.size();
        else std::copy(cand.begin(), cand.end(), h->indices.begin() + ptr[(size_t)i]);
      }
    });
    if (pass == 0) {
      ptr.assign((size_t)size + 1, 0);
      for (int64_t i = 0; i < size; ++i) ptr[(size_t)i + 1] = ptr[(size_t)i] + deg[(size_t)i];
      if (ptr[(size_t)size] >= (int64_t(1) << 31)) {
        delete h;
        set_error("pattern has 2^31 or more entries");
        return TATVA_EINVAL;
      }
      h->indices.assign((size_t)ptr[(size_t)size], 0);
    }
  }
  h->indptr.assign(ptr.begin(), ptr.end());
  *out = h;
  return TATVA_OK;
}

extern "C" int tatva_pattern_from_mesh(const int32_t *elements, int64_t E, int nne, int64_t N,
                                       int ndpn, const int32_t *dof_map, int64_t n_rows,
                                       tatva_csr_host **out) {
  if (!out) {
    set_error("out is NULL");
    return TATVA_EINVAL;
  }
  *out = nullptr;
  if (!elements || E < 0 || nne < 1 || N < 1 || ndpn < 1) {
    set_error("invalid arguments");
    return TATVA_EINVAL;
  }
  for (int64_t i = 0; i < E * nne; ++i)
    if (elements[i] < 0 || elements[i] >= N) {
      set_error("element connectivity references nodes outside [0, n_nodes)");
      return TATVA_EINVAL;
    }
  const int64_t n_full = N * ndpn;
  if (!dof_map) {
    if (n_rows != n_full) {
      set_error("n_rows must equal n_nodes * ndpn without a dof map");
      return TATVA_EINVAL;
    }
  } else {
    // the map must be strictly increasing on kept dofs and cover [0, n_rows)
    int64_t next = 0;
    for (int64_t d = 0; d < n_full; ++d) {
      const int32_t r = dof_map[d];
      if (r < 0) continue;
      if (r != next) {
        set_error("dof_map must number the kept dofs 0, 1, 2, ... in ascending dof order");
        return TATVA_EINVAL;
      }
      ++next;
    }
    if (next != n_rows) {
      set_error("dof_map keeps a different number of dofs than n_rows");
      return TATVA_EINVAL;
    }
  }
  std::vector<int64_t> gptr;
  std::vector<int32_t> gadj;
  int rc = node_graph(elements, E, nne, N, gptr, gadj);
  if (rc) return rc;
  auto map = [&](int64_t d) -> int64_t { return dof_map ? (int64_t)dof_map[d] : d; };
  // row lengths
  std::vector<int64_t> rlen((size_t)n_rows, 0);
  parallel_for(N, [&](int64_t b, int64_t e, int) {
    for (int64_t i = b; i < e; ++i) {
      int64_t cols = 0;
      for (int64_t p = gptr[(size_t)i]; p < gptr[(size_t)i + 1]; ++p)
        for (int c = 0; c < ndpn; ++c) cols += map((int64_t)gadj[(size_t)p] * ndpn + c) >= 0;
      for (int c = 0; c < ndpn; ++c) {
        const int64_t r = map(i * ndpn + c);
        if (r >= 0) rlen[(size_t)r] = cols;
      }
    }
  });
  auto *csr = new tatva_csr_host();
  csr->n_rows = n_rows;
  csr->indptr.assign((size_t)n_rows + 1, 0);
  int64_t total = 0;
  for (int64_t r = 0; r < n_rows; ++r) {
    total += rlen[(size_t)r];
    if (total >= (int64_t(1) << 31)) {
      delete csr;
      set_error("pattern has 2^31 or more non-zeros (int32 CSR)");
      return TATVA_EINVAL;
    }
    csr->indptr[(size_t)r + 1] = (int32_t)total;
  }
  csr->indices.assign((size_t)total, 0);
  parallel_for(N, [&](int64_t b, int64_t e, int) {
    for (int64_t i = b; i < e; ++i)
      for (int c = 0; c < ndpn; ++c) {
        const int64_t r = map(i * ndpn + c);
        if (r < 0) continue;
        int64_t o = csr->indptr[(size_t)r];
        for (int64_t p = gptr[(size_t)i]; p < gptr[(size_t)i + 1]; ++p)
          for (int cc = 0; cc < ndpn; ++cc) {
            const int64_t s = map((int64_t)gadj[(size_t)p] * ndpn + cc);
            if (s >= 0) csr->indices[(size_t)o++] = (int32_t)s;
          }
      }
  });
  *out = csr;
  return TATVA_OK;
}

extern "C" int tatva_csr_submatrix(int64_t n, const int32_t *indptr, const int32_t *indices,
                                   const int32_t *keep_map, int64_t n_new, tatva_csr_host **out) {
  if (!out || !indptr || !indices || !keep_map || n < 0 || n_new < 0) {
    set_error("invalid arguments");
    return TATVA_EINVAL;
  }
  *out = nullptr;
  int64_t next = 0;
  for (int64_t i = 0; i < n; ++i) {
    if (keep_map[i] < 0) continue;
    if (keep_map[i] != next) {
      set_error("keep_map must number kept rows 0, 1, 2, ... in ascending order");
      return TATVA_EINVAL;
    }
    ++next;
  }
  if (next != n_new) {
    set_error("keep_map keeps a different number of rows than n_new");
    return TATVA_EINVAL;
  }
  auto *csr = new tatva_csr_host();
  csr->n_rows = n_new;
  csr->indptr.assign((size_t)n_new + 1, 0);
  for (int64_t i = 0; i < n; ++i) {
    const int32_t r = keep_map[i];
    if (r < 0) continue;
    int32_t cnt = 0;
    for (int32_t p = indptr[i]; p < indptr[i + 1]; ++p) cnt += keep_map[indices[p]] >= 0;
    csr->indptr[(size_t)r + 1] = cnt;
  }
  for (int64_t r = 0; r < n_new; ++r) csr->indptr[(size_t)r + 1] += csr->indptr[(size_t)r];
  csr->indices.assign((size_t)csr->indptr[(size_t)n_new], 0);
  for (int64_t i = 0; i < n; ++i) {
    const int32_t r = keep_map[i];
    if (r < 0) continue;
    int32_t o = csr->indptr[(size_t)r];
    for (int32_t p = indptr[i]; p < indptr[i + 1]; ++p) {
      const int32_t c = keep_map[indices[p]];
      if (c >= 0) csr->indices[(size_t)o++] = c;
    }
  }
  *out = csr;
  return TATVA_OK;
}

extern "C" int64_t tatva_csr_host_rows(const tatva_csr_host *c) { return c ? c->n_rows : -1; }
extern "C" int64_t tatva_csr_host_nnz(const tatva_csr_host *c) {
  return c ? (int64_t)c->indices.size() : -1;
}
extern "C" int tatva_csr_host_copy(const tatva_csr_host *c, int32_t *indptr, int32_t *indices) {
  if (!c || !indptr || (!indices && !c->indices.empty())) {
    set_error("invalid arguments");
    return TATVA_EINVAL;
  }
  std::memcpy(indptr, c->indptr.data(), c->indptr.size() * sizeof(int32_t));
  if (!c->indices.empty()) std::memcpy(indices, c->indices.data(), c->indices.size() * sizeof(int32_t));
  return TATVA_OK;
}
extern "C" void tatva_csr_host_free(tatva_csr_host *c) { delete c; }

// Natural-order greedy on the rows of A^2: color[i] = smallest color not used by
// an already colored j with (A^2)_ij != 0.
//
// Exact acceleration: consecutive rows with identical column sets that each
// contain their own diagonal form a "run" (the dofs of one node in an FEM
// pattern).  Members of a run are mutual A^2 neighbours and share every other
// A^2 neighbour, so the greedy gives the run's members, in order, the smallest
// colors not used by earlier neighbours; and rows k of one run share A(k), so
// each run among the neighbours is scanned once.
extern "C" int tatva_distance2_colors(int64_t n, const int32_t *indptr, const int32_t *indices,
                                      int32_t *colors, int32_t *n_colors_out) {
  if (n < 0 || !indptr || !colors || (n > 0 && !indices)) {
    set_error("invalid arguments");
    return TATVA_EINVAL;
  }
  for (int64_t i = 0; i < n; ++i) {
    if (indptr[i + 1] < indptr[i]) {
      set_error("indptr must be non-decreasing");
      return TATVA_EINVAL;
    }
  }
  for (int64_t p = 0; p < (int64_t)indptr[n]; ++p)
    if (indices[p] < 0 || indices[p] >= n) {
      set_error("column index out of range");
      return TATVA_EINVAL;
    }
  // run id per row
  std::vector<int32_t> run_of((size_t)n);
  std::vector<int64_t> run_start;
  run_start.reserve((size_t)n / 2 + 1);
  auto has_diag = [&](int64_t i) {
    for (int32_t p = indptr[i]; p < indptr[i + 1]; ++p)
      if (indices[p] == i) return true;
    return false;
  };
  auto same_row = [&](int64_t a, int64_t b) {
    const int32_t la = indptr[a + 1] - indptr[a], lb = indptr[b + 1] - indptr[b];
    if (la != lb) return false;
    return std::memcmp(indices + indptr[a], indices + indptr[b], (size_t)la * sizeof(int32_t)) == 0;
  };
  for (int64_t i = 0; i < n; ++i) {
    bool join = false;
    if (i > 0) {
      const int64_t s = run_start.back();
      join = same_row(s, i) && has_diag(i) && has_diag(s);
    }
    if (!join) run_start.push_back(i);
    run_of[(size_t)i] = (int32_t)(run_start.size() - 1);
  }
  const int64_t n_runs = (int64_t)run_start.size();
  run_start.push_back(n);
  for (int64_t i = 0; i < n; ++i) colors[i] = -1;
  std::vector<int64_t> used(64, 0);      // used[c] == mark: color c taken
  std::vector<int64_t> run_seen((size_t)n_runs, 0);
  int32_t max_color = -1;
  for (int64_t R = 0; R < n_runs; ++R) {
    const int64_t s = run_start[(size_t)R], m = run_start[(size_t)R + 1] - s;
    const int64_t mark = R + 1;
    for (int32_t p = indptr[s]; p < indptr[s + 1]; ++p) {
      const int32_t k = indices[p];
      const int32_t kr = run_of[(size_t)k];
      if (run_seen[(size_t)kr] == mark) continue;
      run_seen[(size_t)kr] = mark;
      for (int32_t q = indptr[k]; q < indptr[k + 1]; ++q) {
        const int32_t c = colors[indices[q]];
        if (c < 0) continue;
        if ((size_t)c >= used.size()) used.resize(std::max<size_t>(used.size() * 2, (size_t)c + 1), 0);
        used[(size_t)c] = mark;
      }
    }
    int32_t c = 0;
    for (int64_t t = 0; t < m; ++t) {
      while ((size_t)c < used.size() && used[(size_t)c] == mark) ++c;
      colors[s + t] = c;
      max_color = std::max(max_color, c);
      ++c;
    }
  }
  if (n_colors_out) *n_colors_out = max_color + 1;
  return TATVA_OK;
}

// Distance-2 validity of a coloring for the colored decompression: every row's
// columns carry distinct colors (then J_compressed[row, color] holds exactly
// one entry of the row).  Host arrays; *bad_row = the first row that has two
// columns of one color, or -1.  O(nnz) with one color stamp per color.
extern "C" int tatva_check_distance2_colors(int64_t n, const int32_t *indptr,
                                            const int32_t *indices, const int32_t *colors,
                                            int64_t *bad_row) {
  if (n < 0 || !indptr || !colors || !bad_row || (n > 0 && !indices)) {
    set_error("invalid arguments");
    return TATVA_EINVAL;
  }
  *bad_row = -1;
  int32_t max_color = -1;
  for (int64_t i = 0; i < n; ++i) {
    if (colors[i] < 0) {
      set_error("colors must be non-negative");
      return TATVA_EINVAL;
    }
    max_color = std::max(max_color, colors[i]);
  }
  std::vector<int64_t> stamp((size_t)max_color + 1, -1);
  for (int64_t r = 0; r < n; ++r) {
    for (int64_t k = indptr[r]; k < indptr[r + 1]; ++k) {
      const int32_t j = indices[k];
      if (j < 0 || j >= n) {
        set_error("column index out of range");
        return TATVA_EINVAL;
      }
      int64_t &st = stamp[(size_t)colors[j]];
      if (st == r) {
        *bad_row = r;
        return TATVA_OK;
      }
      st = r;
    }
  }
  return TATVA_OK;
}
