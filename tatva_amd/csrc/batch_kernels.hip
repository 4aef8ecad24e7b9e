// Batched matrix-free FP32 Jacobi-PCG: one launch per kernel step covers every
// active solve of a batch (the serving loop's requests, each with its own mesh,
// material, constraints and tolerance), so requests of very different sizes
// share the launches instead of queueing behind each other.
//
// Per iteration (tick t), with one class = slots of equal (element kind,
// density kind, quadrature rule, affine Hex8 flag):
//   per class  A  tile assembly of H(u_s) z_s into the slot's tile partials
//   per class  G  slot gather -> w_s = H z_s (reduced) and the shares of z.w
//   all slots  S  the single-reduction CG step of each slot (the scalars from
//                 its partials, p, s, x, r, z updates, shares of r.z and r.r)
// with the recurrences of the single-solve PCG (solver_kernels.hip).  A block
// finds its slot through a per-block table (one load) and works on that
// slot's arrays only; the slot's own blocks re-sum its partials across the
// kernel boundary.  Every slot runs the same block decomposition whatever else
// is in the batch, so its arithmetic, and its result, do not depend on the
// mix.  The parity of the partial buffers and scalars is (t - start tick of the
// slot) mod 2.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "assembly.hpp"
#include "common.hpp"
#include "mesh.hpp"
#include "pcg.hpp"

namespace tatva {

// Gather blocks of one slot: grid-stride over nodes beyond kMaxReduceBlocks.
inline int gather_grid(int64_t n_nodes) {
  int64_t b = (n_nodes + kBlock - 1) / kBlock;
  if (b < 1) b = 1;
  if (b > kMaxReduceBlocks) b = kMaxReduceBlocks;
  return (int)b;
}

struct BatchSlot {
  const double *coords;
  const int32_t *conn;
  TileSet tiles;
  int64_t n_nodes;
  int64_t n;
  tatva_density dens;
  const int32_t *dof_map;  // full -> reduced (or NULL: identity)
  const float *u;          // FP32 full state (linearisation point)
  float *partial;          // tile partials, n_slots * ncomp
  float *x, *r, *z, *p, *s, *w, *dinv;
  double *pap, *rz0, *rz1, *rr0, *rr1, *bb;
  PcgState *st;
  int64_t start_tick;  // batch tick of the slot's first iteration
  int32_t g_blocks, v_blocks;
  int32_t a_blk0, g_blk0, v_blk0;
};

template <class Dens>
__device__ __forceinline__ Dens make_density(const tatva_density &d) {
  if constexpr (Dens::kSource) return Dens{float(d.p0), d.elem_params};
  else return Dens{float(d.p0), float(d.p1), d.elem_params};
}

template <int KIND, class Dens, bool AFFINE>
__global__ __launch_bounds__(kTile) void k_batch_assemble(const BatchSlot *__restrict__ slots,
                                                          const int32_t *__restrict__ bmap,
                                                          QuadRule quad) {
  const int b = (int)xcd_block(blockIdx.x, gridDim.x);
  const BatchSlot &S = slots[bmap[b]];
  if (S.st->done) return;
  extern __shared__ unsigned char lds_raw[];
  const Dens dens = make_density<Dens>(S.dens);
  // AFFINE: a Hex8 parallelepiped mesh (no geometry cache here)
  tile_assemble<KIND, Dens, float, kModeHvp,
                (AFFINE && KIND == TATVA_HEX8) ? kGeomAffine : kGeomGeneral>(
      b - S.a_blk0, S.coords, quad, dens, S.u, S.z, S.dof_map, S.tiles, S.partial, lds_raw);
}

template <int NC>
__global__ __launch_bounds__(kBlock) void k_batch_gather(const BatchSlot *__restrict__ slots,
                                                         const int32_t *__restrict__ bmap) {
  const int b = (int)xcd_block(blockIdx.x, gridDim.x);
  const BatchSlot &S = slots[bmap[b]];
  if (S.st->done) return;
  const int lb = b - S.g_blk0;
  const int32_t *__restrict__ dof_map = S.dof_map;
  const float *__restrict__ z = S.z;
  float *__restrict__ w = S.w;
  double dot = 0.0;
  for (int64_t i = (int64_t)lb * kBlock + threadIdx.x; i < S.n_nodes;
       i += (int64_t)S.g_blocks * kBlock) {
    float acc[NC];
    slot_sum<NC, float>(i, S.tiles.slot_ptr, S.partial, acc);
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int64_t dof = i * NC + c;
      const int64_t r = dof_map ? (int64_t)dof_map[dof] : dof;
      if (r >= 0) {
        w[r] = acc[c];
        dot += (double)acc[c] * (double)z[r];
      }
    }
  }
  dot = block_sum(dot);
  if (threadIdx.x == 0) S.pap[lb] = dot;
}

// Iteration k of a batch launch sequence runs at batch tick *tick + k (the
// tick advances by the sequence length at its end, k_batch_tick), so the
// graph of a sequence replays with no per-launch arguments.
// Elements per thread the batched CG step loads before its share sums.
constexpr int kBatchPrefetch = 4;

__global__ __launch_bounds__(kBlock) void k_batch_step(const BatchSlot *__restrict__ slots,
                                                       const int32_t *__restrict__ bmap,
                                                       const int64_t *__restrict__ tick, int k) {
  const int blk = (int)xcd_block(blockIdx.x, gridDim.x);
  const BatchSlot &S = slots[bmap[blk]];
  PcgState *st = S.st;
  if (st->done) return;
  const int lb = blk - S.v_blk0;
  const int q = (int)((*tick + k - S.start_tick) & 1);
  const int iter = st->it[q ^ 1];
  float *__restrict__ x = S.x;
  float *__restrict__ r = S.r;
  float *__restrict__ z = S.z;
  float *__restrict__ p = S.p;
  float *__restrict__ s = S.s;
  const float *__restrict__ w = S.w;
  const float *__restrict__ dinv = S.dinv;
  // the first kBatchPrefetch elements of this thread are loaded before the
  // shares are summed, so their latency overlaps the sums (p and s hold no
  // values before iteration 0: selected away below, never scaled)
  const int64_t stride = (int64_t)S.v_blocks * kBlock;
  const int64_t i0 = (int64_t)lb * kBlock + threadIdx.x;
  float rz[kBatchPrefetch], rp[kBatchPrefetch], rw[kBatchPrefetch], rs[kBatchPrefetch];
  float rx[kBatchPrefetch], rv[kBatchPrefetch], rd[kBatchPrefetch];
#pragma unroll
  for (int j = 0; j < kBatchPrefetch; ++j) {
    const int64_t i = i0 + j * stride;
    if (i < S.n) {
      rz[j] = z[i];
      rp[j] = p[i];
      rw[j] = w[i];
      rs[j] = s[i];
      rx[j] = x[i];
      rv[j] = r[i];
      rd[j] = dinv[i];
    }
  }
  double gamma, rr, delta;
  sum_partials3(q ? S.rz0 : S.rz1, S.v_blocks, q ? S.rr0 : S.rr1, S.v_blocks, S.pap, S.g_blocks,
                gamma, rr, delta);
  const bool first = iter == 0;
  const double beta = first ? 0.0 : gamma / st->gamma[q ^ 1];
  const double denom = first ? delta : delta - beta * gamma / st->alpha[q ^ 1];
  const bool converged = rr <= st->tol2;
  const bool spent = iter >= st->maxiter;
  const bool bad = !(denom > 0.0) || !isfinite(denom) || !isfinite(gamma) || !isfinite(rr);
  if (converged || spent || bad) {
    if (lb == 0 && threadIdx.x == 0) {
      st->done = 1;
      st->rr = rr;
      st->iter = iter;
      st->status = converged ? TATVA_OK : spent ? TATVA_ENOTCONV : TATVA_EBREAKDOWN;
    }
    return;
  }
  const double alpha = gamma / denom;
  if (lb == 0 && threadIdx.x == 0) {
    st->alpha[q] = alpha;
    st->gamma[q] = gamma;
    st->it[q] = iter + 1;
    st->iter = iter + 1;
    st->rr = rr;
  }
  const float a = float(alpha), bt = float(beta);
  double sg = 0.0, srr = 0.0;
#pragma unroll
  for (int j = 0; j < kBatchPrefetch; ++j) {
    const int64_t i = i0 + j * stride;
    if (i < S.n) {
      const float pi = first ? rz[j] : rz[j] + bt * rp[j];
      const float si = first ? rw[j] : rw[j] + bt * rs[j];
      p[i] = pi;
      s[i] = si;
      x[i] = rx[j] + a * pi;
      const float ri = rv[j] - a * si;
      r[i] = ri;
      const float zi = rd[j] * ri;
      z[i] = zi;
      sg += (double)ri * (double)zi;
      srr += (double)ri * (double)ri;
    }
  }
  for (int64_t i = i0 + kBatchPrefetch * stride; i < S.n; i += stride) {
    const float pi = first ? z[i] : z[i] + bt * p[i];
    const float si = first ? w[i] : w[i] + bt * s[i];
    p[i] = pi;
    s[i] = si;
    x[i] += a * pi;
    const float ri = r[i] - a * si;
    r[i] = ri;
    const float zi = dinv[i] * ri;
    z[i] = zi;
    sg += (double)ri * (double)zi;
    srr += (double)ri * (double)ri;
  }
  block_sum2(sg, srr);
  if (threadIdx.x == 0) {
    (q ? S.rz1 : S.rz0)[lb] = sg;
    (q ? S.rr1 : S.rr0)[lb] = srr;
  }
}

__global__ void k_batch_tick(int64_t *__restrict__ tick, int iters) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *tick += iters;
}

// Per-slot workspace: z, p, s, w, r, dinv (n floats each), tile partials
// (n_slots * ncomp floats), the partial sums pap (gather blocks) and
// rz0/rz1/rr0/rr1/bb (vector blocks).
struct SlotLayout {
  size_t z, p, s, w, r, dinv, partial, pap, rz0, rz1, rr0, rr1, bb, total;
};

static size_t al256(size_t x) { return (x + 255) & ~size_t(255); }

static SlotLayout slot_layout(const tatva_mesh *m, int ncomp, int64_t n) {
  SlotLayout L{};
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t at = o;
    o = al256(o + bytes);
    return at;
  };
  L.z = take(n * sizeof(float));
  L.p = take(n * sizeof(float));
  L.s = take(n * sizeof(float));
  L.w = take(n * sizeof(float));
  L.r = take(n * sizeof(float));
  L.dinv = take(n * sizeof(float));
  L.partial = take((size_t)std::max<int64_t>(1, m->n_slots) * ncomp * sizeof(float));
  L.pap = take(gather_grid(m->n_nodes) * sizeof(double));
  const size_t vb = step_grid(n, kBatchPrefetch) * sizeof(double);
  L.rz0 = take(vb);
  L.rz1 = take(vb);
  L.rr0 = take(vb);
  L.rr1 = take(vb);
  L.bb = take(vb);
  L.total = o;
  return L;
}

}  // namespace tatva

using namespace tatva;

struct tatva_batch {
  struct Meta {
    bool active = false;
    int kind = 0, dkind = 0, ncomp = 0, affine = 0;
    QuadRule quad{};
    size_t lds = 0;
    tatva_mesh *mesh = nullptr;
  };
  struct Launch {
    int kind, dkind, ncomp, affine;
    QuadRule quad;
    size_t lds;
    int a_off, a_n, g_off, g_n;
  };
  int device = 0;
  int capacity = 0;
  std::mutex mu;
  std::vector<BatchSlot> slots;  // host copy of the device table
  std::vector<Meta> meta;
  std::vector<Launch> launches;
  int v_off = 0, v_n = 0;
  bool dirty = true;
  int64_t tick = 0;
  BatchSlot *d_slots = nullptr;
  PcgState *d_states = nullptr;
  PcgState *h_states = nullptr;  // pinned
  int32_t *d_bmap = nullptr;
  size_t bmap_cap = 0;
  BatchSlot *h_stage_slots = nullptr;  // pinned staging of the uploads
  int32_t *h_stage_bmap = nullptr;
  size_t stage_bmap_cap = 0;
  hipEvent_t uploaded = nullptr;
  bool upload_pending = false;
  int64_t *d_tick = nullptr;  // device copy of tick (advanced by the launch sequence)
  // the launch sequence of `iters` iterations, built once per launch plan
  hipStream_t cap = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<int64_t> plan;  // what the captured sequence was built for
  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    plan.clear();
  }
};

namespace {

int batch_fail(hipError_t e, const char *what) {
  set_error(std::string(what) + ": " + hipGetErrorString(e));
  return TATVA_EHIP;
}

bool same_quad(const QuadRule &a, const QuadRule &b) {
  return std::memcmp(&a, &b, sizeof(QuadRule)) == 0;
}

// Re-plans the launches over the active slots and uploads the slot table and
// the block map (pinned staging; the previous upload must have completed
// before the staging is rewritten).
int batch_upload(tatva_batch *B, hipStream_t s) {
  if (!B->dirty) return TATVA_OK;
  if (B->upload_pending) {
    TATVA_TRY_HIP(hipEventSynchronize(B->uploaded));
    B->upload_pending = false;
  }
  B->launches.clear();
  std::vector<int32_t> bmap;
  // classes in slot order of their first member
  for (int i = 0; i < B->capacity; ++i) {
    const auto &mi = B->meta[i];
    if (!mi.active) continue;
    bool found = false;
    for (auto &L : B->launches)
      if (L.kind == mi.kind && L.dkind == mi.dkind && L.affine == mi.affine &&
          same_quad(L.quad, mi.quad))
        found = true;
    if (found) continue;
    tatva_batch::Launch L{mi.kind, mi.dkind, mi.ncomp, mi.affine, mi.quad, 0, 0, 0, 0, 0};
    B->launches.push_back(L);
  }
  for (auto &L : B->launches) {
    L.a_off = (int)bmap.size();
    for (int i = 0; i < B->capacity; ++i) {
      const auto &mi = B->meta[i];
      if (!mi.active || mi.kind != L.kind || mi.dkind != L.dkind || mi.affine != L.affine ||
          !same_quad(mi.quad, L.quad))
        continue;
      B->slots[i].a_blk0 = (int32_t)bmap.size() - L.a_off;
      bmap.insert(bmap.end(), B->slots[i].tiles.n_tiles, i);
      L.lds = std::max(L.lds, mi.lds);
    }
    L.a_n = (int)bmap.size() - L.a_off;
    L.g_off = (int)bmap.size();
    for (int i = 0; i < B->capacity; ++i) {
      const auto &mi = B->meta[i];
      if (!mi.active || mi.kind != L.kind || mi.dkind != L.dkind || mi.affine != L.affine ||
          !same_quad(mi.quad, L.quad))
        continue;
      B->slots[i].g_blk0 = (int32_t)bmap.size() - L.g_off;
      bmap.insert(bmap.end(), B->slots[i].g_blocks, i);
    }
    L.g_n = (int)bmap.size() - L.g_off;
  }
  B->v_off = (int)bmap.size();
  for (int i = 0; i < B->capacity; ++i) {
    if (!B->meta[i].active) continue;
    B->slots[i].v_blk0 = (int32_t)bmap.size() - B->v_off;
    bmap.insert(bmap.end(), B->slots[i].v_blocks, i);
  }
  B->v_n = (int)bmap.size() - B->v_off;
  if (bmap.size() >= (size_t)1 << 31) {
    set_error("batch block map too large");
    return TATVA_EINVAL;
  }
  if (bmap.size() > B->bmap_cap) {
    if (B->d_bmap) (void)hipFree(B->d_bmap);
    if (B->h_stage_bmap) (void)hipHostFree(B->h_stage_bmap);
    B->d_bmap = nullptr;
    B->h_stage_bmap = nullptr;
    B->bmap_cap = 0;
    const size_t cap = std::max<size_t>(bmap.size() * 2, 4096);
    hipError_t e = hipMalloc(&B->d_bmap, cap * sizeof(int32_t));
    if (e != hipSuccess) return batch_fail(e, "hipMalloc block map");
    e = hipHostMalloc(&B->h_stage_bmap, cap * sizeof(int32_t), hipHostMallocDefault);
    if (e != hipSuccess) return batch_fail(e, "hipHostMalloc block map");
    B->bmap_cap = cap;
  }
  std::memcpy(B->h_stage_slots, B->slots.data(), B->capacity * sizeof(BatchSlot));
  if (!bmap.empty()) std::memcpy(B->h_stage_bmap, bmap.data(), bmap.size() * sizeof(int32_t));
  TATVA_TRY_HIP(hipMemcpyAsync(B->d_slots, B->h_stage_slots, B->capacity * sizeof(BatchSlot),
                               hipMemcpyHostToDevice, s));
  if (!bmap.empty())
    TATVA_TRY_HIP(hipMemcpyAsync(B->d_bmap, B->h_stage_bmap, bmap.size() * sizeof(int32_t),
                                 hipMemcpyHostToDevice, s));
  TATVA_TRY_HIP(hipEventRecord(B->uploaded, s));
  B->upload_pending = true;
  B->dirty = false;
  return TATVA_OK;
}

template <class F>
int with_class(int kind, int dkind, F &&f) {
  auto dens = [&](auto kc) -> int {
    constexpr int K = decltype(kc)::value;
    constexpr int D = Elem<K>::D;
    switch (dkind) {
      case TATVA_DENSITY_POISSON: return f(kc, Poisson<D, float>{0.0f});
      case TATVA_DENSITY_LINEAR_ELASTIC: return f(kc, LinearElastic<D, float>{0.0f, 0.0f});
      case TATVA_DENSITY_NEO_HOOKEAN: return f(kc, NeoHookean<D, float>{0.0f, 0.0f});
      default: set_error("unknown density kind"); return TATVA_EINVAL;
    }
  };
  switch (kind) {
    case TATVA_TRI3: return dens(std::integral_constant<int, TATVA_TRI3>{});
    case TATVA_TET4: return dens(std::integral_constant<int, TATVA_TET4>{});
    case TATVA_HEX8: return dens(std::integral_constant<int, TATVA_HEX8>{});
    default: set_error("unknown element kind"); return TATVA_EINVAL;
  }
}

}  // namespace

extern "C" int tatva_batch_create(int device, int32_t capacity, tatva_batch **out) {
  TATVA_REQUIRE(out != nullptr, "out is NULL");
  *out = nullptr;
  TATVA_REQUIRE(capacity >= 1 && capacity <= 4096, "capacity must be in [1, 4096]");
  TATVA_TRY_HIP(hipSetDevice(device));
  tatva_batch *B = new tatva_batch();
  B->device = device;
  B->capacity = capacity;
  B->slots.assign(capacity, BatchSlot{});
  B->meta.assign(capacity, tatva_batch::Meta{});
  auto fail = [&](hipError_t e, const char *what) {
    const int rc = batch_fail(e, what);
    tatva_batch_destroy(B);
    return rc;
  };
  hipError_t e;
  if ((e = hipMalloc(&B->d_slots, capacity * sizeof(BatchSlot))) != hipSuccess)
    return fail(e, "hipMalloc slots");
  if ((e = hipMalloc(&B->d_states, capacity * sizeof(PcgState))) != hipSuccess)
    return fail(e, "hipMalloc states");
  if ((e = hipMemset(B->d_states, 0, capacity * sizeof(PcgState))) != hipSuccess)
    return fail(e, "hipMemset states");
  if ((e = hipHostMalloc(&B->h_states, capacity * sizeof(PcgState), hipHostMallocDefault)) != hipSuccess)
    return fail(e, "hipHostMalloc states");
  if ((e = hipHostMalloc(&B->h_stage_slots, capacity * sizeof(BatchSlot), hipHostMallocDefault)) != hipSuccess)
    return fail(e, "hipHostMalloc slots");
  if ((e = hipEventCreateWithFlags(&B->uploaded, hipEventDisableTiming)) != hipSuccess)
    return fail(e, "hipEventCreate");
  if ((e = hipEventCreateWithFlags(&B->ev_in, hipEventDisableTiming)) != hipSuccess)
    return fail(e, "hipEventCreate");
  if ((e = hipEventCreateWithFlags(&B->ev_out, hipEventDisableTiming)) != hipSuccess)
    return fail(e, "hipEventCreate");
  if ((e = hipStreamCreateWithFlags(&B->cap, hipStreamNonBlocking)) != hipSuccess)
    return fail(e, "hipStreamCreate");
  if ((e = hipMalloc(&B->d_tick, sizeof(int64_t))) != hipSuccess) return fail(e, "hipMalloc tick");
  if ((e = hipMemcpy(B->d_tick, &B->tick, sizeof(int64_t), hipMemcpyHostToDevice)) != hipSuccess)
    return fail(e, "hipMemcpy tick");
  std::memset(B->h_states, 0, capacity * sizeof(PcgState));
  *out = B;
  return TATVA_OK;
}

extern "C" int tatva_batch_destroy(tatva_batch *B) {
  if (!B) return TATVA_OK;
  (void)hipSetDevice(B->device);
  (void)hipDeviceSynchronize();
  if (B->d_slots) (void)hipFree(B->d_slots);
  if (B->d_states) (void)hipFree(B->d_states);
  if (B->d_bmap) (void)hipFree(B->d_bmap);
  if (B->h_states) (void)hipHostFree(B->h_states);
  if (B->h_stage_slots) (void)hipHostFree(B->h_stage_slots);
  if (B->h_stage_bmap) (void)hipHostFree(B->h_stage_bmap);
  if (B->uploaded) (void)hipEventDestroy(B->uploaded);
  B->drop_graph();
  if (B->ev_in) (void)hipEventDestroy(B->ev_in);
  if (B->ev_out) (void)hipEventDestroy(B->ev_out);
  if (B->cap) (void)hipStreamDestroy(B->cap);
  if (B->d_tick) (void)hipFree(B->d_tick);
  delete B;
  return TATVA_OK;
}

extern "C" int64_t tatva_batch_slot_workspace(tatva_mesh *m, int32_t ncomp, int64_t n) {
  if (!m || n <= 0 || ncomp < 1 || ncomp > 3) return -1;
  return (int64_t)slot_layout(m, ncomp, n).total;
}

extern "C" int tatva_batch_start(tatva_batch *B, int32_t slot, tatva_mesh *m,
                                 const tatva_density *d, const float *u, const int32_t *dof_map,
                                 int64_t n, const float *diag, const float *b, float *x,
                                 double tol, int32_t maxiter, void *workspace,
                                 int64_t workspace_bytes, void *stream) {
  TATVA_REQUIRE(B && m && d && u && diag && b && x && n > 0, "invalid arguments");
  TATVA_REQUIRE(slot >= 0 && slot < B->capacity, "slot out of range");
  TATVA_REQUIRE(m->device == B->device, "mesh and batch live on different devices");
  TATVA_REQUIRE(d->kind >= TATVA_DENSITY_POISSON && d->kind <= TATVA_DENSITY_NEO_HOOKEAN,
                "unknown density kind");
  const int ncomp = density_ncomp(m, d);
  const int64_t n_full = m->n_nodes * ncomp;
  TATVA_REQUIRE(n <= n_full, "reduced size exceeds the mesh dofs");
  TATVA_REQUIRE(dof_map || n == n_full, "without dof_map n must equal the number of dofs");
  const SlotLayout L = slot_layout(m, ncomp, n);
  TATVA_REQUIRE(workspace && workspace_bytes >= (int64_t)L.total, "batch slot workspace too small");
  const size_t lds = tile_lds_bytes(m->kind, m->tiles.max_local, m->dim, m->nne, ncomp, sizeof(float),
                                    kModeHvp);
  TATVA_REQUIRE(lds <= kTileLdsMax, "mesh tiles need more LDS than a workgroup holds");
  std::lock_guard<std::mutex> lock(B->mu);
  TATVA_TRY_HIP(hipSetDevice(B->device));
  hipStream_t s = (hipStream_t)stream;
  char *base = (char *)workspace;
  BatchSlot S{};
  S.coords = m->coords;
  S.conn = m->conn;
  S.tiles = m->tiles;
  S.n_nodes = m->n_nodes;
  S.n = n;
  S.dens = *d;
  S.dof_map = dof_map;
  S.u = u;
  S.partial = (float *)(base + L.partial);
  S.x = x;
  S.r = (float *)(base + L.r);
  S.z = (float *)(base + L.z);
  S.p = (float *)(base + L.p);
  S.s = (float *)(base + L.s);
  S.w = (float *)(base + L.w);
  S.dinv = (float *)(base + L.dinv);
  S.pap = (double *)(base + L.pap);
  S.rz0 = (double *)(base + L.rz0);
  S.rz1 = (double *)(base + L.rz1);
  S.rr0 = (double *)(base + L.rr0);
  S.rr1 = (double *)(base + L.rr1);
  S.bb = (double *)(base + L.bb);
  S.st = B->d_states + slot;
  S.g_blocks = gather_grid(m->n_nodes);
  S.v_blocks = step_grid(n, kBatchPrefetch);
  S.start_tick = B->tick;
  // r = b - H x, z = D^-1 r, shares of r.z, r.r, b.b (w is scratch here)
  const int rc = matfree_pcg_setup(m, d, u, dof_map, n, diag, b, x, S.r, S.z, S.w, S.dinv, S.rz1,
                                   S.rr1, S.bb, S.v_blocks, tol, maxiter, S.st, s);
  if (rc) return rc;
  B->slots[slot] = S;
  auto &mi = B->meta[slot];
  mi.active = true;
  mi.kind = m->kind;
  mi.dkind = d->kind;
  mi.ncomp = ncomp;
  mi.quad = m->quad;
  mi.affine = m->hex_affine;
  mi.lds = lds;
  mi.mesh = m;
  B->dirty = true;
  return TATVA_OK;
}

extern "C" int tatva_batch_release(tatva_batch *B, int32_t slot) {
  TATVA_REQUIRE(B, "batch is NULL");
  TATVA_REQUIRE(slot >= 0 && slot < B->capacity, "slot out of range");
  std::lock_guard<std::mutex> lock(B->mu);
  if (B->meta[slot].active) {
    B->meta[slot] = tatva_batch::Meta{};
    B->slots[slot] = BatchSlot{};
    B->dirty = true;
  }
  return TATVA_OK;
}

namespace {

// Builds the graph of `iters` batch iterations (the launch plan of
// batch_upload) node by node: a chain of kernel nodes, no stream capture, so
// HIP calls made meanwhile by other host threads cannot disturb it.
int build_iterations(tatva_batch *B, int iters, hipGraph_t *out) {
  hipGraph_t g = nullptr;
  TATVA_TRY_HIP(hipGraphCreate(&g, 0));
  hipGraphNode_t prev = nullptr;
  const BatchSlot *slots = B->d_slots;
  int64_t *tick = B->d_tick;
  auto add = [&](const void *fn, int grid, int block, size_t lds, void **args) -> hipError_t {
    hipKernelNodeParams p{};
    p.func = const_cast<void *>(fn);
    p.gridDim = dim3((unsigned)grid);
    p.blockDim = dim3((unsigned)block);
    p.sharedMemBytes = (unsigned)lds;
    p.kernelParams = args;
    p.extra = nullptr;
    hipGraphNode_t node = nullptr;
    const hipError_t e = hipGraphAddKernelNode(&node, g, prev ? &prev : nullptr, prev ? 1 : 0, &p);
    if (e == hipSuccess) prev = node;
    return e;
  };
  hipError_t err = hipSuccess;
  int rc = TATVA_OK;
  for (int it = 0; it < iters && err == hipSuccess && rc == TATVA_OK; ++it) {
    for (const auto &L : B->launches) {
      rc = with_class(L.kind, L.dkind, [&](auto kc, auto dens_tag) -> int {
        constexpr int K = decltype(kc)::value;
        using Dens = decltype(dens_tag);
        const int32_t *bm_a = B->d_bmap + L.a_off;
        const int32_t *bm_g = B->d_bmap + L.g_off;
        QuadRule quad = L.quad;
        void *a_args[] = {(void *)&slots, (void *)&bm_a, (void *)&quad};
        const void *fa = (K == TATVA_HEX8 && L.affine)
                             ? reinterpret_cast<const void *>(&k_batch_assemble<K, Dens, true>)
                             : reinterpret_cast<const void *>(&k_batch_assemble<K, Dens, false>);
        err = add(fa, L.a_n, kTile, L.lds, a_args);
        if (err != hipSuccess) return TATVA_OK;
        void *g_args[] = {(void *)&slots, (void *)&bm_g};
        err = add(reinterpret_cast<const void *>(&k_batch_gather<Dens::NC>), L.g_n, kBlock, 0, g_args);
        return TATVA_OK;
      });
      if (rc || err != hipSuccess) break;
    }
    if (rc || err != hipSuccess) break;
    const int32_t *bm_v = B->d_bmap + B->v_off;
    int k = it;
    void *s_args[] = {(void *)&slots, (void *)&bm_v, (void *)&tick, (void *)&k};
    err = add(reinterpret_cast<const void *>(&k_batch_step), B->v_n, kBlock, 0, s_args);
  }
  if (err == hipSuccess && rc == TATVA_OK) {
    int n_it = iters;
    void *t_args[] = {(void *)&tick, (void *)&n_it};
    err = add(reinterpret_cast<const void *>(&k_batch_tick), 1, 1, 0, t_args);
  }
  if (err != hipSuccess || rc != TATVA_OK) {
    (void)hipGraphDestroy(g);
    if (rc) return rc;
    set_error(std::string("hipGraphAddKernelNode: ") + hipGetErrorString(err));
    return TATVA_EHIP;
  }
  *out = g;
  return TATVA_OK;
}

// Everything the graph of a launch sequence depends on besides device memory.
std::vector<int64_t> launch_plan(const tatva_batch *B, int iters) {
  std::vector<int64_t> p{iters, B->v_off, B->v_n, (int64_t)(intptr_t)B->d_bmap,
                         (int64_t)(intptr_t)B->d_slots};
  for (const auto &L : B->launches) {
    p.insert(p.end(), {L.kind, L.dkind, L.ncomp, L.affine, (int64_t)L.lds, L.a_off, L.a_n,
                       L.g_off, L.g_n});
    const int64_t *q = reinterpret_cast<const int64_t *>(&L.quad);
    p.insert(p.end(), q, q + sizeof(QuadRule) / sizeof(int64_t));
  }
  return p;
}

}  // namespace

extern "C" int tatva_batch_iterate(tatva_batch *B, int32_t iters, void *stream) {
  TATVA_REQUIRE(B, "batch is NULL");
  TATVA_REQUIRE(iters >= 0, "iters must be >= 0");
  std::lock_guard<std::mutex> lock(B->mu);
  TATVA_TRY_HIP(hipSetDevice(B->device));
  hipStream_t s = (hipStream_t)stream;
  int rc = batch_upload(B, s);
  if (rc) return rc;
  if (B->v_n == 0 || iters == 0) return TATVA_OK;
  // the sequence runs as a graph on the batch's own stream, joined to s with
  // events; the graph is rebuilt when the launch plan changes (slots joining
  // or leaving)
  std::vector<int64_t> plan = launch_plan(B, iters);
  if (!B->exec || plan != B->plan) {
    B->drop_graph();
    rc = build_iterations(B, iters, &B->graph);
    if (rc) return rc;
    TATVA_TRY_HIP(hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0));
    B->plan = std::move(plan);
  }
  TATVA_TRY_HIP(hipEventRecord(B->ev_in, s));
  TATVA_TRY_HIP(hipStreamWaitEvent(B->cap, B->ev_in, 0));
  TATVA_TRY_HIP(hipGraphLaunch(B->exec, B->cap));
  TATVA_TRY_HIP(hipEventRecord(B->ev_out, B->cap));
  TATVA_TRY_HIP(hipStreamWaitEvent(s, B->ev_out, 0));
  B->tick += iters;
  return TATVA_OK;
}

extern "C" int tatva_batch_poll(tatva_batch *B, tatva_pcg_stats *stats, int32_t *n_running,
                                void *stream) {
  TATVA_REQUIRE(B && stats, "invalid arguments");
  std::lock_guard<std::mutex> lock(B->mu);
  TATVA_TRY_HIP(hipSetDevice(B->device));
  hipStream_t s = (hipStream_t)stream;
  TATVA_TRY_HIP(hipMemcpyAsync(B->h_states, B->d_states, B->capacity * sizeof(PcgState),
                               hipMemcpyDeviceToHost, s));
  TATVA_TRY_HIP(hipStreamSynchronize(s));
  int running = 0;
  for (int i = 0; i < B->capacity; ++i) {
    const PcgState &h = B->h_states[i];
    tatva_pcg_stats &o = stats[i];
    std::memset(&o, 0, sizeof(o));
    if (!B->meta[i].active) {
      o.status = TATVA_OK;
      continue;
    }
    if (!h.done) ++running;
    o.iterations = h.iter;
    o.status = h.done ? h.status : TATVA_ENOTCONV;
    o.b_norm = std::sqrt(h.bb);
    o.rel_residual = h.bb > 0 ? std::sqrt(h.rr / h.bb) : 0.0;
    o.reserved = h.done;
  }
  if (n_running) *n_running = running;
  return TATVA_OK;
}
