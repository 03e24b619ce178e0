// akx_api.cpp -- the C ABI of include/akx.h: context, buffers, renumbering,
// CSR build and the step driver that enqueues the HIP kernels of
// akx_kernels.hip.
//
// Internal numbering.  Unless the context is created with
// AKX_CTX_KEEP_NUMBERING, nodes and elements are renumbered along a Morton
// (Z-order) curve of their positions / centroids when the mesh is set, so that
// the elements that share a node, and the nodes of an element, sit close
// together in memory (the node gather and the displacement gather then hit
// the same cache lines).  Every host array crossing the ABI stays in the
// caller's numbering; the node -> slot CSR keeps the caller's element order,
// so the FP64 sums are unchanged by the renumbering.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <new>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/akx.h"
#include "akx_internal.hpp"

using namespace akx;

namespace {

int fail(akx_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIP_TRY(c, expr)                                                            \
  do {                                                                              \
    hipError_t _e = (expr);                                                         \
    if (_e != hipSuccess)                                                           \
      return fail((c), AKX_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

template <typename T>
hipError_t dalloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  return hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T));
}

template <typename T>
void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

// The mass CSR aliases the force CSR when both orders coincide: release the
// alias first so that no buffer is freed twice.
void free_csr(akx_ctx* c) {
  if (c->mass_off == c->csr_off) c->mass_off = nullptr; else dfree(c->mass_off);
  if (c->mass_slot == c->csr_slot) c->mass_slot = nullptr; else dfree(c->mass_slot);
  dfree(c->csr_off);
  dfree(c->csr_slot);
}

void free_derived(akx_ctx* c) {
  dfree(c->B64); dfree(c->J64); dfree(c->B32); dfree(c->J32);
  free_csr(c);
  dfree(c->slot_pos);
  dfree(c->csr2_off); dfree(c->csr2_slot); dfree(c->slot2_pos);
  dfree(c->pnode_cnt); dfree(c->pnode_ids); dfree(c->run_start); dfree(c->run_base);
  dfree(c->lconn); dfree(c->lrank); dfree(c->lcr);
  dfree(c->pcsr_off); dfree(c->pentry); dfree(c->ppos);
  dfree(c->tab64_d);
  dfree(c->node_inv_d); dfree(c->elem_inv_d);
  dfree(c->fel_off); dfree(c->fel); dfree(c->flslot); dfree(c->fch_off); dfree(c->fcbase);
  dfree(c->u_alt); dfree(c->pcoef); dfree(c->frec);
  c->fused_ok = false;
  c->patch_ok = false;
  dfree(c->slots); dfree(c->elem_mat_d); dfree(c->mats_d); dfree(c->damage); dfree(c->Yd);
  // one entry per material: sized by the material count at its first use
  dfree(c->mat_dt);
  dfree(c->partial); dfree(c->scalar); dfree(c->gradu); dfree(c->stress); dfree(c->flag);
  for (auto& g : c->groups) dfree(g.filter);
  c->groups.clear();
  c->initialized = false;
}

void free_graph(akx_ctx* c) {
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->graph_exec = nullptr;
  c->graph = nullptr;
  c->graph_key = akx_ctx::GraphKey();
}

void free_mesh_buffers(akx_ctx* c) {
  free_graph(c);
  dfree(c->dt_node);
  free_derived(c);
  dfree(c->X); dfree(c->conn_soa);
  dfree(c->u); dfree(c->v); dfree(c->a); dfree(c->fext); dfree(c->fint); dfree(c->mass);
  dfree(c->prev); dfree(c->inc); dfree(c->blocked); dfree(c->work);
}

int elem_dim(int et) { return (et == AKX_TETRAHEDRON_4 || et == AKX_HEXAHEDRON_8) ? 3 : 2; }
int elem_nodes(int et) { return et == AKX_TRIANGLE_3 ? 3 : et == AKX_HEXAHEDRON_8 ? 8 : 4; }
int elem_qp(int et) { return et == AKX_QUADRANGLE_4 ? 4 : et == AKX_HEXAHEDRON_8 ? 8 : 1; }

// Gauss points, weights, N and dN/dxi of the element (host FP64: glibc sqrt).
// src/fe_engine/gauss_integration.cc:55-56,72-73,122-123 and the element classes
// element_class_{triangle_3,quadrangle_4,tetrahedron_4,hexahedron_8}_inline_impl.hh.
ElemTables make_tables(int et) {
  ElemTables t;
  std::memset(&t, 0, sizeof t);
  const double g = 1. / std::sqrt(3.);
  const int d = elem_dim(et), Q = elem_qp(et);
  for (int q = 0; q < Q; ++q) {
    double c[3] = {0, 0, 0};
    double* D = t.dnds[q];
    auto set = [&](int i, int a, double v) { D[i + a * d] = v; };
    switch (et) {
    case AKX_TRIANGLE_3:
      c[0] = 1. / 3.; c[1] = 1. / 3.; t.w[q] = 1. / 2.;
      set(0, 0, -1.); set(0, 1, 1.); set(0, 2, 0.); set(1, 0, -1.); set(1, 1, 0.); set(1, 2, 1.);
      t.N[q][0] = 1 - c[0] - c[1]; t.N[q][1] = c[0]; t.N[q][2] = c[1];
      break;
    case AKX_QUADRANGLE_4:
      c[0] = (q & 1) ? g : -g; c[1] = (q & 2) ? g : -g; t.w[q] = 1.;
      set(0, 0, -.25 * (1 - c[1])); set(0, 1, .25 * (1 - c[1]));
      set(0, 2, .25 * (1 + c[1])); set(0, 3, -.25 * (1 + c[1]));
      set(1, 0, -.25 * (1 - c[0])); set(1, 1, -.25 * (1 + c[0]));
      set(1, 2, .25 * (1 + c[0])); set(1, 3, .25 * (1 - c[0]));
      t.N[q][0] = .25 * (1 - c[0]) * (1 - c[1]); t.N[q][1] = .25 * (1 + c[0]) * (1 - c[1]);
      t.N[q][2] = .25 * (1 + c[0]) * (1 + c[1]); t.N[q][3] = .25 * (1 - c[0]) * (1 + c[1]);
      break;
    case AKX_TETRAHEDRON_4: {
      c[0] = c[1] = c[2] = 1. / 4.; t.w[q] = 1. / 6.;
      const double v[3][4] = {{-1, 1, 0, 0}, {-1, 0, 1, 0}, {-1, 0, 0, 1}};
      for (int i = 0; i < 3; ++i)
        for (int a = 0; a < 4; ++a) set(i, a, v[i][a]);
      t.N[q][0] = 1 - c[0] - c[1] - c[2]; t.N[q][1] = c[1]; t.N[q][2] = c[2]; t.N[q][3] = c[0];
    } break;
    case AKX_HEXAHEDRON_8: {
      c[0] = (q & 1) ? g : -g; c[1] = (q & 2) ? g : -g; c[2] = (q & 4) ? g : -g; t.w[q] = 1.;
      const int sx[8] = {-1, 1, 1, -1, -1, 1, 1, -1}, sy[8] = {-1, -1, 1, 1, -1, -1, 1, 1},
                sz[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
      for (int a = 0; a < 8; ++a) {
        auto f = [&](int s, double x) { return s > 0 ? (1 + x) : (1 - x); };
        set(0, a, (sx[a] > 0 ? .125 : -.125) * f(sy[a], c[1]) * f(sz[a], c[2]));
        set(1, a, (sy[a] > 0 ? .125 : -.125) * f(sx[a], c[0]) * f(sz[a], c[2]));
        set(2, a, (sz[a] > 0 ? .125 : -.125) * f(sx[a], c[0]) * f(sy[a], c[1]));
        t.N[q][a] = .125 * f(sx[a], c[0]) * f(sy[a], c[1]) * f(sz[a], c[2]);
      }
    } break;
    }
  }
  return t;
}

// ---------------------------------------------------------------------------
// Morton renumbering
// ---------------------------------------------------------------------------
uint64_t spread3(uint64_t v) {
  v &= 0x1FFFFF;
  v = (v | (v << 32)) & 0x1F00000000FFFFull;
  v = (v | (v << 16)) & 0x1F0000FF0000FFull;
  v = (v | (v << 8)) & 0x100F00F00F00F00Full;
  v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
  v = (v | (v << 2)) & 0x1249249249249249ull;
  return v;
}
uint64_t spread2(uint64_t v) {
  v &= 0xFFFFFFFF;
  v = (v | (v << 16)) & 0x0000FFFF0000FFFFull;
  v = (v | (v << 8)) & 0x00FF00FF00FF00FFull;
  v = (v | (v << 4)) & 0x0F0F0F0F0F0F0F0Full;
  v = (v | (v << 2)) & 0x3333333333333333ull;
  v = (v | (v << 1)) & 0x5555555555555555ull;
  return v;
}

// perm[new] = old, ordered by the Morton key of the points P (n x dim).
std::vector<uint32_t> morton_order(const std::vector<double>& P, size_t n, int dim,
                                   const double* lo, const double* hi) {
  const double levels = dim == 3 ? double((1u << 21) - 1) : double(0xFFFFFFFFu);
  std::vector<std::pair<uint64_t, uint32_t>> key(n);
  for (size_t i = 0; i < n; ++i) {
    uint64_t k = 0;
    for (int j = 0; j < dim; ++j) {
      const double span = hi[j] > lo[j] ? hi[j] - lo[j] : 1.;
      double t = (P[i * dim + j] - lo[j]) / span;
      t = std::min(1., std::max(0., t));
      const uint64_t q = uint64_t(t * levels);
      k |= (dim == 3 ? spread3(q) : spread2(q)) << j;
    }
    key[i] = {k, uint32_t(i)};
  }
  std::sort(key.begin(), key.end());
  std::vector<uint32_t> perm(n);
  for (size_t i = 0; i < n; ++i) perm[i] = key[i].second;
  return perm;
}

// node -> slot CSR in internal numbering; `order` lists caller element ids in
// assembly order.
void build_csr(const akx_ctx* c, const std::vector<uint32_t>& order, std::vector<uint32_t>& off,
               std::vector<uint32_t>& slot) {
  const int n = c->n;
  const int64_t N = c->N;
  off.assign(N + 1, 0);
  for (uint32_t e : order)
    for (int a = 0; a < n; ++a) off[c->node_inv[c->hconn[size_t(e) * n + a]] + 1]++;
  for (int64_t i = 0; i < N; ++i) off[i + 1] += off[i];
  slot.assign(off[N], 0);
  std::vector<uint32_t> fill(off.begin(), off.end() - 1);
  for (uint32_t e : order) {
    const uint32_t ei = c->elem_inv[e];
    for (int a = 0; a < n; ++a) {
      const uint32_t node = c->node_inv[c->hconn[size_t(e) * n + a]];
      slot[fill[node]++] = ei * uint32_t(n) + uint32_t(a);
    }
  }
}

void update_params(HostMaterial& m, int dim) {
  // material_elastic.cc:83-107, material_neohookean.cc:83-89
  m.lambda = m.nu * m.E / ((1 + m.nu) * (1 - 2 * m.nu));
  m.mu = m.E / (2 * (1 + m.nu));
  if (dim == 2 && m.plane_stress && m.kind == AKX_MATERIAL_ELASTIC)
    m.lambda = m.nu * m.E / ((1 + m.nu) * (1 - m.nu));
}

size_t field_elems(const akx_ctx* c, int field) {
  switch (field) {
  case AKX_GRAD_U:
  case AKX_STRESS: return size_t(c->E) * c->Q * c->d * c->d;
  case AKX_DAMAGE:
  case AKX_JACOBIANS: return size_t(c->E) * c->Q;
  case AKX_SHAPES_DERIVATIVES: return size_t(c->E) * c->Q * c->d * c->n;
  default: return size_t(c->N) * c->d;
  }
}

double* nodal_ptr(akx_ctx* c, int field) {
  switch (field) {
  case AKX_DISPLACEMENT: return c->u;
  case AKX_VELOCITY: return c->v;
  case AKX_ACCELERATION: return c->a;
  case AKX_EXTERNAL_FORCE: return c->fext;
  case AKX_INTERNAL_FORCE: return c->fint;
  case AKX_PREVIOUS_DISPLACEMENT: return c->prev;
  case AKX_DISPLACEMENT_INCREMENT: return c->inc;
  default: return nullptr;
  }
}

// caller (N, d) <-> internal (N, d)
template <typename T>
void to_internal_nodal(const akx_ctx* c, const T* src, T* dst) {
  const int d = c->d;
  for (int64_t k = 0; k < c->N; ++k)
    for (int i = 0; i < d; ++i) dst[k * d + i] = src[size_t(c->node_perm[k]) * d + i];
}
template <typename T>
void to_caller_nodal(const akx_ctx* c, const T* src, T* dst) {
  const int d = c->d;
  for (int64_t k = 0; k < c->N; ++k)
    for (int i = 0; i < d; ++i) dst[size_t(c->node_perm[k]) * d + i] = src[k * d + i];
}
// per-point arrays with `comp` values per point, points element-major
template <typename T>
void to_caller_points(const akx_ctx* c, const T* src, T* dst, int comp) {
  const size_t per = size_t(c->Q) * comp;
  for (int64_t e = 0; e < c->E; ++e)
    std::memcpy(dst + size_t(c->elem_perm[e]) * per, src + size_t(e) * per, per * sizeof(T));
}
template <typename T>
void to_internal_points(const akx_ctx* c, const T* src, T* dst, int comp) {
  const size_t per = size_t(c->Q) * comp;
  for (int64_t e = 0; e < c->E; ++e)
    std::memcpy(dst + size_t(e) * per, src + size_t(c->elem_perm[e]) * per, per * sizeof(T));
}

// Step fused over node blocks (slot layout 6): one launch per step, no force
// slots; stateless laws of a single kind on elements of up to 4 nodes, FP64
// or mixed.
bool uses_fused(const akx_ctx* c, int /*precision*/) {
  return c->node_ordered == 6 && c->fused_ok && c->groups.size() == 1 && !c->groups[0].filter &&
         !(c->sur.enabled && c->groups[0].kind == AKX_MATERIAL_NEOHOOKEAN_DAMAGE);
}

// Mixed path with per-patch partial sums (slot layouts 3, 4 and 5; 6 runs as
// 5 where the fused step does not apply): the element kernel reduces
// each node's contributions inside its patch in FP64 and stores one FP64 (3)
// or FP32 (4, 5) partial per (patch, node), in patch order (3, 4) or at the
// node's place in a node-major list (5); the node kernel sums a node's
// partials in FP64.  Needs the LDS patches and a single launch group.
bool uses_partials(const akx_ctx* c, int precision) {
  return precision == AKX_MIXED && c->node_ordered >= 3 && c->node_ordered <= 6 &&
         !uses_fused(c, precision) &&
         c->patch_ok && c->groups.size() == 1 &&
         !c->groups[0].filter &&
         // the partials (padded stride) live in the slot buffer
         c->patch_blocks * int64_t(c->patch_max_nodes) <= c->E * int64_t(c->n);
}
// arithmetic of the node kernel's slots: FP64 for the FP64 path and for
// per-patch partials, FP32 otherwise
int node_precision(const akx_ctx* c, int precision) {
  return uses_partials(c, precision) ? AKX_FP64 : precision;
}

ElemArgs elem_args(akx_ctx* c, const KindGroup& g, int precision) {
  ElemArgs A;
  A.count = g.count;
  A.filter = g.filter;
  A.E = c->E;
  A.conn = c->conn_soa;
  A.B = precision == AKX_FP64 ? static_cast<const void*>(c->B64) : c->B32;
  A.J = precision == AKX_FP64 ? static_cast<const void*>(c->J64) : c->J32;
  A.u = c->u;
  A.slots = c->slots;
  A.elem_mat = c->elem_mat_d;
  A.mats = c->mats_d;
  A.damage = c->damage;
  A.Yd = c->Yd;
  A.xcd_remap = c->xcd_remap;
  A.recompute = c->mixed_recompute;
  A.X = c->X;
  A.slot_pos = !c->node_ordered ? nullptr : precision == AKX_MIXED ? c->slot2_pos : c->slot_pos;
  A.patch = (precision == AKX_MIXED && c->node_ordered >= 2 && c->patch_ok && !g.filter) ? 1 : 0;
  A.partials = (A.patch && uses_partials(c, precision)) ? (c->node_ordered == 3 ? 1 : 2) : 0;
  A.part_pos = (A.partials && c->node_ordered >= 5) ? c->ppos : nullptr;
  A.patch_max_nodes = c->patch_max_nodes;
  A.patch_lds = c->patch_lds;
  A.patch_blocks = c->patch_blocks;
  A.pnode_cnt = c->pnode_cnt;
  A.pnode_ids = c->pnode_ids;
  A.run_start = c->run_start;
  A.run_base = c->run_base;
  A.lconn = c->lconn;
  A.lrank = c->lrank;
  A.lcr = c->lcr;
  A.tabf = c->tabf;
  A.tab64 = c->tab64_d;
  A.sur_img = nullptr;
  A.sur_b = nullptr;
  A.sur_bf16 = false;
  A.sur_eps = A.sur_sS = A.sur_sW = 0.f;
  A.sur_fallbacks = nullptr;
  A.sur_x = A.sur_y = nullptr;
  A.net_ev0 = A.net_ev1 = nullptr;
  if (precision == AKX_MIXED && c->sur.enabled && g.kind == AKX_MATERIAL_NEOHOOKEAN_DAMAGE) {
    A.sur_img = c->sur.img;
    A.sur_b = c->sur.imgb;
    A.sur_bf16 = c->sur.infer_bf16;
    A.sur_eps = float(c->sur.eps);
    A.sur_sS = float(c->sur.sS);
    A.sur_sW = float(c->sur.sW);
    A.sur_fallbacks = c->sur.fallbacks;
    A.sur_x = c->sur.pt_x;
    A.sur_y = c->sur.pt_y;
  }
  return A;
}

NodeArgs node_args(akx_ctx* c, int mode, int precision) {
  NodeArgs A;
  A.N = c->N;
  A.csr_off = precision == AKX_MIXED ? c->csr2_off : c->csr_off;
  A.csr_slot = precision == AKX_MIXED ? c->csr2_slot : c->csr_slot;
  A.slots = c->slots;
  A.u = c->u; A.v = c->v; A.a = c->a;
  A.u_out = nullptr;
  A.u_prev = nullptr;
  A.pcoef = c->pcoef;
  A.fext = c->fext;
  A.mass = c->mass;
  A.blocked = c->blocked;
  A.fint_out = c->fint;
  A.prev = c->prev;
  A.inc = c->inc;
  A.dt = c->dt;
  A.dt_node = c->dt_node;
  for (int i = 0; i < 3; ++i) A.bc_inc[i] = c->bc_inc[i];
  A.mode = mode;
  A.xcd_remap = c->xcd_remap;
  A.node_ordered = c->node_ordered;
  A.stage_lds = 0;
  A.partials = 0;
  if (uses_partials(c, precision)) {
    A.partials = c->node_ordered == 3 ? 1 : 2;
    // one partial sum per (patch, node): a short gather per node, or a
    // contiguous run per node when the partials are node-ordered
    A.csr_off = c->pcsr_off;
    A.csr_slot = c->pentry;
    A.node_ordered = c->node_ordered >= 5 ? 1 : 0;
    return A;
  }
  if (c->node_ordered && c->node_stage) {
    // static LDS of k_node is at most 4 * 256 * 3 * 8 + 256 * 3 bytes (< 26 KiB)
    const size_t bytes = size_t(c->span_max[c->node_block / 64 - 1]) * c->d *
                         (precision == AKX_FP64 ? 8 : 4);
    if (bytes > 0 && bytes <= 38 * 1024) A.stage_lds = int(bytes);
  }
  return A;
}

int check_ready(akx_ctx* c) {
  if (!c) return AKX_ERR_ARGUMENT;
  if (!c->initialized) return fail(c, AKX_ERR_STATE, "akx_init has not been called");
  return AKX_OK;
}

// time_net: record events around the surrogate network launch (if any);
// u: the displacement the elements read (nullptr: c->u)
int run_elements(akx_ctx* c, int precision, bool time_net = false, const double* u = nullptr) {
  for (auto& g : c->groups) {
    ElemArgs A = elem_args(c, g, precision);
    if (u) A.u = u;
    if (time_net && A.sur_img) {
      for (auto& ev : c->net_ev)
        if (!ev) HIP_TRY(c, hipEventCreate(&ev));
      A.net_ev0 = c->net_ev[0];
      A.net_ev1 = c->net_ev[1];
      c->net_timed = true;
    }
    HIP_TRY(c, launch_element(c->et, g.kind, precision, c->stream, A, c->elem_block));
  }
  return AKX_OK;
}

// Patches of PATCH consecutive internal elements for k_element_patch: the
// distinct nodes of each patch, where each node's run of mixed-CSR slots
// starts, and for every (e, a) its patch-local node and rank in the run.
// `slot2` is the mixed-path CSR (node-major, element id ascending).  Leaves
// c->patch_ok false when a mesh does not fit the 16/8-bit local indices or
// the LDS budget.
constexpr int PATCH = 256;
int build_patches(akx_ctx* c, const std::vector<uint32_t>& slot2) {
  c->patch_ok = false;
  const int64_t E = c->E;
  const int n = c->n, d = c->d;
  if (E == 0) return AKX_OK;
  // patches need one constitutive law over all elements (a single launch group)
  for (int64_t e = 1; e < E; ++e)
    if (c->mats[c->elem_mat[e]].kind != c->mats[c->elem_mat[0]].kind) return AKX_OK;
  std::vector<uint32_t> pos(size_t(E) * n);  // CSR position of (e, a)
  for (size_t p = 0; p < slot2.size(); ++p) pos[size_t(slot2[p] % n) * E + slot2[p] / n] = uint32_t(p);
  const int64_t nb = (E + PATCH - 1) / PATCH;
  std::vector<uint32_t> poff(nb + 1, 0), pids, rstart;
  std::vector<uint16_t> rbase, lconn(size_t(E) * n);
  std::vector<uint8_t> lrank(size_t(E) * n);
  int maxm = 0;
  std::vector<uint32_t> nodes;
  for (int64_t b = 0; b < nb; ++b) {
    const int64_t e0 = b * PATCH, e1 = std::min<int64_t>(E, e0 + PATCH);
    nodes.clear();
    for (int64_t e = e0; e < e1; ++e)
      for (int a = 0; a < n; ++a)
        nodes.push_back(c->node_inv[c->hconn[size_t(c->elem_perm[e]) * n + a]]);
    std::sort(nodes.begin(), nodes.end());
    nodes.erase(std::unique(nodes.begin(), nodes.end()), nodes.end());
    const int m = int(nodes.size());
    if (m > 65535) return AKX_OK;
    maxm = std::max(maxm, m);
    std::vector<uint32_t> lo(m, std::numeric_limits<uint32_t>::max()), hi(m, 0), cnt(m, 0);
    for (int64_t e = e0; e < e1; ++e)
      for (int a = 0; a < n; ++a) {
        const uint32_t node = c->node_inv[c->hconn[size_t(c->elem_perm[e]) * n + a]];
        const int k = int(std::lower_bound(nodes.begin(), nodes.end(), node) - nodes.begin());
        const uint32_t p = pos[size_t(a) * E + e];
        lo[k] = std::min(lo[k], p);
        hi[k] = std::max(hi[k], p);
        cnt[k]++;
        lconn[size_t(a) * E + e] = uint16_t(k);
      }
    // patch-local numbering by decreasing run length (then node id): the
    // threads summing the runs of one wavefront loop about equally long
    std::vector<int> ord(m), newk(m);
    for (int k = 0; k < m; ++k) ord[k] = k;
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return cnt[x] > cnt[y]; });
    for (int j = 0; j < m; ++j) newk[ord[j]] = j;
    uint32_t acc = 0;
    for (int j = 0; j < m; ++j) {
      const int k = ord[j];
      if (hi[k] - lo[k] + 1 != cnt[k] || cnt[k] > 255) return AKX_OK;  // run not contiguous
      pids.push_back(nodes[k]);
      rstart.push_back(lo[k]);
      rbase.push_back(uint16_t(acc));
      acc += cnt[k];
    }
    rbase.push_back(uint16_t(acc));
    for (int64_t e = e0; e < e1; ++e)
      for (int a = 0; a < n; ++a) {
        const int k = lconn[size_t(a) * E + e];
        lrank[size_t(a) * E + e] = uint8_t(pos[size_t(a) * E + e] - lo[k]);
        lconn[size_t(a) * E + e] = uint16_t(newk[k]);
      }
    poff[b + 1] = poff[b] + uint32_t(m);
  }
  // LDS of k_element_patch: the node lists (patch_force_offset) and the
  // forces, d floats per contribution, plus the run-sum padding
  const size_t lds = patch_force_offset(maxm, d) + size_t(PATCH * n + RUNCHUNK - 1) * d * 4;
  if (lds > 64 * 1024) return AKX_OK;
  // per-patch lists padded to the stride S = maxm; padding repeats the
  // patch's first node (a valid address) and its run data
  const int S = maxm;
  if (uint64_t(nb) * uint64_t(S + 1) >= (uint64_t(1) << 32)) return AKX_OK;
  std::vector<uint32_t> cnt(nb), pids_p(size_t(nb) * S), rstart_p(size_t(nb) * S);
  std::vector<uint16_t> rbase_p(size_t(nb) * (S + 1));
  for (int64_t b = 0; b < nb; ++b) {
    const uint32_t m = poff[b + 1] - poff[b];
    cnt[b] = m;
    for (int k = 0; k < S; ++k) {
      const uint32_t src = poff[b] + (uint32_t(k) < m ? uint32_t(k) : 0u);
      pids_p[size_t(b) * S + k] = pids[src];
      rstart_p[size_t(b) * S + k] = rstart[src];
    }
    for (int k = 0; k <= S; ++k)
      rbase_p[size_t(b) * (S + 1) + k] = rbase[poff[b] + b + std::min<uint32_t>(uint32_t(k), m)];
  }
  HIP_TRY(c, dalloc(&c->pnode_cnt, cnt.size()));
  HIP_TRY(c, dalloc(&c->pnode_ids, pids_p.size()));
  HIP_TRY(c, dalloc(&c->run_start, rstart_p.size()));
  HIP_TRY(c, dalloc(&c->run_base, rbase_p.size()));
  // patch-local node and rank packed in 16 bits per (e, a) when every patch
  // has at most 256 nodes: one vector load per element
  if (maxm <= 256) {
    const int np = n == 3 ? 4 : n;
    std::vector<uint16_t> lcr(size_t(E) * np, 0);
    for (int64_t e = 0; e < E; ++e)
      for (int a = 0; a < n; ++a)
        lcr[size_t(e) * np + a] =
            uint16_t(lconn[size_t(a) * E + e] | (uint16_t(lrank[size_t(a) * E + e]) << 8));
    HIP_TRY(c, dalloc(&c->lcr, lcr.size()));
    HIP_TRY(c, hipMemcpy(c->lcr, lcr.data(), lcr.size() * 2, hipMemcpyHostToDevice));
  } else {
    HIP_TRY(c, dalloc(&c->lconn, lconn.size()));
    HIP_TRY(c, dalloc(&c->lrank, lrank.size()));
    HIP_TRY(c, hipMemcpy(c->lconn, lconn.data(), lconn.size() * 2, hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->lrank, lrank.data(), lrank.size(), hipMemcpyHostToDevice));
  }
  HIP_TRY(c, hipMemcpy(c->pnode_cnt, cnt.data(), cnt.size() * 4, hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(c->pnode_ids, pids_p.data(), pids_p.size() * 4, hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(c->run_start, rstart_p.data(), rstart_p.size() * 4, hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(c->run_base, rbase_p.data(), rbase_p.size() * 2, hipMemcpyHostToDevice));
  // node -> patch-node entries b * S + k (ascending patch), for the per-patch
  // partial sums
  {
    std::vector<uint32_t> off(size_t(c->N) + 1, 0), ent(pids.size());
    for (uint32_t v : pids) off[v + 1]++;
    for (int64_t k = 0; k < c->N; ++k) off[k + 1] += off[k];
    std::vector<uint32_t> fill(off.begin(), off.end() - 1);
    for (int64_t b = 0; b < nb; ++b)
      for (uint32_t k = 0; k < cnt[b]; ++k)
        ent[fill[pids[poff[b] + k]]++] = uint32_t(b * S + k);
    // the inverse, for node-ordered partials: entry b * S + k -> its position
    // in the node-major list (padding entries are never written)
    std::vector<uint32_t> ppos(size_t(nb) * S, 0);
    for (size_t q = 0; q < ent.size(); ++q) ppos[ent[q]] = uint32_t(q);
    HIP_TRY(c, dalloc(&c->pcsr_off, off.size()));
    HIP_TRY(c, dalloc(&c->pentry, ent.size()));
    HIP_TRY(c, dalloc(&c->ppos, ppos.size()));
    HIP_TRY(c, hipMemcpy(c->pcsr_off, off.data(), off.size() * 4, hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->pentry, ent.data(), ent.size() * 4, hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->ppos, ppos.data(), ppos.size() * 4, hipMemcpyHostToDevice));
  }
  c->patch_max_nodes = maxm;
  c->patch_nodes_total = int64_t(pids.size());
  c->patch_lds = int(lds);
  c->patch_blocks = nb;
  c->patch_ok = true;
  return AKX_OK;
}

// Node blocks of the fused FP64 step: for block b, the elements touching its
// nodes in assembly order, the LDS slot of each (element, node) of the block
// and, per chunk of FUSED_CH elements, where each node's run starts.  `order`
// is the assembly order (caller element ids), `off`/`slot` the FP64 CSR.
// Leaves c->fused_ok false for laws with history (the step recomputes
// elements), for several laws in one mesh and for the 8-node hexahedron
// (its FP64 element state alone takes most of the registers; fused, it
// spills and runs one wave per SIMD).
int build_fused(akx_ctx* c, const std::vector<uint32_t>& order, const std::vector<uint32_t>& off,
                const std::vector<uint32_t>& slot) {
  c->fused_ok = false;
  const int64_t E = c->E, N = c->N;
  const int n = c->n, d = c->d;
  if (E == 0) return AKX_OK;
  const int kind = c->mats[c->elem_mat[0]].kind;
  if (kind != AKX_MATERIAL_ELASTIC && kind != AKX_MATERIAL_NEOHOOKEAN) return AKX_OK;
  for (int64_t e = 1; e < E; ++e)
    if (c->mats[c->elem_mat[e]].kind != kind) return AKX_OK;
  if (n > 4 || int64_t(FUSED_CH) * n > 0xFFFF) return AKX_OK;
  std::vector<uint32_t> rank(E);  // assembly rank of each internal element
  for (int64_t r = 0; r < E; ++r) rank[c->elem_inv[order[r]]] = uint32_t(r);
  const int64_t nblk = (N + FUSED_NB - 1) / FUSED_NB;
  std::vector<uint32_t> el_off(nblk + 1, 0), ch_off(nblk + 1, 0), el;
  std::vector<uint16_t> lslot, cbase;
  std::vector<uint32_t> L;
  std::vector<uint32_t> cnt, run;
  for (int64_t b = 0; b < nblk; ++b) {
    const int64_t n0 = b * FUSED_NB, n1 = std::min<int64_t>(N, n0 + FUSED_NB);
    L.clear();
    for (int64_t k = n0; k < n1; ++k)
      for (uint32_t p = off[k]; p < off[k + 1]; ++p) L.push_back(slot[p] / n);
    std::sort(L.begin(), L.end(), [&](uint32_t x, uint32_t y) { return rank[x] < rank[y]; });
    L.erase(std::unique(L.begin(), L.end()), L.end());
    const size_t nch = (L.size() + FUSED_CH - 1) / FUSED_CH;
    auto pos = [&](uint32_t ei) {  // index of ei in L (sorted by rank)
      return size_t(std::lower_bound(L.begin(), L.end(), ei,
                                     [&](uint32_t x, uint32_t y) { return rank[x] < rank[y]; }) -
                    L.begin());
    };
    cnt.assign(nch * FUSED_NB, 0);
    for (int64_t k = n0; k < n1; ++k)
      for (uint32_t p = off[k]; p < off[k + 1]; ++p)
        cnt[(pos(slot[p] / n) / FUSED_CH) * FUSED_NB + size_t(k - n0)]++;
    const size_t cb0 = cbase.size();
    for (size_t ch = 0; ch < nch; ++ch) {
      uint32_t acc = 0;
      for (int k = 0; k < FUSED_NB; ++k) {
        cbase.push_back(uint16_t(acc));
        acc += cnt[ch * FUSED_NB + k];
      }
      cbase.push_back(uint16_t(acc));
    }
    const size_t l0 = lslot.size();
    lslot.resize(l0 + L.size() * n, 0xFFFF);
    // each node's entries in CSR (assembly) order fill its run of each chunk
    run.assign(nch * FUSED_NB, 0);
    for (int64_t k = n0; k < n1; ++k)
      for (uint32_t p = off[k]; p < off[k + 1]; ++p) {
        const uint32_t ei = slot[p] / n, a = slot[p] % n;
        const size_t j = pos(ei), ch = j / FUSED_CH, kl = size_t(k - n0);
        const uint32_t at = cbase[cb0 + ch * (FUSED_NB + 1) + kl] + run[ch * FUSED_NB + kl]++;
        lslot[l0 + j * n + a] = uint16_t(at);
      }
    el.insert(el.end(), L.begin(), L.end());
    el_off[b + 1] = uint32_t(el.size());
    ch_off[b + 1] = ch_off[b] + uint32_t(nch);
  }
  if (el.size() >= (size_t(1) << 32)) return AKX_OK;
  HIP_TRY(c, dalloc(&c->fel_off, el_off.size()));
  HIP_TRY(c, dalloc(&c->fel, el.size()));
  HIP_TRY(c, dalloc(&c->flslot, lslot.size()));
  HIP_TRY(c, dalloc(&c->fch_off, ch_off.size()));
  HIP_TRY(c, dalloc(&c->fcbase, cbase.size()));
  HIP_TRY(c, dalloc(&c->u_alt, size_t(N) * d));
  HIP_TRY(c, hipMemcpy(c->fel_off, el_off.data(), el_off.size() * 4, hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(c->fel, el.data(), el.size() * 4, hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(c->flslot, lslot.data(), lslot.size() * 2, hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(c->fch_off, ch_off.data(), ch_off.size() * 4, hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(c->fcbase, cbase.data(), cbase.size() * 2, hipMemcpyHostToDevice));
  c->fused_blocks = nblk;
  c->fused_elements = int64_t(el.size());
  c->fused_lds = FUSED_CH * n * d * 8;
  c->fused_ok = true;
  return AKX_OK;
}

int read_blocked(akx_ctx* c, std::vector<uint8_t>& internal) {
  const size_t nd = size_t(c->N) * c->d;
  internal.resize(nd);
  HIP_TRY(c, hipMemcpyAsync(internal.data(), c->blocked, nd, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return AKX_OK;
}

// Whether the surrogate set up on this context can run the current mesh and
// materials; sizes its per-point buffers for the current element count.
int surrogate_check(akx_ctx* c) {
  if (!c->sur.setup) return fail(c, AKX_ERR_STATE, "akx_surrogate_setup has not been called");
  if (c->et != AKX_TETRAHEDRON_4)
    return fail(c, AKX_ERR_UNSUPPORTED, "the surrogate element kernel is for tetrahedron_4");
  if (c->sur.mat < 0 || c->sur.mat >= int(c->mats.size()) ||
      c->mats[c->sur.mat].kind != AKX_MATERIAL_NEOHOOKEAN_DAMAGE)
    return fail(c, AKX_ERR_STATE, "the surrogate's material is not a neohookean_damage material");
  for (const auto& m : c->mats)
    if (m.kind == AKX_MATERIAL_NEOHOOKEAN_DAMAGE &&
        (m.lambda != c->sur.lambda || m.mu != c->sur.mu || m.Sd != c->sur.Sd))
      return fail(c, AKX_ERR_UNSUPPORTED,
                  "every neohookean_damage material must be the one the surrogate was set up for");
  (void)hipSetDevice(c->device);
  HIP_TRY(c, c->sur.ensure_points(c->E));
  return AKX_OK;
}

}  // namespace

/* ======================================================================== */
extern "C" {

int akx_ctx_create(int device, unsigned flags, akx_ctx** out) {
  if (!out) return AKX_ERR_ARGUMENT;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return AKX_ERR_HIP;
  if (device < 0 || device >= count) return AKX_ERR_ARGUMENT;
  auto* c = new (std::nothrow) akx_ctx();
  if (!c) return AKX_ERR_ARGUMENT;
  c->device = device;
  c->reorder = !(flags & AKX_CTX_KEEP_NUMBERING);
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
    delete c;
    return AKX_ERR_HIP;
  }
  c->own_stream = true;
  *out = c;
  return AKX_OK;
}

int akx_ctx_destroy(akx_ctx* c) {
  if (!c) return AKX_ERR_ARGUMENT;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  free_mesh_buffers(c);
  for (auto& e : c->pool) (void)hipEventDestroy(e);
  for (auto& e : c->net_ev)
    if (e) (void)hipEventDestroy(e);
  if (c->copy_stream) (void)hipStreamSynchronize(c->copy_stream);
  for (auto& sl : c->aslots) {
    dfree(sl.buf);
    if (sl.snap) (void)hipEventDestroy(sl.snap);
    if (sl.done) (void)hipEventDestroy(sl.done);
  }
  if (c->copy_stream) (void)hipStreamDestroy(c->copy_stream);
  if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return AKX_OK;
}

const char* akx_last_error(const akx_ctx* c) { return c ? c->err.c_str() : "null context"; }

int akx_set_stream(akx_ctx* c, void* s) {
  if (!c) return AKX_ERR_ARGUMENT;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  if (c->own_stream) (void)hipStreamDestroy(c->stream);
  if (s) {
    c->stream = static_cast<hipStream_t>(s);
    c->own_stream = false;
  } else {
    HIP_TRY(c, hipStreamCreate(&c->stream));
    c->own_stream = true;
  }
  return AKX_OK;
}

int akx_mesh_set(akx_ctx* c, int et, int dim, int64_t nb_nodes, const double* nodes,
                 int64_t nb_elements, const uint32_t* conn) {
  if (!c) return AKX_ERR_ARGUMENT;
  if (et < AKX_TRIANGLE_3 || et > AKX_HEXAHEDRON_8)
    return fail(c, AKX_ERR_ARGUMENT, "unknown element type");
  if (dim != elem_dim(et)) return fail(c, AKX_ERR_ARGUMENT, "dimension does not match element type");
  if (nb_nodes <= 0 || nb_elements < 0 || !nodes || (nb_elements > 0 && !conn))
    return fail(c, AKX_ERR_ARGUMENT, "empty or null mesh arrays");
  const int n = elem_nodes(et);
  if (nb_nodes >= int64_t(std::numeric_limits<uint32_t>::max()) ||
      nb_elements * n >= int64_t(std::numeric_limits<uint32_t>::max()))
    return fail(c, AKX_ERR_ARGUMENT, "mesh too large for 32-bit slot indices");
  for (int64_t k = 0; k < nb_elements * n; ++k)
    if (int64_t(conn[k]) >= nb_nodes)
      return fail(c, AKX_ERR_ARGUMENT, "connectivity refers to a missing node");
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  free_mesh_buffers(c);
  c->et = et; c->d = dim; c->n = n; c->Q = elem_qp(et);
  c->N = nb_nodes; c->E = nb_elements;
  c->hX.assign(nodes, nodes + nb_nodes * dim);
  c->hconn.assign(conn, conn + nb_elements * n);
  c->elem_mat.assign(nb_elements, 0);
  c->mats.clear();
  // the surrogate was set up for a material of the previous mesh
  c->sur.enabled = false;
  c->sur.setup = false;
  c->sur.mat = -1;
  c->has_fext = false;
  c->has_bc_inc = false;
  for (double& x : c->bc_inc) x = 0;
  c->dt = 0;
  c->mesh_set = true;
  c->fint_valid = false;

  // numbering
  c->node_perm.resize(nb_nodes);
  c->elem_perm.resize(nb_elements);
  if (c->reorder && nb_elements > 0) {
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int j = 0; j < dim; ++j) {
      lo[j] = std::numeric_limits<double>::max();
      hi[j] = -std::numeric_limits<double>::max();
    }
    for (int64_t i = 0; i < nb_nodes; ++i)
      for (int j = 0; j < dim; ++j) {
        lo[j] = std::min(lo[j], nodes[i * dim + j]);
        hi[j] = std::max(hi[j], nodes[i * dim + j]);
      }
    c->node_perm = morton_order(c->hX, size_t(nb_nodes), dim, lo, hi);
    std::vector<double> cent(size_t(nb_elements) * dim, 0.);
    for (int64_t e = 0; e < nb_elements; ++e)
      for (int a = 0; a < n; ++a)
        for (int j = 0; j < dim; ++j) cent[e * dim + j] += nodes[size_t(conn[e * n + a]) * dim + j] / n;
    c->elem_perm = morton_order(cent, size_t(nb_elements), dim, lo, hi);
  } else {
    std::iota(c->node_perm.begin(), c->node_perm.end(), 0u);
    std::iota(c->elem_perm.begin(), c->elem_perm.end(), 0u);
  }
  c->node_inv.resize(nb_nodes);
  c->elem_inv.resize(nb_elements);
  for (int64_t k = 0; k < nb_nodes; ++k) c->node_inv[c->node_perm[k]] = uint32_t(k);
  for (int64_t k = 0; k < nb_elements; ++k) c->elem_inv[c->elem_perm[k]] = uint32_t(k);

  const size_t nd = size_t(nb_nodes) * dim;
  HIP_TRY(c, dalloc(&c->X, nd));
  HIP_TRY(c, dalloc(&c->u, nd)); HIP_TRY(c, dalloc(&c->v, nd)); HIP_TRY(c, dalloc(&c->a, nd));
  HIP_TRY(c, dalloc(&c->fext, nd)); HIP_TRY(c, dalloc(&c->fint, nd));
  HIP_TRY(c, dalloc(&c->prev, nd)); HIP_TRY(c, dalloc(&c->inc, nd));
  HIP_TRY(c, dalloc(&c->blocked, nd)); HIP_TRY(c, dalloc(&c->mass, size_t(nb_nodes)));
  HIP_TRY(c, dalloc(&c->work, nd));
  for (double* p : {c->u, c->v, c->a, c->fext, c->fint, c->prev, c->inc})
    HIP_TRY(c, hipMemsetAsync(p, 0, nd * sizeof(double), c->stream));
  HIP_TRY(c, hipMemsetAsync(c->blocked, 0, nd, c->stream));
  std::vector<double> Xi(nd);
  to_internal_nodal(c, nodes, Xi.data());
  HIP_TRY(c, hipMemcpyAsync(c->X, Xi.data(), nd * sizeof(double), hipMemcpyHostToDevice, c->stream));
  std::vector<uint32_t> soa(size_t(nb_elements) * n);
  for (int64_t e = 0; e < nb_elements; ++e) {
    const size_t eo = c->elem_perm[e];
    for (int a = 0; a < n; ++a) soa[size_t(a) * nb_elements + e] = c->node_inv[conn[eo * n + a]];
  }
  HIP_TRY(c, dalloc(&c->conn_soa, soa.size()));
  HIP_TRY(c, hipMemcpyAsync(c->conn_soa, soa.data(), soa.size() * sizeof(uint32_t),
                            hipMemcpyHostToDevice, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return AKX_OK;
}

int akx_material_add(akx_ctx* c, int kind, double E, double nu, double rho, int plane_stress,
                     int* mat_id) {
  if (!c) return AKX_ERR_ARGUMENT;
  if (!c->mesh_set) return fail(c, AKX_ERR_STATE, "set the mesh before adding materials");
  if (kind < AKX_MATERIAL_ELASTIC || kind > AKX_MATERIAL_NEOHOOKEAN_DAMAGE)
    return fail(c, AKX_ERR_ARGUMENT, "unknown material kind");
  if (!(rho > 0) || !(E > 0)) return fail(c, AKX_ERR_ARGUMENT, "E and rho must be positive");
  if (kind != AKX_MATERIAL_ELASTIC && c->d == 2 && plane_stress)
    return fail(c, AKX_ERR_UNSUPPORTED, "plane-stress finite-strain laws are not supported");
  if (c->mats.size() >= 65535) return fail(c, AKX_ERR_UNSUPPORTED, "too many materials");
  HostMaterial m;
  m.kind = kind; m.E = E; m.nu = nu; m.rho = rho; m.plane_stress = plane_stress;
  update_params(m, c->d);
  c->mats.push_back(m);
  c->initialized = false;
  if (mat_id) *mat_id = int(c->mats.size()) - 1;
  return AKX_OK;
}

int akx_material_damage(akx_ctx* c, int mat_id, double Sd, const double* Yd) {
  if (!c) return AKX_ERR_ARGUMENT;
  if (mat_id < 0 || mat_id >= int(c->mats.size())) return fail(c, AKX_ERR_ARGUMENT, "bad material id");
  if (c->mats[mat_id].kind != AKX_MATERIAL_NEOHOOKEAN_DAMAGE)
    return fail(c, AKX_ERR_ARGUMENT, "material has no damage");
  if (!(Sd > 0) || !Yd) return fail(c, AKX_ERR_ARGUMENT, "Sd must be positive and Yd given");
  c->mats[mat_id].Sd = Sd;
  c->mats[mat_id].Yd.assign(Yd, Yd + size_t(c->E) * c->Q);
  c->initialized = false;
  return AKX_OK;
}

int akx_assign_materials(akx_ctx* c, const int32_t* em) {
  if (!c) return AKX_ERR_ARGUMENT;
  if (!c->mesh_set) return fail(c, AKX_ERR_STATE, "no mesh");
  for (int64_t e = 0; e < c->E; ++e) {
    const int k = em ? em[e] : 0;
    if (k < 0 || k >= int(c->mats.size())) return fail(c, AKX_ERR_ARGUMENT, "bad material index");
    c->elem_mat[e] = k;
  }
  c->initialized = false;
  return AKX_OK;
}

int akx_init(akx_ctx* c) {
  if (!c) return AKX_ERR_ARGUMENT;
  if (!c->mesh_set) return fail(c, AKX_ERR_STATE, "no mesh");
  if (c->mats.empty()) return fail(c, AKX_ERR_STATE, "no material");
  (void)hipSetDevice(c->device);
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  const int64_t E = c->E, N = c->N;
  const int n = c->n, Q = c->Q, d = c->d, dn = d * n;
  for (int64_t e = 0; e < E; ++e)
    if (c->elem_mat[e] >= int(c->mats.size()))
      return fail(c, AKX_ERR_ARGUMENT, "element assigned to a missing material");
  if (c->sur.enabled) {
    // materials or element count may have changed since akx_surrogate_enable
    const int src = surrogate_check(c);
    if (src) {
      c->sur.enabled = false;
      return src;
    }
  }
  // a graph captured before holds the buffers freed here: the next graph
  // step captures again
  free_graph(c);
  free_derived(c);

  // assembly order: material by material (creation order), ascending element
  // ids within a material (solid_mechanics_model.cc:396-399,
  // material.cc:190-237), node order within an element.
  const int nm = int(c->mats.size());
  std::vector<std::vector<uint32_t>> filt(nm);
  for (int64_t e = 0; e < E; ++e) filt[c->elem_mat[e]].push_back(uint32_t(e));
  std::vector<uint32_t> order;
  order.reserve(E);
  for (auto& f : filt) order.insert(order.end(), f.begin(), f.end());
  std::vector<uint32_t> off, slot;
  build_csr(c, order, off, slot);
  bool ascending = true;
  for (int64_t e = 0; e < E && ascending; ++e) ascending = (order[e] == uint32_t(e));
  for (int k = 0; k < 4; ++k) {
    const int64_t nb = 64 * (k + 1);
    uint32_t mx = 0;
    for (int64_t n0 = 0; n0 < N; n0 += nb) mx = std::max(mx, off[std::min(N, n0 + nb)] - off[n0]);
    c->span_max[k] = mx;
  }
  HIP_TRY(c, dalloc(&c->csr_off, off.size()));
  HIP_TRY(c, dalloc(&c->csr_slot, slot.size()));
  HIP_TRY(c, hipMemcpy(c->csr_off, off.data(), off.size() * 4, hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(c->csr_slot, slot.data(), slot.size() * 4, hipMemcpyHostToDevice));
  auto upload_pos = [&](const std::vector<uint32_t>& sl, uint32_t** dst) -> int {
    // inverse map for the node-ordered slot layout: CSR position of (e, a)
    std::vector<uint32_t> pos(size_t(E) * n);
    for (size_t p = 0; p < sl.size(); ++p) {
      const uint32_t s = sl[p];
      pos[size_t(s % n) * E + s / n] = uint32_t(p);
    }
    HIP_TRY(c, dalloc(dst, pos.size()));
    HIP_TRY(c, hipMemcpy(*dst, pos.data(), pos.size() * 4, hipMemcpyHostToDevice));
    return AKX_OK;
  };
  if (int rc = upload_pos(slot, &c->slot_pos)) return rc;
  if (int rc = build_fused(c, order, off, slot)) return rc;
  {
    // The mixed path needs a fixed summation order but not the reference's:
    // its CSR lists each node's contributions by internal element id, so that
    // neighbouring elements of one block write neighbouring slots.
    std::vector<uint32_t> off2, slot2;
    build_csr(c, c->elem_perm, off2, slot2);
    HIP_TRY(c, dalloc(&c->csr2_off, off2.size()));
    HIP_TRY(c, dalloc(&c->csr2_slot, slot2.size()));
    HIP_TRY(c, hipMemcpy(c->csr2_off, off2.data(), off2.size() * 4, hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->csr2_slot, slot2.data(), slot2.size() * 4, hipMemcpyHostToDevice));
    if (int rc = upload_pos(slot2, &c->slot2_pos)) return rc;
    if (int rc = build_patches(c, slot2)) return rc;
  }
  if (ascending) {
    c->mass_off = c->csr_off;
    c->mass_slot = c->csr_slot;
  } else {
    // lumped mass assembles every element in ascending order (assembleFieldLumped)
    std::vector<uint32_t> asc(E), moff, mslot;
    std::iota(asc.begin(), asc.end(), 0u);
    build_csr(c, asc, moff, mslot);
    HIP_TRY(c, dalloc(&c->mass_off, moff.size()));
    HIP_TRY(c, dalloc(&c->mass_slot, mslot.size()));
    HIP_TRY(c, hipMemcpy(c->mass_off, moff.data(), moff.size() * 4, hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->mass_slot, mslot.data(), mslot.size() * 4, hipMemcpyHostToDevice));
  }

  // launch groups: one per constitutive law present (internal element ids)
  std::vector<int> kinds;
  for (int k = 0; k < nm; ++k)
    if (!filt[k].empty() &&
        std::find(kinds.begin(), kinds.end(), c->mats[k].kind) == kinds.end())
      kinds.push_back(c->mats[k].kind);
  for (int kind : kinds) {
    KindGroup g{kind, 0, nullptr};
    std::vector<uint32_t> ids;
    for (int k = 0; k < nm; ++k)
      if (c->mats[k].kind == kind)
        for (uint32_t e : filt[k]) ids.push_back(c->elem_inv[e]);
    std::sort(ids.begin(), ids.end());
    g.count = int64_t(ids.size());
    if (kinds.size() > 1) {
      HIP_TRY(c, dalloc(&g.filter, ids.size()));
      HIP_TRY(c, hipMemcpy(g.filter, ids.data(), ids.size() * 4, hipMemcpyHostToDevice));
    }
    c->groups.push_back(g);
  }

  // materials
  std::vector<MatParams> mp(nm);
  bool any_damage = false;
  for (int k = 0; k < nm; ++k) {
    const auto& m = c->mats[k];
    MatParams p;
    std::memset(&p, 0, sizeof p);
    p.lambda = m.lambda; p.mu = m.mu; p.rho = m.rho; p.Sd = m.Sd;
    p.lambda_f = float(m.lambda); p.mu_f = float(m.mu); p.Sd_f = float(m.Sd);
    p.inv_Sd_f = m.Sd > 0 ? float(1. / m.Sd) : 0.f;
    p.kind = m.kind;
    mp[k] = p;
    if (m.kind == AKX_MATERIAL_NEOHOOKEAN_DAMAGE) {
      any_damage = true;
      if (m.Yd.size() != size_t(E) * Q)
        return fail(c, AKX_ERR_STATE, "damage material without thresholds (akx_material_damage)");
    }
  }
  HIP_TRY(c, dalloc(&c->mats_d, nm));
  HIP_TRY(c, hipMemcpy(c->mats_d, mp.data(), nm * sizeof(MatParams), hipMemcpyHostToDevice));
  if (nm > 1) {
    std::vector<uint16_t> em(E);
    for (int64_t e = 0; e < E; ++e) em[e] = uint16_t(c->elem_mat[c->elem_perm[e]]);
    HIP_TRY(c, dalloc(&c->elem_mat_d, size_t(E)));
    HIP_TRY(c, hipMemcpy(c->elem_mat_d, em.data(), size_t(E) * 2, hipMemcpyHostToDevice));
  }
  // the second displacement buffer: the mixed path's previous displacement
  // (position form), the fused step's output
  if (!c->u_alt) HIP_TRY(c, dalloc(&c->u_alt, size_t(N) * d));
  HIP_TRY(c, dalloc(&c->pcoef, size_t(N)));
  // the fused mixed step on triangles streams one 32-byte record per
  // evaluation (built here, after the material indices it carries are on
  // the device; X is fixed after akx_init): node ids in 23 bits and material
  // indices in 17, else it gathers through element ids
  if (c->fused_ok && c->et == AKX_TRIANGLE_3 && N <= FUSED_REC_MAX_NODES &&
      int64_t(nm) <= FUSED_REC_MAX_MATERIALS) {
    HIP_TRY(c, dalloc(&c->frec, size_t(c->fused_elements) * 2));
    float* geo = nullptr;  // per element, only to build the records
    HIP_TRY(c, dalloc(&geo, size_t(E) * fused_geo_stride(n, d)));
    hipError_t er = launch_fused_geometry(c->et, c->stream, E, c->X, c->conn_soa, geo);
    if (er == hipSuccess)
      er = launch_fused_records(c->stream, c->fused_elements, E, c->fel, c->conn_soa, geo,
                                c->flslot, c->elem_mat_d, c->frec);
    const hipError_t es = hipStreamSynchronize(c->stream);
    dfree(geo);
    HIP_TRY(c, er);
    HIP_TRY(c, es);
  }
  if (any_damage) {
    std::vector<double> yd(size_t(E) * Q, 0.), ydi(size_t(E) * Q);
    for (int64_t e = 0; e < E; ++e) {
      const auto& m = c->mats[c->elem_mat[e]];
      if (m.kind == AKX_MATERIAL_NEOHOOKEAN_DAMAGE)
        for (int q = 0; q < Q; ++q) yd[size_t(e) * Q + q] = m.Yd[size_t(e) * Q + q];
    }
    to_internal_points(c, yd.data(), ydi.data(), 1);
    HIP_TRY(c, dalloc(&c->Yd, ydi.size()));
    HIP_TRY(c, dalloc(&c->damage, ydi.size()));
    HIP_TRY(c, hipMemcpy(c->Yd, ydi.data(), ydi.size() * 8, hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemset(c->damage, 0, ydi.size() * 8));
  }

  // tables
  const size_t bq = size_t(E) * Q * dn, jq = size_t(E) * Q;
  HIP_TRY(c, dalloc(&c->B64, bq)); HIP_TRY(c, dalloc(&c->B32, bq));
  HIP_TRY(c, dalloc(&c->J64, jq)); HIP_TRY(c, dalloc(&c->J32, jq));
  HIP_TRY(c, hipMalloc(&c->slots, std::max<size_t>(1, size_t(E) * dn) * sizeof(double)));
  HIP_TRY(c, dalloc(&c->flag, 1));
  HIP_TRY(c, hipMemset(c->flag, 0, sizeof(int)));
  c->partial_blocks = 4096;
  HIP_TRY(c, dalloc(&c->partial, size_t(c->partial_blocks)));
  HIP_TRY(c, dalloc(&c->scalar, 4));
  const ElemTables tab = make_tables(c->et);
  for (int q = 0; q < 8; ++q) {
    for (int k = 0; k < 24; ++k) c->tabf.dnds[q][k] = float(tab.dnds[q][k]);
    c->tabf.w[q] = float(tab.w[q]);
  }
  HIP_TRY(c, dalloc(&c->tab64_d, 1));
  HIP_TRY(c, hipMemcpy(c->tab64_d, &tab, sizeof(ElemTables), hipMemcpyHostToDevice));
  HIP_TRY(c, launch_precompute(c->et, c->stream, E, N, c->X, c->conn_soa, tab, c->B64, c->J64,
                               c->B32, c->J32, c->flag));
  int bad = 0;
  HIP_TRY(c, hipMemcpyAsync(&bad, c->flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (bad)
    return fail(c, AKX_ERR_JACOBIAN,
                "Negative jacobian computed, possible problem in the element node ordering");
  HIP_TRY(c, launch_mass(c->et, c->stream, N, E, c->mass_off, c->mass_slot, tab, c->J64,
                         c->elem_mat_d, c->mats_d, c->mass));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  c->initialized = true;
  c->fint_valid = false;
  return AKX_OK;
}

int akx_field_upload(akx_ctx* c, int field, const void* host, size_t bytes) {
  if (!c) return AKX_ERR_ARGUMENT;
  if (!c->mesh_set) return fail(c, AKX_ERR_STATE, "no mesh");
  if (!host) return fail(c, AKX_ERR_ARGUMENT, "null host pointer");
  const size_t nd = size_t(c->N) * c->d;
  (void)hipSetDevice(c->device);
  if (field == AKX_BLOCKED_DOFS) {
    if (bytes != nd) return fail(c, AKX_ERR_ARGUMENT, "blocked: size mismatch");
    std::vector<uint8_t> cur, h(nd);
    int rc = read_blocked(c, cur);
    if (rc) return rc;
    to_internal_nodal(c, static_cast<const uint8_t*>(host), h.data());
    // bit0: blocked; bit1 (per-step Dirichlet increment) survives on DOFs that stay blocked
    for (size_t k = 0; k < nd; ++k) cur[k] = uint8_t(h[k] ? (1 | (cur[k] & 2)) : 0);
    HIP_TRY(c, hipMemcpy(c->blocked, cur.data(), nd, hipMemcpyHostToDevice));
    return AKX_OK;
  }
  if (field == AKX_DAMAGE) {
    if (!c->damage) return fail(c, AKX_ERR_STATE, "no damage material");
    if (bytes != size_t(c->E) * c->Q * 8) return fail(c, AKX_ERR_ARGUMENT, "damage: size mismatch");
    std::vector<double> tmp(size_t(c->E) * c->Q);
    to_internal_points(c, static_cast<const double*>(host), tmp.data(), 1);
    HIP_TRY(c, hipMemcpy(c->damage, tmp.data(), bytes, hipMemcpyHostToDevice));
    return AKX_OK;
  }
  double* p = nodal_ptr(c, field);
  if (!p || field == AKX_DISPLACEMENT_INCREMENT)
    return fail(c, AKX_ERR_ARGUMENT, "field cannot be uploaded");
  // the internal force is writable as the reference's getInternalForce() is:
  // Material::assembleInternalForces adds into it
  if (bytes != nd * 8) return fail(c, AKX_ERR_ARGUMENT, "size mismatch");
  std::vector<double> tmp(nd);
  to_internal_nodal(c, static_cast<const double*>(host), tmp.data());
  HIP_TRY(c, hipMemcpyAsync(p, tmp.data(), bytes, hipMemcpyHostToDevice, c->stream));
  if (field == AKX_EXTERNAL_FORCE) {
    c->has_fext = false;
    for (size_t k = 0; k < nd && !c->has_fext; ++k) c->has_fext = (tmp[k] != 0.);
  }
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return AKX_OK;
}

int akx_field_download(akx_ctx* c, int field, void* host, size_t bytes) {
  if (!c) return AKX_ERR_ARGUMENT;
  if (!c->mesh_set) return fail(c, AKX_ERR_STATE, "no mesh");
  if (!host) return fail(c, AKX_ERR_ARGUMENT, "null host pointer");
  (void)hipSetDevice(c->device);
  const size_t nd = size_t(c->N) * c->d;
  if (field == AKX_BLOCKED_DOFS) {
    if (bytes != nd) return fail(c, AKX_ERR_ARGUMENT, "blocked: size mismatch");
    std::vector<uint8_t> cur;
    int rc = read_blocked(c, cur);
    if (rc) return rc;
    for (auto& b : cur) b &= 1;
    to_caller_nodal(c, cur.data(), static_cast<uint8_t*>(host));
    return AKX_OK;
  }
  const size_t ne = field_elems(c, field);
  if (bytes != ne * 8) return fail(c, AKX_ERR_ARGUMENT, "size mismatch");
  if (field == AKX_MASS) {
    if (!c->initialized) return fail(c, AKX_ERR_STATE, "akx_init has not been called");
    std::vector<double> m(c->N);
    HIP_TRY(c, hipMemcpyAsync(m.data(), c->mass, c->N * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    double* h = static_cast<double*>(host);
    for (int64_t k = 0; k < c->N; ++k)
      for (int i = 0; i < c->d; ++i) h[size_t(c->node_perm[k]) * c->d + i] = m[k];
    return AKX_OK;
  }
  std::vector<double> tmp(ne);
  if (field == AKX_CURRENT_POSITION) {
    HIP_TRY(c, launch_add(c->stream, int64_t(nd), c->X, c->u, c->work));
    HIP_TRY(c, hipMemcpyAsync(tmp.data(), c->work, bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    to_caller_nodal(c, tmp.data(), static_cast<double*>(host));
    return AKX_OK;
  }
  if (field == AKX_DAMAGE) {
    if (!c->damage) return fail(c, AKX_ERR_STATE, "no damage material");
    HIP_TRY(c, hipMemcpyAsync(tmp.data(), c->damage, bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    to_caller_points(c, tmp.data(), static_cast<double*>(host), 1);
    return AKX_OK;
  }
  if (field == AKX_SHAPES_DERIVATIVES || field == AKX_JACOBIANS) {
    int rc = check_ready(c);
    if (rc) return rc;
    const int comp = field == AKX_JACOBIANS ? 1 : c->d * c->n;
    HIP_TRY(c, hipMemcpyAsync(tmp.data(), field == AKX_JACOBIANS ? c->J64 : c->B64, bytes,
                              hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    std::vector<double> pts(ne);
    const int64_t E = c->E;
    for (int64_t e = 0; e < E; ++e)
      for (int q = 0; q < c->Q; ++q)
        for (int k = 0; k < comp; ++k)
          pts[(size_t(e) * c->Q + q) * comp + k] = tmp[(size_t(q) * comp + k) * E + e];
    to_caller_points(c, pts.data(), static_cast<double*>(host), comp);
    return AKX_OK;
  }
  if (field == AKX_GRAD_U || field == AKX_STRESS) {
    int rc = check_ready(c);
    if (rc) return rc;
    if (!c->gradu) {
      HIP_TRY(c, dalloc(&c->gradu, ne));
      HIP_TRY(c, dalloc(&c->stress, ne));
    }
    for (auto& g : c->groups) {
      ElemArgs A = elem_args(c, g, AKX_FP64);
      HIP_TRY(c, launch_point_fields(c->et, g.kind, c->stream, A, c->gradu, c->stress));
    }
    HIP_TRY(c, hipMemcpyAsync(tmp.data(), field == AKX_GRAD_U ? c->gradu : c->stress, bytes,
                              hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    to_caller_points(c, tmp.data(), static_cast<double*>(host), c->d * c->d);
    return AKX_OK;
  }
  double* p = nodal_ptr(c, field);
  if (!p) return fail(c, AKX_ERR_ARGUMENT, "unknown field");
  HIP_TRY(c, hipMemcpyAsync(tmp.data(), p, bytes, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  to_caller_nodal(c, tmp.data(), static_cast<double*>(host));
  return AKX_OK;
}

int akx_host_alloc(size_t bytes, void** ptr) {
  if (!ptr) return AKX_ERR_ARGUMENT;
  *ptr = nullptr;
  if (bytes == 0) return AKX_ERR_ARGUMENT;
  return hipHostMalloc(ptr, bytes, hipHostMallocDefault) == hipSuccess ? AKX_OK : AKX_ERR_HIP;
}

int akx_host_free(void* ptr) {
  if (!ptr) return AKX_OK;
  return hipHostFree(ptr) == hipSuccess ? AKX_OK : AKX_ERR_HIP;
}

int akx_field_download_async(akx_ctx* c, int field, void* host, size_t bytes, uint64_t* ticket) {
  if (!c || !ticket) return AKX_ERR_ARGUMENT;
  *ticket = 0;
  if (!c->initialized) return fail(c, AKX_ERR_STATE, "akx_init has not been called");
  if (!host) return fail(c, AKX_ERR_ARGUMENT, "null host pointer");
  (void)hipSetDevice(c->device);
  // caller -> internal numbering on the device, built on first use
  if (c->reorder && !c->node_inv_d) {
    HIP_TRY(c, dalloc(&c->node_inv_d, c->node_inv.size()));
    HIP_TRY(c, hipMemcpy(c->node_inv_d, c->node_inv.data(), c->node_inv.size() * 4,
                         hipMemcpyHostToDevice));
    if (!c->elem_inv.empty()) {
      HIP_TRY(c, dalloc(&c->elem_inv_d, c->elem_inv.size()));
      HIP_TRY(c, hipMemcpy(c->elem_inv_d, c->elem_inv.data(), c->elem_inv.size() * 4,
                           hipMemcpyHostToDevice));
    }
  }
  const uint32_t* ninv = c->reorder ? c->node_inv_d : nullptr;
  const uint32_t* einv = c->reorder ? c->elem_inv_d : nullptr;
  const int d = c->d;
  const void* src = nullptr;
  int64_t rows = c->N;
  int dw = d, sw = d, eb = 8;
  const uint32_t* inv = ninv;
  switch (field) {
  case AKX_BLOCKED_DOFS: src = c->blocked; eb = 1; break;
  case AKX_MASS: src = c->mass; sw = 1; break;
  case AKX_CURRENT_POSITION:
    HIP_TRY(c, launch_add(c->stream, c->N * d, c->X, c->u, c->work));
    src = c->work;
    break;
  case AKX_DAMAGE:
    if (!c->damage) return fail(c, AKX_ERR_STATE, "no damage material");
    src = c->damage; rows = c->E; dw = sw = c->Q; inv = einv;
    break;
  case AKX_GRAD_U:
  case AKX_STRESS: {
    if (int rc = check_ready(c)) return rc;
    const size_t ne = field_elems(c, field);
    if (!c->gradu) {
      HIP_TRY(c, dalloc(&c->gradu, ne));
      HIP_TRY(c, dalloc(&c->stress, ne));
    }
    for (auto& g : c->groups) {
      ElemArgs A = elem_args(c, g, AKX_FP64);
      HIP_TRY(c, launch_point_fields(c->et, g.kind, c->stream, A, c->gradu, c->stress));
    }
    src = field == AKX_GRAD_U ? c->gradu : c->stress;
    rows = c->E; dw = sw = c->Q * d * d; inv = einv;
    break;
  }
  case AKX_SHAPES_DERIVATIVES:
  case AKX_JACOBIANS:
    return fail(c, AKX_ERR_ARGUMENT, "field not available asynchronously (akx_field_download)");
  default:
    src = nodal_ptr(c, field);
    if (!src) return fail(c, AKX_ERR_ARGUMENT, "unknown field");
  }
  const size_t need = size_t(rows) * dw * eb;
  if (bytes != need) return fail(c, AKX_ERR_ARGUMENT, "size mismatch");
  if (!c->copy_stream) HIP_TRY(c, hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
  std::lock_guard<std::mutex> lk(c->amu);
  // a free slot (its copy done), else a new one, else wait for the oldest
  constexpr size_t MAX_SLOTS = 16;
  int pick = -1;
  for (size_t k = 0; k < c->aslots.size() && pick < 0; ++k) {
    auto& sl = c->aslots[k];
    if (sl.busy && hipEventQuery(sl.done) == hipSuccess) sl.busy = false;
    if (!sl.busy) pick = int(k);
  }
  if (pick < 0 && c->aslots.size() < MAX_SLOTS) {
    c->aslots.emplace_back();
    auto& sl = c->aslots.back();
    HIP_TRY(c, hipEventCreateWithFlags(&sl.snap, hipEventDisableTiming));
    HIP_TRY(c, hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
    pick = int(c->aslots.size()) - 1;
  }
  if (pick < 0) {
    pick = 0;
    for (size_t k = 1; k < c->aslots.size(); ++k)
      if (c->aslots[k].ticket < c->aslots[pick].ticket) pick = int(k);
    HIP_TRY(c, hipEventSynchronize(c->aslots[pick].done));
    c->aslots[pick].busy = false;
  }
  auto& sl = c->aslots[pick];
  if (sl.cap < need) {
    dfree(sl.buf);
    sl.cap = 0;
    HIP_TRY(c, hipMalloc(&sl.buf, need));
    sl.cap = need;
  }
  HIP_TRY(c, launch_snapshot(c->stream, rows, dw, sw, inv, src, sl.buf, eb));
  HIP_TRY(c, hipEventRecord(sl.snap, c->stream));
  HIP_TRY(c, hipStreamWaitEvent(c->copy_stream, sl.snap, 0));
  HIP_TRY(c, hipMemcpyAsync(host, sl.buf, need, hipMemcpyDeviceToHost, c->copy_stream));
  HIP_TRY(c, hipEventRecord(sl.done, c->copy_stream));
  sl.ticket = c->next_ticket++;
  sl.busy = true;
  *ticket = sl.ticket;
  return AKX_OK;
}

int akx_download_query(akx_ctx* c, uint64_t ticket, int* done) {
  if (!c || !done) return AKX_ERR_ARGUMENT;
  std::lock_guard<std::mutex> lk(c->amu);
  if (ticket == 0 || ticket >= c->next_ticket) return AKX_ERR_ARGUMENT;
  *done = 1;
  for (auto& sl : c->aslots)
    if (sl.busy && sl.ticket == ticket) {
      const hipError_t q = hipEventQuery(sl.done);
      if (q != hipSuccess && q != hipErrorNotReady) return AKX_ERR_HIP;
      *done = q == hipSuccess;
    }
  return AKX_OK;
}

int akx_download_wait(akx_ctx* c, uint64_t ticket) {
  if (!c) return AKX_ERR_ARGUMENT;
  hipEvent_t ev = nullptr;
  {
    std::lock_guard<std::mutex> lk(c->amu);
    if (ticket == 0 || ticket >= c->next_ticket) return AKX_ERR_ARGUMENT;
    for (auto& sl : c->aslots)
      if (sl.busy && sl.ticket == ticket) ev = sl.done;
  }
  // not found: completed and its slot recycled.  A slot is recycled only
  // after its copy completed, so waiting on a re-recorded event waits for a
  // later copy, never for less than this one.
  if (!ev) return AKX_OK;
  (void)hipSetDevice(c->device);
  return hipEventSynchronize(ev) == hipSuccess ? AKX_OK : AKX_ERR_HIP;
}

int akx_field_device_pointer(akx_ctx* c, int field, void** ptr) {
  if (!c || !ptr) return AKX_ERR_ARGUMENT;
  if (field == AKX_BLOCKED_DOFS) { *ptr = c->blocked; return AKX_OK; }
  if (field == AKX_MASS) { *ptr = c->mass; return AKX_OK; }
  double* p = nodal_ptr(c, field);
  if (!p) return fail(c, AKX_ERR_ARGUMENT, "field has no nodal device buffer");
  *ptr = p;
  return AKX_OK;
}

int akx_node_numbering(akx_ctx* c, int64_t* internal_to_caller) {
  if (!c || !internal_to_caller) return AKX_ERR_ARGUMENT;
  if (!c->mesh_set) return fail(c, AKX_ERR_STATE, "no mesh");
  for (int64_t k = 0; k < c->N; ++k) internal_to_caller[k] = c->node_perm[k];
  return AKX_OK;
}

int akx_set_time_step(akx_ctx* c, double dt) {
  if (!c) return AKX_ERR_ARGUMENT;
  if (!(dt > 0) || !std::isfinite(dt)) return fail(c, AKX_ERR_ARGUMENT, "time step must be > 0");
  c->dt = dt;
  return AKX_OK;
}

namespace {

int check_step_args(akx_ctx* c, int precision) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (precision != AKX_FP64 && precision != AKX_MIXED)
    return fail(c, AKX_ERR_ARGUMENT, "unknown precision");
  if (!(c->dt > 0) && !c->dt_node) return fail(c, AKX_ERR_STATE, "time step not set");
  return AKX_OK;
}

// The launches of nsteps steps on c->stream (recorded into a graph when the
// stream is being captured; timing events only when `timing`).
int enqueue_steps(akx_ctx* c, int nsteps, int precision, bool timing) {
  const int d = c->d;
  int rc = AKX_OK;
  if (timing) {
    const size_t need = 2 * size_t(nsteps) + 1;
    while (c->pool.size() < need) {
      hipEvent_t e;
      HIP_TRY(c, hipEventCreate(&e));
      c->pool.push_back(e);
    }
  }
  // the mixed path steps in the position form: u_{n-1} in place of v and a
  // between the first predictor and the last corrector of a run, u_n and
  // u_{n-1} in c->u and c->u_alt by turns
  const int position = precision == AKX_MIXED ? NODE_POSITION : 0;
  double* ucur = c->u;
  double* uprev = c->u_alt;
  // predictor of step 1 (the position form writes u_1 to the other buffer)
  {
    NodeArgs A = node_args(c, NODE_PREDICT | (nsteps == 1 ? NODE_SAVE_PREV : 0) | position,
                           precision);
    if (position) {
      A.u_out = c->u_alt;
      std::swap(ucur, uprev);
    }
    HIP_TRY(c, launch_node(d, node_precision(c, precision), c->has_fext, c->stream, A, c->node_block));
  }
  if (uses_fused(c, precision)) {
    // one fused launch per step; u alternates between c->u and c->u_alt
    double* unext = uprev;
    const KindGroup& g = c->groups[0];
    for (int s = 1; s <= nsteps; ++s) {
      const bool last = (s == nsteps);
      if (timing && s == 1) HIP_TRY(c, hipEventRecord(c->pool[0], c->stream));
      int mode = NODE_CORRECT | position;
      if (last) mode |= NODE_STORE_FINT | NODE_STORE_INC;
      else mode |= NODE_PREDICT | (s + 1 == nsteps ? NODE_SAVE_PREV : 0);
      NodeArgs NA = node_args(c, mode, precision);
      NA.u = ucur;
      ElemArgs EA = elem_args(c, g, precision);
      FusedArgs F;
      F.el_off = c->fel_off;
      F.el = c->fel;
      F.lslot = c->flslot;
      F.ch_off = c->fch_off;
      F.cbase = c->fcbase;
      F.u_out = unext;
      // B rebuilt from the streamed records (mixed_recompute 1) or from
      // coordinates gathered through element ids (0): the same forces bit
      // for bit
      F.rec = c->mixed_recompute ? c->frec : nullptr;
      F.nblk = c->fused_blocks;
      HIP_TRY(c, launch_fused(c->et, g.kind, precision, c->mixed_recompute != 0, c->has_fext,
                              c->stream, EA, NA, F, c->fused_lds));
      if (timing) {
        HIP_TRY(c, hipEventRecord(c->pool[2 * s - 1], c->stream));
        HIP_TRY(c, hipEventRecord(c->pool[2 * s], c->stream));
      }
      std::swap(ucur, unext);
    }
    // the displacement ends where the last step wrote it
    if (ucur != c->u)
      HIP_TRY(c, launch_snapshot(c->stream, int64_t(c->N) * d, 1, 1, nullptr, ucur, c->u, 8));
  } else
  for (int s = 1; s <= nsteps; ++s) {
    const bool last = (s == nsteps);
    if (timing && s == 1) HIP_TRY(c, hipEventRecord(c->pool[0], c->stream));
    rc = run_elements(c, precision, timing && last, ucur);
    if (rc) return rc;
    if (timing) HIP_TRY(c, hipEventRecord(c->pool[2 * s - 1], c->stream));
    int mode = NODE_CORRECT | position;
    if (last) mode |= NODE_STORE_FINT | NODE_STORE_INC;
    else mode |= NODE_PREDICT | (s + 1 == nsteps ? NODE_SAVE_PREV : 0);
    NodeArgs A = node_args(c, mode, precision);
    if (position) {
      // u_{n+1} over u_{n-1}; the last step leaves u_n in c->u
      A.u = ucur;
      A.u_prev = uprev;
      A.u_out = last ? c->u : uprev;
      if (!last) std::swap(ucur, uprev);
    }
    HIP_TRY(c, launch_node(d, node_precision(c, precision), c->has_fext, c->stream, A, c->node_block));
    if (timing) HIP_TRY(c, hipEventRecord(c->pool[2 * s], c->stream));
  }
  if (timing) {
    // element kernel of step s: pool[2s-2] -> pool[2s-1]; node kernel: pool[2s-1] -> pool[2s]
    HIP_TRY(c, hipEventSynchronize(c->pool[2 * nsteps]));
    double el = 0., nd = 0.;
    for (int s = 1; s <= nsteps; ++s) {
      float a = 0.f, b = 0.f;
      HIP_TRY(c, hipEventElapsedTime(&a, c->pool[2 * s - 2], c->pool[2 * s - 1]));
      HIP_TRY(c, hipEventElapsedTime(&b, c->pool[2 * s - 1], c->pool[2 * s]));
      el += a;
      nd += b;
    }
    c->last_elem_ms = el / nsteps;
    c->last_node_ms = nd / nsteps;
    if (c->net_timed) {
      float a = 0.f;
      HIP_TRY(c, hipEventElapsedTime(&a, c->net_ev[0], c->net_ev[1]));
      c->last_net_ms = a;
      c->net_timed = false;
    }
  }
  return AKX_OK;
}

}  // namespace

int akx_solve_step(akx_ctx* c, int nsteps, int precision) {
  int rc = check_step_args(c, precision);
  if (rc) return rc;
  if (nsteps <= 0) return AKX_OK;
  (void)hipSetDevice(c->device);
  rc = enqueue_steps(c, nsteps, precision, c->timing);
  if (rc) return rc;
  c->fint_valid = true;
  return AKX_OK;
}

int akx_solve_step_graph(akx_ctx* c, int nsteps, int precision) {
  int rc = check_step_args(c, precision);
  if (rc) return rc;
  if (nsteps <= 0) return AKX_OK;
  (void)hipSetDevice(c->device);
  akx_ctx::GraphKey k;
  k.nsteps = nsteps; k.precision = precision;
  k.eb = c->elem_block; k.nb = c->node_block; k.x = c->xcd_remap; k.rc = c->mixed_recompute;
  k.lay = c->node_ordered; k.stage = c->node_stage;
  k.fext = c->has_fext; k.bc = c->has_bc_inc; k.dt = c->dt;
  for (int i = 0; i < 3; ++i) k.bc_inc[i] = c->bc_inc[i];
  k.dt_node = c->dt_node;
  k.surrogate = c->sur.enabled;
  if (!c->graph_exec || !(k == c->graph_key)) {
    free_graph(c);
    HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    rc = enqueue_steps(c, nsteps, precision, false);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return rc ? rc : fail(c, AKX_ERR_HIP, std::string("stream capture: ") + hipGetErrorString(e));
    }
    c->graph = g;
    HIP_TRY(c, hipGraphInstantiate(&c->graph_exec, c->graph, nullptr, nullptr, 0));
    c->graph_key = k;
  }
  HIP_TRY(c, hipGraphLaunch(c->graph_exec, c->stream));
  c->fint_valid = true;
  return AKX_OK;
}

int akx_set_node_time_steps(akx_ctx* c, const double* dt) {
  int rc = check_ready(c);
  if (rc) return rc;
  (void)hipSetDevice(c->device);
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (!dt) {
    dfree(c->dt_node);
    return AKX_OK;
  }
  std::vector<double> h(size_t(c->N));
  for (int64_t k = 0; k < c->N; ++k) {
    const double v = dt[c->node_perm[k]];
    if (!(v > 0) || !std::isfinite(v)) return fail(c, AKX_ERR_ARGUMENT, "time steps must be > 0");
    h[k] = v;
  }
  if (!c->dt_node) HIP_TRY(c, dalloc(&c->dt_node, size_t(c->N)));
  HIP_TRY(c, hipMemcpy(c->dt_node, h.data(), h.size() * 8, hipMemcpyHostToDevice));
  return AKX_OK;
}

int akx_stable_time_step_per_material(akx_ctx* c, double* dt) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (!dt) return AKX_ERR_ARGUMENT;
  (void)hipSetDevice(c->device);
  const int nm = int(c->mats.size());
  if (!c->mat_dt) HIP_TRY(c, dalloc(&c->mat_dt, size_t(nm)));
  HIP_TRY(c, launch_stable_dt_material(c->et, c->stream, c->E, c->X, c->u, c->conn_soa,
                                       c->elem_mat_d, c->mats_d, nm, c->mat_dt));
  std::vector<unsigned long long> bits(nm);
  HIP_TRY(c, hipMemcpyAsync(bits.data(), c->mat_dt, size_t(nm) * 8, hipMemcpyDeviceToHost,
                            c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  for (int m = 0; m < nm; ++m) std::memcpy(dt + m, &bits[m], 8);
  return AKX_OK;
}

int akx_assemble_internal_forces(akx_ctx* c, int precision) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (precision != AKX_FP64 && precision != AKX_MIXED)
    return fail(c, AKX_ERR_ARGUMENT, "unknown precision");
  (void)hipSetDevice(c->device);
  rc = run_elements(c, precision);
  if (rc) return rc;
  NodeArgs A = node_args(c, NODE_GATHER | NODE_STORE_FINT, precision);
  HIP_TRY(c, launch_node(c->d, node_precision(c, precision), c->has_fext, c->stream, A, c->node_block));
  c->fint_valid = true;
  return AKX_OK;
}

int akx_stable_time_step(akx_ctx* c, double* dt) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (!dt) return AKX_ERR_ARGUMENT;
  (void)hipSetDevice(c->device);
  HIP_TRY(c, launch_stable_dt(c->et, c->stream, c->E, c->X, c->u, c->conn_soa, c->elem_mat_d,
                              c->mats_d, c->partial, c->partial_blocks, c->scalar));
  HIP_TRY(c, hipMemcpyAsync(dt, c->scalar, 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return AKX_OK;
}

int akx_energy(akx_ctx* c, int kind, double* value) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (!value) return AKX_ERR_ARGUMENT;
  (void)hipSetDevice(c->device);
  if (kind == AKX_ENERGY_KINETIC || kind == AKX_ENERGY_EXTERNAL_WORK) {
    HIP_TRY(c, launch_energy(kind, c->et, 0, c->stream, c->N, c->E, c->u, c->v, c->mass, c->fext,
                             c->fint, c->blocked, c->conn_soa, c->B64, c->J64, c->elem_mat_d,
                             c->mats_d, c->damage, nullptr, 0, c->partial, c->partial_blocks,
                             c->scalar));
    double r = 0;
    HIP_TRY(c, hipMemcpyAsync(&r, c->scalar, 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    *value = kind == AKX_ENERGY_KINETIC ? r * .5 : r * c->dt;
    return AKX_OK;
  }
  if (kind != AKX_ENERGY_POTENTIAL) return fail(c, AKX_ERR_ARGUMENT, "unknown energy");
  double total = 0.;
  for (const auto& g : c->groups) {
    HIP_TRY(c, launch_energy(kind, c->et, g.kind, c->stream, c->N, c->E, c->u, c->v, c->mass,
                             c->fext, c->fint, c->blocked, c->conn_soa, c->B64, c->J64,
                             c->elem_mat_d, c->mats_d, c->damage, g.filter, g.count, c->partial,
                             c->partial_blocks, c->scalar + 1));
    double r = 0;
    HIP_TRY(c, hipMemcpyAsync(&r, c->scalar + 1, 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    total += r;
  }
  *value = total;
  return AKX_OK;
}

int akx_set_dirichlet_increment(akx_ctx* c, int64_t nb, const int64_t* nodes, int axis,
                                double increment) {
  if (!c) return AKX_ERR_ARGUMENT;
  if (!c->mesh_set) return fail(c, AKX_ERR_STATE, "no mesh");
  if (axis < 0 || axis >= c->d) return fail(c, AKX_ERR_ARGUMENT, "bad axis");
  if (nb < 0 || (nb > 0 && !nodes)) return fail(c, AKX_ERR_ARGUMENT, "bad node list");
  if (c->has_bc_inc && c->bc_inc[axis] != 0. && c->bc_inc[axis] != increment)
    return fail(c, AKX_ERR_UNSUPPORTED, "one increment value per axis");
  for (int64_t k = 0; k < nb; ++k)
    if (nodes[k] < 0 || nodes[k] >= c->N) return fail(c, AKX_ERR_ARGUMENT, "node out of range");
  (void)hipSetDevice(c->device);
  std::vector<uint8_t> cur;
  int rc = read_blocked(c, cur);
  if (rc) return rc;
  for (int64_t k = 0; k < nb; ++k) cur[size_t(c->node_inv[nodes[k]]) * c->d + axis] = 3;
  HIP_TRY(c, hipMemcpy(c->blocked, cur.data(), cur.size(), hipMemcpyHostToDevice));
  c->bc_inc[axis] = increment;
  c->has_bc_inc = true;
  return AKX_OK;
}

int akx_synchronize(akx_ctx* c) {
  if (!c) return AKX_ERR_ARGUMENT;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return AKX_OK;
}

int akx_enable_timing(akx_ctx* c, int on) {
  if (!c) return AKX_ERR_ARGUMENT;
  c->timing = on != 0;
  return AKX_OK;
}

int akx_last_step_timing(akx_ctx* c, double* element_ms, double* node_ms) {
  if (!c) return AKX_ERR_ARGUMENT;
  if (element_ms) *element_ms = c->last_elem_ms;
  if (node_ms) *node_ms = c->last_node_ms;
  return AKX_OK;
}

int akx_query(akx_ctx* c, const char* key, double* value) {
  if (!c || !key || !value) return AKX_ERR_ARGUMENT;
  const std::string k(key);
  if (k == "nb_nodes") *value = double(c->N);
  else if (k == "nb_elements") *value = double(c->E);
  else if (k == "csr_entries") *value = double(c->E) * c->n;
  else if (k == "patch_blocks") *value = double(c->patch_ok ? c->patch_blocks : 0);
  else if (k == "patch_nodes_total") *value = double(c->patch_ok ? c->patch_nodes_total : 0);
  else if (k == "patch_max_nodes") *value = double(c->patch_ok ? c->patch_max_nodes : 0);
  else if (k == "patch_available") *value = c->patch_ok ? 1. : 0.;
  else if (k == "patch_packed") *value = (c->patch_ok && c->lcr) ? 1. : 0.;
  else if (k == "fused_available") *value = c->fused_ok ? 1. : 0.;
  else if (k == "fused_records") *value = c->frec ? 1. : 0.;
  else if (k == "fused_elements") *value = double(c->fused_ok ? c->fused_elements : 0);
  else if (k == "element_block") *value = c->elem_block;
  else if (k == "node_block") *value = c->node_block;
  else if (k == "xcd_remap") *value = c->xcd_remap;
  else if (k == "mixed_recompute") *value = c->mixed_recompute;
  else if (k == "slot_layout") *value = c->node_ordered;
  else if (k == "node_stage") *value = c->node_stage;
  else if (k == "surrogate_eps") *value = c->sur.eps;
  else if (k == "surrogate_stress_scale") *value = c->sur.sS;
  else if (k == "surrogate_energy_scale") *value = c->sur.sW;
  else if (k == "surrogate_samples") *value = double(c->sur.nsamples);
  else if (k == "surrogate_parameters") *value = double(akx::mlp::NPARAM);
  else if (k == "surrogate_enabled") *value = c->sur.enabled ? 1. : 0.;
  else if (k == "surrogate_format") *value = c->sur.infer_bf16 ? AKX_SURROGATE_BF16 : AKX_SURROGATE_FP16;
  else if (k == "surrogate_network_ms") *value = c->last_net_ms;
  else return fail(c, AKX_ERR_ARGUMENT, "unknown query key " + k);
  return AKX_OK;
}

// ---------------------------------------------------------------------------
// constitutive surrogate

int akx_surrogate_setup(akx_ctx* c, int mat_id, double eps_max) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (mat_id < 0 || mat_id >= int(c->mats.size()))
    return fail(c, AKX_ERR_ARGUMENT, "no such material");
  const auto& m = c->mats[mat_id];
  if (m.kind != AKX_MATERIAL_NEOHOOKEAN_DAMAGE)
    return fail(c, AKX_ERR_UNSUPPORTED, "the surrogate stands in for neohookean_damage only");
  if (c->d != 3) return fail(c, AKX_ERR_UNSUPPORTED, "the surrogate is three-dimensional");
  if (!(eps_max > 0) || eps_max > 0.5) return fail(c, AKX_ERR_ARGUMENT, "eps_max in (0, 0.5]");
  (void)hipSetDevice(c->device);
  free_graph(c);
  auto& S = c->sur;
  S.mat = mat_id;
  S.lambda = m.lambda;
  S.mu = m.mu;
  S.Sd = m.Sd;
  S.eps = eps_max;
  S.sS = (m.lambda + 2 * m.mu) * eps_max * eps_max;
  S.sW = (m.lambda + 2 * m.mu) * eps_max * eps_max * eps_max;
  HIP_TRY(c, S.ensure_params());
  HIP_TRY(c, S.refresh(c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  S.setup = true;
  return AKX_OK;
}

int akx_surrogate_generate(akx_ctx* c, int64_t paths, int steps, double inc_std,
                           double yd_scale, double yd_shape, uint64_t seed) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (!c->sur.setup) return fail(c, AKX_ERR_STATE, "akx_surrogate_setup has not been called");
  if (paths <= 0 || steps <= 0 || !(inc_std > 0) || !(yd_scale > 0) || !(yd_shape > 0))
    return fail(c, AKX_ERR_ARGUMENT, "bad training-set parameters");
  (void)hipSetDevice(c->device);
  HIP_TRY(c, c->sur.generate(c->stream, paths, steps, inc_std, yd_scale, yd_shape, seed));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return AKX_OK;
}

int akx_surrogate_train(akx_ctx* c, int iterations, int batch, double lr, uint64_t seed,
                        double* final_loss) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (!c->sur.setup || c->sur.nsamples == 0)
    return fail(c, AKX_ERR_STATE, "no training set (akx_surrogate_generate)");
  if (iterations <= 0 || batch <= 0 || batch % 256 != 0 || !(lr > 0))
    return fail(c, AKX_ERR_ARGUMENT, "iterations > 0, batch a multiple of 256, lr > 0");
  (void)hipSetDevice(c->device);
  HIP_TRY(c, c->sur.train(c->stream, iterations, batch, lr, seed, final_loss));
  return AKX_OK;
}

int akx_surrogate_params(akx_ctx* c, float* params, int64_t count, int upload) {
  if (!c || !params) return AKX_ERR_ARGUMENT;
  if (count != akx::mlp::NPARAM)
    return fail(c, AKX_ERR_ARGUMENT, "the network has " + std::to_string(akx::mlp::NPARAM) +
                                         " parameters");
  (void)hipSetDevice(c->device);
  if (upload) HIP_TRY(c, c->sur.set_params(c->stream, params));
  else HIP_TRY(c, c->sur.get_params(c->stream, params));
  return AKX_OK;
}

int akx_surrogate_gradients(akx_ctx* c, int64_t n, const float* x, const float* t, float* grad,
                            double* loss) {
  if (!c || !x || !t || !grad) return AKX_ERR_ARGUMENT;
  if (n <= 0 || n % 256 != 0 || n > (int64_t(1) << 24))
    return fail(c, AKX_ERR_ARGUMENT, "batch must be a positive multiple of 256");
  (void)hipSetDevice(c->device);
  double l = 0;
  HIP_TRY(c, c->sur.gradients(c->stream, n, x, t, grad, &l));
  if (loss) *loss = l;
  return AKX_OK;
}

int akx_surrogate_infer(akx_ctx* c, int64_t n, const float* x, float* y) {
  if (!c || (n > 0 && (!x || !y)) || n < 0) return AKX_ERR_ARGUMENT;
  (void)hipSetDevice(c->device);
  HIP_TRY(c, c->sur.infer(c->stream, n, x, y));
  return AKX_OK;
}

int akx_surrogate_dataset(akx_ctx* c, int64_t first, int64_t count, float* x, float* t) {
  if (!c || !x || !t) return AKX_ERR_ARGUMENT;
  if (first < 0 || count < 0 || first + count > c->sur.nsamples)
    return fail(c, AKX_ERR_ARGUMENT, "sample range outside the training set");
  (void)hipSetDevice(c->device);
  HIP_TRY(c, hipMemcpy(x, c->sur.X + first * 7, size_t(count) * 28, hipMemcpyDeviceToHost));
  HIP_TRY(c, hipMemcpy(t, c->sur.T + first * 7, size_t(count) * 28, hipMemcpyDeviceToHost));
  return AKX_OK;
}

int akx_surrogate_enable(akx_ctx* c, int on) {
  int rc = check_ready(c);
  if (rc) return rc;
  // the captured steps hold the surrogate switch and its scales
  free_graph(c);
  if (!on) {
    c->sur.enabled = false;
    return AKX_OK;
  }
  rc = surrogate_check(c);
  if (rc) return rc;
  c->sur.enabled = true;
  return AKX_OK;
}

int akx_surrogate_format(akx_ctx* c, int format) {
  if (!c) return AKX_ERR_ARGUMENT;
  if (format != AKX_SURROGATE_FP16 && format != AKX_SURROGATE_BF16)
    return fail(c, AKX_ERR_ARGUMENT, "surrogate format: AKX_SURROGATE_FP16 or AKX_SURROGATE_BF16");
  (void)hipSetDevice(c->device);
  // the captured steps hold the network kernel of the previous format
  free_graph(c);
  auto& S = c->sur;
  S.infer_bf16 = format == AKX_SURROGATE_BF16;
  if (S.params) {
    // the image in the new format, from the same master weights
    HIP_TRY(c, S.refresh(c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
  }
  return AKX_OK;
}

int akx_surrogate_fallbacks(akx_ctx* c, uint64_t* count, int reset) {
  if (!c || !count) return AKX_ERR_ARGUMENT;
  if (!c->sur.fallbacks) {
    *count = 0;
    return AKX_OK;
  }
  (void)hipSetDevice(c->device);
  unsigned long long v = 0;
  HIP_TRY(c, hipMemcpyAsync(&v, c->sur.fallbacks, 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  *count = v;
  if (reset) HIP_TRY(c, hipMemset(c->sur.fallbacks, 0, 8));
  return AKX_OK;
}

int akx_autotune(akx_ctx* c, int precision, int nsteps, int* chosen, double* best_ms) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (precision != AKX_FP64 && precision != AKX_MIXED)
    return fail(c, AKX_ERR_ARGUMENT, "unknown precision");
  if (!(c->dt > 0) && !c->dt_node) return fail(c, AKX_ERR_STATE, "time step not set");
  if (nsteps <= 0) nsteps = 10;
  (void)hipSetDevice(c->device);
  // save the state every step touches
  const size_t nd = size_t(c->N) * c->d;
  std::vector<std::pair<double*, double*>> saved;
  std::vector<std::pair<double*, size_t>> arrays = {{c->u, nd}, {c->v, nd}, {c->a, nd},
                                                    {c->fint, nd}, {c->prev, nd}, {c->inc, nd}};
  if (c->damage) arrays.push_back({c->damage, size_t(c->E) * c->Q});
  for (auto& pa : arrays) {
    double* copy = nullptr;
    HIP_TRY(c, dalloc(&copy, pa.second));
    HIP_TRY(c, hipMemcpyAsync(copy, pa.first, pa.second * 8, hipMemcpyDeviceToDevice, c->stream));
    saved.push_back({pa.first, copy});
  }
  auto restore = [&]() -> int {
    for (size_t k = 0; k < saved.size(); ++k)
      HIP_TRY(c, hipMemcpyAsync(saved[k].first, saved[k].second, arrays[k].second * 8,
                                hipMemcpyDeviceToDevice, c->stream));
    return AKX_OK;
  };
  hipEvent_t t0, t1;
  HIP_TRY(c, hipEventCreate(&t0));
  HIP_TRY(c, hipEventCreate(&t1));
  const int keep[6] = {c->elem_block, c->node_block, c->xcd_remap, c->mixed_recompute,
                       c->node_ordered, c->node_stage};
  int best[6] = {keep[0], keep[1], keep[2], keep[3], keep[4], keep[5]};
  float best_t = std::numeric_limits<float>::max();
  const bool was_timing = c->timing;
  c->timing = false;
  // candidates: slot layout x node staging x table/recompute x XCD mapping x block sizes
  struct Shape { int eb, nb, x, rcmp, lay, stage; };
  std::vector<Shape> cand;
  const int blocks[3] = {64, 128, 256};
  // layouts 2 to 5 (LDS patches) exist on the mixed path only; staging the
  // node kernel's slot range applies to node-ordered slots (layouts 1, 2)
  const int nlay = (precision == AKX_MIXED && c->patch_ok) ? 6 : 2;
  for (int lay = 0; lay < nlay; ++lay)
    for (int stage = 0; stage < ((lay == 1 || lay == 2) ? 2 : 1); ++stage)
      for (int rcmp = (lay >= 2 ? 1 : 0); rcmp < 2; ++rcmp)
        for (int x = 0; x < 2; ++x)
          for (int eb : blocks) {
            if (lay >= 2 && eb != 256) continue;  // the patch kernel has its own block shape
            for (int nb : blocks) {
              if (stage && nb != 256) continue;  // LDS-bound occupancy below 256 threads
              cand.push_back(Shape{eb, nb, x, rcmp, lay, stage});
            }
          }
  // the fused step (layout 6) has its own block shapes; on the mixed path it
  // rebuilds B, from per-evaluation records (mixed_recompute 1, triangles)
  // or from coordinates gathered through element ids (0)
  if (c->fused_ok)
    for (int rcmp = (precision == AKX_MIXED && !c->frec ? 1 : 0); rcmp < 2; ++rcmp)
      for (int x = 0; x < 2; ++x) cand.push_back(Shape{256, 256, x, rcmp, 6, 0});
  for (const Shape& sh : cand) {
          const int eb = sh.eb, nb = sh.nb, x = sh.x, lay = sh.lay;
          c->elem_block = eb; c->node_block = nb; c->xcd_remap = x; c->node_ordered = lay;
          c->node_stage = sh.stage;
          c->mixed_recompute = sh.rcmp;
          if ((rc = restore())) break;
          if ((rc = akx_solve_step(c, 1, precision))) break;  // warm-up
          if (hipEventRecord(t0, c->stream) != hipSuccess) { rc = AKX_ERR_HIP; break; }
          if ((rc = akx_solve_step(c, nsteps, precision))) break;
          if (hipEventRecord(t1, c->stream) != hipSuccess ||
              hipEventSynchronize(t1) != hipSuccess) { rc = AKX_ERR_HIP; break; }
          float ms = 0.f;
          (void)hipEventElapsedTime(&ms, t0, t1);
          if (ms < best_t) {
            best_t = ms;
            best[0] = eb; best[1] = nb; best[2] = x; best[3] = c->mixed_recompute;
            best[4] = lay; best[5] = sh.stage;
          }
  }
  c->timing = was_timing;
  c->elem_block = best[0]; c->node_block = best[1]; c->xcd_remap = best[2];
  c->mixed_recompute = best[3]; c->node_ordered = best[4]; c->node_stage = best[5];
  int rc2 = restore();
  (void)hipStreamSynchronize(c->stream);
  for (auto& s : saved) dfree(s.second);
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  if (rc) return rc;
  if (rc2) return rc2;
  if (chosen) for (int k = 0; k < 6; ++k) chosen[k] = best[k];
  if (best_ms) *best_ms = double(best_t) / nsteps;
  return AKX_OK;
}

int akx_set_launch(akx_ctx* c, int element_block, int node_block, int xcd_remap,
                   int mixed_recompute, int slot_layout, int node_stage) {
  if (!c) return AKX_ERR_ARGUMENT;
  auto ok = [](int b) { return b == 0 || (b >= 64 && b <= 256 && b % 64 == 0); };
  if (!ok(element_block) || !ok(node_block))
    return fail(c, AKX_ERR_ARGUMENT, "block size must be 64, 128, 192 or 256");
  if (slot_layout > 6) return fail(c, AKX_ERR_ARGUMENT, "slot layout is 0 to 6");
  if (element_block) c->elem_block = element_block;
  if (node_block) c->node_block = node_block;
  if (xcd_remap >= 0) c->xcd_remap = xcd_remap ? 1 : 0;
  if (mixed_recompute >= 0) c->mixed_recompute = mixed_recompute ? 1 : 0;
  if (slot_layout >= 0) c->node_ordered = slot_layout;
  if (node_stage >= 0) c->node_stage = node_stage ? 1 : 0;
  return AKX_OK;
}

}  // extern "C"

/* ======================================================================== */
/* FEEngine and Material operators on caller fields (akx_fe.hip)            */
/* ======================================================================== */
namespace {

// A device buffer released at the end of an operator call (the calls
// synchronise before they return).
struct TempBuf {
  void* p = nullptr;
  ~TempBuf() {
    if (p) (void)hipFree(p);
  }
  template <typename T>
  T* get() { return static_cast<T*>(p); }
};

hipError_t temp_alloc(TempBuf& b, size_t bytes) {
  return hipMalloc(&b.p, std::max<size_t>(bytes, 8));
}

// The caller's element list (nullptr: every element, caller order) as
// internal ids on the device; returns the row count through *count.
int fe_element_list(akx_ctx* c, const uint32_t* ids, int64_t nb, TempBuf& dev, int64_t* count) {
  if (ids && nb < 0) return fail(c, AKX_ERR_ARGUMENT, "negative element count");
  const int64_t cnt = ids ? nb : c->E;
  std::vector<uint32_t> internal(size_t(std::max<int64_t>(cnt, 0)));
  for (int64_t t = 0; t < cnt; ++t) {
    const uint32_t e = ids ? ids[t] : uint32_t(t);
    if (int64_t(e) >= c->E) return fail(c, AKX_ERR_ARGUMENT, "element id out of range");
    internal[size_t(t)] = c->elem_inv[e];
  }
  HIP_TRY(c, temp_alloc(dev, internal.size() * 4));
  HIP_TRY(c, hipMemcpy(dev.p, internal.data(), internal.size() * 4, hipMemcpyHostToDevice));
  *count = cnt;
  return AKX_OK;
}

FeArgs fe_args(akx_ctx* c, const uint32_t* filt, int64_t count, int nb) {
  FeArgs A;
  std::memset(&A, 0, sizeof A);
  A.count = count;
  A.filt = filt;
  A.E = c->E;
  A.conn = c->conn_soa;
  A.B = c->B64;
  A.J = c->J64;
  A.tab = make_tables(c->et);
  A.nb = nb;
  return A;
}

// A nodal caller array (N, nb) in the internal node numbering, on the device.
int upload_nodal(akx_ctx* c, const double* host, int nb, TempBuf& dev) {
  const size_t N = size_t(c->N);
  std::vector<double> h(N * nb);
  for (size_t k = 0; k < N; ++k)
    for (int i = 0; i < nb; ++i) h[k * nb + i] = host[size_t(c->node_perm[k]) * nb + i];
  HIP_TRY(c, temp_alloc(dev, h.size() * 8));
  HIP_TRY(c, hipMemcpy(dev.p, h.data(), h.size() * 8, hipMemcpyHostToDevice));
  return AKX_OK;
}

int upload_rows(akx_ctx* c, const double* host, size_t n, TempBuf& dev) {
  HIP_TRY(c, temp_alloc(dev, n * 8));
  if (n) HIP_TRY(c, hipMemcpy(dev.p, host, n * 8, hipMemcpyHostToDevice));
  return AKX_OK;
}

int download_rows(akx_ctx* c, const TempBuf& dev, size_t n, double* host) {
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (n) HIP_TRY(c, hipMemcpy(host, dev.p, n * 8, hipMemcpyDeviceToHost));
  return AKX_OK;
}

// One FEEngine operator: rows_in / rows_out values per listed element.
int fe_op(akx_ctx* c, int op, const double* in, int nb, bool nodal_in, size_t in_per_elem,
          size_t out_per_elem, const uint32_t* ids, int64_t nb_ids, double* out) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (!in || !out) return fail(c, AKX_ERR_ARGUMENT, "null array");
  (void)hipSetDevice(c->device);
  TempBuf filt, din, dout;
  int64_t cnt = 0;
  if ((rc = fe_element_list(c, ids, nb_ids, filt, &cnt))) return rc;
  if (nodal_in) rc = upload_nodal(c, in, nb, din);
  else rc = upload_rows(c, in, size_t(cnt) * in_per_elem, din);
  if (rc) return rc;
  HIP_TRY(c, temp_alloc(dout, size_t(cnt) * out_per_elem * 8));
  FeArgs A = fe_args(c, filt.get<uint32_t>(), cnt, nb);
  A.in = din.get<double>();
  A.out = dout.get<double>();
  HIP_TRY(c, launch_fe(op, c->et, c->stream, A));
  return download_rows(c, dout, size_t(cnt) * out_per_elem, out);
}

MatParams mat_params(const HostMaterial& m) {
  MatParams p;
  std::memset(&p, 0, sizeof p);
  p.lambda = m.lambda; p.mu = m.mu; p.rho = m.rho; p.Sd = m.Sd;
  p.lambda_f = float(m.lambda); p.mu_f = float(m.mu); p.Sd_f = float(m.Sd);
  p.inv_Sd_f = m.Sd > 0 ? float(1. / m.Sd) : 0.f;
  p.kind = m.kind;
  return p;
}

// Caller ids of a material's elements, ascending (its element filter).
std::vector<uint32_t> material_filter(const akx_ctx* c, int mat) {
  std::vector<uint32_t> ids;
  for (int64_t e = 0; e < c->E; ++e)
    if (c->elem_mat[e] == mat) ids.push_back(uint32_t(e));
  return ids;
}

int material_setup(akx_ctx* c, int mat, TempBuf& filt, int64_t* cnt, MatArgs* M) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (mat < 0 || mat >= int(c->mats.size())) return fail(c, AKX_ERR_ARGUMENT, "no such material");
  (void)hipSetDevice(c->device);
  const std::vector<uint32_t> ids = material_filter(c, mat);
  if ((rc = fe_element_list(c, ids.data(), int64_t(ids.size()), filt, cnt))) return rc;
  const size_t ne = size_t(c->E) * c->Q * c->d * c->d;
  if (!c->gradu) {
    HIP_TRY(c, dalloc(&c->gradu, ne));
    HIP_TRY(c, dalloc(&c->stress, ne));
  }
  std::memset(M, 0, sizeof *M);
  M->mat = mat_params(c->mats[mat]);
  M->gradu = c->gradu;
  M->stress = c->stress;
  M->damage = c->damage;
  M->Yd = c->Yd;
  return AKX_OK;
}

}  // namespace

extern "C" {

int akx_fe_gradient_on_integration_points(akx_ctx* c, const double* u, int nb_dof,
                                          const uint32_t* elements, int64_t nb_elements,
                                          double* nablauq) {
  if (!c) return AKX_ERR_ARGUMENT;
  if (nb_dof < 1 || nb_dof > FE_MAX_DOF) return fail(c, AKX_ERR_ARGUMENT, "1 <= nb_dof <= 9");
  return fe_op(c, FE_GRADIENT, u, nb_dof, true, 0, size_t(c->Q) * nb_dof * c->d, elements,
               nb_elements, nablauq);
}

int akx_fe_interpolate_on_integration_points(akx_ctx* c, const double* u, int nb_dof,
                                             const uint32_t* elements, int64_t nb_elements,
                                             double* uq) {
  if (!c) return AKX_ERR_ARGUMENT;
  if (nb_dof < 1 || nb_dof > FE_MAX_DOF) return fail(c, AKX_ERR_ARGUMENT, "1 <= nb_dof <= 9");
  return fe_op(c, FE_INTERPOLATE, u, nb_dof, true, 0, size_t(c->Q) * nb_dof, elements,
               nb_elements, uq);
}

int akx_fe_compute_btd(akx_ctx* c, const double* Ds, int rows, const uint32_t* elements,
                       int64_t nb_elements, double* BtDs) {
  if (!c) return AKX_ERR_ARGUMENT;
  if (rows < 1 || rows > FE_MAX_DOF) return fail(c, AKX_ERR_ARGUMENT, "1 <= rows <= 9");
  return fe_op(c, FE_BTD, Ds, rows, false, size_t(c->Q) * rows * c->d,
               size_t(c->Q) * rows * c->n, elements, nb_elements, BtDs);
}

int akx_fe_compute_ntb(akx_ctx* c, const double* bs, int nb_dof, const uint32_t* elements,
                       int64_t nb_elements, double* Ntbs) {
  if (!c) return AKX_ERR_ARGUMENT;
  if (nb_dof < 1 || nb_dof > 24) return fail(c, AKX_ERR_ARGUMENT, "1 <= nb_dof <= 24");
  return fe_op(c, FE_NTB, bs, nb_dof, false, size_t(c->Q) * nb_dof, size_t(c->Q) * nb_dof * c->n,
               elements, nb_elements, Ntbs);
}

int akx_fe_integrate(akx_ctx* c, const double* f, int nb_dof, const uint32_t* elements,
                     int64_t nb_elements, double* intf) {
  if (!c) return AKX_ERR_ARGUMENT;
  if (nb_dof < 1 || nb_dof > 192) return fail(c, AKX_ERR_ARGUMENT, "1 <= nb_dof <= 192");
  return fe_op(c, FE_INTEGRATE, f, nb_dof, false, size_t(c->Q) * nb_dof, size_t(nb_dof),
               elements, nb_elements, intf);
}

int akx_fe_integrate_on_integration_points(akx_ctx* c, const double* f, int nb_dof,
                                           const uint32_t* elements, int64_t nb_elements,
                                           double* out) {
  if (!c) return AKX_ERR_ARGUMENT;
  if (nb_dof < 1 || nb_dof > 192) return fail(c, AKX_ERR_ARGUMENT, "1 <= nb_dof <= 192");
  return fe_op(c, FE_INTEGRATE_POINTS, f, nb_dof, false, size_t(c->Q) * nb_dof,
               size_t(c->Q) * nb_dof, elements, nb_elements, out);
}

int akx_fe_integrate_scalar(akx_ctx* c, const double* f, const uint32_t* elements,
                            int64_t nb_elements, double* value) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (!f || !value) return fail(c, AKX_ERR_ARGUMENT, "null array");
  (void)hipSetDevice(c->device);
  TempBuf filt, din, per, tmp, sort, res;
  int64_t cnt = 0;
  if ((rc = fe_element_list(c, elements, nb_elements, filt, &cnt))) return rc;
  if (cnt == 0) {
    *value = 0.;
    return AKX_OK;
  }
  if (cnt >= int64_t(1) << 31) return fail(c, AKX_ERR_ARGUMENT, "too many elements");
  if ((rc = upload_rows(c, f, size_t(cnt) * c->Q, din))) return rc;
  HIP_TRY(c, temp_alloc(per, size_t(cnt) * 8));
  HIP_TRY(c, temp_alloc(tmp, size_t(cnt) * 8));
  HIP_TRY(c, temp_alloc(res, 8));
  FeArgs A = fe_args(c, filt.get<uint32_t>(), cnt, 1);
  A.in = din.get<double>();
  A.out = per.get<double>();
  HIP_TRY(c, launch_fe(FE_INTEGRATE, c->et, c->stream, A));
  size_t sort_bytes = 0;
  HIP_TRY(c, launch_reduce_sorted(c->stream, cnt, per.get<double>(), tmp.get<double>(),
                                  res.get<double>(), nullptr, &sort_bytes));
  HIP_TRY(c, temp_alloc(sort, sort_bytes));
  HIP_TRY(c, launch_reduce_sorted(c->stream, cnt, per.get<double>(), tmp.get<double>(),
                                  res.get<double>(), sort.p, &sort_bytes));
  return download_rows(c, res, 1, value);
}

int akx_fe_assemble_field_lumped(akx_ctx* c, const double* field, int nb_dof, double* lumped) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (!field || !lumped) return fail(c, AKX_ERR_ARGUMENT, "null array");
  if (nb_dof < 1 || nb_dof > 24) return fail(c, AKX_ERR_ARGUMENT, "1 <= nb_dof <= 24");
  (void)hipSetDevice(c->device);
  const int64_t E = c->E, N = c->N;
  const int Q = c->Q, n = c->n;
  // the field in internal element order, the lumped vector in internal node order
  std::vector<double> fi(size_t(E) * Q * nb_dof);
  for (int64_t e = 0; e < E; ++e)
    std::memcpy(&fi[size_t(e) * Q * nb_dof], field + size_t(c->elem_perm[e]) * Q * nb_dof,
                sizeof(double) * Q * nb_dof);
  TempBuf din, integ, dl;
  if ((rc = upload_rows(c, fi.data(), fi.size(), din))) return rc;
  if ((rc = upload_nodal(c, lumped, nb_dof, dl))) return rc;
  HIP_TRY(c, temp_alloc(integ, size_t(E) * n * nb_dof * 8));
  FeArgs A = fe_args(c, nullptr, E, nb_dof);
  A.in = din.get<double>();
  A.out = integ.get<double>();
  HIP_TRY(c, launch_fe(FE_LUMPED_ELEMENTS, c->et, c->stream, A));
  HIP_TRY(c, launch_fe_lumped_nodes(c->stream, N, nb_dof, n, c->mass_off, c->mass_slot,
                                    integ.get<double>(), dl.get<double>()));
  std::vector<double> h(size_t(N) * nb_dof);
  if ((rc = download_rows(c, dl, h.size(), h.data()))) return rc;
  for (int64_t k = 0; k < N; ++k)
    for (int i = 0; i < nb_dof; ++i)
      lumped[size_t(c->node_perm[k]) * nb_dof + i] = h[size_t(k) * nb_dof + i];
  return AKX_OK;
}

int akx_material_elements(akx_ctx* c, int mat, int64_t* count, uint32_t* ids) {
  if (!c || !count) return AKX_ERR_ARGUMENT;
  if (mat < 0 || mat >= int(c->mats.size())) return fail(c, AKX_ERR_ARGUMENT, "no such material");
  const std::vector<uint32_t> f = material_filter(c, mat);
  *count = int64_t(f.size());
  if (ids) std::copy(f.begin(), f.end(), ids);
  return AKX_OK;
}

int akx_material_compute_stress(akx_ctx* c, int mat, const double* grad_u, double* stress) {
  if (!c) return AKX_ERR_ARGUMENT;
  if (!grad_u) return fail(c, AKX_ERR_ARGUMENT, "null gradient");
  TempBuf filt, din, dout;
  int64_t cnt = 0;
  MatArgs M;
  int rc = material_setup(c, mat, filt, &cnt, &M);
  if (rc) return rc;
  const size_t rows = size_t(cnt) * c->Q * c->d * c->d;
  if ((rc = upload_rows(c, grad_u, rows, din))) return rc;
  if (stress) HIP_TRY(c, temp_alloc(dout, rows * 8));
  FeArgs A = fe_args(c, filt.get<uint32_t>(), cnt, c->d * c->d);
  A.in = din.get<double>();
  A.out = stress ? dout.get<double>() : nullptr;
  HIP_TRY(c, launch_material(MAT_STRESS, c->et, c->mats[mat].kind, c->stream, A, M));
  if (stress) return download_rows(c, dout, rows, stress);
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return AKX_OK;
}

int akx_material_compute_all_stresses(akx_ctx* c, int mat, double* grad_u, double* stress) {
  if (!c) return AKX_ERR_ARGUMENT;
  TempBuf filt;
  int64_t cnt = 0;
  MatArgs M;
  int rc = material_setup(c, mat, filt, &cnt, &M);
  if (rc) return rc;
  FeArgs A = fe_args(c, filt.get<uint32_t>(), cnt, c->d);
  A.in = c->u;
  HIP_TRY(c, launch_material(MAT_ALL_STRESSES, c->et, c->mats[mat].kind, c->stream, A, M));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (grad_u || stress) {
    // the material's rows, gathered from the internal point fields
    const int Q = c->Q, dd = c->d * c->d;
    const size_t ne = size_t(c->E) * Q * dd;
    std::vector<double> g(ne), s(ne);
    HIP_TRY(c, hipMemcpy(g.data(), c->gradu, ne * 8, hipMemcpyDeviceToHost));
    HIP_TRY(c, hipMemcpy(s.data(), c->stress, ne * 8, hipMemcpyDeviceToHost));
    const std::vector<uint32_t> ids = material_filter(c, mat);
    for (size_t t = 0; t < ids.size(); ++t) {
      const size_t e = c->elem_inv[ids[t]];
      if (grad_u) std::memcpy(grad_u + t * Q * dd, &g[e * Q * dd], sizeof(double) * Q * dd);
      if (stress) std::memcpy(stress + t * Q * dd, &s[e * Q * dd], sizeof(double) * Q * dd);
    }
  }
  return AKX_OK;
}

int akx_material_assemble_internal_forces(akx_ctx* c, int mat) {
  if (!c) return AKX_ERR_ARGUMENT;
  TempBuf filt;
  int64_t cnt = 0;
  MatArgs M;
  int rc = material_setup(c, mat, filt, &cnt, &M);
  if (rc) return rc;
  FeArgs A = fe_args(c, filt.get<uint32_t>(), cnt, c->d);
  A.out = static_cast<double*>(c->slots);
  HIP_TRY(c, launch_material(MAT_FORCES, c->et, c->mats[mat].kind, c->stream, A, M));
  HIP_TRY(c, launch_material_gather(c->stream, c->N, c->d, c->n, mat, c->csr_off, c->csr_slot,
                                    c->elem_mat_d, static_cast<double*>(c->slots), c->fint));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  c->fint_valid = true;
  return AKX_OK;
}

int akx_material_potential_energy(akx_ctx* c, int mat, double* value) {
  if (!c || !value) return AKX_ERR_ARGUMENT;
  TempBuf filt;
  int64_t cnt = 0;
  MatArgs M;
  int rc = material_setup(c, mat, filt, &cnt, &M);
  if (rc) return rc;
  if (cnt == 0) {
    *value = 0.;
    return AKX_OK;
  }
  HIP_TRY(c, launch_energy(AKX_ENERGY_POTENTIAL, c->et, c->mats[mat].kind, c->stream, c->N, c->E,
                           c->u, c->v, c->mass, c->fext, c->fint, c->blocked, c->conn_soa, c->B64,
                           c->J64, c->elem_mat_d, c->mats_d, c->damage, filt.get<uint32_t>(), cnt,
                           c->partial, c->partial_blocks, c->scalar + 1));
  double r = 0;
  HIP_TRY(c, hipMemcpyAsync(&r, c->scalar + 1, 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  *value = r;
  return AKX_OK;
}

}  // extern "C"
