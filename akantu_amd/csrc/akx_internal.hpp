// akx_internal.hpp -- host-side context of libakx.so and the kernel launchers.
#pragma once

#include <hip/hip_runtime.h>

#include <stdint.h>
#include <string>
#include <mutex>
#include <vector>

#include "akx_device.hpp"
#include "akx_mlp.hpp"

namespace akx {

struct HostMaterial {
  int kind = ELASTIC;
  double E = 0, nu = 0, rho = 0, Sd = 0;
  int plane_stress = 0;
  double lambda = 0, mu = 0;
  std::vector<double> Yd;  // per point of the whole mesh (only own points used)
};

// A contiguous launch of the element kernel: one constitutive law.
struct KindGroup {
  int kind;
  int64_t count;        // elements in the group
  uint32_t* filter;     // device element ids, nullptr = all elements 0..E-1
};

// Arguments of the element kernels.
struct ElemArgs {
  int64_t count;             // threads' elements
  const uint32_t* filter;    // nullptr = identity
  int64_t E;                 // total elements (SoA stride)
  const uint32_t* conn;      // [a][E]
  const void* B;             // [(q*dn+k)][E], double or float
  const void* J;             // [q][E], double or float
  const double* u;           // (N, d)
  void* slots;               // [e][dn], double or float
  const uint16_t* elem_mat;  // nullptr = material 0
  const MatParams* mats;
  double* damage;            // [e*Q + q]
  const double* Yd;          // [e*Q + q]
  int xcd_remap;
  int recompute;             // mixed path: rebuild B, det(J) w from X
  const double* X;           // (N, d) reference coordinates
  const uint32_t* slot_pos;  // [a][E] CSR position of (e, a); nullptr = element-ordered slots
  // LDS-patch variant of the mixed kernel (k_element_patch)
  int patch;                 // 1: use it
  int partials;              // write one sum per (patch, node) instead of the runs: 1 FP64, 2 FP32
  const uint32_t* part_pos;  // [b * S + k] node-major position of the partial, nullptr = b * S + k
  int patch_max_nodes;       // most distinct nodes of one patch = stride S of the lists
  int patch_lds;             // dynamic LDS bytes per block
  int64_t patch_blocks;      // number of patches
  const uint32_t* pnode_cnt; // [patch] distinct nodes m of the patch
  const uint32_t* pnode_ids; // [b * S + k] patch nodes (internal ids, ascending; padded)
  const uint32_t* run_start; // [b * S + k] first mixed-CSR position of the node's run
  const uint16_t* run_base;  // [b * (S + 1) + k] prefix of run lengths (padded)
  // patch-local node | rank << 8 per (e, a), element-major with stride
  // lcr_stride (n, 4 for 3-node elements); nullptr: lconn / lrank instead
  const uint16_t* lcr;
  const uint16_t* lconn;     // [a][E] patch-local node index
  const uint8_t* lrank;      // [a][E] rank of (e, a) inside the node's run
  ElemTablesF tabf;
  const ElemTables* tab64;   // FP64 element tables in device memory (FP64 recompute)
  // constitutive surrogate (k_sur_strain / network / k_sur_forces); sur_img == nullptr: plain law
  const void* sur_img;
  const float* sur_b;
  bool sur_bf16;             // image and activations in BF16 (else FP16)
  float sur_eps, sur_sS, sur_sW;
  unsigned long long* sur_fallbacks;
  float* sur_x;              // [count][8]: scaled network inputs + envelope flag
  float* sur_y;              // [count][8]: network outputs
  hipEvent_t net_ev0, net_ev1;  // recorded around the network launch when set
};

struct NodeArgs {
  int64_t N;
  const uint32_t* csr_off;
  const uint32_t* csr_slot;
  const void* slots;
  double* u;
  double* v;
  double* a;
  double* u_out;          // where the kernel writes u (nullptr: u itself)
  const double* u_prev;   // NODE_POSITION: the previous step's displacement
  double* pcoef;          // NODE_POSITION: dt^2 / M per node, written by the first predictor
  const double* fext;
  const double* mass;     // one value per node
  const uint8_t* blocked; // bit0 blocked, bit1 per-step Dirichlet increment
  double* fint_out;
  double* prev;
  double* inc;
  double dt;
  const double* dt_node;  // nullptr = dt for every node
  double bc_inc[3];
  int mode;
  int xcd_remap;
  int node_ordered;       // slots stored in CSR order (contiguous per node)
  int stage_lds;          // > 0: bytes of LDS for the block's slot range (node-ordered only)
  int partials;           // slots are per-patch partial sums, 1 FP64, 2 FP32 (csr_off/csr_slot: node -> entries)
};

// Fused FP64 step over node blocks (stateless laws, slot layout 6): block b
// owns nodes [b * FUSED_NB, (b + 1) * FUSED_NB) and recomputes the elements
// around them in chunks of FUSED_CH, in the reference's assembly order.
constexpr int FUSED_NB = 256, FUSED_CH = 256;
struct FusedArgs {
  const uint32_t* el_off;  // [nblk + 1] offsets into el
  const uint32_t* el;      // internal element ids of each block, assembly order
  const uint16_t* lslot;   // [j * n + a] LDS slot of (element j, node a); 0xFFFF: not a block node
  const uint32_t* ch_off;  // [nblk + 1] first chunk table of each block
  const uint16_t* cbase;   // [chunk * (FUSED_NB + 1) + k] first slot of block node k in the chunk
  double* u_out;           // displacement written by the step (read from NodeArgs::u)
  // mixed path on triangles: per evaluation j (entry of el), everything the
  // element needs but u, 32 bytes read as a stream (nullptr: element ids and
  // gathers instead): rec[2 j] = the element's coordinate differences
  // x_a - x_0, a = 1, 2 (4 FP32); rec[2 j + 1] = fused_record_ids(...)
  const uint4* rec;
  int64_t nblk;
};

// Second word of a fused triangle record: the three node ids in the low 23
// bits of x, y, z; LDS slots 0 and 1 in w (16 bits each, 0xFFFF: not a block
// node); LDS slot 2 in 10 bits (0x3FF: not a block node) and the material in
// 17 bits spread over the high 9 bits of x, y, z.
constexpr int FUSED_REC_NODE_BITS = 23;
constexpr int64_t FUSED_REC_MAX_NODES = int64_t(1) << FUSED_REC_NODE_BITS;
constexpr int FUSED_REC_MAX_MATERIALS = 1 << 17;
constexpr uint32_t FUSED_REC_SLOT2_NONE = 0x3FFu;
__host__ __device__ inline uint4 fused_record_ids(uint32_t n0, uint32_t n1, uint32_t n2,
                                                  uint32_t s0, uint32_t s1, uint32_t s2,
                                                  uint32_t mat) {
  const uint32_t q = s2 == 0xFFFFu ? FUSED_REC_SLOT2_NONE : s2;
  return make_uint4(n0 | (q & 0x1FFu) << 23, n1 | (q >> 9) << 23 | (mat & 0xFFu) << 24,
                    n2 | (mat >> 8) << 23, s0 | s1 << 16);
}

// Floats per element of the fused mixed step's geometry record: (n - 1) d
// coordinate differences, padded to whole 16-byte vectors.
constexpr int fused_geo_stride(int n, int d) { return ((n - 1) * d + 3) & ~3; }

enum NodeMode {
  NODE_CORRECT = 1,     // gather forces, residual, lumped solve, corrector
  NODE_PREDICT = 2,     // Dirichlet increments, save previous, predictor
  NODE_STORE_FINT = 4,  // write the assembled internal force
  NODE_SAVE_PREV = 8,   // previous = u before the predictor
  NODE_STORE_INC = 16,  // increment = u - previous after the corrector
  NODE_GATHER = 32,     // gather forces only (with NODE_STORE_FINT)
  // Position form (mixed path): between the first predictor and the last
  // corrector of a run neither v nor a is read or written, the previous
  // displacement u_{n-1} (NodeArgs::u_prev) is read instead:
  // a_n = (f_ext + f_int) / M, u_{n+1} = u_n + ((u_n - u_{n-1}) + dt (dt a_n))
  // -- the same central difference as the predictor / corrector pair with
  // v_{n+1/2} = (u_{n+1} - u_n) / dt eliminated, two arrays less of traffic;
  // dt (dt a_n) is formed as (f_ext + f_int) dt^2 / M with dt^2 / M per node
  // (NodeArgs::pcoef) in place of the mass and time step.  The first
  // predictor writes u_1 = u_0 + dt (v_0 + dt/2 a_0) to another buffer
  // (NodeArgs::u_out), keeps u_0 and writes dt^2 / M; the last corrector
  // returns v_n = (u_n - u_{n-1}) / dt + dt/2 a_n and a_n.
  NODE_POSITION = 64
};

// Byte offset of the forces in the LDS of k_element_patch with list stride S:
// S nodes x 2d floats, S run starts (u32), S + 1 run bases (u16), 16-aligned.
// Entries the patch kernel reads per step when it sums a node's run; the LDS
// force area carries RUNCHUNK - 1 entries of padding after its end.
constexpr int RUNCHUNK = 4;
inline __host__ __device__ size_t patch_force_offset(int S, int d) {
  return (size_t(S) * (2 * d * 4 + 4) + size_t(S + 1) * 2 + 15) & ~size_t(15);
}

// Network evaluation on n points: scaled inputs x (sx floats per point) ->
// outputs y (sy floats per point) (akx_surrogate.hip).
hipError_t launch_sur_mlp(hipStream_t s, const void* img, const float* imgb, bool bf16,
                          int64_t n, const float* x, int sx, float* y, int sy);

// Launchers (akx_kernels.hip).  All enqueue on `stream` and return the launch status.
hipError_t launch_precompute(int et, hipStream_t s, int64_t E, int64_t N, const double* X,
                             const uint32_t* conn_soa, const ElemTables& tab, double* B64,
                             double* J64, float* B32, float* J32, int* bad_jacobian);
hipError_t launch_mass(int et, hipStream_t s, int64_t N, int64_t E, const uint32_t* off,
                       const uint32_t* slot, const ElemTables& tab, const double* J64,
                       const uint16_t* elem_mat, const MatParams* mats, double* mass);
hipError_t launch_element(int et, int kind, int precision, hipStream_t s, const ElemArgs& A,
                          int block);
hipError_t launch_fused(int et, int kind, int precision, bool recompute, bool has_fext,
                        hipStream_t s, const ElemArgs& EA, const NodeArgs& NA, const FusedArgs& F,
                        int lds);
hipError_t launch_node(int d, int precision, bool has_fext, hipStream_t s, const NodeArgs& A,
                       int block);
hipError_t launch_fused_records(hipStream_t s, int64_t J, int64_t E, const uint32_t* el,
                                const uint32_t* conn_soa, const float* geo, const uint16_t* lslot,
                                const uint16_t* elem_mat, uint4* rec);
hipError_t launch_fused_geometry(int et, hipStream_t s, int64_t E, const double* X,
                                 const uint32_t* conn_soa, float* geo);
hipError_t launch_stable_dt(int et, hipStream_t s, int64_t E, const double* X, const double* u,
                            const uint32_t* conn_soa, const uint16_t* elem_mat,
                            const MatParams* mats, double* partial, int nblocks_max,
                            double* out);
hipError_t launch_stable_dt_material(int et, hipStream_t s, int64_t E, const double* X,
                                     const double* u, const uint32_t* conn,
                                     const uint16_t* elem_mat, const MatParams* mats, int nmat,
                                     unsigned long long* out);
hipError_t launch_energy(int kind_energy, int et, int mat_kind_uniform, hipStream_t s,
                         int64_t N, int64_t E, const double* u, const double* v,
                         const double* mass, const double* fext, const double* fint,
                         const uint8_t* blocked, const uint32_t* conn_soa, const double* B64,
                         const double* J64, const uint16_t* elem_mat, const MatParams* mats,
                         const double* damage, const uint32_t* filter, int64_t count,
                         double* partial, int nblocks_max, double* out);
hipError_t launch_point_fields(int et, int kind, hipStream_t s, const ElemArgs& A, double* gradu,
                               double* stress);
hipError_t launch_add(hipStream_t s, int64_t n, const double* x, const double* y, double* out);
hipError_t launch_snapshot(hipStream_t s, int64_t rows, int dw, int sw, const uint32_t* inv,
                           const void* src, void* dst, int elem_bytes);

// FEEngine / Material operators on caller fields (akx_fe.hip), FP64 in the
// reference's operation order.
constexpr int FE_MAX_DOF = 9;  // components of a nodal field (gradient, interpolation)
enum FeOp {
  FE_GRADIENT, FE_INTERPOLATE, FE_BTD, FE_NTB, FE_INTEGRATE, FE_INTEGRATE_POINTS,
  FE_LUMPED_ELEMENTS
};
enum MatOp { MAT_STRESS, MAT_ALL_STRESSES, MAT_FORCES };
struct FeArgs {
  int64_t count;          // listed elements (rows of the per-element arrays)
  const uint32_t* filt;   // their internal ids, nullptr = 0 .. count-1
  int64_t E;              // total elements (stride of the SoA tables)
  const uint32_t* conn;   // [a][E]
  const double* B;        // [(q*dn + k)][E]
  const double* J;        // [q][E]
  ElemTables tab;         // N of the quadrature points
  const double* in;       // operand: nodal (internal numbering) or per-element rows
  double* out;            // result rows
  int nb;                 // components of the operand (rows of D for computeBtD)
};
struct MatArgs {
  MatParams mat;
  double* gradu;          // [(e*Q + q)][d*d], internal element ids
  double* stress;
  double* damage;         // [e*Q + q]
  const double* Yd;
};
hipError_t launch_fe(int op, int et, hipStream_t s, const FeArgs& A);
hipError_t launch_fe_lumped_nodes(hipStream_t s, int64_t N, int nb, int n, const uint32_t* off,
                                  const uint32_t* slot, const double* integ, double* lumped);
hipError_t launch_reduce_sorted(hipStream_t s, int64_t m, double* a, double* tmp, double* out,
                                void* sort_tmp, size_t* sort_bytes);
hipError_t launch_material(int op, int et, int kind, hipStream_t s, const FeArgs& A,
                           const MatArgs& M);
hipError_t launch_material_gather(hipStream_t s, int64_t N, int d, int n, int mat,
                                  const uint32_t* off, const uint32_t* slot,
                                  const uint16_t* elem_mat, const double* fe, double* fint);

// The constitutive surrogate of one context: training set, FP32 master
// weights with Adam moments, BF16 working copies, batch buffers
// (akx_surrogate.hip).
struct Surrogate {
  // plain law it stands in for and the scaling of inputs / outputs
  bool setup = false, enabled = false;
  int mat = -1;
  double lambda = 0, mu = 0, Sd = 0, eps = 0, sS = 0, sW = 0;
  // parameters
  float *params = nullptr, *m1 = nullptr, *m2 = nullptr, *grad = nullptr;
  int64_t step = 0;
  __bf16* Wbf[5] = {};
  __bf16* WT[5] = {};
  mlp::f16x8* img = nullptr;  // 16-byte units; FP16, or BF16 when infer_bf16
  float* imgb = nullptr;
  bool infer_bf16 = false;
  unsigned long long* fallbacks = nullptr;
  // training set
  float *X = nullptr, *T = nullptr;
  int64_t nsamples = 0;
  // batch buffers
  int bcap = 0;
  __bf16 *X0 = nullptr, *X0T = nullptr, *dY = nullptr, *dYT = nullptr;
  float *Tb = nullptr, *Yb = nullptr, *loss_part = nullptr;
  __bf16* H[5] = {};
  __bf16* HT[5] = {};
  float* part[5] = {};
  float* Zt[4] = {};
  __bf16* dZ[4] = {};
  __bf16* dZT[4] = {};
  float *io_x = nullptr, *io_y = nullptr;
  int64_t io_cap = 0;
  // per-point inputs / outputs of the element step (8 floats per point)
  float *pt_x = nullptr, *pt_y = nullptr;
  int64_t pt_cap = 0;

  ~Surrogate();
  void release();
  void release_batch();
  hipError_t ensure_params();
  hipError_t ensure_batch(int B);
  hipError_t ensure_io(int64_t n);
  hipError_t ensure_points(int64_t n);
  hipError_t refresh(hipStream_t s);
  hipError_t pass(hipStream_t s, int B, bool grad_only, float lr);
  hipError_t batch_loss(hipStream_t s, int B, double* loss);
  hipError_t generate(hipStream_t s, int64_t paths, int steps, double inc_std, double yd_scale,
                      double yd_shape, uint64_t seed);
  hipError_t train(hipStream_t s, int iterations, int B, double lr, uint64_t seed,
                   double* final_loss);
  hipError_t gradients(hipStream_t s, int64_t n, const float* x, const float* t, float* grad,
                       double* loss);
  hipError_t infer(hipStream_t s, int64_t n, const float* x, float* y);
  hipError_t set_params(hipStream_t s, const float* host);
  hipError_t get_params(hipStream_t s, float* host);
};

}  // namespace akx

// The context behind the opaque handle of include/akx.h.
struct akx_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::string err;

  // mesh
  int et = -1, d = 0, n = 0, Q = 0;
  int64_t N = 0, E = 0;
  std::vector<double> hX;
  std::vector<uint32_t> hconn;

  // materials
  std::vector<akx::HostMaterial> mats;
  std::vector<int32_t> elem_mat;
  std::vector<akx::KindGroup> groups;

  // device buffers
  double* X = nullptr;
  uint32_t* conn_soa = nullptr;
  double* B64 = nullptr;
  double* J64 = nullptr;
  float* B32 = nullptr;
  float* J32 = nullptr;
  double *u = nullptr, *v = nullptr, *a = nullptr, *fext = nullptr, *fint = nullptr;
  double* mass = nullptr;  // per node
  double *prev = nullptr, *inc = nullptr;
  uint8_t* blocked = nullptr;
  uint32_t *csr_off = nullptr, *csr_slot = nullptr;
  uint32_t *mass_off = nullptr, *mass_slot = nullptr;  // element-ascending order
  void* slots = nullptr;
  uint16_t* elem_mat_d = nullptr;
  akx::MatParams* mats_d = nullptr;
  double* damage = nullptr;
  double* Yd = nullptr;
  double* partial = nullptr;
  double* scalar = nullptr;
  int partial_blocks = 0;
  double* gradu = nullptr;
  double* stress = nullptr;
  double* work = nullptr;  // N*d scratch (current position)
  int* flag = nullptr;

  bool has_fext = false;
  bool has_bc_inc = false;
  double bc_inc[3] = {0, 0, 0};
  double dt = 0.;
  bool mesh_set = false, initialized = false;
  bool fint_valid = false;

  int elem_block = 256, node_block = 256, xcd_remap = 1, mixed_recompute = 1;
  int node_ordered = 0;             // slot layout (see store_slots)
  int node_stage = 1;               // node kernel copies its slot range to LDS when it fits
  uint32_t span_max[4] = {0, 0, 0, 0};  // largest slot range of a node block of 64, 128, 192, 256
  uint32_t* slot_pos = nullptr;     // [a][E] CSR position of (e, a)
  // mixed path: CSR by internal element id (any fixed order will do there)
  uint32_t *csr2_off = nullptr, *csr2_slot = nullptr, *slot2_pos = nullptr;
  // LDS element patches (slot layout 2)
  bool patch_ok = false;
  int patch_max_nodes = 0, patch_lds = 0;
  int64_t patch_blocks = 0, patch_nodes_total = 0;
  uint32_t *pnode_cnt = nullptr, *pnode_ids = nullptr, *run_start = nullptr;
  uint16_t *run_base = nullptr, *lconn = nullptr, *lcr = nullptr;
  uint8_t* lrank = nullptr;
  // node -> entries of the per-patch partial sums (slot layout 3)
  uint32_t *pcsr_off = nullptr, *pentry = nullptr, *ppos = nullptr;
  // fused FP64 step (slot layout 6)
  bool fused_ok = false;
  int fused_lds = 0;
  int64_t fused_blocks = 0, fused_elements = 0;
  uint32_t *fel_off = nullptr, *fel = nullptr, *fch_off = nullptr;
  uint16_t *flslot = nullptr, *fcbase = nullptr;
  double* u_alt = nullptr;  // second displacement buffer (position form, fused step)
  double* pcoef = nullptr;  // position form: dt^2 / M per node
  uint4* frec = nullptr;     // per-evaluation records of the fused mixed step (triangles)
  akx::ElemTablesF tabf;
  akx::ElemTables* tab64_d = nullptr;
  // internal numbering: perm[internal] = caller id, inv[caller id] = internal
  bool reorder = true;
  std::vector<uint32_t> node_perm, node_inv, elem_perm, elem_inv;

  // one time step per node (internal numbering), nullptr = dt everywhere
  double* dt_node = nullptr;
  unsigned long long* mat_dt = nullptr;  // per-material stable time step (bit patterns)
  // captured step sequence (akx_solve_step_graph), re-captured when any
  // launch argument changes
  struct GraphKey {
    int nsteps = 0, precision = -1, eb = 0, nb = 0, x = 0, rc = 0, lay = 0, stage = 0;
    bool fext = false, bc = false;
    double dt = 0, bc_inc[3] = {0, 0, 0};
    const double* dt_node = nullptr;
    bool surrogate = false;
    bool operator==(const GraphKey& o) const {
      return nsteps == o.nsteps && precision == o.precision && eb == o.eb && nb == o.nb &&
             x == o.x && rc == o.rc && lay == o.lay && stage == o.stage && fext == o.fext &&
             bc == o.bc && dt == o.dt && bc_inc[0] == o.bc_inc[0] && bc_inc[1] == o.bc_inc[1] &&
             bc_inc[2] == o.bc_inc[2] && dt_node == o.dt_node && surrogate == o.surrogate;
    }
  } graph_key;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;

  akx::Surrogate sur;

  // asynchronous downloads: a device snapshot taken on `stream`, copied to
  // page-locked host memory on `copy_stream`; slots are recycled once their
  // copy has completed.  `amu` guards the slot table (akx_download_wait may be
  // called from another host thread).
  struct AsyncSlot {
    void* buf = nullptr;
    size_t cap = 0;
    hipEvent_t snap = nullptr, done = nullptr;
    uint64_t ticket = 0;
    bool busy = false;
  };
  std::vector<AsyncSlot> aslots;
  hipStream_t copy_stream = nullptr;
  uint64_t next_ticket = 1;
  std::mutex amu;
  uint32_t *node_inv_d = nullptr, *elem_inv_d = nullptr;  // caller -> internal

  std::vector<hipEvent_t> pool;  // per-step timing events
  bool timing = false;
  double last_elem_ms = 0, last_node_ms = 0;
  hipEvent_t net_ev[2] = {nullptr, nullptr};  // surrogate network of the last timed step
  bool net_timed = false;
  double last_net_ms = 0;
};
