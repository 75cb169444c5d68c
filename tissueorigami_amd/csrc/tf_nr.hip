// Host driver of libtissue_nr.so: the C ABI declared in include/tissue_nr.h.
//
// tf_nr_create builds, once per mesh, the block pattern of the Jacobian from
// the prism connectivity and the element -> block slot map that makes the
// assembly a deterministic gather; tf_nr_step runs the NR loop of
// Simulation::updateStepNR (Simulation.cpp:6605-6746) on the device.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "tf_internal.h"

namespace tf {
namespace {
thread_local std::string g_err;
}
void set_error(const std::string& s) { g_err = s; }
}  // namespace tf

using namespace tf;

// One Young's modulus time series (YoungsModulusModifier.h:9-40): grids at
// n_t time points, the change rate of each interval (host, (n_t - 1) x ny x nx
// and on the device), the prisms it applies to, and their relative (x, y)
// position in the tissue's bounding box.
struct StiffSeries {
  int32_t n_t = 0, nx = 0, ny = 0, method = 0, linear = 1;
  std::vector<double> times;
  double* rate = nullptr;
  uint8_t* mask = nullptr;
  double* relpos = nullptr;
};

// Growth functions evaluated on the device each step (Simulation::calculateGrowth
// with uniform and grid-based functions): per-prism data, the relative
// positions in the bounding box (current and pinned), the functions and the
// compatible corner data of their grids, the pinning times and the column
// volume conservation.
struct GrowthState {
  bool prisms = false;
  GrowthPrisms gp;
  int32_t *type = nullptr, *place = nullptr, *apical_of = nullptr, *prism_tissue = nullptr,
          *node_rng = nullptr;
  uint8_t *flags = nullptr, *node_excl = nullptr;
  double *weights = nullptr, *rel_z = nullptr;
  int symmetric_y = 0;
  double *bb = nullptr, *raw = nullptr, *relcur = nullptr, *relinit = nullptr, *upos = nullptr;
  bool have_cur = false, have_init = false;
  std::vector<GrowthFn> fns;
  GrowthFn* d_fns = nullptr;
  std::vector<double> h_cg, h_ca;
  std::vector<uint8_t> h_ce;
  double *cg = nullptr, *ca = nullptr;
  uint8_t* ce = nullptr;
  int interp = 1;
  bool pinned = false;
  std::vector<double> pin_times;
  std::vector<char> pin_done;
  int32_t n_columns = 0, layers = 0;
  int32_t* columns = nullptr;
  GrowthStage stage;
  double* vref = nullptr;
};

struct tf_handle {
  int32_t Nn = 0, Ne = 0;  // internal node count (tissues padded to 64-node slices) and prisms
  // tissues of the handle: one for tf_nr_create, several for tf_batch_create
  int32_t n_tissues = 1;
  bool batch = false;
  int32_t n_user = 0;                    // caller's node count (concatenated, unpadded)
  std::vector<int32_t> user_off, int_off;  // per-tissue node offsets (user / internal)
  std::vector<int32_t> elem_off;           // per-tissue prism offsets
  // tiles of the fused solver
  TileMap tm;
  int32_t *d_tile_slice = nullptr, *d_tile_tissue = nullptr, *d_tissue_tile = nullptr;
  int spt = 8;
  int autotuned_spt = 0;
  // Krylov method per precision [FP64, mixed]
  int krylov_method[2] = {KRYLOV_BICGSTAB, KRYLOV_CG};
  int device = 0;
  // this device's copy of the library's static scratch array (see scratch_resolve)
  void* scratch = nullptr;
  // laser ablation (tf_nr_set_ablation): per-prism flags and their list, and
  // per-node flags of the ablated nodes (internal numbering)
  uint8_t* abl_el = nullptr;
  int32_t* abl_list = nullptr;
  uint8_t* abl_node = nullptr;
  std::vector<uint8_t> h_abl_node;  // per internal node, for the DOF flags
  int32_t n_abl = 0;
  bool any_abl_node = false;
  hipStream_t st = nullptr;
  // warm start of a step's first Krylov solve (tf_nr_set_warm_start): per
  // tissue, the number of the handle's own last steps (0..2) whose
  // displacements tf_nr_advance keeps; positions handed in reset it
  bool warm_start = true;
  std::vector<int32_t> n_hist;
  // forcing-term safeguard of the NR loop (tf_nr_set_forcing_safeguard)
  bool forcing_safeguard = true;
  DevPattern pat;
  std::vector<int32_t> h_slice_base, h_col, h_row_len, h_rowptr, h_nbr;
  // contact capacity and element width per slice (node-node packing)
  std::vector<int32_t> h_slice_cap, h_slice_ew;
  PackDev pack;
  // inputs
  int32_t* conn = nullptr;
  double* refpos = nullptr;
  double* invFg = nullptr;
  double* lam = nullptr;
  double* mu = nullptr;
  double* eta = nullptr;
  // Young's modulus time series (YoungsModulusModifier): E and nu as set, the
  // per-prism time-series multiplier, the series in application order
  double* E_d = nullptr;
  double* nu_d = nullptr;
  double* stm = nullptr;
  bool stm_touched = false;
  std::vector<StiffSeries> series;
  double sim_time = 0.0;
  double* extv = nullptr;
  double* area = nullptr;
  uint8_t* dof_flags = nullptr;
  std::vector<uint8_t> h_flags_base;  // dof_flags without the ablation bits
  int32_t* pairs = nullptr;
  int32_t n_pairs = 0;
  std::vector<int32_t> h_pairs;
  // explicit bound operator for slave/master binding, and its buffers
  BoundOp bound;
  std::vector<void*> bound_allocs;
  // two-level preconditioner of the FP32 CG solve (single-tissue handles):
  // cell size of the aggregates (< 0: automatic, sized from the stored blocks;
  // 0: off)
  CoarseOp coarse;
  double coarse_cell = -1.0;
  int32_t coarse_bands = 3;  // bands through the thickness per (x, y) cell
  bool coarse_lag = true;    // coarse inverse of a step's first NR iteration kept for the step
  std::vector<double> h_X;  // node reference coordinates (first prism holding the node)
  std::vector<int32_t> bound_pairs;  // pairs the bound operator was built for
  bool any_visc = false, have_material = false, have_growth = false, have_area = false;
  bool assembled = false;
  int last_mixed = 0;
  // ws.uk / ws.un still hold the inputs of the last assembly (tf_nr_assemble_ex),
  // and last_dt its time step: tf_nr_element_outputs recomputes mixed tangents from them
  bool assembly_inputs_kept = false;
  double last_dt = 0.0;
  // ws.strain holds the last Gauss point's strain of the last assembly of
  // every prism; ptissue maps prisms to tissues in a batch (else null)
  bool have_strain = false;
  int32_t* ptissue = nullptr;
  Workspace ws;
  Launch L;
  Prof prof;
  double prof_el_ms = 0.0, prof_as_ms = 0.0;
  int64_t prof_assemblies = 0;
  hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  std::vector<void*> allocs;
  // device-resident time stepping: growth tensor, rates and surface flags
  double* Fg = nullptr;
  double* rates = nullptr;
  int32_t *surf_place = nullptr, *surf_type = nullptr;
  uint8_t* surf_spans = nullptr;
  bool have_rates = false, have_flags = false, have_positions = false;
  GrowthState gr;
  // adjoint state: position checkpoints of the last tf_adj_forward
  double* adj_ckpt = nullptr;
  int adj_cap = 0, adj_steps = 0;
  double adj_dt = 0.0;
  int32_t *adj_place = nullptr, *adj_type = nullptr;
  uint8_t* adj_spans = nullptr;
  double *adj_s = nullptr, *adj_b = nullptr, *adj_dinvFg = nullptr, *adj_valT = nullptr,
         *adj_minvT = nullptr, *adj_a = nullptr;
};

namespace {

#define TF_CHECK(call)                                                              \
  do {                                                                              \
    hipError_t _e = (call);                                                         \
    if (_e != hipSuccess) {                                                         \
      set_error(std::string(#call) + ": " + hipGetErrorString(_e));                 \
      return TF_ERR_HIP;                                                            \
    }                                                                               \
  } while (0)

template <typename T>
hipError_t dalloc(tf_handle* h, T** p, size_t count) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
  if (e == hipSuccess) {
    h->allocs.push_back(q);
    *p = (T*)q;
  }
  return e;
}

// Frees an array allocated by dalloc and clears the pointer.
template <typename T>
void dfree(tf_handle* h, T** p) {
  if (!*p) return;
  auto it = std::find(h->allocs.begin(), h->allocs.end(), (void*)*p);
  if (it != h->allocs.end()) h->allocs.erase(it);
  (void)hipFree(*p);
  *p = nullptr;
}

tf_status layout_pattern(tf_handle* h, const int32_t* conn);

// Scratch of the allocation-free entry point tf_nr_update_visc_surface_noalloc:
// the per-prism placement, type and spans flags staged on the device (9 bytes
// per prism, each array on a 256-byte boundary).  Every device the library runs
// on holds its own copy of the array until the process ends.  Host code never
// uses the array's name as an address: scratch_resolve gets this device's
// address with hipGetSymbolAddress once, with the device current, and every
// launch and copy uses that pointer.  All calls on one device share the array.
constexpr size_t kScratchBytes = size_t(32) << 20;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// This device's address of the scratch array, resolved on the first call for
// the device (the caller has made it current) and kept per device.
tf_status scratch_resolve(int dev, void** out) {
  static std::mutex mu;
  static void* table[kMaxDevices] = {};
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("device index " + std::to_string(dev) + " beyond the scratch table");
    return TF_ERR_NO_DEVICE;
  }
  std::lock_guard<std::mutex> lock(mu);
  if (!table[dev]) {
    void* p = nullptr;
    TF_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) {
      set_error("hipGetSymbolAddress returned no address for the scratch array");
      return TF_ERR_HIP;
    }
    table[dev] = p;
  }
  *out = table[dev];
  return TF_OK;
}

constexpr size_t align256(size_t n) { return (n + 255) / 256 * 256; }

tf_status check_device(int* dev) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    set_error("no HIP device visible");
    return TF_ERR_NO_DEVICE;
  }
  TF_CHECK(hipGetDevice(dev));
  hipDeviceProp_t prop;
  TF_CHECK(hipGetDeviceProperties(&prop, *dev));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    set_error(std::string("device is ") + prop.gcnArchName + ", library is built for gfx950");
    return TF_ERR_NO_DEVICE;
  }
  return TF_OK;
}

// Block pattern, SELL-64 layout and slot map (host side, once per mesh).
tf_status build_pattern(tf_handle* h, const int32_t* conn) {
  const int Nn = h->Nn, Ne = h->Ne;
  // node -> (element, local node) incidence in element order
  std::vector<int32_t> node_ptr(Nn + 1, 0), node_src((size_t)Ne * 6);
  for (int64_t i = 0; i < (int64_t)Ne * 6; ++i) {
    const int nd = conn[i];
    if (nd < 0 || nd >= Nn) {
      set_error("connectivity references node " + std::to_string(nd) + " outside [0, n_nodes)");
      return TF_ERR_ARG;
    }
    node_ptr[nd + 1]++;
  }
  for (int i = 0; i < Nn; ++i) node_ptr[i + 1] += node_ptr[i];
  {
    std::vector<int32_t> pos(node_ptr.begin(), node_ptr.end() - 1);
    for (int e = 0; e < Ne; ++e)
      for (int a = 0; a < 6; ++a) node_src[pos[conn[6 * e + a]]++] = 6 * e + a;
  }
  // neighbours of every node (ascending), including itself
  std::vector<int32_t>& rowptr = h->h_rowptr;
  std::vector<int32_t>& nbr = h->h_nbr;
  rowptr.assign(Nn + 1, 0);
  nbr.clear();
  nbr.reserve((size_t)Nn * 24);
  std::vector<int32_t> cand;
  for (int R = 0; R < Nn; ++R) {
    cand.clear();
    cand.push_back(R);
    for (int t = node_ptr[R]; t < node_ptr[R + 1]; ++t) {
      const int e = node_src[t] / 6;
      for (int b = 0; b < 6; ++b) cand.push_back(conn[6 * e + b]);
    }
    std::sort(cand.begin(), cand.end());
    cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
    nbr.insert(nbr.end(), cand.begin(), cand.end());
    rowptr[R + 1] = (int32_t)nbr.size();
  }
  DevPattern& P0 = h->pat;
  TF_CHECK(dalloc(h, &P0.node_ptr, Nn + 1));
  TF_CHECK(dalloc(h, &P0.node_src, (size_t)Ne * 6));
  TF_CHECK(hipMemcpy(P0.node_ptr, node_ptr.data(), sizeof(int32_t) * (Nn + 1), hipMemcpyHostToDevice));
  TF_CHECK(hipMemcpy(P0.node_src, node_src.data(), sizeof(int32_t) * node_src.size(), hipMemcpyHostToDevice));
  return layout_pattern(h, conn);
}

// SELL-64 layout of the block pattern, transpose slot map and element slot
// map.  Slices with a contact capacity (h_slice_cap, node-node packing) get
// that many extra columns after their element columns.  Replaces the
// slot-indexed device arrays of an earlier layout.
tf_status layout_pattern(tf_handle* h, const int32_t* conn) {
  const int Nn = h->Nn, Ne = h->Ne;
  const std::vector<int32_t>& rowptr = h->h_rowptr;
  const std::vector<int32_t>& nbr = h->h_nbr;
  const int n_slices = (Nn + kSlice - 1) / kSlice;
  const bool contact = !h->h_slice_cap.empty();
  std::vector<int32_t> ew(n_slices, 0);
  std::vector<int32_t>& sb = h->h_slice_base;
  sb.assign(n_slices + 1, 0);
  std::vector<int32_t>& rl = h->h_row_len;
  rl.assign(Nn, 0);
  int64_t total = 0;
  for (int s = 0; s < n_slices; ++s) {
    int w = 0;
    for (int R = s * kSlice; R < std::min(Nn, (s + 1) * kSlice); ++R) {
      rl[R] = rowptr[R + 1] - rowptr[R];
      w = std::max(w, rl[R]);
    }
    ew[s] = w;
    if (contact) w += h->h_slice_cap[s];
    sb[s] = (int32_t)total;
    total += (int64_t)w * kSlice;
    if (total > INT32_MAX) {
      set_error("mesh too large for 32-bit slot indices");
      return TF_ERR_ARG;
    }
  }
  sb[n_slices] = (int32_t)total;
  std::vector<int32_t>& col = h->h_col;
  col.assign(total, 0);
  std::vector<int32_t> diag(Nn, -1);
  for (int s = 0; s < n_slices; ++s) {
    const int w = (sb[s + 1] - sb[s]) / kSlice;
    for (int lane = 0; lane < kSlice; ++lane) {
      const int R = s * kSlice + lane;
      for (int k = 0; k < w; ++k) {
        const int64_t slot = sb[s] + (int64_t)k * kSlice + lane;
        if (R < Nn && k < rl[R]) {
          col[slot] = nbr[rowptr[R] + k];
          if (col[slot] == R) diag[R] = (int32_t)slot;
        } else {
          col[slot] = std::min(R, Nn - 1);
        }
      }
    }
  }
  auto slot_of = [&](int R, int C) -> int64_t {
    const int32_t* b = nbr.data() + rowptr[R];
    const int32_t* e = nbr.data() + rowptr[R + 1];
    const int32_t* p = std::lower_bound(b, e, C);
    const int k = (int)(p - b);
    return sb[R / kSlice] + (int64_t)k * kSlice + (R % kSlice);
  };
  // transpose slot map: slot of (R, k-th column C) -> slot of (C, R)
  std::vector<int32_t> tslot(total, 0);
  for (int R = 0; R < Nn; ++R)
    for (int k = rowptr[R]; k < rowptr[R + 1]; ++k) {
      const int C = nbr[k];
      const int64_t s_rc = sb[R / kSlice] + (int64_t)(k - rowptr[R]) * kSlice + (R % kSlice);
      tslot[s_rc] = (int32_t)slot_of(C, R);
    }
  for (int s = 0; s < n_slices; ++s) {
    const int w = (sb[s + 1] - sb[s]) / kSlice;
    for (int lane = 0; lane < kSlice; ++lane) {
      const int R = s * kSlice + lane;
      for (int k = (R < Nn ? rl[R] : 0); k < w; ++k) {
        const int64_t slot = sb[s] + (int64_t)k * kSlice + lane;
        tslot[slot] = (int32_t)slot;  // padding: zero block, maps onto itself
      }
    }
  }
  // element block -> J slot contributions, in element order per slot
  std::vector<int32_t> cptr(total + 1, 0);
  std::vector<int32_t> slot_cache((size_t)Ne * 36);
  for (int e = 0; e < Ne; ++e)
    for (int A = 0; A < 6; ++A)
      for (int B = 0; B < 6; ++B) {
        const int R = conn[6 * e + B], C = conn[6 * e + A];
        const int64_t slot = slot_of(R, C);
        slot_cache[(size_t)e * 36 + B * 6 + A] = (int32_t)slot;
        cptr[slot + 1]++;
      }
  for (int64_t i = 0; i < total; ++i) cptr[i + 1] += cptr[i];
  std::vector<int32_t> csrc((size_t)Ne * 36);
  {
    std::vector<int32_t> pos(cptr.begin(), cptr.end() - 1);
    for (int e = 0; e < Ne; ++e)
      for (int ab = 0; ab < 36; ++ab) csrc[pos[slot_cache[(size_t)e * 36 + ab]]++] = e * 36 + ab;
  }
  // upload (replacing the arrays of an earlier layout)
  DevPattern& P = h->pat;
  int32_t** old_arrays[] = {&P.slice_base, &P.col, &P.diag_slot, &P.row_len, &P.tslot,
                            &P.contrib_ptr, &P.contrib_src, &P.slice_ew, &P.slice_cap};
  for (int32_t** a : old_arrays) dfree(h, a);
  P.n_slices = n_slices;
  P.n_slots = total;
  TF_CHECK(dalloc(h, &P.slice_base, n_slices + 1));
  TF_CHECK(dalloc(h, &P.col, total));
  TF_CHECK(dalloc(h, &P.diag_slot, Nn));
  TF_CHECK(dalloc(h, &P.row_len, Nn));
  TF_CHECK(dalloc(h, &P.tslot, total));
  TF_CHECK(dalloc(h, &P.contrib_ptr, total + 1));
  TF_CHECK(dalloc(h, &P.contrib_src, (size_t)Ne * 36));
  if (contact) {
    TF_CHECK(dalloc(h, &P.slice_ew, n_slices));
    TF_CHECK(dalloc(h, &P.slice_cap, n_slices));
    TF_CHECK(hipMemcpy(P.slice_ew, ew.data(), sizeof(int32_t) * n_slices, hipMemcpyHostToDevice));
    TF_CHECK(hipMemcpy(P.slice_cap, h->h_slice_cap.data(), sizeof(int32_t) * n_slices,
                       hipMemcpyHostToDevice));
  }
  h->h_slice_ew = ew;
  TF_CHECK(hipMemcpy(P.slice_base, sb.data(), sizeof(int32_t) * (n_slices + 1), hipMemcpyHostToDevice));
  TF_CHECK(hipMemcpy(P.col, col.data(), sizeof(int32_t) * total, hipMemcpyHostToDevice));
  TF_CHECK(hipMemcpy(P.diag_slot, diag.data(), sizeof(int32_t) * Nn, hipMemcpyHostToDevice));
  TF_CHECK(hipMemcpy(P.row_len, rl.data(), sizeof(int32_t) * Nn, hipMemcpyHostToDevice));
  TF_CHECK(hipMemcpy(P.tslot, tslot.data(), sizeof(int32_t) * total, hipMemcpyHostToDevice));
  TF_CHECK(hipMemcpy(P.contrib_ptr, cptr.data(), sizeof(int32_t) * (total + 1), hipMemcpyHostToDevice));
  TF_CHECK(hipMemcpy(P.contrib_src, csrc.data(), sizeof(int32_t) * csrc.size(), hipMemcpyHostToDevice));
  // operator values follow the slot count
  Workspace& w = h->ws;
  if (w.val) {
    dfree(h, &w.val);
    TF_CHECK(dalloc(h, &w.val, (size_t)total * 9));
  }
  if (w.val32) {
    dfree(h, &w.val32);
    TF_CHECK(dalloc(h, &w.val32, (size_t)total * 9));
  }
  return TF_OK;
}

tf_status alloc_workspace(tf_handle* h) {
  Workspace& w = h->ws;
  const size_t n = 3 * (size_t)h->Nn;
  TF_CHECK(dalloc(h, &w.elemKe, (size_t)h->Ne * 324));
  TF_CHECK(dalloc(h, &w.elemFe, (size_t)h->Ne * 18));
  TF_CHECK(dalloc(h, &w.elemFv, (size_t)h->Ne * 18));
  TF_CHECK(dalloc(h, &w.strain, (size_t)h->Ne * 6));
  if (h->n_tissues > 1) {
    std::vector<int32_t> pt((size_t)h->Ne);
    for (int T = 0; T < h->n_tissues; ++T)
      for (int e = h->elem_off[T]; e < h->elem_off[T + 1]; ++e) pt[e] = T;
    TF_CHECK(dalloc(h, &h->ptissue, (size_t)h->Ne));
    TF_CHECK(hipMemcpy(h->ptissue, pt.data(), sizeof(int32_t) * pt.size(), hipMemcpyHostToDevice));
  }
  TF_CHECK(dalloc(h, &w.val, (size_t)h->pat.n_slots * 9));
  TF_CHECK(dalloc(h, &w.minv, (size_t)h->Nn * 9));
  double** vecs[] = {&w.un, &w.uk, &w.g, &w.du, &w.internal, &w.r, &w.rh, &w.p, &w.v,
                     &w.ph, &w.s, &w.sh, &w.t, &w.x, &w.b, &w.tmp, &w.w};
  for (double** v : vecs) TF_CHECK(dalloc(h, v, n));
  TF_CHECK(hipMemset(w.elemFv, 0, sizeof(double) * (size_t)h->Ne * 18));
  return TF_OK;
}

tf_status ensure_visc(tf_handle* h) {
  if (!h->ws.elemKv) TF_CHECK(dalloc(h, &h->ws.elemKv, (size_t)h->Ne * 324));
  return TF_OK;
}

tf_status ensure_mixed(tf_handle* h) {
  Workspace& w = h->ws;
  if (w.val32) return TF_OK;
  const size_t n = 3 * (size_t)h->Nn;
  TF_CHECK(dalloc(h, &w.elemKe32, (size_t)h->Ne * 324));
  TF_CHECK(dalloc(h, &w.elemKv32, (size_t)h->Ne * 324));
  TF_CHECK(dalloc(h, &w.val32, (size_t)h->pat.n_slots * 9));
  TF_CHECK(dalloc(h, &w.minv32, (size_t)h->Nn * 9));
  float** vecs[] = {&w.r32, &w.rh32, &w.p32, &w.v32, &w.ph32, &w.s32, &w.sh32, &w.t32, &w.x32,
                    &w.w32};
  for (float** v : vecs) TF_CHECK(dalloc(h, v, n));
  return TF_OK;
}

// Workspace of the FP64 GMRES solver (allocated at its first use).
tf_status ensure_gmres(tf_handle* h) {
  Workspace& w = h->ws;
  const int m = kGmresRestart, nt = h->n_tissues;
  if (!w.gV) {
    TF_CHECK(dalloc(h, &w.gV, (size_t)(m + 1) * 3 * h->Nn));
    TF_CHECK(dalloc(h, &w.gsum, (size_t)2 * (m + 2) * nt));
    TF_CHECK(dalloc(h, &w.gfac, nt));
    TF_CHECK(dalloc(h, &w.gy, (size_t)m * nt));
    TF_CHECK(dalloc(h, &w.gact, nt));
    TF_CHECK(dalloc(h, &w.gk, nt));
  }
  if (w.gpart_stride < w.tpart_stride) {
    dfree(h, &w.gpart);
    TF_CHECK(dalloc(h, &w.gpart, (size_t)(m + 2) * w.tpart_stride));
    w.gpart_stride = w.tpart_stride;
  }
  return TF_OK;
}

// Buffers of the warm start and of the forcing-term safeguard (first use).
tf_status ensure_warm(tf_handle* h) {
  Workspace& w = h->ws;
  if (w.x0) return TF_OK;
  const size_t n = 3 * (size_t)h->Nn, nt = (size_t)h->n_tissues;
  TF_CHECK(dalloc(h, &w.x0, n));
  TF_CHECK(dalloc(h, &w.r0, n));
  TF_CHECK(dalloc(h, &w.hist0, n));
  TF_CHECK(dalloc(h, &w.hist1, n));
  TF_CHECK(dalloc(h, &w.rscale, nt));
  TF_CHECK(dalloc(h, &w.guess_use, nt));
  TF_CHECK(dalloc(h, &w.hist_order, nt));
  TF_CHECK(hipHostMalloc((void**)&w.host_rscale, sizeof(double) * nt, hipHostMallocDefault));
  return TF_OK;
}

void reset_history(tf_handle* h) { h->n_hist.assign(h->n_tissues, 0); }

// Tiles of spt slices that never cross a tissue boundary (tissues start on
// slice boundaries).  Rebuilding replaces the device tile table.
tf_status build_tiles(tf_handle* h, int spt) {
  std::vector<int32_t> ts, tt, tissue_tile(h->n_tissues + 1, 0);
  ts.push_back(0);
  for (int T = 0; T < h->n_tissues; ++T) {
    const int s_beg = h->int_off[T] / kSlice;
    const int s_end = (h->int_off[T + 1] + kSlice - 1) / kSlice;
    tissue_tile[T] = (int32_t)tt.size();
    for (int s = s_beg; s < s_end; s += spt) {
      ts.back() = s;
      ts.push_back(std::min(s + spt, s_end));
      tt.push_back(T);
    }
  }
  const int n_tiles = (int)tt.size();
  tissue_tile[h->n_tissues] = n_tiles;
  if (h->d_tile_slice) {
    TF_CHECK(hipFree(h->d_tile_slice));
    TF_CHECK(hipFree(h->d_tile_tissue));
    TF_CHECK(hipFree(h->d_tissue_tile));
  }
  TF_CHECK(hipMalloc(&h->d_tile_slice, sizeof(int32_t) * (n_tiles + 1)));
  TF_CHECK(hipMalloc(&h->d_tile_tissue, sizeof(int32_t) * std::max(n_tiles, 1)));
  TF_CHECK(hipMalloc(&h->d_tissue_tile, sizeof(int32_t) * (h->n_tissues + 1)));
  TF_CHECK(hipMemcpy(h->d_tile_slice, ts.data(), sizeof(int32_t) * (n_tiles + 1), hipMemcpyHostToDevice));
  TF_CHECK(hipMemcpy(h->d_tile_tissue, tt.data(), sizeof(int32_t) * n_tiles, hipMemcpyHostToDevice));
  TF_CHECK(hipMemcpy(h->d_tissue_tile, tissue_tile.data(), sizeof(int32_t) * (h->n_tissues + 1),
                     hipMemcpyHostToDevice));
  Workspace& w = h->ws;
  if (w.tpart_stride < n_tiles) {
    if (w.tpart) TF_CHECK(hipFree(w.tpart));
    TF_CHECK(hipMalloc(&w.tpart, sizeof(double) * 5 * (size_t)n_tiles));
    w.tpart_stride = n_tiles;
  }
  h->tm.n_tiles = n_tiles;
  h->tm.grid = std::max(8, (n_tiles + 7) / 8 * 8);
  h->tm.n_tissues = h->n_tissues;
  h->tm.slices_per_tile = spt;
  h->tm.tile_slice = h->d_tile_slice;
  h->tm.tile_tissue = h->d_tile_tissue;
  h->tm.tissue_tile = h->d_tissue_tile;
  h->spt = spt;
  return TF_OK;
}

tf_status alloc_tissue_state(tf_handle* h) {
  Workspace& w = h->ws;
  const int nt = h->n_tissues;
  TF_CHECK(dalloc(h, &w.tstate, 2 * (size_t)nt));
  TF_CHECK(dalloc(h, &w.tscal, 8 * (size_t)nt));
  TF_CHECK(dalloc(h, &w.tactive, nt));
  TF_CHECK(dalloc(h, &w.tnorm, nt));
  TF_CHECK(hipHostMalloc((void**)&w.host_tstate, 2 * sizeof(int32_t) * nt, hipHostMallocDefault));
  TF_CHECK(hipHostMalloc((void**)&w.host_tnorm, sizeof(double) * nt, hipHostMallocDefault));
  TF_CHECK(dalloc(h, &w.gnorm, nt));
  TF_CHECK(hipHostMalloc((void**)&w.host_gnorm, sizeof(double) * nt, hipHostMallocDefault));
  std::vector<int32_t> ones(nt, 1);
  TF_CHECK(hipMemcpy(w.tactive, ones.data(), sizeof(int32_t) * nt, hipMemcpyHostToDevice));
  return TF_OK;
}

// user (concatenated, unpadded) per-node arrays <-> internal padded layout
template <typename T>
std::vector<T> to_internal(const tf_handle* h, const T* user, int comps, T fill) {
  std::vector<T> out((size_t)h->Nn * comps, fill);
  for (int t = 0; t < h->n_tissues; ++t) {
    const int nu = h->user_off[t + 1] - h->user_off[t];
    std::memcpy(out.data() + (size_t)h->int_off[t] * comps, user + (size_t)h->user_off[t] * comps,
                sizeof(T) * nu * comps);
  }
  return out;
}

template <typename T>
void from_internal(const tf_handle* h, const T* internal, T* user, int comps) {
  for (int t = 0; t < h->n_tissues; ++t) {
    const int nu = h->user_off[t + 1] - h->user_off[t];
    std::memcpy(user + (size_t)h->user_off[t] * comps, internal + (size_t)h->int_off[t] * comps,
                sizeof(T) * nu * comps);
  }
}

// Copies a caller node vector (3 per node) to the device in the internal layout.
tf_status upload_nodes(tf_handle* h, double* dst, const double* user) {
  if (!h->batch) {
    TF_CHECK(hipMemcpyAsync(dst, user, sizeof(double) * 3 * h->Nn, hipMemcpyHostToDevice, h->st));
    return TF_OK;
  }
  std::vector<double> tmp = to_internal<double>(h, user, 3, 0.0);
  TF_CHECK(hipStreamSynchronize(h->st));
  TF_CHECK(hipMemcpy(dst, tmp.data(), sizeof(double) * tmp.size(), hipMemcpyHostToDevice));
  return TF_OK;
}

tf_status download_nodes(tf_handle* h, double* user, const double* src) {
  if (!h->batch) {
    TF_CHECK(hipMemcpyAsync(user, src, sizeof(double) * 3 * h->Nn, hipMemcpyDeviceToHost, h->st));
    TF_CHECK(hipStreamSynchronize(h->st));
    return TF_OK;
  }
  std::vector<double> tmp((size_t)3 * h->Nn);
  TF_CHECK(hipStreamSynchronize(h->st));
  TF_CHECK(hipMemcpy(tmp.data(), src, sizeof(double) * tmp.size(), hipMemcpyDeviceToHost));
  from_internal<double>(h, tmp.data(), user, 3);
  return TF_OK;
}

void free_bound(tf_handle* h) {
  for (void* p : h->bound_allocs) (void)hipFree(p);
  h->bound_allocs.clear();
  h->bound = BoundOp();
  h->coarse.built = false;  // its block lists follow the operator's pattern
  h->coarse.active = false;
}

template <typename T>
hipError_t balloc(tf_handle* h, T** p, size_t count, const T* init) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
  if (e != hipSuccess) return e;
  h->bound_allocs.push_back(q);
  *p = (T*)q;
  if (init && count) e = hipMemcpy(q, init, sizeof(T) * count, hipMemcpyHostToDevice);
  return e;
}

// Pattern, slot map and exception lists of A_b = N^T J N + I_slave for the
// handle's slave/master pairs (internal numbering; in a batch both DOF of a
// pair lie in one tissue, so A_b keeps the tissues' rows apart and keeps J's
// slices).  A master is never a slave and a DOF is the slave of one master at
// most, as the reference's binding lists guarantee (Simulation.cpp:1081-1113).
tf_status build_bound(tf_handle* h) {
  free_bound(h);
  const int Nn = h->Nn, n = 3 * Nn;
  const int np = h->n_pairs;
  std::vector<int32_t> M(n);
  std::iota(M.begin(), M.end(), 0);
  for (int p = 0; p < np; ++p) {
    const int sd = h->h_pairs[2 * p], md = h->h_pairs[2 * p + 1];
    if (M[sd] != sd) {
      set_error("slave DOF " + std::to_string(sd) + " is bound twice");
      return TF_ERR_ARG;
    }
    M[sd] = md;
  }
  for (int p = 0; p < np; ++p)
    if (M[h->h_pairs[2 * p + 1]] != h->h_pairs[2 * p + 1]) {
      set_error("a master DOF is itself a slave");
      return TF_ERR_ARG;
    }
  // M^-1(i): i itself (when not a slave) and its slaves, ascending
  std::vector<std::vector<int32_t>> Minv(n);
  std::vector<uint8_t> node_bound(Nn, 0);  // node has a slave or master DOF
  for (int d = 0; d < n; ++d) {
    if (M[d] == d) Minv[d].push_back(d);
  }
  for (int d = 0; d < n; ++d)
    if (M[d] != d) {
      Minv[M[d]].push_back(d);
      node_bound[d / 3] = 1;
      node_bound[M[d] / 3] = 1;
    }
  for (auto& v : Minv) std::sort(v.begin(), v.end());
  const std::vector<int32_t>& rp = h->h_rowptr;
  const std::vector<int32_t>& nb = h->h_nbr;
  auto slotJ = [&](int R, int C) -> int64_t {
    const int32_t* b = nb.data() + rp[R];
    const int32_t* e = nb.data() + rp[R + 1];
    const int32_t* q = std::lower_bound(b, e, C);
    if (q == e || *q != C) return -1;
    return h->h_slice_base[R / kSlice] + (int64_t)(q - b) * kSlice + (R % kSlice);
  };
  // block pattern of A_b
  std::vector<int32_t> brp(Nn + 1, 0), bnb;
  bnb.reserve(nb.size() + 64);
  std::vector<int32_t> cand;
  for (int R = 0; R < Nn; ++R) {
    cand.clear();
    cand.push_back(R);
    if (!node_bound[R]) {
      for (int k = rp[R]; k < rp[R + 1]; ++k)
        for (int q = 0; q < 3; ++q) cand.push_back(M[3 * nb[k] + q] / 3);
    } else {
      for (int pp = 0; pp < 3; ++pp)
        for (int d : Minv[3 * R + pp])
          for (int k = rp[d / 3]; k < rp[d / 3 + 1]; ++k)
            for (int q = 0; q < 3; ++q) cand.push_back(M[3 * nb[k] + q] / 3);
    }
    std::sort(cand.begin(), cand.end());
    cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
    bnb.insert(bnb.end(), cand.begin(), cand.end());
    brp[R + 1] = (int32_t)bnb.size();
  }
  // SELL-64 of A_b (same slices as J)
  const int n_slices = (Nn + kSlice - 1) / kSlice;
  std::vector<int32_t> sb(n_slices + 1, 0), rl(Nn, 0), diag(Nn, -1);
  int64_t total = 0;
  for (int sl = 0; sl < n_slices; ++sl) {
    int w = 0;
    for (int R = sl * kSlice; R < std::min(Nn, (sl + 1) * kSlice); ++R) {
      rl[R] = brp[R + 1] - brp[R];
      w = std::max(w, rl[R]);
    }
    sb[sl] = (int32_t)total;
    total += (int64_t)w * kSlice;
    if (total > INT32_MAX) {
      set_error("bound pattern too large for 32-bit slot indices");
      return TF_ERR_ARG;
    }
  }
  sb[n_slices] = (int32_t)total;
  std::vector<int32_t> col(total, 0), tsl(total, 0), bmap(total, 0);
  auto slotB = [&](int R, int C) -> int64_t {
    const int32_t* b = bnb.data() + brp[R];
    const int32_t* e = bnb.data() + brp[R + 1];
    const int32_t* q = std::lower_bound(b, e, C);
    return sb[R / kSlice] + (int64_t)(q - b) * kSlice + (R % kSlice);
  };
  std::vector<int32_t> ex_ptr(1, 0), ex_src;
  std::vector<uint8_t> ex_one;
  int n_ex = 0;
  const int64_t nsJ = h->pat.n_slots;
  for (int sl = 0; sl < n_slices; ++sl) {
    const int w = (sb[sl + 1] - sb[sl]) / kSlice;
    for (int lane = 0; lane < kSlice; ++lane) {
      const int R = sl * kSlice + lane;
      for (int k = 0; k < w; ++k) {
        const int64_t slot = sb[sl] + (int64_t)k * kSlice + lane;
        if (R >= Nn || k >= rl[R]) {  // padding
          col[slot] = std::min(R, Nn - 1);
          tsl[slot] = (int32_t)slot;
          bmap[slot] = 0;
          continue;
        }
        const int C = bnb[brp[R] + k];
        col[slot] = C;
        if (C == R) diag[R] = (int32_t)slot;
        {
          const int32_t* b = bnb.data() + brp[C];
          const int32_t* e = bnb.data() + brp[C + 1];
          const int32_t* q = std::lower_bound(b, e, R);
          if (q == e || *q != R) {
            set_error("internal: bound pattern is not symmetric");
            return TF_ERR_ARG;
          }
        }
        tsl[slot] = (int32_t)slotB(C, R);
        if (!node_bound[R] && !node_bound[C]) {
          bmap[slot] = (int32_t)slotJ(R, C);
          continue;
        }
        // exception block: per entry, the J values of M^-1(i) x M^-1(j)
        for (int pp = 0; pp < 3; ++pp)
          for (int qq = 0; qq < 3; ++qq) {
            const int i = 3 * R + pp, j = 3 * C + qq;
            uint8_t one = 0;
            if (M[i] != i || M[j] != j) {
              one = (i == j) ? 1 : 0;  // slave row / column: identity only
            } else {
              for (int d : Minv[i]) {
                const int Rp = d / 3, pr = d % 3;
                for (int kk = rp[Rp]; kk < rp[Rp + 1]; ++kk)
                  for (int q2 = 0; q2 < 3; ++q2) {
                    const int e = 3 * nb[kk] + q2;
                    if (M[e] != j) continue;
                    ex_src.push_back((int32_t)((3 * pr + q2) * nsJ + slotJ(Rp, nb[kk])));
                  }
              }
            }
            ex_one.push_back(one);
            ex_ptr.push_back((int32_t)ex_src.size());
          }
        bmap[slot] = -1 - n_ex;
        ++n_ex;
      }
    }
  }
  BoundOp& B = h->bound;
  B.pat.n_slices = n_slices;
  B.pat.n_slots = total;
  B.h_slice_base = sb;
  B.h_col = col;
  B.h_row_len = rl;
  TF_CHECK(balloc(h, &B.pat.slice_base, sb.size(), sb.data()));
  TF_CHECK(balloc(h, &B.pat.col, col.size(), col.data()));
  TF_CHECK(balloc(h, &B.pat.diag_slot, diag.size(), diag.data()));
  TF_CHECK(balloc(h, &B.pat.row_len, rl.size(), rl.data()));
  TF_CHECK(balloc(h, &B.pat.tslot, tsl.size(), tsl.data()));
  TF_CHECK(balloc(h, &B.blockmap, bmap.size(), bmap.data()));
  TF_CHECK(balloc(h, &B.ex_ptr, ex_ptr.size(), ex_ptr.data()));
  TF_CHECK(balloc(h, &B.ex_src, ex_src.size(), ex_src.empty() ? nullptr : ex_src.data()));
  TF_CHECK(balloc(h, &B.ex_one, ex_one.size(), ex_one.empty() ? nullptr : ex_one.data()));
  B.n_ex = n_ex;
  TF_CHECK(balloc<double>(h, &B.val, (size_t)total * 9, nullptr));
  TF_CHECK(balloc<float>(h, &B.val32, (size_t)total * 9, nullptr));
  TF_CHECK(balloc<double>(h, &B.minv, (size_t)Nn * 9, nullptr));
  TF_CHECK(balloc<float>(h, &B.minv32, (size_t)Nn * 9, nullptr));
  // the pairs grouped by master, masters in order of first appearance, each
  // master's slaves in pair order
  std::vector<int32_t> first(n, -1), masters, cnt;
  for (int p = 0; p < np; ++p) {
    const int md = h->h_pairs[2 * p + 1];
    if (first[md] < 0) {
      first[md] = (int32_t)masters.size();
      masters.push_back(md);
      cnt.push_back(0);
    }
    ++cnt[first[md]];
  }
  std::vector<int32_t> bptr(masters.size() + 1, 0), bsl(np);
  for (size_t i = 0; i < masters.size(); ++i) bptr[i + 1] = bptr[i] + cnt[i];
  std::vector<int32_t> fill(bptr.begin(), bptr.end() - 1);
  for (int p = 0; p < np; ++p) bsl[fill[first[h->h_pairs[2 * p + 1]]]++] = h->h_pairs[2 * p];
  TF_CHECK(balloc(h, &B.bm_ptr, bptr.size(), bptr.data()));
  TF_CHECK(balloc(h, &B.bm_master, masters.size(), masters.data()));
  TF_CHECK(balloc(h, &B.bm_slave, bsl.size(), bsl.data()));
  B.n_masters = (int32_t)masters.size();
  B.ready = true;
  return TF_OK;
}

// Common tail of tf_nr_create and tf_batch_create: stream, pattern, slot map,
// device buffers, tiles.  conn holds internal node ids.  Frees h on failure.
// Interleaves the bits of two 16-bit coordinates (Z order).
uint32_t morton2(uint32_t x, uint32_t y) {
  auto spread = [](uint32_t v) {
    v &= 0xffff;
    v = (v | (v << 8)) & 0x00ff00ff;
    v = (v | (v << 4)) & 0x0f0f0f0f;
    v = (v | (v << 2)) & 0x33333333;
    v = (v | (v << 1)) & 0x55555555;
    return v;
  };
  return spread(x) | (spread(y) << 1);
}

// DevPattern::slice_order from the nodes' reference positions (h_X): within
// each tissue, slices by the Morton code of their (x, y) centroid on a
// 2^16 x 2^16 grid over the tissue's extent, ties by slice index.
tf_status slice_order(tf_handle* h) {
  const int ns = h->pat.n_slices;
  std::vector<uint64_t> key(ns);
  for (int T = 0; T < h->n_tissues; ++T) {
    const int r0 = h->int_off[T], r1 = h->int_off[T] + (h->user_off[T + 1] - h->user_off[T]);
    double lo[2] = {1e300, 1e300}, hi[2] = {-1e300, -1e300};
    for (int i = r0; i < r1; ++i)
      for (int k = 0; k < 2; ++k) {
        lo[k] = std::min(lo[k], h->h_X[3 * (size_t)i + k]);
        hi[k] = std::max(hi[k], h->h_X[3 * (size_t)i + k]);
      }
    for (int sl = h->int_off[T] / kSlice; sl < (h->int_off[T + 1] + kSlice - 1) / kSlice; ++sl) {
      double c[2] = {0.0, 0.0};
      int n = 0;
      for (int i = sl * kSlice; i < std::min((sl + 1) * kSlice, r1); ++i, ++n)
        for (int k = 0; k < 2; ++k) c[k] += h->h_X[3 * (size_t)i + k];
      uint32_t q[2] = {0, 0};
      for (int k = 0; k < 2 && n > 0; ++k) {
        const double ext = hi[k] - lo[k];
        const double f = ext > 0.0 ? (c[k] / n - lo[k]) / ext : 0.0;
        q[k] = (uint32_t)std::min(65535.0, std::max(0.0, f * 65535.0));
      }
      key[sl] = ((uint64_t)T << 32) | morton2(q[0], q[1]);
    }
  }
  std::vector<int32_t> order(ns);
  for (int i = 0; i < ns; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return key[a] < key[b]; });
  TF_CHECK(dalloc(h, &h->pat.slice_order, ns));
  TF_CHECK(hipMemcpy(h->pat.slice_order, order.data(), sizeof(int32_t) * ns,
                     hipMemcpyHostToDevice));
  return TF_OK;
}

tf_status finish_create(tf_handle* h, const int32_t* conn, const double* ref_pos, tf_handle** out) {
  tf_status s = TF_OK;
  auto fail = [&](tf_status st) {
    tf_nr_destroy(h);
    return st;
  };
  if (hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking) != hipSuccess) {
    set_error("hipStreamCreate failed");
    return fail(TF_ERR_HIP);
  }
  for (auto& e : h->ev)
    if (hipEventCreate(&e) != hipSuccess) return fail(TF_ERR_HIP);
  if ((s = scratch_resolve(h->device, &h->scratch)) != TF_OK) return fail(s);
  if ((s = build_pattern(h, conn)) != TF_OK) return fail(s);
  if ((s = alloc_workspace(h)) != TF_OK) return fail(s);
  const size_t Ne = h->Ne, Nn = h->Nn;
  if (dalloc(h, &h->conn, Ne * 6) != hipSuccess || dalloc(h, &h->refpos, Ne * 18) != hipSuccess ||
      dalloc(h, &h->invFg, Ne * 9) != hipSuccess || dalloc(h, &h->lam, Ne) != hipSuccess ||
      dalloc(h, &h->mu, Ne) != hipSuccess || dalloc(h, &h->eta, Ne) != hipSuccess ||
      dalloc(h, &h->extv, Nn * 3) != hipSuccess || dalloc(h, &h->area, Nn) != hipSuccess ||
      dalloc(h, &h->dof_flags, Nn * 3) != hipSuccess || dalloc(h, &h->E_d, Ne) != hipSuccess ||
      dalloc(h, &h->nu_d, Ne) != hipSuccess || dalloc(h, &h->stm, Ne) != hipSuccess) {
    set_error("hipMalloc failed");
    return fail(TF_ERR_HIP);
  }
  // padding nodes of a batch are fixed in every direction: identity rows
  std::vector<uint8_t> fu((size_t)h->n_user * 3, 0);
  std::vector<uint8_t> flags = to_internal<uint8_t>(h, fu.data(), 3, (uint8_t)1);
  std::vector<double> zero((size_t)Nn * 3, 0.0);
  if (hipMemcpy(h->conn, conn, Ne * 6 * sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->refpos, ref_pos, Ne * 18 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->dof_flags, flags.data(), Nn * 3, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->extv, zero.data(), Nn * 3 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->area, zero.data(), Nn * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
    set_error("upload failed");
    return fail(TF_ERR_HIP);
  }
  h->L.stream = h->st;
  h->h_X.assign(3 * (size_t)Nn, 0.0);
  {
    std::vector<uint8_t> seen(Nn, 0);
    for (size_t e = 0; e < Ne; ++e)
      for (int a = 0; a < 6; ++a) {
        const int nd = conn[6 * e + a];
        if (nd < 0 || nd >= (int)Nn || seen[nd]) continue;
        seen[nd] = 1;
        for (int k = 0; k < 3; ++k) h->h_X[3 * (size_t)nd + k] = ref_pos[18 * e + 3 * a + k];
      }
  }
  if ((s = slice_order(h)) != TF_OK) return fail(s);
  reset_history(h);
  if ((s = alloc_tissue_state(h)) != TF_OK) return fail(s);
  if ((s = build_tiles(h, 8)) != TF_OK) return fail(s);
  *out = h;
  return TF_OK;
}

KrylovArgs krylov_args(tf_handle* h) {
  KrylovArgs k;
  k.n_nodes = h->Nn;
  k.pat = h->pat;
  k.val = h->ws.val;
  k.val32 = h->ws.val32;
  k.minv = h->ws.minv;
  k.minv32 = h->ws.minv32;
  k.pairs = h->pairs;
  k.n_pairs = h->n_pairs;
  if (h->bound.ready) {
    // bound system: the explicit A_b, solved like any unbound operator
    k.pat = h->bound.pat;
    k.val = h->bound.val;
    k.val32 = h->bound.val32;
    k.minv = h->bound.minv;
    k.minv32 = h->bound.minv32;
    k.n_pairs = 0;
  }
  k.ws = &h->ws;
  k.L = h->L;
  k.prof = &h->prof;
  k.tm = h->tm;
  k.dof_flags = h->dof_flags;
  k.coarse = h->coarse.active ? &h->coarse : nullptr;
  return k;
}

// The coarse space of the FP32 CG preconditioner for the operator just
// assembled (mixed precision): one segment per tissue, each with its own
// aggregates and dense coarse inverse, so that a tissue of a batch gets the
// correction it gets when stepped alone.  Aggregates and block lists are
// built when the operator's pattern changed, then A_c and its inverse.  With
// lagging on (the default), A_c and its inverse are formed at the first NR
// iteration of a step only and kept for the step's later iterations: the
// preconditioner stays a fixed SPD operator within every solve, and a step
// repeated from the same state repeats its bits.
// The coarse space of the CG solve and its factorisation from the step's
// operator (FP32 for the mixed step, FP64 otherwise).
tf_status coarse_factor_op(tf_handle* h, const int32_t* run, int mixed) {
  CoarseOp& c = h->coarse;
  if (h->bound.ready) {
    if (mixed)
      TF_CHECK(coarse_factor(c, h->bound.pat.col, h->bound.val32, h->bound.pat.n_slots,
                             h->dof_flags, run, h->st));
    else
      TF_CHECK(coarse_factor(c, h->bound.pat.col, (const double*)h->bound.val,
                             h->bound.pat.n_slots, h->dof_flags, run, h->st));
  } else if (mixed) {
    TF_CHECK(coarse_factor(c, h->pat.col, h->ws.val32, h->pat.n_slots, h->dof_flags, run, h->st));
  } else {
    TF_CHECK(coarse_factor(c, h->pat.col, (const double*)h->ws.val, h->pat.n_slots, h->dof_flags,
                           run, h->st));
  }
  return TF_OK;
}

tf_status coarse_prepare(tf_handle* h, int nr_iteration, const int32_t* run, int mixed) {
  CoarseOp& c = h->coarse;
  const bool bound = h->bound.ready;
  if (h->coarse_lag && nr_iteration > 0 && c.built && c.for_bound == bound &&
      h->coarse_cell != 0.0) {
    // lagged: a segment whose factorisation succeeded at the step's first
    // iteration keeps it; a running segment whose factorisation failed (a
    // singular A_c, or Newton-Schulz steps that do not converge) tries again,
    // as its tissue does when stepped alone
    std::vector<int32_t> retry(h->n_tissues, 0);
    bool any = false;
    for (int T = 0; T < h->n_tissues; ++T) {
      retry[T] = (!run || run[T]) && !c.seg_active[T] ? 1 : 0;
      any |= retry[T] != 0;
    }
    if (!any) return TF_OK;
    return coarse_factor_op(h, retry.data(), mixed);
  }
  c.active = false;
  if (h->coarse_cell == 0.0) return TF_OK;
  if (!c.built || c.for_bound != bound) {
    std::vector<CoarseSeg> segs(h->n_tissues);
    for (int T = 0; T < h->n_tissues; ++T) {
      CoarseSeg& g = segs[T];
      g.begin = h->int_off[T];
      g.end = h->int_off[T] + (h->user_off[T + 1] - h->user_off[T]);
      g.cell = h->coarse_cell;
      if (g.cell < 0.0) {
        // automatic: n_c = 6 x aggregates close to sqrt(stored blocks of the
        // tissue), so that the dense coarse product (4 n_c^2 bytes) stays near
        // a tenth of an FP32 SpMV (40 bytes per block); the (x, y) bounding
        // box is cut into n_c / (6 bands) square cells.  Bound systems keep
        // whole columns and at most about 400 cells, whose coarse operator
        // stays well conditioned under circumference binding.
        double x0 = 1e300, x1 = -1e300, y0 = 1e300, y1 = -1e300;
        for (int i = g.begin; i < g.end; ++i) {
          x0 = std::min(x0, h->h_X[3 * i]); x1 = std::max(x1, h->h_X[3 * i]);
          y0 = std::min(y0, h->h_X[3 * i + 1]); y1 = std::max(y1, h->h_X[3 * i + 1]);
        }
        const double nc = std::sqrt((double)(h->h_rowptr[g.end] - h->h_rowptr[g.begin]));
        const double cells = bound ? std::min(400.0, nc / 6.0) : nc / (6.0 * h->coarse_bands);
        g.cell = std::sqrt(std::max((x1 - x0) * (y1 - y0), 1e-30) / std::max(cells, 1.0));
      }
    }
    TF_CHECK(hipStreamSynchronize(h->st));
    tiled_release(&h->ws);  // captured chunks hold the previous coarse buffers
    const BoundOp& B = h->bound;
    // bound systems keep whole columns: binding ties a column's nodes to one
    // master (circumference binding slaves x and y of every node above the
    // basal one), which would leave the upper bands of such columns without
    // the modes they are built from
    hipError_t e = bound ? coarse_build(c, h->Nn, h->h_X.data(), segs, 1, B.h_slice_base,
                                        B.h_col, B.h_row_len)
                         : coarse_build(c, h->Nn, h->h_X.data(), segs, h->coarse_bands,
                                        h->h_slice_base, h->h_col, h->h_row_len);
    if (e == hipErrorInvalidValue) {
      set_error("coarse space: cell size gives too many aggregates");
      return TF_ERR_ARG;
    }
    TF_CHECK(e);
    c.for_bound = bound;
    ++c.gen;
  }
  return coarse_factor_op(h, run, mixed);
}

// Lays the pattern out again with `cap` contact columns in every slice that
// holds a node that may pack (node-node packing).
tf_status packing_relayout(tf_handle* h, int cap) {
  PackDev& p = h->pack;
  const int n_slices = (h->Nn + kSlice - 1) / kSlice;
  std::vector<int32_t> place(h->Nn);
  TF_CHECK(hipMemcpy(place.data(), p.place, sizeof(int32_t) * h->Nn, hipMemcpyDeviceToHost));
  std::vector<int32_t> caps(n_slices, 0);
  for (int k = 0; k < h->Nn; ++k)
    if (place[k] >= 0) caps[k / kSlice] = cap;
  std::vector<int32_t> conn((size_t)h->Ne * 6);
  TF_CHECK(hipStreamSynchronize(h->st));
  TF_CHECK(hipMemcpy(conn.data(), h->conn, sizeof(int32_t) * conn.size(), hipMemcpyDeviceToHost));
  tiled_release(&h->ws);  // captured chunks hold the previous pattern
  h->coarse.built = false;
  h->coarse.active = false;
  h->h_slice_cap = caps;
  tf_status st = layout_pattern(h, conn.data());
  if (st != TF_OK) return st;
  dfree(h, &p.inc);
  TF_CHECK(dalloc(h, &p.inc, (size_t)h->Nn * std::max(cap, 1)));
  p.cap = cap;
  p.detected = false;
  h->assembled = false;
  return TF_OK;
}

// Thresholds and couples at device positions u (Simulation::detectPacingNodes,
// Simulation.cpp:7997-8078), widening the contact columns when a node has
// more couples than they hold, then the contact slots of the pattern.
tf_status packing_detect(tf_handle* h, const double* u) {
  PackDev& p = h->pack;
  if (!h->have_flags) {
    set_error("node-node packing needs tf_nr_set_surface_flags (element placement and type)");
    return TF_ERR_ARG;
  }
  if (!p.have_mass) {
    set_error("node-node packing needs node masses (tf_nr_set_node_mass)");
    return TF_ERR_ARG;
  }
  int32_t flag[2] = {0, 0};
  for (int pass = 0; pass < 2; ++pass) {
    TF_CHECK(launch_packing_detect(p, h->Nn, h->Ne, h->conn, u, h->surf_place, h->surf_type,
                                   h->surf_spans, h->pat, h->st));
    TF_CHECK(hipMemcpyAsync(flag, p.flag, sizeof(flag), hipMemcpyDeviceToHost, h->st));
    TF_CHECK(hipStreamSynchronize(h->st));
    if (flag[0] > kPackLocal) {
      set_error("node-node packing: a node has " + std::to_string(flag[0]) +
                " couples, more than " + std::to_string(kPackLocal));
      return TF_ERR_ARG;
    }
    if (flag[0] <= p.cap) break;
    tf_status st = packing_relayout(h, std::min(kPackLocal, std::max(flag[0], 2 * p.cap)));
    if (st != TF_OK) return st;
  }
  TF_CHECK(launch_contact_slots(p, h->Nn, h->pat, h->st));
  p.n_pairs = flag[1];
  p.max_count = flag[0];
  p.detected = true;
  h->assembled = false;
  return TF_OK;
}

// Element kernel + slot-map assembly + residual and binding of the residual.
// in_step: an NR pass, whose strain output is written only for the tissues
// flagged in ws.tactive (in a batch, a tissue's strain then stays that of its
// own last assembly while the others iterate on)
tf_status assemble(tf_handle* h, double dt, int mixed, bool in_step = false) {
  if (!h->have_material || !h->have_growth || !h->have_area) {
    set_error("tf_nr_set_material, tf_nr_set_growth and a viscosity surface must precede assembly");
    return TF_ERR_ARG;
  }
  if (!(dt > 0.0)) {
    set_error("dt must be positive");
    return TF_ERR_ARG;
  }
  Workspace& w = h->ws;
  if (mixed) {
    tf_status s = ensure_mixed(h);
    if (s != TF_OK) return s;
  }
  ElementArgs ea{h->Ne, h->conn, h->refpos, h->invFg, h->lam, h->mu, h->eta, w.uk, w.un, dt,
                 h->any_visc ? 1 : 0, w.elemKe, w.elemKv, w.elemKe32, w.elemKv32, w.elemFe,
                 w.elemFv};
  ea.strain = w.strain;
  if (in_step) {
    ea.ptissue = h->ptissue;
    ea.tactive = w.tactive;
  }
  TF_CHECK(hipEventRecord(h->ev[0], h->st));
  // mixed: FP32 elastic + viscous blocks summed per K block (one gather each)
  TF_CHECK(launch_element(ea, mixed ? 2 : 0, h->st));
  TF_CHECK(launch_ablate_zero(ea, mixed ? 2 : 0, h->abl_list, h->n_abl, h->st));
  TF_CHECK(hipEventRecord(h->ev[1], h->st));
  AssembleArgs aa;
  aa.n_nodes = h->Nn;
  aa.pat = h->pat;
  aa.Ke = w.elemKe; aa.Kv = w.elemKv; aa.Ke32 = w.elemKe32; aa.Kv32 = w.elemKv32;
  aa.Fe = w.elemFe; aa.Fv = w.elemFv;
  aa.any_visc = h->any_visc ? 1 : 0;
  aa.combined = mixed ? 1 : 0;
  aa.ext_visc = h->extv; aa.area = h->area; aa.dof_flags = h->dof_flags;
  aa.abl_node = h->any_abl_node ? h->abl_node : nullptr;
  aa.uk = w.uk; aa.un = w.un; aa.dt = dt;
  aa.val = mixed ? nullptr : w.val;
  aa.val32 = w.val32;
  aa.minv = w.minv; aa.minv32 = w.minv32;
  aa.g = w.g; aa.internal = w.internal;
  if (h->pack.on) {
    PackDev& pk = h->pack;
    if (!pk.detected) {
      set_error("node-node packing is on: tf_nr_detect_packing must precede the step");
      return TF_ERR_ARG;
    }
    aa.pk_cnt = pk.cnt; aa.pk_inc = pk.inc; aa.pk_cap = pk.cap;
    aa.pk_mass = pk.mass; aa.pk_par = pk.par; aa.pk_force = pk.force;
  }
  TF_CHECK(launch_assemble(aa, mixed, h->st));
  if (h->n_pairs > 0) TF_CHECK(launch_bind_forces(h->bound, w.g, h->st));
  if (h->bound.ready)
    TF_CHECK(launch_bound_gather(h->bound, h->Nn, h->pat.n_slots, w.val, w.val32, mixed, h->st));
  TF_CHECK(hipEventRecord(h->ev[2], h->st));
  h->assembled = true;
  h->have_strain = true;
  h->last_mixed = mixed;
  h->last_dt = dt;
  h->assembly_inputs_kept = false;
  return TF_OK;
}

}  // namespace

extern "C" {

const char* tf_last_error(void) { return tf::g_err.c_str(); }

tf_status tf_nr_create(int32_t n_nodes, int32_t n_prisms, const int32_t* conn,
                       const double* ref_pos, tf_handle** out) {
  if (!out || n_nodes <= 0 || n_prisms <= 0 || !conn || !ref_pos) {
    set_error("tf_nr_create: invalid arguments");
    return TF_ERR_ARG;
  }
  *out = nullptr;
  int dev = 0;
  tf_status s = check_device(&dev);
  if (s != TF_OK) return s;
  tf_handle* h = new tf_handle();
  h->Nn = n_nodes;
  h->Ne = n_prisms;
  h->device = dev;
  h->n_user = n_nodes;
  h->user_off = {0, n_nodes};
  h->int_off = {0, n_nodes};
  h->elem_off = {0, n_prisms};
  return finish_create(h, conn, ref_pos, out);
}

tf_status tf_nr_set_material(tf_handle* h, const double* E, const double* nu, const double* eta,
                             const double* ext_visc, const uint8_t* fixed,
                             const int32_t* slave_master, int32_t n_pairs) {
  if (!h || !E || !nu || !eta || !ext_visc || n_pairs < 0 || (n_pairs > 0 && !slave_master)) {
    set_error("tf_nr_set_material: invalid arguments");
    return TF_ERR_ARG;
  }
  const int Ne = h->Ne, Nn = h->Nn;
  std::vector<double> lam(Ne), mu(Ne), stm(Ne, 1.0);
  if (h->stm_touched)
    TF_CHECK(hipMemcpy(stm.data(), h->stm, sizeof(double) * Ne, hipMemcpyDeviceToHost));
  bool any_visc = false;
  for (int e = 0; e < Ne; ++e) {
    // Prism::updateElasticProperties (Prism.cpp:526-527), E the prism's E
    // times its stiffness multiplier; the time-series multiplier is 1 until
    // a Young's modulus series changes it
    const double v = nu[e];
    lam[e] = stm[e] * E[e] * v / (1 + v) / (1 - 2.0 * v);
    mu[e] = stm[e] * E[e] / 2.0 / (1 + v);
    if (eta[e] != 0) any_visc = true;
  }
  if (!h->stm_touched)
    TF_CHECK(hipMemcpy(h->stm, stm.data(), sizeof(double) * Ne, hipMemcpyHostToDevice));
  TF_CHECK(hipMemcpy(h->E_d, E, sizeof(double) * Ne, hipMemcpyHostToDevice));
  TF_CHECK(hipMemcpy(h->nu_d, nu, sizeof(double) * Ne, hipMemcpyHostToDevice));
  std::vector<uint8_t> fu((size_t)h->n_user * 3, 0);
  if (fixed)
    for (size_t i = 0; i < fu.size(); ++i) fu[i] = fixed[i] ? 1 : 0;
  std::vector<uint8_t> flags = to_internal<uint8_t>(h, fu.data(), 3, (uint8_t)1);
  std::vector<double> ev = to_internal<double>(h, ext_visc, 3, 0.0);
  // pairs in the caller's DOF numbering (tissues concatenated) -> the
  // internal padded numbering; both DOF of a pair must be in one tissue
  h->h_pairs.resize(2 * (size_t)n_pairs);
  const int nU = h->n_user;
  for (int p = 0; p < n_pairs; ++p) {
    const int s = slave_master[2 * p], m = slave_master[2 * p + 1];
    if (s < 0 || m < 0 || s >= 3 * nU || m >= 3 * nU || s == m) {
      set_error("slave/master pair out of range");
      return TF_ERR_ARG;
    }
    const int ts = (int)(std::upper_bound(h->user_off.begin(), h->user_off.end(), s / 3) -
                         h->user_off.begin()) - 1;
    const int tm = (int)(std::upper_bound(h->user_off.begin(), h->user_off.end(), m / 3) -
                         h->user_off.begin()) - 1;
    if (ts != tm) {
      set_error("slave/master pair " + std::to_string(p) + " binds two tissues");
      return TF_ERR_ARG;
    }
    const int shift = 3 * (h->int_off[ts] - h->user_off[ts]);
    h->h_pairs[2 * p] = s + shift;
    h->h_pairs[2 * p + 1] = m + shift;
    flags[s + shift] |= kDofSlave;
  }
  h->h_flags_base = flags;
  for (size_t i = 0; i < h->h_abl_node.size(); ++i)
    if (h->h_abl_node[i])
      for (int d = 0; d < 3; ++d) flags[3 * i + d] |= kDofAblated;
  TF_CHECK(hipMemcpy(h->lam, lam.data(), sizeof(double) * Ne, hipMemcpyHostToDevice));
  TF_CHECK(hipMemcpy(h->mu, mu.data(), sizeof(double) * Ne, hipMemcpyHostToDevice));
  TF_CHECK(hipMemcpy(h->eta, eta, sizeof(double) * Ne, hipMemcpyHostToDevice));
  TF_CHECK(hipMemcpy(h->extv, ev.data(), sizeof(double) * Nn * 3, hipMemcpyHostToDevice));
  TF_CHECK(hipMemcpy(h->dof_flags, flags.data(), flags.size(), hipMemcpyHostToDevice));
  if (n_pairs > h->n_pairs || !h->pairs) {
    int32_t* p = nullptr;
    TF_CHECK(dalloc(h, &p, 2 * (size_t)std::max(n_pairs, 1)));
    h->pairs = p;
  }
  if (n_pairs > 0)
    TF_CHECK(hipMemcpy(h->pairs, h->h_pairs.data(), sizeof(int32_t) * 2 * n_pairs,
                       hipMemcpyHostToDevice));
  if (n_pairs > 0 && h->pack.on) {
    set_error("slave/master binding together with node-node packing is not supported");
    return TF_ERR_ARG;
  }
  const bool pairs_changed = n_pairs != h->n_pairs || !h->bound.ready;
  h->n_pairs = n_pairs;
  h->any_visc = any_visc;
  if (n_pairs == 0) {
    free_bound(h);
  } else if (pairs_changed || h->bound_pairs != h->h_pairs) {
    tf_status bs = build_bound(h);
    if (bs != TF_OK) return bs;
    h->bound_pairs = h->h_pairs;
  }
  if (any_visc) {
    tf_status s = ensure_visc(h);
    if (s != TF_OK) return s;
  }
  h->have_material = true;
  return TF_OK;
}

tf_status tf_nr_set_growth(tf_handle* h, const double* inv_Fg) {
  if (!h || !inv_Fg) {
    set_error("tf_nr_set_growth: invalid arguments");
    return TF_ERR_ARG;
  }
  TF_CHECK(hipMemcpy(h->invFg, inv_Fg, sizeof(double) * 9 * h->Ne, hipMemcpyHostToDevice));
  h->have_growth = true;
  return TF_OK;
}

tf_status tf_nr_set_visc_surface(tf_handle* h, const double* area) {
  if (!h || !area) {
    set_error("tf_nr_set_visc_surface: invalid arguments");
    return TF_ERR_ARG;
  }
  std::vector<double> a = to_internal<double>(h, area, 1, 0.0);
  TF_CHECK(hipMemcpy(h->area, a.data(), sizeof(double) * h->Nn, hipMemcpyHostToDevice));
  h->have_area = true;
  return TF_OK;
}

tf_status tf_nr_set_ablation(tf_handle* h, const uint8_t* ablated_prisms,
                            const int32_t* ablated_nodes, int32_t n_ablated_nodes) {
  if (!h || n_ablated_nodes < 0 || (n_ablated_nodes > 0 && !ablated_nodes)) {
    set_error("tf_nr_set_ablation: invalid arguments");
    return TF_ERR_ARG;
  }
  if (h->pack.on && (ablated_prisms || n_ablated_nodes > 0)) {
    set_error("laser ablation together with node-node packing is not supported");
    return TF_ERR_ARG;
  }
  std::vector<int32_t> list;
  if (ablated_prisms)
    for (int e = 0; e < h->Ne; ++e)
      if (ablated_prisms[e]) list.push_back(e);
  std::vector<uint8_t> nf((size_t)h->n_user, 0);
  for (int i = 0; i < n_ablated_nodes; ++i) {
    const int nd = ablated_nodes[i];
    if (nd < 0 || nd >= h->n_user) {
      set_error("tf_nr_set_ablation: ablated node " + std::to_string(nd) + " out of range");
      return TF_ERR_ARG;
    }
    nf[nd] = 1;
  }
  TF_CHECK(hipStreamSynchronize(h->st));
  if (!h->abl_el) {
    TF_CHECK(dalloc(h, &h->abl_el, (size_t)h->Ne));
    TF_CHECK(dalloc(h, &h->abl_list, (size_t)h->Ne));
    TF_CHECK(dalloc(h, &h->abl_node, (size_t)h->Nn));
  }
  std::vector<uint8_t> ef((size_t)h->Ne, 0);
  for (int32_t e : list) ef[e] = 1;
  std::vector<uint8_t> ni = to_internal<uint8_t>(h, nf.data(), 1, (uint8_t)0);
  TF_CHECK(hipMemcpy(h->abl_el, ef.data(), ef.size(), hipMemcpyHostToDevice));
  if (!list.empty())
    TF_CHECK(hipMemcpy(h->abl_list, list.data(), sizeof(int32_t) * list.size(),
                       hipMemcpyHostToDevice));
  TF_CHECK(hipMemcpy(h->abl_node, ni.data(), ni.size(), hipMemcpyHostToDevice));
  h->n_abl = (int32_t)list.size();
  h->any_abl_node = n_ablated_nodes > 0;
  h->h_abl_node = ni;
  if (h->h_flags_base.empty()) {
    // before tf_nr_set_material: padding nodes of a batch are fixed
    std::vector<uint8_t> fu((size_t)h->n_user * 3, 0);
    h->h_flags_base = to_internal<uint8_t>(h, fu.data(), 3, kDofFixed);
  }
  std::vector<uint8_t> flags = h->h_flags_base;
  for (size_t i = 0; i < ni.size(); ++i)
    if (ni[i])
      for (int d = 0; d < 3; ++d) flags[3 * i + d] |= kDofAblated;
  TF_CHECK(hipMemcpy(h->dof_flags, flags.data(), flags.size(), hipMemcpyHostToDevice));
  h->assembled = false;
  return TF_OK;
}

tf_status tf_nr_update_visc_surface(tf_handle* h, const double* u, const int32_t* elem_placement,
                                    const int32_t* elem_type, const uint8_t* spans,
                                    double* area_out) {
  if (!h || !u || !elem_placement || !elem_type || !spans) {
    set_error("tf_nr_update_visc_surface: invalid arguments");
    return TF_ERR_ARG;
  }
  const size_t Ne = h->Ne;
  int32_t *pl = nullptr, *ty = nullptr;
  uint8_t* sp = nullptr;
  TF_CHECK(hipMalloc(&pl, Ne * sizeof(int32_t)));
  TF_CHECK(hipMalloc(&ty, Ne * sizeof(int32_t)));
  TF_CHECK(hipMalloc(&sp, Ne));
  TF_CHECK(hipMemcpy(pl, elem_placement, Ne * sizeof(int32_t), hipMemcpyHostToDevice));
  TF_CHECK(hipMemcpy(ty, elem_type, Ne * sizeof(int32_t), hipMemcpyHostToDevice));
  TF_CHECK(hipMemcpy(sp, spans, Ne, hipMemcpyHostToDevice));
  {
    tf_status up = upload_nodes(h, h->ws.tmp, u);
    if (up != TF_OK) return up;
    TF_CHECK(hipStreamSynchronize(h->st));
  }
  TF_CHECK(launch_visc_surface(h->Nn, h->Ne, h->conn, h->ws.tmp, h->pat.node_ptr,
                               h->pat.node_src, pl, ty, sp, h->area, h->st,
                               h->n_abl ? h->abl_el : nullptr));
  TF_CHECK(hipStreamSynchronize(h->st));
  if (area_out) {
    std::vector<double> a(h->Nn);
    TF_CHECK(hipMemcpy(a.data(), h->area, sizeof(double) * h->Nn, hipMemcpyDeviceToHost));
    from_internal<double>(h, a.data(), area_out, 1);
  }
  TF_CHECK(hipFree(pl));
  TF_CHECK(hipFree(ty));
  TF_CHECK(hipFree(sp));
  h->have_area = true;
  return TF_OK;
}

tf_status tf_nr_update_visc_surface_noalloc(tf_handle* h, const double* u,
                                            const int32_t* elem_placement,
                                            const int32_t* elem_type, const uint8_t* spans,
                                            double* area_out) {
  if (!h || !u || !elem_placement || !elem_type || !spans) {
    set_error("tf_nr_update_visc_surface_noalloc: invalid arguments");
    return TF_ERR_ARG;
  }
  const size_t Ne = h->Ne;
  const size_t o_ty = align256(Ne * sizeof(int32_t)), o_sp = 2 * o_ty;
  const size_t need = o_sp + align256(Ne);
  if (need > kScratchBytes) {
    set_error("tf_nr_update_visc_surface_noalloc: " + std::to_string(Ne) + " prisms need " +
              std::to_string(need) + " bytes of scratch, more than its " +
              std::to_string(kScratchBytes) + " (about 3.7 million prisms)");
    return TF_ERR_ARG;
  }
  if (!h->scratch) {
    set_error("tf_nr_update_visc_surface_noalloc: no scratch address for the device");
    return TF_ERR_HIP;
  }
  unsigned char* base = static_cast<unsigned char*>(h->scratch);
  int32_t* pl = reinterpret_cast<int32_t*>(base);
  int32_t* ty = reinterpret_cast<int32_t*>(base + o_ty);
  uint8_t* sp = base + o_sp;
  // every staged byte is written before the kernel reads it: nothing of an
  // earlier call's contents is used
  TF_CHECK(hipMemcpyAsync(pl, elem_placement, Ne * sizeof(int32_t), hipMemcpyHostToDevice, h->st));
  TF_CHECK(hipMemcpyAsync(ty, elem_type, Ne * sizeof(int32_t), hipMemcpyHostToDevice, h->st));
  TF_CHECK(hipMemcpyAsync(sp, spans, Ne, hipMemcpyHostToDevice, h->st));
  {
    tf_status up = upload_nodes(h, h->ws.tmp, u);
    if (up != TF_OK) return up;
  }
  TF_CHECK(launch_visc_surface(h->Nn, h->Ne, h->conn, h->ws.tmp, h->pat.node_ptr,
                               h->pat.node_src, pl, ty, sp, h->area, h->st,
                               h->n_abl ? h->abl_el : nullptr));
  TF_CHECK(hipStreamSynchronize(h->st));
  h->have_area = true;
  if (area_out) {
    if (!h->batch) {
      TF_CHECK(hipMemcpy(area_out, h->area, sizeof(double) * h->Nn, hipMemcpyDeviceToHost));
    } else {
      std::vector<double> a(h->Nn);
      TF_CHECK(hipMemcpy(a.data(), h->area, sizeof(double) * h->Nn, hipMemcpyDeviceToHost));
      from_internal<double>(h, a.data(), area_out, 1);
    }
  }
  return TF_OK;
}

tf_status tf_nr_assemble(tf_handle* h, const double* uk, const double* un, double dt,
                         double* g_out, double* internal_out) {
  return tf_nr_assemble_ex(h, uk, un, dt, TF_PREC_FP64, g_out, internal_out);
}

tf_status tf_nr_assemble_ex(tf_handle* h, const double* uk, const double* un, double dt,
                            int32_t precision_mode, double* g_out, double* internal_out) {
  if (h && h->batch) {
    set_error("tf_nr_assemble works on single-tissue handles only");
    return TF_ERR_ARG;
  }
  if (!h || !uk || !un || (precision_mode != TF_PREC_FP64 && precision_mode != TF_PREC_MIXED)) {
    set_error("tf_nr_assemble: invalid arguments");
    return TF_ERR_ARG;
  }
  const size_t n = 3 * (size_t)h->Nn;
  TF_CHECK(hipMemcpyAsync(h->ws.uk, uk, sizeof(double) * n, hipMemcpyHostToDevice, h->st));
  TF_CHECK(hipMemcpyAsync(h->ws.un, un, sizeof(double) * n, hipMemcpyHostToDevice, h->st));
  tf_status s = assemble(h, dt, precision_mode == TF_PREC_MIXED ? 1 : 0);
  if (s != TF_OK) return s;
  h->assembly_inputs_kept = true;
  if (g_out) TF_CHECK(hipMemcpyAsync(g_out, h->ws.g, sizeof(double) * n, hipMemcpyDeviceToHost, h->st));
  if (internal_out)
    TF_CHECK(hipMemcpyAsync(internal_out, h->ws.internal, sizeof(double) * n, hipMemcpyDeviceToHost, h->st));
  TF_CHECK(hipStreamSynchronize(h->st));
  return TF_OK;
}

tf_status tf_nr_solve(tf_handle* h, const double* b, double* x, double rtol, int32_t maxit,
                      int32_t* iterations) {
  if (h && h->batch) {
    set_error("tf_nr_solve works on single-tissue handles only");
    return TF_ERR_ARG;
  }
  if (!h || !b || !x || maxit <= 0) {
    set_error("tf_nr_solve: invalid arguments");
    return TF_ERR_ARG;
  }
  if (!h->assembled || h->last_mixed) {
    set_error("tf_nr_solve needs a preceding FP64 tf_nr_assemble");
    return TF_ERR_ARG;
  }
  const size_t n = 3 * (size_t)h->Nn;
  TF_CHECK(hipMemcpyAsync(h->ws.b, b, sizeof(double) * n, hipMemcpyHostToDevice, h->st));
  KrylovArgs k = krylov_args(h);
  int it = 0, state = 0;
  TF_CHECK(krylov_bicgstab(k, h->ws.b, h->ws.x, rtol > 0 ? rtol : 1e-10, maxit, &it, &state, 16));
  TF_CHECK(hipMemcpyAsync(x, h->ws.x, sizeof(double) * n, hipMemcpyDeviceToHost, h->st));
  TF_CHECK(hipStreamSynchronize(h->st));
  if (iterations) *iterations = it;
  if (state != K_CONVERGED) {
    set_error(state == K_BREAKDOWN ? "BiCGStab breakdown" : "BiCGStab iteration cap reached");
    return TF_ERR_KRYLOV;
  }
  return TF_OK;
}

}  // extern "C"

namespace {

// One Young's modulus series over every prism it applies to
// (YoungsModulusModifier::updateTimeSeriesStiffnessMultiplier,
// YoungsModulusModifier.cpp:158-177): the change rate at the prism's grid
// position (GetElementPosition with ShapeBase::convertRelativePosToGridIndex,
// ShapeBase.cpp:834-860; GetCurrentMultiplierChangeRate, :106-156), the
// multiplier's linear or exponential update, and the Lame parameters
// (Prism::updateElasticProperties, Prism.cpp:526-527).  Contraction is off so
// that every product and sum rounds where the reference's does.
__global__ __launch_bounds__(256) void stiffness_series_kernel(
    int32_t n_e, const uint8_t* __restrict__ mask, const double* __restrict__ relpos,
    const double* __restrict__ rate, int32_t nx, int32_t ny, int32_t method, int32_t linear,
    double dt, const double* __restrict__ E, const double* __restrict__ nu, double* stm,
    double* lam, double* mu) {
#pragma clang fp contract(off)
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= n_e || !mask[e]) return;
  double rx = relpos[2 * e] * (double)(float)(nx - 1);
  double ry = relpos[2 * e + 1] * (double)(float)(ny - 1);
  int ix = (int)floor(rx), iy = (int)floor(ry);
  double fx = rx - ix, fy = ry - iy;
  if (ix >= nx - 1) { ix = nx - 2; fx = 1.0; }
  else if (ix < 0) { ix = 0; fx = 0.0; }
  if (iy >= ny - 1) { iy = ny - 2; fy = 1.0; }
  else if (iy < 0) { iy = 0; fy = 0.0; }
  const double n0 = rate[(size_t)iy * nx + ix], n1 = rate[(size_t)iy * nx + ix + 1];
  const double n2 = rate[(size_t)(iy + 1) * nx + ix], n3 = rate[(size_t)(iy + 1) * nx + ix + 1];
  double r;
  if (method == 0) r = fx > 0.5 ? (fy > 0.5 ? n3 : n1) : (fy > 0.5 ? n2 : n0);
  else if (method == 1)
    r = n0 * (1.0 - fx) * (1.0 - fy) + n1 * fx * (1.0 - fy) + n2 * (1.0 - fx) * fy + n3 * fx * fy;
  else r = n0;
  const double m = linear ? stm[e] + r * dt : stm[e] * exp(r * dt);
  stm[e] = m;
  const double v = nu[e];
  lam[e] = m * E[e] * v / (1 + v) / (1 - 2.0 * v);
  mu[e] = m * E[e] / 2.0 / (1 + v);
}

// Simulation::ApplyAllYoungsModulusModifiers (Simulation.cpp:878-884) at time
// t: every series in the order added, each over its interval of t
// (YoungsModulusModifier::calculateCurrentRateGridIndex,
// YoungsModulusModifier.cpp:48-72; none past the last time point or before
// the first).
tf_status apply_series(tf_handle* h, double t, double dt) {
  for (StiffSeries& q : h->series) {
    if (t > q.times[q.n_t - 1]) continue;
    int k = -1;
    for (int i = 0; i < q.n_t - 1; ++i)
      if (t >= q.times[i] && t < q.times[i + 1]) { k = i; break; }
    if (k < 0) continue;
    // the prisms' current positions in the bounding box when the device keeps
    // them (tf_nr_update_relative_positions, every step of tf_nr_advance),
    // else the positions given with the series
    const double* rp = h->gr.have_cur ? h->gr.relcur : q.relpos;
    if (!rp) {
      set_error("stiffness series: no relative positions (give them with the series, or set the "
                "growth prisms and update the relative positions)");
      return TF_ERR_ARG;
    }
    hipLaunchKernelGGL(stiffness_series_kernel, dim3((h->Ne + 255) / 256), dim3(256), 0, h->st,
                       h->Ne, q.mask, rp, q.rate + (size_t)k * q.ny * q.nx, q.nx, q.ny,
                       q.method, q.linear, dt, h->E_d, h->nu_d, h->stm, h->lam, h->mu);
    TF_CHECK(hipGetLastError());
    h->stm_touched = true;
  }
  return TF_OK;
}

// Mixed precision: FP32 inner solves are never asked for less than this
// relative residual (FP64 refinement sweeps go further), and stop after
// kMixedCap iterations, after which the tissue falls back to FP64.
constexpr double kMixedInnerRtol = 1e-5;
constexpr int kMixedCap = 5000;
// GMRES iterations per solve (each synchronises with the host once)
constexpr int kGmresCap = 5000;
// Inexact Newton (a Krylov rtol of at least kForcingMin is a forcing term):
// a solve that the last two increments predict to be the step's last is not
// oversolved (the safeguard of Eisenstat and Walker's forcing terms).  The
// next increment is predicted as e = |du_{k-1}|^2 / |du_{k-2}| (the observed
// linear contraction), and the solve's relative tolerance is raised to
// min(kForcingCap, max(rtol, 0.5 tol / e)): its increment then still decides
// convergence to within the cap, and a position error of about kForcingCap
// times an increment already below tol is left.
constexpr double kForcingMin = 1e-6;
constexpr double kForcingCap = 0.1;

// The NR loop of Simulation::updateStepNR on device-resident state, run for
// every tissue of the handle with the tissue's own convergence test: w.un
// and w.uk hold u_n on entry; w.uk holds the converged iterates on exit.
// S has one entry per tissue.
tf_status nr_loop(tf_handle* h, double dt, double tol, int32_t maxit, int mixed, double rtol,
                  tf_step_stats* S) {
  Workspace& w = h->ws;
  const int nt = h->n_tissues;
  const int kmax = 50000;
  std::vector<int32_t> act(nt, 1), it(nt, 0), cap(nt, 0), fb(nt, 0);
  std::vector<double> norms(nt, 0.0), dprev1(nt, 0.0), dprev2(nt, 0.0);
  for (int T = 0; T < nt; ++T) std::memset(&S[T], 0, sizeof(tf_step_stats));
  static thread_local TiledResult res, fres;
  for (int pass = 0;; ++pass) {
    bool any = false;
    for (int T = 0; T < nt; ++T) any |= act[T] != 0;
    if (!any) break;
    TF_CHECK(hipMemcpyAsync(w.tactive, act.data(), sizeof(int32_t) * nt, hipMemcpyHostToDevice,
                            h->st));
    tf_status s = assemble(h, dt, mixed, true);
    if (s != TF_OK) return s;
    const int method = h->krylov_method[mixed];
    if (method == KRYLOV_CG) {
      s = coarse_prepare(h, pass, act.data(), mixed);
      if (s != TF_OK) return s;
    } else {
      h->coarse.active = false;
    }
    KrylovArgs k = krylov_args(h);
    // |g| per tissue, read after the solve's synchronisation: a NaN or Inf in
    // the residual (the reference only prints it, NewtonRaphsonSolver.cpp:453-455)
    // ends that tissue's step with TF_ERR_NAN
    TF_CHECK(tiled_residual_norms(k, w.g));
    const double inner = std::max(rtol, kMixedInnerRtol);
    // The step's first solve starts from the displacement that the tissue's
    // last steps extrapolate to (no counterpart in the reference, whose
    // direct solve needs no guess): the solve is handed the guess's residual
    // and still ends at rtol |g_T|, so the increment meets the same test.
    bool warm = false;
    if (pass == 0 && h->warm_start && method != KRYLOV_GMRES && h->ws.x0 &&
        h->n_hist.size() == (size_t)nt)
      for (int T = 0; T < nt; ++T) warm |= h->n_hist[T] > 0;
    const double* rhs = w.g;
    if (warm) {
      TF_CHECK(hipMemcpyAsync(w.hist_order, h->n_hist.data(), sizeof(int32_t) * nt,
                              hipMemcpyHostToDevice, h->st));
      TF_CHECK(tiled_extrapolate(k, w.hist_order, w.hist0, w.hist1, w.x0));
      TF_CHECK(tiled_warm_start(k, mixed, w.g, w.x0, w.r0, w.rscale, w.host_rscale));
      k.rtol_scale = w.rscale;
      k.rtol_scale_h = w.host_rscale;
      rhs = w.r0;
    }
    // a solve predicted to be the step's last is not oversolved
    if (pass >= 2 && h->forcing_safeguard && rtol >= kForcingMin && method != KRYLOV_GMRES) {
      bool loose = false;
      std::vector<double> sc(nt, 1.0);
      for (int T = 0; T < nt; ++T) {
        if (!act[T] || !(dprev1[T] > 0.0) || !(dprev2[T] > 0.0)) continue;
        const double e = dprev1[T] * dprev1[T] / dprev2[T];
        if (!(e > 0.0) || !std::isfinite(e)) continue;
        const double eta = std::min(kForcingCap, std::max(rtol, 0.5 * tol / e));
        sc[T] = eta / rtol;
        loose |= sc[T] > 1.0;
      }
      if (loose) {
        s = ensure_warm(h);
        if (s != TF_OK) return s;
        std::copy(sc.begin(), sc.end(), w.host_rscale);
        TF_CHECK(hipMemcpyAsync(w.rscale, w.host_rscale, sizeof(double) * nt,
                                hipMemcpyHostToDevice, h->st));
        k.rtol_scale = w.rscale;
        k.rtol_scale_h = w.host_rscale;
      }
    }
    if (mixed) {
      TF_CHECK(tiled_mixed(k, rhs, w.du, rtol, inner, kMixedCap, 16, method, &res));
    } else if (method == KRYLOV_GMRES) {
      s = ensure_gmres(h);
      if (s != TF_OK) return s;
      TF_CHECK(tiled_gmres(k, w.g, w.du, rtol, kGmresCap, &res));
    } else if (method == KRYLOV_CG) {
      TF_CHECK(tiled_cg(k, rhs, w.du, rtol, kmax, 16, &res));
    } else {
      TF_CHECK(tiled_bicgstab(k, rhs, w.du, rtol, kmax, 16, &res));
    }
    if (warm) {
      TF_CHECK(tiled_add_active(k, w.x0, w.du));
      k = krylov_args(h);
    }
    TF_CHECK(hipEventRecord(h->ev[3], h->st));
    TF_CHECK(hipEventSynchronize(h->ev[3]));
    float t_el = 0, t_as = 0, t_kr = 0;
    TF_CHECK(hipEventElapsedTime(&t_el, h->ev[0], h->ev[1]));
    TF_CHECK(hipEventElapsedTime(&t_as, h->ev[1], h->ev[2]));
    TF_CHECK(hipEventElapsedTime(&t_kr, h->ev[2], h->ev[3]));
    // tissues with a non-finite residual take no fallback solve
    std::vector<int32_t> bad(nt, 0);
    for (int T = 0; T < nt; ++T) bad[T] = act[T] && !std::isfinite(w.host_gnorm[T]);
    // Tissues whose solve failed are solved again, each on its own, by the
    // next method of the chain CG -> BiCGStab (same precision) -> BiCGStab
    // with the FP64 operator; the other tissues keep their increments.
    auto failed = [&](std::vector<int32_t>& f) {
      bool any_f = false;
      for (int T = 0; T < nt; ++T) {
        f[T] = (act[T] && !bad[T] && res.state[T] != K_CONVERGED) ? 1 : 0;
        any_f |= f[T] != 0;
      }
      return any_f;
    };
    // the FP64 operator is current (FP64 step, or the mixed step's FP64 re-assembly)
    bool have_fp64 = !mixed;
    // stage 0: BiCGStab in the step's precision; 1: BiCGStab, FP64 operator;
    // 2: GMRES, FP64 operator
    auto retry = [&](const std::vector<int32_t>& f, int stage) -> tf_status {
      TF_CHECK(hipEventRecord(h->ev[4], h->st));
      TF_CHECK(hipMemcpyAsync(w.tactive, f.data(), sizeof(int32_t) * nt, hipMemcpyHostToDevice,
                              h->st));
      if (stage > 0 && !have_fp64) {
        tf_status as = assemble(h, dt, 0, true);
        if (as != TF_OK) return as;
        have_fp64 = true;
      }
      KrylovArgs kr = krylov_args(h);
      if (stage == 0 && mixed) {
        TF_CHECK(tiled_mixed(kr, w.g, w.tmp, rtol, inner, kMixedCap, 16, KRYLOV_BICGSTAB, &fres));
      } else if (stage == 2) {
        tf_status gs = ensure_gmres(h);
        if (gs != TF_OK) return gs;
        TF_CHECK(tiled_gmres(kr, w.g, w.tmp, rtol, kGmresCap, &fres));
      } else {
        TF_CHECK(tiled_bicgstab(kr, w.g, w.tmp, rtol, kmax, 16, &fres));
      }
      TF_CHECK(tiled_copy_active(kr, w.tmp, w.du));
      TF_CHECK(hipMemcpyAsync(w.tactive, act.data(), sizeof(int32_t) * nt, hipMemcpyHostToDevice,
                              h->st));
      TF_CHECK(hipEventRecord(h->ev[5], h->st));
      TF_CHECK(hipEventSynchronize(h->ev[5]));
      float t_fb = 0;
      TF_CHECK(hipEventElapsedTime(&t_fb, h->ev[4], h->ev[5]));
      t_kr += t_fb;
      for (int T = 0; T < nt; ++T)
        if (f[T]) {
          res.state[T] = fres.state[T];
          res.iters[T] += fres.iters[T];
        }
      return TF_OK;
    };
    if (method == KRYLOV_CG && failed(fb)) {
      tf_status rs = retry(fb, 0);
      if (rs != TF_OK) return rs;
      for (int T = 0; T < nt; ++T) S[T].cg_fallbacks += fb[T];
    }
    if (mixed && failed(fb)) {
      tf_status rs = retry(fb, 1);
      if (rs != TF_OK) return rs;
      for (int T = 0; T < nt; ++T) S[T].fp64_fallbacks += fb[T];
      k = krylov_args(h);
    }
    if (method != KRYLOV_GMRES && failed(fb)) {
      tf_status rs = retry(fb, 2);
      if (rs != TF_OK) return rs;
      for (int T = 0; T < nt; ++T) S[T].gmres_fallbacks += fb[T];
      k = krylov_args(h);
    }
    TF_CHECK(launch_bind_increment(h->pairs, h->n_pairs, w.du, h->st));
    TF_CHECK(tiled_norm_update(k, w.du, w.uk, norms.data()));
    if (h->prof.on) {
      h->prof_el_ms += t_el;
      h->prof_as_ms += t_as;
      h->prof_assemblies++;
    }
    for (int T = 0; T < nt; ++T) {
      if (!act[T]) continue;
      tf_step_stats& St = S[T];
      St.ms_element += t_el;
      St.ms_assembly += t_as;
      St.ms_krylov += t_kr;
      const double dnorm = norms[T];
      const int kit = res.iters[T];
      if (it[T] < TF_MAX_NR_ITERS) {
        St.krylov_iterations[it[T]] = kit;
        St.du_norms[it[T]] = dnorm;
      }
      St.krylov_iterations_total += kit;
      tf_status result = TF_OK;
      if (bad[T]) {
        set_error("NaN/Inf in the NR residual at iteration " + std::to_string(it[T]) +
                  (nt > 1 ? " of tissue " + std::to_string(T) : std::string()));
        result = TF_ERR_NAN;
      } else if (res.state[T] == K_BREAKDOWN) {
        set_error("BiCGStab breakdown in NR iteration " + std::to_string(it[T]) +
                  (nt > 1 ? " of tissue " + std::to_string(T) : std::string()));
        result = TF_ERR_KRYLOV;
      } else if (res.state[T] != K_CONVERGED) {
        set_error("BiCGStab iteration cap in NR iteration " + std::to_string(it[T]) +
                  (nt > 1 ? " of tissue " + std::to_string(T) : std::string()));
        result = TF_ERR_KRYLOV;
      }
      if (!std::isfinite(dnorm)) {
        set_error("NaN/Inf in the NR increment at iteration " + std::to_string(it[T]));
        result = TF_ERR_NAN;
      }
      // NewtonRaphsonSolver::checkConvergenceViaDeltaU (NewtonRaphsonSolver.cpp:807-825);
      // the increment is applied whether or not it converged (Simulation.cpp:6716-6720)
      const bool conv = !(dnorm > tol);
      dprev2[T] = dprev1[T];
      dprev1[T] = dnorm;
      it[T]++;
      if (result != TF_OK) {
        St.status = result;
        act[T] = 0;
      } else if (conv) {
        St.converged = 1;
        act[T] = 0;
      } else if (it[T] > maxit) {
        cap[T] = 1;
        act[T] = 0;
      }
    }
  }
  tf_status worst = TF_OK;
  for (int T = 0; T < nt; ++T) {
    tf_step_stats& St = S[T];
    St.ms_total = St.ms_element + St.ms_assembly + St.ms_krylov;
    St.nr_iterations = it[T];
    if (St.status == TF_OK && cap[T]) {
      set_error("NR did not converge within maxit");
      St.status = TF_ERR_NOT_CONVERGED;
    }
    if (St.status != TF_OK && (worst == TF_OK || worst == TF_ERR_NOT_CONVERGED)) worst = (tf_status)St.status;
  }
  // every tissue participates again in the next solve
  std::vector<int32_t> ones(nt, 1);
  TF_CHECK(hipMemcpyAsync(w.tactive, ones.data(), sizeof(int32_t) * nt, hipMemcpyHostToDevice,
                          h->st));
  TF_CHECK(hipStreamSynchronize(h->st));
  return worst;
}

}  // namespace

extern "C" {

tf_status tf_nr_step(tf_handle* h, const double* u_n, double dt, double tol, int32_t maxit,
                     int32_t precision_mode, double krylov_rtol, double* u_out,
                     tf_step_stats* stats) {
  if (!h || !u_n || !u_out || maxit < 0 || maxit > TF_MAX_NR_ITERS - 2 ||
      (precision_mode != TF_PREC_FP64 && precision_mode != TF_PREC_MIXED)) {
    set_error("tf_nr_step: invalid arguments");
    return TF_ERR_ARG;
  }
  std::vector<tf_step_stats> local(h->n_tissues);
  tf_step_stats* S = stats ? stats : local.data();
  const int mixed = precision_mode == TF_PREC_MIXED ? 1 : 0;
  const size_t n = 3 * (size_t)h->Nn;
  Workspace& w = h->ws;
  tf_status up = upload_nodes(h, w.un, u_n);
  if (up != TF_OK) return up;
  TF_CHECK(hipMemcpyAsync(w.uk, w.un, sizeof(double) * n, hipMemcpyDeviceToDevice, h->st));
  const double rtol = krylov_rtol > 0 ? krylov_rtol : 1e-10;
  reset_history(h);
  tf_status result = nr_loop(h, dt, tol, maxit, mixed, rtol, S);
  if (result != TF_OK && result != TF_ERR_NOT_CONVERGED && result != TF_ERR_KRYLOV &&
      result != TF_ERR_NAN)
    return result;
  tf_status dn = download_nodes(h, u_out, w.uk);
  if (dn != TF_OK) return dn;
  return result;
}

tf_status tf_nr_set_growth_rates(tf_handle* h, const double* rates) {
  if (!h || !rates) {
    set_error("tf_nr_set_growth_rates: invalid arguments");
    return TF_ERR_ARG;
  }
  const size_t Ne = h->Ne;
  if (!h->Fg) {
    TF_CHECK(dalloc(h, &h->Fg, Ne * 9));
    TF_CHECK(dalloc(h, &h->rates, Ne * 3));
  }
  TF_CHECK(hipMemcpy(h->rates, rates, sizeof(double) * Ne * 3, hipMemcpyHostToDevice));
  std::vector<double> I(Ne * 9, 0.0);
  for (size_t e = 0; e < Ne; ++e) I[9 * e] = I[9 * e + 4] = I[9 * e + 8] = 1.0;
  TF_CHECK(hipMemcpy(h->Fg, I.data(), sizeof(double) * Ne * 9, hipMemcpyHostToDevice));
  TF_CHECK(hipMemcpy(h->invFg, I.data(), sizeof(double) * Ne * 9, hipMemcpyHostToDevice));
  h->have_rates = h->have_growth = true;
  return TF_OK;
}

tf_status tf_nr_growth_tensor(tf_handle* h, double* Fg_out, const double* Fg_in) {
  if (!h || !h->Fg || (!Fg_out && !Fg_in)) {
    set_error("tf_nr_growth_tensor: invalid arguments or no growth rates set");
    return TF_ERR_ARG;
  }
  const size_t Ne = h->Ne;
  if (Fg_out) TF_CHECK(hipMemcpy(Fg_out, h->Fg, sizeof(double) * Ne * 9, hipMemcpyDeviceToHost));
  if (Fg_in) {
    TF_CHECK(hipMemcpy(h->Fg, Fg_in, sizeof(double) * Ne * 9, hipMemcpyHostToDevice));
    TF_CHECK(launch_inverse(h->Ne, h->Fg, h->invFg, h->st));
    TF_CHECK(hipStreamSynchronize(h->st));
    h->have_growth = true;
  }
  return TF_OK;
}

tf_status tf_nr_set_surface_flags(tf_handle* h, const int32_t* elem_placement,
                                  const int32_t* elem_type, const uint8_t* spans) {
  if (!h || !elem_placement || !elem_type || !spans) {
    set_error("tf_nr_set_surface_flags: invalid arguments");
    return TF_ERR_ARG;
  }
  const size_t Ne = h->Ne;
  if (!h->surf_place) {
    TF_CHECK(dalloc(h, &h->surf_place, Ne));
    TF_CHECK(dalloc(h, &h->surf_type, Ne));
    TF_CHECK(dalloc(h, &h->surf_spans, Ne));
  }
  TF_CHECK(hipMemcpy(h->surf_place, elem_placement, sizeof(int32_t) * Ne, hipMemcpyHostToDevice));
  TF_CHECK(hipMemcpy(h->surf_type, elem_type, sizeof(int32_t) * Ne, hipMemcpyHostToDevice));
  TF_CHECK(hipMemcpy(h->surf_spans, spans, Ne, hipMemcpyHostToDevice));
  h->have_flags = true;
  return TF_OK;
}

tf_status tf_nr_add_stiffness_series(tf_handle* h, int32_t n_times, const double* times,
                                     int32_t nx, int32_t ny, const double* grids, int32_t linear,
                                     int32_t method, const uint8_t* applicable,
                                     const double* rel_pos) {
  if (!h || n_times < 2 || !times || nx < 2 || ny < 2 || !grids || method < 0 || method > 2 ||
      !applicable || (!rel_pos && !h->gr.prisms)) {
    set_error("tf_nr_add_stiffness_series: invalid arguments");
    return TF_ERR_ARG;
  }
  if (!h->have_material) {
    set_error("tf_nr_add_stiffness_series: set the material first");
    return TF_ERR_ARG;
  }
  StiffSeries q;
  q.n_t = n_times; q.nx = nx; q.ny = ny; q.method = method; q.linear = linear ? 1 : 0;
  q.times.assign(times, times + n_times);
  // YoungsModulusModifier::CalculateYoungsModulusChangeRate
  // (YoungsModulusModifier.cpp:15-46): per interval and grid point,
  // (G1 / G0 - 1) / dT (linear) or log(G1 / G0) / dT (exponential)
  const size_t g = (size_t)nx * ny;
  std::vector<double> rate((size_t)(n_times - 1) * g);
  for (int t = 0; t < n_times - 1; ++t) {
    const double dT = times[t + 1] - times[t];
    for (size_t k = 0; k < g; ++k) {
      const double ratio = grids[(size_t)(t + 1) * g + k] / grids[(size_t)t * g + k];
      rate[(size_t)t * g + k] = q.linear ? (ratio - 1) / dT : std::log(ratio) / dT;
    }
  }
  std::vector<uint8_t> m(h->Ne);
  for (int e = 0; e < h->Ne; ++e) {
    m[e] = applicable[e] ? 1 : 0;
    if (rel_pos && (!std::isfinite(rel_pos[2 * e]) || !std::isfinite(rel_pos[2 * e + 1]))) {
      set_error("tf_nr_add_stiffness_series: relative positions must be finite");
      return TF_ERR_ARG;
    }
  }
  for (size_t k = 0; k < (size_t)n_times * g; ++k)
    if (!std::isfinite(grids[k]) || !(grids[k] > 0.0)) {
      set_error("tf_nr_add_stiffness_series: grid values must be positive and finite");
      return TF_ERR_ARG;
    }
  TF_CHECK(hipStreamSynchronize(h->st));
  if (dalloc(h, &q.rate, rate.size()) != hipSuccess || dalloc(h, &q.mask, (size_t)h->Ne) != hipSuccess ||
      (rel_pos && dalloc(h, &q.relpos, 2 * (size_t)h->Ne) != hipSuccess)) {
    set_error("hipMalloc failed");
    return TF_ERR_HIP;
  }
  TF_CHECK(hipMemcpy(q.rate, rate.data(), sizeof(double) * rate.size(), hipMemcpyHostToDevice));
  TF_CHECK(hipMemcpy(q.mask, m.data(), m.size(), hipMemcpyHostToDevice));
  if (rel_pos)
    TF_CHECK(hipMemcpy(q.relpos, rel_pos, sizeof(double) * 2 * h->Ne, hipMemcpyHostToDevice));
  h->series.push_back(std::move(q));
  return TF_OK;
}

tf_status tf_nr_set_time(tf_handle* h, double t) {
  if (!h || !std::isfinite(t)) {
    set_error("tf_nr_set_time: invalid arguments");
    return TF_ERR_ARG;
  }
  h->sim_time = t;
  return TF_OK;
}

tf_status tf_nr_get_time(tf_handle* h, double* t) {
  if (!h || !t) {
    set_error("tf_nr_get_time: invalid arguments");
    return TF_ERR_ARG;
  }
  *t = h->sim_time;
  return TF_OK;
}

tf_status tf_nr_apply_stiffness_series(tf_handle* h, double t, double dt) {
  if (!h || !std::isfinite(t) || !std::isfinite(dt)) {
    set_error("tf_nr_apply_stiffness_series: invalid arguments");
    return TF_ERR_ARG;
  }
  tf_status s = apply_series(h, t, dt);
  if (s != TF_OK) return s;
  TF_CHECK(hipStreamSynchronize(h->st));
  return TF_OK;
}

tf_status tf_nr_get_stiffness_multiplier(tf_handle* h, double* out) {
  if (!h || !out) {
    set_error("tf_nr_get_stiffness_multiplier: invalid arguments");
    return TF_ERR_ARG;
  }
  TF_CHECK(hipStreamSynchronize(h->st));
  TF_CHECK(hipMemcpy(out, h->stm, sizeof(double) * h->Ne, hipMemcpyDeviceToHost));
  return TF_OK;
}

}  // extern "C"

namespace {

// F_g and InvFg as identity when no growth tensor was given yet.
tf_status ensure_Fg(tf_handle* h) {
  if (h->Fg) return TF_OK;
  const size_t Ne = h->Ne;
  TF_CHECK(dalloc(h, &h->Fg, Ne * 9));
  std::vector<double> I(Ne * 9, 0.0);
  for (size_t e = 0; e < Ne; ++e) I[9 * e] = I[9 * e + 4] = I[9 * e + 8] = 1.0;
  TF_CHECK(hipMemcpy(h->Fg, I.data(), sizeof(double) * Ne * 9, hipMemcpyHostToDevice));
  TF_CHECK(hipMemcpy(h->invFg, I.data(), sizeof(double) * Ne * 9, hipMemcpyHostToDevice));
  h->have_growth = true;
  return TF_OK;
}

// Device positions of the producers: u (user numbering) staged on the device,
// or the handle's resident positions.
tf_status producer_positions(tf_handle* h, const double* u, const double** out) {
  if (u) {
    tf_status up = upload_nodes(h, h->gr.upos, u);
    if (up != TF_OK) return up;
    *out = h->gr.upos;
    return TF_OK;
  }
  if (!h->have_positions) {
    set_error("no positions on the device: give u");
    return TF_ERR_ARG;
  }
  *out = h->ws.uk;
  return TF_OK;
}

// The relative positions of the prisms at positions u (device): the bounding
// box of every tissue, the prism centres, the apical column copy; the first
// time also the pinned positions (Simulation::initiateSystem sets
// initialRelativePosInBoundingBox from the positions of the initial mesh).
tf_status relpos_update(tf_handle* h, const double* u) {
  GrowthState& g = h->gr;
  TF_CHECK(launch_relative_positions(h->n_tissues, g.node_rng, h->Ne, h->conn, g.prism_tissue, u,
                                     g.node_excl, g.symmetric_y, g.apical_of, g.bb, g.raw,
                                     g.relcur, h->st));
  g.have_cur = true;
  if (!g.have_init) {
    TF_CHECK(hipMemcpyAsync(g.relinit, g.relcur, sizeof(double) * 2 * h->Ne,
                            hipMemcpyDeviceToDevice, h->st));
    g.have_init = true;
  }
  return TF_OK;
}

// The growth of one step at time t from positions u (device): the pinning
// updates due at t (Simulation::checkForPinningPositionsUpdate,
// Simulation.cpp:6169-6187), then every active growth function, the rotation,
// the column volume conservation and F_g.
tf_status growth_run(tf_handle* h, const double* u, double t, double dt) {
  GrowthState& g = h->gr;
  if (!g.have_cur) {
    set_error("growth: update the relative positions first");
    return TF_ERR_ARG;
  }
  if (g.pinned)
    for (size_t i = 0; i < g.pin_times.size(); ++i)
      if (t >= g.pin_times[i] && !g.pin_done[i]) {
        TF_CHECK(hipMemcpyAsync(g.relinit, g.relcur, sizeof(double) * 2 * h->Ne,
                                hipMemcpyDeviceToDevice, h->st));
        g.pin_done[i] = 1;
      }
  tf_status fs = ensure_Fg(h);
  if (fs != TF_OK) return fs;
  const uint8_t* abl = h->n_abl ? h->abl_el : nullptr;
  TF_CHECK(launch_growth(h->Ne, h->conn, h->refpos, u, g.gp, g.d_fns, (int)g.fns.size(), g.cg,
                         g.ca, g.ce, g.pinned ? g.relinit : g.relcur, t, dt, g.interp, abl, h->Fg,
                         h->invFg, g.stage, g.n_columns, g.layers, g.columns, h->st));
  h->have_growth = true;
  return TF_OK;
}

tf_status add_growth_fn(tf_handle* h, GrowthFn f) {
  GrowthState& g = h->gr;
  g.fns.push_back(f);
  TF_CHECK(hipStreamSynchronize(h->st));
  dfree(h, &g.d_fns);
  TF_CHECK(dalloc(h, &g.d_fns, g.fns.size()));
  TF_CHECK(hipMemcpy(g.d_fns, g.fns.data(), sizeof(GrowthFn) * g.fns.size(),
                     hipMemcpyHostToDevice));
  return ensure_Fg(h);
}

tf_status check_fn_args(tf_handle* h, double t0, double t1, int32_t apply, double zmin,
                        double zmax, const char* who) {
  if (!h || !std::isfinite(t0) || !std::isfinite(t1) || apply < 0 || apply > 15 ||
      !std::isfinite(zmin) || !std::isfinite(zmax)) {
    set_error(std::string(who) + ": invalid arguments");
    return TF_ERR_ARG;
  }
  if (!h->gr.prisms) {
    set_error(std::string(who) + ": set the growth prisms first");
    return TF_ERR_ARG;
  }
  return TF_OK;
}

}  // namespace

extern "C" {

tf_status tf_nr_set_growth_prisms(tf_handle* h, const int32_t* tissue_type,
                                  const int32_t* placement, const uint8_t* flags,
                                  const double* weights, const double* rel_z,
                                  const int32_t* apical_of, const uint8_t* node_excluded,
                                  int32_t symmetric_y, int32_t explicit_ecm) {
  if (!h || !tissue_type || !placement || !flags || !rel_z) {
    set_error("tf_nr_set_growth_prisms: invalid arguments");
    return TF_ERR_ARG;
  }
  const int Ne = h->Ne;
  for (int e = 0; e < Ne; ++e) {
    if (tissue_type[e] < 0 || tissue_type[e] > 2) {
      set_error("tf_nr_set_growth_prisms: tissue type of prism " + std::to_string(e) +
                " is not 0, 1 or 2");
      return TF_ERR_ARG;
    }
    if (apical_of && (apical_of[e] < -1 || apical_of[e] >= Ne)) {
      set_error("tf_nr_set_growth_prisms: apical prism of prism " + std::to_string(e) +
                " out of range");
      return TF_ERR_ARG;
    }
    if (!std::isfinite(rel_z[e]) || (weights && (!std::isfinite(weights[2 * e]) ||
                                                 !std::isfinite(weights[2 * e + 1])))) {
      set_error("tf_nr_set_growth_prisms: relative z and weights must be finite");
      return TF_ERR_ARG;
    }
  }
  GrowthState& g = h->gr;
  TF_CHECK(hipStreamSynchronize(h->st));
  if (!g.prisms) {
    TF_CHECK(dalloc(h, &g.type, Ne));
    TF_CHECK(dalloc(h, &g.place, Ne));
    TF_CHECK(dalloc(h, &g.apical_of, Ne));
    TF_CHECK(dalloc(h, &g.prism_tissue, Ne));
    TF_CHECK(dalloc(h, &g.node_rng, 2 * (size_t)h->n_tissues));
    TF_CHECK(dalloc(h, &g.flags, Ne));
    TF_CHECK(dalloc(h, &g.node_excl, (size_t)h->Nn));
    TF_CHECK(dalloc(h, &g.weights, 2 * (size_t)Ne));
    TF_CHECK(dalloc(h, &g.rel_z, Ne));
    TF_CHECK(dalloc(h, &g.bb, 6 * (size_t)h->n_tissues));
    TF_CHECK(dalloc(h, &g.raw, 2 * (size_t)Ne));
    TF_CHECK(dalloc(h, &g.relcur, 2 * (size_t)Ne));
    TF_CHECK(dalloc(h, &g.relinit, 2 * (size_t)Ne));
    TF_CHECK(dalloc(h, &g.upos, 3 * (size_t)h->Nn));
  }
  std::vector<int32_t> ap((size_t)Ne, -1), pt((size_t)Ne, 0), rng(2 * (size_t)h->n_tissues);
  if (apical_of) std::copy(apical_of, apical_of + Ne, ap.begin());
  for (int T = 0; T < h->n_tissues; ++T) {
    for (int e = h->elem_off[T]; e < h->elem_off[T + 1]; ++e) pt[e] = T;
    rng[2 * T] = h->int_off[T];
    rng[2 * T + 1] = h->int_off[T] + (h->user_off[T + 1] - h->user_off[T]);
  }
  std::vector<double> wt(2 * (size_t)Ne, 0.0);
  if (weights) std::copy(weights, weights + 2 * (size_t)Ne, wt.begin());
  std::vector<uint8_t> fl(flags, flags + Ne);
  std::vector<uint8_t> nx((size_t)h->n_user, 0);
  if (node_excluded)
    for (int i = 0; i < h->n_user; ++i) nx[i] = node_excluded[i] ? 1 : 0;
  std::vector<uint8_t> nxi = to_internal<uint8_t>(h, nx.data(), 1, (uint8_t)1);
  TF_CHECK(hipMemcpy(g.type, tissue_type, sizeof(int32_t) * Ne, hipMemcpyHostToDevice));
  TF_CHECK(hipMemcpy(g.place, placement, sizeof(int32_t) * Ne, hipMemcpyHostToDevice));
  TF_CHECK(hipMemcpy(g.apical_of, ap.data(), sizeof(int32_t) * Ne, hipMemcpyHostToDevice));
  TF_CHECK(hipMemcpy(g.prism_tissue, pt.data(), sizeof(int32_t) * Ne, hipMemcpyHostToDevice));
  TF_CHECK(hipMemcpy(g.node_rng, rng.data(), sizeof(int32_t) * rng.size(), hipMemcpyHostToDevice));
  TF_CHECK(hipMemcpy(g.flags, fl.data(), Ne, hipMemcpyHostToDevice));
  TF_CHECK(hipMemcpy(g.node_excl, nxi.data(), nxi.size(), hipMemcpyHostToDevice));
  TF_CHECK(hipMemcpy(g.weights, wt.data(), sizeof(double) * wt.size(), hipMemcpyHostToDevice));
  TF_CHECK(hipMemcpy(g.rel_z, rel_z, sizeof(double) * Ne, hipMemcpyHostToDevice));
  g.gp.type = g.type;
  g.gp.place = g.place;
  g.gp.flags = g.flags;
  g.gp.weights = g.weights;
  g.gp.rel_z = g.rel_z;
  g.gp.explicit_ecm = explicit_ecm ? 1 : 0;
  g.symmetric_y = symmetric_y ? 1 : 0;
  g.prisms = true;
  g.have_cur = g.have_init = false;
  return TF_OK;
}

tf_status tf_nr_add_uniform_growth(tf_handle* h, double t0, double t1, int32_t apply, double zmin,
                                   double zmax, const double* rate, double angle) {
  tf_status a = check_fn_args(h, t0, t1, apply, zmin, zmax, "tf_nr_add_uniform_growth");
  if (a != TF_OK) return a;
  if (!rate || !std::isfinite(rate[0]) || !std::isfinite(rate[1]) || !std::isfinite(rate[2]) ||
      !std::isfinite(angle)) {
    set_error("tf_nr_add_uniform_growth: invalid rates or angle");
    return TF_ERR_ARG;
  }
  GrowthFn f;
  f.type = 1;
  f.apply = apply;
  // GrowthFunctionBase keeps the times as float (GrowthFunctionBase.h:39-40)
  f.t0 = (double)(float)t0;
  f.t1 = (double)(float)t1;
  f.zmin = zmin;
  f.zmax = zmax;
  for (int k = 0; k < 3; ++k) f.rate[k] = rate[k];
  // UniformGrowthFunction's ShearAngleRotationMatrix (GrowthFunctionTypes.h:22-44)
  const double c = std::cos(angle), s = std::sin(angle);
  const double R[9] = {c, -1.0 * s, 0.0, s, c, 0.0, 0.0, 0.0, 1.0};
  for (int q = 0; q < 9; ++q) f.rot[q] = R[q];
  return add_growth_fn(h, f);
}

tf_status tf_nr_add_grid_growth(tf_handle* h, double t0, double t1, int32_t apply, double zmin,
                                double zmax, int32_t nx, int32_t ny, const double* rates,
                                const double* angles_deg) {
  tf_status a = check_fn_args(h, t0, t1, apply, zmin, zmax, "tf_nr_add_grid_growth");
  if (a != TF_OK) return a;
  if (nx < 2 || ny < 2 || (int64_t)nx * ny > (int64_t)1 << 24 || !rates || !angles_deg) {
    set_error("tf_nr_add_grid_growth: the grid needs at least 2 x 2 points, rates and angles");
    return TF_ERR_ARG;
  }
  const size_t np = (size_t)nx * ny;
  for (size_t k = 0; k < np; ++k)
    if (!std::isfinite(angles_deg[k]) || !std::isfinite(rates[3 * k]) ||
        !std::isfinite(rates[3 * k + 1]) || !std::isfinite(rates[3 * k + 2])) {
      set_error("tf_nr_add_grid_growth: rates and angles must be finite");
      return TF_ERR_ARG;
    }
  GrowthState& g = h->gr;
  const size_t cells = (size_t)(nx - 1) * (ny - 1), off = g.h_ca.size() / 4;
  g.h_cg.resize(12 * (off + cells));
  g.h_ca.resize(4 * (off + cells));
  g.h_ce.resize(4 * (off + cells));
  grid_compatible(nx, ny, rates, angles_deg, g.h_cg.data() + 12 * off, g.h_ca.data() + 4 * off,
                  g.h_ce.data() + 4 * off);
  TF_CHECK(hipStreamSynchronize(h->st));
  dfree(h, &g.cg);
  dfree(h, &g.ca);
  dfree(h, &g.ce);
  TF_CHECK(dalloc(h, &g.cg, g.h_cg.size()));
  TF_CHECK(dalloc(h, &g.ca, g.h_ca.size()));
  TF_CHECK(dalloc(h, &g.ce, g.h_ce.size()));
  TF_CHECK(hipMemcpy(g.cg, g.h_cg.data(), sizeof(double) * g.h_cg.size(), hipMemcpyHostToDevice));
  TF_CHECK(hipMemcpy(g.ca, g.h_ca.data(), sizeof(double) * g.h_ca.size(), hipMemcpyHostToDevice));
  TF_CHECK(hipMemcpy(g.ce, g.h_ce.data(), g.h_ce.size(), hipMemcpyHostToDevice));
  GrowthFn f;
  f.type = 3;
  f.apply = apply;
  f.t0 = (double)(float)t0;
  f.t1 = (double)(float)t1;
  f.zmin = zmin;
  f.zmax = zmax;
  f.nx = nx;
  f.ny = ny;
  f.cell_off = (int64_t)off;
  return add_growth_fn(h, f);
}

tf_status tf_nr_set_growth_options(tf_handle* h, int32_t interpolation, int32_t pinned,
                                   int32_t n_pin_times, const double* pin_times) {
  if (!h || interpolation < 0 || interpolation > 1 || n_pin_times < 0 ||
      (n_pin_times > 0 && !pin_times)) {
    set_error("tf_nr_set_growth_options: invalid arguments");
    return TF_ERR_ARG;
  }
  GrowthState& g = h->gr;
  g.interp = interpolation;
  g.pinned = pinned != 0;
  g.pin_times.assign(pin_times, pin_times + n_pin_times);
  g.pin_done.assign(n_pin_times, 0);
  return TF_OK;
}

tf_status tf_nr_set_column_volume_conservation(tf_handle* h, int32_t n_columns, int32_t layers,
                                               const int32_t* columns) {
  if (!h || n_columns < 0 || (n_columns > 0 && (layers < 1 || !columns))) {
    set_error("tf_nr_set_column_volume_conservation: invalid arguments");
    return TF_ERR_ARG;
  }
  GrowthState& g = h->gr;
  std::vector<char> seen((size_t)h->Ne, 0);
  for (int64_t k = 0; k < (int64_t)n_columns * layers; ++k) {
    const int e = columns[k];
    if (e < -1 || e >= h->Ne || (e >= 0 && seen[e])) {
      set_error("tf_nr_set_column_volume_conservation: prism " + std::to_string(e) +
                " out of range or in two columns");
      return TF_ERR_ARG;
    }
    if (e >= 0) seen[e] = 1;
  }
  TF_CHECK(hipStreamSynchronize(h->st));
  dfree(h, &g.columns);
  g.n_columns = n_columns;
  g.layers = layers;
  g.stage.on = n_columns > 0 ? 1 : 0;
  if (!g.stage.on) return TF_OK;
  TF_CHECK(dalloc(h, &g.columns, (size_t)n_columns * layers));
  TF_CHECK(hipMemcpy(g.columns, columns, sizeof(int32_t) * n_columns * layers,
                     hipMemcpyHostToDevice));
  if (!g.vref) {
    const size_t Ne = h->Ne;
    TF_CHECK(dalloc(h, &g.vref, Ne));
    TF_CHECK(dalloc(h, &g.stage.G, 9 * Ne));
    TF_CHECK(dalloc(h, &g.stage.SC, 9 * Ne));
    TF_CHECK(dalloc(h, &g.stage.gsr, 9 * Ne));
    TF_CHECK(dalloc(h, &g.stage.vol, Ne));
    TF_CHECK(dalloc(h, &g.stage.rot, Ne));
    TF_CHECK(launch_ref_volume(h->Ne, h->refpos, g.vref, h->st));
    TF_CHECK(hipStreamSynchronize(h->st));
    g.stage.vref = g.vref;
  }
  return TF_OK;
}

tf_status tf_nr_update_relative_positions(tf_handle* h, const double* u) {
  if (!h || !h->gr.prisms) {
    set_error("tf_nr_update_relative_positions: set the growth prisms first");
    return TF_ERR_ARG;
  }
  const double* ud = nullptr;
  tf_status p = producer_positions(h, u, &ud);
  if (p != TF_OK) return p;
  tf_status r = relpos_update(h, ud);
  if (r != TF_OK) return r;
  TF_CHECK(hipStreamSynchronize(h->st));
  return TF_OK;
}

tf_status tf_nr_relative_positions(tf_handle* h, double* current, double* pinned) {
  if (!h || !h->gr.have_cur || (!current && !pinned)) {
    set_error("tf_nr_relative_positions: no relative positions yet");
    return TF_ERR_ARG;
  }
  TF_CHECK(hipStreamSynchronize(h->st));
  const size_t b = sizeof(double) * 2 * h->Ne;
  if (current) TF_CHECK(hipMemcpy(current, h->gr.relcur, b, hipMemcpyDeviceToHost));
  if (pinned) TF_CHECK(hipMemcpy(pinned, h->gr.relinit, b, hipMemcpyDeviceToHost));
  return TF_OK;
}

tf_status tf_nr_grow(tf_handle* h, const double* u, double t, double dt) {
  if (!h || !std::isfinite(t) || !std::isfinite(dt) || h->gr.fns.empty()) {
    set_error("tf_nr_grow: invalid arguments or no growth functions");
    return TF_ERR_ARG;
  }
  const double* ud = nullptr;
  tf_status p = producer_positions(h, u, &ud);
  if (p != TF_OK) return p;
  tf_status r = growth_run(h, ud, t, dt);
  if (r != TF_OK) return r;
  TF_CHECK(hipStreamSynchronize(h->st));
  return TF_OK;
}

tf_status tf_nr_advance(tf_handle* h, const double* u0, int32_t n_steps, double dt, double tol,
                        int32_t maxit, int32_t precision_mode, double krylov_rtol, double* u_out,
                        tf_step_stats* stats) {
  if (!h || n_steps < 0 || maxit < 0 || maxit > TF_MAX_NR_ITERS - 2 ||
      (precision_mode != TF_PREC_FP64 && precision_mode != TF_PREC_MIXED)) {
    set_error("tf_nr_advance: invalid arguments");
    return TF_ERR_ARG;
  }
  const bool fns = !h->gr.fns.empty();
  if ((!h->have_rates && !fns) || !h->have_flags || (!u0 && !h->have_positions)) {
    set_error("tf_nr_advance needs growth rates or growth functions, surface flags and positions");
    return TF_ERR_ARG;
  }
  const size_t n = 3 * (size_t)h->Nn;
  Workspace& w = h->ws;
  if (u0) {
    tf_status up = upload_nodes(h, w.uk, u0);
    if (up != TF_OK) return up;
    reset_history(h);
  }
  if (h->n_hist.size() != (size_t)h->n_tissues) reset_history(h);
  if (h->warm_start) {
    tf_status ws = ensure_warm(h);
    if (ws != TF_OK) return ws;
  }
  h->have_positions = true;
  const int mixed = precision_mode == TF_PREC_MIXED ? 1 : 0;
  const double rtol = krylov_rtol > 0 ? krylov_rtol : 1e-10;
  tf_status worst = TF_OK;
  std::vector<tf_step_stats> S(h->n_tissues);
  for (int s = 0; s < n_steps; ++s) {
    // pre-step producers on the device, in Simulation::runOneStep's order
    // (Simulation.cpp:5974-6063): the relative positions in the bounding box,
    // the Young's modulus series at the step's start time, growth (functions
    // at the start time, or fixed rates), then the viscosity surfaces from
    // the positions at the start of the step
    if (h->gr.prisms) {
      tf_status rs = relpos_update(h, w.uk);
      if (rs != TF_OK) return rs;
    }
    if (!h->series.empty()) {
      tf_status ss = apply_series(h, h->sim_time, dt);
      if (ss != TF_OK) return ss;
    }
    const uint8_t* abl = h->n_abl ? h->abl_el : nullptr;
    if (fns) {
      tf_status gs = growth_run(h, w.uk, h->sim_time, dt);
      if (gs != TF_OK) return gs;
    } else {
      TF_CHECK(launch_grow(h->Ne, h->rates, dt, h->Fg, h->invFg, h->st, abl));
    }
    h->sim_time += dt;
    TF_CHECK(launch_visc_surface(h->Nn, h->Ne, h->conn, w.uk, h->pat.node_ptr, h->pat.node_src,
                                 h->surf_place, h->surf_type, h->surf_spans, h->area, h->st, abl));
    h->have_area = true;
    if (h->pack.on) {
      // node masses from the grown volumes, then the couples at the step's
      // start positions (Simulation.cpp:6065-6076)
      TF_CHECK(launch_node_mass(h->Nn, h->pat.node_ptr, h->pat.node_src, h->Fg, h->pack.vol,
                                h->pack.mass, h->st));
      h->pack.have_mass = true;
      tf_status ps = packing_detect(h, w.uk);
      if (ps != TF_OK) return ps;
    }
    TF_CHECK(hipMemcpyAsync(w.un, w.uk, sizeof(double) * n, hipMemcpyDeviceToDevice, h->st));
    tf_status r = nr_loop(h, dt, tol, maxit, mixed, rtol, S.data());
    if (stats) std::memcpy(stats + (size_t)s * h->n_tissues, S.data(), sizeof(tf_step_stats) * h->n_tissues);
    if (r != TF_OK && r != TF_ERR_NOT_CONVERGED) {
      reset_history(h);
      return r;
    }
    if (r != TF_OK) worst = r;
    // the step's displacement joins each converged tissue's history
    if (h->warm_start) {
      std::swap(w.hist0, w.hist1);
      TF_CHECK(launch_difference((int64_t)n, w.uk, w.un, w.hist0, h->st));
      for (int T = 0; T < h->n_tissues; ++T)
        h->n_hist[T] = S[T].status == TF_OK && S[T].converged ? std::min(h->n_hist[T] + 1, 2) : 0;
    }
  }
  if (u_out) {
    tf_status dn = download_nodes(h, u_out, w.uk);
    if (dn != TF_OK) return dn;
  }
  TF_CHECK(hipStreamSynchronize(h->st));
  return worst;
}

tf_status tf_batch_create(int32_t n_tissues, const int32_t* nodes_per_tissue,
                          const int32_t* prisms_per_tissue, const int32_t* conn,
                          const double* ref_pos, tf_handle** out) {
  if (!out || n_tissues <= 0 || n_tissues > kMaxTissues || !nodes_per_tissue ||
      !prisms_per_tissue || !conn || !ref_pos) {
    set_error("tf_batch_create: invalid arguments");
    return TF_ERR_ARG;
  }
  *out = nullptr;
  int dev = 0;
  tf_status s = check_device(&dev);
  if (s != TF_OK) return s;
  tf_handle* h = new tf_handle();
  auto fail = [&](tf_status st) {
    tf_nr_destroy(h);
    return st;
  };
  h->batch = true;
  h->n_tissues = n_tissues;
  h->device = dev;
  h->user_off.assign(n_tissues + 1, 0);
  h->int_off.assign(n_tissues + 1, 0);
  h->elem_off.assign(n_tissues + 1, 0);
  for (int t = 0; t < n_tissues; ++t) {
    if (nodes_per_tissue[t] <= 0 || prisms_per_tissue[t] <= 0) {
      set_error("tf_batch_create: empty tissue");
      return fail(TF_ERR_ARG);
    }
    h->user_off[t + 1] = h->user_off[t] + nodes_per_tissue[t];
    // every tissue starts on a slice boundary
    h->int_off[t + 1] = h->int_off[t] + (nodes_per_tissue[t] + kSlice - 1) / kSlice * kSlice;
    h->elem_off[t + 1] = h->elem_off[t] + prisms_per_tissue[t];
  }
  h->n_user = h->user_off[n_tissues];
  h->Nn = h->int_off[n_tissues];
  h->Ne = h->elem_off[n_tissues];
  std::vector<int32_t> iconn((size_t)h->Ne * 6);
  for (int t = 0; t < n_tissues; ++t)
    for (int e = h->elem_off[t]; e < h->elem_off[t + 1]; ++e)
      for (int a = 0; a < 6; ++a) {
        const int nd = conn[6 * (size_t)e + a];
        if (nd < 0 || nd >= nodes_per_tissue[t]) {
          set_error("tf_batch_create: prism node id outside its tissue");
          return fail(TF_ERR_ARG);
        }
        iconn[6 * (size_t)e + a] = nd + h->int_off[t];
      }
  return finish_create(h, iconn.data(), ref_pos, out);
}

tf_status tf_adj_forward(tf_handle* h, int32_t n_steps, const double* u0,
                         const double* inv_Fg_steps, double dt, double tol, int32_t maxit,
                         double krylov_rtol, const int32_t* elem_placement,
                         const int32_t* elem_type, const uint8_t* spans, double* u_final,
                         tf_step_stats* stats) {
  if (!h || n_steps <= 0 || !u0 || !inv_Fg_steps || !u_final || !elem_placement || !elem_type ||
      !spans || maxit < 0 || maxit > TF_MAX_NR_ITERS - 2) {
    set_error("tf_adj_forward: invalid arguments");
    return TF_ERR_ARG;
  }
  if (h->n_pairs > 0 || h->batch || h->pack.on || h->n_abl > 0 || h->any_abl_node) {
    set_error("tf_adj_forward: binding, batches, packing and ablation are not supported by the "
              "adjoint");
    return TF_ERR_ARG;
  }
  const size_t n = 3 * (size_t)h->Nn, Ne = h->Ne;
  Workspace& w = h->ws;
  reset_history(h);  // every forward solve starts from zero
  if (h->adj_cap < n_steps + 1) {
    if (h->adj_ckpt) TF_CHECK(hipFree(h->adj_ckpt));
    h->adj_ckpt = nullptr;
    TF_CHECK(hipMalloc(&h->adj_ckpt, sizeof(double) * n * (n_steps + 1)));
    h->adj_cap = n_steps + 1;
  }
  if (!h->adj_place) {
    TF_CHECK(dalloc(h, &h->adj_place, Ne));
    TF_CHECK(dalloc(h, &h->adj_type, Ne));
    TF_CHECK(dalloc(h, &h->adj_spans, Ne));
    TF_CHECK(dalloc(h, &h->adj_s, Ne));
    TF_CHECK(dalloc(h, &h->adj_b, Ne * 18));
    TF_CHECK(dalloc(h, &h->adj_dinvFg, Ne * 9));
    TF_CHECK(dalloc(h, &h->adj_valT, (size_t)h->pat.n_slots * 9));
    TF_CHECK(dalloc(h, &h->adj_minvT, (size_t)h->Nn * 9));
    TF_CHECK(dalloc(h, &h->adj_a, n));
  }
  TF_CHECK(hipMemcpy(h->adj_place, elem_placement, sizeof(int32_t) * Ne, hipMemcpyHostToDevice));
  TF_CHECK(hipMemcpy(h->adj_type, elem_type, sizeof(int32_t) * Ne, hipMemcpyHostToDevice));
  TF_CHECK(hipMemcpy(h->adj_spans, spans, Ne, hipMemcpyHostToDevice));
  TF_CHECK(hipMemcpy(h->adj_ckpt, u0, sizeof(double) * n, hipMemcpyHostToDevice));
  tf_step_stats total;
  std::memset(&total, 0, sizeof(total));
  const double rtol = krylov_rtol > 0 ? krylov_rtol : 1e-10;
  for (int s = 0; s < n_steps; ++s) {
    const double* prev = h->adj_ckpt + n * s;
    TF_CHECK(hipMemcpyAsync(h->invFg, inv_Fg_steps + 9 * Ne * s, sizeof(double) * 9 * Ne,
                            hipMemcpyHostToDevice, h->st));
    TF_CHECK(launch_visc_surface(h->Nn, h->Ne, h->conn, prev, h->pat.node_ptr, h->pat.node_src,
                                 h->adj_place, h->adj_type, h->adj_spans, h->area, h->st));
    h->have_growth = h->have_area = true;
    TF_CHECK(hipMemcpyAsync(w.un, prev, sizeof(double) * n, hipMemcpyDeviceToDevice, h->st));
    TF_CHECK(hipMemcpyAsync(w.uk, prev, sizeof(double) * n, hipMemcpyDeviceToDevice, h->st));
    tf_step_stats S;
    std::memset(&S, 0, sizeof(S));
    tf_status r = nr_loop(h, dt, tol, maxit, 0, rtol, &S);
    if (r != TF_OK) return r;
    TF_CHECK(hipMemcpyAsync(h->adj_ckpt + n * (s + 1), w.uk, sizeof(double) * n,
                            hipMemcpyDeviceToDevice, h->st));
    total.nr_iterations += S.nr_iterations;
    total.krylov_iterations_total += S.krylov_iterations_total;
    total.ms_element += S.ms_element;
    total.ms_assembly += S.ms_assembly;
    total.ms_krylov += S.ms_krylov;
  }
  TF_CHECK(hipMemcpyAsync(u_final, h->adj_ckpt + n * n_steps, sizeof(double) * n,
                          hipMemcpyDeviceToHost, h->st));
  TF_CHECK(hipStreamSynchronize(h->st));
  total.converged = 1;
  total.ms_total = total.ms_element + total.ms_assembly + total.ms_krylov;
  h->adj_steps = n_steps;
  h->adj_dt = dt;
  if (stats) *stats = total;
  return TF_OK;
}

tf_status tf_adj_backward(tf_handle* h, const double* inv_Fg_steps, const double* dinv_Fg_steps,
                          const double* dL_du_final, double krylov_rtol, double* dL_dr,
                          tf_step_stats* stats) {
  if (!h || !inv_Fg_steps || !dinv_Fg_steps || !dL_du_final || !dL_dr) {
    set_error("tf_adj_backward: invalid arguments");
    return TF_ERR_ARG;
  }
  if (h->adj_steps <= 0) {
    set_error("tf_adj_backward needs a preceding tf_adj_forward");
    return TF_ERR_ARG;
  }
  const size_t n = 3 * (size_t)h->Nn, Ne = h->Ne;
  Workspace& w = h->ws;
  const double dt = h->adj_dt;
  const double rtol = krylov_rtol > 0 ? krylov_rtol : 1e-12;
  std::vector<double> acc(Ne, 0.0), s_host(Ne);
  tf_step_stats total;
  std::memset(&total, 0, sizeof(total));
  TF_CHECK(hipMemcpy(h->adj_a, dL_du_final, sizeof(double) * n, hipMemcpyHostToDevice));
  hipEvent_t e0 = h->ev[4], e1 = h->ev[5];
  for (int s = h->adj_steps - 1; s >= 0; --s) {
    const double* prev = h->adj_ckpt + n * s;
    const double* cur = h->adj_ckpt + n * (s + 1);
    TF_CHECK(hipMemcpyAsync(h->invFg, inv_Fg_steps + 9 * Ne * s, sizeof(double) * 9 * Ne,
                            hipMemcpyHostToDevice, h->st));
    TF_CHECK(hipMemcpyAsync(h->adj_dinvFg, dinv_Fg_steps + 9 * Ne * s, sizeof(double) * 9 * Ne,
                            hipMemcpyHostToDevice, h->st));
    TF_CHECK(launch_visc_surface(h->Nn, h->Ne, h->conn, prev, h->pat.node_ptr, h->pat.node_src,
                                 h->adj_place, h->adj_type, h->adj_spans, h->area, h->st));
    TF_CHECK(hipMemcpyAsync(w.un, prev, sizeof(double) * n, hipMemcpyDeviceToDevice, h->st));
    TF_CHECK(hipMemcpyAsync(w.uk, cur, sizeof(double) * n, hipMemcpyDeviceToDevice, h->st));
    tf_status r = assemble(h, dt, 0);
    if (r != TF_OK) return r;
    TF_CHECK(hipEventRecord(e0, h->st));
    // J^T mu = a with the transposed operator and transposed block-Jacobi
    TF_CHECK(launch_transpose_operator(h->pat.n_slots, h->pat.tslot, w.val, h->adj_valT, h->Nn,
                                       w.minv, h->adj_minvT, h->st));
    KrylovArgs k = krylov_args(h);
    k.val = h->adj_valT;
    k.minv = h->adj_minvT;
    int kit = 0, state = 0;
    TF_CHECK(krylov_bicgstab(k, h->adj_a, w.du, rtol, 50000, &kit, &state, 16));
    if (state != K_CONVERGED) {
      set_error("adjoint BiCGStab did not converge at step " + std::to_string(s));
      return TF_ERR_KRYLOV;
    }
    AdjointLaunch l;
    l.n_prisms = h->Ne; l.n_nodes = h->Nn; l.conn = h->conn; l.refpos = h->refpos;
    l.invFg = h->invFg; l.dinvFg = h->adj_dinvFg; l.lambda = h->lam; l.mu = h->mu;
    l.eta = h->eta; l.ext_visc = h->extv; l.dof_flags = h->dof_flags; l.uk = cur; l.un = prev;
    l.dt = dt; l.m = w.du; l.placement = h->adj_place; l.type = h->adj_type;
    l.spans = h->adj_spans; l.area = h->area; l.node_ptr = h->pat.node_ptr;
    l.node_src = h->pat.node_src; l.s_out = h->adj_s; l.b_out = h->adj_b; l.a_prev = h->adj_a;
    TF_CHECK(launch_adjoint_element(l, h->st));
    TF_CHECK(hipEventRecord(e1, h->st));
    TF_CHECK(hipMemcpyAsync(s_host.data(), h->adj_s, sizeof(double) * Ne, hipMemcpyDeviceToHost,
                            h->st));
    TF_CHECK(hipStreamSynchronize(h->st));
    float ms_as = 0, ms_kr = 0;
    TF_CHECK(hipEventElapsedTime(&ms_as, h->ev[0], h->ev[2]));
    TF_CHECK(hipEventElapsedTime(&ms_kr, e0, e1));
    total.ms_assembly += ms_as;
    total.ms_krylov += ms_kr;
    total.krylov_iterations_total += kit;
    total.nr_iterations += 1;
    for (size_t e = 0; e < Ne; ++e) acc[e] += s_host[e];
  }
  std::memcpy(dL_dr, acc.data(), sizeof(double) * Ne);
  total.converged = 1;
  total.ms_total = total.ms_assembly + total.ms_krylov;
  if (stats) *stats = total;
  return TF_OK;
}

}  // extern "C"

namespace {
// Block columns of every row: the element neighbours and, after a packing
// detection, the couples' partners, ascending.  jslot(C, R) is the J slot
// of block (C, R).
struct HostRows {
  std::vector<int32_t> ptr, cols;
  std::vector<int32_t> cnt, inc;
  int cap = 0;
};
tf_status host_rows(tf_handle* h, HostRows& r) {
  const int Nn = h->Nn;
  const bool pk = h->pack.on && h->pack.detected;
  if (pk) {
    r.cap = h->pack.cap;
    r.cnt.resize(Nn);
    r.inc.resize((size_t)Nn * r.cap);
    TF_CHECK(hipMemcpy(r.cnt.data(), h->pack.cnt, sizeof(int32_t) * Nn, hipMemcpyDeviceToHost));
    TF_CHECK(hipMemcpy(r.inc.data(), h->pack.inc, sizeof(int32_t) * r.inc.size(),
                       hipMemcpyDeviceToHost));
  }
  r.ptr.assign(Nn + 1, 0);
  r.cols.clear();
  for (int R = 0; R < Nn; ++R) {
    const size_t b = r.cols.size();
    r.cols.insert(r.cols.end(), h->h_nbr.begin() + h->h_rowptr[R],
                  h->h_nbr.begin() + h->h_rowptr[R + 1]);
    if (pk) {
      for (int e = 0; e < std::min(r.cnt[R], r.cap); ++e)
        r.cols.push_back(r.inc[(size_t)R * r.cap + e]);
      std::sort(r.cols.begin() + b, r.cols.end());
    }
    r.ptr[R + 1] = (int32_t)r.cols.size();
  }
  return TF_OK;
}
int64_t jslot(const tf_handle* h, const HostRows& r, int C, int R) {
  const int32_t* b = h->h_nbr.data() + h->h_rowptr[C];
  const int32_t* e = h->h_nbr.data() + h->h_rowptr[C + 1];
  const int32_t* p = std::lower_bound(b, e, R);
  const int s = C / kSlice;
  if (p != e && *p == R) return h->h_slice_base[s] + (int64_t)(p - b) * kSlice + (C % kSlice);
  if (r.cnt.empty()) return -1;
  for (int k = 0; k < std::min(r.cnt[C], r.cap); ++k)
    if (r.inc[(size_t)C * r.cap + k] == R)
      return h->h_slice_base[s] + (int64_t)(h->h_slice_ew[s] + k) * kSlice + (C % kSlice);
  return -1;
}
}  // namespace

extern "C" {

tf_status tf_nr_set_packing(tf_handle* h, const int32_t* node_placement,
                            const uint8_t* node_excluded, int32_t symmetric_y,
                            int32_t capacity) {
  if (!h) {
    set_error("tf_nr_set_packing: invalid handle");
    return TF_ERR_ARG;
  }
  PackDev& p = h->pack;
  if (node_placement && (h->n_abl > 0 || h->any_abl_node)) {
    set_error("node-node packing together with laser ablation is not supported");
    return TF_ERR_ARG;
  }
  if (!node_placement) {  // off: every contact slot unused
    if (p.on) {
      TF_CHECK(hipMemsetAsync(p.cnt, 0, sizeof(int32_t) * h->Nn, h->st));
      TF_CHECK(launch_contact_slots(p, h->Nn, h->pat, h->st));
      TF_CHECK(hipStreamSynchronize(h->st));
    }
    p.on = false;
    p.detected = false;
    h->assembled = false;
    return TF_OK;
  }
  if (h->batch || h->n_pairs > 0) {
    set_error("tf_nr_set_packing: single-tissue handles without slave/master binding only");
    return TF_ERR_ARG;
  }
  if (capacity < 1 || capacity > kPackLocal) {
    set_error("tf_nr_set_packing: capacity must lie in [1, " + std::to_string(kPackLocal) + "]");
    return TF_ERR_ARG;
  }
  const int Nn = h->Nn, Ne = h->Ne;
  std::vector<int32_t> place(Nn);
  int n_cand = 0;
  for (int k = 0; k < Nn; ++k) {
    const bool ok = (node_placement[k] == 0 || node_placement[k] == 1) &&
                    !(node_excluded && node_excluded[k]);
    place[k] = ok ? node_placement[k] : -1;
    n_cand += ok ? 1 : 0;
  }
  if (!p.place) {
    TF_CHECK(dalloc(h, &p.place, Nn));
    TF_CHECK(dalloc(h, &p.vol, Ne));
    TF_CHECK(dalloc(h, &p.mass, Nn));
    TF_CHECK(dalloc(h, &p.cnt, Nn));
    TF_CHECK(dalloc(h, &p.par, 64));
    TF_CHECK(dalloc(h, &p.force, 3 * (size_t)Nn));
    TF_CHECK(dalloc(h, &p.next, Nn));
    TF_CHECK(dalloc(h, &p.cell, Ne));
    TF_CHECK(dalloc(h, &p.eval, 3 * (size_t)Ne));
    TF_CHECK(dalloc(h, &p.ecnt, Ne));
    TF_CHECK(dalloc(h, &p.flag, 2));
    p.n_chunks = (Ne + 2047) / 2048;
    TF_CHECK(dalloc(h, &p.part, (size_t)p.n_chunks * 52 * 2));
    int hs = 1024;
    while (hs < 2 * Nn) hs *= 2;
    p.hash_size = hs;
    TF_CHECK(dalloc(h, &p.head, hs));
    TF_CHECK(launch_ref_volume(Ne, h->refpos, p.vol, h->st));
    TF_CHECK(hipMemsetAsync(p.cnt, 0, sizeof(int32_t) * Nn, h->st));
    TF_CHECK(hipMemsetAsync(p.force, 0, sizeof(double) * 3 * Nn, h->st));
  }
  TF_CHECK(hipMemcpy(p.place, place.data(), sizeof(int32_t) * Nn, hipMemcpyHostToDevice));
  p.n_cand = n_cand;
  p.symmetric_y = symmetric_y ? 1 : 0;
  p.on = true;
  p.detected = false;
  // contact columns in the slices of the nodes that may pack
  const int n_slices = (Nn + kSlice - 1) / kSlice;
  std::vector<int32_t> caps(n_slices, 0);
  for (int k = 0; k < Nn; ++k)
    if (place[k] >= 0) caps[k / kSlice] = capacity;
  if (caps != h->h_slice_cap || !p.inc) {
    tf_status st = packing_relayout(h, capacity);
    if (st != TF_OK) return st;
  }
  return TF_OK;
}

tf_status tf_nr_set_node_mass(tf_handle* h, const double* mass) {
  if (!h || !mass || !h->pack.on) {
    set_error("tf_nr_set_node_mass: invalid arguments or packing off");
    return TF_ERR_ARG;
  }
  TF_CHECK(hipMemcpy(h->pack.mass, mass, sizeof(double) * h->Nn, hipMemcpyHostToDevice));
  h->pack.have_mass = true;
  return TF_OK;
}

tf_status tf_nr_update_node_mass(tf_handle* h, const double* Fg, double* mass_out) {
  if (!h || !h->pack.on || (!Fg && !h->Fg)) {
    set_error("tf_nr_update_node_mass: invalid arguments, packing off or no F_g");
    return TF_ERR_ARG;
  }
  PackDev& p = h->pack;
  const double* F = h->Fg;
  if (Fg) {
    double* d = nullptr;
    TF_CHECK(hipMalloc(&d, sizeof(double) * 9 * (size_t)h->Ne));
    hipError_t e = hipMemcpy(d, Fg, sizeof(double) * 9 * (size_t)h->Ne, hipMemcpyHostToDevice);
    if (e == hipSuccess)
      e = launch_node_mass(h->Nn, h->pat.node_ptr, h->pat.node_src, d, p.vol, p.mass, h->st);
    if (e == hipSuccess) e = hipStreamSynchronize(h->st);
    (void)hipFree(d);
    TF_CHECK(e);
  } else {
    TF_CHECK(launch_node_mass(h->Nn, h->pat.node_ptr, h->pat.node_src, F, p.vol, p.mass, h->st));
  }
  if (mass_out)
    TF_CHECK(hipMemcpyAsync(mass_out, p.mass, sizeof(double) * h->Nn, hipMemcpyDeviceToHost,
                            h->st));
  TF_CHECK(hipStreamSynchronize(h->st));
  p.have_mass = true;
  return TF_OK;
}

tf_status tf_nr_detect_packing(tf_handle* h, const double* u, int32_t* n_couples) {
  if (!h || !h->pack.on || (!u && !h->have_positions)) {
    set_error("tf_nr_detect_packing: invalid arguments, packing off or no positions");
    return TF_ERR_ARG;
  }
  const double* src = h->ws.uk;
  if (u) {
    tf_status up = upload_nodes(h, h->ws.tmp, u);
    if (up != TF_OK) return up;
    src = h->ws.tmp;
  }
  tf_status st = packing_detect(h, src);
  if (st != TF_OK) return st;
  if (n_couples) *n_couples = h->pack.n_pairs;
  return TF_OK;
}

tf_status tf_nr_get_packing(tf_handle* h, int32_t* n_couples, int32_t* couples, double* forces,
                            double* thresholds) {
  if (!h || !h->pack.on || !h->pack.detected || !n_couples) {
    set_error("tf_nr_get_packing: invalid arguments or no detection");
    return TF_ERR_ARG;
  }
  PackDev& p = h->pack;
  const int Nn = h->Nn;
  if (thresholds) {
    double par[58];
    TF_CHECK(hipMemcpy(par, p.par, sizeof(par), hipMemcpyDeviceToHost));
    thresholds[0] = par[0];
    for (int q = 0; q < 50; ++q) thresholds[1 + q] = par[2 + q];
    for (int q = 0; q < 6; ++q) thresholds[51 + q] = par[52 + q];
  }
  if (!couples && !forces) {
    *n_couples = p.n_pairs;
    return TF_OK;
  }
  if (*n_couples < p.n_pairs) {
    set_error("tf_nr_get_packing: output too small");
    return TF_ERR_ARG;
  }
  std::vector<int32_t> cnt(Nn), inc((size_t)Nn * p.cap);
  std::vector<double> f(3 * (size_t)Nn);
  TF_CHECK(hipMemcpy(cnt.data(), p.cnt, sizeof(int32_t) * Nn, hipMemcpyDeviceToHost));
  TF_CHECK(hipMemcpy(inc.data(), p.inc, sizeof(int32_t) * inc.size(), hipMemcpyDeviceToHost));
  TF_CHECK(hipMemcpy(f.data(), p.force, sizeof(double) * f.size(), hipMemcpyDeviceToHost));
  int c = 0;
  for (int k = 0; k < Nn; ++k)
    for (int e = 0; e < std::min(cnt[k], p.cap); ++e) {
      const int o = inc[(size_t)k * p.cap + e];
      if (o <= k) continue;
      if (couples) {
        couples[2 * c] = k;
        couples[2 * c + 1] = o;
      }
      if (forces)
        for (int i = 0; i < 3; ++i) {
          forces[6 * c + i] = f[3 * (size_t)k + i];
          forces[6 * c + 3 + i] = f[3 * (size_t)o + i];
        }
      ++c;
    }
  *n_couples = c;
  return TF_OK;
}

tf_status tf_csr_pattern(tf_handle* h, int32_t* ia, int32_t* ja, int64_t* nnz) {
  if (h && h->batch) {
    set_error("tf_csr_pattern works on single-tissue handles only");
    return TF_ERR_ARG;
  }
  if (!h || !nnz) {
    set_error("tf_csr_pattern: invalid arguments");
    return TF_ERR_ARG;
  }
  const int Nn = h->Nn;
  HostRows rows;
  tf_status hs = host_rows(h, rows);
  if (hs != TF_OK) return hs;
  const int64_t total = 9 * (int64_t)rows.cols.size();
  if (!ja) {
    *nnz = total;
    return TF_OK;
  }
  if (*nnz < total) {
    set_error("tf_csr_pattern: ja too small");
    return TF_ERR_ARG;
  }
  int64_t c = 0;
  for (int R = 0; R < Nn; ++R)
    for (int i = 0; i < 3; ++i) {
      if (ia) ia[3 * R + i] = (int32_t)c;
      for (int k = rows.ptr[R]; k < rows.ptr[R + 1]; ++k)
        for (int j = 0; j < 3; ++j) ja[c++] = 3 * rows.cols[k] + j;
    }
  if (ia) ia[3 * Nn] = (int32_t)c;
  *nnz = c;
  return TF_OK;
}

tf_status tf_nr_get_csr(tf_handle* h, double threshold, int32_t* ia, int32_t* ja, double* a,
                        int64_t* nnz) {
  if (h && h->batch) {
    set_error("tf_nr_get_csr works on single-tissue handles only");
    return TF_ERR_ARG;
  }
  if (!h || !nnz || !ia) {
    set_error("tf_nr_get_csr: invalid arguments");
    return TF_ERR_ARG;
  }
  if (!h->assembled) {
    set_error("tf_nr_get_csr needs a preceding tf_nr_assemble");
    return TF_ERR_ARG;
  }
  const int Nn = h->Nn;
  const int64_t S = h->pat.n_slots;
  std::vector<double> val((size_t)S * 9);
  if (h->last_mixed) {  // the FP32 operator of the mixed assembly
    std::vector<float> v32(val.size());
    TF_CHECK(hipMemcpy(v32.data(), h->ws.val32, sizeof(float) * v32.size(),
                       hipMemcpyDeviceToHost));
    for (size_t i = 0; i < val.size(); ++i) val[i] = v32[i];
  } else {
    TF_CHECK(hipMemcpy(val.data(), h->ws.val, sizeof(double) * val.size(), hipMemcpyDeviceToHost));
  }
  HostRows hr;
  tf_status hs = host_rows(h, hr);
  if (hs != TF_OK) return hs;
  // K(3R+i, 3C+j) = J(3C+j, 3R+i) = J block (C, R)[j][i]
  const int n = 3 * Nn;
  std::vector<std::map<int, double>> rows;
  const bool bind = h->n_pairs > 0;
  if (bind) rows.resize(n);
  std::vector<int32_t> tja;
  std::vector<double> ta;
  std::vector<int32_t> tia(n + 1, 0);
  for (int R = 0; R < Nn; ++R)
    for (int i = 0; i < 3; ++i) {
      tia[3 * R + i] = (int32_t)tja.size();
      for (int k = hr.ptr[R]; k < hr.ptr[R + 1]; ++k) {
        const int C = hr.cols[k];
        const int64_t slot = jslot(h, hr, C, R);
        for (int j = 0; j < 3; ++j) {
          const double v = val[(size_t)(3 * j + i) * S + slot];
          if (bind) {
            rows[3 * R + i][3 * C + j] = v;
          } else if (v > threshold || v < -threshold) {
            tja.push_back(3 * C + j);
            ta.push_back(v);
          }
        }
      }
    }
  tia[n] = (int32_t)tja.size();
  if (bind) {
    // calculateBoundKWithSlavesMasterDoF on the sparse rows (NewtonRaphsonSolver.cpp:148-180)
    std::vector<std::map<int, double>> cols(n);
    for (int r = 0; r < n; ++r)
      for (auto& kv : rows[r]) cols[kv.first][r] = kv.second;
    auto get = [&](int r, int c) -> double {
      auto it = rows[r].find(c);
      return it == rows[r].end() ? 0.0 : it->second;
    };
    auto set = [&](int r, int c, double v) {
      rows[r][c] = v;
      cols[c][r] = v;
    };
    for (int p = 0; p < h->n_pairs; ++p) {
      const int s = h->h_pairs[2 * p], m = h->h_pairs[2 * p + 1];
      std::vector<std::pair<int, double>> srow(rows[s].begin(), rows[s].end());
      for (auto& kv : srow) set(m, kv.first, get(m, kv.first) + kv.second);
      for (auto& kv : srow) set(s, kv.first, 0.0);
      std::vector<std::pair<int, double>> scol(cols[s].begin(), cols[s].end());
      for (auto& kv : scol) set(kv.first, m, get(kv.first, m) + kv.second);
      for (auto& kv : scol) set(kv.first, s, 0.0);
      set(s, s, 1.0);
    }
    tja.clear();
    ta.clear();
    for (int r = 0; r < n; ++r) {
      tia[r] = (int32_t)tja.size();
      for (auto& kv : rows[r])
        if (kv.second > threshold || kv.second < -threshold) {
          tja.push_back(kv.first);
          ta.push_back(kv.second);
        }
    }
    tia[n] = (int32_t)tja.size();
  }
  const int64_t total = (int64_t)tja.size();
  std::memcpy(ia, tia.data(), sizeof(int32_t) * (n + 1));
  if (!ja) {
    *nnz = total;
    return TF_OK;
  }
  if (*nnz < total) {
    set_error("tf_nr_get_csr: output too small");
    return TF_ERR_ARG;
  }
  std::memcpy(ja, tja.data(), sizeof(int32_t) * total);
  if (a) std::memcpy(a, ta.data(), sizeof(double) * total);
  *nnz = total;
  return TF_OK;
}

tf_status tf_nr_element_outputs(tf_handle* h, double* ge, double* gv, double* Ke, double* Kv) {
  if (!h) {
    set_error("tf_nr_element_outputs: invalid handle");
    return TF_ERR_ARG;
  }
  if (!h->assembled) {
    set_error("tf_nr_element_outputs needs a preceding assembly");
    return TF_ERR_ARG;
  }
  const size_t Ne = h->Ne;
  if (h->last_mixed) {
    // the mixed assembly keeps only the summed FP32 blocks and the summed
    // forces: the separate forces and FP32 tangents are recomputed from the
    // assembly's inputs
    if (!h->assembly_inputs_kept) {
      set_error("tf_nr_element_outputs: after a mixed assembly it needs tf_nr_assemble_ex");
      return TF_ERR_ARG;
    }
    Workspace& w = h->ws;
    ElementArgs ea{h->Ne, h->conn, h->refpos, h->invFg, h->lam, h->mu, h->eta, w.uk, w.un,
                   h->last_dt, h->any_visc ? 1 : 0, w.elemKe, w.elemKv, w.elemKe32, w.elemKv32,
                   w.elemFe, w.elemFv};
    TF_CHECK(launch_element(ea, 1, h->st));
    TF_CHECK(launch_ablate_zero(ea, 1, h->abl_list, h->n_abl, h->st));
    TF_CHECK(hipStreamSynchronize(h->st));
  }
  if (ge) TF_CHECK(hipMemcpy(ge, h->ws.elemFe, sizeof(double) * Ne * 18, hipMemcpyDeviceToHost));
  if (gv) TF_CHECK(hipMemcpy(gv, h->ws.elemFv, sizeof(double) * Ne * 18, hipMemcpyDeviceToHost));
  std::vector<double> blk(Ne * 324);
  std::vector<float> blk32;
  auto fetch = [&](const double* src64, const float* src32) -> hipError_t {
    if (!h->last_mixed)
      return hipMemcpy(blk.data(), src64, sizeof(double) * blk.size(), hipMemcpyDeviceToHost);
    blk32.resize(blk.size());
    hipError_t e = hipMemcpy(blk32.data(), src32, sizeof(float) * blk32.size(),
                             hipMemcpyDeviceToHost);
    for (size_t i = 0; i < blk.size(); ++i) blk[i] = blk32[i];
    return e;
  };
  auto unpack = [&](double* out) {
    for (size_t e = 0; e < Ne; ++e)
      for (int A = 0; A < 6; ++A)
        for (int B = 0; B < 6; ++B)
          for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c)
              out[e * 324 + (3 * A + r) * 18 + 3 * B + c] = blk[((e * 36) + B * 6 + A) * 9 + 3 * r + c];
  };
  if (Ke) {
    TF_CHECK(fetch(h->ws.elemKe, h->ws.elemKe32));
    unpack(Ke);
  }
  if (Kv) {
    if (h->any_visc) {
      TF_CHECK(fetch(h->ws.elemKv, h->ws.elemKv32));
      unpack(Kv);
    } else {
      std::memset(Kv, 0, sizeof(double) * Ne * 324);
    }
  }
  return TF_OK;
}

tf_status tf_nr_element_strain(tf_handle* h, double* strain) {
  if (!h || !strain) {
    set_error("tf_nr_element_strain: invalid arguments");
    return TF_ERR_ARG;
  }
  if (!h->have_strain) {
    set_error("tf_nr_element_strain needs a preceding assembly or step");
    return TF_ERR_ARG;
  }
  TF_CHECK(hipStreamSynchronize(h->st));
  TF_CHECK(hipMemcpy(strain, h->ws.strain, sizeof(double) * 6 * (size_t)h->Ne,
                     hipMemcpyDeviceToHost));
  return TF_OK;
}

tf_status tf_nr_autotune(tf_handle* h, int32_t precision_mode, int32_t* slices_per_tile,
                         double* ms_per_iteration) {
  if (!h || (precision_mode != TF_PREC_FP64 && precision_mode != TF_PREC_MIXED)) {
    set_error("tf_nr_autotune: invalid arguments");
    return TF_ERR_ARG;
  }
  if (!h->assembled || (precision_mode == TF_PREC_MIXED && !h->ws.val32)) {
    set_error("tf_nr_autotune needs an assembled operator of the requested precision");
    return TF_ERR_ARG;
  }
  const int mixed = precision_mode == TF_PREC_MIXED ? 1 : 0;
  const int cands[] = {2, 4, 8, 16, 32};
  int best = h->spt;
  double best_ms = 1e300;
  for (int spt : cands) {
    tf_status s = build_tiles(h, spt);
    if (s != TF_OK) return s;
    KrylovArgs k = krylov_args(h);
    float ms = 0.f;
    const int method = h->krylov_method[mixed];
    TF_CHECK(tiled_time_iterations(k, mixed, method, 2, &ms));  // warm-up
    TF_CHECK(tiled_time_iterations(k, mixed, method, 10, &ms));
    const double per = ms / 10.0;
    if (per < best_ms) {
      best_ms = per;
      best = spt;
    }
  }
  tf_status s = build_tiles(h, best);
  if (s != TF_OK) return s;
  h->autotuned_spt = best;
  if (slices_per_tile) *slices_per_tile = best;
  if (ms_per_iteration) *ms_per_iteration = best_ms;
  return TF_OK;
}

tf_status tf_nr_set_krylov(tf_handle* h, int32_t precision_mode, int32_t method) {
  if (!h || (precision_mode != TF_PREC_FP64 && precision_mode != TF_PREC_MIXED) ||
      (method != TF_KRYLOV_BICGSTAB && method != TF_KRYLOV_CG && method != TF_KRYLOV_GMRES) ||
      (method == TF_KRYLOV_GMRES && precision_mode == TF_PREC_MIXED)) {
    set_error("tf_nr_set_krylov: invalid arguments (GMRES is FP64 only)");
    return TF_ERR_ARG;
  }
  h->krylov_method[precision_mode == TF_PREC_MIXED ? 1 : 0] =
      method == TF_KRYLOV_CG ? KRYLOV_CG
                             : (method == TF_KRYLOV_GMRES ? KRYLOV_GMRES : KRYLOV_BICGSTAB);
  return TF_OK;
}

tf_status tf_nr_set_tile_size(tf_handle* h, int32_t slices_per_tile) {
  if (!h || slices_per_tile < 1 || slices_per_tile > 4096) {
    set_error("tf_nr_set_tile_size: invalid arguments");
    return TF_ERR_ARG;
  }
  return build_tiles(h, slices_per_tile);
}

tf_status tf_nr_profile(tf_handle* h, int32_t enable) {
  if (!h) {
    set_error("tf_nr_profile: invalid handle");
    return TF_ERR_ARG;
  }
  if (enable && !h->prof.ev[0])
    for (auto& e : h->prof.ev) TF_CHECK(hipEventCreate(&e));
  h->prof.on = enable ? 1 : 0;
  h->prof.used = 0;
  for (int i = 0; i < 4; ++i) {
    h->prof.ms[i] = 0.0;
    h->prof.count[i] = 0;
  }
  h->prof_el_ms = h->prof_as_ms = 0.0;
  h->prof_assemblies = 0;
  return TF_OK;
}

tf_status tf_nr_kernel_times(tf_handle* h, tf_kernel_times* out) {
  if (!h || !out) {
    set_error("tf_nr_kernel_times: invalid arguments");
    return TF_ERR_ARG;
  }
  TF_CHECK(hipStreamSynchronize(h->st));
  prof_harvest(&h->prof);
  for (int i = 0; i < 2; ++i) {
    out->spmv_ms[i] = h->prof.ms[i];
    out->spmv_launches[i] = h->prof.count[i];
    out->spmv_t_ms[i] = h->prof.ms[2 + i];
    out->spmv_t_launches[i] = h->prof.count[2 + i];
  }
  out->element_ms = h->prof_el_ms;
  out->assembly_ms = h->prof_as_ms;
  out->assemblies = h->prof_assemblies;
  return TF_OK;
}

tf_status tf_copy_bandwidth(int64_t bytes, int32_t reps, double* gb_per_s) {
  if (bytes < (1 << 20) || reps < 1 || !gb_per_s) {
    set_error("tf_copy_bandwidth: invalid arguments");
    return TF_ERR_ARG;
  }
  int dev = 0;
  tf_status s = check_device(&dev);
  if (s != TF_OK) return s;
  const int64_t n4 = bytes / 16;
  void *src = nullptr, *dst = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipMalloc(&src, n4 * 16);
  if (e == hipSuccess) e = hipMalloc(&dst, n4 * 16);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = hipMemsetAsync(src, 0, n4 * 16, st);
  double best = 0.0;
  for (int v = 0; v < kCopyVariants && e == hipSuccess; ++v) {
    float ms = 0.f;
    e = launch_copy(src, dst, n4, v, st);  // warm-up
    if (e == hipSuccess) e = hipEventRecord(e0, st);
    for (int r = 0; r < reps && e == hipSuccess; ++r) e = launch_copy(src, dst, n4, v, st);
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e == hipSuccess && ms > 0.f)
      best = std::max(best, 2.0 * (double)(n4 * 16) * reps / (ms * 1e-3) / 1e9);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (src) (void)hipFree(src);
  if (dst) (void)hipFree(dst);
  if (e != hipSuccess) {
    set_error(std::string("tf_copy_bandwidth: ") + hipGetErrorString(e));
    return TF_ERR_HIP;
  }
  *gb_per_s = best;
  return TF_OK;
}

tf_status tf_nr_set_coarse(tf_handle* h, double cell_size) {
  if (!h || std::isnan(cell_size)) {
    set_error("tf_nr_set_coarse: invalid arguments");
    return TF_ERR_ARG;
  }
  TF_CHECK(hipStreamSynchronize(h->st));
  h->coarse_cell = cell_size < 0.0 ? -1.0 : cell_size;
  tiled_release(&h->ws);
  coarse_release(h->coarse);
  return TF_OK;
}

tf_status tf_nr_set_coarse_lag(tf_handle* h, int32_t lag) {
  if (!h) {
    set_error("tf_nr_set_coarse_lag: invalid handle");
    return TF_ERR_ARG;
  }
  h->coarse_lag = lag != 0;
  return TF_OK;
}

tf_status tf_nr_set_warm_start(tf_handle* h, int32_t enable) {
  if (!h) {
    set_error("tf_nr_set_warm_start: invalid handle");
    return TF_ERR_ARG;
  }
  h->warm_start = enable != 0;
  reset_history(h);
  return TF_OK;
}

tf_status tf_nr_set_forcing_safeguard(tf_handle* h, int32_t enable) {
  if (!h) {
    set_error("tf_nr_set_forcing_safeguard: invalid handle");
    return TF_ERR_ARG;
  }
  h->forcing_safeguard = enable != 0;
  return TF_OK;
}

tf_status tf_nr_set_coarse_bands(tf_handle* h, int32_t z_bands) {
  if (!h || z_bands < 1 || z_bands > 64) {
    set_error("tf_nr_set_coarse_bands: invalid arguments");
    return TF_ERR_ARG;
  }
  TF_CHECK(hipStreamSynchronize(h->st));
  h->coarse_bands = z_bands;
  tiled_release(&h->ws);
  coarse_release(h->coarse);
  return TF_OK;
}

tf_status tf_nr_coarse_info(tf_handle* h, int32_t* aggregates, int32_t* active) {
  if (!h) {
    set_error("tf_nr_coarse_info: invalid handle");
    return TF_ERR_ARG;
  }
  if (aggregates) *aggregates = h->coarse.built ? h->coarse.n_agg : 0;
  if (active) *active = h->coarse.active ? 1 : 0;
  return TF_OK;
}

tf_status tf_nr_set_graphs(tf_handle* h, int32_t enable) {
  if (!h) {
    set_error("tf_nr_set_graphs: invalid handle");
    return TF_ERR_ARG;
  }
  TF_CHECK(hipStreamSynchronize(h->st));
  h->ws.graphs_on = enable ? 1 : 0;
  if (!enable) tiled_release(&h->ws);
  return TF_OK;
}

tf_status tf_nr_pattern_info(tf_handle* h, int64_t* n_blocks, int64_t* n_slots) {
  if (!h) {
    set_error("tf_nr_pattern_info: invalid handle");
    return TF_ERR_ARG;
  }
  if (n_blocks) *n_blocks = (int64_t)h->h_nbr.size();
  if (n_slots) *n_slots = h->pat.n_slots;
  return TF_OK;
}

tf_status tf_nr_destroy(tf_handle* h) {
  if (!h) return TF_OK;
  if (h->st) (void)hipStreamSynchronize(h->st);
  tiled_release(&h->ws);
  free_bound(h);
  coarse_release(h->coarse);
  for (auto& e : h->prof.ev)
    if (e) (void)hipEventDestroy(e);
  for (void* p : h->allocs) (void)hipFree(p);
  if (h->adj_ckpt) (void)hipFree(h->adj_ckpt);
  if (h->d_tile_slice) (void)hipFree(h->d_tile_slice);
  if (h->d_tile_tissue) (void)hipFree(h->d_tile_tissue);
  if (h->d_tissue_tile) (void)hipFree(h->d_tissue_tile);
  if (h->ws.tpart) (void)hipFree(h->ws.tpart);
  if (h->ws.host_tstate) (void)hipHostFree(h->ws.host_tstate);
  if (h->ws.host_tnorm) (void)hipHostFree(h->ws.host_tnorm);
  if (h->ws.host_gnorm) (void)hipHostFree(h->ws.host_gnorm);
  if (h->ws.host_rscale) (void)hipHostFree(h->ws.host_rscale);
  for (auto& e : h->ev)
    if (e) (void)hipEventDestroy(e);
  if (h->st) (void)hipStreamDestroy(h->st);
  delete h;
  return TF_OK;
}

}  // extern "C"
