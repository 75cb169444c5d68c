// Node-node packing on the device: the pre-step detection of
// Simulation::runOneStep (Simulation.cpp:6065-6076), i.e. node masses
// (Simulation::updateNodeMasses), the bounding box, the side-length thresholds
// of Simulation::getAverageSideLength and the couple search of
// Simulation::detectPacingNodes, and the contact slots that extend the SELL
// pattern of the Jacobian by the couples' blocks.  The couple force and its
// Jacobian blocks (pack_couple, tf_internal.h) enter the residual and the
// assembly in tf_assemble.hip.
//
// The reference tests every pair of nodes (O(N^2), Simulation.cpp:8035-8078).
// Here every node that may pack is hashed into a uniform grid whose cell is the
// largest detection threshold, and searches only the 27 cells around it.  Every
// node collects its own partners (both the couples where it is the master,
// id0, and those where it is id1) in ascending order, which is the order in
// which the reference's couple list touches that node; the result is therefore
// independent of the hash order and bitwise reproducible.
#include "tf_internal.h"

#pragma clang fp contract(off)

namespace tf {
namespace {

constexpr int kPB = 256;
constexpr int kChunk = 2048;  // prisms per partial sum of the thresholds
constexpr int kNQ = 52;       // 50 grid cells, column sum, column count

int blocks_for(int64_t n) {
  int64_t g = (n + kPB - 1) / kPB;
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, 65535 * 4));
}

// Prism::calculateReferenceVolume (Prism.cpp:888-922) through
// ShapeBase::calculateVolumeForInputShapeStructure (ShapeBase.cpp:3364-3398).
__global__ __launch_bounds__(kPB) void ref_volume_kernel(int32_t n_prisms, const double* refpos,
                                                         double* vol) {
  const int faces[8][3] = {{0, 1, 2}, {3, 4, 5}, {0, 1, 3}, {1, 3, 4},
                           {1, 2, 4}, {2, 4, 5}, {0, 2, 3}, {2, 3, 5}};
  for (int e = blockIdx.x * kPB + threadIdx.x; e < n_prisms; e += gridDim.x * kPB) {
    const double* P = refpos + 18 * (size_t)e;
    double c[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 3; ++j) c[j] += P[3 * i + j];
    for (int j = 0; j < 3; ++j) c[j] /= 6;
    double total = 0;
    for (int f = 0; f < 8; ++f) {
      const double* a = P + 3 * faces[f][0];
      const double* b = P + 3 * faces[f][1];
      const double* d = P + 3 * faces[f][2];
      double v1[3], v2[3], vm[3], n[3];
      for (int j = 0; j < 3; ++j) {
        v1[j] = b[j] - a[j];
        v2[j] = d[j] - a[j];
        vm[j] = c[j] - a[j];
      }
      n[0] = v1[1] * v2[2] - v1[2] * v2[1];
      n[1] = v1[2] * v2[0] - v1[0] * v2[2];
      n[2] = v1[0] * v2[1] - v1[1] * v2[0];
      const double nb = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
      const double area = nb / 2;
      double height = (vm[0] * n[0] + vm[1] * n[1] + vm[2] * n[2]) / nb;
      if (height < 0) height *= (-1.0);
      total += 1.0 / 3.0 * (height * area);
    }
    vol[e] = total;
  }
}

// Node::mass: VolumePerNode = det(F_g) V_ref / 6 of every prism of the node,
// added in prism order (ShapeBase.cpp:1497-1499, 3749-3753); det by cofactors
// (ShapeBase.cpp:2264-2271).
__global__ __launch_bounds__(kPB) void node_mass_kernel(int32_t n_nodes, const int32_t* node_ptr,
                                                        const int32_t* node_src, const double* Fg,
                                                        const double* vol, double* mass) {
  for (int R = blockIdx.x * kPB + threadIdx.x; R < n_nodes; R += gridDim.x * kPB) {
    double m = 0.0;
    for (int t = node_ptr[R]; t < node_ptr[R + 1]; ++t) {
      const int e = node_src[t] / 6;
      const double* M = Fg + 9 * (size_t)e;
      double det = 0.0;
      det = M[0] * (M[4] * M[8] - M[5] * M[7]);
      det -= M[1] * (M[3] * M[8] - M[5] * M[6]);
      det += M[2] * (M[3] * M[7] - M[4] * M[6]);
      const double grown = det * vol[e];
      m += grown / 6;
    }
    mass[R] = m;
  }
}

// Simulation::calculateBoundingBox (Simulation.cpp:8388-8425) in one
// workgroup.  Its `else if` tests a node against the maximum only when the
// node does not lower the minimum; for nodes 1.. that changes the maximum only
// when every later node lies below node 0, and then the sequential rule is run
// as written.  Writes par[52..57] = {min x, y, z, max x, y, z}.
__global__ __launch_bounds__(1024) void bbox_kernel(int32_t n_nodes, const double* u,
                                                    int symmetric_y, double* par) {
  __shared__ double lo_s[1024][3], hi_s[1024][3];
  double lo[3] = {100000.0, 100000.0, 100000.0}, hi[3] = {-1e300, -1e300, -1e300};
  for (int k = threadIdx.x; k < n_nodes; k += 1024)
    for (int i = 0; i < 3; ++i) {
      const double v = u[3 * (size_t)k + i];
      lo[i] = fmin(lo[i], v);
      if (k > 0) hi[i] = fmax(hi[i], v);
    }
  for (int i = 0; i < 3; ++i) {
    lo_s[threadIdx.x][i] = lo[i];
    hi_s[threadIdx.x][i] = hi[i];
  }
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int i = 0; i < 3; ++i) {
        lo_s[threadIdx.x][i] = fmin(lo_s[threadIdx.x][i], lo_s[threadIdx.x + w][i]);
        hi_s[threadIdx.x][i] = fmax(hi_s[threadIdx.x][i], hi_s[threadIdx.x + w][i]);
      }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    for (int i = 0; i < 3; ++i) {
      double l = lo_s[0][i], h = hi_s[0][i];
      const double v0 = u[i];
      if (!(h >= v0) || !(v0 < 100000.0)) {
        // the sequential rule of the reference
        l = 100000.0;
        h = -100000.0;
        for (int k = 0; k < n_nodes; ++k) {
          const double v = u[3 * (size_t)k + i];
          if (v < l) l = v;
          else if (v > h) h = v;
        }
      } else {
        h = fmax(h, -100000.0);
      }
      par[52 + i] = l;
      par[55 + i] = h;
    }
    if (symmetric_y) par[53] = (-1.0) * par[56];
  }
}

// Per prism: the apical side-length average, the basal one for basal or
// whole-tissue prisms (Prism.cpp:1329-1393), and the threshold-grid cell of the
// prism centre's relative position in the box (ShapeBase.cpp:2091-2095,
// Simulation.cpp:3231-3240).  Non-columnar prisms take no part (cell -1).
__device__ double side_average(const double* u, const int32_t* c, const int pr[3][2]) {
  double dsum = 0.0;
  int counter = 0;
  for (int i = 0; i < 3; ++i) {
    const double* a = u + 3 * (size_t)c[pr[i][0]];
    const double* b = u + 3 * (size_t)c[pr[i][1]];
    const double dx = a[0] - b[0];
    const double dy = a[1] - b[1];
    const double dz = a[2] - b[2];
    const double d2 = (dx * dx + dy * dy + dz * dz);
    if (d2 > 0) {
      dsum += sqrt(d2);
      counter++;
    }
  }
  if (counter > 0) dsum /= counter;
  return dsum;
}

__global__ __launch_bounds__(kPB) void elem_stats_kernel(int32_t n_prisms, const int32_t* conn,
                                                         const double* u, const int32_t* place,
                                                         const int32_t* type,
                                                         const uint8_t* spans, const double* par,
                                                         int32_t* cell, double* eval,
                                                         int32_t* ecnt) {
  const int ap[3][2] = {{3, 4}, {3, 5}, {4, 5}};
  const int ba[3][2] = {{0, 1}, {0, 2}, {1, 2}};
  const double bx = par[52], by = par[53];
  const double sx = par[55] - par[52], sy = par[56] - par[53];
  for (int e = blockIdx.x * kPB + threadIdx.x; e < n_prisms; e += gridDim.x * kPB) {
    if (type[e] != 0) {
      cell[e] = -1;
      eval[3 * (size_t)e] = eval[3 * (size_t)e + 1] = eval[3 * (size_t)e + 2] = 0.0;
      ecnt[e] = 0;
      continue;
    }
    const int32_t* c = conn + 6 * (size_t)e;
    int bits = 0;
    double cur = side_average(u, c, ap);
    const double apical = cur;
    if (apical > 0) bits |= 1;
    double basal = 0.0;
    if (place[e] == 0 || spans[e]) {
      basal = side_average(u, c, ba);
      cur = basal;
      if (basal > 0) bits |= 2;
    }
    double ctr[3] = {0.0, 0.0, 0.0};
    for (int a = 0; a < 6; ++a)
      for (int j = 0; j < 3; ++j) ctr[j] += u[3 * (size_t)c[a] + j];
    ctr[0] /= 6;
    ctr[1] /= 6;
    const double rx = (ctr[0] - bx) / sx;
    const double ry = (ctr[1] - by) / sy;
    int relX = (int)fmax(fmin(floor(rx), 9.0), 0.0);
    int relY = (int)fmax(fmin(floor(ry / 2.0), 4.0), 0.0);
    cell[e] = 5 * relX + relY;
    eval[3 * (size_t)e] = cur;
    eval[3 * (size_t)e + 1] = apical;
    eval[3 * (size_t)e + 2] = basal;
    ecnt[e] = bits;
  }
}

// Partial sums per chunk of kChunk prisms, in prism order: lane q < 50 sums
// the values of grid cell q (sum, count), lane 50 the column side lengths,
// lane 51 their count.  part[(chunk * kNQ + q) * 2 + {0, 1}].
__global__ __launch_bounds__(64) void stats_partial_kernel(int32_t n_prisms, const int32_t* cell,
                                                           const double* eval,
                                                           const int32_t* ecnt, double* part) {
  const int q = threadIdx.x;
  if (q >= kNQ) return;
  const int beg = blockIdx.x * kChunk, end = min(n_prisms, beg + kChunk);
  double s = 0.0, n = 0.0;
  for (int e = beg; e < end; ++e) {
    const int c = cell[e];
    if (c < 0) continue;
    if (q < 50) {
      if (c == q) {
        s += eval[3 * (size_t)e];
        n += 1.0;
      }
    } else {
      const int bits = ecnt[e];
      if (bits & 1) { s += eval[3 * (size_t)e + 1]; n += 1.0; }
      if (bits & 2) { s += eval[3 * (size_t)e + 2]; n += 1.0; }
    }
  }
  part[((size_t)blockIdx.x * kNQ + q) * 2] = s;
  part[((size_t)blockIdx.x * kNQ + q) * 2 + 1] = n;
}

// Chunk sums in chunk order, then the thresholds of detectPacingNodes
// (Simulation.cpp:8002-8020): par[0] = 0.4 column mean, par[2 + q] the squared
// 0.4 cell mean, par[1] the largest detection threshold.
__global__ __launch_bounds__(64) void stats_final_kernel(int32_t n_chunks, const double* part,
                                                         double* par) {
  __shared__ double g[64];
  const int q = threadIdx.x;
  double s = 0.0, n = 0.0;
  if (q < kNQ)
    for (int c = 0; c < n_chunks; ++c) {
      s += part[((size_t)c * kNQ + q) * 2];
      n += part[((size_t)c * kNQ + q) * 2 + 1];
    }
  g[q] = 0.0;
  if (q < 50) {
    double v = s;
    if (n > 0) v /= n;
    v = 0.4 * v;
    g[q] = v;
    par[2 + q] = v * v;
  } else if (q == 50) {
    par[0] = 0.4 * (s / n);
  }
  __syncthreads();
  if (q == 0) {
    double mx = 0.0;
    for (int i = 0; i < 50; ++i) mx = fmax(mx, g[i]);
    par[1] = mx;
  }
}

__device__ inline bool can_pack(const int32_t* place, const double* mass, int k) {
  return place[k] >= 0 && mass[k] != 0.0;
}

__device__ inline void hash_cell(const double* p, const double* par, double h, int64_t c[3]) {
  for (int i = 0; i < 3; ++i) c[i] = (int64_t)floor((p[i] - par[52 + i]) / h);
}

__device__ inline uint32_t hash_of(int64_t x, int64_t y, int64_t z, uint32_t mask) {
  uint64_t k = (uint64_t)x * 0x9E3779B97F4A7C15ull ^ (uint64_t)y * 0xC2B2AE3D27D4EB4Full ^
               (uint64_t)z * 0x165667B19E3779F9ull;
  k ^= k >> 29;
  return (uint32_t)k & mask;
}

__global__ __launch_bounds__(kPB) void fill_kernel(int32_t* a, int64_t n, int32_t v) {
  for (int64_t i = blockIdx.x * (int64_t)kPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kPB)
    a[i] = v;
}

// Every node that may pack is pushed onto its cell's bucket list.
__global__ __launch_bounds__(kPB) void hash_insert_kernel(int32_t n_nodes, const double* u,
                                                          const int32_t* place,
                                                          const double* mass, const double* par,
                                                          uint32_t mask, int32_t* head,
                                                          int32_t* next) {
  const double h = par[1];
  for (int k = blockIdx.x * kPB + threadIdx.x; k < n_nodes; k += gridDim.x * kPB) {
    if (!(h > 0.0) || !can_pack(place, mass, k)) continue;
    int64_t c[3];
    hash_cell(u + 3 * (size_t)k, par, h, c);
    const uint32_t b = hash_of(c[0], c[1], c[2], mask);
    next[k] = atomicExch(&head[b], k);
  }
}

// The reference's detection threshold of a master node: its cell of the
// 10 x 5 grid, indexed by its relative position in the box without scaling
// (Simulation.cpp:8042-8044), i.e. cell (0, 0) inside the box and cells of
// index 1 on its upper faces.
__device__ inline double master_t2(const double* p, const double* par) {
  const double sx = par[55] - par[52], sy = par[56] - par[53];
  const int gx = (int)fmax(fmin(floor((p[0] - par[52]) / sx), 9.0), 0.0);
  const int gy = (int)fmax(fmin(floor((p[1] - par[53]) / sy), 4.0), 0.0);
  return par[2 + 5 * gx + gy];
}

// One thread per node: its partners, ascending (the couples where it is id1,
// then those where it is id0: the order of the reference's couple list), the
// first cap of them into inc, the count into cnt, the largest count into
// flag[0] and the number of couples into flag[1].
__global__ __launch_bounds__(kPB) void detect_kernel(int32_t n_nodes, const double* u,
                                                     const int32_t* place, const double* mass,
                                                     const double* par, uint32_t mask,
                                                     const int32_t* head, const int32_t* next,
                                                     DevPattern pat, int cap, int32_t* cnt,
                                                     int32_t* inc, int32_t* flag) {
  const double h = par[1];
  for (int k = blockIdx.x * kPB + threadIdx.x; k < n_nodes; k += gridDim.x * kPB) {
    if (!(h > 0.0) || !can_pack(place, mass, k)) {
      cnt[k] = 0;
      continue;
    }
    const double* pk = u + 3 * (size_t)k;
    const double t2k = master_t2(pk, par);
    int64_t c[3];
    hash_cell(pk, par, h, c);
    uint32_t bk[27];
    int nb = 0;
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const uint32_t b = hash_of(c[0] + dx, c[1] + dy, c[2] + dz, mask);
          bool seen = false;
          for (int i = 0; i < nb; ++i) seen |= bk[i] == b;
          if (!seen) bk[nb++] = b;
        }
    int list[kPackLocal];
    int n = 0;
    const int R0 = (k / kSlice), lane = k % kSlice;
    const int64_t rbase = pat.slice_base[R0] + lane;
    const int rl = pat.row_len[k];
    for (int i = 0; i < nb; ++i)
      for (int o = head[bk[i]]; o >= 0; o = next[o]) {
        if (o == k || place[o] != place[k]) continue;
        const int lo = min(k, o), hi = max(k, o);
        const double* pl = u + 3 * (size_t)lo;
        const double* ph = u + 3 * (size_t)hi;
        const double dx = pl[0] - ph[0];
        const double dy = pl[1] - ph[1];
        const double dz = pl[2] - ph[2];
        const double d2 = dx * dx + dy * dy + dz * dz;
        const double t2 = (lo == k) ? t2k : master_t2(pl, par);
        if (!(d2 < t2)) continue;
        // immediate neighbours (a shared prism) never pack: the element row of k
        bool neig = false;
        for (int j = 0; j < rl && !neig; ++j) neig = pat.col[rbase + (int64_t)j * kSlice] == o;
        if (neig) continue;
        if (n < kPackLocal) {
          int j = n;
          while (j > 0 && list[j - 1] > o) {
            list[j] = list[j - 1];
            --j;
          }
          list[j] = o;
        }
        ++n;
      }
    cnt[k] = n;
    for (int e = 0; e < min(n, cap); ++e) inc[(size_t)k * cap + e] = list[e];
    if (n > 0) {
      atomicMax(&flag[0], n);
      int as_master = 0;  // couples (k, o > k)
      for (int e = 0; e < min(n, kPackLocal); ++e) as_master += list[e] > k ? 1 : 0;
      atomicAdd(&flag[1], as_master);
    }
  }
}

// Column and transpose slot of every contact slot: the e-th couple of row R
// sits in column slice_ew + e; its transpose is the slot of R in the partner's
// list.  Unused contact slots point at R (zero block, transposed onto itself).
__global__ __launch_bounds__(kPB) void contact_slots_kernel(int32_t n_nodes, DevPattern pat,
                                                            const int32_t* cnt,
                                                            const int32_t* inc, int cap) {
  const int64_t total = (int64_t)pat.n_slices * kSlice;
  for (int64_t t = blockIdx.x * (int64_t)kPB + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * kPB) {
    const int s = (int)(t / kSlice), lane = (int)(t % kSlice);
    const int ncap = pat.slice_cap[s];
    if (ncap == 0) continue;
    const int R = s * kSlice + lane;
    const int ew = pat.slice_ew[s];
    const int nc = R < n_nodes ? min(cnt[R], cap) : 0;
    for (int e = 0; e < ncap; ++e) {
      const int64_t slot = pat.slice_base[s] + (int64_t)(ew + e) * kSlice + lane;
      if (e < nc) {
        const int o = inc[(size_t)R * cap + e];
        const int no = min(cnt[o], cap);
        int eo = 0;
        while (eo < no && inc[(size_t)o * cap + eo] != R) ++eo;
        const int so = o / kSlice;
        pat.col[slot] = o;
        pat.tslot[slot] =
            (int32_t)(pat.slice_base[so] + (int64_t)(pat.slice_ew[so] + eo) * kSlice + o % kSlice);
      } else {
        pat.col[slot] = min(R, n_nodes - 1);
        pat.tslot[slot] = (int32_t)slot;
      }
    }
  }
}

}  // namespace

hipError_t launch_ref_volume(int32_t n_prisms, const double* refpos, double* vol, hipStream_t st) {
  hipLaunchKernelGGL(ref_volume_kernel, dim3(blocks_for(n_prisms)), dim3(kPB), 0, st, n_prisms,
                     refpos, vol);
  return hipGetLastError();
}

hipError_t launch_node_mass(int32_t n_nodes, const int32_t* node_ptr, const int32_t* node_src,
                            const double* Fg, const double* vol, double* mass, hipStream_t st) {
  hipLaunchKernelGGL(node_mass_kernel, dim3(blocks_for(n_nodes)), dim3(kPB), 0, st, n_nodes,
                     node_ptr, node_src, Fg, vol, mass);
  return hipGetLastError();
}

hipError_t launch_packing_detect(PackDev& p, int32_t n_nodes, int32_t n_prisms,
                                 const int32_t* conn, const double* u, const int32_t* elem_place,
                                 const int32_t* elem_type, const uint8_t* spans,
                                 const DevPattern& pat, hipStream_t st) {
  hipLaunchKernelGGL(bbox_kernel, dim3(1), dim3(1024), 0, st, n_nodes, u, p.symmetric_y, p.par);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(elem_stats_kernel, dim3(blocks_for(n_prisms)), dim3(kPB), 0, st, n_prisms,
                     conn, u, elem_place, elem_type, spans, p.par, p.cell, p.eval, p.ecnt);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(stats_partial_kernel, dim3(p.n_chunks), dim3(64), 0, st, n_prisms, p.cell,
                     p.eval, p.ecnt, p.part);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(stats_final_kernel, dim3(1), dim3(64), 0, st, p.n_chunks, p.part, p.par);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(fill_kernel, dim3(blocks_for(p.hash_size)), dim3(kPB), 0, st, p.head,
                     (int64_t)p.hash_size, -1);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(fill_kernel, dim3(1), dim3(kPB), 0, st, p.flag, (int64_t)2, 0);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  const uint32_t mask = (uint32_t)p.hash_size - 1;
  hipLaunchKernelGGL(hash_insert_kernel, dim3(blocks_for(n_nodes)), dim3(kPB), 0, st, n_nodes, u,
                     p.place, p.mass, p.par, mask, p.head, p.next);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(detect_kernel, dim3(blocks_for(n_nodes)), dim3(kPB), 0, st, n_nodes, u,
                     p.place, p.mass, p.par, mask, p.head, p.next, pat, p.cap, p.cnt, p.inc,
                     p.flag);
  return hipGetLastError();
}

hipError_t launch_contact_slots(const PackDev& p, int32_t n_nodes, const DevPattern& pat,
                                hipStream_t st) {
  hipLaunchKernelGGL(contact_slots_kernel, dim3(blocks_for((int64_t)pat.n_slices * kSlice)),
                     dim3(kPB), 0, st, n_nodes, pat, p.cnt, p.inc, p.cap);
  return hipGetLastError();
}

}  // namespace tf
