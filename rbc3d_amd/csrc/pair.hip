// Real-space part of the Ewald-split matvec on cell surfaces:
//   cell list (HashTable_Build, common/ModHashTable.F90:23-58),
//   pair sums with same-surface masking (AddIntOnRbcs, common/ModIntOnRbcs.F90:47-112),
//   neighbouring-cell lists (NbrRbcList_Insert, common/ModRbcSingInt.F90:341-364),
//   self-patch singular quadrature (RBC_SingInt, common/ModRbcSingInt.F90:29-90),
//   near-singular corrections (RBC_NearSingInt and helpers, :103-310,
//     Spline_FindProjection common/ModSpline.F90:203-269),
//   the periodic linear double-layer term (AddLinearInt, common/ModIntOnRbcs.F90:162-201).
//
// Layout: sources are sorted by cell-list bin (stable, so ascending index inside a
// bin); targets are binned the same way.  The pair kernel gives one workgroup to a
// chunk of up to 256 targets of one bin and streams the sources of the 27
// neighbouring bins through LDS tiles.  Each lane owns one target, accumulates in
// registers and writes once: no atomics, a fixed summation order per target.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <mutex>

#include "internal.h"
#include "surf_core.h"

#define NBR_T 256

// ---------------------------------------------------------------------------
// cell list
__device__ __forceinline__ void hash_index3(const double x[3], const int Nc[3], const double iLbNc[3],
                                            int &i1, int &i2, int &i3) {
  // HashTable_Index (common/ModHashTable.F90:67-91), one rank; 1-based bins
  i1 = imod_d((int)floor(x[0] * iLbNc[0]), Nc[0]) + 1;
  i2 = imod_d((int)floor(x[1] * iLbNc[1]), Nc[1]) + 1;
  i3 = imod_d((int)floor(x[2] * iLbNc[2]), Nc[2]) + 1;
}

struct HashGeom {
  int Nc[3];
  double iLbNc[3];
};

__global__ void k_bins(int n, const double *__restrict__ x, int ldx, HashGeom hg, int *bin, int *idx) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double xi[3] = {x[i], x[i + ldx], x[i + 2 * (size_t)ldx]};
  int i1, i2, i3;
  hash_index3(xi, hg.Nc, hg.iLbNc, i1, i2, i3);
  bin[i] = (i1 - 1) + hg.Nc[0] * ((i2 - 1) + hg.Nc[1] * (i3 - 1));
  idx[i] = i;
}

// bin_start[b] = first sorted position with bin >= b (sorted keys), b = 0..nbins
__global__ void k_bin_start(int n, const int *__restrict__ skey, int nbins, int *bin_start) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b > nbins) return;
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (skey[mid] < b)
      lo = mid + 1;
    else
      hi = mid;
  }
  bin_start[b] = lo;
}

static int radix_bits(int nbins) {
  int bits = 1;
  while ((1 << bits) < nbins) bits++;
  return bits;
}

// Bytes of radix-sort working storage for n (bin, index) pairs (a host-side query).
static int hash_sort_tmp_bytes(int n, int nbins, hipStream_t st, size_t *bytes) {
  *bytes = 0;
  if (n <= 0) return RBC3D_OK;
  HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, *bytes, (const int *)nullptr, (int *)nullptr,
                                             (const int *)nullptr, (int *)nullptr, n, 0,
                                             radix_bits(nbins), st));
  return RBC3D_OK;
}

// Sort n points by bin (stable: ascending index inside a bin) into caller-provided
// buffers, asynchronously on st: d_bin receives the bin of each point in ORIGINAL order,
// d_order the sorted point indices, d_bin_start[0..nbins] the first sorted position of
// each bin.  keys_out and idx_in hold n ints each; tmp holds tmp_bytes (at least
// hash_sort_tmp_bytes).  Allocates nothing.
static int hash_sort_bufs(hipStream_t st, const HashGeom &hg, int nbins, int n, const double *d_x,
                          int ldx, int *d_bin, int *d_order, int *d_bin_start, int *keys_out,
                          int *idx_in, void *tmp, size_t tmp_bytes) {
  if (n > 0) {
    k_bins<<<(n + 255) / 256, 256, 0, st>>>(n, d_x, ldx, hg, d_bin, idx_in);
    HIP_TRY(hipGetLastError());
    size_t need = 0;
    RBC_TRY(hash_sort_tmp_bytes(n, nbins, st, &need));
    if (need > tmp_bytes) {
      rbc_set_error("cell-list sort: working storage too small");
      return RBC3D_EINVAL;
    }
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, need, d_bin, keys_out, idx_in, d_order, n, 0,
                                               radix_bits(nbins), st));
  }
  k_bin_start<<<(nbins + 1 + 255) / 256, 256, 0, st>>>(n, keys_out, nbins, d_bin_start);
  HIP_TRY(hipGetLastError());
  return RBC3D_OK;
}

// As hash_sort_bufs with the working buffers allocated here, then synchronised.
static int hash_sort_grid(rbc3d_ctx *c, const HashGeom &hg, int nbins, int n, const double *d_x,
                          int ldx, int *d_bin, int *d_order, int *d_bin_start) {
  hipStream_t st = c->stream;
  DBuf<int> keys_out, idx_in;
  RBC_TRY(keys_out.ensure(n + 1));
  RBC_TRY(idx_in.ensure(n + 1));
  size_t tmp = 0;
  RBC_TRY(hash_sort_tmp_bytes(n, nbins, st, &tmp));
  RBC_TRY(c->d_cub_tmp.ensure(tmp + 16));
  RBC_TRY(hash_sort_bufs(st, hg, nbins, n, d_x, ldx, d_bin, d_order, d_bin_start, keys_out.p,
                         idx_in.p, c->d_cub_tmp.p, tmp + 16));
  HIP_TRY(hipStreamSynchronize(st));
  keys_out.release();
  idx_in.release();
  return RBC3D_OK;
}

// the reference's cell list (coarse bins)
int rbc_hash_sort(rbc3d_ctx *c, int n, const double *d_x, int ldx, int *d_bin, int *d_order,
                  int *d_bin_start) {
  HashGeom hg;
  for (int d = 0; d < 3; d++) {
    hg.Nc[d] = c->Nc[d];
    hg.iLbNc[d] = c->iLbNc[d];
  }
  return hash_sort_grid(c, hg, c->nbins, n, d_x, ldx, d_bin, d_order, d_bin_start);
}

// the pair kernel's grid (fine bins)
static int fine_sort(rbc3d_ctx *c, int n, const double *d_x, int ldx, int *d_bin, int *d_order,
                     int *d_bin_start) {
  HashGeom hg;
  for (int d = 0; d < 3; d++) {
    hg.Nc[d] = c->Nf[d];
    hg.iLbNc[d] = c->iLbNf[d];
  }
  return hash_sort_grid(c, hg, c->nfbins, n, d_x, ldx, d_bin, d_order, d_bin_start);
}

// hoc/next from the sorted list, with the ghost layers of HashTable_Build
__global__ void k_hoc_next(int nbins, HashGeom hg, const int *__restrict__ bin_start,
                           const int *__restrict__ order, int *hoc, int *next) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbins) return;
  int b1 = b % hg.Nc[0], b2 = (b / hg.Nc[0]) % hg.Nc[1], b3 = b / (hg.Nc[0] * hg.Nc[1]);
  int s = bin_start[b], e = bin_start[b + 1];
  // push-front insertion in index order: head = largest index, next -> next smaller
  int head = e > s ? order[e - 1] + 1 : -1;
  for (int k = s; k < e; k++) next[order[k]] = k > s ? order[k - 1] + 1 : -1;
  int n1 = hg.Nc[0] + 2, n2 = hg.Nc[1] + 2;
  // write the real bin and every ghost image of it: position 0 holds bin Nc, position
  // Nc+1 holds bin 1 (per axis, corners included, as the slab copies leave them)
  int pos[3][3], npos[3];
  int bb[3] = {b1, b2, b3};
  for (int ax = 0; ax < 3; ax++) {
    npos[ax] = 0;
    pos[ax][npos[ax]++] = bb[ax] + 1;
    if (bb[ax] + 1 == hg.Nc[ax]) pos[ax][npos[ax]++] = 0;
    if (bb[ax] == 0) pos[ax][npos[ax]++] = hg.Nc[ax] + 1;
  }
  for (int a3 = 0; a3 < npos[2]; a3++)
    for (int a2 = 0; a2 < npos[1]; a2++)
      for (int a1 = 0; a1 < npos[0]; a1++) hoc[pos[0][a1] + n1 * (pos[1][a2] + n2 * pos[2][a3])] = head;
}

__global__ void k_fill_u64(unsigned long long *p, int n, unsigned long long v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_i(int *p, int n, int v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

extern "C" int rbc3d_hashtable_build(const int Nc[3], const double iLbNc[3], int nPoint,
                                     const double *x, int ldx, int *hoc, int *next) {
  if (!Nc || !iLbNc || nPoint < 0 || (nPoint > 0 && (!x || ldx < nPoint)) || !hoc || !next)
    return RBC3D_EINVAL;
  for (int d = 0; d < 3; d++)
    if (Nc[d] < 1) return RBC3D_EINVAL;
  // the ghost-layer copies of HashTable_Build equal the periodic images only when
  // Nc >= 2 in each direction; Nc = 1 stays valid (the ghosts alias the single bin)
  rbc3d_ctx tmp;
  for (int d = 0; d < 3; d++) {
    tmp.Nc[d] = Nc[d];
    tmp.iLbNc[d] = iLbNc[d];
  }
  tmp.nbins = Nc[0] * Nc[1] * Nc[2];
  HIP_TRY(hipStreamCreateWithFlags(&tmp.stream, hipStreamNonBlocking));
  DBuf<double> dx;
  DBuf<int> bin, order, start, dhoc, dnext;
  int n1 = Nc[0] + 2, n2 = Nc[1] + 2, n3 = Nc[2] + 2, nh = n1 * n2 * n3;
  int r = RBC3D_OK;
  if (!r) r = dx.ensure((size_t)3 * nPoint + 1);
  if (!r) r = bin.ensure(nPoint + 1);
  if (!r) r = order.ensure(nPoint + 1);
  if (!r) r = start.ensure(tmp.nbins + 1);
  if (!r) r = dhoc.ensure(nh);
  if (!r) r = dnext.ensure(nPoint + 1);
  if (!r && nPoint > 0)
    for (int d = 0; d < 3 && !r; d++)
      if (hipMemcpy(dx.p + (size_t)nPoint * d, x + (size_t)ldx * d, sizeof(double) * nPoint,
                    hipMemcpyHostToDevice) != hipSuccess)
        r = RBC3D_EHIP;
  if (!r) r = rbc_hash_sort(&tmp, nPoint, dx.p, nPoint, bin.p, order.p, start.p);
  if (!r) {
    HashGeom hg;
    for (int d = 0; d < 3; d++) {
      hg.Nc[d] = Nc[d];
      hg.iLbNc[d] = iLbNc[d];
    }
    k_fill_i<<<(nh + 255) / 256, 256, 0, tmp.stream>>>(dhoc.p, nh, -1);
    k_hoc_next<<<(tmp.nbins + 255) / 256, 256, 0, tmp.stream>>>(tmp.nbins, hg, start.p, order.p,
                                                                dhoc.p, dnext.p);
    if (hipStreamSynchronize(tmp.stream) != hipSuccess) r = RBC3D_EHIP;
  }
  if (!r && hipMemcpy(hoc, dhoc.p, sizeof(int) * nh, hipMemcpyDeviceToHost) != hipSuccess) r = RBC3D_EHIP;
  if (!r && nPoint > 0 &&
      hipMemcpy(next, dnext.p, sizeof(int) * nPoint, hipMemcpyDeviceToHost) != hipSuccess)
    r = RBC3D_EHIP;
  dx.release();
  bin.release();
  order.release();
  start.release();
  dhoc.release();
  dnext.release();
  tmp.d_cub_tmp.release();
  (void)hipStreamDestroy(tmp.stream);
  tmp.stream = nullptr;
  return r;
}

// ---------------------------------------------------------------------------
// HashTable_Build without allocation: every working buffer is carved from `scratch`, a
// file-scope device array that each GPU holds for the life of the process.  Host code
// never passes the array's name to a kernel (on the host it is only a stand-in symbol):
// its device address is looked up once per device with hipGetSymbolAddress and kept.

static __device__ __align__(256) unsigned char scratch[RBC3D_NOALLOC_SCRATCH_BYTES];

static std::mutex g_scratch_mu;
static unsigned char *g_scratch_ptr[64];

// The calling thread's current device's copy of `scratch`.
static int scratch_address(unsigned char **out) {
  int dev = -1;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) {
    rbc_set_error("no-allocation scratch: device index out of range");
    return RBC3D_EINVAL;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) {
      rbc_set_error("no-allocation scratch: no device address");
      return RBC3D_EHIP;
    }
    g_scratch_ptr[dev] = static_cast<unsigned char *>(p);
  }
  *out = g_scratch_ptr[dev];
  return RBC3D_OK;
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

extern "C" int rbc3d_hashtable_build_noalloc(const int Nc[3], const double iLbNc[3], int nPoint,
                                             const double *d_x, int ldx, int *d_hoc, int *d_next,
                                             void *stream) {
  if (!Nc || !iLbNc || nPoint < 0 || (nPoint > 0 && (!d_x || ldx < nPoint || !d_next)) || !d_hoc) {
    rbc_set_error("rbc3d_hashtable_build_noalloc: invalid arguments");
    return RBC3D_EINVAL;
  }
  for (int d = 0; d < 3; d++)
    if (Nc[d] < 1) {
      rbc_set_error("rbc3d_hashtable_build_noalloc: Nc must be >= 1");
      return RBC3D_EINVAL;
    }
  long long nbins_ll = (long long)Nc[0] * Nc[1] * Nc[2];
  if (nPoint > RBC3D_NOALLOC_MAX_POINTS || nbins_ll > RBC3D_NOALLOC_MAX_BINS) {
    rbc_set_error("rbc3d_hashtable_build_noalloc: more than " +
                  std::to_string(RBC3D_NOALLOC_MAX_POINTS) + " points or " +
                  std::to_string(RBC3D_NOALLOC_MAX_BINS) + " bins (the scratch maximum)");
    return RBC3D_EINVAL;
  }
  int nbins = (int)nbins_ll;
  hipStream_t st = static_cast<hipStream_t>(stream);
  unsigned char *base = nullptr;
  RBC_TRY(scratch_address(&base));
  // carve: bin, order, keys_out, idx_in (n each), bin_start (nbins + 1), sort storage
  size_t nb = align256(sizeof(int) * ((size_t)nPoint + 1));
  size_t off = 0;
  int *bin = reinterpret_cast<int *>(base + off);
  off += nb;
  int *order = reinterpret_cast<int *>(base + off);
  off += nb;
  int *keys_out = reinterpret_cast<int *>(base + off);
  off += nb;
  int *idx_in = reinterpret_cast<int *>(base + off);
  off += nb;
  int *bin_start = reinterpret_cast<int *>(base + off);
  off += align256(sizeof(int) * ((size_t)nbins + 1));
  size_t need = 0;
  RBC_TRY(hash_sort_tmp_bytes(nPoint, nbins, st, &need));
  if (off + need > (size_t)RBC3D_NOALLOC_SCRATCH_BYTES) {
    rbc_set_error("rbc3d_hashtable_build_noalloc: scratch too small for this size");
    return RBC3D_EINVAL;
  }
  HashGeom hg;
  for (int d = 0; d < 3; d++) {
    hg.Nc[d] = Nc[d];
    hg.iLbNc[d] = iLbNc[d];
  }
  RBC_TRY(hash_sort_bufs(st, hg, nbins, nPoint, d_x, ldx, bin, order, bin_start, keys_out, idx_in,
                         base + off, (size_t)RBC3D_NOALLOC_SCRATCH_BYTES - off));
  int nh = (Nc[0] + 2) * (Nc[1] + 2) * (Nc[2] + 2);
  k_fill_i<<<(nh + 255) / 256, 256, 0, st>>>(d_hoc, nh, -1);
  HIP_TRY(hipGetLastError());
  k_hoc_next<<<(nbins + 255) / 256, 256, 0, st>>>(nbins, hg, bin_start, order, d_hoc, d_next);
  HIP_TRY(hipGetLastError());
  return RBC3D_OK;
}

// ---------------------------------------------------------------------------
// source list on cells

// sorted source records, SoA with stride np: x,y,z, f(3), g(3), a3(3), Bcoef  (13)
#define SREC 13

__global__ void k_src_records(int np, const int *__restrict__ order, const double *__restrict__ x,
                              const double *__restrict__ fs, const double *__restrict__ gs,
                              const double *__restrict__ a3, const CellRec *__restrict__ cells,
                              const int *__restrict__ node_cell, const int *__restrict__ node_ll,
                              double *__restrict__ rec, int *__restrict__ rec_meta) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= np) return;
  int p = order[k];
  for (int d = 0; d < 3; d++) {
    rec[k + (size_t)np * d] = x[p + (size_t)np * d];
    rec[k + (size_t)np * (3 + d)] = fs[p + (size_t)np * d];
    rec[k + (size_t)np * (6 + d)] = gs[p + (size_t)np * d];
    rec[k + (size_t)np * (9 + d)] = a3[p + (size_t)np * d];
  }
  rec[k + (size_t)np * 12] = cells[node_cell[p]].Bcoef;
  // sorted-position metadata [3][np]: node, cell, (ilat | ilon << 16)
  rec_meta[k] = p;
  rec_meta[k + (size_t)np] = node_cell[p];
  rec_meta[k + 2 * (size_t)np] = node_ll[p];
}

// slist%f = f dS, slist%g = g dS (SourceList_UpdateDensity, common/ModSourceList.F90:160-187)
__global__ void k_density(int np, const double *__restrict__ src, const double *__restrict__ ds,
                          double *__restrict__ dst) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  double s = ds[p];
  for (int d = 0; d < 3; d++) dst[p + (size_t)np * d] = src[p + (size_t)np * d] * s;
}

// bounding sphere of each cell (centroid of nodes, max node distance)
__global__ void k_cell_sphere(int nrbc, const CellRec *__restrict__ cells, const double *__restrict__ x,
                              int np, double *sph) {
  int ci = blockIdx.x;
  if (ci >= nrbc) return;
  const CellRec cr = cells[ci];
  int npt = cr.nlat * cr.nlon;
  __shared__ double red[4][256];
  double s[3] = {0, 0, 0};
  for (int t = threadIdx.x; t < npt; t += blockDim.x)
    for (int d = 0; d < 3; d++) s[d] += x[cr.off + t + (size_t)np * d];
  for (int d = 0; d < 3; d++) red[d][threadIdx.x] = s[d];
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int d = 0; d < 3; d++) red[d][threadIdx.x] += red[d][threadIdx.x + w];
    __syncthreads();
  }
  double cen[3] = {red[0][0] / npt, red[1][0] / npt, red[2][0] / npt};
  __syncthreads();
  double rmax = 0;
  for (int t = threadIdx.x; t < npt; t += blockDim.x) {
    double r2 = 0;
    for (int d = 0; d < 3; d++) {
      double dd = x[cr.off + t + (size_t)np * d] - cen[d];
      r2 += dd * dd;
    }
    rmax = fmax(rmax, sqrt(r2));
  }
  red[3][threadIdx.x] = rmax;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[3][threadIdx.x] = fmax(red[3][threadIdx.x], red[3][threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    sph[4 * ci] = cen[0];
    sph[4 * ci + 1] = cen[1];
    sph[4 * ci + 2] = cen[2];
    sph[4 * ci + 3] = red[3][0] * (1 + 1e-12) + 1e-12;
  }
}

// bounding spheres of each cell's consecutive 64-node chunks: one wavefront per chunk
__global__ void k_chunk_sphere(int nrbc, const CellRec *__restrict__ cells,
                               const double *__restrict__ x, int np, double *sph) {
  const int ci = blockIdx.x;
  if (ci >= nrbc) return;
  const CellRec cr = cells[ci];
  const int npt = cr.nlat * cr.nlon, nch = (npt + 63) / 64;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int k = wv; k < nch; k += nw) {
    const int q = 64 * k + lane, cnt = min(64, npt - 64 * k);
    double p[3] = {0, 0, 0};
    if (q < npt)
      for (int d = 0; d < 3; d++) p[d] = x[cr.off + q + (size_t)np * d];
    double cen[3];
    for (int d = 0; d < 3; d++) {
      double v = p[d];
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
      cen[d] = v / cnt;
    }
    double r2 = 0;
    if (q < npt)
      for (int d = 0; d < 3; d++) r2 += (p[d] - cen[d]) * (p[d] - cen[d]);
    for (int o = 32; o > 0; o >>= 1) r2 = fmax(r2, __shfl_xor(r2, o, 64));
    if (lane == 0) {
      double *s = sph + 4 * (size_t)(cr.choff + k);
      s[0] = cen[0];
      s[1] = cen[1];
      s[2] = cen[2];
      s[3] = sqrt(r2) * (1 + 1e-12) + 1e-12;
    }
  }
}

// coarse cell-list coordinates of every node, HashTable_Index (common/ModHashTable.F90:67-91)
__global__ void k_coarse_coords(int np, const double *__restrict__ x, HashGeom hg, int *cc) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  double xp[3] = {x[p], x[p + (size_t)np], x[p + 2 * (size_t)np]};
  int i1, i2, i3;
  hash_index3(xp, hg.Nc, hg.iLbNc, i1, i2, i3);
  cc[p] = (i1 - 1) | ((i2 - 1) << 10) | ((i3 - 1) << 20);
}

int rbc_build_source_list(rbc3d_ctx *c) {
  int np = c->np;
  RBC_TRY(c->d_src_bin.ensure(np + 1));
  RBC_TRY(c->d_src_order.ensure(np + 1));
  RBC_TRY(c->d_bin_start.ensure(c->nbins + 1));
  RBC_TRY(rbc_hash_sort(c, np, c->d_x.p, np, c->d_src_bin.p, c->d_src_order.p, c->d_bin_start.p));
  RBC_TRY(c->d_tmp_f.ensure(np + 1));
  RBC_TRY(c->d_fbin_start.ensure(c->nfbins + 1));
  RBC_TRY(fine_sort(c, np, c->d_x.p, np, c->d_tmp_f.p, c->d_src_order.p, c->d_fbin_start.p));
  RBC_TRY(c->d_src_rec.ensure((size_t)SREC * np + 1));
  RBC_TRY(c->d_cell_sph.ensure((size_t)4 * c->nrbc + 4));
  k_cell_sphere<<<c->nrbc, 256, 0, c->stream>>>(c->nrbc, c->d_cells.p, c->d_x.p, np, c->d_cell_sph.p);
  HIP_TRY(hipGetLastError());
  RBC_TRY(c->d_src_cc.ensure(np + 1));
  {
    HashGeom hg;
    for (int d = 0; d < 3; d++) {
      hg.Nc[d] = c->Nc[d];
      hg.iLbNc[d] = c->iLbNc[d];
    }
    k_coarse_coords<<<(np + 255) / 256, 256, 0, c->stream>>>(np, c->d_x.p, hg, c->d_src_cc.p);
    HIP_TRY(hipGetLastError());
  }
  RBC_TRY(c->d_chunk_sph.ensure((size_t)4 * c->nchunks + 4));
  k_chunk_sphere<<<c->nrbc, 256, 0, c->stream>>>(c->nrbc, c->d_cells.p, c->d_x.p, np,
                                                 c->d_chunk_sph.p);
  HIP_TRY(hipGetLastError());
  return rbc_update_density(c, 0);
}

int rbc_update_density(rbc3d_ctx *c, int which) {
  int np = c->np;
  hipStream_t st = c->stream;
  if (which == 1 || which == 2) {
    RBC_TRY(rbc_sh_build_splines(c, which));
    k_density<<<(np + 255) / 256, 256, 0, st>>>(np, which == 1 ? c->d_f.p : c->d_g.p, c->d_ds.p,
                                                which == 1 ? c->d_fs.p : c->d_gs.p);
    HIP_TRY(hipGetLastError());
  }
  RBC_TRY(c->d_work.ensure(3 * (size_t)np + 1));
  k_src_records<<<(np + 255) / 256, 256, 0, st>>>(np, c->d_src_order.p, c->d_x.p, c->d_fs.p,
                                                  c->d_gs.p, c->d_a3.p, c->d_cells.p,
                                                  c->d_node_cell.p, c->d_node_ll.p, c->d_src_rec.p,
                                                  c->d_work.p);
  HIP_TRY(hipGetLastError());
  return RBC3D_OK;
}

// ---------------------------------------------------------------------------
// pair kernel
//
// Targets and sources are binned on a grid twice as fine as the reference's cell
// list (width >= rc/2).  With one rank the reference's visited set contains every
// minimum-image pair with rr <= rc (its bins are at least rc wide, or there are
// three of them and all are visited), so the pair set is {rr <= rc} and any
// conservative enumeration reproduces it: here the source bins within +-R of the
// target bin whose box distance does not exceed rc.  A workgroup is one wavefront
// owning up to 64 targets of one fine bin; sources stream through LDS in tiles.

#define PT 64    // targets per workgroup, one per lane
#define PW 4     // wavefronts per workgroup, each taking every PW-th source of a tile
#define PS 256   // sources per LDS tile

struct PairArgs {
  double Lb[3], iLb[3], iLbNf[3];
  int Nf[3];
  int img;  // 1: periodic images from the visited bins (every dimension has >= 2R+3 bins)
  double rc, rc2hi, rc2lo, reps2, tabscale, c1, c2;
  int np, nT, nfbins, nrbc, nrows;
  const int *rows;       // x-rows of source bins: packed (alo, ahi, o2, o3)
  const double *rec;     // sources sorted by fine bin (SREC x np)
  const int *src_node;   // sorted position -> node, then cell, then (ilat | ilon << 16)
  const int *bin_start;  // sources per fine bin
  const int *node_cell, *node_ll;
  const CellRec *cells;
  const double *sl1, *sl2, *dl1;
  const double *const *mask;  // per grid type same-surface mask tables
  const double *tx, *tA;
  const int *tindx, *tact, *torder, *tbin_start, *chunk_off;
  double *vpair;
  unsigned long long *stat;  // [0] in-cutoff pairs, [1] pair-set checksum
};

__global__ void k_chunks(int nbins, const int *__restrict__ tbin_start, int *chunk_cnt) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b > nbins) return;
  chunk_cnt[b] = b < nbins ? (tbin_start[b + 1] - tbin_start[b] + PT - 1) / PT : 0;
}

// SL: single-layer term present (c1 != 0); DL: double-layer term present (c2 != 0).
// The workgroup's four wavefronts hold the same 64 targets and split every source
// tile between them; their partial sums are added in wavefront order at the end.
// Periodic images: the reference takes the minimum image of every candidate,
// d - NINT(d / Lb) Lb.  When every dimension has at least 2R+3 fine bins, the image of
// an in-cutoff pair is fixed by the bin walk: the source bin's wrap count k_run, the
// target's box index kt (targets may lie outside [0, Lb)) and the source's (zero for a
// source inside the box) give d = fl(xs - xi) + fl((k_run + kt) Lb), the same two
// roundings as the reference's fl(fl(xs - xi) - fl(NINT * Lb)) with NINT = -(k_run +
// kt); a candidate whose image differs is farther than rc either way.  Tiles holding a
// source outside the box take the reference's formula.  In the image path the loader
// keeps only the sources whose image comes within rc (plus a margin) of the chunk's
// bounding sphere, packed in their order: about 30% of the candidates of a dense
// packing pair with some target of the chunk, and the rest would each cost every lane
// a distance test.  The pair set is unchanged.  The tiles take 26 KB of LDS,
// so six workgroups fit a compute unit; the register budget is set to match (six
// wavefronts per SIMD with a few spills measured 10% faster than four without).
// a value every lane of the wavefront holds, moved to scalar registers
__device__ __forceinline__ double wave_uniform(double v) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)u);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(u >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// FILT: the image-path loader drops the sources beyond reach of the chunk (above); it
// pays where bins are dense (a 20% packing: 10% less time) and costs where they are
// sparse (a tube with ~8 sources per fine bin: 6% more), so the host chooses it by the
// mean number of sources per fine bin.
template <bool SL, bool DL, bool FILT>
__global__ void __launch_bounds__(PT *PW) __attribute__((amdgpu_waves_per_eu(6))) k_pair(PairArgs a) {
  constexpr int NP = (SL ? 3 : 0) + (DL ? 6 : 0);
  constexpr int PD = SL ? 3 : 0;  // first DL payload slot
  __shared__ double sx[PS], sy[PS], sz[PS];
  __shared__ double sp[NP > 0 ? NP : 1][PS];
  __shared__ int snode[PS], scell[PS], sll[PS];
  __shared__ double red[PW - 1][3][PT];
  __shared__ int wout[PW], wkeep[PW];  // per wavefront: a source outside the box; sources kept
  static_assert(PS == PT * PW, "one source per thread and tile");
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int nb = a.nfbins, blk = blockIdx.x;
  if (blk >= a.chunk_off[nb]) return;
  int lo = 0, hi = nb;  // chunk_off[lo] <= blk < chunk_off[hi]
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (a.chunk_off[mid] <= blk)
      lo = mid;
    else
      hi = mid;
  }
  const int bin = lo, chunk = blk - a.chunk_off[lo];
  const int t0 = a.tbin_start[bin] + chunk * PT, t1 = a.tbin_start[bin + 1];
  const int k = t0 + lane;
  const bool have = k < t1;
  const int ti = have ? a.torder[k] : 0;
  const bool act = have && a.tact[ti] != 0;
  double xi0 = 0, xi1 = 0, xi2 = 0, cA1 = 0, cA2 = 0;
  int ilon_i = 0, nn_i = 1, cell_i = -1;
  const double *mtab = nullptr;  // the target's row of the same-surface mask table
  const int Nf0 = a.Nf[0], Nf1 = a.Nf[1], Nf2 = a.Nf[2];
  const int b1 = bin % Nf0, b2 = (bin / Nf0) % Nf1, b3 = bin / (Nf0 * Nf1);
  int kt0 = 0, kt1 = 0, kt2 = 0;  // the target's box index per dimension
  if (have) {
    xi0 = a.tx[ti];
    xi1 = a.tx[ti + (size_t)a.nT];
    xi2 = a.tx[ti + 2 * (size_t)a.nT];
    // the binning's own arithmetic (k_bins): unwrapped bin = wrapped bin + kt Nf
    kt0 = ((int)floor(xi0 * a.iLbNf[0]) - b1) / Nf0;
    kt1 = ((int)floor(xi1 * a.iLbNf[1]) - b2) / Nf1;
    kt2 = ((int)floor(xi2 * a.iLbNf[2]) - b3) / Nf2;
    double Ai = a.tA[ti];
    cA1 = a.c1 / Ai;
    cA2 = a.c2 / Ai;
    int surf_i = a.tindx[ti];
    if (surf_i >= 1 && surf_i <= a.nrbc) {
      const CellRec &cr = a.cells[surf_i - 1];
      cell_i = surf_i - 1;
      nn_i = cr.nlon;
      const int ilat_i = a.tindx[ti + a.nT] - 1;
      mtab = a.mask[cr.gtype] + (size_t)ilat_i * cr.nlat * cr.nlon;
      ilon_i = a.tindx[ti + 2 * (size_t)a.nT] - 1;
    }
  }
  const double L0 = a.Lb[0], L1 = a.Lb[1], L2 = a.Lb[2];
  const double iL0 = a.iLb[0], iL1 = a.iLb[1], iL2 = a.iLb[2];
  // The chunk's bounding sphere over its active targets, each taken back into the box
  // (x - kt Lb): in the image path a candidate's d is (xs + k_run Lb) - (xi - kt Lb) up
  // to a few roundings, so a source whose image lies farther than rc + rt (plus a margin
  // far above those roundings) from the centre is beyond rc of every target and the
  // tile loader leaves it out.  Every wavefront holds the same targets and computes the
  // same sphere.
  double ct0 = 0, ct1 = 0, ct2 = 0, cut2 = -1.;
  if (FILT) {
    const double w0 = act ? xi0 - kt0 * L0 : 0., w1 = act ? xi1 - kt1 * L1 : 0.,
                 w2 = act ? xi2 - kt2 * L2 : 0.;
    double s0 = w0, s1 = w1, s2 = w2, n = act ? 1. : 0.;
    for (int o = 32; o > 0; o >>= 1) {
      s0 += __shfl_xor(s0, o, 64);
      s1 += __shfl_xor(s1, o, 64);
      s2 += __shfl_xor(s2, o, 64);
      n += __shfl_xor(n, o, 64);
    }
    if (n > 0.) {
      ct0 = s0 / n;
      ct1 = s1 / n;
      ct2 = s2 / n;
      double r2 = act ? (w0 - ct0) * (w0 - ct0) + (w1 - ct1) * (w1 - ct1) +
                            (w2 - ct2) * (w2 - ct2)
                      : 0.;
      for (int o = 32; o > 0; o >>= 1) r2 = fmax(r2, __shfl_xor(r2, o, 64));
      const double cut = a.rc * (1. + 1e-9) + sqrt(r2) + 1e-9 * (1. + L0 + L1 + L2);
      cut2 = cut * cut;
    }
    ct0 = wave_uniform(ct0);
    ct1 = wave_uniform(ct1);
    ct2 = wave_uniform(ct2);
    cut2 = wave_uniform(cut2);
  }
  double acc0 = 0, acc1 = 0, acc2 = 0;
  // pair statistics: the count, and sum (node + 1) per target (the pair-set checksum
  // sum (ti + 1)(node + 1) is (ti + 1) times it, modulo 2^64)
  unsigned npairs = 0;
  unsigned long long nsum = 0;
  const size_t np = a.np;

  // one candidate: cut-off test on (d0, d1, d2, s2), then the kernels
  auto pair = [&](int j, double d0, double d1, double d2, double s2) {
    if (s2 > a.rc2hi) return;
    // sqrt(s2) <= rc is certain below rc^2 (1 - 1e-12); in the thin shell above, the
    // correctly rounded square root decides, as in the reference
    double rr, ir;
    if (s2 >= a.rc2lo) {
      rr = sqrt(s2);
      if (rr > a.rc) return;
      ir = rcp_nr(rr);
    } else if (s2 >= a.reps2) {
      ir = rsq_nr(s2);
      rr = s2 * ir;
    } else {
      rr = 0.;  // below r_eps the Ewald coefficients are zero
      ir = 0.;
    }
    const int node = snode[j];
    double wm = 1.;
    if (scell[j] == cell_i) {
      const int ll = sll[j];
      const int jl = ll & 0xffff, jn = ll >> 16;
      int dl = jn - ilon_i;
      dl += dl < 0 ? nn_i : 0;
      wm = 1. - mtab[jl * nn_i + dl];
    }
    npairs++;
    nsum += (unsigned)(node + 1);
    if (SL && ir != 0.) {
      double EA, EB;
      ewald_sl_ir(a.sl1, a.sl2, rr, ir, a.tabscale, EA, EB);
      const double f0 = sp[0][j], f1 = sp[1][j], f2 = sp[2][j];
      const double xf = d0 * f0 + d1 * f1 + d2 * f2;
      const double cc = wm * cA1;
      acc0 += cc * (EA * d0 * xf + EB * f0);
      acc1 += cc * (EA * d1 * xf + EB * f1);
      acc2 += cc * (EA * d2 * xf + EB * f2);
    }
    if (DL && ir != 0.) {
      const double EA = ewald_dl_ir(a.dl1, rr, ir, a.tabscale);
      const double xg = d0 * sp[PD][j] + d1 * sp[PD + 1][j] + d2 * sp[PD + 2][j];
      const double xn = d0 * sp[PD + 3][j] + d1 * sp[PD + 4][j] + d2 * sp[PD + 5][j];
      const double cc = wm * cA2 * EA * xg * xn;
      acc0 += cc * d0;
      acc1 += cc * d1;
      acc2 += cc * d2;
    }
  };

  for (int ri = 0; ri < a.nrows; ri++) {
    const unsigned code = (unsigned)a.rows[ri];
    const int alo = (int)(code & 255) - 128, ahi = (int)((code >> 8) & 255) - 128;
    const int o2 = (int)((code >> 16) & 255) - 128, o3 = (int)(code >> 24) - 128;
    const int rowbase = Nf0 * (imod_d(b2 + o2, Nf1) + Nf1 * imod_d(b3 + o3, Nf2));
    const int k2 = b2 + o2 < 0 ? -1 : (b2 + o2 >= Nf1 ? 1 : 0);
    const int k3 = b3 + o3 < 0 ? -1 : (b3 + o3 >= Nf2 ? 1 : 0);
    // the row's x-bins as one or two runs (a range shorter than the row wraps at most
    // once), with the x wrap count of each
    int rlo[2], rhi[2], rk[2] = {0, 0}, nrun = 1;
    if (ahi - alo + 1 >= Nf0) {
      rlo[0] = 0;
      rhi[0] = Nf0 - 1;
    } else {
      const int x0 = b1 + alo, x1 = b1 + ahi;
      if (x0 < 0) {
        rlo[0] = x0 + Nf0, rhi[0] = Nf0 - 1, rlo[1] = 0, rhi[1] = x1, nrun = 2;
        rk[0] = -1;
      } else if (x1 >= Nf0) {
        rlo[0] = x0, rhi[0] = Nf0 - 1, rlo[1] = 0, rhi[1] = x1 - Nf0, nrun = 2;
        rk[1] = 1;
      } else {
        rlo[0] = x0, rhi[0] = x1;
      }
    }
    const double sh1 = (double)(k2 + kt1) * L1, sh2 = (double)(k3 + kt2) * L2;
    const double rsh1 = (double)k2 * L1, rsh2 = (double)k3 * L2;
    for (int ru = 0; ru < nrun; ru++) {
      const double sh0 = (double)(rk[ru] + kt0) * L0, rsh0 = (double)rk[ru] * L0;
      const int s0 = a.bin_start[rowbase + rlo[ru]], s1 = a.bin_start[rowbase + rhi[ru] + 1];
      for (int base = s0; base < s1; base += PS) {
        const int cnt = min(PS, s1 - base);
        // one source per thread: whether it lies outside the box, and (image path)
        // whether its image can come within rc of the chunk's targets
        const int t = threadIdx.x;
        const bool hv = t < cnt;
        const int q = base + t;
        double x0 = 0., x1 = 0., x2 = 0.;
        bool out = false, keep = hv;
        if (hv) {
          x0 = a.rec[q];
          x1 = a.rec[q + np];
          x2 = a.rec[q + 2 * np];
          if (a.img) {
            const double f0 = floor(x0 * a.iLbNf[0]), f1 = floor(x1 * a.iLbNf[1]),
                         f2 = floor(x2 * a.iLbNf[2]);
            out = f0 < 0 || f0 >= Nf0 || f1 < 0 || f1 >= Nf1 || f2 < 0 || f2 >= Nf2;
          }
          if (FILT && a.img) {
            const double e0 = x0 + rsh0 - ct0, e1 = x1 + rsh1 - ct1, e2 = x2 + rsh2 - ct2;
            keep = e0 * e0 + e1 * e1 + e2 * e2 <= cut2;
          }
        }
        const unsigned long long bo = __ballot(out), bk = FILT ? __ballot(keep) : 0ull;
        __syncthreads();  // the previous tile is consumed
        if (lane == 0) {
          wout[wv] = bo != 0ull;
          if (FILT) wkeep[wv] = __popcll(bk);
        }
        __syncthreads();
        int tout = 0, off = 0, nkeep = 0;
        for (int w = 0; w < PW; w++) {
          tout |= wout[w];
          if (FILT) {
            off += w < wv ? wkeep[w] : 0;
            nkeep += wkeep[w];
          }
        }
        // the image path keeps only the sources that can pair, in their order; a tile
        // holding a source outside the box keeps all of them for the reference's formula
        const bool compact = FILT && a.img && !tout;
        const int pos = compact ? off + __popcll(bk & ((1ull << lane) - 1ull)) : t;
        const int ncnt = compact ? nkeep : cnt;
        if (hv && (!compact || keep)) {
          sx[pos] = x0;
          sy[pos] = x1;
          sz[pos] = x2;
          if (SL)
            for (int d = 0; d < 3; d++) sp[d][pos] = a.rec[q + np * (3 + d)];
          if (DL) {
            const double Bj = a.rec[q + np * 12];
            for (int d = 0; d < 3; d++) {
              sp[PD + d][pos] = a.rec[q + np * (6 + d)] * Bj;
              sp[PD + 3 + d][pos] = a.rec[q + np * (9 + d)];
            }
          }
          snode[pos] = a.src_node[q];
          scell[pos] = a.src_node[q + np];
          sll[pos] = a.src_node[q + 2 * np];
        }
        __syncthreads();
        if (!act) continue;
        if (a.img && !tout) {
#pragma unroll 1
          for (int j = wv; j < ncnt; j += PW) {
            double d0, d1, d2, s2;
            {
#pragma clang fp contract(off)
              d0 = (sx[j] - xi0) + sh0;
              d1 = (sy[j] - xi1) + sh1;
              d2 = (sz[j] - xi2) + sh2;
              s2 = d0 * d0 + d1 * d1;
              s2 = s2 + d2 * d2;
            }
            pair(j, d0, d1, d2, s2);
          }
        } else {
#pragma unroll 1
          for (int j = wv; j < cnt; j += PW) {
            // reference distance without FMA contraction (common/ModIntOnRbcs.F90:68-74);
            // rint differs from NINT only at |x/Lb| = 1/2, where |x| = Lb/2 > rc
            double d0, d1, d2, s2;
            {
#pragma clang fp contract(off)
              d0 = sx[j] - xi0;
              d1 = sy[j] - xi1;
              d2 = sz[j] - xi2;
              d0 = d0 - rint(d0 * iL0) * L0;
              d1 = d1 - rint(d1 * iL1) * L1;
              d2 = d2 - rint(d2 * iL2) * L2;
              s2 = d0 * d0 + d1 * d1;
              s2 = s2 + d2 * d2;
            }
            pair(j, d0, d1, d2, s2);
          }
        }
      }
    }
  }
  if (wv > 0) {
    red[wv - 1][0][lane] = acc0;
    red[wv - 1][1][lane] = acc1;
    red[wv - 1][2][lane] = acc2;
  }
  __syncthreads();
  if (wv == 0 && act) {
    for (int w = 0; w < PW - 1; w++) {
      acc0 += red[w][0][lane];
      acc1 += red[w][1][lane];
      acc2 += red[w][2][lane];
    }
    a.vpair[ti] = acc0;
    a.vpair[ti + (size_t)a.nT] = acc1;
    a.vpair[ti + 2 * (size_t)a.nT] = acc2;
  }
  // one atomic per wavefront for the pair statistics
  unsigned long long npl = npairs, csum = (unsigned long long)(ti + 1) * nsum;
  for (int o = 32; o > 0; o >>= 1) {
    npl += __shfl_xor(npl, o, 64);
    csum += __shfl_xor(csum, o, 64);
  }
  if (lane == 0) {
    atomicAdd(&a.stat[0], npl);
    atomicAdd(&a.stat[1], csum);
  }
}

// ---------------------------------------------------------------------------
// Split form of the pair kernel: FP32 arithmetic and accumulation.
//
// Coordinates are taken relative to the centre of the workgroup's target bin in FP64
// (minimum image about that centre) and split into an FP32 head and an FP32 tail, so
// the FP32 difference of two nearby points is as accurate as the FP64 one; the
// kernel values, products and sums are FP32.  Pairs closer than RBC_SPLIT_NEAR rc
// (about 1.5% of the pairs of a dense packing) are evaluated in FP64 into an FP64
// accumulator: their terms are the largest, and the near-singular correction
// subtracts their FP64 values again, so an FP32 error there would not cancel.  Each
// wavefront first tests its 64 sources of a tile against its target with the heads
// alone and keeps the in-cutoff ones as two bit masks (far, near); it then evaluates
// only those, so the loop that evaluates the Stokeslet / stresslet runs as many
// times as the busiest lane has in-cutoff sources instead of once per candidate.  The
// Ewald tables (FP32) are held in LDS; a workgroup takes one target chunk (the
// kernel also strides over chunks when launched with fewer workgroups), and each XCD
// takes one contiguous range of chunks (neighbouring chunks share source tiles in
// that XCD's L2).
#define PS32 (64 * PW)  // sources per tile: 64 per wavefront
#define RUN_MAX 128     // source runs per chunk: two per row of source bins
#define RBC_SPLIT_NEAR 0.25  // pairs with r < RBC_SPLIT_NEAR rc run in FP64

struct Pair32Args {
  double Lb[3], iLb[3], w[3];  // box, 1/box, fine-bin width
  float fLb[3], fiLb[3];
  int Nf[3];
  float rc2, reps2, tabscale, c1, c2;
  double c1d, c2d, rcd;
  float near2;    // (RBC_SPLIT_NEAR rc)^2
  double reps2d;  // r_eps^2 for the FP64 near pairs
  const double *sl1, *sl2, *dl1;  // FP64 tables for the near pairs
  int wrap;  // 1: the box is too small for the minimum image about the bin centre
  int np, nT, nfbins, nrbc, nrows;
  const int *rows;
  const double *rec;
  const int *src_node;
  const int *bin_start;
  const CellRec *cells;
  const float *tab;  // sl1 | sl2 | dl1 (FP32), RBC_NTAB + 1 each
  const double *const *mask;
  const float *const *maskf;
  const double *tx, *tA;
  const int *tindx, *tact, *torder, *tbin_start, *chunk_off;
  double *vpair;
  unsigned long long *stat;
};

// split x (FP64) into head + tail (FP32)
__device__ __forceinline__ void split_f(double x, float &h, float &l) {
  h = (float)x;
  l = (float)(x - (double)h);
}

// Ewald table value at s = r N / rc by linear interpolation from an LDS table; s is
// clamped to the last interval (at r = rc the coefficient is erfc(xi rc) ~ 1e-8 of
// the near field, where the FP64 form returns zero)
__device__ __forceinline__ float tab_interp(const float *t, float st) {
  const int it = min((int)st, RBC_NTAB - 1);
  const float fr = st - (float)it;
  return fmaf(fr, t[it + 1] - t[it], t[it]);
}

// The pass-1 masks of one 32-source half slice.  Positions beyond the tile are
// parked far away (+1e30) and inactive targets at -1e30, so without the wrap no
// bound test is needed; with the wrap (tiny boxes) both are tested explicitly.
template <bool WRAP>
__device__ __forceinline__ void pair_masks(const float4 *__restrict__ sph, int j0, int jn, bool act,
                                           float xh0, float xh1, float xh2, float rc2, float near2,
                                           const float *fLb, const float *fiLb, unsigned &mf,
                                           unsigned &mn) {
  mf = 0;
  mn = 0;
#pragma unroll
  for (int q = 0; q < 32; q++) {
    const float4 p = sph[j0 + q];
    float d0 = p.x - xh0, d1 = p.y - xh1, d2 = p.z - xh2;
    if (WRAP) {
      d0 -= rintf(d0 * fiLb[0]) * fLb[0];
      d1 -= rintf(d1 * fiLb[1]) * fLb[1];
      d2 -= rintf(d2 * fiLb[2]) * fLb[2];
    }
    const float s2 = fmaf(d2, d2, fmaf(d1, d1, d0 * d0));
    bool in = s2 <= rc2;
    if (WRAP) in = in && act && q < jn;
    mf |= (in && s2 >= near2) ? (1u << q) : 0u;
    mn |= (in && s2 < near2) ? (1u << q) : 0u;
  }
}

template <bool SL, bool DL, bool WRAP>
__global__ void __launch_bounds__(PT *PW) k_pair32(Pair32Args a) {
  constexpr int NTB = (SL ? 2 : 0) + (DL ? 1 : 0);
  constexpr int TDL = SL ? 2 : 0;  // LDS slot of the DL table
  __shared__ float tab[NTB > 0 ? NTB : 1][RBC_NTAB + 1];
  // head of the relative position (w: source cell as int bits) and its tail
  // (w: ilat | ilon << 16 as int bits)
  __shared__ float4 spos[2][PS32];
  float4 *sph = spos[0], *spl = spos[1];
  __shared__ float4 sfv[SL ? PS32 : 1];   // f dS
  __shared__ float4 sgn[DL ? 2 * PS32 : 1];  // (g dS Bcoef).xyz, a3.x | a3.yz
  // the wavefronts' partial sums reuse the position buffers once a chunk's tiles are done
  static_assert(sizeof(double) * (PW - 1) * 3 * PT <= sizeof(spos), "red in spos");
  double(*red)[3][PT] = reinterpret_cast<double(*)[3][PT]>(&spos[0][0]);
  __shared__ int run_start[RUN_MAX], run_cnt[RUN_MAX], run_cum[RUN_MAX + 1];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int nb = a.nfbins;
  const int nchunk = a.chunk_off[nb];
  const int xcd = (int)(blockIdx.x % RBC_NXCD), nper = (int)(gridDim.x / RBC_NXCD);
  const int per = (nchunk + RBC_NXCD - 1) / RBC_NXCD;
  const int c_lo = xcd * per, c_hi = min(nchunk, c_lo + per);
  if (c_lo + (int)(blockIdx.x / RBC_NXCD) >= c_hi) return;  // no chunk for this workgroup
  for (int t = threadIdx.x; t < NTB * (RBC_NTAB + 1); t += PT * PW) {
    const int q = t / (RBC_NTAB + 1), e = t - q * (RBC_NTAB + 1);
    const int g = SL ? q : 2;  // global table: sl1, sl2, dl1
    tab[q][e] = a.tab[(size_t)g * (RBC_NTAB + 1) + e];
  }
  const int Nf0 = a.Nf[0], Nf1 = a.Nf[1];
  const size_t np = a.np;
  const double tsd = (double)RBC_NTAB / a.rcd;
  unsigned long long npairs = 0;
  for (int blk = c_lo + (int)(blockIdx.x / RBC_NXCD); blk < c_hi; blk += nper) {
    int lo = 0, hi = nb;  // chunk_off[lo] <= blk < chunk_off[hi]
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (a.chunk_off[mid] <= blk)
        lo = mid;
      else
        hi = mid;
    }
    const int bin = lo, chunk = blk - a.chunk_off[lo];
    const int b1 = bin % Nf0, b2 = (bin / Nf0) % Nf1, b3 = bin / (Nf0 * Nf1);
    const double ctr0 = (b1 + 0.5) * a.w[0], ctr1 = (b2 + 0.5) * a.w[1], ctr2 = (b3 + 0.5) * a.w[2];
    const int t0 = a.tbin_start[bin] + chunk * PT, t1 = a.tbin_start[bin + 1];
    const int k = t0 + lane;
    const bool have = k < t1;
    const int ti = have ? a.torder[k] : 0;
    const bool act = have && a.tact[ti] != 0;
    float xh0 = -1e30f, xh1 = -1e30f, xh2 = -1e30f, xl0 = 0, xl1 = 0, xl2 = 0, cA1 = 0, cA2 = 0;
    double cA1d = 0, cA2d = 0;
    int ilat_i = 0, ilon_i = 0, nl_i = 1, nn_i = 1, cell_i = -1;
    const float *mtab = nullptr;
    const double *mtabd = nullptr;
    if (act) {
      double r0 = a.tx[ti] - ctr0, r1 = a.tx[ti + (size_t)a.nT] - ctr1,
             r2 = a.tx[ti + 2 * (size_t)a.nT] - ctr2;
      r0 -= rint(r0 * a.iLb[0]) * a.Lb[0];
      r1 -= rint(r1 * a.iLb[1]) * a.Lb[1];
      r2 -= rint(r2 * a.iLb[2]) * a.Lb[2];
      split_f(r0, xh0, xl0);
      split_f(r1, xh1, xl1);
      split_f(r2, xh2, xl2);
      const double Ai = a.tA[ti];
      cA1d = a.c1d / Ai;
      cA2d = a.c2d / Ai;
      cA1 = (float)cA1d;
      cA2 = (float)cA2d;
      const int surf_i = a.tindx[ti];
      if (surf_i >= 1 && surf_i <= a.nrbc) {
        const CellRec &cr = a.cells[surf_i - 1];
        cell_i = surf_i - 1;
        nl_i = cr.nlat;
        nn_i = cr.nlon;
        mtab = a.maskf[cr.gtype];
        mtabd = a.mask[cr.gtype];
        ilat_i = a.tindx[ti + a.nT] - 1;
        ilon_i = a.tindx[ti + 2 * (size_t)a.nT] - 1;
      }
    }
    float acc0 = 0, acc1 = 0, acc2 = 0;
    double accd0 = 0, accd1 = 0, accd2 = 0;
    // The chunk's source runs (the x-row ranges of the sorted sources, in the order the
    // FP64 kernel visits them) are packed back to back into full tiles: every source
    // is stored relative to the bin centre, so one tile may hold several runs, and
    // sparse bins do not leave most of a tile (and three of the four wavefronts) idle.
    __syncthreads();  // the previous chunk is done with the run table
    {
      const int t = threadIdx.x;
      int st = 0, n = 0;
      if (t < 2 * a.nrows) {
        const int ri = t >> 1, ru = t & 1;
        const unsigned code = (unsigned)a.rows[ri];
        const int alo = (int)(code & 255) - 128, ahi = (int)((code >> 8) & 255) - 128;
        const int o2 = (int)((code >> 16) & 255) - 128, o3 = (int)(code >> 24) - 128;
        const int rowbase = Nf0 * (imod_d(b2 + o2, Nf1) + Nf1 * imod_d(b3 + o3, a.Nf[2]));
        int rlo[2], rhi[2], nrun = 1;
        if (ahi - alo + 1 >= Nf0) {
          rlo[0] = 0;
          rhi[0] = Nf0 - 1;
        } else {
          const int x0 = b1 + alo, x1 = b1 + ahi;
          if (x0 < 0) {
            rlo[0] = x0 + Nf0, rhi[0] = Nf0 - 1, rlo[1] = 0, rhi[1] = x1, nrun = 2;
          } else if (x1 >= Nf0) {
            rlo[0] = x0, rhi[0] = Nf0 - 1, rlo[1] = 0, rhi[1] = x1 - Nf0, nrun = 2;
          } else {
            rlo[0] = x0, rhi[0] = x1;
          }
        }
        if (ru < nrun) {
          st = a.bin_start[rowbase + rlo[ru]];
          n = a.bin_start[rowbase + rhi[ru] + 1] - st;
        }
      }
      if (t < RUN_MAX) {
        run_start[t] = st;
        run_cnt[t] = n;
      }
    }
    __syncthreads();
    if (wv == 0) {  // exclusive prefix of the run lengths, two runs per lane
      const int c0 = run_cnt[2 * lane], c1 = run_cnt[2 * lane + 1];
      int incl = c0 + c1;
      for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(incl, o, 64);
        if (lane >= o) incl += u;
      }
      run_cum[2 * lane] = incl - c0 - c1;
      run_cum[2 * lane + 1] = incl - c1;
      if (lane == 63) run_cum[RUN_MAX] = incl;
    }
    __syncthreads();
    const int total = run_cum[RUN_MAX];
    {
      {
        for (int base = 0; base < total; base += PS32) {
          const int cnt = min(PS32, total - base);
          __syncthreads();
          {
            const int t = threadIdx.x;  // one source per thread (PS32 = threads)
            if (t < cnt) {
              const int g = base + t;
              int rl = 0, rh = RUN_MAX;  // the run holding g: run_cum[rl] <= g < run_cum[rl + 1]
              while (rh - rl > 1) {
                const int mid = (rl + rh) >> 1;
                if (run_cum[mid] <= g)
                  rl = mid;
                else
                  rh = mid;
              }
              const int q = run_start[rl] + (g - run_cum[rl]);
              double r0 = a.rec[q] - ctr0, r1 = a.rec[q + np] - ctr1, r2 = a.rec[q + 2 * np] - ctr2;
              r0 -= rint(r0 * a.iLb[0]) * a.Lb[0];
              r1 -= rint(r1 * a.iLb[1]) * a.Lb[1];
              r2 -= rint(r2 * a.iLb[2]) * a.Lb[2];
              float4 h, l;
              split_f(r0, h.x, l.x);
              split_f(r1, h.y, l.y);
              split_f(r2, h.z, l.z);
              h.w = __int_as_float(a.src_node[q + np]);
              l.w = __int_as_float(a.src_node[q + 2 * np]);
              sph[t] = h;
              spl[t] = l;
              if (SL)
                sfv[t] = make_float4((float)a.rec[q + np * 3], (float)a.rec[q + np * 4],
                                     (float)a.rec[q + np * 5], 0.f);
              if (DL) {
                const double Bj = a.rec[q + np * 12];
                sgn[2 * t] = make_float4((float)(a.rec[q + np * 6] * Bj),
                                         (float)(a.rec[q + np * 7] * Bj),
                                         (float)(a.rec[q + np * 8] * Bj), (float)a.rec[q + np * 9]);
                sgn[2 * t + 1] = make_float4((float)a.rec[q + np * 10], (float)a.rec[q + np * 11],
                                             0.f, 0.f);
              }
            } else {
              sph[t] = make_float4(1e30f, 1e30f, 1e30f, __int_as_float(-2));
            }
          }
          __syncthreads();
          const int jb = wv * 64;
          for (int half = 0; half < 2; half++) {
            const int j0 = jb + 32 * half;
            unsigned mf, mn;
            pair_masks<WRAP>(sph, j0, cnt - j0, act, xh0, xh1, xh2, a.rc2, a.near2, a.fLb, a.fiLb,
                             mf, mn);
            npairs += (unsigned)(__popc(mf) + __popc(mn));
            // far pairs: FP32
            while (mf) {
              const int j = j0 + __builtin_ctz(mf);
              mf &= mf - 1;
              const float4 ph = sph[j], pl = spl[j];
              float d0 = (ph.x - xh0) + (pl.x - xl0), d1 = (ph.y - xh1) + (pl.y - xl1),
                    d2 = (ph.z - xh2) + (pl.z - xl2);
              if (WRAP) {
                d0 -= rintf(d0 * a.fiLb[0]) * a.fLb[0];
                d1 -= rintf(d1 * a.fiLb[1]) * a.fLb[1];
                d2 -= rintf(d2 * a.fiLb[2]) * a.fLb[2];
              }
              const float s2 = fmaf(d2, d2, fmaf(d1, d1, d0 * d0));
              // 1/r from the hardware estimate and one Newton step
              float ir = __builtin_amdgcn_rsqf(s2);
              ir = ir * fmaf(-0.5f * s2 * ir, ir, 1.5f);
              const float st = s2 * ir * a.tabscale;
              float wm = 1.f;
              if (__float_as_int(ph.w) == cell_i) {
                const int ll = __float_as_int(pl.w);
                int dl = (ll >> 16) - ilon_i;
                dl += dl < 0 ? nn_i : 0;
                wm = 1.f - mtab[((size_t)ilat_i * nl_i + (ll & 0xffff)) * nn_i + dl];
              }
              const float ir2 = ir * ir;
              if (SL) {
                const float c1v = tab_interp(tab[0], st), c2v = tab_interp(tab[1], st);
                const float EA = fmaf(c1v * ir, ir2, c2v * ir2), EB = fmaf(c1v, ir, -c2v);
                const float4 f = sfv[j];
                const float xf = fmaf(d2, f.z, fmaf(d1, f.y, d0 * f.x));
                const float cc = wm * cA1, ax = EA * xf;
                acc0 = fmaf(cc, fmaf(ax, d0, EB * f.x), acc0);
                acc1 = fmaf(cc, fmaf(ax, d1, EB * f.y), acc1);
                acc2 = fmaf(cc, fmaf(ax, d2, EB * f.z), acc2);
              }
              if (DL) {
                const float EA = tab_interp(tab[TDL], st) * (ir2 * ir2 * ir);
                const float4 g = sgn[2 * j], n2 = sgn[2 * j + 1];
                const float xg = fmaf(d2, g.z, fmaf(d1, g.y, d0 * g.x));
                const float xn = fmaf(d2, n2.y, fmaf(d1, n2.x, d0 * g.w));
                const float cc = wm * cA2 * EA * xg * xn;
                acc0 = fmaf(cc, d0, acc0);
                acc1 = fmaf(cc, d1, acc1);
                acc2 = fmaf(cc, d2, acc2);
              }
            }
            // near pairs: FP64 (heads and tails recombined exactly), FP64 tables
            while (mn) {
              const int j = j0 + __builtin_ctz(mn);
              mn &= mn - 1;
              const float4 ph = sph[j], pl = spl[j];
              double d0 = ((double)ph.x - (double)xh0) + ((double)pl.x - (double)xl0);
              double d1 = ((double)ph.y - (double)xh1) + ((double)pl.y - (double)xl1);
              double d2 = ((double)ph.z - (double)xh2) + ((double)pl.z - (double)xl2);
              if (WRAP) {
                d0 -= rint(d0 * a.iLb[0]) * a.Lb[0];
                d1 -= rint(d1 * a.iLb[1]) * a.Lb[1];
                d2 -= rint(d2 * a.iLb[2]) * a.Lb[2];
              }
              const double s2 = d0 * d0 + d1 * d1 + d2 * d2;
              if (s2 < a.reps2d) continue;  // below r_eps the Ewald coefficients are zero
              const double ir = rsq_nr(s2), rr = s2 * ir;
              double wm = 1.;
              if (__float_as_int(ph.w) == cell_i) {
                const int ll = __float_as_int(pl.w);
                int dl = (ll >> 16) - ilon_i;
                dl += dl < 0 ? nn_i : 0;
                wm = 1. - mtabd[((size_t)ilat_i * nl_i + (ll & 0xffff)) * nn_i + dl];
              }
              if (SL) {
                double EA, EB;
                ewald_sl_ir(a.sl1, a.sl2, rr, ir, tsd, EA, EB);
                const float4 f = sfv[j];
                const double xf = d0 * f.x + d1 * f.y + d2 * f.z;
                const double cc = wm * cA1d;
                accd0 += cc * (EA * d0 * xf + EB * f.x);
                accd1 += cc * (EA * d1 * xf + EB * f.y);
                accd2 += cc * (EA * d2 * xf + EB * f.z);
              }
              if (DL) {
                const double EA = ewald_dl_ir(a.dl1, rr, ir, tsd);
                const float4 g = sgn[2 * j], n2 = sgn[2 * j + 1];
                const double xg = d0 * g.x + d1 * g.y + d2 * g.z;
                const double xn = d0 * g.w + d1 * n2.x + d2 * n2.y;
                const double cc = wm * cA2d * EA * xg * xn;
                accd0 += cc * d0;
                accd1 += cc * d1;
                accd2 += cc * d2;
              }
            }
          }
        }
      }
    }
    __syncthreads();  // every wavefront is done with the position buffers
    accd0 += acc0;
    accd1 += acc1;
    accd2 += acc2;
    if (wv > 0) {
      red[wv - 1][0][lane] = accd0;
      red[wv - 1][1][lane] = accd1;
      red[wv - 1][2][lane] = accd2;
    }
    __syncthreads();
    if (wv == 0 && act) {
      for (int w = 0; w < PW - 1; w++) {
        accd0 += red[w][0][lane];
        accd1 += red[w][1][lane];
        accd2 += red[w][2][lane];
      }
      a.vpair[ti] = accd0;
      a.vpair[ti + (size_t)a.nT] = accd1;
      a.vpair[ti + 2 * (size_t)a.nT] = accd2;
    }
  }
  for (int o = 32; o > 0; o >>= 1) npairs += __shfl_xor(npairs, o, 64);
  if (lane == 0) atomicAdd(&a.stat[0], npairs);
}

// ---------------------------------------------------------------------------
// neighbouring-cell lists: for every target and every other cell with a node in the
// cut-off, the closest node (ties: first visited, i.e. lowest 27-bin visit position
// of the reference's cell list, then highest source index), NbrRbcList_Insert.
// One wavefront per target: lanes test the cells' bounding spheres, then sweep the
// nodes of each candidate cell and reduce (distance, visit key, node).
struct NbrArgs {
  double Lb[3], iLb[3], iLbNc[3];
  int Nc[3];
  double rc;
  int np, nT, nrbc;
  const double *x;  // node coordinates [d][np]
  const int *src_cc;  // coarse bin coordinates per node, packed 10 bits each
  const CellRec *cells;
  const double *sph, *chunk_sph;
  const double *tx;
  const int *tindx, *tact;
  int *cnt, *slot_cell, *slot_node;
  double *slot_dist;
  int *flag;
  const int *cand, *ccnt;  // per cell: cells within reach of any of its nodes (k_cell_cands)
};

// Candidate neighbouring cells per cell, in increasing index: cell j is a candidate of
// cell i when the minimum-image distance of their bounding-sphere centres does not
// exceed r_i + r_j + rc (padded).  A target within r_i of c_i that passes k_nbr's
// sphere test for cell j (|x - c_j| <= rc + r_j) has |c_i - c_j| <= r_i + rc + r_j, so
// the list holds every cell the full scan would keep for it.  -1: more than
// RBC_CAND_MAX candidates (the cell's targets scan every cell).
#define RBC_CAND_MAX 128
__global__ void __launch_bounds__(NBR_T) k_cell_cands(NbrArgs a) {
  const int lane = threadIdx.x & 63;
  const int ci = (blockIdx.x * NBR_T + (int)threadIdx.x) >> 6;
  if (ci >= a.nrbc) return;
  const double ci3[3] = {a.sph[4 * ci], a.sph[4 * ci + 1], a.sph[4 * ci + 2]};
  const double ri = a.sph[4 * ci + 3];
  int *cand = const_cast<int *>(a.cand) + (size_t)ci * RBC_CAND_MAX;
  int n = 0;
  for (int cb = 0; cb < a.nrbc; cb += 64) {
    const int cj = cb + lane;
    bool pass = false;
    if (cj < a.nrbc && cj != ci) {
      const double cj3[3] = {a.sph[4 * cj], a.sph[4 * cj + 1], a.sph[4 * cj + 2]};
      double dd[3];
      const double r = ref_min_image_dist(ci3, cj3, a.Lb, a.iLb, dd);
      pass = !(r > (a.rc + ri + a.sph[4 * cj + 3]) * (1 + 1e-9) + 1e-9 * a.rc);
    }
    const unsigned long long m = __ballot(pass);
    const int pos = n + __popcll(m & ((1ull << lane) - 1ull));
    if (pass && pos < RBC_CAND_MAX) cand[pos] = cj;
    n += __popcll(m);
  }
  if (lane == 0) const_cast<int *>(a.ccnt)[ci] = n <= RBC_CAND_MAX ? n : -1;
}

__device__ __forceinline__ bool nbr_better(double r1, int k1, int n1, double r2, int k2, int n2) {
  if (n1 < 0) return false;
  if (n2 < 0) return true;
  if (r1 != r2) return r1 < r2;
  if (k1 != k2) return k1 < k2;
  return n1 > n2;
}

// One wavefront per target: lanes test the cells' bounding spheres; for each cell
// that passes they test its 64-node chunk spheres and sweep the nodes of the chunks
// that reach within rc, then reduce (distance, visit key, node) over the wavefront.
// Cells enter the list in increasing index order.
__global__ void __launch_bounds__(NBR_T) k_nbr(NbrArgs a) {
  const int lane = threadIdx.x & 63;
  const int ti = (blockIdx.x * NBR_T + (int)threadIdx.x) >> 6;
  if (ti >= a.nT) return;
  if (!a.tact[ti]) {
    if (lane == 0) a.cnt[ti] = 0;
    return;
  }
  double xi[3] = {a.tx[ti], a.tx[ti + a.nT], a.tx[ti + 2 * (size_t)a.nT]};
  const int surf_i = a.tindx[ti];
  int t1, t2, t3;
  hash_index3(xi, a.Nc, a.iLbNc, t1, t2, t3);
  t1--;
  t2--;
  t3--;
  int n = 0;
  bool ovf = false;
  const size_t np = a.np;
  // the cells to test: the candidate list of the target's own cell when the target
  // lies within that cell's bounding sphere, else every cell
  const int *clist = nullptr;
  int ncl = a.nrbc;
  if (surf_i >= 1 && surf_i <= a.nrbc && a.ccnt[surf_i - 1] >= 0) {
    const int ci = surf_i - 1;
    const double c3[3] = {a.sph[4 * ci], a.sph[4 * ci + 1], a.sph[4 * ci + 2]};
    double dd[3];
    if (!(ref_min_image_dist(xi, c3, a.Lb, a.iLb, dd) > a.sph[4 * ci + 3])) {
      clist = a.cand + (size_t)ci * RBC_CAND_MAX;
      ncl = a.ccnt[ci];
    }
  }
  for (int cbase = 0; cbase < ncl; cbase += 64) {
    const int idx = cbase + lane;
    const int cj = idx < ncl ? (clist ? clist[idx] : idx) : -1;
    bool pass = false;
    if (cj >= 0 && cj != surf_i - 1) {
      double cen[3] = {a.sph[4 * cj], a.sph[4 * cj + 1], a.sph[4 * cj + 2]}, dd[3];
      double rcen = ref_min_image_dist(xi, cen, a.Lb, a.iLb, dd);
      pass = !(rcen > a.rc + a.sph[4 * cj + 3] + 1e-9 * a.rc);
    }
    unsigned long long m = __ballot(pass);
    while (m) {
      int kk = __ffsll((long long)m) - 1;
      m &= m - 1;
      int c2 = __shfl(cj, kk, 64);
      const CellRec cr = a.cells[c2];
      int npt = cr.nlat * cr.nlon, nch = (npt + 63) / 64;
      double best = 0;
      int bkey = 0, bnode = -1;
      // only the 64-node chunks whose bounding sphere reaches within rc of the target
      for (int cb = 0; cb < nch; cb += 64) {
        bool cp = false;
        if (cb + lane < nch) {
          const double *s = a.chunk_sph + 4 * (size_t)(cr.choff + cb + lane);
          double cen[3] = {s[0], s[1], s[2]}, dd[3];
          double rcen = ref_min_image_dist(xi, cen, a.Lb, a.iLb, dd);
          cp = !(rcen > a.rc + s[3] + 1e-9 * a.rc);
        }
        unsigned long long cm = __ballot(cp);
        while (cm) {
          const int kc = cb + __ffsll((long long)cm) - 1;
          cm &= cm - 1;
          const int q = 64 * kc + lane;
          if (q >= npt) continue;
          const int node = cr.off + q;
          double xj[3] = {a.x[node], a.x[node + np], a.x[node + 2 * np]}, xx[3];
          double rr = ref_min_image_dist(xi, xj, a.Lb, a.iLb, xx);
          if (rr > a.rc) continue;
          const int cc = a.src_cc[node];
          int d1 = (cc & 1023) - t1 + 1, d2 = ((cc >> 10) & 1023) - t2 + 1, d3 = (cc >> 20) - t3 + 1;
          d1 += d1 < 0 ? a.Nc[0] : (d1 >= a.Nc[0] ? -a.Nc[0] : 0);
          d2 += d2 < 0 ? a.Nc[1] : (d2 >= a.Nc[1] ? -a.Nc[1] : 0);
          d3 += d3 < 0 ? a.Nc[2] : (d3 >= a.Nc[2] ? -a.Nc[2] : 0);
          if (d1 > 2 || d2 > 2 || d3 > 2) continue;  // not in the 27 visited bins
          const int key = 9 * d1 + 3 * d2 + d3;
          if (nbr_better(rr, key, node, best, bkey, bnode)) {
            best = rr;
            bkey = key;
            bnode = node;
          }
        }
      }
      if (__ballot(bnode >= 0) == 0ull) continue;  // no node of this cell within rc
      for (int o = 32; o > 0; o >>= 1) {
        double ob = __shfl_xor(best, o, 64);
        int ok = __shfl_xor(bkey, o, 64), on = __shfl_xor(bnode, o, 64);
        if (nbr_better(ob, ok, on, best, bkey, bnode)) {
          best = ob;
          bkey = ok;
          bnode = on;
        }
      }
      if (bnode >= 0) {
        if (n < RBC_NBR_MAX) {
          if (lane == 0) {
            size_t s = (size_t)ti * RBC_NBR_MAX + n;
            a.slot_cell[s] = c2;
            a.slot_node[s] = bnode;
            a.slot_dist[s] = best;
          }
          n++;
        } else {
          ovf = true;
        }
      }
    }
  }
  if (lane == 0) {
    a.cnt[ti] = n;
    if (ovf) atomicOr(a.flag, 1);
  }
}

// ---------------------------------------------------------------------------
// helpers for the singular quadratures (FP64 throughout)

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// PolarPatch_Build for a single local point (common/ModPolarPatch.F90:99-148)
// with the local point's sines and cosines given (st, ct of thL; sp, cp of phL)
__device__ __forceinline__ void patch_point_sc(double s0, double c0, double phi0, double st,
                                               double ct, double sp, double cp, double &thG,
                                               double &phG) {
  double x0 = st * cp, x1 = st * sp, x2 = ct;
  double y0 = c0 * x0 + 0. * x1 + s0 * x2;
  double y1 = 0. * x0 + 1. * x1 + 0. * x2;
  double y2 = -s0 * x0 + 0. * x1 + c0 * x2;
  y2 = fmax(-1.0, fmin(1.0, y2));
  double ph = atan2(y1, y0) + phi0;
  thG = acos(y2);
  phG = ph - floor(ph * (1. / RBC_TWO_PI)) * RBC_TWO_PI;
}
__device__ __forceinline__ void patch_point(double s0, double c0, double phi0, double thL,
                                            double phL, double &thG, double &phG) {
  patch_point_sc(s0, c0, phi0, sin(thL), cos(thL), sin(phL), cos(phL), thG, phG);
}

// GauLeg (common/ModQuadRule.F90:135-208), node k (ascending) and its weight: each root's Newton iteration is
// independent of the others, so one lane computes one node, as GauLeg does for it.
__device__ void gauleg_node(double x1, double x2, int n, int k, double &x, double &w) {
  const int m = (n + 1) / 2;
  const int j = k < m ? k : n - 1 - k;
  double xm = 0.5 * (x2 + x1), xl = 0.5 * (x2 - x1);
  double z = cos(RBC_PI * ((j + 1) - 0.25) / (n + 0.5)), z1, p1 = 0, p2, p3, pp = 0;
  bool unfinished = true;
  for (int its = 1; its <= 10 && unfinished; its++) {
    p1 = 1.0;
    p2 = 0.0;
    for (int jj = 1; jj <= n; jj++) {
      p3 = p2;
      p2 = p1;
      p1 = ((2.0 * jj - 1.0) * z * p2 - (jj - 1.0) * p3) / jj;
    }
    pp = n * (z * p1 - p2) / (z * z - 1.0);
    z1 = z;
    z = z1 - p1 / pp;
    unfinished = fabs(z - z1) > 3.e-14;
  }
  // GauLeg writes x(j) then x(n-1-j); for the middle node of odd n the second wins
  x = (k == n - 1 - j) ? xm + xl * z : xm - xl * z;
  w = 2.0 * xl / ((1.0 - z * z) * pp * pp);
}

// node k of GauLeg_Sinh (common/ModQuadRule.F90:223-263)
__device__ void gauleg_sinh_node(double xmin, double xmax, double a, double b, int n, int k,
                                 double &x, double &w) {
  double xm = 0.5 * (xmin + xmax), xl = 0.5 * (xmax - xmin);
  double a0 = (a - xm) / xl, b0 = b * xl;
  double f1 = (1 + a0) / b0, f2 = (1 - a0) / b0;
  double u1 = log(f1 + sqrt(1 + f1 * f1)), u2 = log(f2 + sqrt(1 + f2 * f2));
  double mu = 0.5 * (u1 + u2), eta = 0.5 * (u1 - u2);
  double sk, wk;
  gauleg_node(-1., 1., n, k, sk, wk);
  double xk = a0 + b0 * sinh(mu * sk - eta);
  wk = wk * b0 * mu * cosh(mu * sk - eta);
  w = xl * wk;
  x = xm + xl * xk;
}

struct SingArgs {
  double Lb[3], iLb[3];
  double rc, reps, c1, c2;
  int np, nT, nrbc;
  const CellRec *cells;
  const double *spl;
  const double *const *th, *const *phi, *const *thG, *const *phiG, *const *pwt;
  const int *nrad, *nazm;
  const double *sl1, *sl2, *dl1, *mtab;
  const double *x, *fs, *gs, *a3, *detj, *w_unused;
  const double *tx, *tA;
  const int *tindx, *tact;
  double *dvself;
  // near-singular
  const int *cnt, *slot_cell, *slot_node, *item_off;
  const double *const *proj;               // projection tables per grid type
  const int *item_order, *item_slot;       // items sorted by neighbouring cell; item -> slot
  double *item_rec;  // 12 doubles per item
  int *work, *work_cnt;
  double *dvitem;
};

// RBC_SingInt (common/ModRbcSingInt.F90:29-90): one wavefront per target on a cell
__global__ void __launch_bounds__(256) k_self(SingArgs a) {
  int wid = (blockIdx.x * 256 + (int)threadIdx.x) >> 6, lane = threadIdx.x & 63;
  if (wid >= a.nT) return;
  int ti = wid;
  int surf = a.tindx[ti];
  if (!a.tact[ti] || surf < 1 || surf > a.nrbc) {
    if (lane == 0)
      for (int d = 0; d < 3; d++) a.dvself[ti + (size_t)a.nT * d] = 0.;
    return;
  }
  const CellRec cr = a.cells[surf - 1];
  int gt = cr.gtype, m = 2 * cr.nlat, n = cr.nlon;
  int ilat = a.tindx[ti + a.nT] - 1, ilon = a.tindx[ti + 2 * (size_t)a.nT] - 1;
  const double *blk = a.spl + cr.spl;
  size_t vs = (size_t)m * n * 4;
  double xi[3];
  spline_interp(blk + vs * SPL_X, m, n, a.th[gt][ilat], a.phi[gt][ilon], 3, xi);
  int nrad = a.nrad[gt], nazm = a.nazm[gt], npt = nrad * nazm;
  const double *thG = a.thG[gt] + (size_t)npt * (ilat + cr.nlat * ilon);
  const double *phG = a.phiG[gt] + (size_t)npt * (ilat + cr.nlat * ilon);
  const double *pw = a.pwt[gt];
  double c2m = a.c2 * cr.Bcoef;
  double dv0 = 0, dv1 = 0, dv2 = 0;
  for (int p = lane; p < npt; p += 64) {
    int irad = p % nrad;
    double thj = thG[p], phj = phG[p], xj[3], xx[3];
    spline_interp(blk + vs * SPL_X, m, n, thj, phj, 3, xj);
    for (int d = 0; d < 3; d++) xx[d] = xj[d] - xi[d];
    double rr = sqrt(xx[0] * xx[0] + xx[1] * xx[1] + xx[2] * xx[2]);
    if (rr >= a.rc) continue;
    if (a.c1 != 0) {
      double fj[3], EA, EB;
      spline_interp(blk + vs * SPL_F, m, n, thj, phj, 3, fj);
      for (int d = 0; d < 3; d++) fj[d] = pw[irad] * fj[d];
      ewald_sl(a.sl1, a.sl2, rr, a.reps, a.rc, EA, EB);
      double xf = xx[0] * fj[0] + xx[1] * fj[1] + xx[2] * fj[2];
      dv0 += a.c1 * (EA * xx[0] * xf + EB * fj[0]);
      dv1 += a.c1 * (EA * xx[1] * xf + EB * fj[1]);
      dv2 += a.c1 * (EA * xx[2] * xf + EB * fj[2]);
    }
    if (c2m != 0) {
      double gj[3], nj[3];
      spline_interp(blk + vs * SPL_G, m, n, thj, phj, 3, gj);
      for (int d = 0; d < 3; d++) gj[d] = pw[irad] * gj[d];
      spline_interp(blk + vs * SPL_A3, m, n, thj, phj, 3, nj);
      double EA = ewald_dl(a.dl1, rr, a.reps, a.rc);
      double xg = xx[0] * gj[0] + xx[1] * gj[1] + xx[2] * gj[2];
      double xn = xx[0] * nj[0] + xx[1] * nj[1] + xx[2] * nj[2];
      dv0 += c2m * (EA * xx[0] * xg * xn);
      dv1 += c2m * (EA * xx[1] * xg * xn);
      dv2 += c2m * (EA * xx[2] * xg * xn);
    }
  }
  dv0 = wave_sum(dv0);
  dv1 = wave_sum(dv1);
  dv2 = wave_sum(dv2);
  if (lane == 0) {
    a.dvself[ti] = dv0;
    a.dvself[ti + (size_t)a.nT] = dv1;
    a.dvself[ti + 2 * (size_t)a.nT] = dv2;
  }
}

// Near-singular stage 1: translate the target next to the closest node, project it
// onto the neighbour's spline surface (Spline_FindProjection) and apply the distance
// test of RBC_NearSingInt.  One lane per (target, neighbouring cell) item; items are
// processed in neighbouring-cell order, and consecutive blocks of that order are
// placed on one XCD, so a cell's spline block is fetched into few L2s.  The quadratic
// fit uses the per-grid-type least-squares weights (c = sum_k W_k d2_k).
// item record: xi(3) x0(3) th0 phi0 a30(3) dist
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_proj(SingArgs a, int nitems) {
  const int lblk = xcd_block((nitems + 255) / 256);
  if (lblk < 0) return;
  const int gidx = lblk * 256 + threadIdx.x;
  if (gidx >= nitems) return;
  const int itm = a.item_order[gidx];
  const int s = a.item_slot[itm];
  const int ti = s / RBC_NBR_MAX;
  const int cj = a.slot_cell[s], node = a.slot_node[s];
  const CellRec cr = a.cells[cj];
  const int gt = cr.gtype, m = 2 * cr.nlat, n = cr.nlon;
  const int q = node - cr.off, ilat0 = q % cr.nlat, ilon0 = q / cr.nlat;
  double th0 = a.th[gt][ilat0], phi0 = a.phi[gt][ilon0], x0[3], xt[3];
  const size_t np = a.np;
  for (int d = 0; d < 3; d++) x0[d] = a.x[node + np * d];
  for (int d = 0; d < 3; d++) {
#pragma clang fp contract(off)
    double xx = a.tx[ti + (size_t)a.nT * d] - x0[d];
    double sh = round(xx * a.iLb[d]) * a.Lb[d];
    xx = xx - sh;
    xt[d] = x0[d] + xx;
  }
  const double *blk = a.spl + cr.spl;
  const size_t vs = (size_t)m * n * 4;
  const double *sxp = blk + vs * SPL_X;
  const double *pj = a.proj[gt];
  spline_find_projection(sxp, m, n, pj, xt, th0, phi0, x0);
  double a30[3];
  spline_interp(blk + vs * SPL_A3, m, n, th0, phi0, 3, a30);
  double dist = a30[0] * (xt[0] - x0[0]) + a30[1] * (xt[1] - x0[1]) + a30[2] * (xt[2] - x0[2]);
  double *r = a.item_rec + 12 * (size_t)itm;
  for (int d = 0; d < 3; d++) {
    r[d] = xt[d];
    r[3 + d] = x0[d];
    r[8 + d] = a30[d];
  }
  r[6] = th0;
  r[7] = phi0;
  r[11] = dist;
  for (int d = 0; d < 3; d++) a.dvitem[3 * (size_t)itm + d] = 0.;
  if (dist > 2 * cr.meshSize) return;
  int slot = atomicAdd(a.work_cnt, 1);
  a.work[2 * slot] = itm;
  a.work[2 * slot + 1] = cj;
}

// items in target order: (neighbouring cell, item) keys and the item -> slot map
__global__ void k_items(int nT, const int *__restrict__ cnt, const int *__restrict__ item_off,
                        const int *__restrict__ slot_cell, int *key, int *val, int *item_slot) {
  int ti = blockIdx.x * blockDim.x + threadIdx.x;
  if (ti >= nT) return;
  int o = item_off[ti];
  for (int p = 0; p < cnt[ti]; p++) {
    int s = ti * RBC_NBR_MAX + p;
    key[o + p] = slot_cell[s];
    val[o + p] = o + p;
    item_slot[o + p] = s;
  }
}

// RBC_NearSingInt_ReAdd (common/ModRbcSingInt.F90:232-310), lanes over the 16 x 32 patch
__device__ void readd_wave(const SingArgs &a, const CellRec &cr, const double *blk, int m, int n,
                           const double xi[3], const double x0[3], double th0, double phi0,
                           double radPat, double c1, double c2, double dv[3]) {
  const int nrad = 16, nazm = 32;
  // lane l always takes radial node l % 16 (64 is a multiple of nrad): its node and
  // masked weight are computed once
  const int lane = threadIdx.x & 63, irad = lane & (nrad - 1);
  double d0 = xi[0] - x0[0], d1 = xi[1] - x0[1], d2 = xi[2] - x0[2];
  double dist = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
  double sizePat = radPat * sqrt(cr.area / (4 * RBC_PI));
  double thl, wl;
  if (dist > 2.2250738585072014e-308) {
    dist = dist * (radPat / sizePat);
    gauleg_sinh_node(0., radPat, 0., dist, nrad, irad, thl, wl);
  } else {
    gauleg_node(0., radPat, nrad, irad, thl, wl);
  }
  double w = wl * sin(thl) * (RBC_TWO_PI / nazm);
  w = w * mask_fn(a.mtab, thl / radPat);
  size_t vs = (size_t)m * n * 4;
  double s0 = sin(th0), c0 = cos(th0);
  double a0 = 0, a1 = 0, a2 = 0;
  // sin and cos of the nazm azimuths, once per wavefront (lane j < nazm: azimuth j, the
  // same expression as the point loop's), and of the lane's own radial angle
  __shared__ double azsc[256 / 64][nazm][2];
  const int wsl = (int)(threadIdx.x >> 6);
  if (lane < nazm) {
    const double phL = lane * RBC_TWO_PI / nazm;
    azsc[wsl][lane][0] = sin(phL);
    azsc[wsl][lane][1] = cos(phL);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  const double stl = sin(thl), ctl = cos(thl);
  for (int p = lane; p < nrad * nazm; p += 64) {
    int iazm = p / nrad;
    double thj, phj;
    patch_point_sc(s0, c0, phi0, stl, ctl, azsc[wsl][iazm][0], azsc[wsl][iazm][1], thj, phj);
    double xj[3], xx[3];
    spline_interp(blk + vs * SPL_X, m, n, thj, phj, 3, xj);
    for (int d = 0; d < 3; d++) xx[d] = xj[d] - xi[d];
    double rr = sqrt(xx[0] * xx[0] + xx[1] * xx[1] + xx[2] * xx[2]);
    if (rr > a.rc) continue;
    if (c1 != 0) {
      double fj[3], EA, EB;
      spline_interp(blk + vs * SPL_F, m, n, thj, phj, 3, fj);
      for (int d = 0; d < 3; d++) fj[d] = w * fj[d];
      ewald_sl(a.sl1, a.sl2, rr, a.reps, a.rc, EA, EB);
      double xf = xx[0] * fj[0] + xx[1] * fj[1] + xx[2] * fj[2];
      a0 += c1 * (EA * xx[0] * xf + EB * fj[0]);
      a1 += c1 * (EA * xx[1] * xf + EB * fj[1]);
      a2 += c1 * (EA * xx[2] * xf + EB * fj[2]);
    }
    if (c2 != 0) {
      double gj[3], nj[3];
      spline_interp(blk + vs * SPL_A3, m, n, thj, phj, 3, nj);
      spline_interp(blk + vs * SPL_G, m, n, thj, phj, 3, gj);
      for (int d = 0; d < 3; d++) gj[d] = w * gj[d];
      double EA = ewald_dl(a.dl1, rr, a.reps, a.rc);
      double xg = xx[0] * gj[0] + xx[1] * gj[1] + xx[2] * gj[2];
      double xn = xx[0] * nj[0] + xx[1] * nj[1] + xx[2] * nj[2];
      a0 += c2 * (EA * xx[0] * xg * xn);
      a1 += c2 * (EA * xx[1] * xg * xn);
      a2 += c2 * (EA * xx[2] * xg * xn);
    }
  }
  dv[0] = wave_sum(a0);
  dv[1] = wave_sum(a1);
  dv[2] = wave_sum(a2);
}

// Near-singular stage 2: one wavefront per item that passed the distance test.
// RBC_NearSingInt (:103-166) = subtract (:175-226) + re-add, or the interpolation
// between the re-adds at xi1 and on the surface (with the double-layer jump).
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_near(SingArgs a, int nwork) {
  int wid = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
  if (wid >= nwork) return;
  int it = a.work[2 * wid], cj = a.work[2 * wid + 1];
  const CellRec cr = a.cells[cj];
  int gt = cr.gtype, m = 2 * cr.nlat, n = cr.nlon, nl = cr.nlat, nn = cr.nlon;
  const double *r = a.item_rec + 12 * (size_t)it;
  double xi[3] = {r[0], r[1], r[2]}, x0[3] = {r[3], r[4], r[5]}, a30[3] = {r[8], r[9], r[10]};
  double th0 = r[6], phi0 = r[7], dist = r[11];
  double radPat = cr.radius;
  double c1 = a.c1, c2 = a.c2 * cr.Bcoef;
  const double *blk = a.spl + cr.spl;
  size_t vs = (size_t)m * n * 4;
  double sizePat = radPat * sqrt(cr.area / (4 * RBC_PI));
  double dist1 = copysign(0.01 * sizePat, dist);
  // subtract: nodes of PolarPatch_FindPoints (common/ModPolarPatch.F90:161-209), with
  // their multiplicity when the longitude range wraps more than once
  const double *th = a.th[gt], *ph = a.phi[gt];
  const double eps = 1.e-10;
  double h_phi = RBC_TWO_PI / nn, ih_phi = 1.0 / h_phi, cosr0 = cos(radPat);
  double s0 = 0, s1 = 0, s2 = 0;
  // the longitude range of each latitude row inside the patch depends on the row only:
  // lane i computes row i's (jmin, jmax) once (0 rows: outside the latitude band), and
  // the node loop reads it from the wavefront's slot in LDS
  __shared__ int rowj[256 / 64][64][2];
  const int wsl = (int)(threadIdx.x >> 6);
  if (lane < nl) {
    const int i = lane;
    int jmin = 1, jmax = 0;
    if (!(th[i] <= th0 - radPat) && !(th[i] >= th0 + radPat)) {
      double dth = th[i] - th0;
      double dphi = 1. - (cos(dth) - cosr0) / ((sin(th[i]) + eps) * (sin(th0) + eps));
      dphi = fmax(-1., fmin(1., dphi));
      dphi = acos(dphi);
      if (dphi > RBC_PI - eps) {
        jmin = 1;
        jmax = nn;
      } else {
        jmin = (int)ceil((phi0 - dphi - ph[0]) * ih_phi) + 1;
        jmax = (int)floor((phi0 + dphi - ph[0]) * ih_phi) + 1;
      }
    } else {
      jmin = 1;
      jmax = -(1 << 29);  // no node of this row
    }
    rowj[wsl][i][0] = jmin;
    rowj[wsl][i][1] = jmax;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  // first row that fails "ths >= th0 + r0" ends the scan (ths ascending)
  for (int q = lane; q < nl * nn; q += 64) {
    int i = q % nl, jl = q / nl;
    const int jmin = rowj[wsl][i][0], jmax = rowj[wsl][i][1];
    if (jmax < -(1 << 28)) continue;  // outside the latitude band
    // multiplicity of node jl: #{j in [jmin,jmax] : (j-1) mod nn == jl}
    int lo = jmin - 1 - jl, hi = jmax - 1 - jl;  // need k*nn in [lo, hi]
    if (hi < lo) continue;
    int kmin = (int)ceil((double)lo / nn), kmax = (int)floor((double)hi / nn);
    int mult = kmax - kmin + 1;
    if (mult <= 0) continue;
    int node = cr.off + q;
    size_t np = a.np;
    double xx[3];
    for (int d = 0; d < 3; d++) xx[d] = a.x[node + np * d] - xi[d];
    double rr = sqrt(xx[0] * xx[0] + xx[1] * xx[1] + xx[2] * xx[2]);
    if (rr > a.rc) continue;
    double dthm = dist_on_sphere(th0, phi0, th[i], ph[jl]);
    double mask = mask_fn(a.mtab, dthm / radPat);
    for (int rep = 0; rep < mult; rep++) {
      if (c1 != 0) {
        double fj[3] = {a.fs[node], a.fs[node + np], a.fs[node + 2 * np]}, EA, EB;
        ewald_sl(a.sl1, a.sl2, rr, a.reps, a.rc, EA, EB);
        double xf = xx[0] * fj[0] + xx[1] * fj[1] + xx[2] * fj[2];
        s0 -= c1 * mask * (EA * xx[0] * xf + EB * fj[0]);
        s1 -= c1 * mask * (EA * xx[1] * xf + EB * fj[1]);
        s2 -= c1 * mask * (EA * xx[2] * xf + EB * fj[2]);
      }
      if (c2 != 0) {
        double gj[3] = {a.gs[node], a.gs[node + np], a.gs[node + 2 * np]};
        double nj[3] = {a.a3[node], a.a3[node + np], a.a3[node + 2 * np]};
        double EA = ewald_dl(a.dl1, rr, a.reps, a.rc);
        double xg = xx[0] * gj[0] + xx[1] * gj[1] + xx[2] * gj[2];
        double xn = xx[0] * nj[0] + xx[1] * nj[1] + xx[2] * nj[2];
        s0 -= c2 * mask * (EA * xx[0] * xg * xn);
        s1 -= c2 * mask * (EA * xx[1] * xg * xn);
        s2 -= c2 * mask * (EA * xx[2] * xg * xn);
      }
    }
  }
  double dv[3] = {wave_sum(s0), wave_sum(s1), wave_sum(s2)}, dt[3];
  if (fabs(dist) >= fabs(dist1)) {
    readd_wave(a, cr, blk, m, n, xi, x0, th0, phi0, radPat, c1, c2, dt);
    for (int d = 0; d < 3; d++) dv[d] += dt[d];
  } else {
    double xi1[3], dv1[3], dv0[3];
    for (int d = 0; d < 3; d++) xi1[d] = x0[d] + dist1 * a30[d];
    readd_wave(a, cr, blk, m, n, xi1, x0, th0, phi0, radPat, c1, c2, dv1);
    readd_wave(a, cr, blk, m, n, x0, x0, th0, phi0, radPat, c1, c2, dv0);
    if (c2 != 0) {
      double detJ0, g0[3];
      spline_interp(blk + vs * SPL_DETJ, m, n, th0, phi0, 1, &detJ0);
      spline_interp(blk + vs * SPL_G, m, n, th0, phi0, 3, g0);
      for (int d = 0; d < 3; d++) g0[d] = g0[d] / detJ0;
      for (int d = 0; d < 3; d++) dv0[d] += (dist > 0 ? 1. : -1.) * c2 * 4 * RBC_PI * g0[d];
    }
    for (int d = 0; d < 3; d++) dv[d] += dv0[d] + dist / dist1 * (dv1[d] - dv0[d]);
  }
  if (lane == 0)
    for (int d = 0; d < 3; d++) a.dvitem[3 * (size_t)it + d] = dv[d];
}

// AddLinearInt (common/ModIntOnRbcs.F90:162-201): partial sums per block
__global__ void __launch_bounds__(256) k_linear_partial(int np, const double *__restrict__ x,
                                                        const double *__restrict__ g,
                                                        const double *__restrict__ a3,
                                                        const double *__restrict__ ds,
                                                        const int *__restrict__ node_cell,
                                                        const CellRec *__restrict__ cells,
                                                        double *part) {
  __shared__ double red[3][256];
  double s[3] = {0, 0, 0};
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < np; p += gridDim.x * blockDim.x) {
    double vn = 0;
    for (int d = 0; d < 3; d++) vn += g[p + (size_t)np * d] * a3[p + (size_t)np * d];
    double B = cells[node_cell[p]].Bcoef;
    for (int d = 0; d < 3; d++) s[d] += B * x[p + (size_t)np * d] * vn * ds[p];
  }
  for (int d = 0; d < 3; d++) red[d][threadIdx.x] = s[d];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int d = 0; d < 3; d++) red[d][threadIdx.x] += red[d][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int d = 0; d < 3; d++) part[3 * blockIdx.x + d] = red[d][0];
}

__global__ void k_linear_final(int nb, const double *part, double scale, double *xv) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s[3] = {0, 0, 0};
  for (int b = 0; b < nb; b++)
    for (int d = 0; d < 3; d++) s[d] += part[3 * b + d];
  for (int d = 0; d < 3; d++) xv[d] = scale * s[d];
}

struct CombineArgs {
  int nT;
  double c2;
  const int *tact, *cnt, *item_off;
  const double *tA, *vpair, *dvself, *dvitem, *xv;
  int has_linear;
  double *v;
};

__global__ void k_combine(CombineArgs a) {
  int ti = blockIdx.x * blockDim.x + threadIdx.x;
  if (ti >= a.nT || !a.tact[ti]) return;
  const double A = a.tA[ti];
  double s[3];
  for (int d = 0; d < 3; d++) s[d] = a.vpair[ti + (size_t)a.nT * d] + a.dvself[ti + (size_t)a.nT * d] / A;
  const int off = a.item_off[ti], n = a.cnt[ti];
  for (int p = 0; p < n; p++)
    for (int d = 0; d < 3; d++) s[d] += a.dvitem[3 * (size_t)(off + p) + d] / A;
  for (int d = 0; d < 3; d++) {
    if (a.has_linear) s[d] += a.c2 * a.xv[d] / A;
    a.v[ti + (size_t)a.nT * d] += s[d];
  }
}

// ---------------------------------------------------------------------------
// table pointer arrays per grid type, owned by the context (rebuilt when a grid
// type is added)
int rbc_type_ptrs(rbc3d_ctx *c, TypePtrs *tp) {
  size_t nt = c->gtypes.size();
  if (c->tp_ntypes != nt) {
    std::vector<const double *> ptr(7 * nt);
    std::vector<int> ni(2 * nt);
    for (size_t i = 0; i < nt; i++) {
      GridType *g = c->gtypes[i];
      ptr[0 * nt + i] = g->d_mask.p;
      ptr[1 * nt + i] = g->d_th.p;
      ptr[2 * nt + i] = g->d_phi.p;
      ptr[3 * nt + i] = g->d_thG.p;
      ptr[4 * nt + i] = g->d_phiG.p;
      ptr[5 * nt + i] = g->d_pwt.p;
      ptr[6 * nt + i] = g->d_proj.p;
      ni[i] = g->nrad;
      ni[nt + i] = g->nazm;
    }
    std::vector<const float *> pf(nt);
    for (size_t i = 0; i < nt; i++) pf[i] = c->gtypes[i]->d_maskf.p;
    RBC_TRY(c->d_tp_maskf.ensure(nt));
    HIP_TRY(hipMemcpy(c->d_tp_maskf.p, pf.data(), nt * sizeof(pf[0]), hipMemcpyHostToDevice));
    RBC_TRY(c->d_tp_ptr.ensure(ptr.size()));
    RBC_TRY(c->d_tp_int.ensure(ni.size()));
    HIP_TRY(hipMemcpy(c->d_tp_ptr.p, ptr.data(), ptr.size() * sizeof(ptr[0]), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->d_tp_int.p, ni.data(), ni.size() * sizeof(int), hipMemcpyHostToDevice));
    c->tp_ntypes = nt;
  }
  const double *const *p = c->d_tp_ptr.p;
  tp->mask = p;
  tp->th = p + nt;
  tp->phi = p + 2 * nt;
  tp->thG = p + 3 * nt;
  tp->phiG = p + 4 * nt;
  tp->pwt = p + 5 * nt;
  tp->proj = p + 6 * nt;
  tp->maskf = c->d_tp_maskf.p;
  tp->nrad = c->d_tp_int.p;
  tp->nazm = c->d_tp_int.p + nt;
  return RBC3D_OK;
}

static int exclusive_scan(rbc3d_ctx *c, const int *in, int *out, int n) {
  size_t tmp = 0;
  HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, in, out, n, c->stream));
  RBC_TRY(c->d_cub_tmp.ensure(tmp + 16));
  HIP_TRY(hipcub::DeviceScan::ExclusiveSum(c->d_cub_tmp.p, tmp, in, out, n, c->stream));
  return RBC3D_OK;
}

// launch the split-form pair kernel with the arguments of the FP64 one
template <bool SL, bool DL>
static void launch_pair(bool filt, int nblk, hipStream_t st, const PairArgs &pa) {
  if (filt)
    k_pair<SL, DL, true><<<nblk, PT * PW, 0, st>>>(pa);
  else
    k_pair<SL, DL, false><<<nblk, PT * PW, 0, st>>>(pa);
}

static int launch_pair32(rbc3d_ctx *c, const PairArgs &pa, int nblk_max,
                         const float *const *maskf) {
  Pair32Args a;
  bool wrap = false;
  for (int d = 0; d < 3; d++) {
    a.Lb[d] = pa.Lb[d];
    a.iLb[d] = pa.iLb[d];
    a.w[d] = c->Lb[d] / c->Nf[d];
    a.fLb[d] = (float)pa.Lb[d];
    a.fiLb[d] = (float)pa.iLb[d];
    a.Nf[d] = pa.Nf[d];
    // sources within rc of a target of the bin lie within rc + w/2 of its centre
    wrap = wrap || !(c->Lb[d] - a.w[d] > 2 * c->conf.rc * (1 + 1e-6));
  }
  a.wrap = wrap;
  a.rc2 = (float)(c->conf.rc * c->conf.rc);
  a.reps2 = (float)(c->r_eps * c->r_eps);
  a.tabscale = (float)pa.tabscale;
  a.c1 = (float)pa.c1;
  a.c2 = (float)pa.c2;
  a.c1d = pa.c1;
  a.c2d = pa.c2;
  a.rcd = c->conf.rc;
  a.near2 = (float)(RBC_SPLIT_NEAR * RBC_SPLIT_NEAR * c->conf.rc * c->conf.rc);
  a.reps2d = c->r_eps * c->r_eps;
  a.sl1 = pa.sl1;
  a.sl2 = pa.sl2;
  a.dl1 = pa.dl1;
  a.np = pa.np;
  a.nT = pa.nT;
  a.nfbins = pa.nfbins;
  a.nrbc = pa.nrbc;
  a.nrows = pa.nrows;
  a.rows = pa.rows;
  a.rec = pa.rec;
  a.src_node = pa.src_node;
  a.bin_start = pa.bin_start;
  a.cells = pa.cells;
  a.tab = c->d_tabf.p;
  a.mask = pa.mask;
  a.maskf = maskf;
  a.tx = pa.tx;
  a.tA = pa.tA;
  a.tindx = pa.tindx;
  a.tact = pa.tact;
  a.torder = pa.torder;
  a.tbin_start = pa.tbin_start;
  if (2 * pa.nrows > RUN_MAX) {
    rbc_set_error("split pair kernel: more source-bin rows than its run table holds");
    return RBC3D_EINVAL;
  }
  a.chunk_off = pa.chunk_off;
  a.vpair = pa.vpair;
  a.stat = pa.stat;
  const bool sl = pa.c1 != 0, dl = pa.c2 != 0;
  // one workgroup per target chunk at most (the hardware balances chunks of uneven
  // cost); workgroups without a chunk leave before loading the tables
  int grid = (int)xcd_grid(nblk_max);
  hipStream_t st = c->stream;
  if (wrap) {
    if (sl && dl)
      k_pair32<true, true, true><<<grid, PT * PW, 0, st>>>(a);
    else if (sl)
      k_pair32<true, false, true><<<grid, PT * PW, 0, st>>>(a);
    else if (dl)
      k_pair32<false, true, true><<<grid, PT * PW, 0, st>>>(a);
  } else {
    if (sl && dl)
      k_pair32<true, true, false><<<grid, PT * PW, 0, st>>>(a);
    else if (sl)
      k_pair32<true, false, false><<<grid, PT * PW, 0, st>>>(a);
    else if (dl)
      k_pair32<false, true, false><<<grid, PT * PW, 0, st>>>(a);
  }
  HIP_TRY(hipGetLastError());
  return RBC3D_OK;
}

// cacheable: the targets are the cells' own nodes (the cell-BIE target list).  Target
// binning, neighbouring-cell lists and closest-point projections depend on the
// geometry only; they are kept from the previous cacheable call while the cells have
// not been set again, and any other call overwrites (and so drops) them.
int rbc_add_int_on_rbcs_dev(rbc3d_ctx *c, double c1, double c2, int nT, const double *d_tx,
                            const int *d_tindx, const int *d_act, const double *d_A, double *d_v,
                            bool cacheable) {
  if (c->nrbc == 0 || nT == 0) return RBC3D_OK;
  hipStream_t st = c->stream;
  int nb = c->nfbins;
  TypePtrs tpv, *tp = &tpv;
  RBC_TRY(rbc_type_ptrs(c, tp));
  cacheable = cacheable && c->geom_reuse;
  const bool hit = cacheable && c->cache_epoch == c->geom_epoch && c->cache_nT == nT;
  c->cache_epoch = -1;  // valid again only when this call completes as cacheable
  RBC_TRY(c->d_flag.ensure(2 * (size_t)(nb + 1) + 8));
  int *chunk_cnt = c->d_flag.p + 8, *chunk_off = chunk_cnt + (nb + 1);
  if (!hit) {
    // targets binned like the sources
    RBC_TRY(c->d_tbin.ensure(nT + 1));
    RBC_TRY(c->d_torder.ensure(nT + 1));
    RBC_TRY(c->d_tbin_start.ensure(nb + 1));
    rbc_stage_begin(c, ST_TSORT);
    RBC_TRY(fine_sort(c, nT, d_tx, nT, c->d_tbin.p, c->d_torder.p, c->d_tbin_start.p));
    rbc_stage_end(c, ST_TSORT);
    k_chunks<<<(nb + 1 + 255) / 256, 256, 0, st>>>(nb, c->d_tbin_start.p, chunk_cnt);
    RBC_TRY(exclusive_scan(c, chunk_cnt, chunk_off, nb + 1));
  }
  RBC_TRY(c->d_vacc.ensure((size_t)3 * nT + 1));
  RBC_TRY(c->d_pairstat.ensure(2));
  // cleared by kernels, not memsets: this path is recorded into graphs
  k_fill_u64<<<1, 64, 0, st>>>(c->d_pairstat.p, 2, 0ull);
  k_fill_i<<<1, 64, 0, st>>>(c->d_flag.p, 8, 0);
  PairArgs pa;
  for (int d = 0; d < 3; d++) {
    pa.Lb[d] = c->Lb[d];
    pa.iLb[d] = c->iLb[d];
    pa.iLbNf[d] = c->iLbNf[d];
    pa.Nf[d] = c->Nf[d];
  }
  pa.img = c->pair_img;
  pa.rc = c->conf.rc;
  pa.rc2hi = c->conf.rc * c->conf.rc * (1.0 + 1e-12);
  pa.rc2lo = c->conf.rc * c->conf.rc * (1.0 - 1e-12);
  pa.reps2 = c->r_eps * c->r_eps;
  pa.tabscale = RBC_NTAB / c->conf.rc;
  pa.nrows = (int)c->noffs;
  pa.rows = c->d_offs.p;
  pa.c1 = c1;
  pa.c2 = c2;
  pa.np = c->np;
  pa.nT = nT;
  pa.nfbins = nb;
  pa.nrbc = c->nrbc;
  pa.rec = c->d_src_rec.p;
  pa.src_node = c->d_work.p;
  pa.bin_start = c->d_fbin_start.p;
  pa.node_cell = c->d_node_cell.p;
  pa.node_ll = c->d_node_ll.p;
  pa.cells = c->d_cells.p;
  pa.sl1 = c->d_sl1.p;
  pa.sl2 = c->d_sl2.p;
  pa.dl1 = c->d_dl1.p;
  pa.mask = tp->mask;
  pa.tx = d_tx;
  pa.tA = d_A;
  pa.tindx = d_tindx;
  pa.tact = d_act;
  pa.torder = c->d_torder.p;
  pa.tbin_start = c->d_tbin_start.p;
  pa.chunk_off = chunk_off;
  pa.vpair = c->d_vacc.p;
  pa.stat = c->d_pairstat.p;
  int nblk = nb + (nT + PT - 1) / PT + 1;
  rbc_stage_begin(c, ST_PAIR);
  if (c->prec == RBC3D_PREC_SPLIT && c->split_pair32) {
    RBC_TRY(launch_pair32(c, pa, nblk, tp->maskf));
  } else {
    // the source filter where fine bins hold 32 sources or more on average
    const bool filt = c->pair_img && c->np >= 32 * (long long)nb;
    if (c1 != 0 && c2 != 0)
      launch_pair<true, true>(filt, nblk, st, pa);
    else if (c1 != 0)
      launch_pair<true, false>(filt, nblk, st, pa);
    else if (c2 != 0)
      launch_pair<false, true>(filt, nblk, st, pa);
    else
      launch_pair<false, false>(filt, nblk, st, pa);
  }
  rbc_stage_end(c, ST_PAIR);
  HIP_TRY(hipGetLastError());
  // neighbouring-cell lists
  int *item_off = nullptr, nitems = 0;
  if (!hit) {
  RBC_TRY(c->d_nbr_cnt.ensure(2 * (size_t)nT + 2));
  RBC_TRY(c->d_nbr_cell.ensure((size_t)RBC_NBR_MAX * nT + 1));
  RBC_TRY(c->d_nbr_node.ensure((size_t)RBC_NBR_MAX * nT + 1));
  RBC_TRY(c->d_nbr_dist.ensure((size_t)RBC_NBR_MAX * nT + 1));
  NbrArgs na;
  for (int d = 0; d < 3; d++) {
    na.Lb[d] = c->Lb[d];
    na.iLb[d] = c->iLb[d];
    na.iLbNc[d] = c->iLbNc[d];
    na.Nc[d] = c->Nc[d];
  }
  na.rc = c->conf.rc;
  na.np = c->np;
  na.nT = nT;
  na.nrbc = c->nrbc;
  na.x = c->d_x.p;
  na.src_cc = c->d_src_cc.p;
  na.cells = c->d_cells.p;
  na.sph = c->d_cell_sph.p;
  na.chunk_sph = c->d_chunk_sph.p;
  na.tx = d_tx;
  na.tindx = d_tindx;
  na.tact = d_act;
  na.cnt = c->d_nbr_cnt.p;
  na.slot_cell = c->d_nbr_cell.p;
  na.slot_node = c->d_nbr_node.p;
  na.slot_dist = c->d_nbr_dist.p;
  na.flag = c->d_flag.p;
  RBC_TRY(c->d_cell_cand.ensure((size_t)(RBC_CAND_MAX + 1) * c->nrbc + 1));
  na.cand = c->d_cell_cand.p;
  na.ccnt = c->d_cell_cand.p + (size_t)RBC_CAND_MAX * c->nrbc;
  rbc_stage_begin(c, ST_NBR);
  k_cell_cands<<<(unsigned)(((size_t)c->nrbc * 64 + NBR_T - 1) / NBR_T), NBR_T, 0, st>>>(na);
  k_nbr<<<(unsigned)(((size_t)nT * 64 + NBR_T - 1) / NBR_T), NBR_T, 0, st>>>(na);
  rbc_stage_end(c, ST_NBR);
  HIP_TRY(hipGetLastError());
  item_off = c->d_nbr_cnt.p + nT + 1;
  HIP_TRY(hipMemsetAsync(c->d_nbr_cnt.p + nT, 0, sizeof(int), st));
  RBC_TRY(exclusive_scan(c, c->d_nbr_cnt.p, item_off, nT + 1));
  int hflag[2] = {0, 0};
  HIP_TRY(hipMemcpyAsync(hflag, c->d_flag.p, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(&nitems, item_off + nT, sizeof(int), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (hflag[0]) {
    rbc_set_error("more than 32 neighbouring cells within the cut-off of a target");
    return RBC3D_ENBR;
  }
  } else {
    item_off = c->d_nbr_cnt.p + nT + 1;
    nitems = c->cache_nitems;
  }
  c->last_nitems = nitems;
  if (!hit) c->last_nwork = 0;
  // singular and near-singular corrections
  RBC_TRY(c->d_scratch.ensure((size_t)12 * nitems + (size_t)3 * nT + 16));
  RBC_TRY(c->d_dvitem.ensure((size_t)3 * nitems + 3));
  RBC_TRY(c->d_work_cnt.ensure(2));
  // d_work holds the sorted-position -> node map; the near-singular worklist is separate
  DBuf<int> &wl = c->d_worklist;
  RBC_TRY(wl.ensure(2 * (size_t)nitems + 2));
  k_fill_i<<<1, 64, 0, st>>>(c->d_work_cnt.p, 1, 0);
  SingArgs sa;
  for (int d = 0; d < 3; d++) {
    sa.Lb[d] = c->Lb[d];
    sa.iLb[d] = c->iLb[d];
  }
  sa.rc = c->conf.rc;
  sa.reps = c->r_eps;
  sa.c1 = c1;
  sa.c2 = c2;
  sa.np = c->np;
  sa.nT = nT;
  sa.nrbc = c->nrbc;
  sa.cells = c->d_cells.p;
  sa.spl = c->d_spl.p;
  sa.th = tp->th;
  sa.phi = tp->phi;
  sa.thG = tp->thG;
  sa.phiG = tp->phiG;
  sa.pwt = tp->pwt;
  sa.nrad = tp->nrad;
  sa.nazm = tp->nazm;
  sa.sl1 = c->d_sl1.p;
  sa.sl2 = c->d_sl2.p;
  sa.dl1 = c->d_dl1.p;
  sa.mtab = c->d_mtab.p;
  sa.x = c->d_x.p;
  sa.fs = c->d_fs.p;
  sa.gs = c->d_gs.p;
  sa.a3 = c->d_a3.p;
  sa.detj = c->d_detj.p;
  sa.w_unused = nullptr;
  sa.tx = d_tx;
  sa.tA = d_A;
  sa.tindx = d_tindx;
  sa.tact = d_act;
  double *dvself = c->d_scratch.p + (size_t)12 * nitems;
  sa.dvself = dvself;
  sa.cnt = c->d_nbr_cnt.p;
  sa.slot_cell = c->d_nbr_cell.p;
  sa.slot_node = c->d_nbr_node.p;
  sa.item_off = item_off;
  sa.item_rec = c->d_scratch.p;
  sa.work = wl.p;
  sa.work_cnt = c->d_work_cnt.p;
  sa.dvitem = c->d_dvitem.p;
  rbc_stage_begin(c, ST_SELF);
  k_self<<<(unsigned)(((size_t)nT * 64 + 255) / 256), 256, 0, st>>>(sa);
  rbc_stage_end(c, ST_SELF);
  HIP_TRY(hipGetLastError());
  if (nitems > 0 && hit) {
    const int nwork = c->cache_nwork;
    c->last_nwork = nwork;
    if (nwork > 0) {
      rbc_stage_begin(c, ST_NEAR);
      k_near<<<(unsigned)(((size_t)nwork * 64 + 255) / 256), 256, 0, st>>>(sa, nwork);
      rbc_stage_end(c, ST_NEAR);
      HIP_TRY(hipGetLastError());
    }
  } else if (nitems > 0) {
    rbc_stage_begin(c, ST_PROJ);
    RBC_TRY(c->d_ikey.ensure(2 * (size_t)nitems));
    RBC_TRY(c->d_ival.ensure(2 * (size_t)nitems));
    RBC_TRY(c->d_islot.ensure(nitems));
    int *k_in = c->d_ikey.p, *k_out = k_in + nitems, *v_in = c->d_ival.p, *v_out = v_in + nitems;
    k_items<<<(nT + 255) / 256, 256, 0, st>>>(nT, c->d_nbr_cnt.p, item_off, c->d_nbr_cell.p, k_in,
                                              v_in, c->d_islot.p);
    HIP_TRY(hipGetLastError());
    int bits = 1;
    while ((1 << bits) < c->nrbc) bits++;
    size_t tmp = 0;
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, k_in, k_out, v_in, v_out, nitems, 0,
                                               bits, st));
    RBC_TRY(c->d_cub_tmp.ensure(tmp + 16));
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(c->d_cub_tmp.p, tmp, k_in, k_out, v_in, v_out, nitems,
                                               0, bits, st));
    sa.item_order = v_out;
    sa.item_slot = c->d_islot.p;
    sa.proj = tp->proj;
    k_proj<<<xcd_grid(((size_t)nitems + 255) / 256), 256, 0, st>>>(sa, nitems);
    rbc_stage_end(c, ST_PROJ);
    HIP_TRY(hipGetLastError());
    int nwork = 0;
    HIP_TRY(hipMemcpyAsync(&nwork, c->d_work_cnt.p, sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    c->last_nwork = nwork;
    if (nwork > 0) {
      rbc_stage_begin(c, ST_NEAR);
      k_near<<<(unsigned)(((size_t)nwork * 64 + 255) / 256), 256, 0, st>>>(sa, nwork);
      rbc_stage_end(c, ST_NEAR);
      HIP_TRY(hipGetLastError());
    }
  }
  // linear double-layer term
  double *xv = c->d_scratch.p + (size_t)12 * nitems + (size_t)3 * nT + 8;
  DBuf<double> &part = c->d_linpart;
  if (c2 != 0) {
    int nbk = 256;
    RBC_TRY(part.ensure(3 * nbk));
    rbc_stage_begin(c, ST_LINEAR);
    k_linear_partial<<<nbk, 256, 0, st>>>(c->np, c->d_x.p, c->d_g.p, c->d_a3.p, c->d_ds.p,
                                          c->d_node_cell.p, c->d_cells.p, part.p);
    rbc_stage_end(c, ST_LINEAR);
    double scale = -8 * RBC_PI * (c->iLb[0] * c->iLb[1] * c->iLb[2]);
    rbc_stage_begin(c, ST_LINEAR);
    k_linear_final<<<1, 64, 0, st>>>(nbk, part.p, scale, xv);
    rbc_stage_end(c, ST_LINEAR);
  }
  CombineArgs ca;
  ca.nT = nT;
  ca.c2 = c2;
  ca.tact = d_act;
  ca.cnt = c->d_nbr_cnt.p;
  ca.item_off = item_off;
  ca.tA = d_A;
  ca.vpair = c->d_vacc.p;
  ca.dvself = dvself;
  ca.dvitem = c->d_dvitem.p;
  ca.xv = xv;
  ca.has_linear = c2 != 0;
  ca.v = d_v;
  rbc_stage_begin(c, ST_COMBINE);
  k_combine<<<(nT + 255) / 256, 256, 0, st>>>(ca);
  rbc_stage_end(c, ST_COMBINE);
  HIP_TRY(hipGetLastError());
  // the pair statistics stay on the device (rbc3d_pair_stats reads them)
  if (cacheable) {
    c->cache_epoch = c->geom_epoch;
    c->cache_nT = nT;
    c->cache_nitems = nitems;
    c->cache_nwork = c->last_nwork;
  }
  return RBC3D_OK;
}
